// Shared device helpers of the gfx950 SPH hot path: error plumbing, wave64
// scans, the canonical fp32 neighbour predicate and the SPH kernel functions.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>
#include <string.h>
#include <string>
#include <type_traits>
#include "sph_mi355x.h"

namespace sph {

void set_error(const std::string& msg);
int hip_fail(const char* where);
// Checks the launch that was just issued.
int launch_check(const char* where);

constexpr int kWave = 64;

// Current value of a launch-shape knob (SPH_TUNE_*).
int tuning(int knob);

// Calls f(std::integral_constant<int, KL>) for the lanes-per-row value kl
// (a power of two from 1 to 32; the setter rejects anything else).
template <typename F>
inline void with_lanes(int kl, F&& f) {
  switch (kl) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    case 32: f(std::integral_constant<int, 32>{}); break;
    default: f(std::integral_constant<int, 8>{}); break;
  }
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

int scan_exclusive(const int* in, int n, int* out, void* ws, size_t ws_bytes, hipStream_t s);
size_t scan_ws_bytes(int n);
int fill_i32(int* p, int n, int v, hipStream_t s);
int fill_f32(float* p, int n, float v, hipStream_t s);

#define SPH_REQUIRE(cond, msg)            \
  do {                                    \
    if (!(cond)) {                        \
      ::sph::set_error(msg);              \
      return SPH_E_ARG;                   \
    }                                     \
  } while (0)

// ---------------------------------------------------------------- predicate
// Cell coordinate of a point: ceil(fl(fl(x - qmin) / hmax)) converted to
// int32, exactly the reference formula (neighborSearch.py:46,
// neighSearch_cuda.cu:122-123).  fp32 subtraction and division are IEEE
// correctly rounded under hipcc's defaults (no fast-math).
__device__ __forceinline__ int cell_coord(float x, float qmin, float hmax) {
  return (int)ceilf((x - qmin) / hmax);
}

// Candidate range [a, b) of sorted points in stencil row dy of a query in
// cell (ix, iy): cells qx in [max(ix - R, 0), ix + R] of row qy = iy + dy,
// skipped when qy < 0 (queryHashMap's negative-index rule,
// neighSearch.cpp:95), keys past the scene's table clipped.  Returns false
// when the row, and every later one, lies past the table.
__device__ __forceinline__ bool stencil_row(const SphGridDesc& d, const int* __restrict__ cs,
                                            int ix, int iy, int dy, int R, int& a, int& b) {
  a = b = 0;
  const int qy = iy + dy;
  if (qy < 0) return true;
  const int qx0 = ix - R < 0 ? 0 : ix - R;
  const int qx1 = ix + R;
  if (qx1 < qx0) return true;
  const int64_t ncell = (int64_t)d.cells_x * d.cells_y;
  const int64_t k0 = qx0 + (int64_t)d.cells_x * qy;
  int64_t k1 = qx1 + (int64_t)d.cells_x * qy;
  if (k0 >= ncell) return false;
  if (k1 >= ncell) k1 = ncell - 1;
  a = cs[k0];
  b = cs[k1 + 1];
  return true;
}

// The radius-1 stencil of a query as one sequence of candidates: rows
// dy = -1, 0, 1 are three contiguous runs of sorted points, and candidate p
// is the p-th point of their concatenation.
struct Stencil3 {
  int ra[3], rb[3];
  int n0, n01, ncand;
  __device__ __forceinline__ void init(const SphGridDesc& d, const int* __restrict__ cs, int ix,
                                       int iy) {
    bool more = true;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (more) more = stencil_row(d, cs, ix, iy, r - 1, 1, ra[r], rb[r]);
      if (!more) ra[r] = rb[r] = 0;
    }
    n0 = rb[0] - ra[0];
    n01 = n0 + rb[1] - ra[1];
    ncand = n01 + rb[2] - ra[2];
  }
  // Sorted slot of candidate p, p clamped into [0, ncand) (ncand > 0) so a
  // batch of loads needs no guard.
  __device__ __forceinline__ int slot(int p) const {
    p = p < ncand ? p : ncand - 1;
    return p < n0 ? ra[0] + p : (p < n01 ? ra[1] + (p - n0) : ra[2] + (p - n01));
  }
};

// Wavefront min / max over the active lanes (the device library's DPP
// reductions; LDS atomics in their place are serialised lane by lane).
extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);

// The stencil runs of a chunk of rows, one row per thread of the block: run
// r (stencil row dy = r - 1) is the union [lo_r, hi_r) of the rows' runs, in
// sorted slots, staged at offset off_r of a chunk-wide array of `tot` slots
// (a run no row of the chunk has is empty at lo_r = 0).
struct ChunkRuns {
  int lo0, lo1, lo2, off1, off2, tot;
  // sorted slot of staged slot k < tot
  __device__ __forceinline__ int slot(int k) const {
    return k < off1 ? lo0 + k : (k < off2 ? lo1 + (k - off1) : lo2 + (k - off2));
  }
  // staged slot of the first candidate of run r of a row's stencil
  __device__ __forceinline__ int first(int r, const Stencil3& st) const {
    return r == 0 ? st.ra[0] - lo0 : (r == 1 ? off1 + (st.ra[1] - lo1) : off2 + (st.ra[2] - lo2));
  }
};

// Every thread of the block calls this with its row's stencils in NS grids
// (valid[s] false: no row in grid s); wlo / whi are LDS scratch of 3 NS
// ints per wavefront.  Two wavefront reductions per run and one barrier
// for all NS grids; c[s] are the chunk's runs in grid s.
template <int NS>
__device__ __forceinline__ void chunk_runs_n(const Stencil3 (&st)[NS], const bool (&valid)[NS],
                                             int (*wlo)[3 * NS], int (*whi)[3 * NS],
                                             ChunkRuns (&c)[NS]) {
  int lo[3 * NS], hi[3 * NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const bool on = valid[s] && st[s].ra[r] < st[s].rb[r];
      lo[3 * s + r] = on ? st[s].ra[r] : INT_MAX;
      hi[3 * s + r] = on ? st[s].rb[r] : INT_MIN;
    }
  }
#pragma unroll
  for (int r = 0; r < 3 * NS; ++r) {
    lo[r] = __ockl_wfred_min_i32(lo[r]);
    hi[r] = __ockl_wfred_max_i32(hi[r]);
  }
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
    for (int r = 0; r < 3 * NS; ++r) {
      wlo[w][r] = lo[r];
      whi[w][r] = hi[r];
    }
  }
  __syncthreads();
  const int nw = blockDim.x / kWave;
  for (int v = 0; v < nw; ++v) {
#pragma unroll
    for (int r = 0; r < 3 * NS; ++r) {
      lo[r] = wlo[v][r] < lo[r] ? wlo[v][r] : lo[r];
      hi[r] = whi[v][r] > hi[r] ? whi[v][r] : hi[r];
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int* l = lo + 3 * s;
    const int* h = hi + 3 * s;
    const bool e0 = h[0] > l[0], e1 = h[1] > l[1], e2 = h[2] > l[2];
    // the same in every lane: held in scalar registers
    c[s].lo0 = __builtin_amdgcn_readfirstlane(e0 ? l[0] : 0);
    c[s].lo1 = __builtin_amdgcn_readfirstlane(e1 ? l[1] : 0);
    c[s].lo2 = __builtin_amdgcn_readfirstlane(e2 ? l[2] : 0);
    c[s].off1 = __builtin_amdgcn_readfirstlane(e0 ? h[0] - c[s].lo0 : 0);
    c[s].off2 = __builtin_amdgcn_readfirstlane(c[s].off1 + (e1 ? h[1] - c[s].lo1 : 0));
    c[s].tot = __builtin_amdgcn_readfirstlane(c[s].off2 + (e2 ? h[2] - c[s].lo2 : 0));
  }
}

// chunk_runs_n for one grid.
__device__ __forceinline__ ChunkRuns chunk_runs(const Stencil3& st, bool valid, int (*wlo)[3],
                                                int (*whi)[3]) {
  const Stencil3 sts[1] = {st};
  const bool v[1] = {valid};
  ChunkRuns c[1];
  chunk_runs_n<1>(sts, v, wlo, whi, c);
  return c[0];
}

// Squared distance fl(fl(dx*dx) + fl(dy*dy)) with no FMA contraction.
__device__ __forceinline__ float dist2_nofma(float xq, float yq, float xj, float yj) {
#pragma clang fp contract(off)
  const float dx = xq - xj;
  const float dy = yq - yj;
  return dx * dx + dy * dy;
}

// Canonical neighbour predicate of the reference, fl(sqrt(s)) < h with a
// correctly rounded sqrt (neighSearch.cpp:222-223), evaluated without any
// sqrt: for every float s >= 0,  fl(sqrt(s)) < h  <=>  s <= sq_threshold(h).
// Rounding is monotone, so fl(sqrt(s)) < h iff sqrt(s) < m, the midpoint of
// h and its predecessor hp = mp 2^ep (mp its 24-bit significand):
// m = (2 mp + 1) 2^(ep - 1).  m^2 = (2 mp + 1)^2 2^(2 ep - 2) has an odd
// significand of 49 or 50 bits, so it is no float: sqrt(s) never ties at m,
// and the largest float below m^2 is m^2 truncated to 24 bits.  Integer
// arithmetic only (64-bit products); exact for 2^-62 <= h <= 2^62, where
// the threshold is a normal float.  Outside that range the threshold
// saturates to 0 or the largest float.
__host__ __device__ __forceinline__ float sq_threshold(float h) {
  if (!(h > 0.f)) return -1.f;
  uint32_t hb;
  memcpy(&hb, &h, 4);
  if (hb == 0x7f800000u) return INFINITY;
  hb -= 1u;                                             // hp, the predecessor
  const uint32_t be = hb >> 23;                         // biased exponent of hp
  if (be == 0) return 0.f;                              // hp subnormal
  const uint64_t mp = (hb & 0x7fffffu) | 0x800000u;     // 24-bit significand
  const uint64_t odd = 2u * mp + 1u;                    // m = odd 2^(ep - 1)
  const uint64_t M = odd * odd;                         // < 2^50
  const int bits = 64 - __builtin_clzll(M);             // 49 or 50
  const int k = bits - 24;
  const uint32_t t = (uint32_t)(M >> k);                // in [2^23, 2^24)
  // value t 2^(2 (ep - 1) + k) with ep = be - 150; biased exponent of the
  // float t 2^E is E + 23 + 127
  const int eb = 2 * ((int)be - 151) + k + 150;
  if (eb <= 0) return 0.f;
  if (eb >= 255) return 3.4028234663852886e38f;
  const uint32_t out = ((uint32_t)eb << 23) | (t & 0x7fffffu);
  float f;
  memcpy(&f, &out, 4);
  return f;
}

// ----------------------------------------------------------- SPH kernels
// Wendland C2 in 2D (kernels.py:21-29): W = 7/pi (1-q)^4 (1+4q) / h^2,
// grad W = -rhat 7/pi / h^3 20 q (1-q)^3.  Spiky gradient (kernels.py:32-33):
// -rhat 30/pi / h^3 (1-q)^2.
constexpr float kC7pi = 2.228169203286535f;   // 7 / pi
constexpr float kC30pi = 9.549296585513721f;  // 30 / pi

__device__ __forceinline__ float wendland(float q, float h) {
  float a = 1.f - q;
  float a2 = a * a;
  return a2 * a2 * (1.f + 4.f * q) * kC7pi / (h * h);
}
// magnitude factor g with grad W = -g * rhat
__device__ __forceinline__ float wendland_grad(float q, float h) {
  float a = 1.f - q;
  return kC7pi / (h * h * h) * (20.f * q * a * a * a);
}
__device__ __forceinline__ float spiky_grad(float q, float h) {
  float a = 1.f - q;
  return kC30pi / (h * h * h) * (a * a);
}

// Hardware reciprocal and reciprocal square root (about 1 ulp); the physics
// tolerances are far above that, and they replace the ~10-instruction
// correctly rounded divide sequence in every pair.  The neighbour predicate
// never uses them.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Per-support constants of the kernels, computed once per thread.
struct KConst {
  float inv_h;  // 1 / h
  float w0;     // 7 / (pi h^2)
  float g0;     // 20 * 7 / (pi h^3)
  float s0;     // 30 / (pi h^3)
  float gr;     // -20 * 7 / (pi h^4): grad W = gr (1 - q)^3 (x_i - x_j)
};
__device__ __forceinline__ KConst kconst(float h) {
  KConst k;
  k.inv_h = 1.f / h;
  k.w0 = kC7pi * k.inv_h * k.inv_h;
  k.g0 = 20.f * kC7pi * k.inv_h * k.inv_h * k.inv_h;
  k.s0 = kC30pi * k.inv_h * k.inv_h * k.inv_h;
  k.gr = -k.g0 * k.inv_h;
  return k;
}
__device__ __forceinline__ float wendland_k(float q, const KConst& k) {
  const float a = 1.f - q;
  const float a2 = a * a;
  return a2 * a2 * (1.f + 4.f * q) * k.w0;
}
__device__ __forceinline__ float wendland_grad_k(float q, const KConst& k) {
  const float a = 1.f - q;
  return k.g0 * q * a * a * a;
}
__device__ __forceinline__ float spiky_grad_k(float q, const KConst& k) {
  const float a = 1.f - q;
  return k.s0 * a * a;
}

// Pair geometry as the reference builds it (neighborSearch.py:218-223):
// r = x_i - x_j, d = |r|, rhat = r / d (zero when d < 1e-7), q = d / h.
struct PairGeo {
  float rx, ry;    // x_i - x_j
  float d2;        // |x_i - x_j|^2
  float d;         // |x_i - x_j|
  float hx, hy;    // rhat
  float q;
};
__device__ __forceinline__ PairGeo pair_geo(float xi, float yi, float xj, float yj, float inv_h) {
  PairGeo g;
  g.rx = xi - xj;
  g.ry = yi - yj;
  g.d2 = g.rx * g.rx + g.ry * g.ry;
  const float inv_d = g.d2 > 0.f ? fast_rsq(g.d2) : 0.f;
  g.d = g.d2 * inv_d;
  const float s = g.d >= 1e-7f ? inv_d : 0.f;
  g.hx = g.rx * s;
  g.hy = g.ry * s;
  g.q = g.d * inv_h;
  return g;
}

// x^y for x > 0 in fp32 arithmetic only (a few ulp); 0 for x == 0 and
// y > 0, as pow gives.
__device__ __forceinline__ float pow_f32(float x, float y) {
  if (x == 0.f) return y > 0.f ? 0.f : INFINITY;
  return exp2f(y * log2f(x));
}

// Equation of state (pressure.py:99-111).
__device__ __forceinline__ float eos(const SphPhysics& P, float rhohat) {
  if (P.eos == 0) {
    float r2 = rhohat * rhohat;
    float r4 = r2 * r2;
    float r7 = r4 * r2 * rhohat;
    // gamma is 7 in every shipped configuration; other exponents go through
    // exp2(gamma log2(rho)) in fp32 (the device's fp32 powf reduces its
    // argument in FP64, which the path keeps for the DFSPH residual alone)
    float rg = (P.gamma == 7.f) ? r7 : pow_f32(rhohat, P.gamma);
    return P.rho0 * P.c0 * P.c0 / P.gamma * (rg - 1.f);
  }
  if (P.eos == 1) return P.c0 * P.c0 * (rhohat * P.rho0 - P.rho0);
  return P.kappa * (rhohat * P.rho0 - P.rho0);
}

// eos() with its constant factor evaluated once per thread, for loops over
// neighbours; the same arithmetic, so the same values.
struct EosK {
  float f, rho0, gamma;
  int kind;
};
__device__ __forceinline__ EosK eos_k(const SphPhysics& P) {
  EosK e;
  e.kind = P.eos;
  e.rho0 = P.rho0;
  e.gamma = P.gamma;
  e.f = P.eos == 0 ? P.rho0 * P.c0 * P.c0 / P.gamma : (P.eos == 1 ? P.c0 * P.c0 : P.kappa);
  return e;
}
__device__ __forceinline__ float eos_eval(const EosK& e, float rhohat) {
  if (e.kind == 0) {
    float r2 = rhohat * rhohat;
    float r4 = r2 * r2;
    float r7 = r4 * r2 * rhohat;
    float rg = (e.gamma == 7.f) ? r7 : pow_f32(rhohat, e.gamma);
    return e.f * (rg - 1.f);
  }
  return e.f * (rhohat * e.rho0 - e.rho0);
}

// ---------------------------------------------------- XCD-aware blocks
// The dispatcher hands consecutive workgroups to the 8 XCDs in turn, and
// each XCD has its own L2.  Rows in cell order share neighbours, so the
// logical block index is remapped to give every XCD one contiguous share of
// the rows: hardware block b runs on XCD b % 8 and becomes logical block
// (XCD's first share block) + b / 8.  The map is a bijection on
// [0, gridDim.x) for any grid size.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

// ------------------------------------------------------ lane groups
// KL consecutive lanes of a wavefront cooperate on one row: lane `sub`
// takes pairs sub, sub + KL, ...; the group's partial sums are combined with
// xor shuffles that stay inside the aligned group.
template <int KL>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = KL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
template <int KL>
__device__ __forceinline__ int group_sum_i(int v) {
#pragma unroll
  for (int o = KL / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
// mask of the lanes of this lane's group
template <int KL>
__device__ __forceinline__ unsigned long long group_mask() {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long ones = (KL == 64) ? ~0ull : ((1ull << KL) - 1ull);
  return ones << (lane & ~(KL - 1));
}
__device__ __forceinline__ unsigned long long lanes_below() {
  const int lane = threadIdx.x & (kWave - 1);
  return (1ull << lane) - 1ull;
}

// Row owner of this lane: row t (before any row order) and the lane's
// position inside its group of KL lanes (blocks remapped per XCD).
template <int KL>
struct RowLane {
  int t, sub;
};
template <int KL>
__device__ __forceinline__ RowLane<KL> row_lane() {
  const int64_t gt = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  return {(int)(gt / KL), (int)(gt % KL)};
}
template <int KL>
__device__ __forceinline__ float group_max(float v) {
#pragma unroll
  for (int o = KL / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}
template <int KL>
__device__ __forceinline__ float group_min(float v) {
#pragma unroll
  for (int o = KL / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}

// The pairs [k0, k1) of one row walked by the KL lanes of a group: lane
// `sub` takes pairs sub, sub + KL, ...; per trip it issues the loads of NB
// pairs before it uses any, so NB neighbour loads are in flight per lane.
// Slots past the row load nothing: these per-lane gathers are bound by the
// texture units' data return, which would serve re-reads as well.  load(j)
// returns the neighbour record, use(j, rec) consumes it, in the lane's pair
// order.
template <int KL, int NB = 3, class Load, class Use>
__device__ __forceinline__ void row_pairs(const int* __restrict__ col, int k0, int k1, int sub,
                                          Load load, Use use) {
  for (int kb = k0 + sub; kb < k1; kb += KL * NB) {
    bool ok[NB];
    int jj[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int k = kb + u * KL;
      ok[u] = k < k1;
      jj[u] = col[ok[u] ? k : k0];
    }
    decltype(load(0)) rec[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (ok[u]) rec[u] = load(jj[u]);
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (ok[u]) use(jj[u], rec[u]);
  }
}

// ------------------------------------------------ wave-staged column slices
// A wavefront's lane groups hold consecutive rows, whose column indices
// form one contiguous slice [rowptr[first], rowptr[last + 1]) of the CSR.
// The wave copies that slice into its own LDS window with coalesced loads
// (64 consecutive indices per instruction) and the row loops read columns
// from LDS; a per-lane strided walk of the CSR would touch one cache line
// per row per instruction.  Pairs beyond the window, and every pair of a
// wave whose rows are not consecutive, are read from global memory.
constexpr int kWaveCols = 1024;

struct WaveCols {
  int base;   // CSR position of window slot 0
  int n;      // staged slots (0: none)
  bool full;  // every pair of the wave's rows is staged (wave-uniform)
};

// Every lane of the wave calls this with the same control flow.  `row` is
// the lane's row (any value when !valid), [k0, k1) its CSR range; valid
// lanes form a prefix of the wave and lane group g holds row first + g.
template <int KL>
__device__ __forceinline__ WaveCols stage_wave_cols(const int* __restrict__ col, int row, int k0,
                                                    int k1, bool valid, int* wcols) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long vm = __ballot(valid);
  WaveCols w = {0, 0, false};
  if (vm == 0) return w;
  const int last = 63 - __clzll(vm);
  const int first_row = __shfl(row, 0, kWave);
  const bool in_seq = !valid || row == first_row + lane / KL;
  if (__ballot(!in_seq) != 0) return w;
  w.base = __shfl(k0, 0, kWave);
  const int end = __shfl(k1, last, kWave);
  w.n = min(end - w.base, kWaveCols);
  w.full = w.n == end - w.base;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int k = lane; k < w.n; k += kWave) wcols[k] = col[w.base + k];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return w;
}

// Column k from the wave's window (STAGED, the wave-uniform w.full case)
// or from global memory.  Callers pick the variant per wave: a per-pair
// choice between the two address spaces compiles to a flat load, whose wait
// drains every outstanding global load and defeats load batching.
template <bool STAGED>
__device__ __forceinline__ int wave_col(const int* __restrict__ col, const WaveCols& w,
                                        const int* wcols, int k) {
  if (STAGED) return wcols[k - w.base];
  return col[k];
}

// --------------------------------------------------------------- wave scans
__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    int t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Block-wide exclusive scan (blockDim.x <= 1024, multiple of 64).  `lds`
// needs blockDim.x / 64 + 1 ints.  Returns the exclusive prefix, total in *tot.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int* tot) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
  int inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) lds[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    int w = lane < nw ? lds[lane] : 0;
    int wi = wave_inclusive_scan(w);
    if (lane < nw) lds[lane] = wi - w;
    if (lane == nw - 1) lds[nw] = wi;
  }
  __syncthreads();
  int res = inc - v + lds[wid];
  *tot = lds[nw];
  __syncthreads();
  return res;
}

// ------------------------------------------------------------- RK4 stage
// One particle of RK4 stage `stage` = 1..4 (simulationBase.py:679-705).
// (ka, kv, kr) are the stage's rates (acceleration, velocity, dp/dt), s0 the
// state at the start of the step, k the particle's five running weighted
// sums.  Stages 1-3 return the next stage's state (c = dt/2, dt/2, dt);
// stage 4 returns the step's result (c = dt/6) and the summed acceleration
// in *asum.  Without a density rate (rate_r false) the density is carried.
// Shared by k_rk4 and the gather that fuses the stage combine; no FMA
// contraction, so both round the same way.
struct Rk4State {
  float x, y, vx, vy, r;
};
__device__ __forceinline__ Rk4State rk4_particle(int stage, const SphPhysics& P, float kax,
                                                 float kay, float kvx, float kvy, float kr,
                                                 bool rate_r, float* __restrict__ k,
                                                 const Rk4State& s0, float2* asum) {
#pragma clang fp contract(off)
  const float w = (stage == 1 || stage == 4) ? 1.f : 2.f;
  float q0, q1, q2, q3, q4;
  if (stage == 1) {
    q0 = kax; q1 = kay; q2 = kvx; q3 = kvy; q4 = kr;
  } else {
    q0 = k[0] + w * kax; q1 = k[1] + w * kay; q2 = k[2] + w * kvx; q3 = k[3] + w * kvy;
    q4 = k[4] + w * kr;
  }
  Rk4State o;
  if (stage < 4) {
    k[0] = q0; k[1] = q1; k[2] = q2; k[3] = q3; k[4] = q4;
    const float c = (stage == 3 ? 1.f : 0.5f) * P.dt;
    o.vx = s0.vx + c * kax;
    o.vy = s0.vy + c * kay;
    o.x = s0.x + c * kvx;
    o.y = s0.y + c * kvy;
    o.r = rate_r ? s0.r + c * kr / P.rho0 : s0.r;
  } else {
    const float c = (1.f / 6.f) * P.dt;
    o.vx = s0.vx + c * q0;
    o.vy = s0.vy + c * q1;
    o.x = s0.x + c * q2;
    o.y = s0.y + c * q3;
    o.r = rate_r ? s0.r + c * q4 / P.rho0 : s0.r;
    if (asum) *asum = make_float2(q0, q1);
  }
  return o;
}

// ------------------------------------------------------------- pinv2x2
// Closed-form 2x2 pseudo-inverse through the SVD angles (util.py:386-433).
__device__ __forceinline__ void pinv2x2_dev(float a, float b, float c, float d, float L[4],
                                            float eig[2]) {
  const float theta = 0.5f * atan2f(2.f * a * c + 2.f * b * d, a * a + b * b - c * c - d * d);
  const float ct = cosf(theta), st = sinf(theta);
  const float S1 = a * a + b * b + c * c + d * d;
  const float e = a * a + b * b - c * c - d * d;
  const float f2 = a * c + b * d;
  const float S2 = sqrtf(e * e + 4.f * f2 * f2);
  const float o1 = sqrtf((S1 + S2) / 2.f);
  const float o2 = sqrtf((S1 - S2) / 2.f);
  const float phi = 0.5f * atan2f(2.f * a * b + 2.f * c * d, a * a - b * b + c * c - d * d);
  const float cp = cosf(phi), sp = sinf(phi);
  const float t11 = (a * ct + c * st) * cp + (b * ct + d * st) * sp;
  const float t22 = (a * st - c * ct) * sp + (-b * st + d * ct) * cp;
  const float s11 = (t11 > 0.f) ? 1.f : ((t11 < 0.f) ? -1.f : 0.f);
  const float s22 = (t22 > 0.f) ? 1.f : ((t22 < 0.f) ? -1.f : 0.f);
  const float V00 = cp * s11, V01 = -sp * s22, V10 = sp * s11, V11 = cp * s22;
  const float U00 = ct, U01 = -st, U10 = st, U11 = ct;
  const float i1 = fabsf(o1) > 1e-5f ? 1.f / o1 : 0.f;
  const float i2 = fabsf(o2) > 1e-5f ? 1.f / o2 : 0.f;
  // (V S^-1 U^T)[r][c] = sum_k V[r][k] i_k U[c][k]
  L[0] = V00 * i1 * U00 + V01 * i2 * U01;
  L[1] = V00 * i1 * U10 + V01 * i2 * U11;
  L[2] = V10 * i1 * U00 + V11 * i2 * U01;
  L[3] = V10 * i1 * U10 + V11 * i2 * U11;
  if (fabsf(o2) > fabsf(o1)) {
    eig[0] = o2;
    eig[1] = o1;
  } else {
    eig[0] = o1;
    eig[1] = o2;
  }
}

}  // namespace sph
