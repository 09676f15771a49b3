// Kernel-sum gathers of the delta-SPH right-hand side and of the boundary
// models.  Every gather is a CSR row gather: a group of KL lanes owns one
// particle, walks its fluid neighbours and then its boundary neighbours
// (lane `sub` takes pairs sub, sub + KL, ...), keeps all sums in registers,
// combines them with shuffles and writes each output once -- no atomics, so
// results do not depend on scheduling.
//
// Two forms of the delta-SPH gather share one pair routine:
//  - sph_dsph_gather reads the structure-of-arrays state directly (the
//    module-level API, any subset of terms);
//  - sph_dsph_pack + sph_dsph_gather_packed first write one 32-byte record
//    per particle {x, y, vx, vy, rho, V, p, rho_hat} so that a neighbour is
//    two 16-byte loads from one cache line, with V = A / rho_hat and the EOS
//    evaluated once per particle instead of once per pair.
#include <climits>
#include <type_traits>

#include "dsph_pair.h"
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;
// neighbour loads a lane keeps in flight: a row loop takes kBatch pairs per
// trip, issues all their loads (slots past the row re-read pair k0), then
// adds the pairs in order -- the order of the one-pair loop
constexpr int kBatch = 4;
// The packed delta-SPH gather takes three pairs per trip and loads only the
// slots inside the row: with ~21 pairs per row over 4 lanes, batches of four
// re-read pair k0 in a third of their slots, and the data-return path of the
// texture units, which bounds this kernel, serves those re-reads too (10%
// faster on the batched benchmark than batches of four with re-reads).
constexpr int kPackedBatch = 3;

__device__ __forceinline__ float2 ld2(const float* p, int i) {
  return reinterpret_cast<const float2*>(p)[i];
}
__device__ __forceinline__ void st2(float* p, int i, float2 v) {
  reinterpret_cast<float2*>(p)[i] = v;
}

// ------------------------------------------------------------- density
template <int KL>
__global__ __launch_bounds__(kBS) void k_fluid_density(SphFluid f, const SphPhysics* __restrict__ phys,
                                                       SphCSR ff, SphBoundary b, SphCSR fb,
                                                       const int* __restrict__ order,
                                                       float* __restrict__ rho_out,
                                                       float* __restrict__ vol_out) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;   // whole groups leave together
  const int i = order ? order[rl.t] : rl.t;
  const KConst K = kconst(phys[f.pscene[i]].h);
  const float2 xi = ld2(f.x, i);
  float sum = 0.f;
  const int k0 = ff.rowptr[i], k1 = ff.rowptr[i + 1];
  for (int kb = k0 + rl.sub; kb < k1; kb += KL * kBatch) {
    bool ok[kBatch];
    int jj[kBatch];
    float2 xj[kBatch];
    float aj[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = kb + u * KL;
      ok[u] = k < k1;
      jj[u] = ff.col[ok[u] ? k : k0];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      xj[u] = ld2(f.x, jj[u]);
      aj[u] = f.area[jj[u]];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (ok[u]) {
        const PairGeo g = pair_geo(xi.x, xi.y, xj[u].x, xj[u].y, K.inv_h);
        sum += aj[u] * wendland_k(g.q, K);
      }
    }
  }
  sum = group_sum<KL>(sum);
  if (vol_out && rl.sub == 0) vol_out[i] = f.area[i] / sum;
  if (fb.rowptr) {
    float bs = 0.f;
    const int e1 = fb.rowptr[i + 1];
    for (int k = fb.rowptr[i] + rl.sub; k < e1; k += KL) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xb.x, xb.y, xi.x, xi.y, K.inv_h);
      bs += wendland_k(g.q, K) * b.vol[j];
    }
    sum += group_sum<KL>(bs);
  }
  if (rl.sub == 0) rho_out[i] = sum;
}

__device__ __forceinline__ void dsph_store(const SphPhysics& P, const Own& o, const Acc& a,
                                           unsigned terms, int i, float* dpdt, float* acc) {
  float dp, ax, ay;
  dsph_rates(P, o, a, terms, dp, ax, ay);
  if (has(terms, SPH_T_ACCUM)) {
    if (dpdt) dpdt[i] += dp;
    if (acc) {
      const float2 v = ld2(acc, i);
      st2(acc, i, make_float2(v.x + ax, v.y + ay));
    }
  } else {
    if (dpdt) dpdt[i] = dp;
    if (acc) st2(acc, i, make_float2(ax, ay));
  }
}

// structure-of-arrays form (any runtime term mask)
template <int KL>
__global__ __launch_bounds__(kBS) void k_dsph_gather(SphFluid f, const SphPhysics* __restrict__ phys,
                                                     SphCSR ff, SphBoundary b, SphCSR fb,
                                                     const float* __restrict__ gradrho,
                                                     unsigned terms, const int* __restrict__ order,
                                                     float* __restrict__ dpdt,
                                                     float* __restrict__ acc) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = order ? order[rl.t] : rl.t;
  const SphPhysics P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const bool dd = has(terms, SPH_T_DDSPH);
  Own o;
  const float2 xi = ld2(f.x, i), vi = ld2(f.v, i);
  o.x = xi.x; o.y = xi.y; o.vx = vi.x; o.vy = vi.y;
  o.rho = f.rho[i] * P.rho0;
  o.p = eos(P, f.rho[i]);
  const float2 gri = dd ? ld2(gradrho, i) : make_float2(0.f, 0.f);
  o.gx = gri.x; o.gy = gri.y;
  Acc a = {};
  if (terms & kFluidPairTerms) {
    const int k1 = ff.rowptr[i + 1];
    for (int k = ff.rowptr[i] + rl.sub; k < k1; k += KL) {
      const int j = ff.col[k];
      const float2 xj = ld2(f.x, j), vj = ld2(f.v, j);
      Nbr n;
      n.x = xj.x; n.y = xj.y; n.vx = vj.x; n.vy = vj.y;
      n.rho = f.rho[j] * P.rho0;
      n.vol = f.vol ? f.vol[j] : f.area[j] / f.rho[j];
      n.p = has(terms, SPH_T_PRESS) ? eos(P, f.rho[j]) : 0.f;
      const float2 grj = dd ? ld2(gradrho, j) : make_float2(0.f, 0.f);
      dsph_pair(P, K, o, n, grj.x, grj.y, true, terms, a);
    }
  }
  if (fb.rowptr && (terms & kBdyPairTerms)) {
    const int e1 = fb.rowptr[i + 1];
    for (int k = fb.rowptr[i] + rl.sub; k < e1; k += KL) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const float2 vb = b.v ? ld2(b.v, j) : make_float2(0.f, 0.f);
      Nbr n;
      n.x = xb.x; n.y = xb.y; n.vx = vb.x; n.vy = vb.y;
      n.rho = b.rho[j] * P.rho0;
      n.vol = b.vol[j] / b.rho[j];
      n.p = b.p ? b.p[j] : 0.f;
      dsph_pair(P, K, o, n, 0.f, 0.f, false, terms, a);
    }
  }
  reduce_acc<KL>(a, terms);
  if (rl.sub == 0) dsph_store(P, o, a, terms, i, dpdt, acc);
}

// A right-hand side without pair terms (gravity alone, as the DFSPH
// prologue asks for): the gather's epilogue on empty sums, one thread per
// row, so dp/dt and a are those of k_dsph_gather bit for bit.
__global__ __launch_bounds__(kBS) void k_dsph_rates_only(SphFluid f,
                                                         const SphPhysics* __restrict__ phys,
                                                         unsigned terms, float* __restrict__ dpdt,
                                                         float* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.np) return;
  const SphPhysics P = phys[f.pscene[i]];
  Own o = {};
  o.rho = f.rho[i] * P.rho0;
  const Acc a = {};
  dsph_store(P, o, a, terms, i, dpdt, acc);
}

// packed records: rec[2 i] = {x, y, vx, vy}, rec[2 i + 1] = {rho, V, p, rho_hat}
__device__ __forceinline__ void pack_fluid(float4* __restrict__ frec, int t, const SphPhysics& P,
                                           float x, float y, float vx, float vy, float rh,
                                           float vol) {
  frec[2 * t] = make_float4(x, y, vx, vy);
  frec[2 * t + 1] = make_float4(rh * P.rho0, vol, eos(P, rh), rh);
}

// Threads t0 + thread id: fluid rows below f.np (skipped when t0 = f.np),
// boundary rows above.
__global__ __launch_bounds__(kBS) void k_dsph_pack(SphFluid f, const SphPhysics* __restrict__ phys,
                                                   SphBoundary b, float4* __restrict__ frec,
                                                   float4* __restrict__ brec, int t0) {
  const int t = t0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (t < f.np) {
    const SphPhysics P = phys[f.pscene[t]];
    const float2 x = ld2(f.x, t), v = ld2(f.v, t);
    const float rh = f.rho[t];
    pack_fluid(frec, t, P, x.x, x.y, v.x, v.y, rh, f.vol ? f.vol[t] : f.area[t] / rh);
  } else if (t < f.np + b.nb) {
    const int k = t - f.np;
    const SphPhysics P = phys[b.bscene[k]];
    const float2 x = ld2(b.x, k);
    const float2 v = b.v ? ld2(b.v, k) : make_float2(0.f, 0.f);
    const float rh = b.rho[k];
    brec[2 * k] = make_float4(x.x, x.y, v.x, v.y);
    brec[2 * k + 1] = make_float4(rh * P.rho0, b.vol[k] / rh, b.p ? b.p[k] : 0.f, rh);
  }
}

// Boundary pairs of fluid row i from the fluid -> boundary list, lane
// `sub` of KL taking pairs sub, sub + KL, ...
template <int KL>
__device__ __forceinline__ void packed_bdy_pairs(const SphPhysics& P, const KConst& K,
                                                 const Own& o, const float4* __restrict__ brec,
                                                 const SphCSR& fb, int i, int sub,
                                                 unsigned terms, Acc& a) {
  if (!fb.rowptr || !(terms & kBdyPairTerms)) return;
  const int k1 = fb.rowptr[i + 1];
  for (int k = fb.rowptr[i] + sub; k < k1; k += KL) {
    const int j = fb.col[k];
    const float4 r0 = brec[2 * j], r1 = brec[2 * j + 1];
    const Nbr n = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z};
    dsph_pair(P, K, o, n, 0.f, 0.f, false, terms, a);
  }
}

// What the RK4 stage combine of row i reads besides the row's sums and the
// running sums K: its state at the step's start and its area (stages 1-3,
// for the next stage's record).  Loaded apart from the combine so that a
// kernel can issue these loads long before it needs them.  (K is read in
// the combine itself: held as well, its five registers spill in the cell
// walk.)
struct Rk4In {
  Rk4State s0;
  float area;
};
__device__ __forceinline__ Rk4In load_rk4_in(const SphRk4Stage& rk, int i) {
  Rk4In r;
  r.s0 = {rk.x0[2 * i], rk.x0[2 * i + 1], rk.v0[2 * i], rk.v0[2 * i + 1], rk.r0[i]};
  r.area = rk.stage < 4 && rk.frec_next ? rk.area[i] : 1.f;
  return r;
}

// Output of row i from its reduced sums.  FUSE: the row's RK4 stage
// combine (rk4_particle) and the next stage's record instead of dp/dt and
// a (see sph_dsph_gather_rk4), with `in` the row's load_rk4_in.
template <bool FUSE>
__device__ __forceinline__ void packed_row_out(const SphPhysics& P, const Own& o, const Acc& a,
                                               unsigned terms, int i, float* __restrict__ dpdt,
                                               float* __restrict__ acc, const SphRk4Stage& rk,
                                               const Rk4In& in) {
  if (!FUSE) {
    dsph_store(P, o, a, terms, i, dpdt, acc);
    return;
  }
  float dp, ax, ay;
  dsph_rates(P, o, a, terms, dp, ax, ay);
  float2 asum;
  const Rk4State n = rk4_particle(rk.stage, P, ax, ay, o.vx, o.vy, dp, true,
                                  rk.K + 5 * (size_t)i, in.s0, &asum);
  st2(rk.x_out, i, make_float2(n.x, n.y));
  if (rk.v_out) st2(rk.v_out, i, make_float2(n.vx, n.vy));
  rk.rho_out[i] = n.r;
  if (rk.stage == 4) {
    if (rk.acc_out) st2(rk.acc_out, i, asum);
  } else if (rk.frec_next) {
    pack_fluid(reinterpret_cast<float4*>(rk.frec_next), i, P, n.x, n.y, n.vx, n.vy, n.r,
               in.area / n.r);
  }
}

// FUSE: the row's RK4 stage combine (rk4_particle) and the next stage's
// record run in the epilogue instead of storing dp/dt and a (see
// sph_dsph_gather_rk4); the record buffers of successive stages alternate,
// since neighbours still read this stage's records.
template <unsigned T, int KL, bool FUSE>
__global__ __launch_bounds__(kBS) void k_dsph_gather_packed(
    const float4* __restrict__ frec, const float4* __restrict__ brec,
    const int* __restrict__ pscene, int np, const SphPhysics* __restrict__ phys, SphCSR ff,
    SphCSR fb, const float* __restrict__ gradrho, unsigned rterms, const int* __restrict__ order,
    float* __restrict__ dpdt, float* __restrict__ acc, SphRk4Stage rk,
    const float* __restrict__ smask) {
  const unsigned terms = T ? T : rterms;
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= np) return;
  const int i = order ? order[rl.t] : rl.t;
  const SphPhysics P = phys[pscene[i]];
  const KConst K = kconst(P.h);
  const bool dd = has(terms, SPH_T_DDSPH);
  const float4 a0 = frec[2 * i], a1 = frec[2 * i + 1];
  Own o;
  o.x = a0.x; o.y = a0.y; o.vx = a0.z; o.vy = a0.w;
  o.rho = a1.x;
  o.p = a1.z;
  // delta-plus switch (pressure.py:40-52): p_j - p_i for a row with negative
  // pressure in the bulk class; the pair term adds o.p + p_j, so the row's
  // own pressure enters with its sign flipped
  if (smask && !(o.p >= 0.f || smask[i] < 1.5f)) o.p = -o.p;
  const float2 gri = dd ? ld2(gradrho, i) : make_float2(0.f, 0.f);
  o.gx = gri.x; o.gy = gri.y;
  Acc a = {};
  if (terms & kFluidPairTerms) {
    const int k0 = ff.rowptr[i], k1 = ff.rowptr[i + 1];
    for (int kb = k0 + rl.sub; kb < k1; kb += KL * kPackedBatch) {
      bool ok[kPackedBatch];
      int jj[kPackedBatch];
      float4 r0[kPackedBatch], r1[kPackedBatch];
      float2 grj[kPackedBatch];
#pragma unroll
      for (int u = 0; u < kPackedBatch; ++u) {
        const int k = kb + u * KL;
        ok[u] = k < k1;
        jj[u] = ff.col[ok[u] ? k : k0];
      }
#pragma unroll
      for (int u = 0; u < kPackedBatch; ++u) {
        if (ok[u]) {
          r0[u] = frec[2 * jj[u]];
          r1[u] = frec[2 * jj[u] + 1];
        } else {
          r0[u] = make_float4(0.f, 0.f, 0.f, 0.f);
          r1[u] = r0[u];
        }
        grj[u] = dd ? ld2(gradrho, jj[u]) : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < kPackedBatch; ++u) {
        if (ok[u]) {
          const Nbr n = {r0[u].x, r0[u].y, r0[u].z, r0[u].w, r1[u].x, r1[u].y, r1[u].z};
          dsph_pair(P, K, o, n, grj[u].x, grj[u].y, true, terms, a);
        }
      }
    }
  }
  packed_bdy_pairs<KL>(P, K, o, brec, fb, i, rl.sub, terms, a);
  reduce_acc<KL>(a, terms);
  if (rl.sub == 0) {
    Rk4In in = {};
    if (FUSE) in = load_rk4_in(rk, i);
    packed_row_out<FUSE>(P, o, a, terms, i, dpdt, acc, rk, in);
  }
}

// ------------------------------------------- cell-walk delta-SPH gather
// The packed gather with its fluid pairs found in-kernel instead of read
// from a pair list (the default MOG delta-SPH right-hand side needs no list
// for anything else).  A workgroup takes chunks of kWalkRows consecutive
// sorted points of one scene.  The three stencil rows of a run of sorted
// points are three contiguous runs of sorted points (each about the run
// plus a cell at either end), so the chunk works in three phases:
//  1. stage: the positions and point indices of those runs are copied into
//     LDS, coalesced;
//  2. test: one lane per row walks its stencil in LDS in the order the
//     pair-list fill visits it (Stencil3, the same fp32 predicate) and
//     writes the hits into the row's list in LDS, up to kWalkCap;
//  3. gather: KL lanes per row walk that list as the packed gather walks a
//     CSR row -- the same pairs, in the same order, on the same lanes -- so
//     dp/dt and a are bit for bit those of sph_dsph_gather_packed over the
//     list of sph_neighbor_fill.
// A row with more hits than its list holds walks its stencil again from
// global memory, the lanes of its group testing KL candidates at a time and
// each lane taking the unlisted hit of its rank; a chunk whose stencil does
// not fit in LDS (crowded cells, points outside the cell table) does all
// its rows that way.
constexpr int kWalkRows = kBS;                  // rows per chunk
// staged stencil points per chunk: three stencil rows of a chunk's kWalkRows
// points, with room for neighbouring cell rows half again as full as the
// chunk's own (cells of one row hold 2 or 3 lattice rows at the support of
// ~20 neighbours) and a cell at either end
constexpr int kWalkRecs = 3 * kWalkRows * 3 / 2 + 256;
constexpr int kWalkCap = 27;                    // listed hits per row
constexpr int kWalkBatch = 4;                   // candidates tested per trip
constexpr int kBdyBatch = 2;                    // boundary candidates per trip
// The stage loads its slots in rounds of kStageRound per thread: a round
// issues all its loads before it writes any to LDS, so the stage waits on
// memory once per round instead of once per slot (six slots per round
// spill registers at the kernel's 96).
constexpr int kStageRound = 3;

struct WalkShared {
  // sorted positions of the staged runs; the test loads up to
  // kWalkBatch - 1 slots past a run's end (no hit), hence the padding
  float2 pos[kWalkRecs + kWalkBatch];
  int ridx[kWalkRecs];             // their point indices (record rows)
  unsigned short list[kWalkRows][kWalkCap + 1];   // slot kWalkCap: scratch
  unsigned short cnt[kWalkRows];   // listed hits, kWalkCap + 1 = more
  int wlo[kBS / kWave][3], whi[kBS / kWave][3];   // chunk_runs scratch
};

// Listed fluid pairs of a row: entry k (rank k) on lane k mod KL, in
// batches of kPackedBatch.  Every slot of a batch loads (slots past the row
// re-read its last entry), so no divergent branch surrounds the loads.
template <int KL>
__device__ __forceinline__ void walk_list(const SphPhysics& P, const KConst& K, const Own& o,
                                          unsigned terms, const unsigned short* lst, int n,
                                          int sub, const WalkShared& S,
                                          const float4* __restrict__ frec, Acc& a) {
  for (int kb = sub; kb < n; kb += KL * kPackedBatch) {
    float4 r0[kPackedBatch], r1[kPackedBatch];
#pragma unroll
    for (int u = 0; u < kPackedBatch; ++u) {
      const int k = kb + u * KL;
      const int j = S.ridx[lst[k < n ? k : n - 1]];
      r0[u] = frec[2 * j];
      r1[u] = frec[2 * j + 1];
    }
#pragma unroll
    for (int u = 0; u < kPackedBatch; ++u) {
      if (kb + u * KL < n) {
        const Nbr nb = {r0[u].x, r0[u].y, r0[u].z, r0[u].w, r1[u].x, r1[u].y, r1[u].z};
        dsph_pair(P, K, o, nb, 0.f, 0.f, true, terms, a);
      }
    }
  }
}

// Fluid pairs of rank >= skip of the rows with `slow` set, from global
// memory: the group tests KL candidates per trip and lane L takes the hit of
// rank congruent to L mod KL.  Every lane of the wavefront runs the loop
// (ballots and shuffles), so call it converged.
template <int KL>
__device__ __forceinline__ void walk_slow(bool slow, const SphGrid& g, int sc, int sp, float thr,
                                          int skip, int sub, const SphPhysics& P,
                                          const KConst& K, const Own& o, unsigned terms,
                                          const float4* __restrict__ frec, Acc& a) {
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  Stencil3 st;
  int ncand = 0;
  float xq = 0.f, yq = 0.f;
  if (slow) {
    const SphGridDesc d = g.scenes[sc];
    const float2 q = sx[sp];
    xq = q.x;
    yq = q.y;
    st.init(d, g.cell_start + d.cell_begin, cell_coord(xq, d.qmin_x, d.hmax),
            cell_coord(yq, d.qmin_y, d.hmax));
    ncand = st.ncand;
  }
  const unsigned long long gmask = group_mask<KL>();
  const int lane = threadIdx.x & (kWave - 1);
  int base = 0;   // rank of the trip's first hit
  for (int p0 = 0; __ballot(slow && p0 < ncand) != 0; p0 += KL) {
    const int p = p0 + sub;
    bool hit = false;
    int slot = 0;
    if (slow && p < ncand) {
      slot = st.slot(p);
      const float2 xy = sx[slot];
      hit = dist2_nofma(xq, yq, xy.x, xy.y) <= thr;
    }
    const unsigned long long m = __ballot(hit) & gmask;
    const int nh = __popcll(m);
    const int want = (sub - base) & (KL - 1);   // rank of this lane's hit in the trip
    unsigned long long mm = m;
    for (int i = 0; i < want && mm; ++i) mm &= mm - 1;
    const int src = mm ? __ffsll((long long)mm) - 1 : lane;
    const int js = __shfl(slot, src, kWave);
    if (slow && want < nh && base + want >= skip) {
      const int j = g.sorted_idx[js];
      const float4 r0 = frec[2 * j], r1 = frec[2 * j + 1];
      const Nbr nb = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z};
      dsph_pair(P, K, o, nb, 0.f, 0.f, true, terms, a);
    }
    base += nh;
  }
}

// Boundary pairs of the fluid point at (xq, yq) from the boundary grid bg:
// the boundary points of its stencil st in bg (ncand candidates, 0: no
// row) in bg's sorted order that pass the predicate -- the row of a fluid
// -> boundary search against bg (sph_neighbor_fill with the fluid points as
// queries) -- walked with the packed gather's lane assignment (the hit of
// rank k on lane k mod KL).  Every lane of the wavefront runs the loop
// (ballots and shuffles), so call it converged.
template <int KL>
__device__ __forceinline__ void walk_bdy(const Stencil3& st, int ncand, const SphGrid& bg,
                                         float xq, float yq, float thr, int sub,
                                         const SphPhysics& P, const KConst& K, const Own& o,
                                         unsigned terms, const float4* __restrict__ brec,
                                         Acc& a) {
  const float2* __restrict__ bsx = reinterpret_cast<const float2*>(bg.sorted_x);
  if constexpr (KL == 1) {
    // kBdyBatch candidates per trip: their positions and indices loaded
    // together, then the records of the trip's hits together, and the hits
    // added in candidate order.  (One candidate per trip made a row by a
    // wall a chain of ~40 dependent loads that held its whole wavefront.)
    for (int p0 = 0; p0 < ncand; p0 += kBdyBatch) {
      float2 xy[kBdyBatch];
      int bi[kBdyBatch];
#pragma unroll
      for (int u = 0; u < kBdyBatch; ++u) {
        const int slot = st.slot(p0 + u);   // clamped into the candidates
        xy[u] = bsx[slot];
        bi[u] = bg.sorted_idx[slot];
      }
      bool hit[kBdyBatch];
      float4 r0[kBdyBatch], r1[kBdyBatch];
#pragma unroll
      for (int u = 0; u < kBdyBatch; ++u) {
        hit[u] = p0 + u < ncand && dist2_nofma(xq, yq, xy[u].x, xy[u].y) <= thr;
        if (hit[u]) {
          r0[u] = brec[2 * bi[u]];
          r1[u] = brec[2 * bi[u] + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < kBdyBatch; ++u) {
        if (hit[u]) {
          const Nbr nb = {r0[u].x, r0[u].y, r0[u].z, r0[u].w, r1[u].x, r1[u].y, r1[u].z};
          dsph_pair(P, K, o, nb, 0.f, 0.f, false, terms, a);
        }
      }
    }
  } else {
    const unsigned long long gmask = group_mask<KL>();
    const int lane = threadIdx.x & (kWave - 1);
    int base = 0;   // rank of the trip's first hit
    for (int p0 = 0; __ballot(p0 < ncand) != 0; p0 += KL) {
      const int p = p0 + sub;
      bool hit = false;
      int slot = 0;
      if (p < ncand) {
        slot = st.slot(p);
        const float2 xy = bsx[slot];
        hit = dist2_nofma(xq, yq, xy.x, xy.y) <= thr;
      }
      const unsigned long long m = __ballot(hit) & gmask;
      const int nh = __popcll(m);
      const int want = (sub - base) & (KL - 1);   // rank of this lane's hit in the trip
      unsigned long long mm = m;
      for (int i = 0; i < want && mm; ++i) mm &= mm - 1;
      const int src = mm ? __ffsll((long long)mm) - 1 : lane;
      const int js = __shfl(slot, src, kWave);
      if (want < nh) {
        const int b = bg.sorted_idx[js];
        const float4 r0 = brec[2 * b], r1 = brec[2 * b + 1];
        const Nbr nb = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z};
        dsph_pair(P, K, o, nb, 0.f, 0.f, false, terms, a);
      }
      base += nh;
    }
  }
}

// The stencil of the point at (xq, yq) in the boundary grid's scene bd;
// (fix, fiy) is its cell in the fluid grid, taken when bd has the fluid
// grid's origin and cell size (`same`).
__device__ __forceinline__ void bdy_stencil(const SphGrid& bg, const SphGridDesc& bd, bool same,
                                            float xq, float yq, int fix, int fiy, Stencil3& st) {
  st.init(bd, bg.cell_start + bd.cell_begin, same ? fix : cell_coord(xq, bd.qmin_x, bd.hmax),
          same ? fiy : cell_coord(yq, bd.qmin_y, bd.hmax));
}

// The chunks of one workgroup's segment [s0, s1) of sorted slots.  UNI:
// every row of the segment is in scene usc, whose grid, constants and
// threshold are then the same in every lane: loaded once, into scalar
// registers, which leaves the vector registers to the pair loop.
template <unsigned T, int KL, bool FUSE, bool UNI>
__device__ __forceinline__ void cellwalk_segment(WalkShared& S, int s0, int s1, int usc,
                                                 const SphGrid& g, const float4* __restrict__ frec,
                                                 const float4* __restrict__ brec,
                                                 const SphPhysics* __restrict__ phys,
                                                 const SphCSR& fb, const SphGrid& bg,
                                                 unsigned terms, float* __restrict__ dpdt,
                                                 float* __restrict__ acc, const SphRk4Stage& rk) {
  const int t = threadIdx.x;
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  SphGridDesc ud = {};
  float uthr = 0.f;
  if (UNI) {
    ud = g.scenes[usc];
    uthr = sq_threshold(phys[usc].h);
  }
  for (int c0 = s0; c0 < s1; c0 += kWalkRows) {
    // ---- the stencil of row t (one lane per row)
    const int sp = c0 + t;
    const bool valid = sp < s1;
    Stencil3 st;
    SphGridDesc d = ud;
    float xq = 0.f, yq = 0.f, thr = uthr;
    int qi = 0, sc = UNI ? usc : 0, ix = 0, iy = 0;
    float4 a0 = {}, a1 = {};
    Rk4In rin = {};
    if (valid) {
      qi = g.sorted_idx[sp];
      if (!UNI) {
        sc = g.nscenes == 1 ? 0 : g.pscene[qi];
        d = g.scenes[sc];
        thr = sq_threshold(phys[sc].h);
      }
      const float2 q = sx[sp];
      xq = q.x;
      yq = q.y;
      ix = cell_coord(xq, d.qmin_x, d.hmax);
      iy = cell_coord(yq, d.qmin_y, d.hmax);
      st.init(d, g.cell_start + d.cell_begin, ix, iy);
      // one lane per row: the row's own record and RK4 inputs are loaded
      // here, their wait hidden behind the stage's
      if (KL == 1) {
        a0 = frec[2 * qi];
        a1 = frec[2 * qi + 1];
        if (FUSE) rin = load_rk4_in(rk, qi);
      }
    }
    // ---- stage: the chunk's three stencil runs, positions and indices
    const ChunkRuns cr = chunk_runs(st, valid, S.wlo, S.whi);
    const int tot = cr.tot;
    const bool staged = tot <= kWalkRecs;
    if (staged) {
      // a wavefront runs a round when any of its slots is staged; slots
      // past the end re-read the last one, so no divergent branch surrounds
      // the loads and their values stay in registers
      const int w0 = __builtin_amdgcn_readfirstlane(t);
      for (int u0 = 0; w0 + u0 * kBS < tot; u0 += kStageRound) {
        float2 pv[kStageRound];
        int iv[kStageRound];
#pragma unroll
        for (int u = 0; u < kStageRound; ++u) {
          const int k = t + (u0 + u) * kBS;
          const int sl = cr.slot(k < tot ? k : tot - 1);
          pv[u] = sx[sl];
          iv[u] = g.sorted_idx[sl];
        }
#pragma unroll
        for (int u = 0; u < kStageRound; ++u) {
          const int k = t + (u0 + u) * kBS;
          if (k < tot) {
            S.pos[k] = pv[u];
            S.ridx[k] = iv[u];
          }
        }
      }
    }
    __syncthreads();
    // ---- test: row t's hits, in candidate order; whole batches of
    // kWalkBatch LDS loads need no bound, a run's last partial batch is
    // masked
    if (valid && staged) {
      unsigned short* lst = S.list[t];
      int cnt = 0;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int b = cr.first(r, st), e = b + (st.rb[r] - st.ra[r]);
        int k = b;
        for (; k + kWalkBatch <= e; k += kWalkBatch) {
          float2 xy[kWalkBatch];
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u) xy[u] = S.pos[k + u];
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u) {
            const float d2 = dist2_nofma(xq, yq, xy[u].x, xy[u].y);
            lst[cnt < kWalkCap ? cnt : kWalkCap] = (unsigned short)(k + u);
            cnt += d2 <= thr ? 1 : 0;
          }
        }
        if (k < e) {
          float2 xy[kWalkBatch];
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u) xy[u] = S.pos[k + u];
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u) {
            const float d2 = dist2_nofma(xq, yq, xy[u].x, xy[u].y);
            const bool hit = (d2 <= thr) & (k + u < e);
            lst[cnt < kWalkCap ? cnt : kWalkCap] = (unsigned short)(k + u);
            cnt += hit ? 1 : 0;
          }
        }
      }
      S.cnt[t] = (unsigned short)(cnt > kWalkCap ? kWalkCap + 1 : cnt);
    }
    // with one lane per row a row's list is its own thread's: no wait
    if (KL > 1) __syncthreads();
    // ---- gather: KL lanes per row.  With one lane per row, row rr is the
    // thread's own row of the test, whose index, scene, cell and threshold
    // are still in registers.
    const int nrow = s1 - c0 < kWalkRows ? s1 - c0 : kWalkRows;
    const int sub = t % KL;
    for (int rr = t / KL; rr < nrow; rr += kBS / KL) {
      const int spr = c0 + rr;
      if (KL > 1) {
        qi = g.sorted_idx[spr];
        if (!UNI) {
          sc = g.nscenes == 1 ? 0 : g.pscene[qi];
          d = g.scenes[sc];
        }
        const float2 q = sx[spr];
        xq = q.x;
        yq = q.y;
        ix = cell_coord(xq, d.qmin_x, d.hmax);
        iy = cell_coord(yq, d.qmin_y, d.hmax);
      }
      const SphPhysics P = phys[sc];
      const KConst K = kconst(P.h);
      if (KL > 1 && !UNI) thr = sq_threshold(P.h);
      if (KL > 1) {
        a0 = frec[2 * qi];
        a1 = frec[2 * qi + 1];
      }
      Own o;
      o.x = a0.x; o.y = a0.y; o.vx = a0.z; o.vy = a0.w;
      o.rho = a1.x;
      o.p = a1.z;
      o.gx = 0.f;
      o.gy = 0.f;
      Acc a = {};
      if (terms & kFluidPairTerms) {
        const int n = staged ? S.cnt[rr] : 0;
        walk_list<KL>(P, K, o, terms, S.list[rr], n < kWalkCap ? n : kWalkCap, sub, S, frec, a);
        const bool slow = !staged || n > kWalkCap;
        if (__ballot(slow) != 0)
          walk_slow<KL>(slow, g, sc, spr, thr, staged ? kWalkCap : 0, sub, P, K, o, terms, frec,
                        a);
      }
      if (bg.cell_start) {
        if (terms & kBdyPairTerms) {
          const SphGridDesc bd = bg.scenes[bg.nscenes == 1 ? 0 : sc];
          const bool same = bd.qmin_x == d.qmin_x && bd.qmin_y == d.qmin_y && bd.hmax == d.hmax;
          Stencil3 bst;
          bdy_stencil(bg, bd, same, xq, yq, ix, iy, bst);
          walk_bdy<KL>(bst, bst.ncand, bg, xq, yq, thr, sub, P, K, o, terms, brec, a);
        }
      } else {
        packed_bdy_pairs<KL>(P, K, o, brec, fb, qi, sub, terms, a);
      }
      reduce_acc<KL>(a, terms);
      if (sub == 0) {
        if (FUSE && KL > 1) rin = load_rk4_in(rk, qi);
        packed_row_out<FUSE>(P, o, a, terms, qi, dpdt, acc, rk, rin);
      }
    }
    // the next chunk restages the LDS arrays
    if (c0 + kWalkRows < s1) __syncthreads();
  }
}

// ---- one lane per row: the neighbours' records staged in LDS
// With one lane per row the test and the pair sums of a row run in the same
// thread, so a row's hits stay in registers, as one 64-bit mask over its
// three stencil runs' candidates laid end to end, and the stage copies,
// besides the runs' positions, the fields of their packed records the pair
// routine reads: {vx, vy, rho, V} and p.  The pair loop then reads LDS
// only, each record once per chunk from memory instead of once per pair.
// Positions come from the grid (sorted_x), the same values as the records'
// own.  A row with more than 64 candidates, or a chunk whose runs do not
// fit, takes the global-memory walk (walk_slow) for all its fluid pairs.
constexpr int kWalk1Cap = 1088;        // staged stencil points per chunk
constexpr int kWalk1MaxCand = 64;      // candidates of a row's three runs a mask holds
constexpr int kWalk1Round = 4;         // staged slots per thread in flight

struct Walk1Shared {
  float2 pos[kWalk1Cap + kWalkBatch];  // the test reads past a run's end
  float4 rec[kWalk1Cap];               // {vx, vy, rho, V}
  float p[kWalk1Cap];
  int wlo[kBS / kWave][3], whi[kBS / kWave][3];
};

template <unsigned T, bool FUSE, bool UNI>
__device__ __forceinline__ void cellwalk_segment1(Walk1Shared& S, int s0, int s1, int usc,
                                                  const SphGrid& g,
                                                  const float4* __restrict__ frec,
                                                  const float4* __restrict__ brec,
                                                  const SphPhysics* __restrict__ phys,
                                                  const SphCSR& fb, const SphGrid& bg,
                                                  unsigned terms, float* __restrict__ dpdt,
                                                  float* __restrict__ acc,
                                                  const SphRk4Stage& rk) {
  const int t = threadIdx.x;
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  SphGridDesc ud = {};
  float uthr = 0.f;
  if (UNI) {
    ud = g.scenes[usc];
    uthr = sq_threshold(phys[usc].h);
  }
  for (int c0 = s0; c0 < s1; c0 += kWalkRows) {
    // ---- the stencil of row t, its own record and RK4 inputs
    const int sp = c0 + t;
    const bool valid = sp < s1;
    Stencil3 st;
    SphGridDesc d = ud;
    float xq = 0.f, yq = 0.f, thr = uthr;
    int qi = 0, sc = UNI ? usc : 0, ix = 0, iy = 0;
    float4 a0 = {}, a1 = {};
    Rk4In rin = {};
    if (valid) {
      qi = g.sorted_idx[sp];
      if (!UNI) {
        sc = g.nscenes == 1 ? 0 : g.pscene[qi];
        d = g.scenes[sc];
        thr = sq_threshold(phys[sc].h);
      }
      const float2 q = sx[sp];
      xq = q.x;
      yq = q.y;
      ix = cell_coord(xq, d.qmin_x, d.hmax);
      iy = cell_coord(yq, d.qmin_y, d.hmax);
      st.init(d, g.cell_start + d.cell_begin, ix, iy);
      a0 = frec[2 * qi];
      a1 = frec[2 * qi + 1];
      if (FUSE) rin = load_rk4_in(rk, qi);
    }
    // ---- stage: the runs' positions and the fields of their records
    const ChunkRuns cr = chunk_runs(st, valid, S.wlo, S.whi);
    const int tot = cr.tot;
    const bool staged = tot <= kWalk1Cap;
    if (staged) {
      const int w0 = __builtin_amdgcn_readfirstlane(t);
      for (int u0 = 0; w0 + u0 * kBS < tot; u0 += kWalk1Round) {
        float2 pv[kWalk1Round];
        int jv[kWalk1Round];
#pragma unroll
        for (int u = 0; u < kWalk1Round; ++u) {
          const int k = t + (u0 + u) * kBS;
          const int sl = cr.slot(k < tot ? k : tot - 1);
          pv[u] = sx[sl];
          jv[u] = g.sorted_idx[sl];
        }
        float4 r0[kWalk1Round], r1[kWalk1Round];
#pragma unroll
        for (int u = 0; u < kWalk1Round; ++u) {
          r0[u] = frec[2 * jv[u]];
          r1[u] = frec[2 * jv[u] + 1];
        }
#pragma unroll
        for (int u = 0; u < kWalk1Round; ++u) {
          const int k = t + (u0 + u) * kBS;
          if (k < tot) {
            S.pos[k] = pv[u];
            S.rec[k] = make_float4(r0[u].z, r0[u].w, r1[u].x, r1[u].y);
            S.p[k] = r1[u].z;
          }
        }
      }
    }
    __syncthreads();
    // ---- test: the row's hits as bits of one mask over its three runs'
    // candidates, run after run: bit t1 starts the second run, t2 the third
    unsigned long long m = 0ull;
    // (a lane past the segment has no stencil: nothing of it is read)
    const int len0 = valid ? st.rb[0] - st.ra[0] : 0, len1 = valid ? st.rb[1] - st.ra[1] : 0;
    const int t1 = len0, t2 = len0 + len1;
    bool slow = !staged || (valid && t2 + (st.rb[2] - st.ra[2]) > kWalk1MaxCand);
    if (valid && !slow) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int n = st.rb[r] - st.ra[r];
        const int b = cr.first(r, st);
        const int off = r == 0 ? 0 : (r == 1 ? t1 : t2);
        for (int k = 0; k < n; k += kWalkBatch) {
          float2 xy[kWalkBatch];
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u) xy[u] = S.pos[b + k + u];
          unsigned nib = 0;   // the batch's hits, then set four at a time
#pragma unroll
          for (int u = 0; u < kWalkBatch; ++u)
            nib |= ((k + u < n) && dist2_nofma(xq, yq, xy[u].x, xy[u].y) <= thr) ? 1u << u : 0u;
          m |= (unsigned long long)nib << (off + k);
        }
      }
    }
    // ---- gather: the row's own pairs
    if (valid) {
      const SphPhysics P = phys[sc];
      const KConst K = kconst(P.h);
      Own o;
      o.x = a0.x; o.y = a0.y; o.vx = a0.z; o.vy = a0.w;
      o.rho = a1.x;
      o.p = a1.z;
      o.gx = 0.f;
      o.gy = 0.f;
      Acc a = {};
      if (terms & kFluidPairTerms) {
        if (!slow) {
          // one stream over the three runs (the runs' pairs in order), so
          // a wavefront pads to its longest row, not to its longest run
          // three times over
          const int b0 = cr.first(0, st), b1 = cr.first(1, st) - t1, b2 = cr.first(2, st) - t2;
          unsigned long long mm = m;
          {
            while (mm) {
              int k[kPackedBatch];
              bool ok[kPackedBatch];
#pragma unroll
              for (int u = 0; u < kPackedBatch; ++u) {
                ok[u] = mm != 0ull;
                const int c = ok[u] ? (int)__builtin_ctzll(mm) : 0;
                k[u] = ok[u] ? c + (c < t1 ? b0 : (c < t2 ? b1 : b2)) : 0;
                mm &= mm - 1ull;
              }
              float2 xy[kPackedBatch];
              float4 rc[kPackedBatch];
              float pj[kPackedBatch];
#pragma unroll
              for (int u = 0; u < kPackedBatch; ++u) {
                xy[u] = S.pos[k[u]];
                rc[u] = S.rec[k[u]];
                pj[u] = S.p[k[u]];
              }
#pragma unroll
              for (int u = 0; u < kPackedBatch; ++u) {
                if (ok[u]) {
                  const Nbr nb = {xy[u].x, xy[u].y, rc[u].x, rc[u].y, rc[u].z, rc[u].w, pj[u]};
                  dsph_pair(P, K, o, nb, 0.f, 0.f, true, terms, a);
                }
              }
            }
          }
        }
        // a row with a long run, or every row of an unstaged chunk: all its
        // fluid pairs from memory
        if (__ballot(slow) != 0)
          walk_slow<1>(slow, g, sc, sp, thr, 0, 0, P, K, o, terms, frec, a);
      }
      if (bg.cell_start) {
        if (terms & kBdyPairTerms) {
          const SphGridDesc bd = bg.scenes[bg.nscenes == 1 ? 0 : sc];
          const bool same = bd.qmin_x == d.qmin_x && bd.qmin_y == d.qmin_y && bd.hmax == d.hmax;
          Stencil3 bst;
          bdy_stencil(bg, bd, same, xq, yq, ix, iy, bst);
          walk_bdy<1>(bst, bst.ncand, bg, xq, yq, thr, 0, P, K, o, terms, brec, a);
        }
      } else {
        packed_bdy_pairs<1>(P, K, o, brec, fb, qi, 0, terms, a);
      }
      packed_row_out<FUSE>(P, o, a, terms, qi, dpdt, acc, rk, rin);
    }
    // the next chunk restages the LDS arrays
    if (c0 + kWalkRows < s1) __syncthreads();
  }
}

// segs[b] = {first, end} sorted slots of workgroup b.  Five wavefronts per
// SIMD is what the LDS of the stage allows; the register budget is held to
// it (96 VGPRs), since 100 registers left four.
template <unsigned T, int KL, bool FUSE>
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(5))) void k_dsph_cellwalk(
    SphGrid g, const int2* __restrict__ segs, const float4* __restrict__ frec,
    const float4* __restrict__ brec, const SphPhysics* __restrict__ phys, SphCSR fb, SphGrid bg,
    unsigned rterms, float* __restrict__ dpdt, float* __restrict__ acc, SphRk4Stage rk) {
  const unsigned terms = T ? T : rterms;
  __shared__ typename std::conditional<KL == 1, Walk1Shared, WalkShared>::type S;
  const int2 seg = segs[xcd_block()];
  const int s0 = seg.x > 0 ? seg.x : 0, s1 = seg.y < g.np ? seg.y : g.np;
  // A segment inside one scene (scenes are contiguous in sorted order, so
  // its first and last rows tell) takes the uniform form.
  int usc = -1;
  if (g.nscenes == 1) {
    usc = 0;
  } else if (s0 < s1) {
    const int a = g.pscene[g.sorted_idx[s0]], z = g.pscene[g.sorted_idx[s1 - 1]];
    usc = a == z ? a : -1;
  }
  usc = __builtin_amdgcn_readfirstlane(usc);
  if constexpr (KL == 1) {
    if (usc >= 0)
      cellwalk_segment1<T, FUSE, true>(S, s0, s1, usc, g, frec, brec, phys, fb, bg, terms, dpdt,
                                       acc, rk);
    else
      cellwalk_segment1<T, FUSE, false>(S, s0, s1, -1, g, frec, brec, phys, fb, bg, terms, dpdt,
                                        acc, rk);
  } else {
    if (usc >= 0)
      cellwalk_segment<T, KL, FUSE, true>(S, s0, s1, usc, g, frec, brec, phys, fb, bg, terms,
                                          dpdt, acc, rk);
    else
      cellwalk_segment<T, KL, FUSE, false>(S, s0, s1, -1, g, frec, brec, phys, fb, bg, terms,
                                           dpdt, acc, rk);
  }
}

// ------------------------------------------ renormalised density gradient
__global__ __launch_bounds__(kBS) void k_pinv2x2(const float* __restrict__ M, int n,
                                                 float* __restrict__ L, float* __restrict__ eig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float l[4], e[2];
  pinv2x2_dev(M[4 * i], M[4 * i + 1], M[4 * i + 2], M[4 * i + 3], l, e);
  for (int k = 0; k < 4; ++k) L[4 * i + k] = l[k];
  if (eig) {
    eig[2 * i] = e[0];
    eig[2 * i + 1] = e[1];
  }
}

// M_i = sum_j V_j (x_j - x_i) (x) gradW_ij (densityDiffusion.py:30-43) with
// the boundary part (solidBoundary.py:684-687); L_i = pinv2x2(M_i).  A
// group of KL lanes walks each row.
struct XVol {
  float x, y, v;
};
template <int KL>
__global__ __launch_bounds__(kBS) void k_dsph_normalization(SphFluid f, const SphPhysics* __restrict__ phys,
                                                            SphCSR ff, SphBoundary b, SphCSR fb,
                                                            const int* __restrict__ order,
                                                            float* __restrict__ Mout,
                                                            float* __restrict__ Lout,
                                                            float* __restrict__ eig) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = order ? order[rl.t] : rl.t;
  const KConst K = kconst(phys[f.pscene[i]].h);
  const float2 xi = ld2(f.x, i);
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f;
  auto use = [&](int, const XVol& r) {
    const PairGeo g = pair_geo(xi.x, xi.y, r.x, r.y, K.inv_h);
    const float gw = -wendland_grad_k(g.q, K);
    const float gx = gw * g.hx, gy = gw * g.hy;
    const float rx = -g.rx, ry = -g.ry;
    m00 += rx * gx * r.v;
    m01 += rx * gy * r.v;
    m10 += ry * gx * r.v;
    m11 += ry * gy * r.v;
  };
  row_pairs<KL>(
      ff.col, ff.rowptr[i], ff.rowptr[i + 1], rl.sub,
      [&](int j) {
        const float2 xj = ld2(f.x, j);
        return XVol{xj.x, xj.y, f.vol ? f.vol[j] : f.area[j] / f.rho[j]};
      },
      use);
  if (fb.rowptr) {
    for (int k = fb.rowptr[i] + rl.sub; k < fb.rowptr[i + 1]; k += KL) {
      const int j = fb.col[k];
      const float2 xj = ld2(b.x, j);
      use(j, XVol{xj.x, xj.y, b.vol[j] / b.rho[j]});
    }
  }
  m00 = group_sum<KL>(m00);
  m01 = group_sum<KL>(m01);
  m10 = group_sum<KL>(m10);
  m11 = group_sum<KL>(m11);
  if (rl.sub != 0) return;
  if (Mout) {
    Mout[4 * i] = m00;
    Mout[4 * i + 1] = m01;
    Mout[4 * i + 2] = m10;
    Mout[4 * i + 3] = m11;
  }
  float l[4], e[2];
  pinv2x2_dev(m00, m01, m10, m11, l, e);
  for (int k = 0; k < 4; ++k) Lout[4 * i + k] = l[k];
  if (eig) {
    eig[2 * i] = e[0];
    eig[2 * i + 1] = e[1];
  }
}

// <grad rho>_i = -sum_j 2 (rho_j - rho_i) V_j L_i gradW_ij
// (densityDiffusion.py:45-61, solidBoundary.py:704-709)
__global__ __launch_bounds__(kBS) void k_dsph_density_gradient(SphFluid f,
                                                               const SphPhysics* __restrict__ phys,
                                                               SphCSR ff, SphBoundary b, SphCSR fb,
                                                               const float* __restrict__ L,
                                                               const int* __restrict__ order,
                                                               float* __restrict__ gradrho) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const SphPhysics P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const float2 xi = ld2(f.x, i);
  const float rho_i = f.rho[i] * P.rho0;
  const float l00 = L[4 * i], l01 = L[4 * i + 1], l10 = L[4 * i + 2], l11 = L[4 * i + 3];
  float sx = 0.f, sy = 0.f;
  for (int pass = 0; pass < 2; ++pass) {
    const SphCSR& c = pass == 0 ? ff : fb;
    if (!c.rowptr) continue;
    for (int k = c.rowptr[i]; k < c.rowptr[i + 1]; ++k) {
      const int j = c.col[k];
      const float2 xj = pass == 0 ? ld2(f.x, j) : ld2(b.x, j);
      const float rj = pass == 0 ? f.rho[j] : b.rho[j];
      const float V = pass == 0 ? f.area[j] / rj : b.vol[j] / rj;
      const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, K.inv_h);
      const float gw = -wendland_grad_k(g.q, K);
      const float gx = gw * g.hx, gy = gw * g.hy;
      const float lx = l00 * gx + l01 * gy, ly = l10 * gx + l11 * gy;
      const float c2 = (rj * P.rho0 - rho_i) * V * 2.f;
      sx += c2 * lx;
      sy += c2 * ly;
    }
  }
  st2(gradrho, i, make_float2(-sx, -sy));
}

// ----------------------------------------------------------------- mDBC
// 3x3 Moore-Penrose inverse via one-sided Jacobi SVD in fp32, singular
// values at or below rtol * sigma_max dropped with torch.linalg.pinv's
// default rtol = max(m, n) * eps(float32) (solidBoundary.py:603).  Column
// pairs are rotated until they are orthogonal to fp32 precision.
__device__ void pinv3x3f(const float A[9], float X[9]) {
  float U[9], V[9];
  for (int k = 0; k < 9; ++k) {
    U[k] = A[k];
    V[k] = (k % 4 == 0) ? 1.f : 0.f;
  }
  for (int sweep = 0; sweep < 20; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        float al = 0.f, be = 0.f, ga = 0.f;
        for (int r = 0; r < 3; ++r) {
          al += U[3 * r + p] * U[3 * r + p];
          be += U[3 * r + q] * U[3 * r + q];
          ga += U[3 * r + p] * U[3 * r + q];
        }
        if (!(fabsf(ga) > 1.2e-7f * sqrtf(al * be))) continue;
        rotated = true;
        const float zeta = (be - al) / (2.f * ga);
        const float tt = (zeta >= 0.f ? 1.f : -1.f) / (fabsf(zeta) + sqrtf(1.f + zeta * zeta));
        const float c = 1.f / sqrtf(1.f + tt * tt), sn = c * tt;
        for (int r = 0; r < 3; ++r) {
          const float up = U[3 * r + p], uq = U[3 * r + q];
          U[3 * r + p] = c * up - sn * uq;
          U[3 * r + q] = sn * up + c * uq;
          const float vp = V[3 * r + p], vq = V[3 * r + q];
          V[3 * r + p] = c * vp - sn * vq;
          V[3 * r + q] = sn * vp + c * vq;
        }
      }
    }
    if (!rotated) break;
  }
  float sig[3], smax = 0.f;
  for (int k = 0; k < 3; ++k) {
    sig[k] = sqrtf(U[k] * U[k] + U[3 + k] * U[3 + k] + U[6 + k] * U[6 + k]);
    smax = fmaxf(smax, sig[k]);
  }
  const float cut = 3.f * 1.1920928955078125e-07f * smax;
  for (int k = 0; k < 9; ++k) X[k] = 0.f;
  for (int k = 0; k < 3; ++k) {
    if (!(sig[k] > cut) || sig[k] == 0.f) continue;
    const float inv = 1.f / (sig[k] * sig[k]);  // u_k = U[:,k] / sig, times 1/sig
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) X[3 * r + c] += V[3 * r + k] * inv * U[3 * c + k];
  }
}

// mDBC sums of one ghost over its fluid neighbours (solidBoundary.py:561-
// 601), added pair by pair in the order of the ghost's row.
struct MdbcSums {
  float num = 0.f, den = 0.f, gsx = 0.f, gsy = 0.f, vgx = 0.f, vgy = 0.f, psx = 0.f, psy = 0.f;
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f;
  int cnt = 0;
};

__device__ __forceinline__ void mdbc_add(MdbcSums& s, const SphPhysics& P, const KConst& K,
                                         float2 g, float2 xf, float area, float rho_hat) {
  // distances of a foreign query: (ghost - fluid) / |.|, q = |.| / h_query
  const PairGeo geo = pair_geo(g.x, g.y, xf.x, xf.y, K.inv_h);
  const float W = wendland_k(geo.q, K);
  const float gw = -wendland_grad_k(geo.q, K);
  const float gWx = gw * geo.hx, gWy = gw * geo.hy;
  const float mass = area * P.rho0;
  const float V = area / rho_hat;
  s.num += mass * W;
  s.den += V * W;
  s.gsx += mass * gWx;
  s.gsy += mass * gWy;
  s.vgx += V * gWx;
  s.vgy += V * gWy;
  const float xijx = xf.x - g.x, xijy = xf.y - g.y;
  s.psx += V * W * xijx;
  s.psy += V * W * xijy;
  s.m00 += V * xijx * gWx;
  s.m01 += V * xijx * gWy;
  s.m10 += V * xijy * gWx;
  s.m11 += V * xijy * gWy;
  ++s.cnt;
}

// The boundary density from the sums (solidBoundary.py:603-616): MLS for
// more than 10 neighbours, Shepard for 1-10, 1 without; clamped below at 1.
__device__ __forceinline__ float mdbc_rho(const MdbcSums& s, const SphPhysics& P,
                                          const KConst& K, float2 xb, float2 g) {
  float rho = 1.f;
  if (s.cnt > 10) {
    // The MLS system A [rho, grad rho] = [num, sum m grad W] in units of the
    // support: unknowns z = [rho, h grad rho] and the gradient rows times h,
    //   [ den      psx / h  psy / h ] [ rho            ]   [ num   ]
    //   [ h vgx    m00      m01     ] [ h d rho / dx   ] = [ h gsx ]
    //   [ h vgy    m10      m11     ] [ h d rho / dy   ]   [ h gsy ]
    // whose entries are all O(1), so the fp32 solve loses no accuracy to
    // the 1/h^2 spread of the raw system.  For every system whose singular
    // values stay above the cutoff (a full set of more than 10 neighbours)
    // the result is pinv(A) rhs of solidBoundary.py:603-612.  The solve is
    // for the correction dz to the Shepard value z0 = [num / den, 0, 0]:
    // A dz = rhs - A z0, z = z0 + dz.  The same solution in exact
    // arithmetic; in fp32 the pseudo-inverse's rounding then scales with dz
    // (the density's variation over a support) instead of with rho itself,
    // which matters because the Tait EOS turns rho_b - 1 of a few ulp at
    // rest into the wall pressure.
    const float h = P.h, ih = K.inv_h;
    const float A[9] = {s.den, s.psx * ih, s.psy * ih, s.vgx * h, s.m00, s.m01,
                        s.vgy * h, s.m10, s.m11};
    float X[9];
    pinv3x3f(A, X);
    const float rs = s.num / s.den;
    const float rhs[3] = {s.num - s.den * rs, (s.gsx - s.vgx * rs) * h, (s.gsy - s.vgy * rs) * h};
    float dz[3];
    for (int r = 0; r < 3; ++r)
      dz[r] = X[3 * r] * rhs[0] + X[3 * r + 1] * rhs[1] + X[3 * r + 2] * rhs[2];
    const float dx = xb.x - g.x, dy = xb.y - g.y;
    rho = (rs + dz[0] + (dx * dz[1] + dy * dz[2]) * ih) / P.rho0;
  } else if (s.cnt > 0) {
    rho = s.num / s.den / P.rho0;
  }
  return rho < 1.f ? 1.f : rho;  // solidBoundary.py:615
}

__global__ __launch_bounds__(kBS) void k_mdbc_density(SphFluid f, const SphPhysics* __restrict__ phys,
                                                      const float* __restrict__ bx,
                                                      const float* __restrict__ gxp,
                                                      const int* __restrict__ bscene, int nb,
                                                      SphCSR gf, float* __restrict__ rho_b,
                                                      float* __restrict__ p_b,
                                                      int* __restrict__ counts) {
  const int bI = blockIdx.x * blockDim.x + threadIdx.x;
  if (bI >= nb) return;
  const SphPhysics P = phys[bscene[bI]];
  const KConst K = kconst(P.h);
  const float2 g = ld2(gxp, bI);
  MdbcSums s;
  const int a0 = gf.rowptr[bI], a1 = gf.rowptr[bI + 1];
  for (int k = a0; k < a1; ++k) {
    const int j = gf.col[k];
    mdbc_add(s, P, K, g, ld2(f.x, j), f.area[j], f.rho[j]);
  }
  const float rho = mdbc_rho(s, P, K, ld2(bx, bI), g);
  rho_b[bI] = rho;
  if (p_b) p_b[bI] = eos(P, rho);
  if (counts) counts[bI] = s.cnt;
}

// mDBC with the ghosts' fluid neighbours found in the kernel (no ghost ->
// fluid list): one lane per ghost walks its stencil in the fluid grid g in
// the fill's candidate order, with the fill's predicate (|ghost - x_j| <
// gh), and adds each hit as k_mdbc_density adds its row, so rho_b, p_b and
// the counts are those of sph_mdbc_density over the list of sph_neighbor_fill
// bit for bit.  Four candidate loads in flight per lane.  brec: the ghost's
// boundary record as k_dsph_pack writes it from rho_b and p_b (velocity bv,
// zero when NULL; volume bvol), so the stage needs no separate pack.
__global__ __launch_bounds__(kBS) void k_mdbc_walk(SphFluid f, const SphPhysics* __restrict__ phys,
                                                   const float* __restrict__ bx,
                                                   const float* __restrict__ gxp,
                                                   const float* __restrict__ gh,
                                                   const int* __restrict__ bscene, int nb,
                                                   SphGrid grid, float* __restrict__ rho_b,
                                                   float* __restrict__ p_b,
                                                   int* __restrict__ counts,
                                                   const float* __restrict__ bv,
                                                   const float* __restrict__ bvol,
                                                   float4* __restrict__ brec) {
  const int bI = xcd_block() * blockDim.x + threadIdx.x;
  if (bI >= nb) return;
  const int sc = bscene[bI];
  const SphPhysics P = phys[sc];
  const KConst K = kconst(P.h);
  const float2 g = ld2(gxp, bI);
  const SphGridDesc d = grid.scenes[grid.nscenes == 1 ? 0 : sc];
  Stencil3 st;
  st.init(d, grid.cell_start + d.cell_begin, cell_coord(g.x, d.qmin_x, d.hmax),
          cell_coord(g.y, d.qmin_y, d.hmax));
  const float thr = sq_threshold(gh[bI]);
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(grid.sorted_x);
  MdbcSums s;
  for (int p0 = 0; p0 < st.ncand; p0 += kBatch) {
    float2 xf[kBatch];
    int slot[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      slot[u] = st.slot(p0 + u);   // clamped into the candidates
      xf[u] = sx[slot[u]];
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (p0 + u < st.ncand && dist2_nofma(g.x, g.y, xf[u].x, xf[u].y) <= thr) {
        const int j = grid.sorted_idx[slot[u]];
        mdbc_add(s, P, K, g, xf[u], f.area[j], f.rho[j]);
      }
    }
  }
  const float2 xb = ld2(bx, bI);
  const float rho = mdbc_rho(s, P, K, xb, g);
  const float p = eos(P, rho);
  rho_b[bI] = rho;
  if (p_b) p_b[bI] = p;
  if (counts) counts[bI] = s.cnt;
  if (brec) {
    const float2 v = bv ? ld2(bv, bI) : make_float2(0.f, 0.f);
    brec[2 * bI] = make_float4(xb.x, xb.y, v.x, v.y);
    brec[2 * bI + 1] = make_float4(rho * P.rho0, bvol[bI] / rho, p, rho);
  }
}

__global__ __launch_bounds__(kBS) void k_eos(const float* __restrict__ rho,
                                             const int* __restrict__ pscene,
                                             const SphPhysics* __restrict__ phys, int n,
                                             float* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = eos(phys[pscene[i]], rho[i]);
}

__global__ __launch_bounds__(kBS) void k_akinci_bdensity(SphFluid f, const SphPhysics* __restrict__ phys,
                                                         const float* __restrict__ bx,
                                                         const int* __restrict__ bscene, int nb,
                                                         const float* __restrict__ term,
                                                         SphCSR bf, float* __restrict__ rho_b) {
  const int bI = blockIdx.x * blockDim.x + threadIdx.x;
  if (bI >= nb) return;
  const SphPhysics P = phys[bscene[bI]];
  const KConst K = kconst(P.h);
  const float2 xb = ld2(bx, bI);
  float s = 0.f;
  for (int k = bf.rowptr[bI]; k < bf.rowptr[bI + 1]; ++k) {
    const int j = bf.col[k];
    const float2 xf = ld2(f.x, j);
    const PairGeo g = pair_geo(xb.x, xb.y, xf.x, xf.y, K.inv_h);
    s += wendland_k(g.q, K) * f.area[j];
  }
  rho_b[bI] = term[bI] + s + P.beta;
}

}  // namespace

static bool fluid_ok(const SphFluid* f) {
  return f && f->x && f->rho && f->area && f->pscene && f->np >= 0;
}
static SphCSR csr_or_null(const SphCSR* c) {
  SphCSR r = {nullptr, nullptr, 0};
  if (c) r = *c;
  return r;
}
static SphBoundary bdy_or_null(const SphBoundary* b) {
  SphBoundary r = {};
  if (b) r = *b;
  return r;
}

int fluid_density(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                  const SphCSR* fb, const int* order, float* rho_out, float* vol_out,
                  hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && ff && ff->rowptr && rho_out, "sph_fluid_density: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || (b && b->x && b->vol), "sph_fluid_density: boundary list without boundary");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int bs = tuning(SPH_TUNE_BLOCK);
    hipLaunchKernelGGL((k_fluid_density<KL>), dim3(blocks_for((int64_t)f->np * KL, bs)),
                       dim3(bs), 0, s, *f, phys, *ff, bdy_or_null(b), csr_or_null(fb), order,
                       rho_out, vol_out);
  });
  return launch_check("sph_fluid_density");
}

static int check_bdy_terms(unsigned terms, const SphBoundary* b, const SphCSR* fb, const char* who) {
  const bool wants_b = terms & (SPH_T_CONT_B | SPH_T_PRESS_B | SPH_T_VISC_B);
  if (wants_b && fb && fb->rowptr && !(b && b->x && b->rho && b->vol)) {
    set_error(std::string(who) + ": boundary terms without boundary state");
    return SPH_E_ARG;
  }
  return SPH_OK;
}

int dsph_gather(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                const SphCSR* fb, const float* gradrho, unsigned terms, const int* order,
                float* dpdt, float* acc, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && f->v && phys && ff && ff->rowptr, "sph_dsph_gather: bad arguments");
  SPH_REQUIRE(!(terms & SPH_T_DDSPH) || gradrho, "sph_dsph_gather: DDSPH needs gradrho");
  int rc = check_bdy_terms(terms, b, fb, "sph_dsph_gather");
  if (rc) return rc;
  if (f->np == 0) return SPH_OK;
  if (!(terms & (kFluidPairTerms | kBdyPairTerms))) {
    hipLaunchKernelGGL(k_dsph_rates_only, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f,
                       phys, terms, dpdt, acc);
    return launch_check("sph_dsph_gather");
  }
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int bs = tuning(SPH_TUNE_BLOCK);
    hipLaunchKernelGGL((k_dsph_gather<KL>), dim3(blocks_for((int64_t)f->np * KL, bs)),
                       dim3(bs), 0, s, *f, phys, *ff, bdy_or_null(b), csr_or_null(fb), gradrho,
                       terms, order, dpdt, acc);
  });
  return launch_check("sph_dsph_gather");
}

int dsph_pack(const SphFluid* f, const SphPhysics* phys, const SphBoundary* b, float* frec,
              float* brec, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && f->v && phys, "sph_dsph_pack: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bb.nb == 0 || (bb.x && bb.rho && bb.vol && bb.bscene && brec),
              "sph_dsph_pack: bad boundary");
  const int t0 = frec ? 0 : f->np;   // no fluid records: boundary rows only
  const int n = f->np + bb.nb - t0;
  if (n <= 0) return SPH_OK;
  hipLaunchKernelGGL(k_dsph_pack, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, phys, bb,
                     reinterpret_cast<float4*>(frec), reinterpret_cast<float4*>(brec), t0);
  return launch_check("sph_dsph_pack");
}

// term masks with a compiled specialisation; others use the runtime form
constexpr unsigned kMaskMOG = SPH_T_CONT | SPH_T_CONT_B | SPH_T_MOG | SPH_T_PRESS | SPH_T_PRESS_B |
                              SPH_T_VISC | SPH_T_VISC_B | SPH_T_LAM | SPH_T_GRAV;
constexpr unsigned kMaskVL = SPH_T_VISC | SPH_T_LAM | SPH_T_ACCUM;

template <bool FUSE>
static void launch_gather(const float4* fr, const float4* br, const int* pscene, int np,
                          const SphPhysics* phys, const SphCSR& ff, const SphCSR& fb,
                          const float* gradrho, unsigned terms, const int* order, float* dpdt,
                          float* acc, const SphRk4Stage& rk, hipStream_t s,
                          const float* smask = nullptr) {
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int bs = tuning(SPH_TUNE_BLOCK);
    const dim3 grid(blocks_for((int64_t)np * KL, bs)), block(bs);
    if (terms == kMaskMOG)
      hipLaunchKernelGGL((k_dsph_gather_packed<kMaskMOG, KL, FUSE>), grid, block, 0, s, fr, br,
                         pscene, np, phys, ff, fb, gradrho, terms, order, dpdt, acc, rk, smask);
    else if (!FUSE && terms == kMaskVL)
      hipLaunchKernelGGL((k_dsph_gather_packed<kMaskVL, KL, false>), grid, block, 0, s, fr, br,
                         pscene, np, phys, ff, fb, gradrho, terms, order, dpdt, acc, rk, smask);
    else
      hipLaunchKernelGGL((k_dsph_gather_packed<0, KL, FUSE>), grid, block, 0, s, fr, br, pscene,
                         np, phys, ff, fb, gradrho, terms, order, dpdt, acc, rk, smask);
  });
}

int dsph_gather_packed(const float* frec, const float* brec, const int* pscene, int np,
                       const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                       const float* gradrho, unsigned terms, const int* order, float* dpdt,
                       float* acc, hipStream_t s) {
  SPH_REQUIRE(frec && pscene && phys && ff && ff->rowptr && np >= 0,
              "sph_dsph_gather_packed: bad arguments");
  SPH_REQUIRE(!(terms & SPH_T_DDSPH) || gradrho, "sph_dsph_gather_packed: DDSPH needs gradrho");
  SPH_REQUIRE(!(terms & kBdyPairTerms) || !fb || !fb->rowptr || brec,
              "sph_dsph_gather_packed: boundary terms without boundary records");
  if (np == 0) return SPH_OK;
  const SphRk4Stage none = {};
  launch_gather<false>(reinterpret_cast<const float4*>(frec), reinterpret_cast<const float4*>(brec),
                       pscene, np, phys, *ff, csr_or_null(fb), gradrho, terms, order, dpdt, acc,
                       none, s);
  return launch_check("sph_dsph_gather_packed");
}

int dsph_gather_packed_dplus(const float* frec, const float* brec, const int* pscene, int np,
                             const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                             const float* gradrho, unsigned terms, const int* order,
                             const float* surface_mask, float* dpdt, float* acc, hipStream_t s) {
  SPH_REQUIRE(frec && pscene && phys && ff && ff->rowptr && np >= 0 && surface_mask,
              "sph_dsph_gather_packed_dplus: bad arguments");
  SPH_REQUIRE(!(terms & SPH_T_DDSPH) || gradrho,
              "sph_dsph_gather_packed_dplus: DDSPH needs gradrho");
  SPH_REQUIRE(!(terms & kBdyPairTerms) || !fb || !fb->rowptr || brec,
              "sph_dsph_gather_packed_dplus: boundary terms without boundary records");
  if (np == 0) return SPH_OK;
  const SphRk4Stage none = {};
  launch_gather<false>(reinterpret_cast<const float4*>(frec), reinterpret_cast<const float4*>(brec),
                       pscene, np, phys, *ff, csr_or_null(fb), gradrho, terms, order, dpdt, acc,
                       none, s, surface_mask);
  return launch_check("sph_dsph_gather_packed_dplus");
}

static int check_rk4(const SphRk4Stage* rk, const float* frec, unsigned terms,
                     const std::string& who) {
  SPH_REQUIRE(rk, who + ": bad arguments");
  SPH_REQUIRE(rk->stage >= 1 && rk->stage <= 4, who + ": stage must be 1..4");
  SPH_REQUIRE(rk->x0 && rk->v0 && rk->r0 && rk->K && rk->x_out && rk->rho_out &&
                  (rk->v_out || (rk->stage < 4 && rk->frec_next)),
              who + ": missing state arrays");
  SPH_REQUIRE(!rk->frec_next || (rk->area && rk->frec_next != frec),
              who + ": next records need the areas and a buffer of their own");
  SPH_REQUIRE(!(terms & SPH_T_ACCUM), who + ": SPH_T_ACCUM has no meaning here");
  return SPH_OK;
}

int dsph_gather_rk4(const float* frec, const float* brec, const int* pscene, int np,
                    const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                    const float* gradrho, unsigned terms, const int* order, const SphRk4Stage* rk,
                    hipStream_t s) {
  SPH_REQUIRE(frec && pscene && phys && ff && ff->rowptr && np >= 0 && rk,
              "sph_dsph_gather_rk4: bad arguments");
  int rc = check_rk4(rk, frec, terms, "sph_dsph_gather_rk4");
  if (rc) return rc;
  SPH_REQUIRE(!(terms & SPH_T_DDSPH) || gradrho, "sph_dsph_gather_rk4: DDSPH needs gradrho");
  SPH_REQUIRE(!(terms & kBdyPairTerms) || !fb || !fb->rowptr || brec,
              "sph_dsph_gather_rk4: boundary terms without boundary records");
  if (np == 0) return SPH_OK;
  launch_gather<true>(reinterpret_cast<const float4*>(frec), reinterpret_cast<const float4*>(brec),
                      pscene, np, phys, *ff, csr_or_null(fb), gradrho, terms, order, nullptr,
                      nullptr, *rk, s);
  return launch_check("sph_dsph_gather_rk4");
}

template <bool FUSE>
static void launch_cellwalk(const SphGrid& g, const int* segs, int nseg, const float* frec,
                            const float* brec, const SphPhysics* phys, const SphCSR& fb,
                            const SphGrid& bg, unsigned terms, float* dpdt, float* acc,
                            const SphRk4Stage& rk, hipStream_t s) {
  const int2* sg = reinterpret_cast<const int2*>(segs);
  const float4* fr = reinterpret_cast<const float4*>(frec);
  const float4* br = reinterpret_cast<const float4*>(brec);
  with_lanes(tuning(SPH_TUNE_WALK_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    if (terms == kMaskMOG)
      hipLaunchKernelGGL((k_dsph_cellwalk<kMaskMOG, KL, FUSE>), dim3(nseg), dim3(kBS), 0, s, g,
                         sg, fr, br, phys, fb, bg, terms, dpdt, acc, rk);
    else
      hipLaunchKernelGGL((k_dsph_cellwalk<0, KL, FUSE>), dim3(nseg), dim3(kBS), 0, s, g, sg, fr,
                         br, phys, fb, bg, terms, dpdt, acc, rk);
  });
}

static SphGrid grid_or_null(const SphGrid* g) {
  SphGrid r = {};
  if (g) r = *g;
  return r;
}

static int check_cellwalk(const SphGrid* g, const int* segs, int nseg, const float* frec,
                          const float* brec, const SphPhysics* phys, const SphCSR* fb,
                          const SphGrid* bg, unsigned terms, const std::string& who) {
  SPH_REQUIRE(g && g->scenes && g->sorted_idx && g->sorted_x && g->cell_start && g->np >= 0 &&
                  (g->nscenes == 1 || g->pscene),
              who + ": the grid is not built");
  SPH_REQUIRE(nseg >= 0 && (nseg == 0 || segs) && frec && phys, who + ": bad arguments");
  SPH_REQUIRE(!(terms & SPH_T_DDSPH),
              who + ": the DDSPH term needs its pair list (sph_dsph_gather_packed)");
  SPH_REQUIRE(!(terms & kBdyPairTerms) || ((!fb || !fb->rowptr) && !bg) || brec,
              who + ": boundary terms without boundary records");
  SPH_REQUIRE(!bg || (bg->scenes && bg->cell_start && bg->sorted_idx && bg->sorted_x &&
                      bg->nscenes == g->nscenes),
              who + ": the boundary grid is not built over the fluid grid's scenes");
  SPH_REQUIRE(!(bg && fb && fb->rowptr), who + ": boundary pairs from a grid or a list, not both");
  return SPH_OK;
}

int dsph_cellwalk(const SphGrid* g, const int* segs, int nseg, const float* frec,
                  const float* brec, const SphPhysics* phys, const SphCSR* fb,
                  const SphGrid* bg, unsigned terms, float* dpdt, float* acc, hipStream_t s) {
  int rc = check_cellwalk(g, segs, nseg, frec, brec, phys, fb, bg, terms, "sph_dsph_cellwalk");
  if (rc) return rc;
  if (nseg == 0 || g->np == 0) return SPH_OK;
  const SphRk4Stage none = {};
  launch_cellwalk<false>(*g, segs, nseg, frec, brec, phys, csr_or_null(fb), grid_or_null(bg),
                         terms, dpdt, acc, none, s);
  return launch_check("sph_dsph_cellwalk");
}

int dsph_cellwalk_rk4(const SphGrid* g, const int* segs, int nseg, const float* frec,
                      const float* brec, const SphPhysics* phys, const SphCSR* fb,
                      const SphGrid* bg, unsigned terms, const SphRk4Stage* rk, hipStream_t s) {
  int rc = check_cellwalk(g, segs, nseg, frec, brec, phys, fb, bg, terms,
                          "sph_dsph_cellwalk_rk4");
  if (rc) return rc;
  if ((rc = check_rk4(rk, frec, terms, "sph_dsph_cellwalk_rk4"))) return rc;
  if (nseg == 0 || g->np == 0) return SPH_OK;
  launch_cellwalk<true>(*g, segs, nseg, frec, brec, phys, csr_or_null(fb), grid_or_null(bg),
                        terms, nullptr, nullptr, *rk, s);
  return launch_check("sph_dsph_cellwalk_rk4");
}

int dsph_normalization(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const int* order, float* M,
                       float* L, float* eig, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && ff && ff->rowptr && L, "sph_dsph_normalization: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || (b && b->x && b->rho && b->vol),
              "sph_dsph_normalization: boundary list without boundary state");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_dsph_normalization<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)),
                       dim3(kBS), 0, s, *f, phys, *ff, bdy_or_null(b), csr_or_null(fb), order, M,
                       L, eig);
  });
  return launch_check("sph_dsph_normalization");
}

int dsph_density_gradient(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                          const SphBoundary* b, const SphCSR* fb, const float* L,
                          const int* order, float* gradrho, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && ff && ff->rowptr && L && gradrho,
              "sph_dsph_density_gradient: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || (b && b->x && b->rho && b->vol),
              "sph_dsph_density_gradient: boundary list without boundary state");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_dsph_density_gradient, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f,
                     phys, *ff, bdy_or_null(b), csr_or_null(fb), L, order, gradrho);
  return launch_check("sph_dsph_density_gradient");
}

int pinv2x2(const float* M, int n, float* L, float* eig, hipStream_t s) {
  SPH_REQUIRE(M && L && n >= 0, "sph_pinv2x2: bad arguments");
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_pinv2x2, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, M, n, L, eig);
  return launch_check("sph_pinv2x2");
}

int mdbc_density(const SphFluid* f, const SphPhysics* phys, const float* bx, const float* gx,
                 const int* bscene, int nb, const SphCSR* gf, float* rho_b, float* p_b,
                 int* counts, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && bx && gx && bscene && nb >= 0 && gf && gf->rowptr && rho_b,
              "sph_mdbc_density: bad arguments");
  if (nb == 0) return SPH_OK;
  hipLaunchKernelGGL(k_mdbc_density, dim3(blocks_for(nb, kBS)), dim3(kBS), 0, s, *f, phys, bx, gx,
                     bscene, nb, *gf, rho_b, p_b, counts);
  return launch_check("sph_mdbc_density");
}

int mdbc_density_walk(const SphFluid* f, const SphPhysics* phys, const float* bx,
                      const float* gx, const float* gh, const int* bscene, int nb,
                      const SphGrid* g, float* rho_b, float* p_b, int* counts, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && bx && gx && gh && bscene && nb >= 0 && rho_b,
              "sph_mdbc_density_walk: bad arguments");
  SPH_REQUIRE(g && g->scenes && g->cell_start && g->sorted_idx && g->sorted_x && g->np == f->np,
              "sph_mdbc_density_walk: the fluid grid is not built");
  if (nb == 0) return SPH_OK;
  hipLaunchKernelGGL(k_mdbc_walk, dim3(blocks_for(nb, kBS)), dim3(kBS), 0, s, *f, phys, bx, gx,
                     gh, bscene, nb, *g, rho_b, p_b, counts, (const float*)nullptr,
                     (const float*)nullptr, (float4*)nullptr);
  return launch_check("sph_mdbc_density_walk");
}

int mdbc_density_walk_records(const SphFluid* f, const SphPhysics* phys, const float* bx,
                              const float* bv, const float* bvol, const float* gx,
                              const float* gh, const int* bscene, int nb, const SphGrid* g,
                              float* rho_b, float* p_b, int* counts, float* brec, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && bx && bvol && gx && gh && bscene && nb >= 0 && rho_b &&
                  brec,
              "sph_mdbc_density_walk_records: bad arguments");
  SPH_REQUIRE(g && g->scenes && g->cell_start && g->sorted_idx && g->sorted_x && g->np == f->np,
              "sph_mdbc_density_walk_records: the fluid grid is not built");
  if (nb == 0) return SPH_OK;
  hipLaunchKernelGGL(k_mdbc_walk, dim3(blocks_for(nb, kBS)), dim3(kBS), 0, s, *f, phys, bx, gx,
                     gh, bscene, nb, *g, rho_b, p_b, counts, bv, bvol,
                     reinterpret_cast<float4*>(brec));
  return launch_check("sph_mdbc_density_walk_records");
}

int eos_all(const float* rho, const int* pscene, const SphPhysics* phys, int n, float* p,
            hipStream_t s) {
  SPH_REQUIRE(rho && pscene && phys && p && n >= 0, "sph_eos: bad arguments");
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_eos, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, rho, pscene, phys, n, p);
  return launch_check("sph_eos");
}

int akinci_boundary_density(const SphFluid* f, const SphPhysics* phys, const float* bx,
                            const int* bscene, int nb, const float* term, const SphCSR* bf,
                            float* rho_b, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && bx && bscene && term && bf && bf->rowptr && rho_b && nb >= 0,
              "sph_akinci_boundary_density: bad arguments");
  if (nb == 0) return SPH_OK;
  hipLaunchKernelGGL(k_akinci_bdensity, dim3(blocks_for(nb, kBS)), dim3(kBS), 0, s, *f, phys, bx,
                     bscene, nb, term, *bf, rho_b);
  return launch_check("sph_akinci_boundary_density");
}

}  // namespace sph

extern "C" {
int sph_fluid_density(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphBoundary* b, const SphCSR* fb, const int* order, float* rho_out,
                      float* vol_out, hipStream_t s) {
  return sph::fluid_density(f, phys, ff, b, fb, order, rho_out, vol_out, s);
}
int sph_dsph_gather(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                    const SphBoundary* b, const SphCSR* fb, const float* gradrho, unsigned terms,
                    const int* order, float* dpdt, float* acc, hipStream_t s) {
  return sph::dsph_gather(f, phys, ff, b, fb, gradrho, terms, order, dpdt, acc, s);
}
int sph_dsph_pack(const SphFluid* f, const SphPhysics* phys, const SphBoundary* b, float* frec,
                  float* brec, hipStream_t s) {
  return sph::dsph_pack(f, phys, b, frec, brec, s);
}
int sph_dsph_gather_packed(const float* frec, const float* brec, const int* pscene, int np,
                           const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                           const float* gradrho, unsigned terms, const int* order, float* dpdt,
                           float* acc, hipStream_t s) {
  return sph::dsph_gather_packed(frec, brec, pscene, np, phys, ff, fb, gradrho, terms, order, dpdt,
                                 acc, s);
}
int sph_dsph_gather_packed_dplus(const float* frec, const float* brec, const int* pscene,
                                 int np, const SphPhysics* phys, const SphCSR* ff,
                                 const SphCSR* fb, const float* gradrho, unsigned terms,
                                 const int* order, const float* surface_mask, float* dpdt,
                                 float* acc, hipStream_t s) {
  return sph::dsph_gather_packed_dplus(frec, brec, pscene, np, phys, ff, fb, gradrho, terms,
                                       order, surface_mask, dpdt, acc, s);
}
int sph_dsph_gather_rk4(const float* frec, const float* brec, const int* pscene, int np,
                        const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                        const float* gradrho, unsigned terms, const int* order,
                        const SphRk4Stage* rk, hipStream_t s) {
  return sph::dsph_gather_rk4(frec, brec, pscene, np, phys, ff, fb, gradrho, terms, order, rk, s);
}
int sph_dsph_cellwalk(const SphGrid* g, const int* segs, int nseg, const float* frec,
                      const float* brec, const SphPhysics* phys, const SphCSR* fb,
                      const SphGrid* bg, unsigned terms, float* dpdt, float* acc, hipStream_t s) {
  return sph::dsph_cellwalk(g, segs, nseg, frec, brec, phys, fb, bg, terms, dpdt, acc, s);
}
int sph_dsph_cellwalk_rk4(const SphGrid* g, const int* segs, int nseg, const float* frec,
                          const float* brec, const SphPhysics* phys, const SphCSR* fb,
                          const SphGrid* bg, unsigned terms, const SphRk4Stage* rk,
                          hipStream_t s) {
  return sph::dsph_cellwalk_rk4(g, segs, nseg, frec, brec, phys, fb, bg, terms, rk, s);
}
int sph_dsph_normalization(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                           const SphBoundary* b, const SphCSR* fb, const int* order, float* M,
                           float* L, float* eig, hipStream_t s) {
  return sph::dsph_normalization(f, phys, ff, b, fb, order, M, L, eig, s);
}
int sph_dsph_density_gradient(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                              const SphBoundary* b, const SphCSR* fb, const float* L,
                              const int* order, float* gradrho, hipStream_t s) {
  return sph::dsph_density_gradient(f, phys, ff, b, fb, L, order, gradrho, s);
}
int sph_pinv2x2(const float* M, int n, float* L, float* eig, hipStream_t s) {
  return sph::pinv2x2(M, n, L, eig, s);
}
int sph_mdbc_density(const SphFluid* f, const SphPhysics* phys, const float* bx, const float* gx,
                     const int* bscene, int nb, const SphCSR* gf, float* rho_b, float* p_b,
                     int* counts, hipStream_t s) {
  return sph::mdbc_density(f, phys, bx, gx, bscene, nb, gf, rho_b, p_b, counts, s);
}
int sph_mdbc_density_walk(const SphFluid* f, const SphPhysics* phys, const float* bx,
                          const float* gx, const float* gh, const int* bscene, int nb,
                          const SphGrid* g, float* rho_b, float* p_b, int* counts,
                          hipStream_t s) {
  return sph::mdbc_density_walk(f, phys, bx, gx, gh, bscene, nb, g, rho_b, p_b, counts, s);
}
int sph_mdbc_density_walk_records(const SphFluid* f, const SphPhysics* phys, const float* bx,
                                  const float* bv, const float* bvol, const float* gx,
                                  const float* gh, const int* bscene, int nb, const SphGrid* g,
                                  float* rho_b, float* p_b, int* counts, float* brec,
                                  hipStream_t s) {
  return sph::mdbc_density_walk_records(f, phys, bx, bv, bvol, gx, gh, bscene, nb, g, rho_b, p_b,
                                        counts, brec, s);
}
int sph_eos(const float* rho, const int* pscene, const SphPhysics* phys, int n, float* p,
            hipStream_t s) {
  return sph::eos_all(rho, pscene, phys, n, p, s);
}
int sph_akinci_boundary_density(const SphFluid* f, const SphPhysics* phys, const float* bx,
                                const int* bscene, int nb, const float* term, const SphCSR* bf,
                                float* rho_b, hipStream_t s) {
  return sph::akinci_boundary_density(f, phys, bx, bscene, nb, term, bf, rho_b, s);
}
}
