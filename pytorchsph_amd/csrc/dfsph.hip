// DFSPH relaxed-Jacobi density solve (modules/dfsph.py:25-311) with the
// Akinci PBSPH boundary hooks (akinciBoundary.py:166-389), batched over
// scenes.
//
// The reference syncs to the host once per iteration to read an fp32 mean
// residual.  Here the residual is summed in FP64 on the device, and a
// one-workgroup control kernel decides per scene whether to continue and
// rebuilds a compact list of the scenes still iterating.  The two gathers of
// an iteration stride over that list only, so scenes that have converged cost
// nothing while slower ones finish; the host issues iterations in chunks and
// polls `any_active` between chunks.
//
// Each iteration reads neighbours through 16-byte records written by the
// previous kernel: xrec = {x, y, p2 / (rho rho0)^2, A rho0} for the
// pressure acceleration, arec = {a*, dt^2 A^act} for the pressure update.
//
// Spiky gradient in the fluid-row convention: G_ij = -spiky(q) rhat_ij with
// rhat_ij = (x_i - x_j) / |x_i - x_j|.  Boundary neighbours enter fluid rows
// like fluid neighbours with their own area (V_b), actual area
// (V_b / rho_b), predicted velocity (0) and pressure acceleration (0).
#include "dsph_pair.h"
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;
constexpr int kPairBatch = 4;   // neighbour loads a lane keeps in flight in the row loops

__device__ __forceinline__ float2 ld2(const float* p, int i) {
  return reinterpret_cast<const float2*>(p)[i];
}
__device__ __forceinline__ void st2(float* p, int i, float2 v) {
  reinterpret_cast<float2*>(p)[i] = v;
}
__device__ __forceinline__ float4 ld4(const float* p, int i) {
  return reinterpret_cast<const float4*>(p)[i];
}
__device__ __forceinline__ void st4(float* p, int i, float4 v) {
  reinterpret_cast<float4*>(p)[i] = v;
}

// Calls f(load(col(k))) for this lane's pairs k = k0 + sub, k0 + sub + KL,
// ... < k1, in that order (the order of the one-pair loop), with kPairBatch
// neighbour loads in flight: loads are issued unconditionally (a slot past
// the row loads pair k0 again) and one batch ahead of the pairs being
// evaluated.  An empty row loads nothing.  k0, k1 are the same in every lane
// of a group.
template <int KL, typename Col, typename Load, typename F>
__device__ __forceinline__ void pair_pipeline_with(int k0, int k1, int sub, Col col, Load load,
                                                   F f) {
  if (k0 >= k1) return;
  using R = decltype(load(0));
  R rj[kPairBatch], rn[kPairBatch];
  int kb = k0 + sub;
#pragma unroll
  for (int u = 0; u < kPairBatch; ++u) {
    const int k = kb + u * KL;
    rj[u] = load(col(k < k1 ? k : k0));
  }
  for (; kb < k1; kb += KL * kPairBatch) {
    const int kn = kb + KL * kPairBatch;
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u) {
      const int k = kn + u * KL;
      rn[u] = load(col(k < k1 ? k : k0));
    }
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u)
      if (kb + u * KL < k1) f(rj[u]);
#pragma unroll
    for (int u = 0; u < kPairBatch; ++u) rj[u] = rn[u];
  }
}

// The same with one 16-byte record per neighbour, rec[col(k)].
template <int KL, typename Col, typename F>
__device__ __forceinline__ void pair_pipeline(int k0, int k1, int sub, const float* rec, Col col,
                                              F f) {
  pair_pipeline_with<KL>(k0, k1, sub, col, [&](int j) { return ld4(rec, j); }, f);
}

// Two records of one neighbour.
struct Rec2 {
  float4 a, b;
};

__device__ __forceinline__ float2 spiky_G(float xi, float yi, float xj, float yj, const KConst& K) {
  const PairGeo g = pair_geo(xi, yi, xj, yj, K.inv_h);
  const float s = -spiky_grad_k(g.q, K);
  return make_float2(s * g.hx, s * g.hy);
}

// Maps active work item t to (scene, index): pre[0..nact] are the prefix
// sums of the active scenes' counts; returns the rank k with
// pre[k] <= t < pre[k + 1].
__device__ __forceinline__ int find_rank(const int* __restrict__ pre, int nact, int t) {
  int lo = 0, hi = nact;  // invariant: pre[lo] <= t < pre[hi]
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (pre[m] <= t) lo = m; else hi = m;
  }
  return lo;
}
// The same rank found by walking forward from the rank r of an item
// t0 <= t.  A workgroup's items are consecutive, so one binary search with
// a workgroup-uniform argument (scalar loads) plus this walk, usually zero
// or one step, replaces a binary search per lane.
__device__ __forceinline__ int advance_rank(const int* __restrict__ pre, int nact, int r, int t) {
  while (r + 1 < nact && pre[r + 1] <= t) ++r;
  return r;
}

// Row epilogues shared by the CSR and stencil-list forms of the gathers
// below, so that both compute the same bits from the same sums.

// alpha_i, the source s_i and W_i of fluid row i from its sums (dfsph.py:58-100).
// (The row inputs of these epilogues are loaded apart, *_in, so that a
// kernel can issue those loads before its pair sums.)
struct AlphaIn {
  float aact, mass, rho;
};
__device__ __forceinline__ AlphaIn alpha_in(const SphFluid& f, const SphDfsph& d, int i) {
  return AlphaIn{d.aact[i], f.area[i] * f.rest[i], f.rho[i]};
}
__device__ __forceinline__ void alpha_store(const SphPhysics& P, const SphDfsph& d, int i,
                                            const AlphaIn& in, float k1x, float k1y, float k2,
                                            float src) {
  const float fac = -P.dt * P.dt * in.aact;
  float al = fac / in.mass * (k1x * k1x + k1y * k1y) + fac * k2;
  al = al < -1.f ? -1.f : (al > -1e-7f ? -1e-7f : al);  // dfsph.py:82
  d.alpha[i] = al;
  // divergenceSolve: the source is the divergence term only (dfsph.py:161)
  d.src[i] = d.divergence ? src : 1.f - in.rho + src;
  // W_i = dt^2 k1_i: the part of the pressure-update sum that does not
  // change during the solve
  st2(d.wvec, i, make_float2(P.dt * P.dt * k1x, P.dt * P.dt * k1y));
}

// alpha and source of boundary row k0 over its fluid neighbours
// (akinciBoundary.py:317-352); lane `sub` of a group of KL.
template <int KL>
__device__ __forceinline__ void alpha_boundary_row(const SphPhysics* __restrict__ phys,
                                                   const SphBoundary& b, const SphCSR& bf,
                                                   const SphDfsph& d, int k0, int sub) {
  const SphPhysics P = phys[b.bscene[k0]];
  const KConst K = kconst(P.h);
  const float2 xb = ld2(b.x, k0);
  float s2 = 0.f, src = 0.f;
  pair_pipeline_with<KL>(
      bf.rowptr[k0], bf.rowptr[k0 + 1], sub, [&](int k) { return bf.col[k]; },
      [&](int j) { return Rec2{ld4(d.xrec, j), ld4(d.prec, j)}; }, [&](const Rec2& r) {
        const float2 G = spiky_G(xb.x, xb.y, r.a.x, r.a.y, K);
        const float aa = r.b.z;
        s2 += aa * aa / r.a.w * (G.x * G.x + G.y * G.y);
        src += aa * ((0.f - r.b.x) * G.x + (0.f - r.b.y) * G.y);
      });
  s2 = group_sum<KL>(s2);
  src = group_sum<KL>(src);
  if (sub != 0) return;
  d.b_alpha[k0] = -P.dt * P.dt * d.b_aact[k0] * s2;   // unclamped (akinciBoundary.py:340)
  d.b_src[k0] = 1.f - b.rho[k0] - P.dt * src;
}

// What the pressure update of fluid row i reads besides its fluid sum:
// loaded apart so that a kernel can issue the loads before the sum.
struct PressIn {
  float2 ai, wi;         // a*_i, W_i
  float src, p2, alpha;  // s_i, the current iterate, alpha_i
  float rho;             // rho_hat rho0
};
__device__ __forceinline__ PressIn press_in(const SphFluid& f, const SphDfsph& d, int i) {
  PressIn r;
  r.ai = ld2(d.arec, 2 * i);
  r.wi = ld2(d.wvec, i);
  r.src = d.src[i];
  r.p2 = d.p2[i];
  r.alpha = d.alpha[i];
  r.rho = f.rho[i] * f.rest[i];
  return r;
}

// The relaxed update of fluid row i from its fluid sum ks and the clamped
// residual for the scene sum (dfsph.py:44-57,164-184):
//   r_i = a*_i . W_i - ks - s_i.
__device__ __forceinline__ double pressure_row_update(const SphDfsph& d, const PressIn& in,
                                                      float omega, float tau, int s, int i,
                                                      float ks) {
  const float rr = (in.ai.x * in.wi.x + in.ai.y * in.wi.y) - ks - in.src;
  d.p[i] = in.p2;
  float pn = in.p2 - omega * rr / in.alpha;
  // densitySolve clamps at 0 (clamp(min = 0) keeps NaN); divergenceSolve
  // does not (dfsph.py:177)
  if (!d.divergence) pn = pn < 0.f ? 0.f : pn;
  if (isnan(pn) || isinf(pn)) atomicOr(&d.status[s], SPH_ST_NAN);
  d.p2[i] = pn;
  d.resid[i] = rr;
  d.xrec[4 * i + 2] = pn / (in.rho * in.rho);
  return (double)(rr < -tau ? -tau : rr);
}

// The pressure update of boundary item tb of the active list
// (akinciBoundary.py:280-315): its scene into *sc and, on the group's first
// lane, *valid and the clamped residual.
template <int KL>
__device__ __forceinline__ double pressure_boundary_row(const SphBoundary& b, const SphCSR& bf,
                                                        const SphDfsph& d,
                                                        const SphPhysics* __restrict__ phys,
                                                        const int* preb, int nact, int r0b,
                                                        int tb, int sub, int* sc, bool* valid) {
  const int r = advance_rank(preb, nact, r0b, tb);
  const int s = d.act_scene[r];
  *sc = s;
  const int k0 = d.bbegin[s] + (tb - preb[r]);
  const SphPhysics P = phys[s];
  const KConst K = kconst(P.h);
  const float2 xb = ld2(d.bxrec, 2 * k0);
  float ks = 0.f;
  // sum_f w_f a*_f . (-G_bf) with {x_f, y_f, w_f a*_f} from prec
  pair_pipeline<KL>(bf.rowptr[k0], bf.rowptr[k0 + 1], sub, d.prec,
                    [&](int k) { return bf.col[k]; }, [&](const float4& pf) {
                      const float2 G = spiky_G(xb.x, xb.y, pf.x, pf.y, K);
                      ks -= pf.z * G.x + pf.w * G.y;
                    });
  ks = group_sum<KL>(ks);
  if (sub != 0) return 0.0;
  const float rr = ks - d.b_src[k0];
  float pn = d.b_p[k0] - P.omega * rr / d.b_alpha[k0];
  pn = pn < 0.f ? 0.f : pn;
  d.b_p[k0] = pn;
  d.b_resid[k0] = rr;
  const float rb = b.rho[k0] * b.rest[k0];
  d.bxrec[4 * k0 + 2] = pn / (rb * rb);
  *valid = true;
  return (double)(rr < -P.tau ? -P.tau : rr);
}

// The end of the solve for fluid row i with its final pressure
// acceleration a: a*, p, v* and (acc != NULL) acc += a.
struct FinishIn {
  float2 vp, acc;   // v*_i, acc_i (0 without acc)
  float p2;
};
__device__ __forceinline__ FinishIn finish_in(const SphDfsph& d, int i,
                                              const float* __restrict__ acc) {
  return FinishIn{ld2(d.vpred, i), acc ? ld2(acc, i) : make_float2(0.f, 0.f), d.p2[i]};
}
__device__ __forceinline__ void finish_store(const SphPhysics& P, const SphDfsph& d, int i,
                                             const FinishIn& in, float2 a,
                                             float* __restrict__ acc) {
  st2(d.apred, i, a);
  d.p[i] = in.p2;
  st2(d.vpred, i, make_float2(in.vp.x + P.dt * a.x, in.vp.y + P.dt * a.y));
  if (acc) st2(acc, i, make_float2(in.acc.x + a.x, in.acc.y + a.y));
}

// finish_store with the diffusion terms: acc = (acc + a) + their
// acceleration.
__device__ __forceinline__ void finish_visc_store(const SphPhysics& P, const Own& o,
                                                  const SphDfsph& d, const Acc& va,
                                                  unsigned vterms, int i, const FinishIn& in,
                                                  float2 a, float* __restrict__ acc) {
  st2(d.apred, i, a);
  d.p[i] = in.p2;
  st2(d.vpred, i, make_float2(in.vp.x + P.dt * a.x, in.vp.y + P.dt * a.y));
  float dp, vx, vy;
  dsph_rates(P, o, va, vterms, dp, vx, vy);
  const float2 w = make_float2(in.acc.x + a.x, in.acc.y + a.y);
  st2(acc, i, make_float2(w.x + vx, w.y + vy));
}

__global__ __launch_bounds__(kBS) void k_prepare(SphFluid f, const float* __restrict__ acc,
                                                 const SphPhysics* __restrict__ phys, SphBoundary b,
                                                 SphDfsph d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < f.np) {
    const SphPhysics P = phys[f.pscene[t]];
    const float2 v = ld2(f.v, t);
    const float2 a = acc ? ld2(acc, t) : make_float2(0.f, 0.f);
    st2(d.vpred, t, make_float2(v.x + P.dt * a.x, v.y + P.dt * a.y));
    const float aact = f.area[t] / f.rho[t];
    d.aact[t] = aact;
    d.p[t] = 0.f;
    d.p2[t] = 0.f;
    st2(d.apred, t, make_float2(0.f, 0.f));
    const float2 x = ld2(f.x, t);
    const float mass = f.area[t] * f.rest[t];
    st4(d.xrec, t, make_float4(x.x, x.y, 0.f, mass));
    st4(d.arec, t, make_float4(0.f, 0.f, P.dt * P.dt * aact, 0.f));
    // until the first k_accel rewrites it, prec carries what the alpha /
    // source gather needs besides the position: {v*, A^act, A rho0}
    st4(d.prec, t, make_float4(v.x + P.dt * a.x, v.y + P.dt * a.y, aact, mass));
  } else if (t < f.np + b.nb) {
    const int k = t - f.np;
    d.b_aact[k] = b.vol[k] / b.rho[k];
    d.b_p[k] = 0.f;
    const float2 x = ld2(b.x, k);
    st4(d.bxrec, k, make_float4(x.x, x.y, 0.f, b.vol[k] * b.rest[k]));
  }
  if (t < SPH_DFSPH_ERR_SLOTS * SPH_DFSPH_ERR_STRIDE(d.nscenes)) d.err_acc[t] = 0.0;
  if (t < d.nscenes) {
    d.iters[t] = 0;
    d.active[t] = 1;
  }
  if (t == 0) *d.any_active = 1;
}

// Rebuilds the active scene list and its prefix sums (one workgroup).
__device__ void build_active_lists(SphDfsph& d) {
  __shared__ int lds[kBS / kWave + 1];
  int carry_s = 0, carry_f = 0, carry_b = 0;
  for (int base = 0; base < d.nscenes; base += blockDim.x) {
    const int s = base + threadIdx.x;
    const int act = (s < d.nscenes) ? d.active[s] : 0;
    const int nf = act ? d.nfluid[s] : 0;
    const int nb = (act && !d.divergence) ? d.nbdy[s] : 0;
    int tot;
    const int rank = block_exclusive_scan(act, lds, &tot) + carry_s;
    const int tot_s = tot;
    const int pf = block_exclusive_scan(nf, lds, &tot) + carry_f;
    const int tot_f = tot;
    const int pb = block_exclusive_scan(nb, lds, &tot) + carry_b;
    const int tot_b = tot;
    if (act) {
      d.act_scene[rank] = s;
      d.act_pre[rank] = pf;
      d.act_pre[d.nscenes + 1 + rank] = pb;
    }
    carry_s += tot_s;
    carry_f += tot_f;
    carry_b += tot_b;
  }
  if (threadIdx.x == 0) {
    d.act_pre[carry_s] = carry_f;
    d.act_pre[d.nscenes + 1 + carry_s] = carry_b;
    d.act_n[0] = carry_f;
    d.act_n[1] = carry_b;
    d.act_n[2] = carry_s;
    *d.any_active = carry_s > 0;
  }
}

__global__ __launch_bounds__(kBS) void k_active_lists(SphDfsph d) { build_active_lists(d); }

// KL lanes share one row (fluid particle or boundary particle).
template <int KL>
__global__ __launch_bounds__(kBS) void k_alpha_source(SphFluid f, const SphPhysics* __restrict__ phys,
                                                      SphCSR ff, SphBoundary b, SphCSR fb, SphCSR bf,
                                                      SphDfsph d) {
  const int64_t gt = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int t = (int)(gt / KL), sub = (int)(gt % KL);
  if (t < f.np) {
    const int i = t;
    const SphPhysics P = phys[f.pscene[i]];
    const KConst K = kconst(P.h);
    const float2 xi = ld2(f.x, i), vi = ld2(d.vpred, i);
    float k1x = 0.f, k1y = 0.f, k2 = 0.f, src = 0.f;
    // fluid neighbours: xrec = {x, y, -, A rho0}, prec = {v*, A^act, A rho0}
    pair_pipeline_with<KL>(
        ff.rowptr[i], ff.rowptr[i + 1], sub, [&](int k) { return ff.col[k]; },
        [&](int j) { return Rec2{ld4(d.xrec, j), ld4(d.prec, j)}; }, [&](const Rec2& r) {
          const float2 G = spiky_G(xi.x, xi.y, r.a.x, r.a.y, K);
          const float aa = r.b.z;
          k1x += aa * G.x;
          k1y += aa * G.y;
          k2 += aa * aa / r.a.w * (G.x * G.x + G.y * G.y);
          src += -P.dt * aa * ((vi.x - r.b.x) * G.x + (vi.y - r.b.y) * G.y);
        });
    if (fb.rowptr) {
      // boundary neighbours: bxrec = {x, y, -, V rho0}; predicted velocity 0
      // (static walls)
      pair_pipeline_with<KL>(
          fb.rowptr[i], fb.rowptr[i + 1], sub, [&](int k) { return fb.col[k]; },
          [&](int j) { return Rec2{ld4(d.bxrec, j), make_float4(d.b_aact[j], 0.f, 0.f, 0.f)}; },
          [&](const Rec2& r) {
            const float2 G = spiky_G(xi.x, xi.y, r.a.x, r.a.y, K);
            const float aa = r.b.x;
            k1x += aa * G.x;
            k1y += aa * G.y;
            k2 += aa * aa / r.a.w * (G.x * G.x + G.y * G.y);
            src += -P.dt * aa * (vi.x * G.x + vi.y * G.y);
          });
    }
    k1x = group_sum<KL>(k1x);
    k1y = group_sum<KL>(k1y);
    k2 = group_sum<KL>(k2);
    src = group_sum<KL>(src);
    if (sub != 0) return;
    alpha_store(P, d, i, alpha_in(f, d, i), k1x, k1y, k2, src);
  } else if (t < f.np + b.nb) {
    alpha_boundary_row<KL>(phys, b, bf, d, t - f.np, sub);
  }
}

template <int KL, bool STAGED>
__device__ __forceinline__ float2 pressure_accel_t(const SphCSR& ff, const SphCSR& fb,
                                                   const SphDfsph& d, const KConst& K, int i,
                                                   int sub, const WaveCols& w, const int* wcols) {
  const float4 ri = ld4(d.xrec, i);
  float ax = 0.f, ay = 0.f;
  auto add = [&](const float4& rj) {
    const float2 G = spiky_G(ri.x, ri.y, rj.x, rj.y, K);
    const float c = -rj.w * (ri.z + rj.z);
    ax += c * G.x;
    ay += c * G.y;
  };
  pair_pipeline<KL>(ff.rowptr[i], ff.rowptr[i + 1], sub, d.xrec,
                    [&](int k) { return wave_col<STAGED>(ff.col, w, wcols, k); }, add);
  if (fb.rowptr)
    pair_pipeline<KL>(fb.rowptr[i], fb.rowptr[i + 1], sub, d.bxrec,
                      [&](int k) { return fb.col[k]; }, add);
  return make_float2(group_sum<KL>(ax), group_sum<KL>(ay));
}

// a*_i = sum_j -(A_j rho0_j)(pt_i + pt_j) G_ij + boundary (dfsph.py:25-42,
// akinciBoundary.py:229-252), pt = p / (rho rho0)^2.  Lane `sub` of a group
// of KL takes pairs sub, sub + KL, ...; every lane returns the group's sum.
template <int KL>
__device__ __forceinline__ float2 pressure_accel(const SphCSR& ff, const SphCSR& fb,
                                                 const SphDfsph& d, const KConst& K, int i,
                                                 int sub, const WaveCols& w, const int* wcols) {
  if (w.full) return pressure_accel_t<KL, true>(ff, fb, d, K, i, sub, w, wcols);
  return pressure_accel_t<KL, false>(ff, fb, d, K, i, sub, w, wcols);
}

// Work item (active fluid row) of this lane group and the stride between
// the items of successive grid passes; group-uniform.
template <int KL>
__device__ __forceinline__ int item0() {
  return (int)(((int64_t)xcd_block() * blockDim.x + threadIdx.x) / KL);
}
template <int KL>
__device__ __forceinline__ int item_stride() {
  return (int)((int64_t)gridDim.x * blockDim.x / KL);
}

// a*_i and the record the pressure update gathers: {x, y, w a*} with
// w = dt^2 A^act.
__device__ __forceinline__ void store_accel(const SphDfsph& d, int i, float2 a, float dt2) {
  const float w = dt2 * d.aact[i];
  st4(d.arec, i, make_float4(a.x, a.y, w, 0.f));
  const float2 x = ld2(d.xrec, 2 * i);
  st4(d.prec, i, make_float4(x.x, x.y, w * a.x, w * a.y));
}

template <int KL>
__global__ __launch_bounds__(kBS) void k_accel(SphFluid f, const SphPhysics* __restrict__ phys,
                                               SphCSR ff, SphCSR fb, SphDfsph d) {
  __shared__ int s_cols[kBS / kWave][kWaveCols];
  int* wcols = s_cols[threadIdx.x / kWave];
  const int nact = d.act_n[2];
  const int nitems = d.act_n[0];
  const int sub = threadIdx.x % KL;
  const int per_block = blockDim.x / KL;
  for (int base = xcd_block() * per_block; base < nitems; base += gridDim.x * per_block) {
    const int t = base + threadIdx.x / KL;
    const int r0 = find_rank(d.act_pre, nact, base);
    // a pass inside one active scene takes that scene's data with
    // workgroup-uniform (scalar) loads
    const bool uni = base + per_block <= d.act_pre[r0 + 1];
    const bool valid = t < nitems;
    int s = 0, i = 0, k0 = 0, k1 = 0;
    if (uni) {
      s = d.act_scene[r0];
      i = d.fbegin[s] + (t - d.act_pre[r0]);
    } else if (valid) {
      const int r = advance_rank(d.act_pre, nact, r0, t);
      s = d.act_scene[r];
      i = d.fbegin[s] + (t - d.act_pre[r]);
    }
    if (valid) {
      k0 = ff.rowptr[i];
      k1 = ff.rowptr[i + 1];
    }
    const WaveCols w = stage_wave_cols<KL>(ff.col, i, k0, k1, valid, wcols);
    if (uni) {
      const SphPhysics P = phys[d.act_scene[r0]];
      const KConst K = kconst(P.h);
      const float2 a = pressure_accel<KL>(ff, fb, d, K, i, sub, w, wcols);
      if (sub == 0) store_accel(d, i, a, P.dt * P.dt);
    } else if (valid) {
      const SphPhysics P = phys[s];
      const KConst K = kconst(P.h);
      const float2 a = pressure_accel<KL>(ff, fb, d, K, i, sub, w, wcols);
      if (sub == 0) store_accel(d, i, a, P.dt * P.dt);
    }
  }
}

// Adds v (FP64) into acc[scene] with as few atomics as the scene layout
// allows: one per workgroup when all its valid items belong to one scene,
// else one per wavefront whose valid items share a scene, else one per
// item.  Atomics on one address serialise, so per-wavefront atomics of a
// large single scene would queue tens of thousands deep.  `acc` is the
// workgroup's slot row.  Every thread of the workgroup calls this the same
// number of times (it holds two workgroup barriers).
__device__ __forceinline__ void scene_sum(double v, int s, bool valid, double* acc, double* wsum,
                                          int* wscene) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const int s0 = __shfl(valid ? s : -1, 0, kWave);
  const unsigned long long bv = __ballot(valid);
  const bool wuni = bv != 0 && s0 >= 0 && __ballot(valid && s == s0) == bv;
  const double w = wave_sum(valid ? v : 0.0);
  // per wave: its scene, -1 when it has no valid item, -2 when mixed
  if (lane == 0) {
    wsum[wv] = w;
    wscene[wv] = bv == 0 ? -1 : (wuni ? s0 : -2);
  }
  __syncthreads();
  int bs = -1;
  bool buni = true;
  for (int k = 0; k < (int)(blockDim.x / kWave); ++k) {
    const int ws = wscene[k];
    if (ws == -2 || (ws >= 0 && bs >= 0 && ws != bs)) buni = false;
    if (ws >= 0 && bs < 0) bs = ws;
  }
  if (buni) {
    if (threadIdx.x == 0 && bs >= 0) {
      double tot = 0.0;
      for (int k = 0; k < (int)(blockDim.x / kWave); ++k)
        if (wscene[k] >= 0) tot += wsum[k];
      atomicAdd(&acc[bs], tot);
    }
  } else if (wuni) {
    if (lane == 0) atomicAdd(&acc[s0], w);
  } else if (valid) {
    atomicAdd(&acc[s], v);
  }
  __syncthreads();   // the next pass rewrites wsum / wscene
}

// The fluid-row part of k_pressure (dfsph.py:44-57,164-184 with the Akinci
// terms of akinciBoundary.py:280-315):
//   r_i = sum_j w_j (a*_i - a*_j) . G_ij + sum_b w_b a*_i . G_ib - s_i
//       = a*_i . W_i - sum_j (w_j a*_j) . G_ij - s_i,
// W_i = sum_j w_j G_ij + sum_b w_b G_ib being fixed during the solve (it is
// dt^2 k1_i of the alpha term).  One 16-byte record {x_j, y_j, w_j a*_j}
// per pair; boundary neighbours only enter through W_i.  Then the relaxed
// update of p2 and the clamped residual for the scene sum.
template <int KL, bool STAGED>
__device__ __forceinline__ double pressure_fluid_row(const SphFluid& f, const SphCSR& ff,
                                                     const SphDfsph& d, const KConst& K,
                                                     float omega, float tau, int s, int i, int k0,
                                                     int k1, int sub, const WaveCols& w,
                                                     const int* wcols, bool* wrote) {
  const float2 xi = ld2(d.xrec, 2 * i);          // {x, y} of xrec[i]
  float ks = 0.f;
  pair_pipeline<KL>(k0, k1, sub, d.prec, [&](int k) { return wave_col<STAGED>(ff.col, w, wcols, k); },
                    [&](const float4& pj) {
                      const float2 G = spiky_G(xi.x, xi.y, pj.x, pj.y, K);
                      ks += pj.z * G.x + pj.w * G.y;
                    });
  ks = group_sum<KL>(ks);
  *wrote = false;
  if (sub != 0) return 0.0;
  *wrote = true;
  return pressure_row_update(d, press_in(f, d, i), omega, tau, s, i, ks);
}

template <int KL>
// Eight wavefronts per SIMD: held to 64 VGPRs (a few spill), it runs ~7%
// faster than at 73 VGPRs and six wavefronts.
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(8))) void k_pressure(SphFluid f, const SphPhysics* __restrict__ phys,
                                                  SphCSR ff, SphBoundary b, SphCSR fb, SphCSR bf,
                                                  SphDfsph d) {
  const int nact = d.act_n[2];
  const int nf = d.act_n[0];
  const int nitems = nf + d.act_n[1];
  const int per_block = blockDim.x / KL;
  const int stride = gridDim.x * per_block;
  const int sub = threadIdx.x % KL;
  const int* preb = d.act_pre + d.nscenes + 1;
  double* acc_slot =
      d.err_acc + (size_t)(blockIdx.x % SPH_DFSPH_ERR_SLOTS) * SPH_DFSPH_ERR_STRIDE(d.nscenes);
  __shared__ double s_wsum[kBS / kWave];
  __shared__ int s_wscene[kBS / kWave];
  __shared__ int s_cols[kBS / kWave][kWaveCols];
  int* wcols = s_cols[threadIdx.x / kWave];
  // every lane runs the same number of loop trips so the wavefront-wide
  // reductions inside see all lanes
  for (int base = xcd_block() * per_block; base < nitems; base += stride) {
    const int t = base + threadIdx.x / KL;
    double contrib = 0.0;
    int s = -1;
    bool valid = false;
    // workgroup-uniform ranks of the pass's first fluid and boundary items
    const int r0f = base < nf ? find_rank(d.act_pre, nact, base) : 0;
    const int r0b = base + per_block > nf ? find_rank(preb, nact, base > nf ? base - nf : 0) : 0;
    // a pass of fluid items inside one active scene: scalar scene data
    const bool uni = base + per_block <= nf && base + per_block <= d.act_pre[r0f + 1];
    int i = 0, k0 = 0, k1 = 0;
    if (uni) {
      s = d.act_scene[r0f];
      i = d.fbegin[s] + (t - d.act_pre[r0f]);
    } else if (t < nf) {
      const int r = advance_rank(d.act_pre, nact, r0f, t);
      s = d.act_scene[r];
      i = d.fbegin[s] + (t - d.act_pre[r]);
    }
    if (t < nf) {
      k0 = ff.rowptr[i];
      k1 = ff.rowptr[i + 1];
    }
    const WaveCols w = stage_wave_cols<KL>(ff.col, i, k0, k1, t < nf, wcols);
    if (uni) {
      const SphPhysics P = phys[d.act_scene[r0f]];
      const KConst K = kconst(P.h);
      contrib = w.full ? pressure_fluid_row<KL, true>(f, ff, d, K, P.fluid_omega, P.tau, s, i, k0,
                                                      k1, sub, w, wcols, &valid)
                       : pressure_fluid_row<KL, false>(f, ff, d, K, P.fluid_omega, P.tau, s, i, k0,
                                                       k1, sub, w, wcols, &valid);
    } else if (t < nf) {
      const SphPhysics P = phys[s];
      const KConst K = kconst(P.h);
      contrib = w.full ? pressure_fluid_row<KL, true>(f, ff, d, K, P.fluid_omega, P.tau, s, i, k0,
                                                      k1, sub, w, wcols, &valid)
                       : pressure_fluid_row<KL, false>(f, ff, d, K, P.fluid_omega, P.tau, s, i, k0,
                                                       k1, sub, w, wcols, &valid);
    } else if (t < nitems) {
      contrib = pressure_boundary_row<KL>(b, bf, d, phys, preb, nact, r0b, t - nf, sub, &s,
                                          &valid);
    }
    scene_sum(contrib, s, valid, acc_slot, s_wsum, s_wscene);
  }
}

// Per-scene loop control, densitySolve (dfsph.py:205-207): after iteration
// i the solve continues while (i < minIters or err > tau) and i <= maxIters.
__global__ __launch_bounds__(kBS) void k_control(const SphPhysics* __restrict__ phys, SphDfsph d) {
  for (int s = threadIdx.x; s < d.nscenes; s += blockDim.x) {
    if (!d.active[s]) continue;
    const SphPhysics P = phys[s];
    // densitySolve with the PBSPH boundary averages over fluid and boundary
    // rows (dfsph.py:219-222); divergenceSolve over fluid rows (:253)
    const double n = (double)(d.nfluid[s] + (d.divergence ? 0 : d.nbdy[s]));
    double sum = 0.0;
    const size_t stride = SPH_DFSPH_ERR_STRIDE(d.nscenes);
#pragma unroll 8
    for (int k = 0; k < SPH_DFSPH_ERR_SLOTS; ++k) {
      sum += d.err_acc[k * stride + s];
      d.err_acc[k * stride + s] = 0.0;
    }
    const double err = n > 0 ? sum / n : 0.0;
    const int it = d.iters[s];
    if (it < d.max_hist) d.err_hist[(size_t)s * d.max_hist + it] = err;
    const int i = it + 1;
    d.iters[s] = i;
    const bool go = (i < P.min_iters || err > (double)P.tau) && i <= P.max_iters;
    if (!go) {
      d.active[s] = 0;
      // the divergence solve stopping at its small cap is routine
      if (err > (double)P.tau && !d.divergence) atomicOr(&d.status[s], SPH_ST_ITERCAP);
    }
  }
  __syncthreads();
  build_active_lists(d);
}

template <int KL>
__global__ __launch_bounds__(kBS) void k_finish(SphFluid f, const SphPhysics* __restrict__ phys,
                                                SphCSR ff, SphCSR fb, SphDfsph d,
                                                float* __restrict__ acc) {
  const int i = item0<KL>();
  const int sub = threadIdx.x % KL;
  if (i >= f.np) return;
  const SphPhysics P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const WaveCols none = {0, 0, false};
  const float2 a = pressure_accel<KL>(ff, fb, d, K, i, sub, none, nullptr);
  if (sub != 0) return;
  finish_store(P, d, i, finish_in(d, i, acc), a, acc);
}

// k_finish and the velocity-diffusion and laminar gather of the step
// (diffusion.py:30-44, laminar.py:30-46, the terms sph_dsph_gather_packed
// adds with SPH_T_VISC | SPH_T_LAM | SPH_T_ACCUM) over one walk of the
// fluid list: per pair the pressure record xrec[j] and the packed record
// frec[j] of the step's state (V = A over the fluid-only density).  acc
// becomes (acc + a*) + the diffusion terms, summed in the order and on the
// lanes of the two separate kernels, so with the same lanes per row the
// result is theirs bit for bit.
template <int KL>
__global__ __launch_bounds__(kBS) void k_finish_visc(SphFluid f,
                                                     const SphPhysics* __restrict__ phys,
                                                     SphCSR ff, SphCSR fb, SphDfsph d,
                                                     const float4* __restrict__ frec,
                                                     unsigned vterms, float* __restrict__ acc) {
  const int i = item0<KL>();
  const int sub = threadIdx.x % KL;
  if (i >= f.np) return;
  const SphPhysics P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const float4 ri = ld4(d.xrec, i);
  const float4 a0 = frec[2 * i], a1 = frec[2 * i + 1];
  Own o;
  o.x = a0.x; o.y = a0.y; o.vx = a0.z; o.vy = a0.w;
  o.rho = a1.x;
  o.p = a1.z;
  o.gx = 0.f;
  o.gy = 0.f;
  float ax = 0.f, ay = 0.f;
  Acc va = {};
  struct Rec3 {
    float4 x, r0, r1;
  };
  pair_pipeline_with<KL>(
      ff.rowptr[i], ff.rowptr[i + 1], sub, [&](int k) { return ff.col[k]; },
      [&](int j) { return Rec3{ld4(d.xrec, j), frec[2 * j], frec[2 * j + 1]}; },
      [&](const Rec3& r) {
        const float2 G = spiky_G(ri.x, ri.y, r.x.x, r.x.y, K);
        const float c = -r.x.w * (ri.z + r.x.z);
        ax += c * G.x;
        ay += c * G.y;
        const Nbr nb = {r.r0.x, r.r0.y, r.r0.z, r.r0.w, r.r1.x, r.r1.y, r.r1.z};
        dsph_pair(P, K, o, nb, 0.f, 0.f, true, vterms, va);
      });
  if (fb.rowptr)
    pair_pipeline<KL>(fb.rowptr[i], fb.rowptr[i + 1], sub, d.bxrec,
                      [&](int k) { return fb.col[k]; }, [&](const float4& rj) {
                        const float2 G = spiky_G(ri.x, ri.y, rj.x, rj.y, K);
                        const float c = -rj.w * (ri.z + rj.z);
                        ax += c * G.x;
                        ay += c * G.y;
                      });
  const float2 a = make_float2(group_sum<KL>(ax), group_sum<KL>(ay));
  reduce_acc<KL>(va, vterms);
  if (sub != 0) return;
  finish_visc_store(P, o, d, va, vterms, i, finish_in(d, i, acc), a, acc);
}

// ------------------------------------------------ stencil-list gathers
// The DFSPH gathers over stencil lists (SphStencilList) of a state held in
// its grid's sorted order (point index = sorted slot).  A workgroup takes
// the fluid rows of a pass (kBS consecutive rows, or the pass's rows of the
// active list), stages the records of their three stencil runs in LDS with
// coalesced loads and one wait, and walks each row's mask bits from there,
// one lane per row: the pairs of the CSR row in its order, so every sum is
// the CSR gather's with one lane per row bit for bit.  A pass whose runs do
// not fit in LDS (rows of two scenes, crowded cells) reads the records from
// global memory.  The boundary pairs keep their CSR lists.
template <int NQ>
struct SlRecs {
  float4 a[NQ];
};
template <int NQ>
struct SlCap {
  static constexpr int v = NQ == 1 ? 3 * kBS * 3 / 2 + 256 : 1024;   // staged points
};
template <int NQ>
struct SlShared {
  float4 rec[NQ][SlCap<NQ>::v];
  int wlo[kBS / kWave][3], whi[kBS / kWave][3];
};
constexpr int kSlBatch = 3;   // pairs a lane loads before it adds them

// Row r of a stencil list: its runs (as a Stencil3's ra / rb) and masks.
struct SlRow {
  Stencil3 st;
  unsigned long long m[3];
};
__device__ __forceinline__ SlRow sl_row(const SphStencilList& sl, int r) {
  SlRow w;
  const int4 R = reinterpret_cast<const int4*>(sl.runs)[r];
  w.st.ra[0] = R.x;
  w.st.ra[1] = R.y;
  w.st.ra[2] = R.z;
  w.st.rb[0] = R.x + (R.w & 255);
  w.st.rb[1] = R.y + ((R.w >> 8) & 255);
  w.st.rb[2] = R.z + ((R.w >> 16) & 255);
  w.st.n0 = w.st.n01 = w.st.ncand = 0;
  w.m[0] = sl.masks[r];
  w.m[1] = sl.masks[(size_t)sl.n + r];
  w.m[2] = sl.masks[2 * (size_t)sl.n + r];
  return w;
}
__device__ __forceinline__ SlRow sl_none() {
  SlRow w = {};
  return w;
}

// The chunk's runs and whether they are staged (workgroup-uniform).  Every
// thread of the workgroup calls this (two barriers); load(j) gives the
// records of point j.
template <int NQ, typename Load>
__device__ __forceinline__ bool sl_stage(SlShared<NQ>& S, const Stencil3& st, bool valid,
                                         Load load, ChunkRuns& cr) {
  cr = chunk_runs(st, valid, S.wlo, S.whi);
  const int tot = cr.tot;
  const bool staged = tot <= SlCap<NQ>::v;
  if (staged) {
    // slots per round: all loads of a round before its LDS writes
    constexpr int R = NQ == 1 ? 3 : 2;
    const int t = threadIdx.x;
    const int w0 = __builtin_amdgcn_readfirstlane(t);
    for (int u0 = 0; w0 + u0 * kBS < tot; u0 += R) {
      SlRecs<NQ> v[R];
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const int k = t + (u0 + u) * kBS;
        v[u] = load(cr.slot(k < tot ? k : tot - 1));
      }
#pragma unroll
      for (int u = 0; u < R; ++u) {
        const int k = t + (u0 + u) * kBS;
        if (k < tot) {
#pragma unroll
          for (int q = 0; q < NQ; ++q) S.rec[q][k] = v[u].a[q];
        }
      }
    }
  }
  __syncthreads();
  return staged;
}

// f(records) for each pair of row w in order: from LDS (STAGED) or global
// memory, kSlBatch loads in flight.  When a row's three masks hold 64
// candidates or fewer they are laid end to end in one mask and walked as
// one stream (bit t1 starts the second run, t2 the third), so a wavefront
// pads its loop to its longest row rather than to its longest run three
// times; a row with more walks its runs one after the other, in passes 1
// and 2 of the same loop.
__device__ __forceinline__ unsigned long long shl64(unsigned long long x, int s) {
  return s < 64 ? x << s : 0ull;
}
template <bool STAGED, int NQ, typename Load, typename F>
__device__ __forceinline__ void sl_walk_t(const SlShared<NQ>& S, const ChunkRuns& cr,
                                          const SlRow& w, Load load, F f) {
  int n[3], a[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int len = w.st.rb[r] - w.st.ra[r];
    n[r] = len < 64 ? len : 64;
    a[r] = STAGED ? cr.first(r, w.st) : w.st.ra[r];
  }
  const bool one = n[0] + n[1] + n[2] <= 64;
#pragma unroll 1
  for (int pass = 0; pass < 3; ++pass) {
    unsigned long long mm;
    int t1, t2, b0, b1, b2;
    if (pass == 0) {
      mm = one ? (w.m[0] | shl64(w.m[1], n[0]) | shl64(w.m[2], n[0] + n[1])) : w.m[0];
      t1 = one ? n[0] : 64;
      t2 = one ? n[0] + n[1] : 64;
      b0 = a[0];
      b1 = a[1] - n[0];
      b2 = a[2] - n[0] - n[1];
    } else {
      mm = one ? 0ull : (pass == 1 ? w.m[1] : w.m[2]);
      t1 = t2 = 64;
      b0 = b1 = b2 = pass == 1 ? a[1] : a[2];
    }
    while (mm) {
      int k[kSlBatch];
      bool ok[kSlBatch];
#pragma unroll
      for (int u = 0; u < kSlBatch; ++u) {
        ok[u] = mm != 0ull;
        const int c = ok[u] ? (int)__builtin_ctzll(mm) : 0;
        // a slot past the row's last pair reloads its first (in bounds)
        k[u] = ok[u] ? c + (c < t1 ? b0 : (c < t2 ? b1 : b2)) : k[0];
        mm &= mm - 1ull;
      }
      SlRecs<NQ> v[kSlBatch];
#pragma unroll
      for (int u = 0; u < kSlBatch; ++u) {
        if (STAGED) {
#pragma unroll
          for (int q = 0; q < NQ; ++q) v[u].a[q] = S.rec[q][k[u]];
        } else {
          v[u] = load(k[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < kSlBatch; ++u)
        if (ok[u]) f(v[u]);
    }
  }
}
template <int NQ, typename Load, typename F>
__device__ __forceinline__ void sl_walk(const SlShared<NQ>& S, const ChunkRuns& cr, bool staged,
                                        const SlRow& w, Load load, F f) {
  if (staged)
    sl_walk_t<true, NQ>(S, cr, w, load, f);
  else
    sl_walk_t<false, NQ>(S, cr, w, load, f);
}

// k_fluid_density (gather.hip) over a stencil list, one lane per row.
__global__ __launch_bounds__(kBS) void k_fluid_density_sl(SphFluid f,
                                                          const SphPhysics* __restrict__ phys,
                                                          SphStencilList sl, SphBoundary b,
                                                          SphCSR fb, float* __restrict__ rho_out,
                                                          float* __restrict__ vol_out) {
  __shared__ SlShared<1> S;
  const int i = xcd_block() * kBS + threadIdx.x;
  const bool valid = i < f.np;
  const SlRow w = valid ? sl_row(sl, i) : sl_none();
  auto load = [&](int j) {
    SlRecs<1> r;
    const float2 x = ld2(f.x, j);
    r.a[0] = make_float4(x.x, x.y, f.area[j], 0.f);
    return r;
  };
  // the row's own data, loaded under the stage's wait
  float2 xi = make_float2(0.f, 0.f);
  float hi = 0.f, ai = 0.f;
  if (valid) {
    xi = ld2(f.x, i);
    hi = phys[f.pscene[i]].h;
    ai = f.area[i];
  }
  ChunkRuns cr;
  const bool staged = sl_stage<1>(S, w.st, valid, load, cr);
  if (!valid) return;
  const KConst K = kconst(hi);
  float sum = 0.f;
  sl_walk<1>(S, cr, staged, w, load, [&](const SlRecs<1>& r) {
    const PairGeo g = pair_geo(xi.x, xi.y, r.a[0].x, r.a[0].y, K.inv_h);
    sum += r.a[0].z * wendland_k(g.q, K);
  });
  if (vol_out) vol_out[i] = ai / sum;
  if (fb.rowptr) {
    float bs = 0.f;
    const int e1 = fb.rowptr[i + 1];
    for (int k = fb.rowptr[i]; k < e1; ++k) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xb.x, xb.y, xi.x, xi.y, K.inv_h);
      bs += wendland_k(g.q, K) * b.vol[j];
    }
    sum += bs;
  }
  rho_out[i] = sum;
}

// k_alpha_source's fluid rows over a stencil list; boundary rows (threads
// past the fluid) as in k_alpha_source with one lane per row.
__global__ __launch_bounds__(kBS) void k_alpha_source_sl(SphFluid f,
                                                         const SphPhysics* __restrict__ phys,
                                                         SphStencilList sl, SphBoundary b,
                                                         SphCSR fb, SphCSR bf, SphDfsph d) {
  __shared__ SlShared<2> S;
  const int t = xcd_block() * kBS + threadIdx.x;
  const bool fluid = t < f.np;
  const SlRow w = fluid ? sl_row(sl, t) : sl_none();
  auto load = [&](int j) {
    SlRecs<2> r;
    r.a[0] = ld4(d.xrec, j);
    r.a[1] = ld4(d.prec, j);
    return r;
  };
  // the row's own data, loaded under the stage's wait
  float2 xi = make_float2(0.f, 0.f), vi = xi;
  int sc = 0;
  AlphaIn ain = {};
  if (fluid) {
    xi = ld2(f.x, t);
    vi = ld2(d.vpred, t);
    sc = f.pscene[t];
    ain = alpha_in(f, d, t);
  }
  ChunkRuns cr;
  const bool staged = sl_stage<2>(S, w.st, fluid, load, cr);
  if (fluid) {
    const int i = t;
    const SphPhysics P = phys[sc];
    const KConst K = kconst(P.h);
    float k1x = 0.f, k1y = 0.f, k2 = 0.f, src = 0.f;
    sl_walk<2>(S, cr, staged, w, load, [&](const SlRecs<2>& rr) {
      const Rec2 r = {rr.a[0], rr.a[1]};
      const float2 G = spiky_G(xi.x, xi.y, r.a.x, r.a.y, K);
      const float aa = r.b.z;
      k1x += aa * G.x;
      k1y += aa * G.y;
      k2 += aa * aa / r.a.w * (G.x * G.x + G.y * G.y);
      src += -P.dt * aa * ((vi.x - r.b.x) * G.x + (vi.y - r.b.y) * G.y);
    });
    if (fb.rowptr) {
      pair_pipeline_with<1>(
          fb.rowptr[i], fb.rowptr[i + 1], 0, [&](int k) { return fb.col[k]; },
          [&](int j) { return Rec2{ld4(d.bxrec, j), make_float4(d.b_aact[j], 0.f, 0.f, 0.f)}; },
          [&](const Rec2& r) {
            const float2 G = spiky_G(xi.x, xi.y, r.a.x, r.a.y, K);
            const float aa = r.b.x;
            k1x += aa * G.x;
            k1y += aa * G.y;
            k2 += aa * aa / r.a.w * (G.x * G.x + G.y * G.y);
            src += -P.dt * aa * (vi.x * G.x + vi.y * G.y);
          });
    }
    alpha_store(P, d, i, ain, k1x, k1y, k2, src);
  } else if (t < f.np + b.nb) {
    alpha_boundary_row<1>(phys, b, bf, d, t - f.np, 0);
  }
}

// pressure_accel_t's sum over a stencil list (fluid pairs from the walk,
// boundary pairs from fb), one lane per row.
template <typename Walk>
__device__ __forceinline__ float2 pressure_accel_sl(const SphCSR& fb, const SphDfsph& d,
                                                    const KConst& K, int i, const float4& ri,
                                                    Walk walk) {
  float ax = 0.f, ay = 0.f;
  auto add = [&](const float4& rj) {
    const float2 G = spiky_G(ri.x, ri.y, rj.x, rj.y, K);
    const float c = -rj.w * (ri.z + rj.z);
    ax += c * G.x;
    ay += c * G.y;
  };
  walk([&](const SlRecs<1>& r) { add(r.a[0]); });
  if (fb.rowptr)
    pair_pipeline<1>(fb.rowptr[i], fb.rowptr[i + 1], 0, d.bxrec, [&](int k) { return fb.col[k]; },
                     add);
  return make_float2(ax, ay);
}

// Fluid row i of active item t of the pass of kBS items from `base` (every
// item of which is a fluid item when the pass is full), its scene s, and
// whether the pass lies in one scene su (that scene's data then uniform).
// With every scene active, the items are the fluid rows in order: no
// search of the active list.
struct PassRow {
  int i, s, su;
  bool uni;
};
__device__ __forceinline__ PassRow pass_row(const SphFluid& f, const SphDfsph& d, bool all,
                                            int nact, int nfl, int base, int t) {
  PassRow w = {0, 0, 0, false};
  const bool valid = t < nfl;
  if (all) {
    w.su = __builtin_amdgcn_readfirstlane(f.pscene[base]);
    w.uni = base + kBS <= nfl &&
            __builtin_amdgcn_readfirstlane(f.pscene[base + kBS - 1]) == w.su;
    w.i = t;
    w.s = valid ? (w.uni ? w.su : f.pscene[t]) : 0;
    return w;
  }
  const int r0 = find_rank(d.act_pre, nact, base);
  w.uni = base + kBS <= nfl && base + kBS <= d.act_pre[r0 + 1];
  w.su = d.act_scene[r0];
  if (w.uni) {
    w.s = w.su;
    w.i = d.fbegin[w.s] + (t - d.act_pre[r0]);
  } else if (valid) {
    const int r = advance_rank(d.act_pre, nact, r0, t);
    w.s = d.act_scene[r];
    w.i = d.fbegin[w.s] + (t - d.act_pre[r]);
  }
  return w;
}

// k_accel over a stencil list.
__global__ __launch_bounds__(kBS) void k_accel_sl(SphFluid f, const SphPhysics* __restrict__ phys,
                                                  SphStencilList sl, SphCSR fb, SphDfsph d) {
  __shared__ SlShared<1> S;
  const int nact = d.act_n[2];
  const int nitems = d.act_n[0];
  auto load = [&](int j) {
    SlRecs<1> r;
    r.a[0] = ld4(d.xrec, j);
    return r;
  };
  const bool all = nact == d.nscenes;
  for (int base = xcd_block() * kBS; base < nitems; base += gridDim.x * kBS) {
    const int t = base + threadIdx.x;
    const bool valid = t < nitems;
    const PassRow pr = pass_row(f, d, all, nact, nitems, base, t);
    const int i = pr.i;
    // the row's list, record and area, loaded under the stage's wait
    const SlRow w = valid ? sl_row(sl, i) : sl_none();
    const float4 ri = valid ? ld4(d.xrec, i) : make_float4(0.f, 0.f, 0.f, 0.f);
    const float aact = valid ? d.aact[i] : 0.f;
    ChunkRuns cr;
    const bool staged = sl_stage<1>(S, w.st, valid, load, cr);
    if (!valid) continue;
    const SphPhysics P = phys[pr.uni ? pr.su : pr.s];
    const KConst K = kconst(P.h);
    const float2 a = pressure_accel_sl(fb, d, K, i, ri, [&](auto f2) {
      sl_walk<1>(S, cr, staged, w, load, f2);
    });
    // store_accel with the row's area and position loaded above
    const float wv = P.dt * P.dt * aact;
    st4(d.arec, i, make_float4(a.x, a.y, wv, 0.f));
    st4(d.prec, i, make_float4(ri.x, ri.y, wv * a.x, wv * a.y));
  }
}

// k_pressure over a stencil list: the fluid rows' sums from the walk, the
// boundary rows as in k_pressure with one lane per row.
// Held to 71 VGPRs, seven wavefronts per SIMD (what its LDS allows): at 88
// registers it ran five (C3 70 -> 65 us per launch).
__global__ __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(7))) void k_pressure_sl(SphFluid f, const SphPhysics* __restrict__ phys,
                                                     SphStencilList sl, SphBoundary b, SphCSR fb,
                                                     SphCSR bf, SphDfsph d) {
  __shared__ SlShared<1> S;
  __shared__ double s_wsum[kBS / kWave];
  __shared__ int s_wscene[kBS / kWave];
  const int nact = d.act_n[2];
  const int nf = d.act_n[0];
  const int nitems = nf + d.act_n[1];
  const int stride = gridDim.x * kBS;
  const int* preb = d.act_pre + d.nscenes + 1;
  double* acc_slot =
      d.err_acc + (size_t)(blockIdx.x % SPH_DFSPH_ERR_SLOTS) * SPH_DFSPH_ERR_STRIDE(d.nscenes);
  auto load = [&](int j) {
    SlRecs<1> r;
    r.a[0] = ld4(d.prec, j);
    return r;
  };
  const bool all = nact == d.nscenes;
  for (int base = xcd_block() * kBS; base < nitems; base += stride) {
    const int t = base + threadIdx.x;
    double contrib = 0.0;
    int s = -1;
    bool valid = false;
    const int r0b = base + kBS > nf ? find_rank(preb, nact, base > nf ? base - nf : 0) : 0;
    PassRow pr = {0, 0, 0, false};
    if (base < nf) pr = pass_row(f, d, all, nact, nf, base, t);
    const int i = pr.i;
    const bool fl = t < nf;
    if (fl) s = pr.s;
    // the row's list, position and update inputs, loaded under the stage's wait
    const SlRow w = fl ? sl_row(sl, i) : sl_none();
    float2 xi = make_float2(0.f, 0.f);
    PressIn pin = {};
    if (fl) {
      xi = ld2(d.xrec, 2 * i);
      pin = press_in(f, d, i);
    }
    ChunkRuns cr;
    const bool staged = sl_stage<1>(S, w.st, fl, load, cr);
    if (fl) {
      const SphPhysics P = phys[pr.uni ? pr.su : s];
      const KConst K = kconst(P.h);
      float ks = 0.f;
      sl_walk<1>(S, cr, staged, w, load, [&](const SlRecs<1>& r) {
        const float4 pj = r.a[0];
        const float2 G = spiky_G(xi.x, xi.y, pj.x, pj.y, K);
        ks += pj.z * G.x + pj.w * G.y;
      });
      contrib = pressure_row_update(d, pin, P.fluid_omega, P.tau, s, i, ks);
      valid = true;
    } else if (t < nitems) {
      contrib = pressure_boundary_row<1>(b, bf, d, phys, preb, nact, r0b, t - nf, 0, &s, &valid);
    }
    scene_sum(contrib, s, valid, acc_slot, s_wsum, s_wscene);
  }
}

// k_finish over a stencil list.
__global__ __launch_bounds__(kBS) void k_finish_sl(SphFluid f, const SphPhysics* __restrict__ phys,
                                                   SphStencilList sl, SphCSR fb, SphDfsph d,
                                                   float* __restrict__ acc) {
  __shared__ SlShared<1> S;
  const int i = xcd_block() * kBS + threadIdx.x;
  const bool valid = i < f.np;
  const SlRow w = valid ? sl_row(sl, i) : sl_none();
  auto load = [&](int j) {
    SlRecs<1> r;
    r.a[0] = ld4(d.xrec, j);
    return r;
  };
  // the row's own data, loaded under the stage's wait
  float4 ri = make_float4(0.f, 0.f, 0.f, 0.f);
  int sc = 0;
  FinishIn fin = {};
  if (valid) {
    ri = ld4(d.xrec, i);
    sc = f.pscene[i];
    fin = finish_in(d, i, acc);
  }
  ChunkRuns cr;
  const bool staged = sl_stage<1>(S, w.st, valid, load, cr);
  if (!valid) return;
  const SphPhysics P = phys[sc];
  const KConst K = kconst(P.h);
  const float2 a = pressure_accel_sl(fb, d, K, i, ri, [&](auto f2) {
    sl_walk<1>(S, cr, staged, w, load, f2);
  });
  finish_store(P, d, i, fin, a, acc);
}

// k_finish_visc over a stencil list: per neighbour its pressure record and
// {vx, vy, rho, V} of its packed record (the position is the pressure
// record's, the same values; the pressure is not read: SPH_T_VISC and
// SPH_T_LAM only).
__global__ __launch_bounds__(kBS) void k_finish_visc_sl(SphFluid f,
                                                        const SphPhysics* __restrict__ phys,
                                                        SphStencilList sl, SphCSR fb, SphDfsph d,
                                                        const float4* __restrict__ frec,
                                                        unsigned vterms, float* __restrict__ acc) {
  __shared__ SlShared<2> S;
  const int i = xcd_block() * kBS + threadIdx.x;
  const bool valid = i < f.np;
  const SlRow w = valid ? sl_row(sl, i) : sl_none();
  auto load = [&](int j) {
    SlRecs<2> r;
    r.a[0] = ld4(d.xrec, j);
    const float4 p0 = frec[2 * j], p1 = frec[2 * j + 1];
    r.a[1] = make_float4(p0.z, p0.w, p1.x, p1.y);
    return r;
  };
  // the row's own data, loaded under the stage's wait
  float4 ri = make_float4(0.f, 0.f, 0.f, 0.f), a0 = ri, a1 = ri;
  int sc = 0;
  FinishIn fin = {};
  if (valid) {
    ri = ld4(d.xrec, i);
    a0 = frec[2 * i];
    a1 = frec[2 * i + 1];
    sc = f.pscene[i];
    fin = finish_in(d, i, acc);
  }
  ChunkRuns cr;
  const bool staged = sl_stage<2>(S, w.st, valid, load, cr);
  if (!valid) return;
  const SphPhysics P = phys[sc];
  const KConst K = kconst(P.h);
  Own o;
  o.x = a0.x; o.y = a0.y; o.vx = a0.z; o.vy = a0.w;
  o.rho = a1.x;
  o.p = a1.z;
  o.gx = 0.f;
  o.gy = 0.f;
  float ax = 0.f, ay = 0.f;
  Acc va = {};
  sl_walk<2>(S, cr, staged, w, load, [&](const SlRecs<2>& r) {
    const float2 G = spiky_G(ri.x, ri.y, r.a[0].x, r.a[0].y, K);
    const float c = -r.a[0].w * (ri.z + r.a[0].z);
    ax += c * G.x;
    ay += c * G.y;
    const Nbr nb = {r.a[0].x, r.a[0].y, r.a[1].x, r.a[1].y, r.a[1].z, r.a[1].w, 0.f};
    dsph_pair(P, K, o, nb, 0.f, 0.f, true, vterms, va);
  });
  if (fb.rowptr)
    pair_pipeline<1>(fb.rowptr[i], fb.rowptr[i + 1], 0, d.bxrec, [&](int k) { return fb.col[k]; },
                     [&](const float4& rj) {
                       const float2 G = spiky_G(ri.x, ri.y, rj.x, rj.y, K);
                       const float c = -rj.w * (ri.z + rj.z);
                       ax += c * G.x;
                       ay += c * G.y;
                     });
  finish_visc_store(P, o, d, va, vterms, i, fin, make_float2(ax, ay), acc);
}

}  // namespace

static bool dfsph_ok(const SphDfsph* d) {
  return d && d->vpred && d->aact && d->alpha && d->src && d->p && d->p2 && d->apred &&
         d->resid && d->xrec && d->arec && d->prec && d->wvec && d->err_acc && d->iters && d->active &&
         d->any_active && d->status && d->nfluid && d->nbdy && d->fbegin && d->bbegin &&
         d->act_scene && d->act_pre && d->act_n && d->nscenes > 0 && d->grid_blocks > 0;
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
static bool bdy_ok(const SphBoundary& bb, const SphDfsph* d) {
  return bb.nb == 0 || (bb.x && bb.vol && bb.rho && bb.rest && bb.bscene && d->b_aact &&
                        d->b_alpha && d->b_src && d->b_p && d->b_resid && d->bxrec);
}

int dfsph_prepare(const SphFluid* f, const float* acc, const SphPhysics* phys, const SphBoundary* b,
                  SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(f && f->x && f->v && f->area && f->rho && f->rest && f->pscene && phys &&
                  dfsph_ok(d),
              "sph_dfsph_prepare: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bdy_ok(bb, d), "sph_dfsph_prepare: bad boundary");
  SPH_REQUIRE(!d->divergence || bb.nb == 0,
              "sph_dfsph_prepare: the divergence solve takes no boundary");
  int n = f->np + bb.nb;
  const int nacc = SPH_DFSPH_ERR_SLOTS * SPH_DFSPH_ERR_STRIDE(d->nscenes);
  if (n < nacc) n = nacc;
  if (n < 1) n = 1;
  hipLaunchKernelGGL(k_prepare, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, acc, phys, bb, *d);
  hipLaunchKernelGGL(k_active_lists, dim3(1), dim3(kBS), 0, s, *d);
  return launch_check("sph_dfsph_prepare");
}

int dfsph_alpha_source(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                       hipStream_t s) {
  SPH_REQUIRE(f && phys && ff && ff->rowptr && dfsph_ok(d), "sph_dfsph_alpha_source: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bdy_ok(bb, d) && (bb.nb == 0 || (bf && bf->rowptr)),
              "sph_dfsph_alpha_source: boundary needs its pair list");
  SPH_REQUIRE(!d->divergence || (bb.nb == 0 && (!fb || !fb->rowptr)),
              "sph_dfsph_alpha_source: the divergence solve takes no boundary");
  const int n = f->np + bb.nb;
  if (n == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_SOLVE_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_alpha_source<KL>), dim3(blocks_for((int64_t)n * KL, kBS)), dim3(kBS), 0,
                       s, *f, phys, *ff, bb, csr_or_null(fb), csr_or_null(bf), *d);
  });
  return launch_check("sph_dfsph_alpha_source");
}

int dfsph_accel(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(f && phys && ff && ff->rowptr && dfsph_ok(d), "sph_dfsph_accel: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || d->bxrec, "sph_dfsph_accel: boundary list without records");
  SPH_REQUIRE(!d->divergence || !fb || !fb->rowptr,
              "sph_dfsph_accel: the divergence solve takes no boundary");
  with_lanes(tuning(SPH_TUNE_SOLVE_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_accel<KL>), dim3(d->grid_blocks), dim3(kBS), 0, s, *f, phys, *ff,
                       csr_or_null(fb), *d);
  });
  return launch_check("sph_dfsph_accel");
}

int dfsph_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                   const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                   hipStream_t s) {
  SPH_REQUIRE(f && f->rho && f->rest && phys && ff && ff->rowptr && dfsph_ok(d),
              "sph_dfsph_pressure: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bdy_ok(bb, d) && (bb.nb == 0 || (bf && bf->rowptr)),
              "sph_dfsph_pressure: bad boundary");
  SPH_REQUIRE(!d->divergence || (bb.nb == 0 && (!fb || !fb->rowptr)),
              "sph_dfsph_pressure: the divergence solve takes no boundary");
  with_lanes(tuning(SPH_TUNE_SOLVE_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_pressure<KL>), dim3(d->grid_blocks), dim3(kBS), 0, s, *f, phys, *ff,
                       bb, csr_or_null(fb), csr_or_null(bf), *d);
  });
  return launch_check("sph_dfsph_pressure");
}

int dfsph_control(const SphPhysics* phys, SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(phys && dfsph_ok(d), "sph_dfsph_control: bad arguments");
  hipLaunchKernelGGL(k_control, dim3(1), dim3(kBS), 0, s, phys, *d);
  return launch_check("sph_dfsph_control");
}

int dfsph_iterate(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                  const SphCSR* fb, const SphCSR* bf, SphDfsph* d, hipStream_t s) {
  int rc;
  if ((rc = dfsph_accel(f, phys, ff, fb, d, s))) return rc;
  if ((rc = dfsph_pressure(f, phys, ff, b, fb, bf, d, s))) return rc;
  return dfsph_control(phys, d, s);
}

int dfsph_finish(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                 SphDfsph* d, float* acc, hipStream_t s) {
  SPH_REQUIRE(f && f->pscene && phys && ff && ff->rowptr && dfsph_ok(d),
              "sph_dfsph_finish: bad arguments");
  SPH_REQUIRE(!d->divergence || !fb || !fb->rowptr,
              "sph_dfsph_finish: the divergence solve takes no boundary");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_SOLVE_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_finish<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)), dim3(kBS), 0, s,
                       *f, phys, *ff, csr_or_null(fb), *d, acc);
  });
  return launch_check("sph_dfsph_finish");
}

int dfsph_finish_visc(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphCSR* fb, SphDfsph* d, const float* frec, unsigned vterms,
                      float* acc, hipStream_t s) {
  SPH_REQUIRE(f && f->pscene && phys && ff && ff->rowptr && dfsph_ok(d) && frec && acc,
              "sph_dfsph_finish_visc: bad arguments");
  SPH_REQUIRE(!d->divergence, "sph_dfsph_finish_visc: not for the divergence solve");
  SPH_REQUIRE((vterms & ~(unsigned)(SPH_T_VISC | SPH_T_LAM)) == 0,
              "sph_dfsph_finish_visc: only SPH_T_VISC and SPH_T_LAM");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_SOLVE_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_finish_visc<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)), dim3(kBS),
                       0, s, *f, phys, *ff, csr_or_null(fb), *d,
                       reinterpret_cast<const float4*>(frec), vterms, acc);
  });
  return launch_check("sph_dfsph_finish_visc");
}

// ---------------------------------------------- stencil-list entry points
static bool sl_ok(const SphStencilList* sl, int np) {
  return sl && sl->n == np && (np == 0 || (sl->runs && sl->masks));
}

int fluid_density_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                     const SphBoundary* b, const SphCSR* fb, float* rho_out, float* vol_out,
                     hipStream_t s) {
  SPH_REQUIRE(f && f->x && f->area && f->pscene && f->np >= 0 && phys && rho_out &&
                  sl_ok(sl, f->np),
              "sph_fluid_density_sl: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || (b && b->x && b->vol),
              "sph_fluid_density_sl: boundary list without boundary");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_fluid_density_sl, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys,
                     *sl, bdy_or_null(b), csr_or_null(fb), rho_out, vol_out);
  return launch_check("sph_fluid_density_sl");
}

int dfsph_alpha_source_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                          const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                          hipStream_t s) {
  SPH_REQUIRE(f && f->x && f->area && f->rest && f->rho && f->pscene && phys && dfsph_ok(d) &&
                  sl_ok(sl, f->np),
              "sph_dfsph_alpha_source_sl: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bdy_ok(bb, d) && (bb.nb == 0 || (bf && bf->rowptr)),
              "sph_dfsph_alpha_source_sl: boundary needs its pair list");
  SPH_REQUIRE(!d->divergence || (bb.nb == 0 && (!fb || !fb->rowptr)),
              "sph_dfsph_alpha_source_sl: the divergence solve takes no boundary");
  const int n = f->np + bb.nb;
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_alpha_source_sl, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, phys, *sl,
                     bb, csr_or_null(fb), csr_or_null(bf), *d);
  return launch_check("sph_dfsph_alpha_source_sl");
}

int dfsph_accel_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                   const SphCSR* fb, SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(f && phys && dfsph_ok(d) && sl_ok(sl, f->np), "sph_dfsph_accel_sl: bad arguments");
  SPH_REQUIRE(!fb || !fb->rowptr || d->bxrec, "sph_dfsph_accel_sl: boundary list without records");
  SPH_REQUIRE(!d->divergence || !fb || !fb->rowptr,
              "sph_dfsph_accel_sl: the divergence solve takes no boundary");
  hipLaunchKernelGGL(k_accel_sl, dim3(d->grid_blocks), dim3(kBS), 0, s, *f, phys, *sl,
                     csr_or_null(fb), *d);
  return launch_check("sph_dfsph_accel_sl");
}

int dfsph_pressure_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                      const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                      hipStream_t s) {
  SPH_REQUIRE(f && f->rho && f->rest && phys && dfsph_ok(d) && sl_ok(sl, f->np),
              "sph_dfsph_pressure_sl: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bdy_ok(bb, d) && (bb.nb == 0 || (bf && bf->rowptr)),
              "sph_dfsph_pressure_sl: bad boundary");
  SPH_REQUIRE(!d->divergence || (bb.nb == 0 && (!fb || !fb->rowptr)),
              "sph_dfsph_pressure_sl: the divergence solve takes no boundary");
  hipLaunchKernelGGL(k_pressure_sl, dim3(d->grid_blocks), dim3(kBS), 0, s, *f, phys, *sl, bb,
                     csr_or_null(fb), csr_or_null(bf), *d);
  return launch_check("sph_dfsph_pressure_sl");
}

int dfsph_iterate_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                     const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                     hipStream_t s) {
  int rc;
  if ((rc = dfsph_accel_sl(f, phys, sl, fb, d, s))) return rc;
  if ((rc = dfsph_pressure_sl(f, phys, sl, b, fb, bf, d, s))) return rc;
  return dfsph_control(phys, d, s);
}

int dfsph_finish_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                    const SphCSR* fb, SphDfsph* d, float* acc, hipStream_t s) {
  SPH_REQUIRE(f && f->pscene && phys && dfsph_ok(d) && sl_ok(sl, f->np),
              "sph_dfsph_finish_sl: bad arguments");
  SPH_REQUIRE(!d->divergence || !fb || !fb->rowptr,
              "sph_dfsph_finish_sl: the divergence solve takes no boundary");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_finish_sl, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys, *sl,
                     csr_or_null(fb), *d, acc);
  return launch_check("sph_dfsph_finish_sl");
}

int dfsph_finish_visc_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                         const SphCSR* fb, SphDfsph* d, const float* frec, unsigned vterms,
                         float* acc, hipStream_t s) {
  SPH_REQUIRE(f && f->pscene && phys && dfsph_ok(d) && frec && acc && sl_ok(sl, f->np),
              "sph_dfsph_finish_visc_sl: bad arguments");
  SPH_REQUIRE(!d->divergence, "sph_dfsph_finish_visc_sl: not for the divergence solve");
  SPH_REQUIRE((vterms & ~(unsigned)(SPH_T_VISC | SPH_T_LAM)) == 0,
              "sph_dfsph_finish_visc_sl: only SPH_T_VISC and SPH_T_LAM");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_finish_visc_sl, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys,
                     *sl, csr_or_null(fb), *d, reinterpret_cast<const float4*>(frec), vterms, acc);
  return launch_check("sph_dfsph_finish_visc_sl");
}

}  // namespace sph

extern "C" {
int sph_dfsph_prepare(const SphFluid* f, const float* acc, const SphPhysics* phys,
                      const SphBoundary* b, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_prepare(f, acc, phys, b, d, s);
}
int sph_dfsph_alpha_source(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                           const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                           hipStream_t s) {
  return sph::dfsph_alpha_source(f, phys, ff, b, fb, bf, d, s);
}
int sph_dfsph_iterate(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                      hipStream_t s) {
  return sph::dfsph_iterate(f, phys, ff, b, fb, bf, d, s);
}
int sph_dfsph_accel(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                    const SphCSR* fb, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_accel(f, phys, ff, fb, d, s);
}
int sph_dfsph_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                       hipStream_t s) {
  return sph::dfsph_pressure(f, phys, ff, b, fb, bf, d, s);
}
int sph_dfsph_control(const SphPhysics* phys, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_control(phys, d, s);
}
int sph_dfsph_finish(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                     const SphCSR* fb, SphDfsph* d, float* acc, hipStream_t s) {
  return sph::dfsph_finish(f, phys, ff, fb, d, acc, s);
}
int sph_dfsph_finish_visc(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                          const SphCSR* fb, SphDfsph* d, const float* frec, unsigned vterms,
                          float* acc, hipStream_t s) {
  return sph::dfsph_finish_visc(f, phys, ff, fb, d, frec, vterms, acc, s);
}
int sph_fluid_density_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                         const SphBoundary* b, const SphCSR* fb, float* rho_out, float* vol_out,
                         hipStream_t s) {
  return sph::fluid_density_sl(f, phys, sl, b, fb, rho_out, vol_out, s);
}
int sph_dfsph_alpha_source_sl(const SphFluid* f, const SphPhysics* phys,
                              const SphStencilList* sl, const SphBoundary* b, const SphCSR* fb,
                              const SphCSR* bf, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_alpha_source_sl(f, phys, sl, b, fb, bf, d, s);
}
int sph_dfsph_accel_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                       const SphCSR* fb, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_accel_sl(f, phys, sl, fb, d, s);
}
int sph_dfsph_pressure_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                          const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                          hipStream_t s) {
  return sph::dfsph_pressure_sl(f, phys, sl, b, fb, bf, d, s);
}
int sph_dfsph_iterate_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                         const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                         hipStream_t s) {
  return sph::dfsph_iterate_sl(f, phys, sl, b, fb, bf, d, s);
}
int sph_dfsph_finish_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                        const SphCSR* fb, SphDfsph* d, float* acc, hipStream_t s) {
  return sph::dfsph_finish_sl(f, phys, sl, fb, d, acc, s);
}
int sph_dfsph_finish_visc_sl(const SphFluid* f, const SphPhysics* phys,
                             const SphStencilList* sl, const SphCSR* fb, SphDfsph* d,
                             const float* frec, unsigned vterms, float* acc, hipStream_t s) {
  return sph::dfsph_finish_visc_sl(f, phys, sl, fb, d, frec, vterms, acc, s);
}
}
