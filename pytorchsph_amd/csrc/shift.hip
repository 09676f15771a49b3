// delta-plus particle shifting: free-surface detection through the smaller
// singular value lambda of the renormalisation matrix, the fluid normal from
// grad lambda, the shift vector and its surface projection
// (modules/deltaPlus.py:81-251), and the delta-plus pressure switch
// (pressure.py:40-52, solidBoundary.py:882-887).
//
// Each row is walked by a group of lanes (the gather lanes knob) with
// several neighbour loads in flight per lane, as in the delta-SPH gathers.
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;

// deltaPlus.py:86-92 and 215-216
constexpr float kShiftCFL = 1.5f;
constexpr float kShiftR = 0.2f;
constexpr float kShiftClamp = 0.05f;
// deltaPlus.py:195, 204: thresholds compared in fp32
constexpr float kPi6 = 0.52359877559829887f;
constexpr float kLamSurface = 0.6f;
constexpr float kLamKeep = 0.5f;
constexpr float kTwoThirds = 0.66666666666666663f;

__device__ __forceinline__ float2 ld2(const float* p, int i) {
  return reinterpret_cast<const float2*>(p)[i];
}
__device__ __forceinline__ void st2(float* p, int i, float2 v) {
  reinterpret_cast<float2*>(p)[i] = v;
}

__device__ __forceinline__ float fluid_vol(const SphFluid& f, int j) {
  return f.vol ? f.vol[j] : f.area[j] / f.rho[j];
}

struct XV {
  float x, y, v;
};
struct XVL {
  float x, y, v, l;
};

// M_q over two lists (densityDiffusion.py:30-43), L = pinv2x2(M).
template <int KL>
__global__ __launch_bounds__(kBS) void k_norm_rows(const float* __restrict__ xq,
                                                   const int* __restrict__ qscene, int nq,
                                                   const SphPhysics* __restrict__ phys, SphCSR a,
                                                   const float* __restrict__ xa,
                                                   const float* __restrict__ va, SphCSR b,
                                                   const float* __restrict__ xb,
                                                   const float* __restrict__ vb,
                                                   float* __restrict__ Mout,
                                                   float* __restrict__ Lout,
                                                   float* __restrict__ eig) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= nq) return;
  const int q = rl.t;
  const KConst K = kconst(phys[qscene[q]].h);
  const float2 xi = ld2(xq, q);
  float m[4] = {0.f, 0.f, 0.f, 0.f};
  for (int pass = 0; pass < 2; ++pass) {
    const SphCSR& c = pass == 0 ? a : b;
    if (!c.rowptr) continue;
    const float* __restrict__ xs = pass == 0 ? xa : xb;
    const float* __restrict__ vs = pass == 0 ? va : vb;
    float s00 = 0.f, s01 = 0.f, s10 = 0.f, s11 = 0.f;
    row_pairs<KL>(
        c.col, c.rowptr[q], c.rowptr[q + 1], rl.sub,
        [&](int j) {
          const float2 xj = ld2(xs, j);
          return XV{xj.x, xj.y, vs[j]};
        },
        [&](int, const XV& r) {
          const PairGeo g = pair_geo(xi.x, xi.y, r.x, r.y, K.inv_h);
          const float gw = -wendland_grad_k(g.q, K);
          const float gx = gw * g.hx, gy = gw * g.hy;
          const float rx = -g.rx, ry = -g.ry;
          s00 += rx * gx * r.v;
          s01 += rx * gy * r.v;
          s10 += ry * gx * r.v;
          s11 += ry * gy * r.v;
        });
    m[0] += group_sum<KL>(s00);
    m[1] += group_sum<KL>(s01);
    m[2] += group_sum<KL>(s10);
    m[3] += group_sum<KL>(s11);
  }
  if (rl.sub != 0) return;
  if (Mout) *reinterpret_cast<float4*>(Mout + 4 * q) = make_float4(m[0], m[1], m[2], m[3]);
  float l[4], e[2];
  pinv2x2_dev(m[0], m[1], m[2], m[3], l, e);
  *reinterpret_cast<float4*>(Lout + 4 * q) = make_float4(l[0], l[1], l[2], l[3]);
  if (eig) st2(eig, q, make_float2(e[0], e[1]));
}

// g_i = -sum V_j (lam_j - lam_i) L_i gradW_ij (deltaPlus.py:132-161), summed
// by the KL lanes of row i's group; every lane of the group gets the sum.
template <int KL>
__device__ __forceinline__ float2 lambda_grad(const SphFluid& f, const KConst& K, const SphCSR& ff,
                                              const float* __restrict__ Lm,
                                              const float* __restrict__ eig,
                                              const float* __restrict__ bx,
                                              const float* __restrict__ beig,
                                              const float* __restrict__ bvol_n, const SphCSR& fb,
                                              int i, int sub, float2 xi) {
  const float lam_i = eig[2 * i + 1];
  const float4 L4 = *reinterpret_cast<const float4*>(Lm + 4 * i);
  auto pair = [&](float& sx, float& sy, const XVL& r) {
    const PairGeo pg = pair_geo(xi.x, xi.y, r.x, r.y, K.inv_h);
    const float gw = -wendland_grad_k(pg.q, K);
    const float gx = gw * pg.hx, gy = gw * pg.hy;
    const float cx = L4.x * gx + L4.y * gy, cy = L4.z * gx + L4.w * gy;
    const float w = -(r.v * (r.l - lam_i));
    sx += w * cx;
    sy += w * cy;
  };
  float fx = 0.f, fy = 0.f;
  row_pairs<KL>(
      ff.col, ff.rowptr[i], ff.rowptr[i + 1], sub,
      [&](int j) {
        const float2 xj = ld2(f.x, j);
        return XVL{xj.x, xj.y, fluid_vol(f, j), eig[2 * j + 1]};
      },
      [&](int, const XVL& r) { pair(fx, fy, r); });
  fx = group_sum<KL>(fx);
  fy = group_sum<KL>(fy);
  float bxs = 0.f, bys = 0.f;
  if (fb.rowptr) {
    for (int k = fb.rowptr[i] + sub; k < fb.rowptr[i + 1]; k += KL) {
      const int j = fb.col[k];
      const float2 xj = ld2(bx, j);
      pair(bxs, bys, XVL{xj.x, xj.y, bvol_n[j], beig[2 * j + 1]});
    }
    bxs = group_sum<KL>(bxs);
    bys = group_sum<KL>(bys);
  }
  return make_float2(fx + bxs, fy + bys);
}

// The angle test of detectFluidSurface (deltaPlus.py:165-195) for row i with
// normal n: min / max of (-rhat_ij) . n over the non-self fluid pairs, max
// of rhat_bi . n over the boundary pairs, an empty max or min being 0
// (torch_scatter's fill).  Lane 0 of the group writes the mask.
template <int KL>
__device__ __forceinline__ void angle_test(const SphFluid& f, const KConst& K, const SphCSR& ff,
                                           const float* __restrict__ bx, const SphCSR& fb, int i,
                                           int sub, float2 xi, float2 n, float lam_i,
                                           float* __restrict__ mask, float* __restrict__ amin,
                                           float* __restrict__ amax) {
  float dmin = INFINITY, dmax = -INFINITY;
  row_pairs<KL>(
      ff.col, ff.rowptr[i], ff.rowptr[i + 1], sub, [&](int j) { return ld2(f.x, j); },
      [&](int j, const float2& xj) {
        if (j == i) return;
        const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, K.inv_h);
        const float d = -g.hx * n.x + -g.hy * n.y;   // (-rhat_ij) . n_i
        dmin = fminf(dmin, d);
        dmax = fmaxf(dmax, d);
      });
  dmin = group_min<KL>(dmin);
  dmax = group_max<KL>(dmax);
  float bmax = -INFINITY;
  if (fb.rowptr) {
    for (int k = fb.rowptr[i] + sub; k < fb.rowptr[i + 1]; k += KL) {
      const float2 xb = ld2(bx, fb.col[k]);
      const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, K.inv_h);
      bmax = fmaxf(bmax, -g.hx * n.x + -g.hy * n.y);   // rhat_bi . n_i
    }
    bmax = group_max<KL>(bmax);
  }
  if (sub != 0) return;
  if (dmax == -INFINITY) {   // no non-self pair
    dmin = 0.f;
    dmax = 0.f;
  }
  if (bmax == -INFINITY) bmax = 0.f;
  const float ang = acosf(fmaxf(dmax, bmax));
  mask[i] = (ang > kPi6 && lam_i < kLamSurface) ? 1.f : 0.f;
  if (amin) amin[i] = acosf(dmin);
  if (amax) amax[i] = acosf(dmax);
}

// n_i = g_i / (|g_i| + eps) (deltaPlus.py:163).
__device__ __forceinline__ float2 normal_of(float2 g, float eps) {
  const float inv = 1.f / (sqrtf(g.x * g.x + g.y * g.y) + eps);
  return make_float2(g.x * inv, g.y * inv);
}

template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_normal(SphFluid f, const SphPhysics* __restrict__ phys,
                                                      SphCSR ff, const float* __restrict__ Lm,
                                                      const float* __restrict__ eig,
                                                      const float* __restrict__ bx,
                                                      const float* __restrict__ beig,
                                                      const float* __restrict__ bvol_n, SphCSR fb,
                                                      float* __restrict__ grad,
                                                      float* __restrict__ normal) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = rl.t;
  const SphPhysics& P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const float2 xi = ld2(f.x, i);
  const float2 g = lambda_grad<KL>(f, K, ff, Lm, eig, bx, beig, bvol_n, fb, i, rl.sub, xi);
  if (rl.sub != 0) return;
  if (grad) st2(grad, i, g);
  st2(normal, i, normal_of(g, P.eps));
}

template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_angle(SphFluid f, const SphPhysics* __restrict__ phys,
                                                     SphCSR ff, const float* __restrict__ normal,
                                                     const float* __restrict__ eig,
                                                     const float* __restrict__ bx, SphCSR fb,
                                                     float* __restrict__ mask,
                                                     float* __restrict__ amin,
                                                     float* __restrict__ amax) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = rl.t;
  const KConst K = kconst(phys[f.pscene[i]].h);
  angle_test<KL>(f, K, ff, bx, fb, i, rl.sub, ld2(f.x, i), ld2(normal, i), eig[2 * i + 1], mask,
                 amin, amax);
}

// The normal and the angle test in one pass over the row: the group holds
// n_i after the lambda-gradient sum and walks the same pairs again, their
// positions now cached.
template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_normal_angle(
    SphFluid f, const SphPhysics* __restrict__ phys, SphCSR ff, const float* __restrict__ Lm,
    const float* __restrict__ eig, const float* __restrict__ bx, const float* __restrict__ beig,
    const float* __restrict__ bvol_n, SphCSR fb, float* __restrict__ grad,
    float* __restrict__ normal, float* __restrict__ mask, float* __restrict__ amin,
    float* __restrict__ amax) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = rl.t;
  const SphPhysics& P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const float2 xi = ld2(f.x, i);
  const float2 g = lambda_grad<KL>(f, K, ff, Lm, eig, bx, beig, bvol_n, fb, i, rl.sub, xi);
  const float2 n = normal_of(g, P.eps);
  if (rl.sub == 0) {
    if (grad) st2(grad, i, g);
    st2(normal, i, n);
  }
  angle_test<KL>(f, K, ff, bx, fb, i, rl.sub, xi, n, eig[2 * i + 1], mask, amin, amax);
}

// Second pass (deltaPlus.py:197-209): one-ring spread of the mask and the
// final classes 0 / 2/3 / 1 / 2.
template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_classify(int np, SphCSR ff,
                                                        const float* __restrict__ mask,
                                                        const float* __restrict__ eig,
                                                        float* __restrict__ out) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= np) return;
  const int i = rl.t;
  float m2 = 0.f;
  row_pairs<KL>(
      ff.col, ff.rowptr[i], ff.rowptr[i + 1], rl.sub, [&](int j) { return mask[j]; },
      [&](int j, float mj) {
        if (j != i) m2 = fmaxf(m2, mj);
      });
  m2 = group_max<KL>(m2);
  if (rl.sub != 0) return;
  float fm = 0.f;
  if (m2 > 0.f) fm = kTwoThirds;
  if (mask[i] > 0.f) fm = 1.f;
  float r = fm;
  if (eig[2 * i + 1] >= kLamKeep && fm > 0.f) r = 1.f;
  if (fm > 0.7f) r = 0.f;
  if (fm < 0.1f) r = 2.f;
  out[i] = r;
}

__global__ void k_clear_u32(unsigned* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// max_k |v_k| per scene; non-negative floats order like their bit patterns.
// A block takes kSpeedRows consecutive rows (coalesced, kBS apart); a lane
// keeps a running maximum per scene and flushes it with one atomic when its
// rows change scene; at the end a wavefront whose lanes share one scene
// reduces with shuffles and issues one atomic.  Same-address atomics
// serialise, so the count per scene is kept near n / kSpeedRows * 4.
constexpr int kSpeedIters = 32;
constexpr int kSpeedRows = kBS * kSpeedIters;
__global__ __launch_bounds__(kBS) void k_speed_max(SphFluid f, unsigned* __restrict__ vmax) {
  const int64_t base = (int64_t)blockIdx.x * kSpeedRows + threadIdx.x;
  int sc = -1;
  float m = 0.f;
  for (int it = 0; it < kSpeedIters; ++it) {
    const int64_t i = base + (int64_t)it * kBS;
    if (i >= f.np) break;
    const float2 v = ld2(f.v, (int)i);
    const float s = sqrtf(v.x * v.x + v.y * v.y);
    const int si = f.pscene[i];
    if (si != sc) {
      if (sc >= 0 && m > 0.f) atomicMax(vmax + sc, __float_as_uint(m));
      sc = si;
      m = 0.f;
    }
    m = fmaxf(m, s);
  }
  const int sc0 = __shfl(sc, 0);
  if (__all(sc == sc0 || sc < 0)) {
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && sc0 >= 0 && m > 0.f) atomicMax(vmax + sc0, __float_as_uint(m));
  } else if (sc >= 0 && m > 0.f) {
    atomicMax(vmax + sc, __float_as_uint(m));
  }
}

struct XAR {
  float x, y, a, r;
};

// computeShiftAmount + adjustShiftingAmount (deltaPlus.py:81-115, 211-236).
template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_shift(SphFluid f, const SphPhysics* __restrict__ phys,
                                                     const float* __restrict__ packing, SphCSR ff,
                                                     SphBoundary b, SphCSR fb,
                                                     const float* __restrict__ normal,
                                                     const float* __restrict__ smask,
                                                     const unsigned* __restrict__ vmax,
                                                     float* __restrict__ shift) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = rl.t;
  const int sc = f.pscene[i];
  const SphPhysics& P = phys[sc];
  const KConst K = kconst(P.h);
  const float inv_k0 = 1.f / wendland(packing[sc], P.h);
  const float2 xi = ld2(f.x, i);
  const float ri = f.rho[i];
  auto pair = [&](float& sx, float& sy, const XAR& r) {
    const float mass = r.a / (ri + r.r);
    const PairGeo g = pair_geo(xi.x, xi.y, r.x, r.y, K.inv_h);
    const float w = wendland_k(g.q, K) * inv_k0;
    const float w2 = w * w;
    const float kt = 1.f + kShiftR * (w2 * w2);
    const float gw = -wendland_grad_k(g.q, K);
    const float t = kt * mass;
    sx += t * (gw * g.hx);
    sy += t * (gw * g.hy);
  };
  float fx = 0.f, fy = 0.f;
  row_pairs<KL>(
      ff.col, ff.rowptr[i], ff.rowptr[i + 1], rl.sub,
      [&](int j) {
        const float2 xj = ld2(f.x, j);
        return XAR{xj.x, xj.y, f.area[j], f.rho[j]};
      },
      [&](int, const XAR& r) { pair(fx, fy, r); });
  fx = group_sum<KL>(fx);
  fy = group_sum<KL>(fy);
  float bxs = 0.f, bys = 0.f;
  if (fb.rowptr) {
    for (int k = fb.rowptr[i] + rl.sub; k < fb.rowptr[i + 1]; k += KL) {
      const int j = fb.col[k];
      const float2 xj = ld2(b.x, j);
      pair(bxs, bys, XAR{xj.x, xj.y, b.vol[j], b.rho[j]});
    }
    bxs = group_sum<KL>(bxs);
    bys = group_sum<KL>(bys);
  }
  if (rl.sub != 0) return;
  const float Ma = __uint_as_float(vmax[sc]) / P.c0;
  const float c = -kShiftCFL * Ma * (4.f * P.h * P.h);
  float sx = c * fx + c * bxs;
  float sy = c * fy + c * bys;
  const float len = sqrtf(sx * sx + sy * sy);
  const float cap = P.h * kShiftClamp;
  if (len > cap) {
    sx = sx / len * cap;
    sy = sy / len * cap;
  }
  const float m = smask[i];
  if (m < 0.5f) {
    sx = 0.f;
    sy = 0.f;
  } else if (m < 1.5f) {
    const float2 n = ld2(normal, i);
    const float a00 = 1.f - n.x * n.x, a01 = -(n.x * n.y), a10 = -(n.y * n.x), a11 = 1.f - n.y * n.y;
    const float px = a00 * sx + a01 * sy, py = a10 * sx + a11 * sy;
    sx = px;
    sy = py;
  }
  st2(shift, i, make_float2(sx, sy));
}

// computePressureAccelDeltaPlus (pressure.py:40-52) for fluid and boundary
// neighbours.
template <int KL>
__global__ __launch_bounds__(kBS) void k_dplus_pressure(SphFluid f, const SphPhysics* __restrict__ phys,
                                                        SphCSR ff, SphBoundary b, SphCSR fb,
                                                        const float* __restrict__ smask,
                                                        unsigned terms, float* __restrict__ acc) {
  const RowLane<KL> rl = row_lane<KL>();
  if (rl.t >= f.np) return;
  const int i = rl.t;
  const SphPhysics& P = phys[f.pscene[i]];
  const KConst K = kconst(P.h);
  const float2 xi = ld2(f.x, i);
  const EosK E = eos_k(P);
  const float pi = eos_eval(E, f.rho[i]);
  const bool sum = pi >= 0.f || smask[i] < 1.5f;
  auto pair = [&](float& sx, float& sy, float xj, float yj, float pj, float V) {
    const float pij = sum ? pi + pj : pj - pi;
    const PairGeo g = pair_geo(xi.x, xi.y, xj, yj, K.inv_h);
    const float gw = -wendland_grad_k(g.q, K);
    const float t = pij * V;
    sx += t * (gw * g.hx);
    sy += t * (gw * g.hy);
  };
  float fx = 0.f, fy = 0.f, bxs = 0.f, bys = 0.f;
  if (ff.rowptr && (terms & SPH_T_PRESS)) {
    row_pairs<KL>(
        ff.col, ff.rowptr[i], ff.rowptr[i + 1], rl.sub,
        [&](int j) {
          const float2 xj = ld2(f.x, j);
          const float rj = f.rho[j];
          return XAR{xj.x, xj.y, f.vol ? f.vol[j] : f.area[j] / rj, rj};
        },
        [&](int, const XAR& r) { pair(fx, fy, r.x, r.y, eos_eval(E, r.r), r.a); });
    fx = group_sum<KL>(fx);
    fy = group_sum<KL>(fy);
  }
  if (fb.rowptr && (terms & SPH_T_PRESS_B)) {
    for (int k = fb.rowptr[i] + rl.sub; k < fb.rowptr[i + 1]; k += KL) {
      const int j = fb.col[k];
      const float2 xj = ld2(b.x, j);
      pair(bxs, bys, xj.x, xj.y, b.p[j], b.vol[j] / b.rho[j]);
    }
    bxs = group_sum<KL>(bxs);
    bys = group_sum<KL>(bys);
  }
  if (rl.sub != 0) return;
  const float inv_rho = 1.f / (f.rho[i] * P.rho0);
  float ax = -inv_rho * fx + -inv_rho * bxs;
  float ay = -inv_rho * fy + -inv_rho * bys;
  if (terms & SPH_T_ACCUM) {
    const float2 o = ld2(acc, i);
    ax += o.x;
    ay += o.y;
  }
  st2(acc, i, make_float2(ax, ay));
}

bool csr_ok(const SphCSR* c) { return c && c->rowptr && c->col; }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

bool fluid_ok(const SphFluid* f) {
  return f && f->np >= 0 && f->x && f->pscene && (f->np == 0 || (f->rho && f->area));
}

}  // namespace

int normalization_rows(const float* xq, const int* qscene, int nq, const SphPhysics* phys,
                       const SphCSR* a, const float* xa, const float* va, const SphCSR* b,
                       const float* xb, const float* vb, float* M, float* L, float* eig,
                       hipStream_t s) {
  SPH_REQUIRE(nq >= 0 && phys && L && (nq == 0 || (xq && qscene)),
              "sph_normalization_rows: bad arguments");
  SPH_REQUIRE(aligned16(L) && aligned16(M), "sph_normalization_rows: M and L need 16-byte alignment");
  const SphCSR none = {nullptr, nullptr, 0};
  const bool ua = a && a->rowptr, ub = b && b->rowptr;
  SPH_REQUIRE(!ua || (a->col && xa && va), "sph_normalization_rows: list a without its points");
  SPH_REQUIRE(!ub || (b->col && xb && vb), "sph_normalization_rows: list b without its points");
  if (nq == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_norm_rows<KL>), dim3(blocks_for((int64_t)nq * KL, kBS)), dim3(kBS), 0,
                       s, xq, qscene, nq, phys, ua ? *a : none, xa, va, ub ? *b : none, xb, vb,
                       M, L, eig);
  });
  return launch_check("sph_normalization_rows");
}

int dplus_normal(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const float* L,
                 const float* eig, const float* bx, const float* beig, const float* bvol_n,
                 const SphCSR* fb, float* grad, float* normal, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && csr_ok(ff) && L && eig && normal,
              "sph_dplus_normal: bad arguments");
  SPH_REQUIRE(aligned16(L), "sph_dplus_normal: L needs 16-byte alignment");
  const SphCSR none = {nullptr, nullptr, 0};
  const bool ub = fb && fb->rowptr;
  SPH_REQUIRE(!ub || (fb->col && bx && beig && bvol_n),
              "sph_dplus_normal: boundary list without boundary points, lambda or volumes");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_dplus_normal<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)),
                       dim3(kBS), 0, s, *f, phys, *ff, L, eig, bx, beig, bvol_n,
                       ub ? *fb : none, grad, normal);
  });
  return launch_check("sph_dplus_normal");
}

int dplus_surface(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                  const float* normal, const float* eig, const float* bx, const SphCSR* fb,
                  float* mask_tmp, float* amin, float* amax, float* surface_mask,
                  hipStream_t s) {
  SPH_REQUIRE(f && f->np >= 0 && f->x && f->pscene && phys && csr_ok(ff) && normal && eig &&
                  mask_tmp && surface_mask,
              "sph_dplus_surface: bad arguments");
  const SphCSR none = {nullptr, nullptr, 0};
  const bool ub = fb && fb->rowptr;
  SPH_REQUIRE(!ub || (fb->col && bx), "sph_dplus_surface: boundary list without boundary points");
  SPH_REQUIRE(mask_tmp != surface_mask, "sph_dplus_surface: mask_tmp must not alias the output");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int nb = blocks_for((int64_t)f->np * KL, kBS);
    hipLaunchKernelGGL((k_dplus_angle<KL>), dim3(nb), dim3(kBS), 0, s, *f, phys, *ff, normal, eig,
                       bx, ub ? *fb : none, mask_tmp, amin, amax);
    hipLaunchKernelGGL((k_dplus_classify<KL>), dim3(nb), dim3(kBS), 0, s, f->np, *ff, mask_tmp,
                       eig, surface_mask);
  });
  return launch_check("sph_dplus_surface");
}

int dplus_detect(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const float* L,
                 const float* eig, const float* bx, const float* beig, const float* bvol_n,
                 const SphCSR* fb, float* grad, float* normal, float* mask_tmp, float* amin,
                 float* amax, float* surface_mask, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && csr_ok(ff) && L && eig && normal && mask_tmp &&
                  surface_mask,
              "sph_dplus_detect: bad arguments");
  SPH_REQUIRE(aligned16(L), "sph_dplus_detect: L needs 16-byte alignment");
  SPH_REQUIRE(mask_tmp != surface_mask, "sph_dplus_detect: mask_tmp must not alias the output");
  const SphCSR none = {nullptr, nullptr, 0};
  const bool ub = fb && fb->rowptr;
  SPH_REQUIRE(!ub || (fb->col && bx && beig && bvol_n),
              "sph_dplus_detect: boundary list without boundary points, lambda or volumes");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int nb = blocks_for((int64_t)f->np * KL, kBS);
    hipLaunchKernelGGL((k_dplus_normal_angle<KL>), dim3(nb), dim3(kBS), 0, s, *f, phys, *ff, L,
                       eig, bx, beig, bvol_n, ub ? *fb : none, grad, normal, mask_tmp, amin,
                       amax);
    hipLaunchKernelGGL((k_dplus_classify<KL>), dim3(nb), dim3(kBS), 0, s, f->np, *ff, mask_tmp,
                       eig, surface_mask);
  });
  return launch_check("sph_dplus_detect");
}

int dplus_shift(const SphFluid* f, const SphPhysics* phys, int nscenes, const float* packing,
                const SphCSR* ff, const SphBoundary* b, const SphCSR* fb, const float* normal,
                const float* surface_mask, unsigned* vmax_ws, float* shift, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && f->v && phys && nscenes >= 1 && packing && csr_ok(ff) && normal &&
                  surface_mask && vmax_ws && shift,
              "sph_dplus_shift: bad arguments");
  const SphCSR none = {nullptr, nullptr, 0};
  const SphBoundary nob = {};
  const bool ub = fb && fb->rowptr;
  SPH_REQUIRE(!ub || (fb->col && b && b->x && b->vol && b->rho),
              "sph_dplus_shift: boundary list without boundary state");
  hipLaunchKernelGGL(k_clear_u32, dim3(blocks_for(nscenes, kBS)), dim3(kBS), 0, s, vmax_ws,
                     nscenes);
  if (f->np == 0) return launch_check("sph_dplus_shift");
  hipLaunchKernelGGL(k_speed_max, dim3(blocks_for(f->np, kSpeedRows)), dim3(kBS), 0, s, *f,
                     vmax_ws);
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_dplus_shift<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)),
                       dim3(kBS), 0, s, *f, phys, packing, *ff, ub ? *b : nob, ub ? *fb : none,
                       normal, surface_mask, vmax_ws, shift);
  });
  return launch_check("sph_dplus_shift");
}

int dplus_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                   const SphBoundary* b, const SphCSR* fb, const float* surface_mask,
                   unsigned terms, float* acc, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && surface_mask && acc, "sph_dplus_pressure: bad arguments");
  const SphCSR none = {nullptr, nullptr, 0};
  const SphBoundary nob = {};
  const bool uf = (terms & SPH_T_PRESS) != 0, ub = (terms & SPH_T_PRESS_B) != 0;
  SPH_REQUIRE(!uf || csr_ok(ff), "sph_dplus_pressure: SPH_T_PRESS needs the fluid list");
  SPH_REQUIRE(!ub || (csr_ok(fb) && b && b->x && b->vol && b->rho && b->p),
              "sph_dplus_pressure: SPH_T_PRESS_B needs the boundary list and state");
  if (f->np == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_GATHER_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    hipLaunchKernelGGL((k_dplus_pressure<KL>), dim3(blocks_for((int64_t)f->np * KL, kBS)),
                       dim3(kBS), 0, s, *f, phys, uf ? *ff : none, ub ? *b : nob,
                       ub ? *fb : none, surface_mask, terms, acc);
  });
  return launch_check("sph_dplus_pressure");
}

}  // namespace sph

extern "C" {
int sph_normalization_rows(const float* xq, const int* qscene, int nq, const SphPhysics* phys,
                           const SphCSR* a, const float* xa, const float* va, const SphCSR* b,
                           const float* xb, const float* vb, float* M, float* L, float* eig,
                           hipStream_t s) {
  return sph::normalization_rows(xq, qscene, nq, phys, a, xa, va, b, xb, vb, M, L, eig, s);
}
int sph_dplus_normal(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const float* L,
                     const float* eig, const float* bx, const float* beig, const float* bvol_n,
                     const SphCSR* fb, float* grad, float* normal, hipStream_t s) {
  return sph::dplus_normal(f, phys, ff, L, eig, bx, beig, bvol_n, fb, grad, normal, s);
}
int sph_dplus_surface(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const float* normal, const float* eig, const float* bx, const SphCSR* fb,
                      float* mask_tmp, float* angle_min, float* angle_max, float* surface_mask,
                      hipStream_t s) {
  return sph::dplus_surface(f, phys, ff, normal, eig, bx, fb, mask_tmp, angle_min, angle_max,
                            surface_mask, s);
}
int sph_dplus_detect(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const float* L,
                     const float* eig, const float* bx, const float* beig, const float* bvol_n,
                     const SphCSR* fb, float* grad, float* normal, float* mask_tmp,
                     float* angle_min, float* angle_max, float* surface_mask, hipStream_t s) {
  return sph::dplus_detect(f, phys, ff, L, eig, bx, beig, bvol_n, fb, grad, normal, mask_tmp,
                           angle_min, angle_max, surface_mask, s);
}
int sph_dplus_shift(const SphFluid* f, const SphPhysics* phys, int nscenes, const float* packing,
                    const SphCSR* ff, const SphBoundary* b, const SphCSR* fb,
                    const float* normal, const float* surface_mask, unsigned* vmax_ws,
                    float* shift, hipStream_t s) {
  return sph::dplus_shift(f, phys, nscenes, packing, ff, b, fb, normal, surface_mask, vmax_ws,
                          shift, s);
}
int sph_dplus_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const float* surface_mask,
                       unsigned terms, float* acc, hipStream_t s) {
  return sph::dplus_pressure(f, phys, ff, b, fb, surface_mask, terms, acc, s);
}
}
