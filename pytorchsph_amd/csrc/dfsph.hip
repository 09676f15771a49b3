// DFSPH relaxed-Jacobi density solve (modules/dfsph.py:25-311) with the
// Akinci PBSPH boundary hooks (akinciBoundary.py:166-389), batched over
// scenes.  The reference syncs to the host once per iteration to read an
// fp32 mean residual; here the residual is summed in FP64 on the device and
// a one-workgroup control kernel decides, per scene, whether to continue.
// Kernels of scenes that have converged return immediately, so the host
// can issue iterations in chunks and poll `any_active` between chunks.
//
// Spiky gradient in the fluid-row convention: G_ij = -spiky(q) rhat_ij with
// rhat_ij = (x_i - x_j) / |x_i - x_j|.  Boundary neighbours enter fluid rows
// exactly like fluid neighbours with their own area (V_b), actual area
// (V_b / rho_b), predicted velocity (0) and pressure acceleration (0).
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;

__device__ __forceinline__ float2 ld2(const float* p, int i) {
  return reinterpret_cast<const float2*>(p)[i];
}
__device__ __forceinline__ void st2(float* p, int i, float2 v) {
  reinterpret_cast<float2*>(p)[i] = v;
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
    d.aact[t] = f.area[t] / f.rho[t];
    d.p[t] = 0.f;
    d.p2[t] = 0.f;
    st2(d.apred, t, make_float2(0.f, 0.f));
  } else if (t < f.np + b.nb) {
    const int k = t - f.np;
    d.b_aact[k] = b.vol[k] / b.rho[k];
    d.b_p[k] = 0.f;
  }
  if (t < d.nscenes) {
    d.err_acc[t] = 0.0;
    d.iters[t] = 0;
    d.active[t] = 1;
  }
  if (t == 0) *d.any_active = 1;
}

__global__ __launch_bounds__(kBS) void k_alpha_source(SphFluid f, const SphPhysics* __restrict__ phys,
                                                      SphCSR ff, SphBoundary b, SphCSR fb, SphCSR bf,
                                                      const int* __restrict__ order, SphDfsph d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < f.np) {
    const int i = order ? order[t] : t;
    const SphPhysics P = phys[f.pscene[i]];
    const float2 xi = ld2(f.x, i), vi = ld2(d.vpred, i);
    float k1x = 0.f, k1y = 0.f, k2 = 0.f, src = 0.f;
    for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
      const int j = ff.col[k];
      const float2 xj = ld2(f.x, j), vj = ld2(d.vpred, j);
      const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
      const float gs = -spiky_grad(g.q, P.h);
      const float Gx = gs * g.hx, Gy = gs * g.hy;
      const float aa = d.aact[j];
      k1x += aa * Gx;
      k1y += aa * Gy;
      k2 += aa * aa / (f.area[j] * f.rest[j]) * (Gx * Gx + Gy * Gy);
      src += -P.dt * aa * ((vi.x - vj.x) * Gx + (vi.y - vj.y) * Gy);
    }
    if (fb.rowptr) {
      for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
        const int j = fb.col[k];
        const float2 xb = ld2(b.x, j);
        const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, P.h);
        const float gs = -spiky_grad(g.q, P.h);
        const float Gx = gs * g.hx, Gy = gs * g.hy;
        const float aa = d.b_aact[j];
        k1x += aa * Gx;
        k1y += aa * Gy;
        k2 += aa * aa / (b.vol[j] * b.rest[j]) * (Gx * Gx + Gy * Gy);
        // boundary predicted velocity is zero (static walls)
        src += -P.dt * aa * (vi.x * Gx + vi.y * Gy);
      }
    }
    const float ai = d.aact[i];
    const float fac = -P.dt * P.dt * ai;
    const float mass = f.area[i] * f.rest[i];
    float al = fac / mass * (k1x * k1x + k1y * k1y) + fac * k2;
    al = al < -1.f ? -1.f : (al > -1e-7f ? -1e-7f : al);  // dfsph.py:82
    d.alpha[i] = al;
    d.src[i] = 1.f - f.rho[i] + src;
  } else if (t < f.np + b.nb) {
    const int k0 = t - f.np;
    const SphPhysics P = phys[b.bscene[k0]];
    const float2 xb = ld2(b.x, k0);
    float s2 = 0.f, src = 0.f;
    for (int k = bf.rowptr[k0]; k < bf.rowptr[k0 + 1]; ++k) {
      const int j = bf.col[k];
      const float2 xf = ld2(f.x, j), vf = ld2(d.vpred, j);
      const PairGeo g = pair_geo(xb.x, xb.y, xf.x, xf.y, P.h);
      const float gs = -spiky_grad(g.q, P.h);
      const float Gx = gs * g.hx, Gy = gs * g.hy;
      const float aa = d.aact[j];
      s2 += aa * aa / (f.area[j] * f.rest[j]) * (Gx * Gx + Gy * Gy);
      src += aa * ((0.f - vf.x) * Gx + (0.f - vf.y) * Gy);
    }
    d.b_alpha[k0] = -P.dt * P.dt * d.b_aact[k0] * s2;   // unclamped (akinciBoundary.py:340)
    d.b_src[k0] = 1.f - b.rho[k0] - P.dt * src;
  }
}

// pressure acceleration a*_i from p2 (fluid) and p_b (boundary)
__device__ __forceinline__ float2 pressure_accel(const SphFluid& f, const SphPhysics& P,
                                                 const SphCSR& ff, const SphBoundary& b,
                                                 const SphCSR& fb, const SphDfsph& d, int i) {
  const float2 xi = ld2(f.x, i);
  const float ri = f.rho[i] * f.rest[i];
  const float pi = d.p2[i] / (ri * ri);
  float ax = 0.f, ay = 0.f;
  for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
    const int j = ff.col[k];
    const float2 xj = ld2(f.x, j);
    const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
    const float gs = -spiky_grad(g.q, P.h);
    const float rj = f.rho[j] * f.rest[j];
    const float c = -(f.area[j] * f.rest[j]) * (pi + d.p2[j] / (rj * rj));
    ax += c * gs * g.hx;
    ay += c * gs * g.hy;
  }
  if (fb.rowptr) {
    for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, P.h);
      const float gs = -spiky_grad(g.q, P.h);
      const float rb = b.rho[j] * b.rest[j];
      const float c = -(b.vol[j] * b.rest[j]) * (pi + d.b_p[j] / (rb * rb));
      ax += c * gs * g.hx;
      ay += c * gs * g.hy;
    }
  }
  return make_float2(ax, ay);
}

__global__ __launch_bounds__(kBS) void k_accel(SphFluid f, const SphPhysics* __restrict__ phys,
                                               SphCSR ff, SphBoundary b, SphCSR fb,
                                               const int* __restrict__ order, SphDfsph d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const int s = f.pscene[i];
  if (!d.active[s]) return;
  const SphPhysics P = phys[s];
  st2(d.apred, i, pressure_accel(f, P, ff, b, fb, d, i));
}

// Adds v (FP64) into err_acc[scene], one atomic per wavefront when the
// whole wavefront belongs to one scene.
__device__ __forceinline__ void scene_sum(double v, int s, bool valid, double* acc) {
  const int lane = threadIdx.x & (kWave - 1);
  const int s0 = __shfl(valid ? s : -1, 0, kWave);
  const unsigned long long ball = __ballot(valid && s == s0);
  const unsigned long long bv = __ballot(valid);
  if (ball == bv && s0 >= 0) {
    const double w = wave_sum(valid ? v : 0.0);
    if (lane == 0) atomicAdd(&acc[s0], w);
  } else if (valid) {
    atomicAdd(&acc[s], v);
  }
}

__global__ __launch_bounds__(kBS) void k_pressure(SphFluid f, const SphPhysics* __restrict__ phys,
                                                  SphCSR ff, SphBoundary b, SphCSR fb, SphCSR bf,
                                                  const int* __restrict__ order, SphDfsph d) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  int s = -1;
  bool valid = false;
  if (t < f.np) {
    const int i = order ? order[t] : t;
    s = f.pscene[i];
    if (d.active[s]) {
      const SphPhysics P = phys[s];
      const float2 xi = ld2(f.x, i), ai = ld2(d.apred, i);
      const float dt2 = P.dt * P.dt;
      float ks = 0.f;
      for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
        const int j = ff.col[k];
        const float2 xj = ld2(f.x, j), aj = ld2(d.apred, j);
        const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
        const float gs = -spiky_grad(g.q, P.h);
        const float c = dt2 * d.aact[j];
        ks += c * ((ai.x - aj.x) * gs * g.hx + (ai.y - aj.y) * gs * g.hy);
      }
      if (fb.rowptr) {
        for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
          const int j = fb.col[k];
          const float2 xb = ld2(b.x, j);
          const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, P.h);
          const float gs = -spiky_grad(g.q, P.h);
          const float c = dt2 * d.b_aact[j];
          ks += c * (ai.x * gs * g.hx + ai.y * gs * g.hy);
        }
      }
      const float r = ks - d.src[i];
      const float pold = d.p2[i];
      d.p[i] = pold;
      float pn = pold - P.fluid_omega * r / d.alpha[i];
      pn = pn < 0.f ? 0.f : pn;  // clamp(min = 0) keeps NaN
      if (isnan(pn) || isinf(pn)) atomicOr(&d.status[s], SPH_ST_NAN);
      d.p2[i] = pn;
      d.resid[i] = r;
      contrib = (double)(r < -P.tau ? -P.tau : r);
      valid = true;
    }
  } else if (t < f.np + b.nb) {
    const int k0 = t - f.np;
    s = b.bscene[k0];
    if (d.active[s]) {
      const SphPhysics P = phys[s];
      const float2 xb = ld2(b.x, k0);
      const float dt2 = P.dt * P.dt;
      float ks = 0.f;
      for (int k = bf.rowptr[k0]; k < bf.rowptr[k0 + 1]; ++k) {
        const int j = bf.col[k];
        const float2 xf = ld2(f.x, j), af = ld2(d.apred, j);
        const PairGeo g = pair_geo(xb.x, xb.y, xf.x, xf.y, P.h);
        const float gs = -spiky_grad(g.q, P.h);
        const float c = dt2 * d.aact[j];
        ks += c * (af.x * (-gs * g.hx) + af.y * (-gs * g.hy));
      }
      const float r = ks - d.b_src[k0];
      float pn = d.b_p[k0] - P.omega * r / d.b_alpha[k0];
      pn = pn < 0.f ? 0.f : pn;
      d.b_p[k0] = pn;
      d.b_resid[k0] = r;
      contrib = (double)(r < -P.tau ? -P.tau : r);
      valid = true;
    }
  }
  scene_sum(contrib, s, valid, d.err_acc);
}

// Per-scene loop control, densitySolve (dfsph.py:205-207): after iteration
// i the solve continues while (i < minIters or err > tau) and i <= maxIters.
__global__ __launch_bounds__(kBS) void k_control(const SphPhysics* __restrict__ phys, SphDfsph d) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < d.nscenes; s += blockDim.x) {
    if (!d.active[s]) continue;
    const SphPhysics P = phys[s];
    const double n = (double)(d.nfluid[s] + d.nbdy[s]);
    const double err = n > 0 ? d.err_acc[s] / n : 0.0;
    const int it = d.iters[s];
    if (it < d.max_hist) d.err_hist[(size_t)s * d.max_hist + it] = err;
    const int i = it + 1;
    d.iters[s] = i;
    d.err_acc[s] = 0.0;
    const bool go = (i < P.min_iters || err > (double)P.tau) && i <= P.max_iters;
    if (!go) {
      d.active[s] = 0;
      if (err > (double)P.tau) atomicOr(&d.status[s], SPH_ST_ITERCAP);
    } else {
      atomicOr(&any, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *d.any_active = any;
}

__global__ __launch_bounds__(kBS) void k_finish(SphFluid f, const SphPhysics* __restrict__ phys,
                                                SphCSR ff, SphBoundary b, SphCSR fb,
                                                const int* __restrict__ order, SphDfsph d,
                                                float* __restrict__ acc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const SphPhysics P = phys[f.pscene[i]];
  const float2 a = pressure_accel(f, P, ff, b, fb, d, i);
  st2(d.apred, i, a);
  d.p[i] = d.p2[i];
  const float2 vp = ld2(d.vpred, i);
  st2(d.vpred, i, make_float2(vp.x + P.dt * a.x, vp.y + P.dt * a.y));
  if (acc) {
    const float2 o = ld2(acc, i);
    st2(acc, i, make_float2(o.x + a.x, o.y + a.y));
  }
}

}  // namespace

static bool dfsph_ok(const SphDfsph* d) {
  return d && d->vpred && d->aact && d->alpha && d->src && d->p && d->p2 && d->apred &&
         d->resid && d->err_acc && d->iters && d->active && d->any_active && d->status &&
         d->nfluid && d->nbdy && d->nscenes > 0;
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

int dfsph_prepare(const SphFluid* f, const float* acc, const SphPhysics* phys, const SphBoundary* b,
                  SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(f && f->v && f->area && f->rho && f->pscene && phys && dfsph_ok(d),
              "sph_dfsph_prepare: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bb.nb == 0 || (bb.vol && bb.rho && d->b_aact && d->b_p),
              "sph_dfsph_prepare: bad boundary");
  int n = f->np + bb.nb;
  if (n < d->nscenes) n = d->nscenes;
  if (n < 1) n = 1;
  hipLaunchKernelGGL(k_prepare, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, acc, phys, bb, *d);
  return launch_check("sph_dfsph_prepare");
}

int dfsph_alpha_source(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, const int* order,
                       SphDfsph* d, hipStream_t s) {
  SPH_REQUIRE(f && phys && ff && ff->rowptr && dfsph_ok(d), "sph_dfsph_alpha_source: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bb.nb == 0 || (bf && bf->rowptr && d->b_alpha && d->b_src),
              "sph_dfsph_alpha_source: boundary needs its pair list");
  const int n = f->np + bb.nb;
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_alpha_source, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, phys, *ff, bb,
                     csr_or_null(fb), csr_or_null(bf), order, *d);
  return launch_check("sph_dfsph_alpha_source");
}

int dfsph_iterate(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                  const SphCSR* fb, const SphCSR* bf, const int* order, SphDfsph* d,
                  hipStream_t s) {
  SPH_REQUIRE(f && phys && ff && ff->rowptr && dfsph_ok(d), "sph_dfsph_iterate: bad arguments");
  const SphBoundary bb = bdy_or_null(b);
  SPH_REQUIRE(bb.nb == 0 || (bf && bf->rowptr && d->b_resid), "sph_dfsph_iterate: bad boundary");
  if (f->np > 0)
    hipLaunchKernelGGL(k_accel, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys, *ff, bb,
                       csr_or_null(fb), order, *d);
  const int n = f->np + bb.nb;
  if (n > 0)
    hipLaunchKernelGGL(k_pressure, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, phys, *ff, bb,
                       csr_or_null(fb), csr_or_null(bf), order, *d);
  hipLaunchKernelGGL(k_control, dim3(1), dim3(kBS), 0, s, phys, *d);
  return launch_check("sph_dfsph_iterate");
}

int dfsph_finish(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                 const SphCSR* fb, const int* order, SphDfsph* d, float* acc, hipStream_t s) {
  SPH_REQUIRE(f && phys && ff && ff->rowptr && dfsph_ok(d), "sph_dfsph_finish: bad arguments");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_finish, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys, *ff,
                     bdy_or_null(b), csr_or_null(fb), order, *d, acc);
  return launch_check("sph_dfsph_finish");
}

}  // namespace sph

extern "C" {
int sph_dfsph_prepare(const SphFluid* f, const float* acc, const SphPhysics* phys,
                      const SphBoundary* b, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_prepare(f, acc, phys, b, d, s);
}
int sph_dfsph_alpha_source(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                           const SphBoundary* b, const SphCSR* fb, const SphCSR* bf,
                           const int* order, SphDfsph* d, hipStream_t s) {
  return sph::dfsph_alpha_source(f, phys, ff, b, fb, bf, order, d, s);
}
int sph_dfsph_iterate(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, const int* order,
                      SphDfsph* d, hipStream_t s) {
  return sph::dfsph_iterate(f, phys, ff, b, fb, bf, order, d, s);
}
int sph_dfsph_finish(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                     const SphBoundary* b, const SphCSR* fb, const int* order, SphDfsph* d,
                     float* acc, hipStream_t s) {
  return sph::dfsph_finish(f, phys, ff, b, fb, order, d, acc, s);
}
}
