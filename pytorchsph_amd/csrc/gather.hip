// Kernel-sum gathers of the delta-SPH right-hand side and of the boundary
// models.  Every gather is a CSR row gather: one thread owns one particle,
// walks its fluid neighbours and then its boundary neighbours, keeps all
// sums in registers and writes each output once -- no atomics, so results
// do not depend on scheduling.  Threads visit particles in the grid's sorted
// order (`order`) so that the neighbours of a wavefront are spatially close
// and their state is served from L2.
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

// ------------------------------------------------------------- density
__global__ __launch_bounds__(kBS) void k_fluid_density(SphFluid f, const SphPhysics* __restrict__ phys,
                                                       SphCSR ff, SphBoundary b, SphCSR fb,
                                                       const int* __restrict__ order,
                                                       float* __restrict__ rho_out,
                                                       float* __restrict__ vol_out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const SphPhysics P = phys[f.pscene[i]];
  const float2 xi = ld2(f.x, i);
  float sum = 0.f;
  for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
    const int j = ff.col[k];
    const float2 xj = ld2(f.x, j);
    const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
    sum += f.area[j] * wendland(g.q, P.h);
  }
  if (vol_out) vol_out[i] = f.area[i] / sum;
  if (fb.rowptr) {
    float bs = 0.f;
    for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xb.x, xb.y, xi.x, xi.y, P.h);
      bs += wendland(g.q, P.h) * b.vol[j];
    }
    sum += bs;
  }
  rho_out[i] = sum;
}

// ------------------------------------------------- delta-SPH fused gather
struct Nbr {
  float2 x, v;
  float rhoh;  // normalised density
  float vol;   // volume A / rhoh (fluid) or V_b / rhoh_b (boundary)
  float p;     // pressure
};

struct Acc {
  float cont, mog, dd;
  float px, py, vx, vy, lx, ly;
};

template <unsigned T>
__device__ __forceinline__ void dsph_pair(const SphPhysics& P, float2 xi, float2 vi, float rho_i,
                                          float p_i, float2 gri, const Nbr& n, float2 grj,
                                          bool fluid, unsigned terms, Acc& a) {
  const PairGeo g = pair_geo(xi.x, xi.y, n.x.x, n.x.y, P.h);
  const float gw = -wendland_grad(g.q, P.h);
  const float gWx = gw * g.hx, gWy = gw * g.hy;
  const float xjix = -g.rx, xjiy = -g.ry;           // x_j - x_i
  const float rji2 = g.d * g.d + P.eps;
  const float vjix = n.v.x - vi.x, vjiy = n.v.y - vi.y;
  const float rho_j = n.rhoh * P.rho0;
  const bool cont = fluid ? (terms & SPH_T_CONT) : (terms & SPH_T_CONT_B);
  const bool press = fluid ? (terms & SPH_T_PRESS) : (terms & SPH_T_PRESS_B);
  const bool visc = fluid ? (terms & SPH_T_VISC) : (terms & SPH_T_VISC_B);
  if (cont) a.cont += (vjix * gWx + vjiy * gWy) * n.vol * rho_j;
  if (fluid && (terms & SPH_T_MOG))
    a.mog += 2.f * (rho_j - rho_i) / rji2 * (xjix * gWx + xjiy * gWy) * n.vol;
  if (fluid && (terms & SPH_T_DDSPH)) {
    const float c = 2.f * (rho_j - rho_i) / rji2;
    const float psx = c * xjix - (gri.x + grj.x);
    const float psy = c * xjiy - (gri.y + grj.y);
    a.dd += (psx * gWx + psy * gWy) * n.vol;
  }
  if (press) {
    const float c = (p_i + n.p) * n.vol;
    a.px += c * gWx;
    a.py += c * gWy;
  }
  if (visc) {
    float pij = vjix * xjix + vjiy * xjiy;
    pij = pij > 0.f ? 0.f : pij;
    pij = pij / rji2;
    const float c = pij * n.vol * rho_j / (rho_i + rho_j);
    a.vx += c * gWx;
    a.vy += c * gWy;
  }
  if (fluid && (terms & SPH_T_LAM)) {
    // laminar.py:35-45: mu_i + mu_j = 2 rho_i nu (muj indexes rhoj by i)
    const float mj = rho_j * n.vol;
    const float num = 4.f * mj * (2.f * rho_i * P.nu) * (xjix * gWx + xjiy * gWy);
    const float rs = rho_i + rho_j;
    const float term = num / (rs * rs * rji2);
    a.lx += term * (vi.x - n.v.x);
    a.ly += term * (vi.y - n.v.y);
  }
}

template <unsigned T>
__global__ __launch_bounds__(kBS) void k_dsph_gather(SphFluid f, const SphPhysics* __restrict__ phys,
                                                     SphCSR ff, SphBoundary b, SphCSR fb,
                                                     const float* __restrict__ gradrho,
                                                     unsigned terms, const int* __restrict__ order,
                                                     float* __restrict__ dpdt,
                                                     float* __restrict__ acc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const SphPhysics P = phys[f.pscene[i]];
  const float2 xi = ld2(f.x, i), vi = ld2(f.v, i);
  const float rhoh_i = f.rho[i];
  const float rho_i = rhoh_i * P.rho0;
  const float p_i = eos(P, rhoh_i);
  const bool dd = terms & SPH_T_DDSPH;
  const float2 gri = dd ? ld2(gradrho, i) : make_float2(0.f, 0.f);
  Acc a = {};
  const bool pair_terms = terms & (SPH_T_CONT | SPH_T_MOG | SPH_T_DDSPH | SPH_T_PRESS | SPH_T_VISC | SPH_T_LAM);
  const int kend = pair_terms ? ff.rowptr[i + 1] : 0;
  for (int k = pair_terms ? ff.rowptr[i] : 0; k < kend; ++k) {
    const int j = ff.col[k];
    Nbr n;
    n.x = ld2(f.x, j);
    n.v = ld2(f.v, j);
    n.rhoh = f.rho[j];
    n.vol = f.vol ? f.vol[j] : f.area[j] / n.rhoh;
    n.p = (terms & SPH_T_PRESS) ? eos(P, n.rhoh) : 0.f;
    const float2 grj = dd ? ld2(gradrho, j) : make_float2(0.f, 0.f);
    dsph_pair<T>(P, xi, vi, rho_i, p_i, gri, n, grj, true, terms, a);
  }
  if (fb.rowptr && (terms & (SPH_T_CONT_B | SPH_T_PRESS_B | SPH_T_VISC_B))) {
    for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
      const int j = fb.col[k];
      Nbr n;
      n.x = ld2(b.x, j);
      n.v = b.v ? ld2(b.v, j) : make_float2(0.f, 0.f);
      n.rhoh = b.rho[j];
      n.vol = b.vol[j] / n.rhoh;
      n.p = b.p ? b.p[j] : 0.f;
      dsph_pair<T>(P, xi, vi, rho_i, p_i, gri, n, make_float2(0.f, 0.f), false, terms, a);
    }
  }
  const float dp = -a.cont + P.h * P.delta * P.c0 * (a.mog + a.dd);
  float ax = -a.px / rho_i + P.h * P.alpha * P.c0 * a.vx - a.lx;
  float ay = -a.py / rho_i + P.h * P.alpha * P.c0 * a.vy - a.ly;
  if (terms & SPH_T_GRAV) {
    ax += P.gx;
    ay += P.gy;
  }
  if (terms & SPH_T_ACCUM) {
    if (dpdt) dpdt[i] += dp;
    if (acc) {
      float2 o = ld2(acc, i);
      st2(acc, i, make_float2(o.x + ax, o.y + ay));
    }
  } else {
    if (dpdt) dpdt[i] = dp;
    if (acc) st2(acc, i, make_float2(ax, ay));
  }
}

// ------------------------------------------ renormalised density gradient
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

__global__ __launch_bounds__(kBS) void k_dsph_normalization(SphFluid f, const SphPhysics* __restrict__ phys,
                                                            SphCSR ff, SphBoundary b, SphCSR fb,
                                                            const int* __restrict__ order,
                                                            float* __restrict__ Mout,
                                                            float* __restrict__ Lout,
                                                            float* __restrict__ eig) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= f.np) return;
  const int i = order ? order[t] : t;
  const SphPhysics P = phys[f.pscene[i]];
  const float2 xi = ld2(f.x, i);
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f;
  // M_i = sum_j V_j (x_j - x_i) (x) gradW_ij  (densityDiffusion.py:30-43)
  for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
    const int j = ff.col[k];
    const float2 xj = ld2(f.x, j);
    const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
    const float gw = -wendland_grad(g.q, P.h);
    const float gx = gw * g.hx, gy = gw * g.hy;
    const float V = f.area[j] / f.rho[j];
    const float rx = -g.rx, ry = -g.ry;
    m00 += rx * gx * V;
    m01 += rx * gy * V;
    m10 += ry * gx * V;
    m11 += ry * gy * V;
  }
  if (fb.rowptr) {
    for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, P.h);
      const float gw = -wendland_grad(g.q, P.h);
      const float gx = gw * g.hx, gy = gw * g.hy;
      const float V = b.vol[j] / b.rho[j];
      const float rx = -g.rx, ry = -g.ry;
      m00 += rx * gx * V;
      m01 += rx * gy * V;
      m10 += ry * gx * V;
      m11 += ry * gy * V;
    }
  }
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
  const float2 xi = ld2(f.x, i);
  const float rho_i = f.rho[i] * P.rho0;
  const float l00 = L[4 * i], l01 = L[4 * i + 1], l10 = L[4 * i + 2], l11 = L[4 * i + 3];
  float sx = 0.f, sy = 0.f;
  // <grad rho>_i = -sum_j 2 (rho_j - rho_i) V_j L_i gradW_ij (densityDiffusion.py:45-61)
  for (int k = ff.rowptr[i]; k < ff.rowptr[i + 1]; ++k) {
    const int j = ff.col[k];
    const float2 xj = ld2(f.x, j);
    const PairGeo g = pair_geo(xi.x, xi.y, xj.x, xj.y, P.h);
    const float gw = -wendland_grad(g.q, P.h);
    const float gx = gw * g.hx, gy = gw * g.hy;
    const float lx = l00 * gx + l01 * gy, ly = l10 * gx + l11 * gy;
    const float c = (f.rho[j] * P.rho0 - rho_i) * (f.area[j] / f.rho[j]) * 2.f;
    sx += c * lx;
    sy += c * ly;
  }
  if (fb.rowptr) {
    for (int k = fb.rowptr[i]; k < fb.rowptr[i + 1]; ++k) {
      const int j = fb.col[k];
      const float2 xb = ld2(b.x, j);
      const PairGeo g = pair_geo(xi.x, xi.y, xb.x, xb.y, P.h);
      const float gw = -wendland_grad(g.q, P.h);
      const float gx = gw * g.hx, gy = gw * g.hy;
      const float lx = l00 * gx + l01 * gy, ly = l10 * gx + l11 * gy;
      const float c = (b.rho[j] * P.rho0 - rho_i) * (b.vol[j] / b.rho[j]) * 2.f;
      sx += c * lx;
      sy += c * ly;
    }
  }
  st2(gradrho, i, make_float2(-sx, -sy));
}

// ----------------------------------------------------------------- mDBC
// 3x3 Moore-Penrose inverse via one-sided Jacobi SVD in FP64, singular
// values below rtol * sigma_max dropped (torch.linalg.pinv default rtol =
// max(m, n) * eps of float32).
__device__ void pinv3x3(const double A[9], double X[9]) {
  double U[9], V[9];
  for (int k = 0; k < 9; ++k) {
    U[k] = A[k];
    V[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int r = 0; r < 3; ++r) {
          al += U[3 * r + p] * U[3 * r + p];
          be += U[3 * r + q] * U[3 * r + q];
          ga += U[3 * r + p] * U[3 * r + q];
        }
        if (fabs(ga) <= 1e-300 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
        off = fmax(off, fabs(ga) / sqrt(al * be));
        const double zeta = (be - al) / (2.0 * ga);
        const double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
        for (int r = 0; r < 3; ++r) {
          const double up = U[3 * r + p], uq = U[3 * r + q];
          U[3 * r + p] = c * up - s * uq;
          U[3 * r + q] = s * up + c * uq;
          const double vp = V[3 * r + p], vq = V[3 * r + q];
          V[3 * r + p] = c * vp - s * vq;
          V[3 * r + q] = s * vp + c * vq;
        }
      }
    }
    if (off < 1e-15) break;
  }
  double sig[3], smax = 0.0;
  for (int k = 0; k < 3; ++k) {
    sig[k] = sqrt(U[k] * U[k] + U[3 + k] * U[3 + k] + U[6 + k] * U[6 + k]);
    smax = fmax(smax, sig[k]);
  }
  const double cut = 3.0 * 1.1920928955078125e-07 * smax;
  for (int k = 0; k < 9; ++k) X[k] = 0.0;
  for (int k = 0; k < 3; ++k) {
    if (!(sig[k] > cut) || sig[k] == 0.0) continue;
    const double inv = 1.0 / (sig[k] * sig[k]);  // u_k = U[:,k] / sig, times 1/sig
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) X[3 * r + c] += V[3 * r + k] * inv * U[3 * c + k];
  }
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
  const float2 g = ld2(gxp, bI);
  float num = 0.f, den = 0.f, gsx = 0.f, gsy = 0.f, vgx = 0.f, vgy = 0.f, psx = 0.f, psy = 0.f;
  float m00 = 0.f, m01 = 0.f, m10 = 0.f, m11 = 0.f;
  const int a0 = gf.rowptr[bI], a1 = gf.rowptr[bI + 1];
  for (int k = a0; k < a1; ++k) {
    const int j = gf.col[k];
    const float2 xf = ld2(f.x, j);
    // distances of a foreign query: (ghost - fluid) / |.|, q = |.| / h_query
    const PairGeo geo = pair_geo(g.x, g.y, xf.x, xf.y, P.h);
    const float W = wendland(geo.q, P.h);
    const float gw = -wendland_grad(geo.q, P.h);
    const float gWx = gw * geo.hx, gWy = gw * geo.hy;
    const float mass = f.area[j] * P.rho0;
    const float V = f.area[j] / f.rho[j];
    num += mass * W;
    den += V * W;
    gsx += mass * gWx;
    gsy += mass * gWy;
    vgx += V * gWx;
    vgy += V * gWy;
    const float xijx = xf.x - g.x, xijy = xf.y - g.y;
    psx += V * W * xijx;
    psy += V * W * xijy;
    m00 += V * xijx * gWx;
    m01 += V * xijx * gWy;
    m10 += V * xijy * gWx;
    m11 += V * xijy * gWy;
  }
  const int cnt = a1 - a0;
  float rho = 1.f;
  if (cnt > 10) {
    const double A[9] = {den, psx, psy, vgx, m00, m01, vgy, m10, m11};
    double X[9];
    pinv3x3(A, X);
    const double rhs[3] = {num, gsx, gsy};
    double res[3];
    for (int r = 0; r < 3; ++r) res[r] = X[3 * r] * rhs[0] + X[3 * r + 1] * rhs[1] + X[3 * r + 2] * rhs[2];
    const float2 xb = ld2(bx, bI);
    const double dx = (double)xb.x - g.x, dy = (double)xb.y - g.y;
    rho = (float)((res[0] + dx * res[1] + dy * res[2]) / P.rho0);
  } else if (cnt > 0) {
    rho = num / den / P.rho0;
  }
  rho = rho < 1.f ? 1.f : rho;  // solidBoundary.py:615
  rho_b[bI] = rho;
  if (p_b) p_b[bI] = eos(P, rho);
  if (counts) counts[bI] = cnt;
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
  const float2 xb = ld2(bx, bI);
  float s = 0.f;
  for (int k = bf.rowptr[bI]; k < bf.rowptr[bI + 1]; ++k) {
    const int j = bf.col[k];
    const float2 xf = ld2(f.x, j);
    const PairGeo g = pair_geo(xb.x, xb.y, xf.x, xf.y, P.h);
    s += wendland(g.q, P.h) * f.area[j];
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
  hipLaunchKernelGGL(k_fluid_density, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys, *ff,
                     bdy_or_null(b), csr_or_null(fb), order, rho_out, vol_out);
  return launch_check("sph_fluid_density");
}

int dsph_gather(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff, const SphBoundary* b,
                const SphCSR* fb, const float* gradrho, unsigned terms, const int* order,
                float* dpdt, float* acc, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && f->v && phys && ff && ff->rowptr, "sph_dsph_gather: bad arguments");
  SPH_REQUIRE(!(terms & SPH_T_DDSPH) || gradrho, "sph_dsph_gather: DDSPH needs gradrho");
  const bool wants_b = terms & (SPH_T_CONT_B | SPH_T_PRESS_B | SPH_T_VISC_B);
  SPH_REQUIRE(!wants_b || !fb || !fb->rowptr || (b && b->x && b->rho && b->vol),
              "sph_dsph_gather: boundary terms without boundary state");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_dsph_gather<0>, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys, *ff,
                     bdy_or_null(b), csr_or_null(fb), gradrho, terms, order, dpdt, acc);
  return launch_check("sph_dsph_gather");
}

int dsph_normalization(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const int* order, float* M,
                       float* L, float* eig, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && ff && ff->rowptr && L, "sph_dsph_normalization: bad arguments");
  if (f->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_dsph_normalization, dim3(blocks_for(f->np, kBS)), dim3(kBS), 0, s, *f, phys,
                     *ff, bdy_or_null(b), csr_or_null(fb), order, M, L, eig);
  return launch_check("sph_dsph_normalization");
}

int dsph_density_gradient(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                          const SphBoundary* b, const SphCSR* fb, const float* L,
                          const int* order, float* gradrho, hipStream_t s) {
  SPH_REQUIRE(fluid_ok(f) && phys && ff && ff->rowptr && L && gradrho,
              "sph_dsph_density_gradient: bad arguments");
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
