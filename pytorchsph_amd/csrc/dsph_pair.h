// The delta-SPH pair routine and its row epilogue, shared by the gathers of
// gather.hip (list, packed and cell-walk forms) and the DFSPH viscosity pass
// of dfsph.hip, so every kernel that evaluates a term computes the same
// bits.
#pragma once
#include "sph_common.h"

namespace sph {
namespace {

struct Nbr {
  float x, y, vx, vy;
  float rho;   // rho_hat * rho0
  float vol;   // volume A / rho_hat (fluid) or V_b / rho_hat_b (boundary)
  float p;     // pressure
};

struct Own {
  float x, y, vx, vy, rho, p;
  float gx, gy;  // renormalised density gradient (DDSPH)
};

// Partial sums of a row.  The pair routine keeps the common factors of the
// reference's terms out of the sums; dsph_rates applies them once per row:
//   cont     sum (v_ji . gradW) V_j rho_j                  (continuity)
//   mog      sum (rho_j - rho_i) F R V_j; MOG sum = -2 mog
//   dd       sum (psi_ij . gradW) V_j                       (delta-SPH diffusion)
//   px, py   sum (p_i + p_j) V_j gradW                      (pressure)
//   vx, vy   sum w c_ij gradW; artificial viscosity sum = -(vx, vy)
//   lx, ly   sum F R m_j / (rho_i + rho_j)^2 v_ji; laminar sum = 8 nu rho_i (lx, ly)
// with gradW_ij = F r_ij (r_ij = x_i - x_j, F = -g(q) / |r_ij|),
// R = |r|^2 / (|r|^2 + eps), w = max(v_ji . r_ij, 0) and
// c_ij = m_j / ((|r|^2 + eps)(rho_i + rho_j)).
struct Acc {
  float cont, mog, dd;
  float px, py, vx, vy, lx, ly;
};

__device__ __forceinline__ bool has(unsigned terms, unsigned bit) { return (terms & bit) != 0; }

// Combines the partial sums of a lane group; only the sums that `terms`
// feeds are exchanged.
template <int KL>
__device__ __forceinline__ void reduce_acc(Acc& a, unsigned terms) {
  if (KL == 1) return;
  if (terms & (SPH_T_CONT | SPH_T_CONT_B)) a.cont = group_sum<KL>(a.cont);
  if (terms & SPH_T_MOG) a.mog = group_sum<KL>(a.mog);
  if (terms & SPH_T_DDSPH) a.dd = group_sum<KL>(a.dd);
  if (terms & (SPH_T_PRESS | SPH_T_PRESS_B)) {
    a.px = group_sum<KL>(a.px);
    a.py = group_sum<KL>(a.py);
  }
  if (terms & (SPH_T_VISC | SPH_T_VISC_B)) {
    a.vx = group_sum<KL>(a.vx);
    a.vy = group_sum<KL>(a.vy);
  }
  if (terms & SPH_T_LAM) {
    a.lx = group_sum<KL>(a.lx);
    a.ly = group_sum<KL>(a.ly);
  }
}

// One pair of the delta-SPH right-hand side (fluid row i, neighbour j).
// The reference's gradient is gradW_ij = -g(q) rhat_ij with
// g(q) = 20 (7 / pi h^3) q (1 - q)^3 and rhat = r / |r| (kernels.py:24-29),
// so gradW = F r with F = -(20 (7 / pi) / h^4) (1 - q)^3: no division by |r|
// and no q factor.  F = 0 for |r|^2 < 1e-14, where the reference zeroes
// rhat (|r| < 1e-7, neighborSearch.py:220-221).  Then
//   x_ji . gradW = -F |r|^2,   v_ji . gradW = F (v_ji . r),
// and the terms' scalar factors share F, R and m_j = rho_j V_j (see Acc).
// The multiply-adds are written out (fmaf), so every kernel that inlines
// this routine computes the same bits.
__device__ __forceinline__ void dsph_pair(const SphPhysics& P, const KConst& K, const Own& o,
                                          const Nbr& n, float grjx, float grjy, bool fluid,
                                          unsigned terms, Acc& a) {
  const float rx = o.x - n.x, ry = o.y - n.y;
  const float d2 = fmaf(ry, ry, rx * rx);
  const float d = __builtin_amdgcn_sqrtf(d2);
  const float om = fmaf(-d, K.inv_h, 1.f);           // 1 - q
  const float F = d2 >= 1e-14f ? om * om * om * K.gr : 0.f;
  const float ie = fast_rcp(d2 + P.eps);             // 1 / (|r|^2 + eps)
  const float R = d2 * ie;
  const float vjix = n.vx - o.vx, vjiy = n.vy - o.vy;
  const float pv = fmaf(vjiy, ry, vjix * rx);        // v_ji . r_ij = -(v_ji . x_ji)
  const float mj = n.vol * n.rho;
  const float FR = F * R;
  const bool cont = fluid ? has(terms, SPH_T_CONT) : has(terms, SPH_T_CONT_B);
  const bool press = fluid ? has(terms, SPH_T_PRESS) : has(terms, SPH_T_PRESS_B);
  const bool visc = fluid ? has(terms, SPH_T_VISC) : has(terms, SPH_T_VISC_B);
  // momentum.py:31-38
  if (cont) a.cont = fmaf(F * pv, mj, a.cont);
  // densityDiffusion.py:74-82: 2 (rho_j - rho_i) (x_ji . gradW) / (|r|^2 + eps) V_j
  if (fluid && has(terms, SPH_T_MOG)) a.mog = fmaf(FR, (n.rho - o.rho) * n.vol, a.mog);
  // densityDiffusion.py:63-71: psi = -2 (rho_j - rho_i) r / (|r|^2 + eps)
  // - (<grad rho>_i + <grad rho>_j)
  if (fluid && has(terms, SPH_T_DDSPH)) {
    const float gs = fmaf(o.gy + grjy, ry, (o.gx + grjx) * rx);
    const float psr = fmaf(-2.f * (n.rho - o.rho), R, -gs);   // psi . r
    a.dd = fmaf(F * psr, n.vol, a.dd);
  }
  // pressure.py:31-38
  if (press) {
    const float c = (o.p + n.p) * n.vol * F;
    a.px = fmaf(c, rx, a.px);
    a.py = fmaf(c, ry, a.py);
  }
  if (visc || (fluid && has(terms, SPH_T_LAM))) {
    const float inv_rs = fast_rcp(o.rho + n.rho);
    // diffusion.py:30-44: min(v_ji . x_ji, 0) = -max(v_ji . r_ij, 0)
    if (visc) {
      const float c = fmaxf(pv, 0.f) * ie * mj * inv_rs * F;
      a.vx = fmaf(c, rx, a.vx);
      a.vy = fmaf(c, ry, a.vy);
    }
    // laminar.py:30-46: mu_i + mu_j = 2 rho_i nu (muj indexes rhoj by i)
    if (fluid && has(terms, SPH_T_LAM)) {
      const float c = FR * mj * inv_rs * inv_rs;
      a.lx = fmaf(c, vjix, a.lx);
      a.ly = fmaf(c, vjiy, a.ly);
    }
  }
}

// dp/dt and acceleration of a row from its sums (the factors kept out of
// the pair sums, see Acc).  Written without FMA contraction, so the values
// are the same whichever kernel (plain or fused with the RK4 combine)
// evaluates them.
__device__ __forceinline__ void dsph_rates(const SphPhysics& P, const Own& o, const Acc& a,
                                           unsigned terms, float& dp, float& ax, float& ay) {
#pragma clang fp contract(off)
  dp = -a.cont + P.h * P.delta * P.c0 * (-2.f * a.mog + a.dd);
  const float inv_rho = 1.f / o.rho;
  const float av = P.h * P.alpha * P.c0;
  const float al = 8.f * P.nu * o.rho;
  ax = -a.px * inv_rho - av * a.vx - al * a.lx;
  ay = -a.py * inv_rho - av * a.vy - al * a.ly;
  if (has(terms, SPH_T_GRAV)) {
    ax += P.gx;
    ay += P.gy;
  }
}

constexpr unsigned kFluidPairTerms =
    SPH_T_CONT | SPH_T_MOG | SPH_T_DDSPH | SPH_T_PRESS | SPH_T_VISC | SPH_T_LAM;
constexpr unsigned kBdyPairTerms = SPH_T_CONT_B | SPH_T_PRESS_B | SPH_T_VISC_B;

}  // namespace
}  // namespace sph
