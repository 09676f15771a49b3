// Time integrators (simulationBase.py:564-569, 647-652, 679-705) and a
// permutation gather.  All elementwise, one thread per particle.
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;

__global__ __launch_bounds__(kBS) void k_rk4(int stage, const int* __restrict__ pscene,
                                             const SphPhysics* __restrict__ phys, int np,
                                             const float* x0,
                                             const float* v0,
                                             const float* r0,
                                             const float* __restrict__ vs,
                                             const float* __restrict__ acc,
                                             const float* __restrict__ dpdt,
                                             float* __restrict__ K, float* xo,
                                             float* vo, float* ro,
                                             float* __restrict__ ao) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const SphPhysics P = phys[pscene[i]];
  const Rk4State s0 = {x0[2 * i], x0[2 * i + 1], v0[2 * i], v0[2 * i + 1], r0[i]};
  float2 asum;
  const Rk4State o = rk4_particle(stage, P, acc[2 * i], acc[2 * i + 1], vs[2 * i], vs[2 * i + 1],
                                  dpdt ? dpdt[i] : 0.f, dpdt != nullptr, K + 5 * (size_t)i, s0,
                                  &asum);
  vo[2 * i] = o.vx;
  vo[2 * i + 1] = o.vy;
  xo[2 * i] = o.x;
  xo[2 * i + 1] = o.y;
  if (dpdt || ro != r0) ro[i] = o.r;
  if (stage == 4 && ao) {
    ao[2 * i] = asum.x;
    ao[2 * i + 1] = asum.y;
  }
}

__global__ __launch_bounds__(kBS) void k_semi(const int* __restrict__ pscene,
                                              const SphPhysics* __restrict__ phys, int np,
                                              const float* x0,
                                              const float* v0,
                                              const float* r0,
                                              const float* vs,
                                              const float* acc,
                                              const float* dpdt,
                                              float* xo, float* vo,
                                              float* ro) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const SphPhysics P = phys[pscene[i]];
  const float dt = P.dt;
  const float ax = acc[2 * i], ay = acc[2 * i + 1];
  const float dxx = vs[2 * i] + dt * ax, dxy = vs[2 * i + 1] + dt * ay;
  vo[2 * i] = v0[2 * i] + dt * ax;
  vo[2 * i + 1] = v0[2 * i + 1] + dt * ay;
  xo[2 * i] = x0[2 * i] + dt * dxx;
  xo[2 * i + 1] = x0[2 * i + 1] + dt * dxy;
  if (ro) ro[i] = dpdt ? r0[i] + dt * dpdt[i] / P.rho0 : r0[i];
}

__global__ __launch_bounds__(kBS) void k_permute(const float* __restrict__ in,
                                                 const int* __restrict__ perm, int n, int width,
                                                 float* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * width) return;
  const int k = (int)(t / width), c = (int)(t % width);
  const int64_t src = perm ? (int64_t)perm[k] : (int64_t)k;
  out[t] = in[src * width + c];
}

// Field q of element perm[k] into tmp's block of field q (fields stored one
// after the other, each [n][width] from a 16-byte boundary): one thread per
// element, each field moved with one load and one store of its width, so a
// wavefront writes each field's rows contiguously.
// Words of tmp a field of width w takes: n * w rounded up to four, so
// every field's block starts on a 16-byte boundary.
__host__ __device__ __forceinline__ int64_t block_words(int n, int w) {
  return ((int64_t)n * w + 3) & ~(int64_t)3;
}

__device__ __forceinline__ void move_words(const float* __restrict__ in, float* __restrict__ out,
                                           int w) {
  // a vector access where both ends are aligned to it (a field block of
  // tmp starts wherever the widths before it put it)
  const uintptr_t al = reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out);
  if (w == 2 && (al & 7) == 0) {
    *reinterpret_cast<float2*>(out) = *reinterpret_cast<const float2*>(in);
  } else if (w == 4 && (al & 15) == 0) {
    *reinterpret_cast<float4*>(out) = *reinterpret_cast<const float4*>(in);
  } else {
    for (int c = 0; c < w; ++c) out[c] = in[c];
  }
}

__global__ __launch_bounds__(kBS) void k_fields_gather(SphFields f, const int* __restrict__ perm,
                                                       int n, float* __restrict__ tmp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t src = perm[k];
  int64_t o = 0;   // words of tmp before field q
  for (int q = 0; q < f.nfields; ++q) {
    const int w = f.width[q];
    move_words(f.ptr[q] + src * w, tmp + o + (int64_t)k * w, w);
    o += block_words(n, w);
  }
}

// tmp's field blocks back into the fields: a flat copy of every field's
// n * width words, four at a time (the host checks that tmp and every field
// are 16-byte aligned; otherwise one word at a time, vec = false).
template <bool VEC>
__global__ __launch_bounds__(kBS) void k_fields_copy(SphFields f, int n,
                                                     const float* __restrict__ tmp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t o = 0;
  for (int q = 0; q < f.nfields; ++q) {
    const int64_t words = (int64_t)n * f.width[q];
    const int64_t units = VEC ? (words + 3) / 4 : words;
    if (t < units) {
      float* __restrict__ out = f.ptr[q];
      const float* __restrict__ in = tmp + o;
      if (VEC && 4 * t + 4 <= words) {
        reinterpret_cast<float4*>(out)[t] = reinterpret_cast<const float4*>(in)[t];
      } else if (VEC) {
        for (int64_t c = 4 * t; c < words; ++c) out[c] = in[c];
      } else {
        out[t] = in[t];
      }
    }
    o += block_words(n, f.width[q]);
  }
}

}  // namespace

int permute_fields(const SphFields* f, const int* perm, int n, float* tmp, hipStream_t s) {
  SPH_REQUIRE(f && perm && tmp && n >= 0 && f->nfields >= 0 &&
                  f->nfields <= SPH_PERMUTE_MAX_FIELDS,
              "sph_permute_fields: bad arguments");
  int rw = 0;
  for (int q = 0; q < f->nfields; ++q) {
    SPH_REQUIRE(f->ptr[q] && f->width[q] >= 1 && f->width[q] <= 4,
                "sph_permute_fields: each field needs a pointer and a width of 1..4");
    rw += f->width[q];
  }
  if (n == 0 || rw == 0) return SPH_OK;
  hipLaunchKernelGGL(k_fields_gather, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, *f, perm, n,
                     tmp);
  // tmp and every field 16-byte aligned (the blocks of tmp then are): four
  // words per copy
  bool vec = (reinterpret_cast<uintptr_t>(tmp) & 15) == 0;
  int64_t most = 0;
  for (int q = 0; q < f->nfields; ++q) {
    const int64_t words = (int64_t)n * f->width[q];
    vec = vec && (reinterpret_cast<uintptr_t>(f->ptr[q]) & 15) == 0;
    most = words > most ? words : most;
  }
  if (vec)
    hipLaunchKernelGGL(k_fields_copy<true>, dim3(blocks_for((most + 3) / 4, kBS)), dim3(kBS), 0, s,
                       *f, n, tmp);
  else
    hipLaunchKernelGGL(k_fields_copy<false>, dim3(blocks_for(most, kBS)), dim3(kBS), 0, s, *f, n,
                       tmp);
  return launch_check("sph_permute_fields");
}

int integrate_rk4(int stage, const int* pscene, const SphPhysics* phys, int np, const float* x0,
                  const float* v0, const float* rho0, const float* vs, const float* acc,
                  const float* dpdt, float* K, float* xo, float* vo, float* ro, float* ao,
                  hipStream_t s) {
  SPH_REQUIRE(stage >= 1 && stage <= 4, "sph_integrate_rk4: stage must be 1..4");
  SPH_REQUIRE(pscene && phys && x0 && v0 && vs && acc && K && xo && vo && np >= 0,
              "sph_integrate_rk4: bad arguments");
  SPH_REQUIRE(!dpdt || (rho0 && ro), "sph_integrate_rk4: density integration needs rho buffers");
  if (np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_rk4, dim3(blocks_for(np, kBS)), dim3(kBS), 0, s, stage, pscene, phys, np,
                     x0, v0, rho0, vs, acc, dpdt, K, xo, vo, ro, ao);
  return launch_check("sph_integrate_rk4");
}

int integrate_semi(const int* pscene, const SphPhysics* phys, int np, const float* x0,
                   const float* v0, const float* rho0, const float* vs, const float* acc,
                   const float* dpdt, float* xo, float* vo, float* ro, hipStream_t s) {
  SPH_REQUIRE(pscene && phys && x0 && v0 && vs && acc && xo && vo && np >= 0,
              "sph_integrate_semi_implicit: bad arguments");
  SPH_REQUIRE(!ro || rho0, "sph_integrate_semi_implicit: rho_out needs rho0");
  if (np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_semi, dim3(blocks_for(np, kBS)), dim3(kBS), 0, s, pscene, phys, np, x0, v0,
                     rho0, vs, acc, dpdt, xo, vo, ro);
  return launch_check("sph_integrate_semi_implicit");
}

int permute_f32(const float* in, const int* perm, int n, int width, float* out, hipStream_t s) {
  SPH_REQUIRE(in && out && n >= 0 && width > 0, "sph_permute_f32: bad arguments");
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_permute, dim3(blocks_for((int64_t)n * width, kBS)), dim3(kBS), 0, s, in,
                     perm, n, width, out);
  return launch_check("sph_permute_f32");
}

}  // namespace sph

extern "C" {
int sph_permute_fields(const SphFields* f, const int* perm, int n, float* tmp, hipStream_t s) {
  return sph::permute_fields(f, perm, n, tmp, s);
}
int sph_integrate_rk4(int stage, const int* pscene, const SphPhysics* phys, int np,
                      const float* x0, const float* v0, const float* rho0, const float* v_stage,
                      const float* acc, const float* dpdt, float* kacc, float* x_out,
                      float* v_out, float* rho_out, float* acc_out, hipStream_t s) {
  return sph::integrate_rk4(stage, pscene, phys, np, x0, v0, rho0, v_stage, acc, dpdt, kacc,
                            x_out, v_out, rho_out, acc_out, s);
}
int sph_integrate_semi_implicit(const int* pscene, const SphPhysics* phys, int np,
                                const float* x0, const float* v0, const float* rho0,
                                const float* v_stage, const float* acc, const float* dpdt,
                                float* x_out, float* v_out, float* rho_out, hipStream_t s) {
  return sph::integrate_semi(pscene, phys, np, x0, v0, rho0, v_stage, acc, dpdt, x_out, v_out,
                             rho_out, s);
}
int sph_permute_f32(const float* in, const int* perm, int n, int width, float* out,
                    hipStream_t s) {
  return sph::permute_f32(in, perm, n, width, out, s);
}
}
