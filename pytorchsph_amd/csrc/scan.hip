// Exclusive prefix sums and fills.  Three launches: per-tile sums, a
// single-workgroup scan of the tile sums, and the tile-local scans.
#include "sph_common.h"

#include <stdint.h>
#include <stdio.h>

#include <atomic>
#include <mutex>

namespace sph {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int hip_fail(const char* where) {
  hipError_t e = hipGetLastError();
  char buf[256];
  snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  g_last_error = buf;
  return SPH_E_HIP;
}

int launch_check(const char* where) {
  hipError_t e = hipPeekAtLastError();
  if (e != hipSuccess) return hip_fail(where);
  return SPH_OK;
}

// Defaults (the shapes the autotuner picks on a 256-scene mixed batch):
// lanes per row of the neighbour count, the gathers, the DFSPH solve, the
// neighbour fill and the cell walk, and the workgroup size of the list
// gathers and the per-query search.
static std::atomic<int> g_tuning[SPH_TUNE_COUNT] = {{2}, {4}, {2}, {8}, {1}, {256}};

int tuning(int knob) { return g_tuning[knob].load(std::memory_order_relaxed); }

static int set_tuning(int knob, int value) {
  if (knob < 0 || knob >= SPH_TUNE_COUNT) {
    set_error("sph_set_tuning: unknown knob");
    return SPH_E_ARG;
  }
  if (knob == SPH_TUNE_BLOCK) {
    if (value != 64 && value != 128 && value != 256) {
      set_error("sph_set_tuning: workgroup size must be 64, 128 or 256");
      return SPH_E_ARG;
    }
  } else if (value < 1 || value > 32 || (value & (value - 1))) {
    set_error("sph_set_tuning: lanes per row must be 1, 2, 4, 8, 16 or 32");
    return SPH_E_ARG;
  }
  g_tuning[knob].store(value, std::memory_order_relaxed);
  return SPH_OK;
}

namespace {
constexpr int kScanBS = 256;
constexpr int kScanItems = 8;   // the vector path moves them as two int4
static_assert(kScanItems == 8, "load_items and k_scan_apply move 8 items as two int4");
constexpr int kTile = kScanBS * kScanItems;

// Each thread owns kScanItems consecutive items.  With 16-byte aligned
// arrays (VEC) it moves them as two int4 (a wave reads 2 KiB contiguous per
// instruction pair); otherwise one int at a time.
template <bool VEC>
__device__ __forceinline__ void load_items(const int* __restrict__ in, int n, int base,
                                           int (&v)[kScanItems]) {
  if (VEC && base + kScanItems <= n) {
    const int4 a = *reinterpret_cast<const int4*>(in + base);
    const int4 b = *reinterpret_cast<const int4*>(in + base + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    return;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0;
}

template <bool VEC>
__global__ __launch_bounds__(kScanBS) void k_scan_reduce(const int* __restrict__ in, int n,
                                                         int* __restrict__ partial) {
  const int base = blockIdx.x * kTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  load_items<VEC>(in, n, base, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += v[k];
  s = wave_sum(s);
  __shared__ int lds[kScanBS / kWave];
  if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanBS / kWave; ++w) t += lds[w];
    partial[blockIdx.x] = t;
  }
}

// Exclusive scan of the nb tile sums in place (one workgroup): thread t owns
// the contiguous segment [t seg, (t + 1) seg), sums it, takes its offset from
// one workgroup scan of those sums and rewrites the segment -- one scan
// whatever nb (a loop of workgroup scans over 1024 sums at a time cost a
// few microseconds of barriers per round).
__global__ __launch_bounds__(1024) void k_scan_partials(int* __restrict__ partial, int nb,
                                                        int* __restrict__ out, int n) {
  __shared__ int lds[1024 / kWave + 1];
  const int seg = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
  const int a = threadIdx.x * seg;
  const int b = a + seg < nb ? a + seg : nb;
  int sum = 0;
  for (int i = a; i < b; ++i) sum += partial[i];
  int tot;
  int run = block_exclusive_scan(sum, lds, &tot);
  for (int i = a; i < b; ++i) {
    const int v = partial[i];
    partial[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) out[n] = tot;
}

template <bool VEC>
__global__ __launch_bounds__(kScanBS) void k_scan_apply(const int* __restrict__ in, int n,
                                                        const int* __restrict__ partial,
                                                        int* __restrict__ out) {
  __shared__ int lds[kScanBS / kWave + 1];
  const int base = blockIdx.x * kTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  load_items<VEC>(in, n, base, v);
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += v[k];
  int tot;
  int ex = block_exclusive_scan(s, lds, &tot) + partial[blockIdx.x];
  int o[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    o[k] = ex;
    ex += v[k];
  }
  if (VEC && base + kScanItems <= n) {
    *reinterpret_cast<int4*>(out + base) = make_int4(o[0], o[1], o[2], o[3]);
    *reinterpret_cast<int4*>(out + base + 4) = make_int4(o[4], o[5], o[6], o[7]);
    return;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) out[base + k] = o[k];
}

template <typename T>
__global__ void k_fill(T* p, int n, T v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Tile partials of the no-allocation scan: one int per tile of kTile items
// plus the total, so n <= kTile * (SPH_SCAN_NOALLOC_PARTIALS - 1).  Each GPU
// the library runs on holds its own copy until the process ends.
__device__ __align__(256) int scratch[SPH_SCAN_NOALLOC_PARTIALS];
}  // namespace

// Device address of `scratch` on each device, looked up once per device with
// that device current.  The array's name is a host-side stand-in; only these
// addresses may be handed to kernels.
constexpr int kMaxDevices = 64;
static std::atomic<int*> g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(int** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return hip_fail("sph_scan_exclusive_noalloc(device)");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("sph_scan_exclusive_noalloc: device index beyond the supported 64");
    return SPH_E_ARG;
  }
  int* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* addr = nullptr;
      if (hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)) != hipSuccess || !addr)
        return hip_fail("sph_scan_exclusive_noalloc(hipGetSymbolAddress)");
      p = static_cast<int*>(addr);
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return SPH_OK;
}

int scan_exclusive_noalloc(const int* in, int n, int* out, hipStream_t s) {
  SPH_REQUIRE(n >= 0 && out && (n == 0 || in), "sph_scan_exclusive_noalloc: bad size or buffers");
  const int64_t max_n = (int64_t)kTile * (SPH_SCAN_NOALLOC_PARTIALS - 1);
  if ((int64_t)n > max_n) {
    set_error("sph_scan_exclusive_noalloc: n exceeds the static partials (at most " +
              std::to_string(max_n) + " items); use sph_scan_exclusive with a workspace");
    return SPH_E_SPACE;
  }
  int* partial = nullptr;
  int rc = scratch_ptr(&partial);
  if (rc) return rc;
  // every partial read is written earlier in the same call (k_scan_reduce
  // writes one per tile, k_scan_partials rescans them), so nothing left by
  // a previous call is read and nothing needs clearing
  return scan_exclusive(in, n, out, partial, sizeof(int) * (size_t)SPH_SCAN_NOALLOC_PARTIALS, s);
}

int scan_exclusive(const int* in, int n, int* out, void* ws, size_t ws_bytes, hipStream_t s) {
  SPH_REQUIRE(n >= 0 && out, "sph_scan_exclusive: bad size or output");
  if (n == 0) {
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, s, out, 1, 0);
    return launch_check("sph_scan_exclusive(fill)");
  }
  const int nb = blocks_for(n, kTile);
  if (ws_bytes < sizeof(int) * (size_t)(nb + 1) || !ws) {
    set_error("sph_scan_exclusive: workspace too small");
    return SPH_E_SPACE;
  }
  int* partial = (int*)ws;
  const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  if (vec)
    hipLaunchKernelGGL(k_scan_reduce<true>, dim3(nb), dim3(kScanBS), 0, s, in, n, partial);
  else
    hipLaunchKernelGGL(k_scan_reduce<false>, dim3(nb), dim3(kScanBS), 0, s, in, n, partial);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, partial, nb, out, n);
  if (vec)
    hipLaunchKernelGGL(k_scan_apply<true>, dim3(nb), dim3(kScanBS), 0, s, in, n, partial, out);
  else
    hipLaunchKernelGGL(k_scan_apply<false>, dim3(nb), dim3(kScanBS), 0, s, in, n, partial, out);
  return launch_check("sph_scan_exclusive");
}

size_t scan_ws_bytes(int n) { return sizeof(int) * (size_t)(blocks_for(n < 1 ? 1 : n, kTile) + 1); }

int fill_i32(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return SPH_OK;
  hipLaunchKernelGGL(k_fill<int>, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  return launch_check("sph_fill_i32");
}
int fill_f32(float* p, int n, float v, hipStream_t s) {
  if (n <= 0) return SPH_OK;
  hipLaunchKernelGGL(k_fill<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  return launch_check("sph_fill_f32");
}

}  // namespace sph

extern "C" {
const char* sph_last_error(void) { return sph::g_last_error.c_str(); }
int sph_version(void) { return 1; }
size_t sph_scan_workspace_bytes(int n) { return sph::scan_ws_bytes(n); }
int sph_scan_exclusive(const int* in, int n, int* out, void* ws, size_t ws_bytes,
                       hipStream_t s) {
  return sph::scan_exclusive(in, n, out, ws, ws_bytes, s);
}
int sph_scan_exclusive_noalloc(const int* in, int n, int* out, hipStream_t s) {
  return sph::scan_exclusive_noalloc(in, n, out, s);
}
int sph_fill_i32(int* p, int n, int value, hipStream_t s) { return sph::fill_i32(p, n, value, s); }
int sph_fill_f32(float* p, int n, float value, hipStream_t s) {
  return sph::fill_f32(p, n, value, s);
}
int sph_set_tuning(int knob, int value) { return sph::set_tuning(knob, value); }
int sph_get_tuning(int knob) {
  if (knob < 0 || knob >= SPH_TUNE_COUNT) {
    sph::set_error("sph_get_tuning: unknown knob");
    return -SPH_E_ARG;
  }
  return sph::tuning(knob);
}
}
