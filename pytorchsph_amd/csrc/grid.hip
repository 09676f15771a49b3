// Neighbour search on a dense per-scene cell table.
//
// The reference sorts particles by a linear cell key and looks cells up
// through a hash table keyed (3 ix + 5 iy) mod P (neighborSearch.py:58-88,
// neighSearch.cpp:93-114).  On a bounded 2D domain the table can simply be
// dense: a counting sort by (scene, key) produces the cell-start array in
// one scan, and every stencil row of 2R+1 cells becomes one contiguous range
// of sorted points.  Pair predicate and cell formula are the reference's,
// so the pair set is the same.
#include "sph_common.h"

namespace sph {
namespace {

constexpr int kBS = 256;

// Cell key and a slot inside the cell.  Points in cell order (the engine
// keeps its state that way) put runs of equal keys in neighbouring lanes:
// each run takes its slots with one atomic, issued by the run's first lane,
// and its lanes take consecutive slots.  (Which points fill which slots of
// a cell still depends on the order the atomics land; k_grid_rank makes it
// deterministic.)
// keys_in == nullptr: the key is computed from the position; else it is
// keys_in[p] (sph_grid_build_keyed).
__global__ __launch_bounds__(kBS) void k_grid_keys(SphGrid g, const int* __restrict__ keys_in) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const bool valid = p < g.np;
  int kg = -1 - lane;          // distinct negative keys for lanes past the end
  int s = 0, k = 0;
  if (valid) {
    s = g.pscene[p];
    const SphGridDesc d = g.scenes[s];
    const float x = g.x[2 * p], y = g.x[2 * p + 1];
    int ix = 0, iy = 0;
    int64_t key;
    if (keys_in) {
      key = keys_in[p];
    } else {
      ix = cell_coord(x, d.qmin_x, d.hmax);
      iy = cell_coord(y, d.qmin_y, d.hmax);
      // reference key: ix + cellsX * iy in int32 (neighborSearch.py:47)
      key = (int64_t)ix + (int64_t)d.cells_x * (int64_t)iy;
    }
    const int64_t ncell = (int64_t)d.cells_x * d.cells_y;
    // a NaN / Inf position has no neighbour under the fp32 predicate; it
    // goes to the outside slot rather than to whatever cell the float to
    // int conversion gives (0 for NaN), where a blown-up scene would pile
    // all its points into one cell that stencils do visit
    const bool finite = isfinite(x) && isfinite(y);
    if (finite && key >= 0 && key < ncell) {
      k = (int)key;
      // a point past a side edge keeps the reference's aliased key (the
      // first cells of the next row, or the last of the previous one) so
      // that pairs stay those of the reference, and is flagged: its own
      // stencil no longer covers its true neighbourhood
      if (!keys_in && (ix < 0 || ix >= d.cells_x || iy < 0 || iy >= d.cells_y))
        atomicOr(&g.status[s], SPH_ST_OUTSIDE);
    } else {
      // the scene's last slot collects points outside its table; stencils
      // never reach it, and the scene stays contiguous in sorted order
      k = (int)ncell;
      atomicOr(&g.status[s], SPH_ST_OUTSIDE);
    }
    kg = d.cell_begin + k;     // global cell slot
    g.key[p] = k;
  }
  // runs of equal global cells among neighbouring lanes
  const int prev = __shfl_up(kg, 1, kWave);
  const bool head = lane == 0 || prev != kg;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(heads & upto);              // first lane of my run
  const unsigned long long after = heads & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : kWave;  // first lane of the next run
  int base = 0;
  if (valid && head) base = atomicAdd(&g.cell_count[kg], end - start);
  base = __shfl(base, start, kWave);
  if (valid) g.slot[p] = base + (lane - start);
}

// Places each point at its atomic slot of its cell: the pair {point, global
// cell} goes into the sorted position, staged in sorted_x's memory (8 bytes
// per slot) until k_grid_rank and k_grid_place have ordered each cell.
__global__ __launch_bounds__(kBS) void k_grid_scatter(SphGrid g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  const int s = g.pscene[p];
  const int kg = g.scenes[s].cell_begin + g.key[p];
  const int pos = g.cell_start[kg] + g.slot[p];
  reinterpret_cast<int2*>(g.sorted_x)[pos] = make_int2(p, kg);
}

// Whether global cell c is its scene's slot of outside points (the scene's
// last slot): stencils never visit it, so its order does not matter.
__device__ bool outside_slot(const SphGrid& g, int c) {
  int lo = 0, hi = g.nscenes;   // scene s with cell_begin[s] <= c
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (g.scenes[m].cell_begin <= c) lo = m; else hi = m;
  }
  const SphGridDesc& d = g.scenes[lo];
  return (int64_t)c == (int64_t)d.cell_begin + (int64_t)d.cells_x * d.cells_y;
}

// An outside slot holding more points than this keeps the atomic order
// (ranking its k points costs k^2 loads, and no stencil reads it).
constexpr int kRankOutsideMax = 64;

// One thread per sorted position: the rank of its point among the points of
// its cell by index (the order the atomics left them in is arbitrary), as
// the point's final slot.  Reads the staged pairs of k_grid_scatter; points
// of the engine's cell-ordered state make these reads nearly contiguous.
__global__ __launch_bounds__(kBS) void k_grid_rank(SphGrid g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.np) return;
  const int2* __restrict__ tmp = reinterpret_cast<const int2*>(g.sorted_x);
  const int2 e = tmp[i];
  const int a = g.cell_start[e.y], b = g.cell_start[e.y + 1];
  int rank = 0;
  if (b - a > kRankOutsideMax && outside_slot(g, e.y)) {
    rank = i - a;
  } else {
    for (int q = a; q < b; ++q) rank += tmp[q].x < e.x ? 1 : 0;
  }
  g.slot[e.x] = a + rank;
}

// One thread per point: the point and its position into their sorted slot.
__global__ __launch_bounds__(kBS) void k_grid_place(SphGrid g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  const int pos = g.slot[p];
  g.sorted_idx[pos] = p;
  reinterpret_cast<float2*>(g.sorted_x)[pos] = reinterpret_cast<const float2*>(g.x)[p];
}

constexpr int kCandBatch = 4;  // candidates each lane loads at once in a search pass

// Count (FILL = false) or emit (FILL = true) the pairs of every query.
// Stencil rows are walked bottom to top and each row is one contiguous range
// of sorted points, so the columns of a row come out in sorted-point order.
//
// KL lanes share one query: lane `sub` takes candidates p = p0 + u KL + sub,
// u < kCandBatch, so the group reads consecutive sorted points (coalesced)
// and every lane has kCandBatch independent loads in flight before its first
// test.  Hits are placed with a ballot and a popcount of the lower lanes, so
// the row is written in candidate order.  Wider stencils (radius > 1) walk
// their rows one load at a time.
//
// Self-sorted queries read their position from the sorted copy (coalesced,
// no wait on the index load), and a one-scene grid needs no scene lookup, so
// the first stencil loads wait on one load only.
template <bool FILL, int KL>
__global__ __launch_bounds__(kBS) void k_neighbors(SphGrid g, SphQuery q,
                                                   const int* __restrict__ rowptr,
                                                   int* __restrict__ out, int64_t cap) {
  const int64_t gt = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const int t = (int)(gt / KL);
  const int sub = (int)(gt % KL);
  if (t >= q.nq) return;   // whole groups leave together
  const int qi = q.self_sorted ? g.sorted_idx[t] : t;
  const int s = g.nscenes == 1 ? 0 : q.qscene[qi];
  const SphGridDesc d = g.scenes[s];
  const float* __restrict__ xp = q.self_sorted ? g.sorted_x + 2 * (int64_t)t : q.x + 2 * (int64_t)qi;
  const float xq = xp[0], yq = xp[1];
  int ix, iy;
  if (q.qcell) {
    // the cell of a given key, in C integer arithmetic as
    // neighSearch.cpp:207-209 derives it
    const int key = q.qcell[qi];
    ix = key % d.cells_x;
    iy = key / d.cells_x;
  } else {
    ix = cell_coord(xq, d.qmin_x, d.hmax);
    iy = cell_coord(yq, d.qmin_y, d.hmax);
  }
  const int R = q.radius;
  const int* __restrict__ cs = g.cell_start + d.cell_begin;
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  const unsigned long long gmask = group_mask<KL>();
  int64_t base = 0;
  if (FILL) base = rowptr[qi];
  int cnt = 0;   // hits so far, identical in every lane of the group
  if (R == 1) {
    Stencil3 st;
    st.init(d, cs, ix, iy);
    const int ncand = st.ncand;
    const float thr = sq_threshold(q.h[qi]);
    for (int p0 = 0; p0 < ncand; p0 += KL * kCandBatch) {
      float2 xj[kCandBatch];
      int idx[kCandBatch];
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        const int j = st.slot(p0 + u * KL + sub);
        xj[u] = sx[j];
        if (FILL) idx[u] = g.sorted_idx[j];
      }
#pragma unroll
      for (int u = 0; u < kCandBatch; ++u) {
        const int p = p0 + u * KL + sub;
        const bool hit = p < ncand && dist2_nofma(xq, yq, xj[u].x, xj[u].y) <= thr;
        const unsigned long long m = __ballot(hit) & gmask;
        if (FILL && hit) {
          const int64_t pos = base + cnt + __popcll(m & lanes_below());
          if (pos < cap) out[pos] = idx[u];
        }
        cnt += __popcll(m);
      }
    }
  } else {
    const float thr = sq_threshold(q.h[qi]);
    for (int dy = -R; dy <= R; ++dy) {
      int a, b;
      if (!stencil_row(d, cs, ix, iy, dy, R, a, b)) break;
      for (int c0 = a; c0 < b; c0 += KL) {
        const int j = c0 + sub;
        bool hit = false;
        if (j < b) {
          const float2 xj = sx[j];
          hit = dist2_nofma(xq, yq, xj.x, xj.y) <= thr;
        }
        const unsigned long long m = __ballot(hit) & gmask;
        if (FILL && hit) {
          const int64_t pos = base + cnt + __popcll(m & lanes_below());
          if (pos < cap) out[pos] = g.sorted_idx[j];
        }
        cnt += __popcll(m);
      }
    }
  }
  if (sub == 0) {
    if (!FILL) {
      out[qi] = cnt;
    } else if (base + cnt > cap) {
      atomicOr(&g.status[s], SPH_ST_OVERFLOW);
    }
  }
}

// Count the pairs of self-sorted, radius-1 queries with the stencil staged
// in LDS.  A workgroup takes 256 consecutive sorted points; the three
// stencil runs of such a chunk are three contiguous runs of sorted points
// (chunk_runs), copied into LDS with coalesced loads, and one lane per row
// tests its candidates from there with the fill's predicate, four loads in
// flight and no ballot per candidate (C2: 139 us against 223 for
// k_neighbors' count).  Chunks whose stencil does not fit in LDS count from
// global memory, one lane per row.  (The fill keeps k_neighbors: the same
// walk storing each hit as found, or listing a chunk's hits in LDS and
// writing them after the walk, wrote C2's randomly ordered rows back from
// the L2 in more pieces and took 15-60% longer.)
constexpr int kSWalkRecs = 3 * kBS * 3 / 2 + 256;   // staged stencil points per chunk
constexpr int kSWalkBatch = 4;                      // candidates tested per trip
constexpr int kSWalkRound = 6;                      // staged slots per thread per round

__global__ __launch_bounds__(kBS) void k_walk_count(SphGrid g, SphQuery q,
                                                    int* __restrict__ counts) {
  __shared__ float2 s_pos[kSWalkRecs + kSWalkBatch];   // padded: a trip reads past a run's end
  __shared__ int s_wlo[kBS / kWave][3], s_whi[kBS / kWave][3];
  const int t = threadIdx.x;
  const int sp = xcd_block() * kBS + t;
  const bool valid = sp < q.nq;
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  int qi = 0;
  float xq = 0.f, yq = 0.f, thr = 0.f;
  Stencil3 st;
  if (valid) {
    qi = g.sorted_idx[sp];
    const int sc = g.nscenes == 1 ? 0 : q.qscene[qi];
    const SphGridDesc d = g.scenes[sc];
    const float2 xy = sx[sp];
    xq = xy.x;
    yq = xy.y;
    int ix, iy;
    if (q.qcell) {
      // the cell of a given key, in C integer arithmetic as
      // neighSearch.cpp:207-209 derives it
      const int key = q.qcell[qi];
      ix = key % d.cells_x;
      iy = key / d.cells_x;
    } else {
      ix = cell_coord(xq, d.qmin_x, d.hmax);
      iy = cell_coord(yq, d.qmin_y, d.hmax);
    }
    st.init(d, g.cell_start + d.cell_begin, ix, iy);
    thr = sq_threshold(q.h[qi]);
  }
  const ChunkRuns cr = chunk_runs(st, valid, s_wlo, s_whi);
  const bool staged = cr.tot <= kSWalkRecs;
  if (staged) {
    // kSWalkRound slots per thread in flight before their LDS writes (one
    // wait per round, not per slot); a wavefront runs a round when any of
    // its slots is staged, slots past the end re-reading the last one
    const int w0 = __builtin_amdgcn_readfirstlane(t);
    const int tot = cr.tot;
    for (int u0 = 0; w0 + u0 * kBS < tot; u0 += kSWalkRound) {
      float2 pv[kSWalkRound];
#pragma unroll
      for (int u = 0; u < kSWalkRound; ++u) {
        const int k = t + (u0 + u) * kBS;
        pv[u] = sx[cr.slot(k < tot ? k : tot - 1)];
      }
#pragma unroll
      for (int u = 0; u < kSWalkRound; ++u) {
        const int k = t + (u0 + u) * kBS;
        if (k < tot) s_pos[k] = pv[u];
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  int cnt = 0;
  if (staged) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int b = cr.first(r, st), e = b + (st.rb[r] - st.ra[r]);
      for (int k = b; k < e; k += kSWalkBatch) {
        float2 xy[kSWalkBatch];
#pragma unroll
        for (int u = 0; u < kSWalkBatch; ++u) xy[u] = s_pos[k + u];
#pragma unroll
        for (int u = 0; u < kSWalkBatch; ++u) {
          const float d2 = dist2_nofma(xq, yq, xy[u].x, xy[u].y);
          cnt += ((d2 <= thr) & (k + u < e)) ? 1 : 0;
        }
      }
    }
  } else {
    for (int p = 0; p < st.ncand; ++p) {
      const float2 xy = sx[st.slot(p)];
      cnt += dist2_nofma(xq, yq, xy.x, xy.y) <= thr ? 1 : 0;
    }
  }
  counts[qi] = cnt;
}

// Stencil lists (sph_stencil_list_build): k_walk_count's chunks, stage and
// test, with each hit set as a bit of its run's mask instead of counted.
// Rows are sorted slots, so a row's candidates are its stencil runs as
// they are and its bits follow the fill's candidate order.
constexpr int kSlMaxRun = 64;   // candidates a run's mask holds

__global__ __launch_bounds__(kBS) void k_stencil_list(SphGrid g, const float* __restrict__ h,
                                                      int4* __restrict__ runs,
                                                      unsigned long long* __restrict__ masks) {
  __shared__ float2 s_pos[kSWalkRecs + kSWalkBatch];   // padded: a trip reads past a run's end
  __shared__ int s_wlo[kBS / kWave][3], s_whi[kBS / kWave][3];
  const int t = threadIdx.x;
  const int sp = xcd_block() * kBS + t;
  const bool valid = sp < g.np;
  const float2* __restrict__ sx = reinterpret_cast<const float2*>(g.sorted_x);
  int sc = 0;
  float xq = 0.f, yq = 0.f, thr = 0.f;
  Stencil3 st;
  if (valid) {
    const int qi = g.sorted_idx[sp];
    sc = g.nscenes == 1 ? 0 : g.pscene[qi];
    const SphGridDesc d = g.scenes[sc];
    const float2 xy = sx[sp];
    xq = xy.x;
    yq = xy.y;
    st.init(d, g.cell_start + d.cell_begin, cell_coord(xq, d.qmin_x, d.hmax),
            cell_coord(yq, d.qmin_y, d.hmax));
    thr = sq_threshold(h[qi]);
  }
  const ChunkRuns cr = chunk_runs(st, valid, s_wlo, s_whi);
  const int tot = cr.tot;
  const bool staged = tot <= kSWalkRecs;
  if (staged) {
    const int w0 = __builtin_amdgcn_readfirstlane(t);
    for (int u0 = 0; w0 + u0 * kBS < tot; u0 += kSWalkRound) {
      float2 pv[kSWalkRound];
#pragma unroll
      for (int u = 0; u < kSWalkRound; ++u) {
        const int k = t + (u0 + u) * kBS;
        pv[u] = sx[cr.slot(k < tot ? k : tot - 1)];
      }
#pragma unroll
      for (int u = 0; u < kSWalkRound; ++u) {
        const int k = t + (u0 + u) * kBS;
        if (k < tot) s_pos[k] = pv[u];
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  unsigned long long m[3] = {0ull, 0ull, 0ull};
  int len[3];
  bool over = false;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    len[r] = st.rb[r] - st.ra[r];
    over |= len[r] > kSlMaxRun;
    const int n = len[r] < kSlMaxRun ? len[r] : kSlMaxRun;
    // LDS or global memory for the whole run (a choice per candidate
    // between the two would compile to flat loads)
    if (staged) {
      const int b = cr.first(r, st);
      for (int k = 0; k < n; k += kSWalkBatch) {
        float2 xy[kSWalkBatch];
#pragma unroll
        for (int u = 0; u < kSWalkBatch; ++u) xy[u] = s_pos[b + k + u];   // padded
        unsigned nib = 0;
#pragma unroll
        for (int u = 0; u < kSWalkBatch; ++u)
          nib |= ((k + u < n) && dist2_nofma(xq, yq, xy[u].x, xy[u].y) <= thr) ? 1u << u : 0u;
        m[r] |= (unsigned long long)nib << k;
      }
    } else {
      for (int k = 0; k < n; ++k) {
        const float2 xy = sx[st.ra[r] + k];
        m[r] |= dist2_nofma(xq, yq, xy.x, xy.y) <= thr ? 1ull << k : 0ull;
      }
    }
  }
  if (over) atomicOr(&g.status[sc], SPH_ST_OVERFLOW);
  const int l0 = len[0] < 255 ? len[0] : 255, l1 = len[1] < 255 ? len[1] : 255,
            l2 = len[2] < 255 ? len[2] : 255;
  runs[sp] = make_int4(st.ra[0], st.ra[1], st.ra[2], l0 | (l1 << 8) | (l2 << 16));
  const size_t n = (size_t)g.np;
  masks[sp] = m[0];
  masks[n + sp] = m[1];
  masks[2 * n + sp] = m[2];
}

__global__ __launch_bounds__(kBS) void k_iota(int* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ __launch_bounds__(kBS) void k_clamp_rowptr(int* __restrict__ rowptr, int n, int64_t cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((int64_t)rowptr[i] > cap) rowptr[i] = (int)cap;
}

__global__ __launch_bounds__(kBS) void k_sort_rows(const int* __restrict__ rowptr, int nrows,
                                                   int* __restrict__ col) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int a = rowptr[r], b = rowptr[r + 1];
  for (int i = a + 1; i < b; ++i) {
    const int v = col[i];
    int j = i - 1;
    while (j >= a && col[j] > v) {
      col[j + 1] = col[j];
      --j;
    }
    col[j + 1] = v;
  }
}

__global__ __launch_bounds__(kBS) void k_tr_count(const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, int nrows,
                                                  int* __restrict__ tcount,
                                                  int* __restrict__ slot) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) slot[k] = atomicAdd(&tcount[col[k]], 1);
}

__global__ __launch_bounds__(kBS) void k_tr_scatter(const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, int nrows,
                                                    const int* __restrict__ slot,
                                                    const int* __restrict__ trowptr,
                                                    int* __restrict__ tcol) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) tcol[trowptr[col[k]] + slot[k]] = r;
}

__global__ __launch_bounds__(kBS) void k_pair_geometry(const int* __restrict__ rowptr,
                                                       const int* __restrict__ col, int nrows,
                                                       const float* __restrict__ xq,
                                                       const float* __restrict__ xg,
                                                       const float* __restrict__ hq,
                                                       float support, float thresh,
                                                       int64_t* __restrict__ rows64,
                                                       int64_t* __restrict__ cols64,
                                                       float* __restrict__ rhat,
                                                       float* __restrict__ qout) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const float xi = xq[2 * r], yi = xq[2 * r + 1];
  const float hdiv = hq ? hq[r] : support;
  for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
    const int j = col[k];
    const float rx = xi - xg[2 * j], ry = yi - xg[2 * j + 1];
    const float d = sqrtf(rx * rx + ry * ry);
    float hx = 0.f, hy = 0.f;
    if (d >= thresh) {
      hx = rx / d;
      hy = ry / d;
    }
    if (rows64) rows64[k] = r;
    if (cols64) cols64[k] = j;
    if (rhat) {
      rhat[2 * (int64_t)k] = hx;
      rhat[2 * (int64_t)k + 1] = hy;
    }
    if (qout) qout[k] = d / hdiv;
  }
}

__global__ __launch_bounds__(kBS) void k_pair_thresholds(const float* __restrict__ h, int n,
                                                         float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sq_threshold(h[i]);
}

// ------------------------------------------------ radix-sort grid build
// For points in arbitrary order (the atomics of k_grid_keys serialise on
// random cells, and the scatters of k_grid_scatter / k_grid_place write one
// word per random line): a least-significant-digit radix sort of the pairs
// (global cell, point index), 8 bits per pass, stable, so the points come
// out in (scene, key, index) order -- the order sph_grid_build gives them.
// Each pass: a per-tile digit histogram (LDS atomics), an exclusive scan of
// the digit-major histograms (every digit's tiles in order), then per tile a
// stable rank of every item among the tile's items of its digit, a reorder
// of the tile in LDS by digit, and coalesced writes of each digit's run.
constexpr int kRsBits = 8;
constexpr int kRsDigits = 1 << kRsBits;
constexpr int kRsItems = 16;                      // items per thread
constexpr int kRsTile = kBS * kRsItems;          // items per tile
constexpr int kRsWaves = kBS / kWave;

// Global cell of every point (k_grid_keys' cell and outside rules, without
// the slot atomics); the local key into g.key.
__global__ __launch_bounds__(kBS) void k_rs_keys(SphGrid g, const int* __restrict__ keys_in,
                                                 int* __restrict__ kout) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.np) return;
  const int s = g.pscene[p];
  const SphGridDesc d = g.scenes[s];
  const float x = g.x[2 * p], y = g.x[2 * p + 1];
  int ix = 0, iy = 0;
  int64_t key;
  if (keys_in) {
    key = keys_in[p];
  } else {
    ix = cell_coord(x, d.qmin_x, d.hmax);
    iy = cell_coord(y, d.qmin_y, d.hmax);
    key = (int64_t)ix + (int64_t)d.cells_x * (int64_t)iy;
  }
  const int64_t ncell = (int64_t)d.cells_x * d.cells_y;
  int k;
  if (isfinite(x) && isfinite(y) && key >= 0 && key < ncell) {
    k = (int)key;
    if (!keys_in && (ix < 0 || ix >= d.cells_x || iy < 0 || iy >= d.cells_y))
      atomicOr(&g.status[s], SPH_ST_OUTSIDE);
  } else {
    k = (int)ncell;
    atomicOr(&g.status[s], SPH_ST_OUTSIDE);
  }
  g.key[p] = k;
  kout[p] = d.cell_begin + k;
}

// Digit histogram of each tile: hist[digit * ntiles + tile].
__global__ __launch_bounds__(kBS) void k_rs_hist(const int* __restrict__ keys, int n, int shift,
                                                 int* __restrict__ hist) {
  __shared__ int cnt[kRsDigits];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRsTile;
#pragma unroll 4
  for (int u = 0; u < kRsItems; ++u) {
    const int i = base + u * kBS + t;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRsDigits - 1)], 1);
  }
  __syncthreads();
  hist[t * gridDim.x + blockIdx.x] = cnt[t];
}

// Lanes of this wavefront whose `d` equals this lane's, among `act`.
__device__ __forceinline__ unsigned long long match_digit(int d, unsigned long long act) {
  unsigned long long m = act;
#pragma unroll
  for (int b = 0; b < kRsBits; ++b) {
    const unsigned long long on = __ballot((d >> b) & 1);
    m &= ((d >> b) & 1) ? on : ~on;
  }
  return m;
}

// One pass: items of tile b move to off[digit * ntiles + b] + (their rank
// among the tile's items of that digit, in tile order).  Wavefront w ranks
// items [w * kRsTile / kRsWaves, (w + 1) * kRsTile / kRsWaves) of the tile,
// 64 consecutive ones per round, with a per-wavefront running count per
// digit.  vals_in == nullptr: the value is the item's index (first pass).
// LAST: the values are point indices, written to g.sorted_idx with their
// positions to g.sorted_x.
template <bool LAST>
__global__ __launch_bounds__(kBS) void k_rs_scatter(const int* __restrict__ keys_in,
                                                    const int* __restrict__ vals_in, int n,
                                                    int shift, const int* __restrict__ off,
                                                    int* __restrict__ keys_out,
                                                    int* __restrict__ vals_out, SphGrid g) {
  __shared__ int s_key[kRsTile];
  __shared__ int s_val[kRsTile];
  __shared__ int s_cnt[kRsWaves][kRsDigits];   // per-wavefront counts, then prefixes
  __shared__ int s_start[kRsDigits];           // tile-local first slot of each digit
  __shared__ int s_goff[kRsDigits];            // global first slot of the tile's digit run
  __shared__ int s_scan[kRsWaves + 1];
  const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
  for (int k = t; k < kRsWaves * kRsDigits; k += kBS) (&s_cnt[0][0])[k] = 0;
  __syncthreads();
  const int base = blockIdx.x * kRsTile + w * (kRsTile / kRsWaves);
  int key[kRsItems], val[kRsItems], rank[kRsItems];
#pragma unroll
  for (int u = 0; u < kRsItems; ++u) {
    const int i = base + u * kWave + lane;
    const bool ok = i < n;
    key[u] = ok ? keys_in[i] : 0;
    val[u] = ok ? (vals_in ? vals_in[i] : i) : 0;
  }
  const unsigned long long lower = (1ull << lane) - 1ull;
#pragma unroll
  for (int u = 0; u < kRsItems; ++u) {
    const bool ok = base + u * kWave + lane < n;
    const int d = (key[u] >> shift) & (kRsDigits - 1);
    const unsigned long long act = __ballot(ok);
    const unsigned long long peers = match_digit(d, act);
    const int before = s_cnt[w][d];
    rank[u] = before + __popcll(peers & lower);
    // the group's last lane carries the count forward
    if (ok && (peers >> lane) == 1ull) s_cnt[w][d] = before + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // digit t: prefixes over the wavefronts, the tile's total, its start in
  // the tile and in the pass's output
  int tot = 0;
#pragma unroll
  for (int v = 0; v < kRsWaves; ++v) {
    const int c = s_cnt[v][t];
    s_cnt[v][t] = tot;
    tot += c;
  }
  int all;
  const int start = block_exclusive_scan(tot, s_scan, &all);
  s_start[t] = start;
  s_goff[t] = off[t * gridDim.x + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kRsItems; ++u) {
    if (base + u * kWave + lane < n) {
      const int d = (key[u] >> shift) & (kRsDigits - 1);
      const int lp = s_start[d] + s_cnt[w][d] + rank[u];
      s_key[lp] = key[u];
      s_val[lp] = val[u];
    }
  }
  __syncthreads();
  const int tn = n - blockIdx.x * kRsTile < kRsTile ? n - blockIdx.x * kRsTile : kRsTile;
  for (int j = t; j < tn; j += kBS) {
    const int k = s_key[j], v = s_val[j];
    const int d = (k >> shift) & (kRsDigits - 1);
    const int gp = s_goff[d] + (j - s_start[d]);
    keys_out[gp] = k;
    if (LAST) {
      g.sorted_idx[gp] = v;
      reinterpret_cast<float2*>(g.sorted_x)[gp] = reinterpret_cast<const float2*>(g.x)[v];
    } else {
      vals_out[gp] = v;
    }
  }
}

// cell_count[c] = the length of cell c's run in the sorted keys (the
// caller zeroed cell_count), one thread per sorted slot, the run's first
// slot writing.
__global__ __launch_bounds__(kBS) void k_rs_runs(const int* __restrict__ skeys, int n,
                                                 int* __restrict__ cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = skeys[i];
  if (i > 0 && skeys[i - 1] == k) return;
  int e = i + 1;
  while (e < n && skeys[e] == k) ++e;
  cell_count[k] = e - i;
}

}  // namespace

int pair_thresholds(const float* h, int n, float* out, hipStream_t s) {
  SPH_REQUIRE(n >= 0 && (n == 0 || (h && out)), "sph_pair_thresholds: bad arguments");
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_pair_thresholds, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, h, n, out);
  return launch_check("sph_pair_thresholds");
}

static int grid_build_impl(const SphGrid* gp, const int* keys, void* ws, size_t ws_bytes,
                           hipStream_t s) {
  SPH_REQUIRE(gp, "sph_grid_build: null grid");
  const SphGrid& g = *gp;
  SPH_REQUIRE(g.np >= 0 && g.ncells > 0 && g.nscenes > 0, "sph_grid_build: bad sizes");
  // point arrays may be NULL for an empty point set
  SPH_REQUIRE(g.scenes && g.cell_count && g.cell_start && g.status &&
                  (g.np == 0 || (g.pscene && g.x && g.key && g.slot && g.sorted_idx && g.sorted_x)),
              "sph_grid_build: null buffer");
  int rc;
  if ((rc = fill_i32(g.cell_count, g.ncells, 0, s))) return rc;
  if (g.np > 0)
    hipLaunchKernelGGL(k_grid_keys, dim3(blocks_for(g.np, kBS)), dim3(kBS), 0, s, g, keys);
  if ((rc = launch_check("sph_grid_build(keys)"))) return rc;
  if ((rc = scan_exclusive(g.cell_count, g.ncells, g.cell_start, ws, ws_bytes, s))) return rc;
  if (g.np > 0) {
    const dim3 grid(blocks_for(g.np, kBS));
    hipLaunchKernelGGL(k_grid_scatter, grid, dim3(kBS), 0, s, g);
    hipLaunchKernelGGL(k_grid_rank, grid, dim3(kBS), 0, s, g);
    hipLaunchKernelGGL(k_grid_place, grid, dim3(kBS), 0, s, g);
  }
  return launch_check("sph_grid_build");
}

// Workspace of the radix-sort build: two key and two value buffers of np
// words, the digit histograms and their offsets (kRsDigits per tile), and
// the scan workspace of the larger of those and the cell table.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
static int rs_tiles(int np) { return (np + kRsTile - 1) / kRsTile; }
size_t grid_sort_ws_bytes(int np, int ncells) {
  if (np < 0 || ncells < 0) return 0;
  const int64_t hist = (int64_t)kRsDigits * rs_tiles(np);
  const int64_t scan_n = hist > ncells ? hist : ncells;
  return 4 * align256(4 * (size_t)(np > 0 ? np : 1)) + 2 * align256(4 * (size_t)(hist + 1)) +
         align256(scan_ws_bytes((int)scan_n));
}

int grid_build_sorted(const SphGrid* gp, const int* keys, void* ws, size_t ws_bytes,
                      hipStream_t s) {
  SPH_REQUIRE(gp, "sph_grid_build_sorted: null grid");
  const SphGrid& g = *gp;
  SPH_REQUIRE(g.np >= 0 && g.ncells > 0 && g.nscenes > 0, "sph_grid_build_sorted: bad sizes");
  SPH_REQUIRE(g.scenes && g.cell_count && g.cell_start && g.status &&
                  (g.np == 0 || (g.pscene && g.x && g.key && g.sorted_idx && g.sorted_x)),
              "sph_grid_build_sorted: null buffer");
  const size_t need = grid_sort_ws_bytes(g.np, g.ncells);
  if (!ws || ws_bytes < need) {
    set_error("sph_grid_build_sorted: workspace smaller than sph_grid_sort_workspace_bytes");
    return SPH_E_SPACE;
  }
  const int np = g.np;
  const int tiles = rs_tiles(np);
  const int64_t hist_n = (int64_t)kRsDigits * tiles;
  char* w = static_cast<char*>(ws);
  int* ka = reinterpret_cast<int*>(w);   w += align256(4 * (size_t)(np > 0 ? np : 1));
  int* va = reinterpret_cast<int*>(w);   w += align256(4 * (size_t)(np > 0 ? np : 1));
  int* kb = reinterpret_cast<int*>(w);   w += align256(4 * (size_t)(np > 0 ? np : 1));
  int* vb = reinterpret_cast<int*>(w);   w += align256(4 * (size_t)(np > 0 ? np : 1));
  int* hist = reinterpret_cast<int*>(w); w += align256(4 * (size_t)(hist_n + 1));
  int* off = reinterpret_cast<int*>(w);  w += align256(4 * (size_t)(hist_n + 1));
  void* sws = w;
  const size_t sws_bytes = need - (size_t)(w - static_cast<char*>(ws));
  int rc;
  if ((rc = fill_i32(g.cell_count, g.ncells, 0, s))) return rc;
  if (np > 0) {
    hipLaunchKernelGGL(k_rs_keys, dim3(blocks_for(np, kBS)), dim3(kBS), 0, s, g, keys, ka);
    if ((rc = launch_check("sph_grid_build_sorted(keys)"))) return rc;
    // digits of the largest global cell
    const int bits = 32 - __builtin_clz((unsigned)g.ncells);
    const int passes = (bits + kRsBits - 1) / kRsBits;
    const int* kin = ka;
    const int* vin = nullptr;
    for (int ps = 0; ps < passes; ++ps) {
      const int shift = ps * kRsBits;
      const bool last = ps == passes - 1;
      int* kout = (kin == ka) ? kb : ka;
      int* vout = (kin == ka) ? vb : va;
      hipLaunchKernelGGL(k_rs_hist, dim3(tiles), dim3(kBS), 0, s, kin, np, shift, hist);
      if ((rc = scan_exclusive(hist, (int)hist_n, off, sws, sws_bytes, s))) return rc;
      if (last)
        hipLaunchKernelGGL(k_rs_scatter<true>, dim3(tiles), dim3(kBS), 0, s, kin, vin, np, shift,
                           off, kout, vout, g);
      else
        hipLaunchKernelGGL(k_rs_scatter<false>, dim3(tiles), dim3(kBS), 0, s, kin, vin, np,
                           shift, off, kout, vout, g);
      if ((rc = launch_check("sph_grid_build_sorted(pass)"))) return rc;
      kin = kout;
      vin = vout;
    }
    hipLaunchKernelGGL(k_rs_runs, dim3(blocks_for(np, kBS)), dim3(kBS), 0, s, kin, np,
                       g.cell_count);
  }
  if ((rc = scan_exclusive(g.cell_count, g.ncells, g.cell_start, sws, sws_bytes, s))) return rc;
  return launch_check("sph_grid_build_sorted");
}

int grid_build(const SphGrid* gp, void* ws, size_t ws_bytes, hipStream_t s) {
  return grid_build_impl(gp, nullptr, ws, ws_bytes, s);
}

int grid_build_keyed(const SphGrid* gp, const int* keys, void* ws, size_t ws_bytes,
                     hipStream_t s) {
  SPH_REQUIRE(gp && (gp->np == 0 || keys), "sph_grid_build_keyed: null keys");
  return grid_build_impl(gp, keys, ws, ws_bytes, s);
}

static int check_query(const SphGrid* g, const SphQuery* q, const char* who) {
  if (!g || !q || q->nq < 0 || q->radius < 0 || (q->nq > 0 && (!q->x || !q->h || !q->qscene))) {
    set_error(std::string(who) + ": bad query");
    return SPH_E_ARG;
  }
  if (q->self_sorted && q->nq != g->np) {
    set_error(std::string(who) + ": self_sorted query must cover the grid's points");
    return SPH_E_ARG;
  }
  return SPH_OK;
}

// Self-sorted radius-1 queries (a scene's own points) are counted with the
// LDS-staged walk; foreign and wider queries by the per-query kernel.
static bool walks(const SphQuery* q) { return q->self_sorted && q->radius == 1; }

int neighbor_count(const SphGrid* g, const SphQuery* q, int* counts, hipStream_t s) {
  int rc = check_query(g, q, "sph_neighbor_count");
  if (rc) return rc;
  SPH_REQUIRE(counts, "sph_neighbor_count: null counts");
  if (q->nq == 0) return SPH_OK;
  if (walks(q)) {
    hipLaunchKernelGGL(k_walk_count, dim3(blocks_for(q->nq, kBS)), dim3(kBS), 0, s, *g, *q,
                       counts);
    return launch_check("sph_neighbor_count");
  }
  with_lanes(tuning(SPH_TUNE_SEARCH_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int bs = tuning(SPH_TUNE_BLOCK);
    hipLaunchKernelGGL((k_neighbors<false, KL>), dim3(blocks_for((int64_t)q->nq * KL, bs)),
                       dim3(bs), 0, s, *g, *q, (const int*)nullptr, counts, (int64_t)0);
  });
  return launch_check("sph_neighbor_count");
}

int neighbor_fill(const SphGrid* g, const SphQuery* q, int* rowptr, int* col, int64_t cap,
                  hipStream_t s) {
  int rc = check_query(g, q, "sph_neighbor_fill");
  if (rc) return rc;
  SPH_REQUIRE(rowptr && (col || cap == 0) && cap >= 0, "sph_neighbor_fill: bad output");
  if (q->nq == 0) return SPH_OK;
  with_lanes(tuning(SPH_TUNE_FILL_LANES), [&](auto kl) {
    constexpr int KL = decltype(kl)::value;
    const int bs = tuning(SPH_TUNE_BLOCK);
    hipLaunchKernelGGL((k_neighbors<true, KL>), dim3(blocks_for((int64_t)q->nq * KL, bs)),
                       dim3(bs), 0, s, *g, *q, rowptr, col, cap);
  });
  hipLaunchKernelGGL(k_clamp_rowptr, dim3(blocks_for(q->nq + 1, kBS)), dim3(kBS), 0, s, rowptr,
                     q->nq + 1, cap);
  return launch_check("sph_neighbor_fill");
}

int stencil_list_build(const SphGrid* g, const float* h, int* runs, unsigned long long* masks,
                       hipStream_t s) {
  SPH_REQUIRE(g && g->scenes && g->cell_start && g->sorted_idx && g->sorted_x && g->status &&
                  (g->nscenes == 1 || g->pscene) && g->np >= 0 && (g->np == 0 || (h && runs && masks)),
              "sph_stencil_list_build: bad arguments (the grid must be built)");
  if (g->np == 0) return SPH_OK;
  hipLaunchKernelGGL(k_stencil_list, dim3(blocks_for(g->np, kBS)), dim3(kBS), 0, s, *g, h,
                     reinterpret_cast<int4*>(runs), masks);
  return launch_check("sph_stencil_list_build");
}

int iota(int* a, int n, hipStream_t s) {
  SPH_REQUIRE(n >= 0 && (n == 0 || a), "sph_iota: bad arguments");
  if (n == 0) return SPH_OK;
  hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, kBS)), dim3(kBS), 0, s, a, n);
  return launch_check("sph_iota");
}

int csr_sort_rows(const SphCSR* a, int* col, hipStream_t s) {
  SPH_REQUIRE(a && a->rowptr && a->nrows >= 0, "sph_csr_sort_rows: bad csr");
  if (a->nrows == 0) return SPH_OK;
  hipLaunchKernelGGL(k_sort_rows, dim3(blocks_for(a->nrows, kBS)), dim3(kBS), 0, s, a->rowptr,
                     a->nrows, col);
  return launch_check("sph_csr_sort_rows");
}

int csr_transpose(const SphCSR* a, int ncols_t, int* t_counts, int* t_slot, int* t_rowptr,
                  int* t_col, void* ws, size_t ws_bytes, hipStream_t s) {
  SPH_REQUIRE(a && a->rowptr && a->nrows >= 0 && ncols_t >= 0 && t_counts && t_rowptr,
              "sph_csr_transpose: bad arguments");
  int rc;
  if ((rc = fill_i32(t_counts, ncols_t, 0, s))) return rc;
  if (a->nrows > 0)
    hipLaunchKernelGGL(k_tr_count, dim3(blocks_for(a->nrows, kBS)), dim3(kBS), 0, s, a->rowptr,
                       a->col, a->nrows, t_counts, t_slot);
  if ((rc = launch_check("sph_csr_transpose(count)"))) return rc;
  if ((rc = scan_exclusive(t_counts, ncols_t, t_rowptr, ws, ws_bytes, s))) return rc;
  if (a->nrows > 0)
    hipLaunchKernelGGL(k_tr_scatter, dim3(blocks_for(a->nrows, kBS)), dim3(kBS), 0, s,
                       a->rowptr, a->col, a->nrows, t_slot, t_rowptr, t_col);
  if ((rc = launch_check("sph_csr_transpose(scatter)"))) return rc;
  if (ncols_t > 0)
    hipLaunchKernelGGL(k_sort_rows, dim3(blocks_for(ncols_t, kBS)), dim3(kBS), 0, s, t_rowptr,
                       ncols_t, t_col);
  return launch_check("sph_csr_transpose");
}

int pair_geometry(const SphCSR* a, const float* xq, const float* xg, const float* hq,
                  float support, float thresh, int64_t* rows64, int64_t* cols64, float* rhat,
                  float* q, hipStream_t s) {
  SPH_REQUIRE(a && a->rowptr && xq && xg && a->nrows >= 0, "sph_pair_geometry: bad arguments");
  if (a->nrows == 0) return SPH_OK;
  hipLaunchKernelGGL(k_pair_geometry, dim3(blocks_for(a->nrows, kBS)), dim3(kBS), 0, s,
                     a->rowptr, a->col, a->nrows, xq, xg, hq, support, thresh, rows64, cols64,
                     rhat, q);
  return launch_check("sph_pair_geometry");
}

}  // namespace sph

extern "C" {
size_t sph_grid_workspace_bytes(int ncells) { return sph::scan_ws_bytes(ncells); }
int sph_grid_build(const SphGrid* g, void* ws, size_t ws_bytes, hipStream_t s) {
  return sph::grid_build(g, ws, ws_bytes, s);
}
int sph_grid_build_keyed(const SphGrid* g, const int* keys, void* ws, size_t ws_bytes,
                         hipStream_t s) {
  return sph::grid_build_keyed(g, keys, ws, ws_bytes, s);
}
size_t sph_grid_sort_workspace_bytes(int np, int ncells) {
  return sph::grid_sort_ws_bytes(np, ncells);
}
int sph_grid_build_sorted(const SphGrid* g, const int* keys, void* ws, size_t ws_bytes,
                          hipStream_t s) {
  return sph::grid_build_sorted(g, keys, ws, ws_bytes, s);
}
int sph_pair_thresholds(const float* h, int n, float* out, hipStream_t s) {
  return sph::pair_thresholds(h, n, out, s);
}
int sph_neighbor_count(const SphGrid* g, const SphQuery* q, int* counts, hipStream_t s) {
  return sph::neighbor_count(g, q, counts, s);
}
int sph_neighbor_fill(const SphGrid* g, const SphQuery* q, int* rowptr, int* col,
                      int64_t cap, hipStream_t s) {
  return sph::neighbor_fill(g, q, rowptr, col, cap, s);
}
int sph_csr_sort_rows(const SphCSR* a, int* col_mut, hipStream_t s) {
  return sph::csr_sort_rows(a, col_mut, s);
}
int sph_stencil_list_build(const SphGrid* g, const float* h, int* runs,
                           unsigned long long* masks, hipStream_t s) {
  return sph::stencil_list_build(g, h, runs, masks, s);
}
int sph_iota(int* a, int n, hipStream_t s) { return sph::iota(a, n, s); }
int sph_csr_transpose(const SphCSR* a, int ncols_t, int* t_counts, int* t_slot, int* t_rowptr,
                      int* t_col, void* ws, size_t ws_bytes, hipStream_t s) {
  return sph::csr_transpose(a, ncols_t, t_counts, t_slot, t_rowptr, t_col, ws, ws_bytes, s);
}
int sph_pair_geometry(const SphCSR* a, const float* xq, const float* xg, const float* hq,
                      float support, float thresh, int64_t* rows64, int64_t* cols64,
                      float* rhat, float* q, hipStream_t s) {
  return sph::pair_geometry(a, xq, xg, hq, support, thresh, rows64, cols64, rhat, q, s);
}
}
