// Torch operators over the C-ABI library (libsph_mi355x.so).
//
// TORCH_LIBRARY(sph, m) registers the reference's three neighbour-search
// entry points with the reference's argument lists and results
// (cppSrc/neighSearch.cpp:143-177 and 261-267, bound by pybind at :413-418,
// called from src/modules/neighborSearch.py:114,138,160,170), so that eager
// Python and TorchScript call them as torch.ops.sph.<name>.  The hash-table
// inputs (hashTable, hashMapLength) are accepted and not read: the search
// runs on the library's dense cell table, built from the reference's own
// cell keys (cellIndices / cumCell / cellSpan), so every point sits in the
// cell the reference put it in and the pair set is the reference's.
//
// Inputs on the host are moved to the current GPU and results moved back
// (the computation always runs on the GPU; there is no CPU path).  Each call
// reads two scalars back (the table height and the pair count), as the
// reference's CUDA path reads its pair count (neighSearch_cuda.cu:267).
// Results are lexicographically ordered (i, j) pairs, the order the
// reference's finalize step produces (neighborSearch.py:153-157).
#include <ATen/ATen.h>
#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <cstring>
#include <vector>

#include "sph_mi355x.h"

namespace {

void check(int rc, const char* what) {
  TORCH_CHECK(rc == SPH_OK, what, " failed: ", sph_last_error());
}

// The device the search runs on: the inputs' GPU, or the current one when
// the inputs live on the host.
at::Device compute_device(const at::Tensor& t) {
  if (t.is_cuda()) return t.device();
  return at::Device(at::kCUDA, c10::hip::current_device());
}

at::Tensor on(const at::Tensor& t, at::Device d, at::ScalarType st) {
  return t.to(d, st).contiguous();
}

void check_points(const at::Tensor& x, const at::Tensor& h, const char* xn, const char* hn) {
  TORCH_CHECK(x.dim() == 2 && x.size(1) == 2, xn, " must be [N, 2]");
  TORCH_CHECK(h.dim() == 1 && h.size(0) == x.size(0), hn, " must be [N] like ", xn);
  TORCH_CHECK(x.scalar_type() == at::kFloat && h.scalar_type() == at::kFloat,
              xn, " and ", hn, " must be float32");
}

struct CellArrays {
  at::Tensor keys;   // [N] int32 cell key of each sorted point
  int cells_y;       // table height covering every key
};

// Per-point cell keys from the reference's cell structures: cell c (in the
// hash order constructHashMap leaves them, neighborSearch.py:70-88) holds
// cumCell[c] points from sorted slot cellSpan[c] on, all with key
// cellIndices[c].
CellArrays point_keys(const at::Tensor& cellIndices, const at::Tensor& cumCell,
                      const at::Tensor& cellSpan, int64_t n, int cells_x, at::Device d) {
  TORCH_CHECK(cellIndices.dim() == 1 && cumCell.sizes() == cellIndices.sizes() &&
                  cellSpan.sizes() == cellIndices.sizes(),
              "cellIndices, cumCell and cellSpan must be [C] alike");
  TORCH_CHECK(cells_x > 0, "cellsX must be positive");
  const auto L = at::TensorOptions().dtype(at::kLong).device(d);
  at::Tensor ck = on(cellIndices, d, at::kLong);
  at::Tensor cnt = on(cumCell, d, at::kLong);
  at::Tensor st = on(cellSpan, d, at::kLong);
  const int64_t C = ck.size(0);
  at::Tensor keys = at::empty({n}, L.dtype(at::kInt));
  int64_t kmax = -1;
  if (C > 0) {
    // one host read: the point total and the largest key (table height)
    at::Tensor stats = at::stack({cnt.sum(), ck.max()}).cpu();
    const int64_t total = stats[0].item<int64_t>();
    kmax = stats[1].item<int64_t>();
    TORCH_CHECK(total == n, "cumCell sums to ", total, " points, sortedParticles holds ", n);
    at::Tensor cell_of = at::repeat_interleave(at::arange(C, L), cnt, 0, n);
    at::Tensor first = at::cumsum(cnt, 0) - cnt;
    at::Tensor pos = st.index_select(0, cell_of) + at::arange(n, L) - first.index_select(0, cell_of);
    keys.index_put_({pos}, ck.index_select(0, cell_of).to(at::kInt));
  } else {
    TORCH_CHECK(n == 0, "no cells for ", n, " points");
  }
  const int64_t cy = kmax < 0 ? 1 : kmax / cells_x + 2;
  TORCH_CHECK(cy * (int64_t)cells_x + 1 < (int64_t)INT32_MAX, "cell table exceeds int32 range");
  return {keys, (int)cy};
}

struct Pairs {
  at::Tensor rowptr;   // [Q + 1] int32
  at::Tensor counts;   // [Q] int32
  at::Tensor col;      // [nnz] int32, grid point (sorted slot) of each pair
  int64_t nnz;
};

// Queries xq / hq against the grid of points xg whose keys are given.
// qcell: optional per-query keys (the reference's CPU self search takes a
// query's cell from its key).  self_sorted: the queries are the grid's
// points.
Pairs run_search(const at::Tensor& xg, const at::Tensor& gkeys, const at::Tensor& xq,
                 const at::Tensor& hq, const at::Tensor* qcell, float qmin_x, float qmin_y,
                 float hmax, int cells_x, int cells_y, int radius, bool self_sorted,
                 at::Device d) {
  TORCH_CHECK(radius >= 0, "searchRadius must be >= 0");
  const int N = (int)xg.size(0), Q = (int)xq.size(0);
  const int ncells = cells_x * cells_y + 1;
  const auto I = at::TensorOptions().dtype(at::kInt).device(d);
  hipStream_t s = c10::hip::getCurrentHIPStream(d.index()).stream();

  SphGridDesc desc{qmin_x, qmin_y, hmax, cells_x, cells_y, 0, 0, N};
  at::Tensor desc_h = at::empty({(int64_t)sizeof(desc)}, at::kByte);
  std::memcpy(desc_h.data_ptr(), &desc, sizeof(desc));
  at::Tensor desc_d = desc_h.to(d);
  at::Tensor pscene = at::zeros({std::max(N, 1)}, I);
  at::Tensor qscene = at::zeros({std::max(Q, 1)}, I);
  at::Tensor cell_count = at::empty({ncells}, I);
  at::Tensor cell_start = at::empty({ncells + 1}, I);
  at::Tensor key = at::empty({std::max(N, 1)}, I);
  at::Tensor slot = at::empty({std::max(N, 1)}, I);
  at::Tensor sorted_idx = at::empty({std::max(N, 1)}, I);
  at::Tensor sorted_x = at::empty({std::max(N, 1), 2}, I.dtype(at::kFloat));
  at::Tensor status = at::zeros({1}, I);
  const size_t gws = sph_grid_workspace_bytes(ncells);
  at::Tensor ws = at::empty({(int64_t)gws}, I.dtype(at::kByte));
  SphGrid g{reinterpret_cast<const SphGridDesc*>(desc_d.data_ptr()), pscene.data_ptr<int>(),
            xg.data_ptr<float>(), N, 1, ncells, cell_count.data_ptr<int>(),
            cell_start.data_ptr<int>(), key.data_ptr<int>(), slot.data_ptr<int>(),
            sorted_idx.data_ptr<int>(), sorted_x.data_ptr<float>(),
            reinterpret_cast<unsigned*>(status.data_ptr<int>())};
  check(sph_grid_build_keyed(&g, gkeys.data_ptr<int>(), ws.data_ptr(), gws, s),
        "sph_grid_build_keyed");
  SphQuery q{xq.data_ptr<float>(), hq.data_ptr<float>(), qscene.data_ptr<int>(), Q, radius,
             self_sorted ? 1 : 0, qcell ? qcell->data_ptr<int>() : nullptr};
  Pairs p;
  p.counts = at::empty({std::max(Q, 1)}, I);
  p.rowptr = at::empty({Q + 1}, I);
  check(sph_neighbor_count(&g, &q, p.counts.data_ptr<int>(), s), "sph_neighbor_count");
  const size_t sws = sph_scan_workspace_bytes(Q);
  at::Tensor ws2 = at::empty({(int64_t)sws}, I.dtype(at::kByte));
  check(sph_scan_exclusive(p.counts.data_ptr<int>(), Q, p.rowptr.data_ptr<int>(), ws2.data_ptr(),
                           sws, s),
        "sph_scan_exclusive");
  p.nnz = p.rowptr[Q].item<int>();
  p.col = at::empty({std::max<int64_t>(p.nnz, 1)}, I);
  check(sph_neighbor_fill(&g, &q, p.rowptr.data_ptr<int>(), p.col.data_ptr<int>(), p.nnz, s),
        "sph_neighbor_fill");
  p.counts = p.counts.narrow(0, 0, Q);
  return p;
}

// Row-ordered pairs (rows ascending): maps the columns from sorted slots to
// original indices through `sort` and orders each row's columns.
std::vector<at::Tensor> rows_by_query(Pairs& p, const at::Tensor& sort, at::Device d,
                                      bool to_host) {
  const int Q = (int)p.counts.size(0);
  const auto L = at::TensorOptions().dtype(at::kLong).device(d);
  at::Tensor rows = at::repeat_interleave(at::arange(Q, L), p.counts.to(at::kLong), 0, p.nnz)
                        .to(at::kInt);
  at::Tensor cols = sort.index_select(0, p.col.narrow(0, 0, p.nnz).to(at::kLong))
                        .to(at::kInt).contiguous();
  if (p.nnz > 0) {
    SphCSR c{p.rowptr.data_ptr<int>(), cols.data_ptr<int>(), Q};
    check(sph_csr_sort_rows(&c, cols.data_ptr<int>(),
                            c10::hip::getCurrentHIPStream(d.index()).stream()),
          "sph_csr_sort_rows");
  }
  if (to_host) return {rows.cpu(), cols.cpu()};
  return {rows, cols};
}

struct Prepared {
  at::Device d;
  bool host;
  at::Tensor sx, sh, sort;
  CellArrays cells;
};

Prepared prepare(const at::Tensor& sortedParticles, const at::Tensor& sortedSupport,
                 const at::Tensor& cellIndices, const at::Tensor& cumCell,
                 const at::Tensor& cellSpan, const at::Tensor& sort, int64_t cellsX) {
  check_points(sortedParticles, sortedSupport, "sortedParticles", "sortedSupport");
  TORCH_CHECK(sort.dim() == 1 && sort.size(0) == sortedParticles.size(0),
              "sort must be [N] like sortedParticles");
  const at::Device d = compute_device(sortedParticles);
  const bool host = !sortedParticles.is_cuda();
  at::Tensor sx = on(sortedParticles, d, at::kFloat);
  at::Tensor sh = on(sortedSupport, d, at::kFloat);
  at::Tensor so = on(sort, d, at::kLong);
  CellArrays cells = point_keys(cellIndices, cumCell, cellSpan, sx.size(0), (int)cellsX, d);
  return {d, host, sx, sh, so, cells};
}

// buildNeighborListCUDA (neighSearch.cpp:143-170): queries in their own
// order against the sorted grid; i = query index, j = sort[slot], pair iff
// |x_i - x_j| < support_i.
std::vector<at::Tensor> buildNeighborListCUDA(
    at::Tensor queryParticles, at::Tensor support, at::Tensor sortedParticles,
    at::Tensor sortedSupport, at::Tensor hashTable, at::Tensor cellIndices, at::Tensor cumCell,
    at::Tensor cellSpan, at::Tensor sort, at::Tensor qMin, double hMax, int64_t cellsX,
    int64_t hashMapLength, int64_t searchRadius) {
  (void)hashTable;
  (void)hashMapLength;
  check_points(queryParticles, support, "queryParticles", "support");
  TORCH_CHECK(qMin.numel() == 2, "qMin must hold 2 values");
  Prepared P = prepare(sortedParticles, sortedSupport, cellIndices, cumCell, cellSpan, sort, cellsX);
  c10::hip::HIPGuard guard(P.d.index());
  at::Tensor xq = on(queryParticles, P.d, at::kFloat);
  at::Tensor hq = on(support, P.d, at::kFloat);
  at::Tensor qm = qMin.to(at::kCPU, at::kFloat).contiguous();
  const float* qmp = qm.data_ptr<float>();
  Pairs p = run_search(P.sx, P.cells.keys, xq, hq, nullptr, qmp[0], qmp[1], (float)hMax,
                       (int)cellsX, P.cells.cells_y, (int)searchRadius, false, P.d);
  return rows_by_query(p, P.sort, P.d, P.host);
}

// buildNeighborListUnsortedPerParticle (neighSearch.cpp:261-331): the same
// search with the reference's other argument order.
std::vector<at::Tensor> buildNeighborListUnsortedPerParticle(
    at::Tensor inputParticles, at::Tensor inputSupport, at::Tensor queryParticles,
    at::Tensor support, at::Tensor hashTable, at::Tensor cellIndices, at::Tensor cumCell,
    at::Tensor cellSpan, at::Tensor sort, int64_t cellsX, int64_t hashMapLength, at::Tensor qMin,
    double hMax, int64_t searchRadius) {
  return buildNeighborListCUDA(inputParticles, inputSupport, queryParticles, support, hashTable,
                               cellIndices, cumCell, cellSpan, sort, qMin, hMax, cellsX,
                               hashMapLength, searchRadius);
}

// buildNeighborList (neighSearch.cpp:172-259): the sorted points searched
// against themselves, each point's cell taken from its key (:207-209);
// rows = sort[i], cols = sort[j], pair iff |x_i - x_j| < support_i.
std::vector<at::Tensor> buildNeighborList(at::Tensor queryParticles, at::Tensor support,
                                          at::Tensor hashTable, at::Tensor cellIndices,
                                          at::Tensor cumCell, at::Tensor cellSpan,
                                          at::Tensor sort, int64_t cellsX, int64_t hashMapLength,
                                          int64_t searchRadius) {
  (void)hashTable;
  (void)hashMapLength;
  Prepared P = prepare(queryParticles, support, cellIndices, cumCell, cellSpan, sort, cellsX);
  c10::hip::HIPGuard guard(P.d.index());
  // no grid origin is passed: every cell comes from a key, so origin and
  // cell size only shape the (unused) position-to-cell mapping
  Pairs p = run_search(P.sx, P.cells.keys, P.sx, P.sh, &P.cells.keys, 0.f, 0.f, 1.f,
                       (int)cellsX, P.cells.cells_y, (int)searchRadius, true, P.d);
  // rows indexed by sorted slot: map both ends, then order (i, j)
  const int64_t n = P.sx.size(0);
  const auto L = at::TensorOptions().dtype(at::kLong).device(P.d);
  at::Tensor slot_rows =
      at::repeat_interleave(at::arange(p.counts.size(0), L), p.counts.to(at::kLong), 0, p.nnz);
  at::Tensor r = P.sort.index_select(0, slot_rows);
  at::Tensor c = P.sort.index_select(0, p.col.narrow(0, 0, p.nnz).to(at::kLong));
  at::Tensor key = std::get<0>(at::sort(r * n + c));
  at::Tensor rows = at::div(key, n, "trunc").to(at::kInt);
  at::Tensor cols = at::remainder(key, n).to(at::kInt);
  if (P.host) return {rows.cpu(), cols.cpu()};
  return {rows, cols};
}

}  // namespace

TORCH_LIBRARY(sph, m) {
  m.def("buildNeighborListCUDA(Tensor queryParticles, Tensor support, Tensor sortedParticles, "
        "Tensor sortedSupport, Tensor hashTable, Tensor cellIndices, Tensor cumCell, "
        "Tensor cellSpan, Tensor sort, Tensor qMin, float hMax, int cellsX, int hashMapLength, "
        "int searchRadius) -> Tensor[]",
        &buildNeighborListCUDA);
  m.def("buildNeighborList(Tensor queryParticles, Tensor support, Tensor hashTable, "
        "Tensor cellIndices, Tensor cumCell, Tensor cellSpan, Tensor sort, int cellsX, "
        "int hashMapLength, int searchRadius) -> Tensor[]",
        &buildNeighborList);
  m.def("buildNeighborListUnsortedPerParticle(Tensor inputParticles, Tensor inputSupport, "
        "Tensor queryParticles, Tensor support, Tensor hashTable, Tensor cellIndices, "
        "Tensor cumCell, Tensor cellSpan, Tensor sort, int cellsX, int hashMapLength, "
        "Tensor qMin, float hMax, int searchRadius) -> Tensor[]",
        &buildNeighborListUnsortedPerParticle);
  m.def("version() -> int", []() -> int64_t { return sph_version(); });
}
