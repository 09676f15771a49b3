/*
 * sph_mi355x.h -- C-ABI of the MI355X (gfx950) per-step particle hot path of
 * a 2D SPH solver: hashed-grid neighbour search into CSR pair lists, Wendland
 * kernel-sum gathers (delta-SPH density / continuity / diffusion / momentum),
 * the DFSPH relaxed-Jacobi density solve with an FP64 residual, and the
 * ghost-particle (mDBC) and Akinci boundary gathers.
 *
 * Conventions
 *  - Every entry point is batched-first: particles of many independent scenes
 *    live in one flat index space; `pscene[p]` names the scene of point p and
 *    per-scene parameters come from device arrays of SphGridDesc / SphPhysics.
 *    A single scene is nscenes == 1.
 *  - All pointers are device pointers unless the parameter says "host".
 *    Positions / velocities / accelerations are float[n][2] (x, y interleaved).
 *  - Nothing here allocates device memory.  Callers pass buffers together with
 *    their capacities; scratch needs are reported by *_workspace_bytes().
 *  - Every function launches on the caller's stream and never synchronises,
 *    so whole steps can be captured into a hipGraph.
 *  - Return value: SPH_OK or an SPH_E_* code; sph_last_error() explains the
 *    last failure (thread-local).  Data-dependent failures that can only be
 *    seen on the device (CSR overflow, particles outside the grid, NaN
 *    pressure, iteration cap) are OR-ed into a per-scene device status word.
 *
 * Reference interfaces replaced (file:line in wi-re/pytorchSPH):
 *   sph_grid_build            <- sortPositions + constructHashMap
 *                                (src/modules/neighborSearch.py:30-88)
 *   sph_neighbor_count/fill   <- buildNeighborList / buildNeighborListCUDA /
 *                                buildNeighborListUnsortedPerParticle
 *                                (cppSrc/neighSearch.cpp:143-331,
 *                                 cppSrc/neighSearch_cuda.cu:101-307)
 *   sph_csr_sort_rows + sph_pair_geometry
 *                             <- finalize sort + pair geometry
 *                                (src/modules/neighborSearch.py:153-157,216-250)
 *   sph_fluid_density         <- computeDensity (src/modules/density.py:26-53)
 *                                + akinci evalBoundaryDensity
 *                                (src/modules/akinciBoundary.py:395-420)
 *   sph_dsph_gather           <- computeDivergenceTerm (momentum.py:31-38),
 *                                computeDensityDiffusionMOG/DeltaSPH
 *                                (densityDiffusion.py:63-82),
 *                                computePressureAccel (pressure.py:31-38),
 *                                computeVelocityDiffusion (diffusion.py:30-44),
 *                                computeLaminarViscosity (laminar.py:30-46),
 *                                gravity (gravity.py:46-57) and the solid
 *                                boundary terms (solidBoundary.py:746-895)
 *   sph_dsph_normalization / sph_dsph_density_gradient
 *                             <- computeNormalizationMatrix + pinv2x2 +
 *                                computeRenormalizedDensityGradient
 *                                (densityDiffusion.py:30-61, util.py:386-433)
 *   sph_mdbc_density          <- computeBoundaryDensitymDBC
 *                                (solidBoundary.py:561-616)
 *   sph_dfsph_*               <- DFSPH / densitySolve and the akinci PBSPH
 *                                hooks (modules/dfsph.py:25-311,
 *                                akinciBoundary.py:166-389)
 *   sph_integrate_*           <- integrateValues / integrate
 *                                (simulationBase.py:564-569,632-705)
 *   sph_normalization_rows    <- computeNormalizationMatrix + pinv2x2 for
 *                                boundary rows (solidBoundary.py:676-693)
 *   sph_dplus_*               <- deltaPlusModule: detectSurface, shift
 *                                (modules/deltaPlus.py:81-251) and
 *                                computePressureAccelDeltaPlus
 *                                (pressure.py:40-52, solidBoundary.py:882-887)
 */
#ifndef SPH_MI355X_H
#define SPH_MI355X_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_OK 0
#define SPH_E_ARG 1      /* bad argument (null pointer, negative size, ...) */
#define SPH_E_HIP 2      /* a HIP runtime call failed                       */
#define SPH_E_SPACE 3    /* a caller buffer is smaller than required        */

/* per-scene device status bits */
#define SPH_ST_OUTSIDE   1u  /* a point lies outside the cell table: its key is out
                                of range or a coordinate is past a side edge */
#define SPH_ST_OVERFLOW  2u  /* CSR capacity exceeded: pair list truncated     */
#define SPH_ST_NAN       4u  /* DFSPH pressure became NaN / Inf                */
#define SPH_ST_ITERCAP   8u  /* DFSPH solve stopped at the iteration cap       */

/* Uniform grid of one scene.  Cells are hmax wide; a point's cell is
 * (ceil((x - qmin_x) / hmax), ceil((y - qmin_y) / hmax)) in fp32 with
 * correctly rounded subtraction and division, its key is
 * ix + cells_x * iy (neighborSearch.py:44-47). */
typedef struct SphGridDesc {
  float qmin_x, qmin_y;   /* minDomain - hmax                              */
  float hmax;             /* max support = cell size                      */
  int cells_x, cells_y;   /* ceil((qmax - qmin) / hmax), qmax = max + 2 hmax */
  int cell_begin;         /* first slot of this scene in the cell table; a
                             scene owns cells_x * cells_y + 1 slots, the last
                             one collecting points outside its table        */
  int p_begin, p_count;   /* this scene's point range                      */
} SphGridDesc;

/* Grid over a point set (many scenes).  All members are device pointers
 * except the sizes. */
typedef struct SphGrid {
  const SphGridDesc* scenes; /* [nscenes]                                  */
  const int* pscene;         /* [np] scene of each point                   */
  const float* x;            /* [np][2] point positions                    */
  int np, nscenes, ncells;   /* ncells = sum of (cells_x*cells_y + 1)      */
  int* cell_count;           /* [ncells] scratch                           */
  int* cell_start;           /* [ncells+1] first sorted slot of each cell  */
  int* key;                  /* [np] cell key of each point                */
  int* slot;                 /* [np] scratch                               */
  int* sorted_idx;           /* [np] point index in (scene, key, index) order */
  float* sorted_x;           /* [np][2] positions in sorted order          */
  unsigned* status;          /* [nscenes] status bits                      */
} SphGrid;

/* A set of query points searched against a grid. */
typedef struct SphQuery {
  const float* x;     /* [nq][2] query positions                           */
  const float* h;     /* [nq] query supports: pair iff |xq - xj| < h[q]    */
  const int* qscene;  /* [nq] scene of each query                          */
  int nq;
  int radius;         /* stencil half width in cells (searchRadius)        */
  int self_sorted;    /* 1: queries are the grid's own points; rows are
                         visited in the grid's sorted order for locality  */
  const int* qcell;   /* optional [nq] linear cell key of each query: its
                         cell is (qcell % cells_x, qcell / cells_x) in C
                         integer arithmetic, as the reference's CPU search
                         derives it from a cell's key (neighSearch.cpp:207-
                         209); NULL: the cell is computed from x            */
} SphQuery;

/* Physical constants of one scene (simulationBase.py:22-111 defaults). */
typedef struct SphPhysics {
  float h;          /* particle support (config particle.support)          */
  float rho0;       /* rest density                                        */
  float c0;         /* speed of sound                                      */
  float gamma;      /* Tait exponent                                       */
  float kappa;      /* compressible EOS stiffness                          */
  float alpha;      /* delta-SPH artificial viscosity                      */
  float delta;      /* density diffusion coefficient                       */
  float nu;         /* kinematic viscosity (laminar)                       */
  float eps;        /* 0.1 h^2                                             */
  float gx, gy;     /* gravity acceleration                                */
  float dt;         /* time step                                           */
  float beta;       /* akinci beta                                         */
  float omega;      /* boundary relaxed-Jacobi factor                      */
  float tau;        /* DFSPH density threshold                             */
  float fluid_omega;/* fluid relaxed-Jacobi factor (0.3 in the reference)  */
  int min_iters, max_iters;
  int eos;          /* 0 Tait, 1 deltaSPH (linear), 2 compressible         */
  int pad;
} SphPhysics;

/* delta-SPH gather terms (bitmask for sph_dsph_gather) */
#define SPH_T_CONT     (1u << 0)  /* continuity, fluid neighbours         */
#define SPH_T_CONT_B   (1u << 1)  /* continuity, boundary neighbours      */
#define SPH_T_MOG      (1u << 2)  /* MOG density diffusion                */
#define SPH_T_DDSPH    (1u << 3)  /* renormalised delta-SPH diffusion     */
#define SPH_T_PRESS    (1u << 4)  /* pressure acceleration, fluid         */
#define SPH_T_PRESS_B  (1u << 5)  /* pressure acceleration, boundary      */
#define SPH_T_VISC     (1u << 6)  /* artificial viscosity, fluid          */
#define SPH_T_VISC_B   (1u << 7)  /* artificial viscosity, boundary       */
#define SPH_T_LAM      (1u << 8)  /* laminar viscosity                    */
#define SPH_T_GRAV     (1u << 9)  /* gravity                              */
#define SPH_T_ACCUM    (1u << 16) /* add into dpdt/acc instead of overwrite */

/* Fluid state (device pointers, [np] or [np][2]). */
typedef struct SphFluid {
  const float* x; const float* v; const float* rho; /* rho = normalised density */
  const float* area; const float* rest;             /* area A, rest density   */
  const float* vol;   /* optional volume override; NULL means A / rho       */
  const int* pscene;
  int np;
} SphFluid;

/* Boundary particle state.  vol is the boundary volume (boundaryVolume);
 * rho is the normalised boundary density; p the boundary pressure. */
typedef struct SphBoundary {
  const float* x; const float* v; const float* rho; const float* vol;
  const float* p; const float* rest; const int* bscene;
  int nb;
} SphBoundary;

/* A CSR pair list: row r holds col[rowptr[r] .. rowptr[r+1]). */
typedef struct SphCSR {
  const int* rowptr; const int* col; int nrows;
} SphCSR;

/* ---------------------------------------------------------------- misc */
const char* sph_last_error(void);
int sph_version(void);
size_t sph_scan_workspace_bytes(int n);
/* out[0] = 0, out[k+1] = in[0] + ... + in[k]; out has n+1 entries. */
int sph_scan_exclusive(const int* in, int n, int* out, void* ws, size_t ws_bytes,
                       hipStream_t s);
/* sph_scan_exclusive without a workspace and without any allocation: the
 * tile partials live in a static device array of SPH_SCAN_NOALLOC_PARTIALS
 * ints per GPU, so n may be at most 2048 * (SPH_SCAN_NOALLOC_PARTIALS - 1)
 * = 134,215,680 items (larger n: SPH_E_SPACE, nothing launched).  The array
 * is shared by every call on that GPU: calls that overlap on one GPU (on
 * different streams, or from different host threads) race and must be
 * ordered by the caller.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc). */
#define SPH_SCAN_NOALLOC_PARTIALS 65536
int sph_scan_exclusive_noalloc(const int* in, int n, int* out, hipStream_t s);
int sph_fill_i32(int* p, int n, int value, hipStream_t s);
int sph_fill_f32(float* p, int n, float value, hipStream_t s);

/* Launch-shape knobs (process-wide; a captured graph keeps the values that
 * were set when it was captured).  The *_LANES knobs give the number of
 * lanes of a wavefront that share one row of a pair list (1, 2, 4, 8, 16 or
 * 32): the lanes walk the row's pairs together and combine their partial
 * sums with shuffles.  Pair sets do not depend on them; sums are re-associated
 * across the lanes.  No reference counterpart: the reference's gathers are
 * torch scatter_add over the pair list (util.py:11-40). */
#define SPH_TUNE_SEARCH_LANES 0   /* sph_neighbor_count                       */
#define SPH_TUNE_GATHER_LANES 1   /* density and delta-SPH gathers            */
#define SPH_TUNE_SOLVE_LANES  2   /* DFSPH alpha/source, iterate and finish   */
#define SPH_TUNE_FILL_LANES   3   /* sph_neighbor_fill                        */
#define SPH_TUNE_WALK_LANES   4   /* sph_dsph_cellwalk's list walk            */
#define SPH_TUNE_BLOCK        5   /* threads per workgroup (64, 128 or 256) of
                                     the list gathers (sph_fluid_density,
                                     sph_dsph_gather*, sph_dsph_gather_packed*)
                                     and the per-query search kernels        */
#define SPH_TUNE_COUNT        6
int sph_set_tuning(int knob, int value);
int sph_get_tuning(int knob);

/* -------------------------------------------------------- neighbour search */
/* Counting sort of the grid's points by (scene, cell key, index) and the
 * dense cell-start table.  Replaces the reference's hash map; the pair set
 * found through it is identical because a stencil cell (qx, qy) with
 * qx, qy >= 0 matches exactly the points whose key is qx + cells_x * qy. */
size_t sph_grid_workspace_bytes(int ncells);
int sph_grid_build(const SphGrid* g, void* ws, size_t ws_bytes, hipStream_t s);
/* sph_grid_build with the cell key of every point given (keys[np]) rather
 * than computed from its position: keys outside [0, cells_x * cells_y) go
 * to the scene's outside slot.  For callers that hold the reference's cell
 * structures (the linear keys of sortPositions, neighborSearch.py:44-48)
 * but not the grid origin they came from (buildNeighborList,
 * neighSearch.cpp:172-177, takes no qMin). */
int sph_grid_build_keyed(const SphGrid* g, const int* keys, void* ws, size_t ws_bytes,
                         hipStream_t s);
/* sph_grid_build (keys NULL) or sph_grid_build_keyed (keys given) for
 * points in any order: the same sorted_idx, sorted_x, cell_start, key and
 * status (every cell, its outside slot included, in index order), by a
 * stable least-significant-digit radix sort of (global cell, index) instead
 * of per-cell atomics and scatters, which serialise and write one word per
 * cache line when the points are not in cell order (a point cloud, the
 * reference's own particle order).  slot is not used.  ws holds
 * sph_grid_sort_workspace_bytes(np, ncells) bytes.  Replaces sortPositions
 * (neighborSearch.py:30-55) like sph_grid_build. */
size_t sph_grid_sort_workspace_bytes(int np, int ncells);
int sph_grid_build_sorted(const SphGrid* g, const int* keys, void* ws, size_t ws_bytes,
                          hipStream_t s);
/* The squared-distance threshold the search tests each support with:
 * out[i] = the largest float T with fl(sqrt(T)) < h[i], so that for every
 * float s >= 0, fl(sqrt(s)) < h[i] (the reference's predicate with a
 * correctly rounded sqrt, neighSearch.cpp:222-223) iff s <= out[i].
 * Computed in integer arithmetic; exact for 2^-62 <= h <= 2^62. */
int sph_pair_thresholds(const float* h, int n, float* out, hipStream_t s);
/* Number of pairs of each query (rows indexed by query number). */
int sph_neighbor_count(const SphGrid* g, const SphQuery* q, int* counts, hipStream_t s);
/* Writes col[rowptr[r] + k] = grid point index; cap = capacity of col.
 * Pairs past the capacity are dropped, the scene is flagged
 * SPH_ST_OVERFLOW and rowptr is clamped to cap so that every consumer of
 * the truncated list stays inside col. */
int sph_neighbor_fill(const SphGrid* g, const SphQuery* q, int* rowptr, int* col,
                      int64_t cap, hipStream_t s);
/* Stencil lists: the pairs of a grid's own points (self-sorted queries,
 * radius 1, supports h[point]) as bitmasks over each row's stencil instead
 * of a CSR column list.  Row r is sorted slot r; its three stencil rows
 * dy = q - 1 (q = 0, 1, 2) are the candidate runs of sorted slots
 * [runs[4r + q], runs[4r + q] + len_q), len_q = byte q of runs[4r + 3]
 * (at most 255), and bit k of masks[q n + r] is set when candidate
 * runs[4r + q] + k is a pair (the search predicate of sph_neighbor_fill).
 * The set bits, run by run, are the row of sph_neighbor_fill in its order.
 * A run of more than 64 candidates keeps its first 64 and flags its scene
 * SPH_ST_OVERFLOW.  12 + 4 + 24 bytes per row and no count / scan pass:
 * a neighbour list the DFSPH gathers walk with the chunk's stencil records
 * staged in LDS (sph_*_sl).  Replaces the fluid-fluid list of
 * buildNeighborList (neighSearch.cpp:172-259) for those gathers. */
typedef struct SphStencilList {
  const int* runs;                   /* [n][4]                             */
  const unsigned long long* masks;   /* [3][n]                             */
  int n;                             /* rows: the grid's np                */
} SphStencilList;
int sph_stencil_list_build(const SphGrid* g, const float* h, int* runs,
                           unsigned long long* masks, hipStream_t s);
/* a[i] = i for i < n: the grid's sorted_idx once its points have been
 * permuted into their sorted order (sph_permute_fields by sorted_idx). */
int sph_iota(int* a, int n, hipStream_t s);
/* Sorts the columns of every row ascending (lexicographic (i, j) order). */
int sph_csr_sort_rows(const SphCSR* a, int* col_mut, hipStream_t s);
/* Transposes an nrows x ncols_t CSR.  t_counts is [ncols_t] scratch,
 * t_slot is [nnz] scratch, t_rowptr is [ncols_t+1], t_col [nnz].  Rows of
 * the transpose are sorted ascending. */
int sph_csr_transpose(const SphCSR* a, int ncols_t, int* t_counts, int* t_slot,
                      int* t_rowptr, int* t_col, void* ws, size_t ws_bytes,
                      hipStream_t s);
/* Reference pair geometry: rows64/cols64 (int64 COO), rhat = (xq_i - x_j)/|.|
 * (0 below thresh), q = |xq_i - x_j| / hdiv, with hdiv = support (fluid-fluid)
 * or the query support when hq != NULL. */
int sph_pair_geometry(const SphCSR* a, const float* xq, const float* xg, const float* hq,
                      float support, float thresh, int64_t* rows64, int64_t* cols64,
                      float* rhat, float* q, hipStream_t s);

/* ------------------------------------------------------ kernel-sum gathers */
/* rho_out[p] = sum_j A_j W_pj (+ sum_b V_b W_pb when fb != NULL, the Akinci
 * boundary density contribution); vol_out[p] = A_p / (fluid-only sum). */
int sph_fluid_density(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphBoundary* b, const SphCSR* fb, const int* order,
                      float* rho_out, float* vol_out, hipStream_t s);
/* delta-SPH right-hand side of every fluid particle over its fluid (ff) and
 * boundary (fb) neighbours; terms = SPH_T_* mask. gradrho [np][2] is needed
 * for SPH_T_DDSPH.  Writes (or adds into, SPH_T_ACCUM) dpdt and acc. */
int sph_dsph_gather(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                    const SphBoundary* b, const SphCSR* fb, const float* gradrho,
                    unsigned terms, const int* order, float* dpdt, float* acc,
                    hipStream_t s);
/* Packs one 32-byte record per particle for sph_dsph_gather_packed:
 * frec[np][8] = {x, y, vx, vy, rho_hat rho0, V, p, rho_hat} with
 * V = A / rho_hat (or f->vol) and p = EOS(rho_hat); brec[nb][8] likewise for
 * the boundary with V = V_b / rho_b and p = b->p.  frec == NULL packs the
 * boundary records only. */
int sph_dsph_pack(const SphFluid* f, const SphPhysics* phys, const SphBoundary* b,
                  float* frec, float* brec, hipStream_t s);
/* sph_dsph_gather on packed records (same terms, same results). */
int sph_dsph_gather_packed(const float* frec, const float* brec, const int* pscene, int np,
                           const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                           const float* gradrho, unsigned terms, const int* order,
                           float* dpdt, float* acc, hipStream_t s);

/* sph_dsph_gather_packed under the delta-plus scheme: the pressure terms
 * (SPH_T_PRESS, SPH_T_PRESS_B) take p_j - p_i for a row with p_i < 0 and
 * surface_mask_i >= 1.5 (the bulk class of sph_dplus_surface), p_i + p_j
 * otherwise (computePressureAccelDeltaPlus, pressure.py:40-52 and
 * solidBoundary.py:882-887); every other term as sph_dsph_gather_packed. */
int sph_dsph_gather_packed_dplus(const float* frec, const float* brec, const int* pscene,
                                 int np, const SphPhysics* phys, const SphCSR* ff,
                                 const SphCSR* fb, const float* gradrho, unsigned terms,
                                 const int* order, const float* surface_mask, float* dpdt,
                                 float* acc, hipStream_t s);

/* One RK4 stage combined into the packed gather (simulationBase.py:679-705
 * applied to the rates of deltaSPH.py:146-200). */
typedef struct SphRk4Stage {
  int stage;            /* 1..4                                             */
  const float* x0;      /* [np][2] state at the start of the step           */
  const float* v0;      /* [np][2]                                          */
  const float* r0;      /* [np]                                             */
  const float* area;    /* [np] particle areas (for frec_next's volumes)    */
  float* K;             /* [np][5] running sums, as sph_integrate_rk4's kacc */
  float* x_out;         /* [np][2] next stage's state; the step's result at */
  float* v_out;         /*         stage 4 (may alias x0 / v0 / r0 there);  */
  float* rho_out;       /* [np]    v_out may be NULL at stages 1-3 when
                           frec_next is given (the next stage reads its
                           velocities from those records)                   */
  float* acc_out;       /* stage 4: summed acceleration (as acc_out of
                           sph_integrate_rk4), or NULL                      */
  float* frec_next;     /* stages 1-3: packed records of the next stage's
                           state (not frec), or NULL                        */
} SphRk4Stage;
/* sph_dsph_gather_packed followed, per particle, by sph_integrate_rk4 of the
 * stage (stage velocity = the record's) and, for stages 1-3, the fluid
 * records of the new state (as sph_dsph_pack with V = A / rho_hat): one
 * launch instead of three, no dp/dt or acceleration arrays.  The same
 * arithmetic as the three calls.  Replaces the integrator step around
 * deltaSPH.timestep (simulationBase.py:679-705); SPH_T_ACCUM not allowed. */
int sph_dsph_gather_rk4(const float* frec, const float* brec, const int* pscene, int np,
                        const SphPhysics* phys, const SphCSR* ff, const SphCSR* fb,
                        const float* gradrho, unsigned terms, const int* order,
                        const SphRk4Stage* rk, hipStream_t s);
/* sph_dsph_gather_packed with the fluid pairs found in the kernel from the
 * grid g (sph_grid_build of the records' positions) instead of read from a
 * list: the same pairs as sph_neighbor_fill of g's own points (radius 1,
 * pair iff fl(|x_i - x_j|) < phys[scene].h, which must not exceed the
 * scene's cell size), walked in the same order on the same lanes, so
 * dp/dt and a are bit for bit those of sph_dsph_gather_packed over that
 * list.  Replaces, for the delta-SPH right-hand side, the fluid pair list
 * of neighborSearch.py:126-157 (neighSearch_cuda.cu:214-307) together with
 * the gathers of momentum.py:31-38, densityDiffusion.py:74-82,
 * pressure.py:31-38, diffusion.py:30-44 and laminar.py:30-46.
 * segs[nseg][2] = {first, end} partitions the sorted slots [0, g->np) into
 * runs of one scene each, one workgroup per run; any partition gives the
 * same results, runs of at most 256 slots keep the stencil records in LDS.
 * Rows of frec are in point order (g->x's).  Boundary pairs come either
 * from the fluid -> boundary list fb (rows in point order) or, when bg is
 * given, from bg, a grid of the boundary points over the same scene
 * geometry (sph_grid_build of the boundary positions): then each row's
 * boundary pairs are found in the kernel, in the order of
 * sph_neighbor_fill for the fluid points as foreign queries against bg
 * (the solid-boundary search of solidBoundary.py:618-642), and are walked
 * as the packed gather walks that list.  SPH_T_DDSPH not allowed (its
 * gradient needs the list). */
int sph_dsph_cellwalk(const SphGrid* g, const int* segs, int nseg, const float* frec,
                      const float* brec, const SphPhysics* phys, const SphCSR* fb,
                      const SphGrid* bg, unsigned terms, float* dpdt, float* acc,
                      hipStream_t s);
/* sph_dsph_cellwalk followed by the RK4 stage, as sph_dsph_gather_rk4. */
int sph_dsph_cellwalk_rk4(const SphGrid* g, const int* segs, int nseg, const float* frec,
                          const float* brec, const SphPhysics* phys, const SphCSR* fb,
                          const SphGrid* bg, unsigned terms, const SphRk4Stage* rk,
                          hipStream_t s);
/* Renormalisation matrix M_i (fluid + boundary neighbours) and L_i =
 * pinv2x2(M_i); M, L are [np][2][2]. */
int sph_dsph_normalization(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                           const SphBoundary* b, const SphCSR* fb, const int* order,
                           float* M, float* L, float* eig, hipStream_t s);
/* Renormalised density gradient <grad rho>_i, [np][2]. */
int sph_dsph_density_gradient(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                              const SphBoundary* b, const SphCSR* fb, const float* L,
                              const int* order, float* gradrho, hipStream_t s);
/* Batched closed-form 2x2 pseudo-inverse (util.py:386-433). */
int sph_pinv2x2(const float* M, int n, float* L, float* eig, hipStream_t s);
/* mDBC boundary density from ghost->fluid pairs (gf rows = boundary index).
 * Writes rho_b, p_b (EOS) and neighbour counts. */
int sph_mdbc_density(const SphFluid* f, const SphPhysics* phys, const float* bx,
                     const float* gx, const int* bscene, int nb, const SphCSR* gf,
                     float* rho_b, float* p_b, int* counts, hipStream_t s);
/* sph_mdbc_density with each ghost's fluid neighbours found in the kernel
 * from the fluid grid g (sph_grid_build of f's positions) instead of read
 * from a list: the pairs of sph_neighbor_fill for the ghost queries gx with
 * supports gh (radius 1), in its order, so rho_b, p_b and counts are those
 * of sph_mdbc_density bit for bit.  Replaces the ghost search of
 * solidBoundary.py:618-642 and computeBoundaryDensitymDBC
 * (solidBoundary.py:561-616) in one launch. */
int sph_mdbc_density_walk(const SphFluid* f, const SphPhysics* phys, const float* bx,
                          const float* gx, const float* gh, const int* bscene, int nb,
                          const SphGrid* g, float* rho_b, float* p_b, int* counts,
                          hipStream_t s);
/* sph_mdbc_density_walk that also writes the boundary records brec [nb][8]
 * of sph_dsph_pack from the densities and pressures it computes ({x, y, vx,
 * vy, rho_b rho0, V / rho_b, p_b, rho_b}; bv [nb][2] may be NULL for walls
 * at rest, bvol [nb] holds the volumes V): the same values with no second
 * pass over the boundary.  Replaces, with sph_mdbc_density_walk, the
 * boundary density and pressure of computeBoundaryDensitymDBC
 * (solidBoundary.py:561-616) that the stage's boundary terms read. */
int sph_mdbc_density_walk_records(const SphFluid* f, const SphPhysics* phys, const float* bx,
                                  const float* bv, const float* bvol, const float* gx,
                                  const float* gh, const int* bscene, int nb, const SphGrid* g,
                                  float* rho_b, float* p_b, int* counts, float* brec,
                                  hipStream_t s);
/* EOS for n particles: p = EOS(rho). */
int sph_eos(const float* rho, const int* pscene, const SphPhysics* phys, int n, float* p,
            hipStream_t s);
/* Akinci boundary density rho_b = term_b + sum_f A_f W_bf + beta (bf rows =
 * boundary index). */
int sph_akinci_boundary_density(const SphFluid* f, const SphPhysics* phys, const float* bx,
                                const int* bscene, int nb, const float* term,
                                const SphCSR* bf, float* rho_b, hipStream_t s);

/* ------------------------------------------------------------------ DFSPH */
/* Solver state of all scenes.  Per-particle arrays are [np] / [np][2]
 * (fluid) or [nb] (boundary). */
typedef struct SphDfsph {
  float* vpred;  /* [np][2] predicted velocity                        */
  float* aact;   /* [np] actual area A / rho                          */
  float* alpha;  /* [np]                                              */
  float* src;    /* [np] source term                                  */
  float* p;      /* [np] pressure (previous iterate)                  */
  float* p2;     /* [np] pressure (current iterate)                   */
  float* apred;  /* [np][2] pressure acceleration (after finish)      */
  float* resid;  /* [np] residual                                     */
  float* xrec;   /* [np][4] {x, y, p2 / (rho rho0)^2, A rho0}         */
  float* arec;   /* [np][4] {a*_x, a*_y, dt^2 A^act, 0}               */
  float* prec;   /* [np][4] {x, y, w a*_x, w a*_y}, w = dt^2 A^act, as
                    written by sph_dfsph_accel and read by
                    sph_dfsph_pressure; from sph_dfsph_prepare until the
                    first accel it holds {v*, A^act, A rho0}            */
  float* wvec;   /* [np][2] W_i = dt^2 sum_j A^act_j gradWs_ij over fluid
                    and boundary neighbours (fixed during the solve)  */
  float* b_aact; /* [nb]                                              */
  float* b_alpha;/* [nb]                                              */
  float* b_src;  /* [nb]                                              */
  float* b_p;    /* [nb] boundary pressure                            */
  float* b_resid;/* [nb]                                              */
  float* bxrec;  /* [nb][4] {x, y, p_b / (rho_b rho0_b)^2, V_b rho0_b} */
  double* err_acc;  /* [SPH_DFSPH_ERR_SLOTS][SPH_DFSPH_ERR_STRIDE(nscenes)]
                       FP64 partial sums of clamped residuals (scene s
                       of slot k at k * stride + s); workgroups spread
                       their atomics over the slots                    */
  double* err_hist; /* [nscenes][max_hist] per-iteration errors        */
  int* iters;       /* [nscenes] iterations done                       */
  int* active;      /* [nscenes] 1 while the scene still iterates      */
  int* any_active;  /* [1]                                            */
  unsigned* status; /* [nscenes]                                       */
  const int* nfluid;    /* [nscenes] fluid count per scene             */
  const int* nbdy;      /* [nscenes] boundary count per scene          */
  const int* fbegin;    /* [nscenes] first fluid index of each scene   */
  const int* bbegin;    /* [nscenes] first boundary index of each scene */
  int* act_scene;   /* [nscenes] scenes still iterating                */
  int* act_pre;     /* [2 (nscenes + 1)] prefix sums of their fluid,
                       then boundary, counts                            */
  int* act_n;       /* [3] active fluid items, boundary items, scenes  */
  int nscenes, max_hist;
  int grid_blocks;  /* grid of the iteration kernels (work is strided) */
  int divergence;   /* 0: densitySolve (dfsph.py:195-234); 1: divergenceSolve
                       (dfsph.py:236-260): source = divergence only (no
                       1 - rho term), pressures not clamped, the residual
                       mean over fluid rows only; boundary arguments must be
                       NULL (the reference's Akinci hooks return zeros when
                       density = False, akinciBoundary.py:167,190,255,319,356).
                       phys carries the divergence threshold and iteration
                       bounds for that call                                */
} SphDfsph;

#define SPH_DFSPH_ERR_SLOTS 64
/* Row stride of err_acc in doubles: a multiple of 16 (128 bytes), so no two
 * slots share a cache line even for one scene. */
#define SPH_DFSPH_ERR_STRIDE(nscenes) ((((nscenes) + 15) / 16) * 16)

/* v* = v + dt a, A^act = A / rho; boundary A^act_b = V_b / rho_b, p_b = 0;
 * pressures zeroed; iteration counters reset; every scene active.  The
 * fluid of scene s must be [fbegin[s], fbegin[s] + nfluid[s]) and its
 * boundary [bbegin[s], bbegin[s] + nbdy[s]). */
int sph_dfsph_prepare(const SphFluid* f, const float* acc, const SphPhysics* phys,
                      const SphBoundary* b, SphDfsph* d, hipStream_t s);
/* alpha and source term of fluid and boundary particles. */
int sph_dfsph_alpha_source(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                           const SphBoundary* b, const SphCSR* fb, const SphCSR* bf,
                           SphDfsph* d, hipStream_t s);
/* One relaxed-Jacobi iteration of every still-active scene: pressure
 * acceleration from p2, then the pressure / residual update and the FP64
 * residual reduction, then the per-scene convergence decision and the new
 * active list.  Scenes that stopped are not touched. */
int sph_dfsph_iterate(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const SphBoundary* b, const SphCSR* fb, const SphCSR* bf,
                      SphDfsph* d, hipStream_t s);
/* The three launches of sph_dfsph_iterate, one entry point each, in the
 * order of the reference loop body (dfsph.py:205-230):
 *   sph_dfsph_accel    <- computeAcceleration (dfsph.py:186-193, 26-42)
 *   sph_dfsph_pressure <- computeUpdatedPressure + the clamped residual sum
 *                         (dfsph.py:164-184, 44-57, 228)
 *   sph_dfsph_control  <- the while condition of densitySolve (dfsph.py:205-207)
 * Between sph_dfsph_control calls, accel and pressure touch only the scenes
 * of the current active list. */
int sph_dfsph_accel(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                    const SphCSR* fb, SphDfsph* d, hipStream_t s);
int sph_dfsph_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const SphCSR* bf,
                       SphDfsph* d, hipStream_t s);
int sph_dfsph_control(const SphPhysics* phys, SphDfsph* d, hipStream_t s);
/* Final pressure acceleration, v* += dt a*, acc += a*, p = p2. */
int sph_dfsph_finish(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                     const SphCSR* fb, SphDfsph* d, float* acc, hipStream_t s);
/* sph_dfsph_finish and the velocity-diffusion / laminar gather that follows
 * it in a DFSPH step (vterms of SPH_T_VISC and SPH_T_LAM: diffusion.py:30-44,
 * laminar.py:30-46; sph_dsph_gather_packed with vterms | SPH_T_ACCUM over
 * ff) in one walk of the fluid list: acc = (acc + a*) + those terms, frec
 * the packed records of the step's state (sph_dsph_pack, V = A over the
 * fluid-only density).  With equal lanes per row for the solve and the
 * gathers, bit for bit the two calls.  No counterpart in the reference,
 * whose timestep evaluates the two modules one after the other
 * (dfsph.py:263-311). */
int sph_dfsph_finish_visc(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                          const SphCSR* fb, SphDfsph* d, const float* frec, unsigned vterms,
                          float* acc, hipStream_t s);

/* The DFSPH gathers over a stencil list sl of the fluid (sph_stencil_list_build
 * of the fluid grid) instead of the fluid-fluid CSR list ff: the same
 * arguments otherwise and the same results as the CSR entry points with one
 * lane per row (SPH_TUNE_SOLVE_LANES / SPH_TUNE_GATHER_LANES = 1), bit for
 * bit.  The fluid state must be stored in the grid's sorted order (point
 * index = sorted slot; sph_permute_fields by sorted_idx after the build,
 * then sph_iota on sorted_idx).  A workgroup stages the records of its rows'
 * stencil runs in LDS and walks the mask bits from there.  Boundary pairs
 * keep their CSR lists (fb, bf).  Each replaces its CSR form's reference
 * function (computeDensity density.py:26-53, dfsph.py:25-100,
 * 164-190, 263-311, akinciBoundary.py:166-389). */
int sph_fluid_density_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                         const SphBoundary* b, const SphCSR* fb, float* rho_out, float* vol_out,
                         hipStream_t s);
int sph_dfsph_alpha_source_sl(const SphFluid* f, const SphPhysics* phys,
                              const SphStencilList* sl, const SphBoundary* b, const SphCSR* fb,
                              const SphCSR* bf, SphDfsph* d, hipStream_t s);
int sph_dfsph_accel_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                       const SphCSR* fb, SphDfsph* d, hipStream_t s);
int sph_dfsph_pressure_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                          const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                          hipStream_t s);
int sph_dfsph_iterate_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                         const SphBoundary* b, const SphCSR* fb, const SphCSR* bf, SphDfsph* d,
                         hipStream_t s);
int sph_dfsph_finish_sl(const SphFluid* f, const SphPhysics* phys, const SphStencilList* sl,
                        const SphCSR* fb, SphDfsph* d, float* acc, hipStream_t s);
int sph_dfsph_finish_visc_sl(const SphFluid* f, const SphPhysics* phys,
                             const SphStencilList* sl, const SphCSR* fb, SphDfsph* d,
                             const float* frec, unsigned vterms, float* acc, hipStream_t s);

/* ------------------------------------------------------------- integrators */
/* RK4 stage s in 1..4 (simulationBase.py:679-705): accumulates
 * k = (acc, v_stage, dpdt) into kacc[np][5] and writes the next stage state
 * (x_out, v_out, rho_out) = perennial + c_s dt k; at s = 4 the final state
 * perennial + dt/6 sum(w k) and acc_out = sum(w k). */
int sph_integrate_rk4(int stage, const int* pscene, const SphPhysics* phys, int np,
                      const float* x0, const float* v0, const float* rho0,
                      const float* v_stage, const float* acc, const float* dpdt,
                      float* kacc, float* x_out, float* v_out, float* rho_out,
                      float* acc_out, hipStream_t s);
/* semi-implicit Euler: v = v0 + dt a; x = x0 + dt (v_stage + dt a);
 * rho = rho0 + dt dpdt / rho0 when dpdt != NULL. */
int sph_integrate_semi_implicit(const int* pscene, const SphPhysics* phys, int np,
                                const float* x0, const float* v0, const float* rho0,
                                const float* v_stage, const float* acc, const float* dpdt,
                                float* x_out, float* v_out, float* rho_out,
                                hipStream_t s);
/* Gathers arrays of 32-bit words by a permutation: out[k] = in[perm[k]]
 * for `width` words per element (perm == NULL copies).  Words are moved
 * bit for bit, so int32 / int64 arrays permute as well. */
int sph_permute_f32(const float* in, const int* perm, int n, int width, float* out,
                    hipStream_t s);

/* Up to SPH_PERMUTE_MAX_FIELDS arrays of 32-bit words permuted in place
 * together: field[k] = old field[perm[k]] for every field, through `tmp`
 * (n * sum of widths + 4 * nfields words).  Two launches whatever the
 * number of fields: one gathers every field of element perm[k] into tmp
 * (each field a block of its own there, starting on a 16-byte boundary),
 * one copies the blocks back.  The engine's per-step
 * move of the particle state into cell order; no counterpart in the
 * reference, which does not reorder. */
#define SPH_PERMUTE_MAX_FIELDS 8
typedef struct SphFields {
  float* ptr[SPH_PERMUTE_MAX_FIELDS];  /* device arrays [n][width[k]]          */
  int width[SPH_PERMUTE_MAX_FIELDS];   /* words per element, 1..4             */
  int nfields;
} SphFields;
int sph_permute_fields(const SphFields* f, const int* perm, int n, float* tmp, hipStream_t s);

/* ------------------------------------------------ delta-plus shifting */
/* Renormalisation matrices of arbitrary query rows over up to two pair
 * lists (computeNormalizationMatrix, densityDiffusion.py:30-43, as the solid
 * boundary calls it for its own rows, solidBoundary.py:676-693):
 *   M_q = sum_{j in a(q)} va_j (xa_j - xq_q) (x) gradW_qj
 *       + sum_{j in b(q)} vb_j (xb_j - xq_q) (x) gradW_qj,
 * L = pinv2x2(M), eig = its two singular values, larger first (util.py:
 * 386-433).  A list whose rowptr is NULL is skipped.  M and eig may be NULL;
 * M and L are 16-byte aligned.  qscene[q] selects the support. */
int sph_normalization_rows(const float* xq, const int* qscene, int nq, const SphPhysics* phys,
                           const SphCSR* a, const float* xa, const float* va,
                           const SphCSR* b, const float* xb, const float* vb,
                           float* M, float* L, float* eig, hipStream_t s);

/* Gradient of lambda and the fluid normal (deltaPlus.py:132-163).  lambda is
 * the smaller singular value of M: eig[i][1] for fluid rows, beig[b][1] for
 * boundary points.
 *   g_i = -sum_{j in ff(i)} V_j (lam_j - lam_i) L_i gradW_ij
 *         -sum_{b in fb(i)} bvol_n[b] (lam_b - lam_i) L_i gradW_ib,
 *   normal_i = g_i / (|g_i| + eps),
 * with V_j = f->vol[j] (or A_j / rho_j) and eps = 0.1 h^2.  bvol_n is the
 * volume the reference uses for boundary neighbour b: fluidArea[b] /
 * fluidDensity[b], the fluid arrays indexed by the boundary index
 * (deltaPlus.py:151); the caller passes it.  grad may be NULL; L ([np][2][2])
 * is 16-byte aligned. */
int sph_dplus_normal(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                     const float* L, const float* eig, const float* bx, const float* beig,
                     const float* bvol_n, const SphCSR* fb, float* grad, float* normal,
                     hipStream_t s);

/* Free-surface classification (deltaPlus.py:165-209).  Over the non-self
 * fluid pairs d_ij = (x_j - x_i)/|x_j - x_i| . n_i and over the boundary
 * pairs d_ib = (x_b - x_i)/|x_b - x_i| . n_i; c_i = max(max_j d_ij,
 * max_b d_ib), each max 0 when the row has no such pair; mask_i = 1 when
 * acos(c_i) > pi/6 and lam_i < 0.6, else 0; m2_i = max over non-self fluid
 * neighbours of mask_j.  surface_mask_i = 0 where mask_i = 1, else 1 where
 * m2_i = 1 and lam_i >= 0.5, else 2/3 where m2_i = 1, else 2.
 * mask_tmp is [np] scratch.  angle_min / angle_max (acos of the fluid-pair
 * min / max, 0 taken for a row without pairs) may be NULL. */
int sph_dplus_surface(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                      const float* normal, const float* eig, const float* bx, const SphCSR* fb,
                      float* mask_tmp, float* angle_min, float* angle_max,
                      float* surface_mask, hipStream_t s);

/* sph_dplus_normal followed by sph_dplus_surface in two launches: the
 * normal pass also runs the angle test of its row (same results). */
int sph_dplus_detect(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                     const float* L, const float* eig, const float* bx, const float* beig,
                     const float* bvol_n, const SphCSR* fb, float* grad, float* normal,
                     float* mask_tmp, float* angle_min, float* angle_max, float* surface_mask,
                     hipStream_t s);

/* Particle shift (deltaPlus.py:81-115 and 211-236):
 *   s_i = -1.5 Ma 4h^2 [ sum_{j in ff(i)} k_ij A_j / (rho_i + rho_j) gradW_ij
 *                      + sum_{b in fb(i)} k_ib V_b / (rho_i + rho_b) gradW_ib ],
 *   k_ij = 1 + 0.2 (W_ij / W(packing))^4,  Ma = max_k |v_k| / c0 over the
 * particle's scene (rho normalised, V_b = b->vol, rho_b = b->rho); then
 * |s_i| is clamped to 0.05 h, s_i <- (I - n_i n_i^T) s_i where
 * surface_mask_i < 1.5 and s_i <- 0 where surface_mask_i < 0.5.
 * packing[scene] is the lattice spacing over the support.  vmax_ws is
 * [nscenes] scratch words (cleared here by a kernel).  b / fb may be NULL. */
int sph_dplus_shift(const SphFluid* f, const SphPhysics* phys, int nscenes, const float* packing,
                    const SphCSR* ff, const SphBoundary* b, const SphCSR* fb,
                    const float* normal, const float* surface_mask, unsigned* vmax_ws,
                    float* shift, hipStream_t s);

/* delta-plus pressure acceleration (pressure.py:40-52; boundary part
 * solidBoundary.py:882-887): a_i = -(1/rho_i) sum_j p_ij V_j gradW_ij with
 * p_ij = p_i + p_j where p_i >= 0 or surface_mask_i < 1.5, else p_j - p_i;
 * p_i = EOS(rho_i), boundary p_b = b->p, V_b = b->vol / b->rho.  terms:
 * SPH_T_PRESS (fluid pairs), SPH_T_PRESS_B (boundary pairs), SPH_T_ACCUM
 * (add into acc). */
int sph_dplus_pressure(const SphFluid* f, const SphPhysics* phys, const SphCSR* ff,
                       const SphBoundary* b, const SphCSR* fb, const float* surface_mask,
                       unsigned terms, float* acc, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif
