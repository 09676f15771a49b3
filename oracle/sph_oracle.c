/*
 * CPU oracle of the pytorchSPH per-step hot path.  TEST INFRASTRUCTURE ONLY:
 * it is used by tests/, by the smoke check and by the CPU-baseline leg of
 * bench.py as the checker; the product path never calls it.
 *
 *  - Neighbour search: a restatement of the reference's hashed-grid search in
 *    fp32 with the reference's exact arithmetic (cell = ceil((x - qMin) /
 *    hMax), key = ix + cellsX * iy, pair iff sqrtf(dx*dx + dy*dy) < h_i),
 *    compiled with -ffp-contract=off.  Cells are looked up by key equality,
 *    the semantics of queryHashMap, through binary search on sorted keys.
 *      self search    : neighSearch.cpp:172-259 (buildNeighborList)
 *      foreign queries: neighSearch.cpp:261-331 (UnsortedPerParticle)
 *      cell keys      : neighborSearch.py:30-55 (sortPositions)
 *    Output rows are in query order with columns ascending, i.e. the
 *    lexicographic (i, j) order of neighborSearch.py:153-157.
 *  - Physics: FP64 restatements of every per-pair sum on the path, each
 *    function citing the reference lines it follows.  State arrays are
 *    passed as double; pair lists are CSR with int64 row pointers.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 *
 * How each part is pinned (tests/test_oracle.py, test_oracle_pins.py,
 * test_oracle_dplus.py, test_oracle_ref.py): the search by the reference's
 * own compiled search and brute force; every physics term by closed forms,
 * symmetries and independent computations that a dropped factor, a flipped
 * sign or a wrong index fails (Laplacians of quadratic densities, gradients
 * of linear pressures, exact mDBC extrapolation of linear fields, dry-wall
 * Akinci densities, two-particle viscosity and alpha, momentum
 * conservation, the predicted-density error a converged DFSPH solve must
 * meet, the RK4 and symplectic-Euler step maps); pinv2x2 and pinv3x3 by
 * textbook pseudo-inverses on their special cases and across the cutoff.
 *
 * Parity unpinned: the fp32 .item() iteration sequence of densitySolve
 * (dfsph.py:205-231).  orc_dfsph sums the residual in FP64 and may stop one
 * iteration earlier or later than an fp32 mean near the threshold; only an
 * import of the reference could pin which.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------ structures */
typedef struct OrcCSR {
  int nrows;
  const int64_t* rowptr;
  const int32_t* col;
} OrcCSR;

typedef struct OrcFluid {
  int n;
  const double *x, *v, *rho, *area, *rest;
  const double* vol; /* optional volume override (NULL: A / rho) */
} OrcFluid;

typedef struct OrcBdy {
  int n;
  const double *x, *v, *rho, *vol, *p, *rest, *gx;
} OrcBdy;

typedef struct OrcPhys {
  double h, rho0, c0, gamma, kappa, alpha, delta, nu, eps, gx, gy, dt, beta, omega, tau,
      fluid_omega;
  int min_iters, max_iters, eos, pad;
} OrcPhys;

/* term bits, same values as the device library */
#define T_CONT (1u << 0)
#define T_CONT_B (1u << 1)
#define T_MOG (1u << 2)
#define T_DDSPH (1u << 3)
#define T_PRESS (1u << 4)
#define T_PRESS_B (1u << 5)
#define T_VISC (1u << 6)
#define T_VISC_B (1u << 7)
#define T_LAM (1u << 8)
#define T_GRAV (1u << 9)

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ====================================================== neighbour search */
typedef struct Grid {
  int np;
  int32_t* key;   /* sorted keys */
  int32_t* perm;  /* point index of each sorted slot */
  float* xs;      /* sorted positions */
  float qmin_x, qmin_y, hmax;
  int cells_x;
} Grid;

static int32_t cell_of(float x, float qmin, float hmax) {
  return (int32_t)ceilf((x - qmin) / hmax);
}

static int32_t key_of(float x, float y, const Grid* g) {
  int64_t ix = cell_of(x, g->qmin_x, g->hmax);
  int64_t iy = cell_of(y, g->qmin_y, g->hmax);
  return (int32_t)(ix + (int64_t)g->cells_x * iy);
}

static const int32_t* g_sort_keys;
static int cmp_slot(const void* a, const void* b) {
  int32_t ia = *(const int32_t*)a, ib = *(const int32_t*)b;
  int32_t ka = g_sort_keys[ia], kb = g_sort_keys[ib];
  if (ka != kb) return ka < kb ? -1 : 1;
  return ia < ib ? -1 : (ia > ib);
}

static void grid_make(Grid* g, const float* xp, int np, float qmin_x, float qmin_y, float hmax,
                      int cells_x) {
  g->np = np;
  g->qmin_x = qmin_x;
  g->qmin_y = qmin_y;
  g->hmax = hmax;
  g->cells_x = cells_x;
  int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
  g->perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
  g->key = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
  g->xs = (float*)malloc(sizeof(float) * 2 * (size_t)(np > 0 ? np : 1));
#pragma omp parallel for schedule(static)
  for (int p = 0; p < np; ++p) {
    keys[p] = key_of(xp[2 * p], xp[2 * p + 1], g);
    g->perm[p] = p;
  }
  g_sort_keys = keys;
  qsort(g->perm, (size_t)np, sizeof(int32_t), cmp_slot);
  for (int k = 0; k < np; ++k) {
    g->key[k] = keys[g->perm[k]];
    g->xs[2 * k] = xp[2 * g->perm[k]];
    g->xs[2 * k + 1] = xp[2 * g->perm[k] + 1];
  }
  free(keys);
}

static void grid_free(Grid* g) {
  free(g->key);
  free(g->perm);
  free(g->xs);
}

/* range of sorted slots whose key equals k (queryHashMap semantics) */
static void key_range(const Grid* g, int32_t k, int* a, int* b) {
  int lo = 0, hi = g->np;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (g->key[m] < k) lo = m + 1; else hi = m;
  }
  *a = lo;
  hi = g->np;
  while (lo < hi) {
    int m = (lo + hi) >> 1;
    if (g->key[m] <= k) lo = m + 1; else hi = m;
  }
  *b = lo;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return x < y ? -1 : (x > y);
}

/* Visits the pairs of query q.  Returns the count; writes columns if out. */
static int64_t query_pairs(const Grid* g, float xq, float yq, float hq, int32_t qx, int32_t qy,
                           int R, int32_t* out) {
  int64_t cnt = 0;
  for (int xx = -R; xx <= R; ++xx) {
    for (int yy = -R; yy <= R; ++yy) {
      int32_t cx = qx + xx, cy = qy + yy;
      if (cx < 0 || cy < 0) continue; /* neighSearch.cpp:95 */
      int32_t lin = (int32_t)((int64_t)cx + (int64_t)g->cells_x * cy);
      int a, b;
      key_range(g, lin, &a, &b);
      for (int j = a; j < b; ++j) {
        float dx = xq - g->xs[2 * j];
        float dy = yq - g->xs[2 * j + 1];
        float dist = sqrtf(dx * dx + dy * dy);
        if (dist < hq) {
          if (out) out[cnt] = g->perm[j];
          ++cnt;
        }
      }
    }
  }
  return cnt;
}

/*
 * Pair list of nq queries against np grid points.
 * self_mode = 1: the queries are the grid points themselves (xq == xp) and
 * the query cell is the decomposition of the point's own key
 * (neighSearch.cpp:204-207); otherwise the query cell is computed from the
 * query position (neighSearch.cpp:286-287).
 * Call with col == NULL to get row counts as an exclusive scan in rowptr
 * (nq + 1 entries; returns nnz); call again with col to fill.
 */
int64_t orc_search(const float* xq, const float* hq, int nq, const float* xp, int np,
                   float qmin_x, float qmin_y, float hmax, int cells_x, int R, int self_mode,
                   int64_t* rowptr, int32_t* col) {
  Grid g;
  grid_make(&g, xp, np, qmin_x, qmin_y, hmax, cells_x);
  int64_t* cnt = NULL;
  if (!col) cnt = (int64_t*)calloc((size_t)nq + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
  for (int q = 0; q < nq; ++q) {
    const float x = xq[2 * q], y = xq[2 * q + 1];
    int32_t qx, qy;
    if (self_mode) {
      int32_t k = key_of(x, y, &g);
      qx = k % cells_x;
      qy = k / cells_x;
      if (k < 0) { /* own cell lookup fails: no pairs (neighSearch.cpp:208) */
        if (cnt) cnt[q] = 0;
        continue;
      }
    } else {
      qx = cell_of(x, qmin_x, hmax);
      qy = cell_of(y, qmin_y, hmax);
    }
    if (!col) {
      cnt[q] = query_pairs(&g, x, y, hq[q], qx, qy, R, NULL);
    } else {
      int32_t* out = col + rowptr[q];
      int64_t c = query_pairs(&g, x, y, hq[q], qx, qy, R, out);
      qsort(out, (size_t)c, sizeof(int32_t), cmp_i32);
    }
  }
  int64_t nnz = 0;
  if (!col) {
    rowptr[0] = 0;
    for (int q = 0; q < nq; ++q) rowptr[q + 1] = rowptr[q] + cnt[q];
    nnz = rowptr[nq];
    free(cnt);
  } else {
    nnz = rowptr[nq];
  }
  grid_free(&g);
  return nnz;
}

/* Brute force O(nq * np) with the same predicate (no grid). */
int64_t orc_brute(const float* xq, const float* hq, int nq, const float* xp, int np,
                  int64_t* rowptr, int32_t* col) {
  if (!col) rowptr[0] = 0;
  for (int q = 0; q < nq; ++q) {
    int64_t c = 0;
    for (int j = 0; j < np; ++j) {
      float dx = xq[2 * q] - xp[2 * j];
      float dy = xq[2 * q + 1] - xp[2 * j + 1];
      if (sqrtf(dx * dx + dy * dy) < hq[q]) {
        if (col) col[rowptr[q] + c] = j;
        ++c;
      }
    }
    if (!col) rowptr[q + 1] = rowptr[q] + c;
  }
  return rowptr[nq];
}

/* reference grid parameters (neighborSearch.py:34-47) in fp32 */
void orc_grid_params(float min_x, float min_y, float max_x, float max_y, float hmax,
                     float* qmin_x, float* qmin_y, int* cells_x, int* cells_y) {
  float qx = min_x - hmax, qy = min_y - hmax;
  float Qx = max_x + 2.f * hmax, Qy = max_y + 2.f * hmax;
  *qmin_x = qx;
  *qmin_y = qy;
  *cells_x = (int)ceilf((Qx - qx) / hmax);
  *cells_y = (int)ceilf((Qy - qy) / hmax);
}

/* ============================================================== kernels */
/* kernels.py:21-33 in FP64 */
static double W(double q, double h) {
  double a = 1.0 - q;
  return a * a * a * a * (1.0 + 4.0 * q) * (7.0 / ORC_PI) / (h * h);
}
static double dW(double q, double h) { /* grad W = -dW * rhat */
  double a = 1.0 - q;
  return (7.0 / ORC_PI) / (h * h * h) * (20.0 * q * a * a * a);
}
static double dWs(double q, double h) { /* spiky: grad = -dWs * rhat */
  double a = 1.0 - q;
  return 30.0 / ORC_PI / (h * h * h) * a * a;
}

double orc_W(double q, double h) { return W(q, h); }
double orc_dW(double q, double h) { return dW(q, h); }
double orc_dWs(double q, double h) { return dWs(q, h); }

typedef struct Geo {
  double rx, ry, d, hx, hy, q;
} Geo;

/* neighborSearch.py:218-223: r = x_i - x_j, rhat zero below 1e-7 */
static Geo geo(double xi, double yi, double xj, double yj, double hdiv) {
  Geo g;
  g.rx = xi - xj;
  g.ry = yi - yj;
  g.d = sqrt(g.rx * g.rx + g.ry * g.ry);
  if (g.d >= 1e-7) {
    g.hx = g.rx / g.d;
    g.hy = g.ry / g.d;
  } else {
    g.hx = g.hy = 0.0;
  }
  g.q = g.d / hdiv;
  return g;
}

/* pressure.py:99-111 */
static double eos(const OrcPhys* P, double r) {
  if (P->eos == 0) return P->rho0 * P->c0 * P->c0 / P->gamma * (pow(r, P->gamma) - 1.0);
  if (P->eos == 1) return P->c0 * P->c0 * (r * P->rho0 - P->rho0);
  return P->kappa * (r * P->rho0 - P->rho0);
}

void orc_eos(const OrcPhys* P, const double* rho, int n, double* p) {
  for (int i = 0; i < n; ++i) p[i] = eos(P, rho[i]);
}

/* density.py:26-30 (+ akinciBoundary.py:414-420 when fb is given):
 * rho_i = sum_j A_j W_ij (+ sum_b V_b W_ib); vol_i = A_i / fluid-only sum
 * (density.py:53 runs before the boundary contribution is added). */
void orc_fluid_density(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                       const OrcCSR* fb, double* rho, double* vol) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n; ++i) {
    double s = 0.0;
    for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
      int j = ff->col[k];
      Geo g = geo(F->x[2 * i], F->x[2 * i + 1], F->x[2 * j], F->x[2 * j + 1], P->h);
      s += F->area[j] * W(g.q, P->h);
    }
    if (vol) vol[i] = F->area[i] / s;
    if (fb && fb->rowptr) {
      for (int64_t k = fb->rowptr[i]; k < fb->rowptr[i + 1]; ++k) {
        int b = fb->col[k];
        Geo g = geo(B->x[2 * b], B->x[2 * b + 1], F->x[2 * i], F->x[2 * i + 1], P->h);
        s += W(g.q, P->h) * B->vol[b];
      }
    }
    rho[i] = s;
  }
}

/* util.py:386-433 closed-form 2x2 pseudo-inverse, FP64 */
static void pinv2(double a, double b, double c, double d, double L[4], double eig[2]) {
  double theta = 0.5 * atan2(2 * a * c + 2 * b * d, a * a + b * b - c * c - d * d);
  double ct = cos(theta), st = sin(theta);
  double S1 = a * a + b * b + c * c + d * d;
  double e = a * a + b * b - c * c - d * d, f = a * c + b * d;
  double S2 = sqrt(e * e + 4 * f * f);
  double o1 = sqrt((S1 + S2) / 2), o2 = sqrt((S1 - S2) / 2);
  double phi = 0.5 * atan2(2 * a * b + 2 * c * d, a * a - b * b + c * c - d * d);
  double cp = cos(phi), sp = sin(phi);
  double t11 = (a * ct + c * st) * cp + (b * ct + d * st) * sp;
  double t22 = (a * st - c * ct) * sp + (-b * st + d * ct) * cp;
  double s11 = t11 > 0 ? 1 : (t11 < 0 ? -1 : 0);
  double s22 = t22 > 0 ? 1 : (t22 < 0 ? -1 : 0);
  double V00 = cp * s11, V01 = -sp * s22, V10 = sp * s11, V11 = cp * s22;
  double U00 = ct, U01 = -st, U10 = st, U11 = ct;
  double i1 = fabs(o1) > 1e-5 ? 1 / o1 : 0, i2 = fabs(o2) > 1e-5 ? 1 / o2 : 0;
  L[0] = V00 * i1 * U00 + V01 * i2 * U01;
  L[1] = V00 * i1 * U10 + V01 * i2 * U11;
  L[2] = V10 * i1 * U00 + V11 * i2 * U01;
  L[3] = V10 * i1 * U10 + V11 * i2 * U11;
  if (fabs(o2) > fabs(o1)) {
    eig[0] = o2;
    eig[1] = o1;
  } else {
    eig[0] = o1;
    eig[1] = o2;
  }
}

void orc_pinv2x2(const double* M, int n, double* L, double* eig) {
  for (int i = 0; i < n; ++i) {
    double e[2];
    pinv2(M[4 * i], M[4 * i + 1], M[4 * i + 2], M[4 * i + 3], L + 4 * i, e);
    if (eig) {
      eig[2 * i] = e[0];
      eig[2 * i + 1] = e[1];
    }
  }
}

/* 3x3 Moore-Penrose inverse by one-sided Jacobi SVD; singular values below
 * rtol * max are dropped (torch.linalg.pinv, solidBoundary.py:603). */
void orc_pinv3x3(const double* A, double rtol, double* X) {
  double U[9], V[9];
  for (int k = 0; k < 9; ++k) {
    U[k] = A[k];
    V[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int r = 0; r < 3; ++r) {
          al += U[3 * r + p] * U[3 * r + p];
          be += U[3 * r + q] * U[3 * r + q];
          ga += U[3 * r + p] * U[3 * r + q];
        }
        if (ga == 0.0 || fabs(ga) <= 1e-16 * sqrt(al * be)) continue;
        double rel = fabs(ga) / sqrt(al * be);
        if (rel > off) off = rel;
        double zeta = (be - al) / (2 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        double c = 1 / sqrt(1 + t * t), s = c * t;
        for (int r = 0; r < 3; ++r) {
          double up = U[3 * r + p], uq = U[3 * r + q];
          U[3 * r + p] = c * up - s * uq;
          U[3 * r + q] = s * up + c * uq;
          double vp = V[3 * r + p], vq = V[3 * r + q];
          V[3 * r + p] = c * vp - s * vq;
          V[3 * r + q] = s * vp + c * vq;
        }
      }
    if (off < 1e-16) break;
  }
  double sig[3], smax = 0;
  for (int k = 0; k < 3; ++k) {
    sig[k] = sqrt(U[k] * U[k] + U[3 + k] * U[3 + k] + U[6 + k] * U[6 + k]);
    if (sig[k] > smax) smax = sig[k];
  }
  for (int k = 0; k < 9; ++k) X[k] = 0;
  for (int k = 0; k < 3; ++k) {
    if (!(sig[k] > rtol * smax) || sig[k] == 0) continue;
    double inv = 1 / (sig[k] * sig[k]);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) X[3 * r + c] += V[3 * r + k] * inv * U[3 * c + k];
  }
}

/* Neighbour j of fluid row i: position, velocity, normalised density,
 * volume, pressure. */
typedef struct Nb {
  double x, y, vx, vy, rhoh, vol, p;
} Nb;

static Nb fluid_nb(const OrcFluid* F, const OrcPhys* P, int j) {
  Nb n;
  n.x = F->x[2 * j];
  n.y = F->x[2 * j + 1];
  n.vx = F->v[2 * j];
  n.vy = F->v[2 * j + 1];
  n.rhoh = F->rho[j];
  n.vol = F->vol ? F->vol[j] : F->area[j] / F->rho[j];
  n.p = eos(P, F->rho[j]);
  return n;
}

/* solidBoundary.py:748, 867: boundary volume V_b / rho_b, boundary pressure
 * from the EOS of the boundary density (solidBoundary.py:808-814). */
static Nb bdy_nb(const OrcBdy* B, const OrcPhys* P, int b) {
  Nb n;
  n.x = B->x[2 * b];
  n.y = B->x[2 * b + 1];
  n.vx = B->v ? B->v[2 * b] : 0.0;
  n.vy = B->v ? B->v[2 * b + 1] : 0.0;
  n.rhoh = B->rho[b];
  n.vol = B->vol[b] / B->rho[b];
  n.p = B->p ? B->p[b] : eos(P, B->rho[b]);
  return n;
}

/* M_i (densityDiffusion.py:30-43) and its boundary part
 * (solidBoundary.py:684-687). */
static void norm_matrix(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                        const OrcCSR* fb, int i, double M[4]) {
  M[0] = M[1] = M[2] = M[3] = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const OrcCSR* c = pass == 0 ? ff : fb;
    if (!c || !c->rowptr) continue;
    for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) {
      int j = c->col[k];
      Nb n = pass == 0 ? fluid_nb(F, P, j) : bdy_nb(B, P, j);
      Geo g = geo(F->x[2 * i], F->x[2 * i + 1], n.x, n.y, P->h);
      double gw = -dW(g.q, P->h), gx = gw * g.hx, gy = gw * g.hy;
      double rx = n.x - F->x[2 * i], ry = n.y - F->x[2 * i + 1];
      M[0] += rx * gx * n.vol;
      M[1] += rx * gy * n.vol;
      M[2] += ry * gx * n.vol;
      M[3] += ry * gy * n.vol;
    }
  }
}

/* delta-SPH right-hand side (deltaSPH.py:173-198 sequence) in FP64. */
void orc_dsph_rhs(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                  const OrcCSR* fb, unsigned terms, double* dpdt, double* acc, double* gradrho_out) {
  const int n = F->n;
  double* grho = NULL;
  if (terms & T_DDSPH) {
    /* densityDiffusion.py:134-150: M, L = pinv2x2(M), <grad rho> */
    double* L = (double*)malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
    grho = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double M[4], e[2];
      norm_matrix(F, P, ff, B, fb, i, M);
      pinv2(M[0], M[1], M[2], M[3], L + 4 * i, e);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double sx = 0, sy = 0;
      double rho_i = F->rho[i] * P->rho0;
      const double* l = L + 4 * i;
      for (int pass = 0; pass < 2; ++pass) {
        const OrcCSR* c = pass == 0 ? ff : fb;
        if (!c || !c->rowptr) continue;
        for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) {
          int j = c->col[k];
          Nb nb = pass == 0 ? fluid_nb(F, P, j) : bdy_nb(B, P, j);
          Geo g = geo(F->x[2 * i], F->x[2 * i + 1], nb.x, nb.y, P->h);
          double gw = -dW(g.q, P->h), gx = gw * g.hx, gy = gw * g.hy;
          double lx = l[0] * gx + l[1] * gy, ly = l[2] * gx + l[3] * gy;
          double c2 = (nb.rhoh * P->rho0 - rho_i) * nb.vol * 2;
          sx += c2 * lx;
          sy += c2 * ly;
        }
      }
      grho[2 * i] = -sx;
      grho[2 * i + 1] = -sy;
    }
    free(L);
    if (gradrho_out) memcpy(gradrho_out, grho, sizeof(double) * 2 * (size_t)n);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double xi = F->x[2 * i], yi = F->x[2 * i + 1];
    const double vix = F->v[2 * i], viy = F->v[2 * i + 1];
    const double rho_i = F->rho[i] * P->rho0;
    const double p_i = eos(P, F->rho[i]);
    double cont = 0, mog = 0, dd = 0, px = 0, py = 0, vx = 0, vy = 0, lx = 0, ly = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const OrcCSR* c = pass == 0 ? ff : fb;
      if (!c || !c->rowptr) continue;
      const int fl = pass == 0;
      for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) {
        int j = c->col[k];
        Nb nb = fl ? fluid_nb(F, P, j) : bdy_nb(B, P, j);
        Geo g = geo(xi, yi, nb.x, nb.y, P->h);
        double gw = -dW(g.q, P->h), gWx = gw * g.hx, gWy = gw * g.hy;
        double xjx = nb.x - xi, xjy = nb.y - yi;
        double rji2 = g.d * g.d + P->eps;
        double vjx = nb.vx - vix, vjy = nb.vy - viy;
        double rho_j = nb.rhoh * P->rho0;
        /* momentum.py:31-38, solidBoundary.py:755-760 */
        if ((fl && (terms & T_CONT)) || (!fl && (terms & T_CONT_B)))
          cont += (vjx * gWx + vjy * gWy) * nb.vol * rho_j;
        /* densityDiffusion.py:74-82 */
        if (fl && (terms & T_MOG)) mog += 2 * (rho_j - rho_i) / rji2 * (xjx * gWx + xjy * gWy) * nb.vol;
        /* densityDiffusion.py:63-71 */
        if (fl && (terms & T_DDSPH)) {
          double c2 = 2 * (rho_j - rho_i) / rji2;
          double psx = c2 * xjx - (grho[2 * i] + grho[2 * j]);
          double psy = c2 * xjy - (grho[2 * i + 1] + grho[2 * j + 1]);
          dd += (psx * gWx + psy * gWy) * nb.vol;
        }
        /* pressure.py:31-38, solidBoundary.py:875-880 */
        if ((fl && (terms & T_PRESS)) || (!fl && (terms & T_PRESS_B))) {
          double cc = (p_i + nb.p) * nb.vol;
          px += cc * gWx;
          py += cc * gWy;
        }
        /* diffusion.py:30-44, solidBoundary.py:829-835 */
        if ((fl && (terms & T_VISC)) || (!fl && (terms & T_VISC_B))) {
          double pij = vjx * xjx + vjy * xjy;
          if (pij > 0) pij = 0;
          pij /= rji2;
          double cc = pij * nb.vol * rho_j / (rho_i + rho_j);
          vx += cc * gWx;
          vy += cc * gWy;
        }
        /* laminar.py:30-46 with mu_j = rho_i nu (quirk of laminar.py:39) */
        if (fl && (terms & T_LAM)) {
          double mj = rho_j * nb.vol;
          double num = 4 * mj * (2 * rho_i * P->nu) * (xjx * gWx + xjy * gWy);
          double rs = rho_i + rho_j;
          double term = num / (rs * rs * rji2);
          lx += term * (vix - nb.vx);
          ly += term * (viy - nb.vy);
        }
      }
    }
    dpdt[i] = -cont + P->h * P->delta * P->c0 * (mog + dd);
    double ax = -px / rho_i + P->h * P->alpha * P->c0 * vx - lx;
    double ay = -py / rho_i + P->h * P->alpha * P->c0 * vy - ly;
    if (terms & T_GRAV) { /* gravity.py:56-57 */
      ax += P->gx;
      ay += P->gy;
    }
    acc[2 * i] = ax;
    acc[2 * i + 1] = ay;
  }
  free(grho);
}

/* mDBC boundary density (solidBoundary.py:561-616) from ghost->fluid pairs
 * (gf rows = boundary index; ghost support = h, MLSscale = 1). */
void orc_mdbc(const OrcFluid* F, const OrcPhys* P, const OrcBdy* B, const OrcCSR* gf,
              double* rho_b, double* p_b, int* counts) {
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n; ++b) {
    const double gxp = B->gx[2 * b], gyp = B->gx[2 * b + 1];
    double num = 0, den = 0, gsx = 0, gsy = 0, vgx = 0, vgy = 0, psx = 0, psy = 0;
    double m00 = 0, m01 = 0, m10 = 0, m11 = 0;
    for (int64_t k = gf->rowptr[b]; k < gf->rowptr[b + 1]; ++k) {
      int f = gf->col[k];
      double xf = F->x[2 * f], yf = F->x[2 * f + 1];
      Geo g = geo(gxp, gyp, xf, yf, P->h);
      double w = W(g.q, P->h), gw = -dW(g.q, P->h);
      double gWx = gw * g.hx, gWy = gw * g.hy;
      double mass = F->area[f] * P->rho0;
      double V = F->area[f] / F->rho[f];
      num += mass * w;
      den += V * w;
      gsx += mass * gWx;
      gsy += mass * gWy;
      vgx += V * gWx;
      vgy += V * gWy;
      double xx = xf - gxp, xy = yf - gyp;
      psx += V * w * xx;
      psy += V * w * xy;
      m00 += V * xx * gWx;
      m01 += V * xx * gWy;
      m10 += V * xy * gWx;
      m11 += V * xy * gWy;
    }
    int cnt = (int)(gf->rowptr[b + 1] - gf->rowptr[b]);
    double rho = 1.0;
    if (cnt > 10) {
      double A[9] = {den, psx, psy, vgx, m00, m01, vgy, m10, m11}, X[9];
      orc_pinv3x3(A, 3 * 1.1920928955078125e-07, X);
      double r0 = X[0] * num + X[1] * gsx + X[2] * gsy;
      double r1 = X[3] * num + X[4] * gsx + X[5] * gsy;
      double r2 = X[6] * num + X[7] * gsx + X[8] * gsy;
      rho = (r0 + (B->x[2 * b] - gxp) * r1 + (B->x[2 * b + 1] - gyp) * r2) / P->rho0;
    } else if (cnt > 0) {
      rho = num / den / P->rho0;
    }
    if (rho < 1.0) rho = 1.0;
    rho_b[b] = rho;
    if (p_b) p_b[b] = eos(P, rho);
    if (counts) counts[b] = cnt;
  }
}

/* akinciBoundary.py:414-419: rho_b = term_b + sum_f A_f W_bf + beta */
void orc_akinci_bdensity(const OrcFluid* F, const OrcPhys* P, const OrcBdy* B,
                         const double* term, const OrcCSR* bf, double* rho_b) {
#pragma omp parallel for schedule(static)
  for (int b = 0; b < B->n; ++b) {
    double s = 0;
    for (int64_t k = bf->rowptr[b]; k < bf->rowptr[b + 1]; ++k) {
      int f = bf->col[k];
      Geo g = geo(B->x[2 * b], B->x[2 * b + 1], F->x[2 * f], F->x[2 * f + 1], P->h);
      s += W(g.q, P->h) * F->area[f];
    }
    rho_b[b] = term[b] + s + P->beta;
  }
}

/* ================================================================ DFSPH */
/* Results of orc_dfsph. */
typedef struct OrcDfsphOut {
  double *alpha, *src, *p2, *apred, *vpred, *b_alpha, *b_src, *b_p, *err_hist;
  int max_hist;
  int iters;
} OrcDfsphOut;

/* Spiky gradient of pair (i at xi, j at xj) in the row convention
 * G = -dWs(q) rhat, rhat = (xi - xj)/|.| (kernels.py:32-33). */
static void G(double xi, double yi, double xj, double yj, double h, double* gx, double* gy) {
  Geo g = geo(xi, yi, xj, yj, h);
  double s = -dWs(g.q, h);
  *gx = s * g.hx;
  *gy = s * g.hy;
}

/* a*_i (dfsph.py:25-42, akinciBoundary.py:229-252 PBSPH) */
static void dfsph_accel(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                        const OrcCSR* fb, const double* p2, const double* bp, double* a) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n; ++i) {
    double xi = F->x[2 * i], yi = F->x[2 * i + 1];
    double ri = F->rho[i] * F->rest[i];
    double pi = p2[i] / (ri * ri);
    double ax = 0, ay = 0;
    for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
      int j = ff->col[k];
      double gx, gy;
      G(xi, yi, F->x[2 * j], F->x[2 * j + 1], P->h, &gx, &gy);
      double rj = F->rho[j] * F->rest[j];
      double c = -(F->area[j] * F->rest[j]) * (pi + p2[j] / (rj * rj));
      ax += c * gx;
      ay += c * gy;
    }
    if (fb && fb->rowptr) {
      for (int64_t k = fb->rowptr[i]; k < fb->rowptr[i + 1]; ++k) {
        int b = fb->col[k];
        double gx, gy;
        G(xi, yi, B->x[2 * b], B->x[2 * b + 1], P->h, &gx, &gy);
        double rb = B->rho[b] * B->rest[b];
        double c = -(B->vol[b] * B->rest[b]) * (pi + bp[b] / (rb * rb));
        ax += c * gx;
        ay += c * gy;
      }
    }
    a[2 * i] = ax;
    a[2 * i + 1] = ay;
  }
}

/*
 * DFSPH solve (modules/dfsph.py:263-311) with Akinci PBSPH boundary
 * (akinciBoundary.py:166-389), FP64.  acc is the acceleration before the
 * solve (gravity); on return out->apred is the final pressure
 * acceleration.  fixed_iters > 0 runs exactly that many iterations.
 *  density = 1: densitySolve (dfsph.py:195-234): source 1 - rho + divergence
 *    (:161), pressure clamped at 0 (:177), error over fluid and boundary
 *    (:219-222);
 *  density = 0: divergenceSolve (dfsph.py:236-260): every Akinci hook
 *    returns zeros (akinciBoundary.py:167,190,255,319,356), so the solve is
 *    fluid-only; source = divergence only, no clamp, error = mean of the
 *    clamped fluid residuals (:253).
 * P->tau, P->min_iters and P->max_iters are the limits of the solve run
 * (densityThreshold / divergenceThreshold and their iteration bounds).
 */
void orc_dfsph_solve(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                     const OrcCSR* fb, const OrcCSR* bf, const double* acc, int fixed_iters,
                     int density, OrcDfsphOut* out) {
  if (!density) {
    B = NULL;
    fb = NULL;
    bf = NULL;
  }
  const int n = F->n, nb = B ? B->n : 0;
  const double dt = P->dt;
  double* aact = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  double* baact = (double*)malloc(sizeof(double) * (size_t)(nb + 1));
  double* p = (double*)calloc((size_t)n + 1, sizeof(double));
  double* p2 = out->p2;
  double* bp = out->b_p;
  double* vp = out->vpred;
  double* ap = out->apred;
  for (int i = 0; i < n; ++i) {
    vp[2 * i] = F->v[2 * i] + dt * acc[2 * i];
    vp[2 * i + 1] = F->v[2 * i + 1] + dt * acc[2 * i + 1];
    aact[i] = F->area[i] / F->rho[i];
    p2[i] = 0;
  }
  for (int b = 0; b < nb; ++b) {
    baact[b] = B->vol[b] / B->rho[b];
    bp[b] = 0;
  }
  /* alpha and source (dfsph.py:58-100, akinciBoundary.py:317-389) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double xi = F->x[2 * i], yi = F->x[2 * i + 1];
    double k1x = 0, k1y = 0, k2 = 0, src = 0;
    for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
      int j = ff->col[k];
      double gx, gy;
      G(xi, yi, F->x[2 * j], F->x[2 * j + 1], P->h, &gx, &gy);
      k1x += aact[j] * gx;
      k1y += aact[j] * gy;
      k2 += aact[j] * aact[j] / (F->area[j] * F->rest[j]) * (gx * gx + gy * gy);
      src += -dt * aact[j] * ((vp[2 * i] - vp[2 * j]) * gx + (vp[2 * i + 1] - vp[2 * j + 1]) * gy);
    }
    if (fb && fb->rowptr) {
      for (int64_t k = fb->rowptr[i]; k < fb->rowptr[i + 1]; ++k) {
        int b = fb->col[k];
        double gx, gy;
        G(xi, yi, B->x[2 * b], B->x[2 * b + 1], P->h, &gx, &gy);
        k1x += baact[b] * gx;
        k1y += baact[b] * gy;
        k2 += baact[b] * baact[b] / (B->vol[b] * B->rest[b]) * (gx * gx + gy * gy);
        src += -dt * baact[b] * (vp[2 * i] * gx + vp[2 * i + 1] * gy);
      }
    }
    double fac = -dt * dt * aact[i];
    double mass = F->area[i] * F->rest[i];
    double al = fac / mass * (k1x * k1x + k1y * k1y) + fac * k2;
    if (al < -1) al = -1;
    if (al > -1e-7) al = -1e-7;
    out->alpha[i] = al;
    out->src[i] = density ? 1.0 - F->rho[i] + src : src;
  }
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    double s2 = 0, src = 0;
    for (int64_t k = bf->rowptr[b]; k < bf->rowptr[b + 1]; ++k) {
      int f = bf->col[k];
      double gx, gy;
      G(B->x[2 * b], B->x[2 * b + 1], F->x[2 * f], F->x[2 * f + 1], P->h, &gx, &gy);
      s2 += aact[f] * aact[f] / (F->area[f] * F->rest[f]) * (gx * gx + gy * gy);
      src += aact[f] * ((0 - vp[2 * f]) * gx + (0 - vp[2 * f + 1]) * gy);
    }
    out->b_alpha[b] = -dt * dt * baact[b] * s2;
    out->b_src[b] = 1.0 - B->rho[b] - dt * src;
  }
  /* densitySolve loop (dfsph.py:195-234) */
  int it = 0;
  double err = 0;
  for (;;) {
    if (fixed_iters > 0) {
      if (it >= fixed_iters) break;
    } else if (!((it < P->min_iters || err > P->tau) && it <= P->max_iters)) {
      break;
    }
    dfsph_accel(F, P, ff, B, fb, p2, bp, ap);
    memcpy(p, p2, sizeof(double) * (size_t)n);
    double esum = 0;
#pragma omp parallel for schedule(static) reduction(+ : esum)
    for (int i = 0; i < n; ++i) {
      double xi = F->x[2 * i], yi = F->x[2 * i + 1];
      double ks = 0;
      for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
        int j = ff->col[k];
        double gx, gy;
        G(xi, yi, F->x[2 * j], F->x[2 * j + 1], P->h, &gx, &gy);
        ks += dt * dt * aact[j] * ((ap[2 * i] - ap[2 * j]) * gx + (ap[2 * i + 1] - ap[2 * j + 1]) * gy);
      }
      if (fb && fb->rowptr) {
        for (int64_t k = fb->rowptr[i]; k < fb->rowptr[i + 1]; ++k) {
          int b = fb->col[k];
          double gx, gy;
          G(xi, yi, B->x[2 * b], B->x[2 * b + 1], P->h, &gx, &gy);
          ks += dt * dt * baact[b] * (ap[2 * i] * gx + ap[2 * i + 1] * gy);
        }
      }
      double r = ks - out->src[i];
      double pn = p[i] - P->fluid_omega * r / out->alpha[i];
      if (density && pn < 0) pn = 0;
      p2[i] = pn;
      esum += r < -P->tau ? -P->tau : r;
    }
    double bsum = 0;
#pragma omp parallel for schedule(static) reduction(+ : bsum)
    for (int b = 0; b < nb; ++b) {
      double ks = 0;
      for (int64_t k = bf->rowptr[b]; k < bf->rowptr[b + 1]; ++k) {
        int f = bf->col[k];
        double gx, gy;
        G(B->x[2 * b], B->x[2 * b + 1], F->x[2 * f], F->x[2 * f + 1], P->h, &gx, &gy);
        ks += dt * dt * aact[f] * (ap[2 * f] * (-gx) + ap[2 * f + 1] * (-gy));
      }
      double r = ks - out->b_src[b];
      double pn = bp[b] - P->omega * r / out->b_alpha[b];
      if (pn < 0) pn = 0;
      bp[b] = pn;
      bsum += r < -P->tau ? -P->tau : r;
    }
    err = (esum + bsum) / (double)(n + nb);
    if (it < out->max_hist) out->err_hist[it] = err;
    ++it;
  }
  out->iters = it;
  /* final acceleration (dfsph.py:304-309) */
  dfsph_accel(F, P, ff, B, fb, p2, bp, ap);
  for (int i = 0; i < n; ++i) {
    vp[2 * i] += dt * ap[2 * i];
    vp[2 * i + 1] += dt * ap[2 * i + 1];
  }
  free(aact);
  free(baact);
  free(p);
}

/* densitySolve (density = 1) of orc_dfsph_solve. */
void orc_dfsph(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
               const OrcCSR* fb, const OrcCSR* bf, const double* acc, int fixed_iters,
               OrcDfsphOut* out) {
  orc_dfsph_solve(F, P, ff, B, fb, bf, acc, fixed_iters, 1, out);
}

/* ====================================================== delta-plus shifting */
/* Renormalisation matrices of arbitrary query rows over two lists
 * (computeNormalizationMatrix, densityDiffusion.py:30-43, called for the
 * solid boundary's own rows at solidBoundary.py:676-693) and pinv2x2. */
void orc_norm_rows(const double* xq, int nq, double h, const OrcCSR* a, const double* xa,
                   const double* va, const OrcCSR* b, const double* xb, const double* vb,
                   double* M, double* L, double* eig) {
#pragma omp parallel for schedule(static)
  for (int q = 0; q < nq; ++q) {
    double m[4] = {0, 0, 0, 0};
    for (int pass = 0; pass < 2; ++pass) {
      const OrcCSR* c = pass == 0 ? a : b;
      const double* xs = pass == 0 ? xa : xb;
      const double* vs = pass == 0 ? va : vb;
      if (!c || !c->rowptr) continue;
      for (int64_t k = c->rowptr[q]; k < c->rowptr[q + 1]; ++k) {
        int j = c->col[k];
        Geo g = geo(xq[2 * q], xq[2 * q + 1], xs[2 * j], xs[2 * j + 1], h);
        double gw = -dW(g.q, h), gx = gw * g.hx, gy = gw * g.hy;
        double rx = xs[2 * j] - xq[2 * q], ry = xs[2 * j + 1] - xq[2 * q + 1];
        m[0] += rx * gx * vs[j];
        m[1] += rx * gy * vs[j];
        m[2] += ry * gx * vs[j];
        m[3] += ry * gy * vs[j];
      }
    }
    double e[2];
    if (M) memcpy(M + 4 * q, m, sizeof m);
    pinv2(m[0], m[1], m[2], m[3], L + 4 * q, e);
    if (eig) {
      eig[2 * q] = e[0];
      eig[2 * q + 1] = e[1];
    }
  }
}

/* Fluid rows: M_i over fluid and boundary neighbours (deltaPlus.py:116-131
 * with solidBoundary.py:684-687), L = pinv2x2(M), eig. */
void orc_fluid_norm(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                    const OrcCSR* fb, double* M, double* L, double* eig) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n; ++i) {
    double m[4], e[2];
    norm_matrix(F, P, ff, B, fb, i, m);
    if (M) memcpy(M + 4 * i, m, sizeof m);
    pinv2(m[0], m[1], m[2], m[3], L + 4 * i, e);
    eig[2 * i] = e[0];
    eig[2 * i + 1] = e[1];
  }
}

/* computeFluidNormal (deltaPlus.py:132-163): lambda = eig[.][1];
 * g_i = -sum_j V_j (lam_j - lam_i) L_i gradW_ij (fluid), the same over the
 * boundary neighbours with volume bvol_n[b] and lambda beig[b][1];
 * n_i = g_i / (|g_i| + eps). */
void orc_dplus_normal(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const double* L,
                      const double* eig, const double* bx, const double* beig,
                      const double* bvol_n, const OrcCSR* fb, double* grad, double* normal) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n; ++i) {
    const double* l = L + 4 * i;
    double lam_i = eig[2 * i + 1], gx = 0, gy = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const OrcCSR* c = pass == 0 ? ff : fb;
      if (!c || !c->rowptr) continue;
      for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) {
        int j = c->col[k];
        double xj, yj, V, lam_j;
        if (pass == 0) {
          xj = F->x[2 * j];
          yj = F->x[2 * j + 1];
          V = F->vol ? F->vol[j] : F->area[j] / F->rho[j];
          lam_j = eig[2 * j + 1];
        } else {
          xj = bx[2 * j];
          yj = bx[2 * j + 1];
          V = bvol_n[j];
          lam_j = beig[2 * j + 1];
        }
        Geo g = geo(F->x[2 * i], F->x[2 * i + 1], xj, yj, P->h);
        double gw = -dW(g.q, P->h), wx = gw * g.hx, wy = gw * g.hy;
        double cx = l[0] * wx + l[1] * wy, cy = l[2] * wx + l[3] * wy;
        gx -= V * (lam_j - lam_i) * cx;
        gy -= V * (lam_j - lam_i) * cy;
      }
    }
    if (grad) {
      grad[2 * i] = gx;
      grad[2 * i + 1] = gy;
    }
    double nrm = sqrt(gx * gx + gy * gy) + P->eps;
    normal[2 * i] = gx / nrm;
    normal[2 * i + 1] = gy / nrm;
  }
}

/* detectFluidSurface (deltaPlus.py:165-209).  cmax_i = max(max over the
 * non-self fluid pairs of (x_j - x_i)/|x_j - x_i| . n_i, max over the
 * boundary pairs of (x_b - x_i)/|x_b - x_i| . n_i), an empty max being 0
 * (torch_scatter's fill); mask_i = [acos(cmax_i) > pi/6 and lam_i < 0.6];
 * m2_i = max of mask_j over the non-self fluid pairs; classes 0, 2/3, 1, 2.
 * The thresholds are the fp32 values the reference compares against. */
void orc_dplus_surface(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff,
                       const double* normal, const double* eig, const double* bx,
                       const OrcCSR* fb, double* cmax, double* mask, double* surface) {
  const int n = F->n;
  const double pi6 = (double)(float)(ORC_PI / 6), lam_s = (double)0.6f, lam_k = (double)0.5f;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double nx = normal[2 * i], ny = normal[2 * i + 1], fmax = 0, bmax = 0;
    int any = 0;
    for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
      int j = ff->col[k];
      if (j == i) continue;
      Geo g = geo(F->x[2 * j], F->x[2 * j + 1], F->x[2 * i], F->x[2 * i + 1], P->h);
      double d = g.hx * nx + g.hy * ny;
      fmax = any ? (d > fmax ? d : fmax) : d;
      any = 1;
    }
    any = 0;
    if (fb && fb->rowptr)
      for (int64_t k = fb->rowptr[i]; k < fb->rowptr[i + 1]; ++k) {
        int b = fb->col[k];
        Geo g = geo(bx[2 * b], bx[2 * b + 1], F->x[2 * i], F->x[2 * i + 1], P->h);
        double d = g.hx * nx + g.hy * ny;
        bmax = any ? (d > bmax ? d : bmax) : d;
        any = 1;
      }
    double c = fmax > bmax ? fmax : bmax;
    cmax[i] = c;
    mask[i] = (acos(c) > pi6 && eig[2 * i + 1] < lam_s) ? 1.0 : 0.0;
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double m2 = 0;
    for (int64_t k = ff->rowptr[i]; k < ff->rowptr[i + 1]; ++k) {
      int j = ff->col[k];
      if (j != i && mask[j] > m2) m2 = mask[j];
    }
    double r;
    if (mask[i] > 0)
      r = 0;
    else if (m2 > 0)
      r = eig[2 * i + 1] >= lam_k ? 1.0 : 2.0 / 3.0;
    else
      r = 2;
    surface[i] = r;
  }
}

/* computeShiftAmount + adjustShiftingAmount (deltaPlus.py:81-115,
 * 211-236).  Ma = max_k |v_k| / c0; k0 = W(packing); raw (optional) is the
 * shift before the clamp and the projection. */
void orc_dplus_shift(const OrcFluid* F, const OrcPhys* P, double packing, const OrcCSR* ff,
                     const OrcBdy* B, const OrcCSR* fb, const double* normal,
                     const double* surface, double* raw, double* shift) {
  const int n = F->n;
  double vmax = 0;
  for (int i = 0; i < n; ++i) {
    double s = sqrt(F->v[2 * i] * F->v[2 * i] + F->v[2 * i + 1] * F->v[2 * i + 1]);
    if (s > vmax) vmax = s;
  }
  const double Ma = vmax / P->c0, k0 = W(packing, P->h), h = P->h;
  const double c = -1.5 * Ma * 4 * h * h;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double sx = 0, sy = 0;
    for (int pass = 0; pass < 2; ++pass) {
      const OrcCSR* cs = pass == 0 ? ff : fb;
      if (!cs || !cs->rowptr) continue;
      for (int64_t k = cs->rowptr[i]; k < cs->rowptr[i + 1]; ++k) {
        int j = cs->col[k];
        double xj, yj, mass;
        if (pass == 0) {
          xj = F->x[2 * j];
          yj = F->x[2 * j + 1];
          mass = F->area[j] / (F->rho[i] + F->rho[j]);
        } else {
          xj = B->x[2 * j];
          yj = B->x[2 * j + 1];
          mass = B->vol[j] / (F->rho[i] + B->rho[j]);
        }
        Geo g = geo(F->x[2 * i], F->x[2 * i + 1], xj, yj, h);
        double w = W(g.q, h) / k0;
        double kt = 1 + 0.2 * w * w * w * w;
        double gw = -dW(g.q, h);
        sx += kt * mass * gw * g.hx;
        sy += kt * mass * gw * g.hy;
      }
    }
    sx *= c;
    sy *= c;
    if (raw) {
      raw[2 * i] = sx;
      raw[2 * i + 1] = sy;
    }
    double len = sqrt(sx * sx + sy * sy), cap = 0.05 * h;
    if (len > cap) {
      sx = sx / len * cap;
      sy = sy / len * cap;
    }
    double m = surface[i];
    if (m < 0.5) {
      sx = sy = 0;
    } else if (m < 1.5) {
      double nx = normal[2 * i], ny = normal[2 * i + 1];
      double px = (1 - nx * nx) * sx - nx * ny * sy;
      double py = -ny * nx * sx + (1 - ny * ny) * sy;
      sx = px;
      sy = py;
    }
    shift[2 * i] = sx;
    shift[2 * i + 1] = sy;
  }
}

/* computePressureAccelDeltaPlus (pressure.py:40-52) for the fluid pairs
 * (T_PRESS) and the boundary pairs (T_PRESS_B, solidBoundary.py:882-887):
 * p_ij = p_i + p_j if p_i >= 0 or surface_i < 1.5, else p_j - p_i. */
void orc_dplus_pressure(const OrcFluid* F, const OrcPhys* P, const OrcCSR* ff, const OrcBdy* B,
                        const OrcCSR* fb, const double* surface, unsigned terms, double* acc) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < F->n; ++i) {
    double pi = eos(P, F->rho[i]), ax = 0, ay = 0;
    int sum = pi >= 0 || surface[i] < 1.5;
    for (int pass = 0; pass < 2; ++pass) {
      const OrcCSR* c = pass == 0 ? ff : fb;
      if (!c || !c->rowptr || !(terms & (pass == 0 ? T_PRESS : T_PRESS_B))) continue;
      for (int64_t k = c->rowptr[i]; k < c->rowptr[i + 1]; ++k) {
        int j = c->col[k];
        Nb nb = pass == 0 ? fluid_nb(F, P, j) : bdy_nb(B, P, j);
        double pij = sum ? pi + nb.p : nb.p - pi;
        Geo g = geo(F->x[2 * i], F->x[2 * i + 1], nb.x, nb.y, P->h);
        double gw = -dW(g.q, P->h);
        ax += pij * nb.vol * gw * g.hx;
        ay += pij * nb.vol * gw * g.hy;
      }
    }
    double rho_i = F->rho[i] * P->rho0;
    acc[2 * i] = -ax / rho_i;
    acc[2 * i + 1] = -ay / rho_i;
  }
}

void orc_free(void* p) { free(p); }
