/*
 * CPU restatement (C + OpenMP, FP64) of jNO's P1 Poisson path -- TEST INFRASTRUCTURE ONLY: the
 * timed CPU baseline of bench.py and a second checker for the tests.  Never linked into jno_amd.
 *
 *   oracle_p1_pattern_count / _fill : structural CSR pattern (sorted unique element pairs), the
 *       index structure of compress_plan (jno/utils/solver/fem_utils.py:4009-4060)
 *   oracle_p1_assemble : K_e = k |T| G G^T summed per slot, load f|T|/4, symmetric Dirichlet
 *       elimination (jno/utils/solver/fem_1d.py:667-686), Jacobi inverse (jno/precond.py:105-109)
 *   oracle_pcg : Jacobi-PCG with jax.scipy.sparse.linalg.cg semantics (jno/solve.py:191-206)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* node -> incident cells (counting sort; ascending cell ids per node) */
static void incidence(const int32_t* cells, int64_t T, int64_t N, int32_t* ptr, int32_t* inc) {
  memset(ptr, 0, sizeof(int32_t) * (N + 1));
  for (int64_t i = 0; i < 4 * T; ++i) ptr[cells[i] + 1]++;
  for (int64_t i = 0; i < N; ++i) ptr[i + 1] += ptr[i];
  int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * N);
  memcpy(cur, ptr, sizeof(int32_t) * N);
  for (int64_t i = 0; i < 4 * T; ++i) inc[cur[cells[i]]++] = (int32_t)i;
  free(cur);
}

static int unique_neighbours(const int32_t* cells, const int32_t* ptr, const int32_t* inc, int64_t r, int32_t* buf) {
  int m = 0;
  for (int32_t q = ptr[r]; q < ptr[r + 1]; ++q) {
    const int32_t e = inc[q] >> 2;
    for (int a = 0; a < 4; ++a) buf[m++] = cells[4 * (int64_t)e + a];
  }
  qsort(buf, m, sizeof(int32_t), cmp_i32);
  int u = 0;
  for (int i = 0; i < m; ++i)
    if (i == 0 || buf[i] != buf[i - 1]) buf[u++] = buf[i];
  return u;
}

/* workspace: ptr (N+1) and inc (4T) int32, owned by the caller */
int64_t oracle_p1_pattern_count(const int32_t* cells, int64_t T, int64_t N, int32_t* inc_ptr, int32_t* inc,
                                int32_t* rowptr) {
  incidence(cells, T, N, inc_ptr, inc);
  int maxdeg = 0;
  for (int64_t r = 0; r < N; ++r) {
    int d = inc_ptr[r + 1] - inc_ptr[r];
    if (d > maxdeg) maxdeg = d;
  }
  rowptr[0] = 0;
#pragma omp parallel
  {
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (4 * maxdeg + 4));
#pragma omp for schedule(static)
    for (int64_t r = 0; r < N; ++r) rowptr[r + 1] = unique_neighbours(cells, inc_ptr, inc, r, buf);
    free(buf);
  }
  int64_t s = 0;
  for (int64_t r = 0; r < N; ++r) {
    s += rowptr[r + 1];
    rowptr[r + 1] = (int32_t)s;
  }
  return s;
}

void oracle_p1_pattern_fill(const int32_t* cells, int64_t N, const int32_t* inc_ptr, const int32_t* inc,
                            const int32_t* rowptr, int32_t* col) {
  int maxdeg = 0;
  for (int64_t r = 0; r < N; ++r) {
    int d = inc_ptr[r + 1] - inc_ptr[r];
    if (d > maxdeg) maxdeg = d;
  }
#pragma omp parallel
  {
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (4 * maxdeg + 4));
#pragma omp for schedule(static)
    for (int64_t r = 0; r < N; ++r) {
      int u = unique_neighbours(cells, inc_ptr, inc, r, buf);
      memcpy(col + rowptr[r], buf, sizeof(int32_t) * u);
    }
    free(buf);
  }
}

static void tet_grads(const double* pts, const int32_t* c, double g[4][3], double* vol) {
  const double* p0 = pts + 3 * (int64_t)c[0];
  double e[3][3];
  for (int v = 0; v < 3; ++v)
    for (int i = 0; i < 3; ++i) e[v][i] = pts[3 * (int64_t)c[v + 1] + i] - p0[i];
  double c23[3] = {e[1][1] * e[2][2] - e[1][2] * e[2][1], e[1][2] * e[2][0] - e[1][0] * e[2][2],
                   e[1][0] * e[2][1] - e[1][1] * e[2][0]};
  double c31[3] = {e[2][1] * e[0][2] - e[2][2] * e[0][1], e[2][2] * e[0][0] - e[2][0] * e[0][2],
                   e[2][0] * e[0][1] - e[2][1] * e[0][0]};
  double c12[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                   e[0][0] * e[1][1] - e[0][1] * e[1][0]};
  double det = e[0][0] * c23[0] + e[0][1] * c23[1] + e[0][2] * c23[2];
  for (int i = 0; i < 3; ++i) {
    g[1][i] = c23[i] / det;
    g[2][i] = c31[i] / det;
    g[3][i] = c12[i] / det;
    g[0][i] = -(g[1][i] + g[2][i] + g[3][i]);
  }
  *vol = fabs(det) / 6.0;
}

/* Row-parallel assembly (each row owned by one thread): vals, b, dinv.  dmask/g nullable. */
void oracle_p1_assemble(const double* pts, const int32_t* cells, int64_t N, const int32_t* inc_ptr, const int32_t* inc,
                        const int32_t* rowptr, const int32_t* col, const double* k, double f, const uint8_t* dmask,
                        const double* g, double* vals, double* b, double* dinv) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < N; ++r) {
    const int32_t rb = rowptr[r], re = rowptr[r + 1];
    for (int32_t j = rb; j < re; ++j) vals[j] = 0.0;
    double bl = 0.0;
    for (int32_t q = inc_ptr[r]; q < inc_ptr[r + 1]; ++q) {
      const int32_t e = inc[q] >> 2, a = inc[q] & 3;
      const int32_t* c = cells + 4 * (int64_t)e;
      double gr[4][3], vol;
      tet_grads(pts, c, gr, &vol);
      const double coef = k[e] * vol;
      for (int bb = 0; bb < 4; ++bb) {
        const double v = coef * (gr[a][0] * gr[bb][0] + gr[a][1] * gr[bb][1] + gr[a][2] * gr[bb][2]);
        /* binary search of the column */
        int32_t lo = rb, hi = re - 1;
        while (lo < hi) {
          int32_t mid = (lo + hi) >> 1;
          if (col[mid] < c[bb]) lo = mid + 1; else hi = mid;
        }
        vals[lo] += v;
      }
      bl += f * vol / 4.0;
    }
    if (dmask) {
      if (dmask[r]) {
        for (int32_t j = rb; j < re; ++j) vals[j] = (col[j] == r) ? 1.0 : 0.0;
        bl = g ? g[r] : 0.0;
      } else {
        double lift = 0.0;
        for (int32_t j = rb; j < re; ++j)
          if (dmask[col[j]]) {
            if (g) lift += vals[j] * g[col[j]];
            vals[j] = 0.0;
          }
        bl -= lift;
      }
    }
    double d = 0.0;
    for (int32_t j = rb; j < re; ++j)
      if (col[j] == r) d += vals[j];
    if (b) b[r] = bl;
    if (dinv) dinv[r] = fabs(d) > 1e-30 ? 1.0 / d : 1.0;
  }
}

static void spmv(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) s += vals[j] * x[col[j]];
    y[r] = s;
  }
}

/* JAX cg semantics: x0 = 0, r = b, z = M r, p = z; while r.r > max(tol^2 b.b, atol^2) and k < maxiter.
 * Returns the iteration count; *relres = ||b - A x|| / ||b|| (true residual). */
int64_t oracle_pcg(int64_t n, const int32_t* rowptr, const int32_t* col, const double* vals, const double* dinv,
                   const double* b, double* x, double tol, double atol, int64_t maxiter, double* relres) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double bb = 0.0, gamma = 0.0, rr = 0.0;
#pragma omp parallel for reduction(+ : bb, gamma, rr) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
    p[i] = dinv[i] * b[i];
    bb += b[i] * b[i];
    gamma += b[i] * p[i];
    rr += b[i] * b[i];
  }
  const double atol2 = fmax(tol * tol * bb, atol * atol);
  int64_t k = 0;
  while (rr > atol2 && k < maxiter) {
    spmv(n, rowptr, col, vals, p, q);
    double pq = 0.0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
    for (int64_t i = 0; i < n; ++i) pq += p[i] * q[i];
    const double alpha = gamma / pq;
    double gn = 0.0, rn = 0.0;
#pragma omp parallel for reduction(+ : gn, rn) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      gn += r[i] * dinv[i] * r[i];
      rn += r[i] * r[i];
    }
    const double beta = gn / gamma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = dinv[i] * r[i] + beta * p[i];
    gamma = gn;
    rr = rn;
    ++k;
  }
  spmv(n, rowptr, col, vals, x, q);
  double res = 0.0;
#pragma omp parallel for reduction(+ : res) schedule(static)
  for (int64_t i = 0; i < n; ++i) res += (b[i] - q[i]) * (b[i] - q[i]);
  if (relres) *relres = bb > 0 ? sqrt(res / bb) : sqrt(res);
  free(r);
  free(p);
  free(q);
  return k;
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
