/*
 * ORACLE — test infrastructure only.
 *
 * A CPU (C + OpenMP) restatement of the filtered-Lanczos path that the
 * reference drives through pEVSL, used as the checker of the HIP library and
 * as the CPU baseline ("port") timed by bench.py.  Only tests/, the smoke test
 * and bench.py's cpu_baseline leg call it; the product never links it.
 *
 * pEVSL (js1019/pEVSL fork of eigs/pEVSL, no pinned version; makefile.in:4-6,
 * INSTALL.md:22-31) is not vendored, so every routine below restates the
 * published EVSL/pEVSL algorithm and cites the reference call site whose
 * behaviour it stands for:
 *   or_spmv          pevsl_parcsrmatvec_f90      src/mod_matvec.f90:453,471,495
 *   or_diagscale     Bdiagscaling/Apdiagscaling  src/mod_matvec.f90:252-441
 *   or_apply_a       sparseAV  w = D A D v       src/mod_matvec.f90:445-458
 *   or_apply_b       sparseBV  w = B~ v          src/mod_matvec.f90:461-472
 *   or_chebiter      pEVSL_SETUP_CHEBITER/chebiter src/mod_matvec.f90:88-93,475-483
 *   or_lanbounds     pEVSL_LANBOUNDS_F90         src/mod_pevsl.f90:84; mod_matvec.f90:85
 *   or_findpol       pEVSL_FINDPOL_F90           src/mod_pevsl.f90:115
 *   or_chebav        filter apply inside ChebLanNr
 *   or_cheblannr     pEVSL_CHEBLANNR_F90         src/mod_pevsl.f90:122-130
 *   or_residuals     residual check              src/mod_pevsl.f90:137-162
 *
 * Parity status of each routine (what pins it) is listed in the oracle's
 * Python front (oracle/oracle.py) and checked by tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct {
  int64_t n_rows, n_cols;
  const int64_t *rowptr;
  const int32_t *col;
  const double *val;
} or_csr;

typedef struct {
  int64_t n;
  or_csr A;          /* stiffness, unscaled when d != NULL */
  const double *d;   /* D = diag(B)^-1/2, or NULL (A already scaled) */
  or_csr B;          /* scaled mass B~; rowptr == NULL for a standard problem */
  double blmin, blmax; /* chebiter interval for B~ */
  int bdeg;          /* chebiter degree (number of B~ products) */
  int amode;         /* 0: operator D A D ; 1: operator B~ */
} or_prob;

/* progress lines on stderr every `or_verbose_every` Lanczos steps (0: silent),
 * for long CPU solves */
#include <stdio.h>
static int or_verbose_every = 0;
void or_set_verbose(int every) { or_verbose_every = every; }
static void or_progress(const char *what, int k) {
  if (or_verbose_every > 0 && k % or_verbose_every == 0) {
    fprintf(stderr, "[oracle] %s: step %d\n", what, k);
    fflush(stderr);
  }
}

/* ------------------------------------------------------------------ */
/* vectors                                                             */
static double dot(int64_t n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}
static void axpy(int64_t n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}
static void scal(int64_t n, double a, double *x) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) x[i] *= a;
}

/* start vector: splitmix64 hash of (seed, index) mapped to [-1, 1) */
static double start_entry(uint64_t seed, uint64_t i) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ULL + i + 0x632BE59BD9B4E019ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
void or_start_vector(int64_t n, uint64_t seed, double *v) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) v[i] = start_entry(seed, (uint64_t)i);
}

/* ------------------------------------------------------------------ */
/* sparse products                                                     */
void or_spmv(const or_csr *M, const double *x, double *y) {
#pragma omp parallel for schedule(static, 256)
  for (int64_t i = 0; i < M->n_rows; ++i) {
    double s = 0.0;
    for (int64_t k = M->rowptr[i]; k < M->rowptr[i + 1]; ++k) s += M->val[k] * x[M->col[k]];
    y[i] = s;
  }
}

/* d_i = 1/sqrt(sign*M_ii);  M_ij <- sign*M_ij*d_i*d_j (square M).  Returns
 * -1 - i for the first row i whose sign*diagonal is not positive or missing. */
int64_t or_diagscale(int64_t n, const int64_t *rowptr, const int32_t *col, double *val, double sign,
                     double *d) {
  for (int64_t i = 0; i < n; ++i) {
    double di = 0.0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == i) di = sign * val[k];
    if (!(di > 0.0)) return -1 - i;
    d[i] = 1.0 / sqrt(di);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) val[k] = sign * val[k] * d[col[k]] * d[i];
  return 0;
}

/* y = op(x); work of n */
void or_apply_a(const or_prob *p, const double *x, double *y, double *work) {
  const int64_t n = p->n;
  if (p->amode == 1) {
    or_spmv(&p->B, x, y);
    return;
  }
  if (p->d) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) work[i] = p->d[i] * x[i];
    or_spmv(&p->A, work, y);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] *= p->d[i];
  } else {
    or_spmv(&p->A, x, y);
  }
}

void or_apply_b(const or_prob *p, const double *x, double *y) {
  if (p->B.rowptr)
    or_spmv(&p->B, x, y);
  else
    memcpy(y, x, sizeof(double) * p->n);
}

/* Chebyshev semi-iteration for B x = b on [lmin, lmax], x_0 = 0, using
 * ``deg`` products with B (Golub & Van Loan sec. 11.2; Saad, Iterative
 * Methods, alg. 12.1).  With th = (lmax+lmin)/2, de = (lmax-lmin)/2,
 * s = th/de:
 *   x_1 = b/th,
 *   x_{k+1} = x_{k-1} + w_{k+1} (x_k - x_{k-1} + (b - B x_k)/th),
 *   w_1 = 2, w_{k+1} = 1/(1 - w_k/(4 s^2)),
 * so that b - B x_k = T_k((th - B)/de) b / T_k(s).  Returns x_{deg+1}.
 * work: 3n. */
void or_chebiter(const or_csr *B, double lmin, double lmax, int deg, const double *b, double *x,
                 double *work) {
  const int64_t n = B->n_rows;
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), s = th / de;
  double *xk = work, *xm = work + n, *bx = work + 2 * n;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    xm[i] = 0.0;
    xk[i] = b[i] / th;
  }
  double w = 2.0;
  for (int k = 1; k <= deg; ++k) {
    w = 1.0 / (1.0 - w / (4.0 * s * s));
    or_spmv(B, xk, bx);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double xn = xm[i] + w * (xk[i] - xm[i] + (b[i] - bx[i]) / th);
      xm[i] = xn; /* xm now holds x_{k+1} */
    }
    double *t = xk;
    xk = xm;
    xm = t;
  }
  memcpy(x, xk, sizeof(double) * n);
}

/* x ~ B~^{-1} b (identity for a standard problem); work 3n */
void or_bsolve(const or_prob *p, const double *b, double *x, double *work) {
  if (!p->B.rowptr) {
    memcpy(x, b, sizeof(double) * p->n);
    return;
  }
  or_chebiter(&p->B, p->blmin, p->blmax, p->bdeg, b, x, work);
}

/* ------------------------------------------------------------------ */
/* dense symmetric helpers                                             */

/* Eigenvalues of the symmetric tridiagonal (d[0..m-1], e[0..m-2]) by Sturm
 * bisection; w ascending. */
void or_tridiag_eigvals(int m, const double *d, const double *e, double *w) {
  double lo = d[0], hi = d[0];
  for (int i = 0; i < m; ++i) {
    double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < m - 1 ? fabs(e[i]) : 0.0);
    if (d[i] - r < lo) lo = d[i] - r;
    if (d[i] + r > hi) hi = d[i] + r;
  }
  double span = hi - lo;
  lo -= 1e-14 * span + 1e-300;
  hi += 1e-14 * span + 1e-300;
#pragma omp parallel for schedule(dynamic, 4)
  for (int k = 0; k < m; ++k) {
    double a = lo, b = hi;
    for (int it = 0; it < 200; ++it) {
      double c = 0.5 * (a + b);
      if (c <= a || c >= b) break;
      /* count eigenvalues < c */
      int cnt = 0;
      double q = d[0] - c;
      if (q < 0) cnt++;
      for (int i = 1; i < m; ++i) {
        if (q == 0.0) q = 1e-300;
        q = d[i] - c - e[i - 1] * e[i - 1] / q;
        if (q < 0) cnt++;
      }
      if (cnt > k) b = c; else a = c;
    }
    w[k] = 0.5 * (a + b);
  }
}

/* Implicit symmetric tridiagonal QR with Wilkinson shift (Golub & Van Loan,
 * alg. 8.3.2-8.3.3), accumulating the rotations into Z (m x m, column j =
 * eigenvector j, column-major).  Eigenvalues sorted ascending on return. */
int or_tridiag_eig(int m, const double *d_in, const double *e_in, double *w, double *Z) {
  double *d = (double *)malloc(sizeof(double) * m);
  double *e = (double *)calloc(m, sizeof(double));
  memcpy(d, d_in, sizeof(double) * m);
  if (m > 1) memcpy(e, e_in, sizeof(double) * (m - 1));
  for (int i = 0; i < m * m; ++i) Z[i] = 0.0;
  for (int i = 0; i < m; ++i) Z[i * m + i] = 1.0;
  int hi = m - 1, iters = 0;
  while (hi > 0) {
    /* deflate */
    for (int i = 0; i < hi; ++i)
      if (fabs(e[i]) <= 2.2e-16 * (fabs(d[i]) + fabs(d[i + 1]))) e[i] = 0.0;
    while (hi > 0 && e[hi - 1] == 0.0) hi--;
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) lo--;
    if (++iters > 60 * m) { free(d); free(e); return -1; }
    /* Wilkinson shift from the trailing 2x2 of the block [lo, hi] */
    double tt = 0.5 * (d[hi - 1] - d[hi]);
    double bb = e[hi - 1];
    double den = tt + (tt >= 0 ? 1.0 : -1.0) * hypot(tt, bb);
    double mu = d[hi] - (den != 0.0 ? bb * bb / den : 0.0);
    double x = d[lo] - mu, z = e[lo];
    for (int k = lo; k < hi; ++k) {
      double r = hypot(x, z);
      double c = (r == 0.0) ? 1.0 : x / r, s = (r == 0.0) ? 0.0 : -z / r;
      if (k > lo) e[k - 1] = r;
      /* apply G(k,k+1,theta)^T T G */
      double a11 = d[k], a12 = e[k], a22 = d[k + 1];
      d[k] = c * c * a11 - 2 * c * s * a12 + s * s * a22;
      d[k + 1] = s * s * a11 + 2 * c * s * a12 + c * c * a22;
      e[k] = c * s * (a11 - a22) + (c * c - s * s) * a12;
      if (k < hi - 1) {
        x = e[k];
        z = -s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      for (int i = 0; i < m; ++i) {
        double zi = Z[k * m + i], zj = Z[(k + 1) * m + i];
        Z[k * m + i] = c * zi - s * zj;
        Z[(k + 1) * m + i] = s * zi + c * zj;
      }
    }
  }
  /* sort ascending */
  int *idx = (int *)malloc(sizeof(int) * m);
  for (int i = 0; i < m; ++i) idx[i] = i;
  for (int i = 1; i < m; ++i) {
    int t = idx[i], j = i - 1;
    while (j >= 0 && d[idx[j]] > d[t]) { idx[j + 1] = idx[j]; j--; }
    idx[j + 1] = t;
  }
  double *Zs = (double *)malloc(sizeof(double) * m * m);
  for (int j = 0; j < m; ++j) {
    w[j] = d[idx[j]];
    memcpy(Zs + (size_t)j * m, Z + (size_t)idx[j] * m, sizeof(double) * m);
  }
  memcpy(Z, Zs, sizeof(double) * m * m);
  free(Zs); free(idx); free(d); free(e);
  return 0;
}

/* Cyclic Jacobi for a dense symmetric S (n x n, overwritten); V column-major
 * eigenvectors, w ascending. */
int or_sym_jacobi(int n, double *S, double *w, double *V) {
  for (int i = 0; i < n * n; ++i) V[i] = 0.0;
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double a = S[i * n + j] * S[i * n + j];
        tot += a;
        if (i != j) off += a;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = S[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        double theta = (S[q * n + q] - S[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double skp = S[k * n + p], skq = S[k * n + q];
          S[k * n + p] = c * skp - s * skq;
          S[k * n + q] = s * skp + c * skq;
        }
        for (int k = 0; k < n; ++k) {
          double spk = S[p * n + k], sqk = S[q * n + k];
          S[p * n + k] = c * spk - s * sqk;
          S[q * n + k] = s * spk + c * sqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[p * n + k], vkq = V[q * n + k];
          V[p * n + k] = c * vkp - s * vkq;
          V[q * n + k] = s * vkp + c * vkq;
        }
      }
  }
  int *idx = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 1; i < n; ++i) {
    int t = idx[i], j = i - 1;
    while (j >= 0 && S[idx[j] * n + idx[j]] > S[t * n + t]) { idx[j + 1] = idx[j]; j--; }
    idx[j + 1] = t;
  }
  double *Vs = (double *)malloc(sizeof(double) * n * n);
  for (int j = 0; j < n; ++j) {
    w[j] = S[idx[j] * n + idx[j]];
    memcpy(Vs + (size_t)j * n, V + (size_t)idx[j] * n, sizeof(double) * n);
  }
  memcpy(V, Vs, sizeof(double) * n * n);
  free(Vs); free(idx);
  return 0;
}

/* Dense generalized symmetric-definite G_A w = mu G_B w via Cholesky of G_B
 * (G_B = L L^T, C = L^-1 G_A L^-T) and Jacobi; W column-major (n x n). */
int or_sygv(int n, const double *GA, const double *GB, double *mu, double *W) {
  double *L = (double *)calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j) {
    double s = GB[j * n + j];
    for (int k = 0; k < j; ++k) s -= L[j * n + k] * L[j * n + k];
    if (!(s > 0.0)) { free(L); return -1; }
    L[j * n + j] = sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      double t = GB[i * n + j];
      for (int k = 0; k < j; ++k) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / L[j * n + j];
    }
  }
  /* X = L^-1 GA (forward solve per column), C = X L^-T */
  double *X = (double *)malloc(sizeof(double) * n * n);
  for (int c = 0; c < n; ++c)
    for (int i = 0; i < n; ++i) {
      double t = GA[i * n + c];
      for (int k = 0; k < i; ++k) t -= L[i * n + k] * X[k * n + c];
      X[i * n + c] = t / L[i * n + i];
    }
  double *C = (double *)malloc(sizeof(double) * n * n);
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < n; ++i) {
      double t = X[r * n + i];
      for (int k = 0; k < i; ++k) t -= L[i * n + k] * C[r * n + k];
      C[r * n + i] = t / L[i * n + i];
    }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double s = 0.5 * (C[i * n + j] + C[j * n + i]);
      C[i * n + j] = C[j * n + i] = s;
    }
  double *Y = (double *)malloc(sizeof(double) * n * n);
  or_sym_jacobi(n, C, mu, Y);
  /* W = L^-T Y */
  for (int j = 0; j < n; ++j)
    for (int i = n - 1; i >= 0; --i) {
      double t = Y[j * n + i];
      for (int k = i + 1; k < n; ++k) t -= L[k * n + i] * W[j * n + k];
      W[j * n + i] = t / L[i * n + i];
    }
  free(L); free(X); free(C); free(Y);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Lanczos workspace: V (B-orthonormal), Z = B V (generalized) */
typedef struct {
  int64_t n;
  int cap;
  double *V, *Z;
  int gen;
} lanbasis;

static void cgs2(const lanbasis *L, int k, double *z, double *h) {
  /* z -= Z(:,0:k) (V(:,0:k)^T z), twice */
  const int64_t n = L->n;
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = 0; j < k; ++j) h[j] = dot(n, L->V + (size_t)j * n, z);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += L->Z[(size_t)j * n + i] * h[j];
      z[i] -= s;
    }
  }
}

/* Lanczos bounds of the operator (B^-1 A when geneig, else A): non-restarted
 * Lanczos with full re-orthogonalisation, convergence of both extreme Ritz
 * pairs checked every 10 steps (residual beta*|s_end| <= tol * spread);
 * returns lmin = theta_min - r_min, lmax = theta_max + r_max.  Restates
 * pEVSL_LANBOUNDS_F90(h, MLAN, LANSTEP, TOL, LMIN, LMAX) (src/mod_pevsl.f90:84,
 * src/mod_matvec.f90:83-86, 160-163). */
int or_lanbounds(const or_prob *p, int geneig, int mlan, int maxit, double tol, uint64_t seed,
                 double *lmin, double *lmax, int *nsteps) {
  const int64_t n = p->n;
  int m = mlan < maxit ? mlan : maxit;
  if (m > n) m = (int)n;
  lanbasis L = {n, m + 1, NULL, NULL, geneig};
  L.V = (double *)malloc(sizeof(double) * n * (m + 1));
  L.Z = geneig ? (double *)malloc(sizeof(double) * n * (m + 1)) : L.V;
  double *work = (double *)malloc(sizeof(double) * 4 * n);
  double *h = (double *)malloc(sizeof(double) * (m + 1));
  double *al = (double *)malloc(sizeof(double) * (m + 1));
  double *be = (double *)malloc(sizeof(double) * (m + 1));
  double *w = (double *)malloc(sizeof(double) * (m + 1));
  double *S = (double *)malloc(sizeof(double) * (size_t)(m + 1) * (m + 1));
  or_start_vector(n, seed, L.V);
  if (geneig) {
    or_apply_b(p, L.V, L.Z);
    double t = 1.0 / sqrt(dot(n, L.V, L.Z));
    scal(n, t, L.V);
    scal(n, t, L.Z);
  } else {
    scal(n, 1.0 / sqrt(dot(n, L.V, L.V)), L.V);
  }
  double beta = 0.0;
  int k, done = 0;
  for (k = 0; k < m; ++k) {
    double *v = L.V + (size_t)k * n, *z = L.Z + (size_t)k * n;
    double *vn = L.V + (size_t)(k + 1) * n, *zn = L.Z + (size_t)(k + 1) * n;
    or_progress("lanbounds", k);
    or_apply_a(p, v, zn, work);
    if (k > 0) axpy(n, -beta, L.Z + (size_t)(k - 1) * n, zn);
    double alpha = dot(n, v, zn);
    axpy(n, -alpha, z, zn);
    cgs2(&L, k + 1, zn, h);
    if (geneig) {
      or_bsolve(p, zn, vn, work);
      beta = sqrt(dot(n, vn, zn));
    } else {
      beta = sqrt(dot(n, zn, zn));
    }
    al[k] = alpha;
    be[k] = beta;
    if (beta == 0.0) { k++; done = 1; break; }
    scal(n, 1.0 / beta, vn);
    if (geneig) scal(n, 1.0 / beta, zn);
    if ((k + 1) % 10 == 0 || k == m - 1) {
      int kd = k + 1;
      or_tridiag_eig(kd, al, be, w, S);
      double rmin = beta * fabs(S[0 * kd + kd - 1]);
      double rmax = beta * fabs(S[(size_t)(kd - 1) * kd + kd - 1]);
      double spread = fabs(w[kd - 1]) + fabs(w[0]);
      *lmin = w[0] - rmin;
      *lmax = w[kd - 1] + rmax;
      if (rmin <= tol * spread && rmax <= tol * spread) { k++; done = 1; break; }
    }
  }
  if (!done || beta == 0.0) {
    int kd = k;
    or_tridiag_eig(kd, al, be, w, S);
    double rmin = beta * fabs(S[0 * kd + kd - 1]);
    double rmax = beta * fabs(S[(size_t)(kd - 1) * kd + kd - 1]);
    *lmin = w[0] - rmin;
    *lmax = w[kd - 1] + rmax;
  }
  *nsteps = k;
  free(L.V);
  if (geneig) free(L.Z);
  free(work); free(h); free(al); free(be); free(w); free(S);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Chebyshev-delta polynomial filter (restates EVSL/pEVSL find_pol, called as
 * pEVSL_FINDPOL_F90(XINTV, THRE_INT, THRE_EXT, POL), src/mod_pevsl.f90:108-115).
 * [lmin,lmax] -> [-1,1] by t = (x-cc)/dd.  p(t) = sum_k mu_k T_k(t) with
 * mu_k = g_k cos(k th_c), g_0 = 1/2, g_k the damping factors (1: Jackson,
 * 2: Lanczos sigma), normalised to p(gam) = 1, gam = cos(th_c).  Extremal
 * intervals (a within IntTol of -1 or b of +1) put gam at the end and use
 * thr_ext; interior ones choose th_c so that p(a) = p(b) (balanced) and use
 * thr_int.  The degree is the smallest m >= min_deg whose endpoint values are
 * <= threshold; bar = min over the interval endpoints of p. */
static void damping_factors(int m, int damping, double *g) {
  g[0] = 0.5;
  const double dm = (double)m;
  for (int k = 1; k <= m; ++k) {
    if (damping == 0) {
      g[k] = 1.0;
    } else if (damping == 1) {
      double thJ = M_PI / (dm + 2.0), a1 = 1.0 / (dm + 2.0), a2 = sin(thJ);
      g[k] = a1 * sin((k + 1) * thJ) / a2 + (1.0 - (k + 1) * a1) * cos(k * thJ);
    } else {
      double thL = M_PI / (dm + 1.0);
      g[k] = sin(k * thL) / (k * thL);
    }
  }
}
static double cheb_eval(int m, const double *mu, double t) {
  /* sum mu_k T_k(t), |t| <= 1 (three-term recurrence) */
  double t0 = 1.0, t1 = t, s = mu[0];
  if (m >= 1) s += mu[1] * t1;
  for (int k = 2; k <= m; ++k) {
    double t2 = 2.0 * t * t1 - t0;
    s += mu[k] * t2;
    t0 = t1;
    t1 = t2;
  }
  return s;
}
static void delta_coefs(int m, const double *g, double thc, double *mu) {
  for (int k = 0; k <= m; ++k) mu[k] = g[k] * cos(k * thc);
  double pg = cheb_eval(m, mu, cos(thc));
  for (int k = 0; k <= m; ++k) mu[k] /= pg;
}

typedef struct {
  int deg;
  int type; /* 0 interior, 1 left end, 2 right end */
  double cc, dd, gam, bar;
  double aa, bb; /* interval in [-1,1] coordinates */
} or_pol;

int or_findpol(const double *xintv, double thr_int, double thr_ext, int min_deg, int max_deg,
               int damping, double *mu_out, or_pol *pol) {
  const double IntTol = 0.0005;
  double a = xintv[0] > xintv[2] ? xintv[0] : xintv[2];
  double b = xintv[1] < xintv[3] ? xintv[1] : xintv[3];
  double lmin = xintv[2], lmax = xintv[3];
  if (!(b > a) || !(lmax > lmin)) return -1;
  double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  double aa = (a - cc) / dd, bb = (b - cc) / dd;
  if (aa < -1.0) aa = -1.0;
  if (bb > 1.0) bb = 1.0;
  double tha = acos(aa), thb = acos(bb);
  double *g = (double *)malloc(sizeof(double) * (max_deg + 1));
  double thr;
  int type;
  if (aa - IntTol <= -1.0) { type = 1; thr = thr_ext; }
  else if (bb + IntTol >= 1.0) { type = 2; thr = thr_ext; }
  else { type = 0; thr = thr_int; }
  int m, found = 0;
  double thc = 0.0;
  for (m = min_deg; m <= max_deg; ++m) {
    damping_factors(m, damping, g);
    if (type == 1) thc = M_PI;
    else if (type == 2) thc = 0.0;
    else {
      /* balance: f(th) = p_th(aa) - p_th(bb) changes sign on [thb, tha] */
      double lo = thb, hi = tha;
      for (int it = 0; it < 100; ++it) {
        double mid = 0.5 * (lo + hi);
        delta_coefs(m, g, mid, mu_out);
        double f = cheb_eval(m, mu_out, aa) - cheb_eval(m, mu_out, bb);
        /* moving the centre towards a (larger theta) raises p(aa) */
        if (f > 0.0) hi = mid; else lo = mid;
        if (hi - lo < 1e-15) break;
      }
      thc = 0.5 * (lo + hi);
    }
    delta_coefs(m, g, thc, mu_out);
    double va = cheb_eval(m, mu_out, aa), vb = cheb_eval(m, mu_out, bb);
    double vmax = type == 1 ? vb : (type == 2 ? va : (va > vb ? va : vb));
    if (vmax <= thr) { found = 1; break; }
  }
  free(g);
  if (!found) return -2;
  double va = cheb_eval(m, mu_out, aa), vb = cheb_eval(m, mu_out, bb);
  pol->deg = m;
  pol->type = type;
  pol->cc = cc;
  pol->dd = dd;
  pol->gam = cos(thc);
  pol->bar = type == 1 ? vb : (type == 2 ? va : (va < vb ? va : vb));
  pol->aa = aa;
  pol->bb = bb;
  return 0;
}

/* y = p((A B^-1 - cc)/dd) z in the z-space of the generalized problem
 * (B p(B^-1 A) v for z = B v), or p((A - cc)/dd) z for a standard one:
 * v_0 = z, v_1 = (M v_0 - cc v_0)/dd, v_{k+1} = 2(M v_k - cc v_k)/dd - v_{k-1},
 * y = sum mu_k v_k, M = A B^-1.  work: 8n */
void or_chebav(const or_prob *p, int geneig, int deg, const double *mu, double cc, double dd,
               const double *z, double *y, double *work) {
  const int64_t n = p->n;
  double *vk = work, *vkm = work + n, *vkp = work + 2 * n, *tmp = work + 3 * n, *w4 = work + 4 * n;
  memcpy(vk, z, sizeof(double) * n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    vkm[i] = 0.0;
    y[i] = mu[0] * vk[i];
  }
  for (int k = 1; k <= deg; ++k) {
    double t = k == 1 ? 1.0 / dd : 2.0 / dd;
    if (geneig) {
      or_bsolve(p, vk, tmp, w4);
      or_apply_a(p, tmp, vkp, w4);
    } else {
      or_apply_a(p, vk, vkp, w4);
    }
    const double muk = mu[k];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double v = t * (vkp[i] - cc * vk[i]) - (k == 1 ? 0.0 : vkm[i]);
      vkp[i] = v;
      y[i] += muk * v;
    }
    double *s = vkm;
    vkm = vk;
    vk = vkp;
    vkp = s;
  }
}

/* Residuals of eigenpairs (x B~-normalised on entry or not): for each j,
 *   res_ref[j] = sqrt(sum_i r_i^2 / n) / |lam_j|   (src/mod_pevsl.f90:137-158)
 *   res2[j]    = ||r||_2 / |lam_j| with x scaled to x^T B~ x = 1
 * r = A~ x - lam B~ x. */
void or_residuals(const or_prob *p, int nev, const double *lam, const double *X, double *res_ref,
                  double *res2) {
  const int64_t n = p->n;
  double *ax = (double *)malloc(sizeof(double) * n), *bx = (double *)malloc(sizeof(double) * n);
  double *work = (double *)malloc(sizeof(double) * n);
  for (int j = 0; j < nev; ++j) {
    const double *x = X + (size_t)j * n;
    or_apply_a(p, x, ax, work);
    or_apply_b(p, x, bx);
    double s = 0.0, xb = 0.0;
#pragma omp parallel for reduction(+ : s, xb) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double r = ax[i] - lam[j] * bx[i];
      s += r * r;
      xb += x[i] * bx[i];
    }
    double al = fabs(lam[j]) > 0 ? fabs(lam[j]) : 1.0;
    res_ref[j] = sqrt(s / (double)n) / al;
    res2[j] = sqrt(s / xb) / al;
  }
  free(ax); free(bx); free(work);
}

/* Non-restarted Chebyshev-filtered Lanczos with full re-orthogonalisation
 * (restates pEVSL ChebLanNr, called as pEVSL_CHEBLANNR_F90(h, XINTV, MAXIT,
 * TOL, POL), src/mod_pevsl.f90:122).  Generalized form works on the pair
 * (v, z = B v) with B-inner products:
 *   z_{k+1} = p(A B^-1) z_k - beta_k z_{k-1} - alpha_k z_k,
 *   alpha_k = v_k^T p(A B^-1) z_k, CGS2 against (Z, V), v_{k+1} = B^-1 z_{k+1},
 *   beta_{k+1} = sqrt(v_{k+1}^T z_{k+1}).
 * Convergence is tested from step ntest on every ``cycle`` steps: the sum of
 * the Ritz values of T that are >= bar has changed by < tol relatively and
 * each of them has Ritz residual beta |s_last| <= rtol.  Extraction: Ritz
 * vectors Y = V S for Ritz values >= bar, Rayleigh-Ritz of (A~, B~) on span(Y),
 * keep mu in [a, b].  Outputs ascending eigenvalues lam[nev], eigenvectors
 * X (n x nev, B~-normalised, scaled coordinates). */
int or_cheblannr(const or_prob *p, int geneig, const double *xintv, int maxit, double tol,
                 double rtol, int ntest, int cycle, int deg, const double *mu, const or_pol *pol,
                 uint64_t seed, int max_nev, int *nev_out, double *lam, double *X, int *nsteps_out,
                 double *ncand_out) {
  const int64_t n = p->n;
  if (maxit > n) maxit = (int)n;
  lanbasis L = {n, maxit + 1, NULL, NULL, geneig};
  L.V = (double *)malloc(sizeof(double) * n * (maxit + 1));
  L.Z = geneig ? (double *)malloc(sizeof(double) * n * (maxit + 1)) : L.V;
  double *work = (double *)malloc(sizeof(double) * 8 * n);
  double *h = (double *)malloc(sizeof(double) * (maxit + 1));
  double *al = (double *)malloc(sizeof(double) * (maxit + 1));
  double *be = (double *)malloc(sizeof(double) * (maxit + 1));
  double *w = (double *)malloc(sizeof(double) * (maxit + 1));
  or_start_vector(n, seed, L.V);
  if (geneig) {
    or_apply_b(p, L.V, L.Z);
    double t = 1.0 / sqrt(dot(n, L.V, L.Z));
    scal(n, t, L.V);
    scal(n, t, L.Z);
  } else {
    scal(n, 1.0 / sqrt(dot(n, L.V, L.V)), L.V);
  }
  const double bar = pol->bar;
  double beta = 0.0, tr0 = 0.0;
  int k, kdim = 0;
  for (k = 0; k < maxit; ++k) {
    double *v = L.V + (size_t)k * n, *z = L.Z + (size_t)k * n;
    double *vn = L.V + (size_t)(k + 1) * n, *zn = L.Z + (size_t)(k + 1) * n;
    or_progress("cheblannr", k);
    or_chebav(p, geneig, deg, mu, pol->cc, pol->dd, z, zn, work);
    if (k > 0) axpy(n, -beta, L.Z + (size_t)(k - 1) * n, zn);
    double alpha = dot(n, v, zn);
    axpy(n, -alpha, z, zn);
    cgs2(&L, k + 1, zn, h);
    if (geneig) {
      or_bsolve(p, zn, vn, work);
      beta = sqrt(dot(n, vn, zn));
    } else {
      beta = sqrt(dot(n, zn, zn));
    }
    al[k] = alpha;
    be[k] = beta;
    kdim = k + 1;
    if (beta == 0.0) break;
    scal(n, 1.0 / beta, vn);
    if (geneig) scal(n, 1.0 / beta, zn);
    if (kdim < ntest || ((kdim - ntest) % cycle) != 0) continue;
    or_tridiag_eigvals(kdim, al, be, w);
    double tr1 = 0.0;
    int cnt = 0;
    for (int i = 0; i < kdim; ++i)
      if (w[i] >= bar) { tr1 += w[i]; cnt++; }
    if (cnt == 0) {
      /* empty interval: the largest Ritz value has converged below bar */
      double *S0 = (double *)malloc(sizeof(double) * (size_t)kdim * kdim);
      double *w0 = (double *)malloc(sizeof(double) * kdim);
      or_tridiag_eig(kdim, al, be, w0, S0);
      int imax = 0;
      for (int i = 1; i < kdim; ++i)
        if (w0[i] > w0[imax]) imax = i;
      const double r0 = beta * fabs(S0[(size_t)imax * kdim + kdim - 1]);
      free(S0); free(w0);
      tr0 = 0.0;
      if (r0 <= rtol) { k++; break; }
      continue;
    }
    int conv = fabs(tr1 - tr0) < tol * fabs(tr1);
    tr0 = tr1;
    if (conv) {
      /* residual check of the candidates */
      double *S = (double *)malloc(sizeof(double) * (size_t)kdim * kdim);
      double *ww = (double *)malloc(sizeof(double) * kdim);
      or_tridiag_eig(kdim, al, be, ww, S);
      double rmax = 0.0;
      for (int i = 0; i < kdim; ++i)
        if (ww[i] >= bar) {
          double r = beta * fabs(S[(size_t)i * kdim + kdim - 1]);
          if (r > rmax) rmax = r;
        }
      free(S); free(ww);
      if (rmax <= rtol) { k++; break; }
    }
  }
  *nsteps_out = kdim;
  /* extraction */
  double *S = (double *)malloc(sizeof(double) * (size_t)kdim * kdim);
  or_tridiag_eig(kdim, al, be, w, S);
  int nc = 0;
  for (int i = 0; i < kdim; ++i)
    if (w[i] >= bar) nc++;
  *ncand_out = nc;
  int nev = 0;
  if (nc > 0) {
    double *Y = (double *)malloc(sizeof(double) * n * nc);
    int c = 0;
    for (int i = 0; i < kdim; ++i) {
      if (w[i] < bar) continue;
      const double *s = S + (size_t)i * kdim;
      double *y = Y + (size_t)c * n;
#pragma omp parallel for schedule(static)
      for (int64_t r = 0; r < n; ++r) {
        double t = 0.0;
        for (int j = 0; j < kdim; ++j) t += L.V[(size_t)j * n + r] * s[j];
        y[r] = t;
      }
      c++;
    }
    /* Rayleigh-Ritz of (A~, B~) on span(Y) */
    double *AY = (double *)malloc(sizeof(double) * n * nc);
    double *BY = (double *)malloc(sizeof(double) * n * nc);
    for (int j = 0; j < nc; ++j) {
      or_apply_a(p, Y + (size_t)j * n, AY + (size_t)j * n, work);
      or_apply_b(p, Y + (size_t)j * n, BY + (size_t)j * n);
    }
    double *GA = (double *)malloc(sizeof(double) * nc * nc);
    double *GB = (double *)malloc(sizeof(double) * nc * nc);
    for (int i = 0; i < nc; ++i)
      for (int j = i; j < nc; ++j) {
        double ga = dot(n, Y + (size_t)i * n, AY + (size_t)j * n);
        double gb = dot(n, Y + (size_t)i * n, BY + (size_t)j * n);
        GA[i * nc + j] = GA[j * nc + i] = ga;
        GB[i * nc + j] = GB[j * nc + i] = gb;
      }
    double *rmu = (double *)malloc(sizeof(double) * nc);
    double *W = (double *)malloc(sizeof(double) * nc * nc);
    if (or_sygv(nc, GA, GB, rmu, W) == 0) {
      for (int j = 0; j < nc; ++j) {
        if (rmu[j] < xintv[0] || rmu[j] > xintv[1]) continue;
        if (nev >= max_nev) break;
        double *x = X + (size_t)nev * n;
        const double *wc = W + (size_t)j * nc;
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; ++r) {
          double t = 0.0;
          for (int i = 0; i < nc; ++i) t += Y[(size_t)i * n + r] * wc[i];
          x[r] = t;
        }
        lam[nev] = rmu[j];
        nev++;
      }
    }
    free(Y); free(AY); free(BY); free(GA); free(GB); free(rmu); free(W);
  }
  *nev_out = nev;
  free(S);
  free(L.V);
  if (geneig) free(L.Z);
  free(work); free(h); free(al); free(be); free(w);
  return 0;
}

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int t) { omp_set_num_threads(t); }
double or_wtime(void) { return omp_get_wtime(); }
