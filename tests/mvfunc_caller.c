/* A pEVSL-style caller that reaches the GPU operators only through the
 * MVFunc-shaped callbacks nmg_amv / nmg_bmv / nmg_bsol (the reference hands
 * sparseAV / sparseBV to pEVSL as such callbacks, src/mod_pevsl.f90:69-80).
 *
 * Usage: mvfunc_caller DIR N M
 *   DIR holds A_rowptr.i64, A_col.i32, A_val.f64, B_* (unscaled CSR) and x.f64.
 * It sets up the operators like setupmatvec (D from B, B~ bounds, Chebyshev
 * B-solve of degree 58, A~ = D A D), then:
 *   - writes amv(x), bmv(x), bsol(x) to DIR/y_amv.f64, y_bmv.f64, y_bsol.f64;
 *   - runs M steps of its own Lanczos on B~^-1 A~ with B~-inner products and
 *     full re-orthogonalisation, calling only the three callbacks, and
 *     writes the tridiagonal's alpha (M) and beta (M-1) to DIR/alpha.f64,
 *     DIR/beta.f64.
 * Exit status 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nmgpu.h"

typedef void (*MVFunc)(double *x, double *y, void *data);

static void *load(const char *dir, const char *name, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  void *p = malloc(bytes ? bytes : 1);
  if (fread(p, 1, bytes, f) != bytes) { fprintf(stderr, "short read %s\n", path); exit(2); }
  fclose(f);
  return p;
}

static size_t fsize(const char *dir, const char *name) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  fseek(f, 0, SEEK_END);
  long s = ftell(f);
  fclose(f);
  return (size_t)s;
}

static void save(const char *dir, const char *name, const double *p, size_t n) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(p, sizeof(double), n, f) != n) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
  fclose(f);
}

#define CHECK(call)                                                              \
  do {                                                                           \
    int rc_ = (call);                                                            \
    if (rc_ != 0) {                                                              \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, nmg_last_error());     \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

static double dot(long n, const double *a, const double *b) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s DIR N M\n", argv[0]); return 2; }
  const char *dir = argv[1];
  const long n = atol(argv[2]);
  const int m = atoi(argv[3]);
  const size_t nnzA = fsize(dir, "A_col.i32") / 4, nnzB = fsize(dir, "B_col.i32") / 4;
  int64_t *arp = load(dir, "A_rowptr.i64", 8 * (n + 1)), *brp = load(dir, "B_rowptr.i64", 8 * (n + 1));
  int32_t *acol = load(dir, "A_col.i32", 4 * nnzA), *bcol = load(dir, "B_col.i32", 4 * nnzB);
  double *aval = load(dir, "A_val.f64", 8 * nnzA), *bval = load(dir, "B_val.f64", 8 * nnzB);
  double *x = load(dir, "x.f64", 8 * n);

  nmg_handle matA, matB, hB, h, cheb;
  double *d = malloc(8 * n);
  CHECK(nmg_csr_create(n, n, arp, acol, aval, &matA));
  CHECK(nmg_csr_create(n, n, brp, bcol, bval, &matB));
  CHECK(nmg_csr_diagscale(matB, 1.0, d));
  double bl, bu;
  CHECK(nmg_start(0, &hB));
  CHECK(nmg_setprobsizes(hB, n, n, -1));
  CHECK(nmg_set_amv(hB, matB));
  CHECK(nmg_lanbounds(hB, 1000, 2000, 1e-12, &bl, &bu));
  CHECK(nmg_finish(hB));
  CHECK(nmg_setup_chebiter(bl, bu, 58, matB, &cheb));
  CHECK(nmg_start(0, &h));
  CHECK(nmg_setprobsizes(h, n, n, -1));
  CHECK(nmg_set_bmv(h, matB));
  CHECK(nmg_set_bsol_chebiter(h, 2, cheb));
  CHECK(nmg_set_amv_dad(h, matA, d));
  CHECK(nmg_set_geneig(h));

  /* from here on only the callbacks, as a pEVSL caller would use them */
  MVFunc amv = nmg_amv, bmv = nmg_bmv, bsol = nmg_bsol;
  void *data = &h;
  double *y = malloc(8 * n);
  amv(x, y, data);
  save(dir, "y_amv.f64", y, n);
  bmv(x, y, data);
  save(dir, "y_bmv.f64", y, n);
  bsol(x, y, data);
  save(dir, "y_bsol.f64", y, n);

  /* Lanczos on B~^-1 A~: v_1 B~-normalised, z_j = B~ v_j,
     w = A~ v_j - beta_j z_{j-1} - alpha_j z_j, CGS2 against Z in the V metric,
     v_{j+1} = B~^-1 w / beta_{j+1}, beta_{j+1} = sqrt(w^T B~^-1 w) */
  double *V = malloc(8 * n * (size_t)(m + 1)), *Z = malloc(8 * n * (size_t)(m + 1));
  double *w = malloc(8 * n), *alpha = malloc(8 * m), *beta = malloc(8 * m);
  for (long i = 0; i < n; ++i) V[i] = 1.0 + 0.5 * sin(0.37 * (double)i);
  bmv(V, Z, data);
  double nb = sqrt(dot(n, V, Z));
  for (long i = 0; i < n; ++i) { V[i] /= nb; Z[i] /= nb; }
  for (int j = 0; j < m; ++j) {
    double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n;
    amv(vj, w, data);
    if (j > 0)
      for (long i = 0; i < n; ++i) w[i] -= beta[j - 1] * Z[(size_t)(j - 1) * n + i];
    alpha[j] = dot(n, vj, w);
    for (long i = 0; i < n; ++i) w[i] -= alpha[j] * zj[i];
    for (int pass = 0; pass < 2; ++pass)
      for (int k = 0; k <= j; ++k) {
        const double c = dot(n, V + (size_t)k * n, w);
        for (long i = 0; i < n; ++i) w[i] -= c * Z[(size_t)k * n + i];
      }
    double *vn = V + (size_t)(j + 1) * n, *zn = Z + (size_t)(j + 1) * n;
    bsol(w, vn, data);
    const double b2 = dot(n, vn, w);
    if (!(b2 > 0.0)) { fprintf(stderr, "breakdown at step %d\n", j); return 1; }
    beta[j] = sqrt(b2);
    for (long i = 0; i < n; ++i) { vn[i] /= beta[j]; zn[i] = w[i] / beta[j]; }
  }
  save(dir, "alpha.f64", alpha, m);
  save(dir, "beta.f64", beta, m - 1);
  CHECK(nmg_finish(h));
  CHECK(nmg_chebiter_free(cheb));
  CHECK(nmg_csr_free(matA));
  CHECK(nmg_csr_free(matB));
  printf("mvfunc caller: n=%ld, %d Lanczos steps through the callbacks\n", n, m);
  return 0;
}
