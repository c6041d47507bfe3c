/*
 * nmgpu — MI355X (gfx950) filtered-Lanczos eigensolver for the generalized
 * problem A~ x = lambda B~ x of planetary normal modes.
 *
 * A C-ABI shaped after the pEVSL Fortran-90 interface that NormalModes calls
 * (src/mod_pevsl.f90, src/mod_matvec.f90).  Each entry names the reference
 * call it replaces.  Conventions:
 *   - every function returns 0 on success or a negative NMG_E* code, and
 *     nmg_last_error() gives the message of the last failure on the calling
 *     thread; no entry aborts the process;
 *   - handles are opaque 64-bit integers (the reference keeps pEVSL handles
 *     in integer*8, src/mod_cg_datatype.f90:92-96);
 *   - CSR input is 0-based (rowdist(1) = 0, col-1; src/mod_matvec.f90:59);
 *   - array arguments are HOST pointers unless the name ends in _dev;
 *   - one HIP stream per solver handle; calls on one handle are serialised by
 *     the caller (pEVSL is not re-entrant either).
 */
#ifndef NMGPU_H
#define NMGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int64_t nmg_handle;

enum {
  NMG_OK = 0,
  NMG_EINVAL = -1,   /* bad argument */
  NMG_EHIP = -2,     /* HIP runtime / library failure */
  NMG_ENOMEM = -3,   /* device allocation failed */
  NMG_ESTATE = -4,   /* call out of order (e.g. no operator set) */
  NMG_ENUM = -5,     /* numerical breakdown (non-SPD B, NaN, no convergence) */
  NMG_ENOTIMPL = -6  /* requested variant not available */
};

/* precision of the Chebyshev filter (nmg_set_precision) */
enum { NMG_FP64 = 0, NMG_FP32_FILTER = 1, NMG_FP32_BSOLVE = 2, NMG_FP32_FILTER_A = 3 };

/* Error message of the last failing call on this thread, library version,
 * and the device's name, CU count and HBM size (no reference counterpart:
 * pEVSL reports no errors). */
const char *nmg_last_error(void);
int nmg_version(int *major, int *minor);
int nmg_device_info(char *name, int name_len, int *cu_count, double *hbm_gbytes);

/* --- solver handle ------------------------------------------------------ */
/* pEVSL_Start_F90(comm, h)      src/mod_pevsl.f90:54; src/mod_matvec.f90:75,152 */
int nmg_start(int comm, nmg_handle *h);
/* PEVSL_FINISH_F90(h)           src/mod_pevsl.f90:220; src/mod_matvec.f90:95 */
int nmg_finish(nmg_handle h);
/* PEVSL_SETPROBSIZES_F90(h, N, n_loc, nfirst)   src/mod_pevsl.f90:57 */
int nmg_setprobsizes(nmg_handle h, int64_t n_glob, int64_t n_loc, int64_t nfirst);

/* --- matrices ----------------------------------------------------------- */
/* PEVSL_PARCSRCREATE_F90(nrow_g, ncol_g, row_starts, col_starts, ia, ja, a,
 * comm, mat)   src/mod_matvec.f90:69-71, 117-119, 146-148, 196-198, 218-220,
 * 242-244.  One process owns all rows (row_starts = col_starts = {0, n}).
 * ia: int32[nrow+1] 0-based offsets, ja: int32[nnz] 0-based columns, a: f64.
 * The matrix is copied to HBM once; the layout (3x3 block CSR, scalar mass
 * x I3, or CSR) is chosen from its structure. */
int nmg_parcsr_create(int nrow_g, int ncol_g, const int *row_starts, const int *col_starts,
                      const int *ia, const int *ja, const double *a, int comm, nmg_handle *mat);
/* the same with int64 row offsets, for nnz beyond 2^31 (no reference
 * counterpart: its rowdist is int32, src/mod_cg_datatype.f90:40) */
int nmg_csr_create(int64_t nrow, int64_t ncol, const int64_t *rowptr, const int32_t *col,
                   const double *val, nmg_handle *mat);
/* releases the device copy (no reference counterpart: the reference never
 * frees its ParCSR handles before PEVSL_FINISH_F90) */
int nmg_csr_free(nmg_handle mat);
/* layout chosen at creation: 0 CSR, 1 3x3-block CSR, 2 scalar mass (x) I3
 * (no reference counterpart) */
int nmg_csr_info(nmg_handle mat, int64_t *nrow, int64_t *ncol, int64_t *nnz, int *layout);
/* pevsl_parcsrmatvec_f90(x, y, mat)  src/mod_matvec.f90:453,471,495,507,509,515 */
int nmg_parcsr_matvec(const double *x, double *y, nmg_handle mat);
/* Bdiagscaling / Apdiagscaling  src/mod_matvec.f90:252-342, 345-441:
 * d_i = 1/sqrt(sign * M_ii), M_ij <- sign * d_i M_ij d_j, done in HBM.
 * sign = +1 for B, -1 for Ap.  d: host f64[nrow] out (may be NULL). */
int nmg_csr_diagscale(nmg_handle mat, double sign, double *d);

/* --- operators of the eigenproblem ------------------------------------- */
/* PEVSL_SETAMV_F90(h, sparseAV, mymatvec): w = D A D v  src/mod_pevsl.f90:77,
 * src/mod_matvec.f90:445-458.  A is used through a scaled copy D A D. */
int nmg_set_amv_dad(nmg_handle h, nmg_handle matA, const double *d);

/* sparsefsAV, src/mod_matvec.f90:498-520 (PEVSL_SETAMV_F90(pevslAB,
 * sparsefsAV, mymatvec), src/mod_pevsl.f90:77-80): the fluid-solid operator
 *   w = D [Ad (D v) + E (Dp C_p (Dp ET (D v)))] = D (Ad - E Ap^-1 ET) D v,
 * C_p = the Chebyshev solve chebAp (nmg_setup_chebiter on Ap~ = Dp (-Ap) Dp,
 * the matrix scaled by nmg_csr_diagscale(matAp, -1, dp) as Apdiagscaling,
 * src/mod_matvec.f90:345-441).  matAd: n x n; matE: n x np; matET: np x n
 * (src/mod_matvec.f90:98-222).  d: the B scaling D (n), dp: Dp (np). */
int nmg_set_amv_fs(nmg_handle h, nmg_handle matAd, nmg_handle matE, nmg_handle matET, nmg_handle chebAp,
                   const double *d, const double *dp);
/* PEVSL_SETAMV_F90(h, sparseBV|sparseApV, ...) with a plain matrix:
 * w = M v  (used for bounds of B~, src/mod_matvec.f90:80, and of Ap~ :157) */
int nmg_set_amv(nmg_handle h, nmg_handle mat);
/* PEVSL_SETBMV_F90(h, sparseBV, mymatvec)  src/mod_pevsl.f90:69 */
int nmg_set_bmv(nmg_handle h, nmg_handle matB);
/* PEVSL_SETBSOL_CHEBITER_F90(h, type, chebB)  src/mod_pevsl.f90:72-73 */
int nmg_set_bsol_chebiter(nmg_handle h, int type, nmg_handle cheb);
/* PEVSL_SET_GENEIG_F90(h)  src/mod_pevsl.f90:82 */
int nmg_set_geneig(nmg_handle h);
/* Filter precision (no reference counterpart: pEVSL is FP64 only).
 * NMG_FP32_FILTER: the Chebyshev B-solve inside the filter reads FP32 values
 * of B~ (FP64 vectors, FP64 row accumulation); A~, the Lanczos recurrence,
 * re-orthogonalisation, Rayleigh-Ritz and residuals stay FP64.
 * NMG_FP32_BSOLVE: as NMG_FP32_FILTER, and the k-column B-solves of the
 * block and slice paths also hold b and their iterates in FP32 (each
 * product and update formed in FP64, rounded once on store): half the
 * vector bytes of every Chebyshev step.
 * NMG_FP32_FILTER_A: as NMG_FP32_FILTER, and the filter's stiffness products
 * (A~, or Ad, E, ET and the Ap~ solve of the fluid-solid operator) read FP32
 * values with FP64 accumulation and FP64 vectors; Rayleigh-Ritz and the
 * residual check stay FP64.  Call after nmg_set_bsol_chebiter and the
 * operator (nmg_set_amv*). */
int nmg_set_precision(nmg_handle h, int mode);
/* Lanczos start vector seed (pEVSL draws a random vinit). */
int nmg_set_seed(nmg_handle h, uint64_t seed);

/* --- Chebyshev iteration for B~^{-1} ----------------------------------- */
/* pEVSL_SETUP_CHEBITER_F90(lmin, lmax, deg, mat, cheb)
 * src/mod_matvec.f90:88-93, 167-174.  deg = number of products with mat. */
int nmg_setup_chebiter(double lmin, double lmax, int deg, nmg_handle mat, nmg_handle *cheb);
/* pevsl_chebiter_f90(type, b, x, cheb)  src/mod_matvec.f90:480, 512 */
int nmg_chebiter(int type, const double *b, double *x, nmg_handle cheb);
/* The same solve on device-resident vectors without allocating: no
 * reference counterpart; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  b_dev and x_dev are device pointers of nrow doubles on the current
 * device, stream a hipStream_t (NULL: the null stream); the call is enqueued
 * on the stream and returns without synchronizing.  The second iterate lives
 * in a fixed 32 MiB device array per device, so nrow <= 4194304 (larger
 * matrices are rejected with NMG_EINVAL).  Every call on one device shares
 * that array: calls that overlap on one device race. */
int nmg_chebiter_noalloc(int type, const double *b_dev, double *x_dev, nmg_handle cheb, void *stream);
/* pEVSL_chebiterstatsprint_f90(cheb)  src/mod_pevsl.f90:132-135 */
int nmg_chebiter_stats(nmg_handle cheb, int64_t *ncalls, int64_t *nmatvec, double *seconds);
/* releases a B-solve (no reference counterpart: chebB lives until the end
 * of the run) */
int nmg_chebiter_free(nmg_handle cheb);
/* How the solve runs (no reference counterpart): mode 0 (default) launches
 * one fused kernel per step; mode 1 runs all deg steps in one persistent
 * cooperative launch when the matrix is a scalar mass x I3 whose rows fit
 * the workgroups' LDS (rows and values are read from HBM once per solve,
 * steps are separated by grid barriers; slower on MI355X at the sizes
 * measured, because of the barrier cost). */
int nmg_chebiter_set_mode(nmg_handle cheb, int mode);
/* --- device FE producer (solid regions) --------------------------------- */
/* The element stiffness and mass of CGE3D_ISO (src/mod_cg_create_matrix.f90
 * :980-1266) formed on the GPU, one wavefront per element, and gathered into
 * the CSR values of matrixstruct's pattern (:1269-1455) in ascending element
 * order per row (deterministic; A exactly symmetric).  Host arrays in and
 * out: nodes [n_nodes*3], elems [n_elem*np] (0-based, reference local
 * order), vslot [4] vertex slots, lam/mu/rho [n_elem*np], g [n_elem*np*3]
 * reference gravity (km/s^2) or NULL (JOB=1), refM [np*np], refD [3*np*np],
 * node adjacency adj_ptr [n_nodes+1] / adj (sorted), the row offsets of A
 * (3x3 blocks over the adjacency) and B (scalar mass per component); fills
 * a_val and b_val.  np = 4 or 10.  Fails on a flat element. */
int nmg_assemble_elastic(int64_t n_nodes, int64_t n_elem, int np, const double *nodes, const int32_t *elems,
                         const int32_t *vslot, const double *lam, const double *mu, const double *rho,
                         const double *g, const double *refM, const double *refD, const int64_t *adj_ptr,
                         const int32_t *adj, const int64_t *a_rowptr, double *a_val, const int64_t *b_rowptr,
                         double *b_val);

/* --- device FE producer (fluid-solid bodies) ------------------------------ */
/* Ad, B, E, ET and Ap of CGFSE3D_ISO (src/mod_cg_create_matrix.f90:103-977)
 * on the layout of matrixstruct_general (:1458-2033), formed on the GPU:
 * per-element geometry and face data and the fluid element matrices one
 * thread per element, the solid elements' stiffness with the solid
 * producer's kernel, and every row gathered one thread per node in
 * ascending element order (deterministic; E = ET^T exactly).  The arguments
 * and outputs are those of the CPU producer nm_assemble_fs (host arrays):
 * nfp face nodes (3 or 6), massF [4*nfp*nfp] and fmask [4*nfp] of the
 * reference faces, efluid [n_elem], neigh [n_elem*4] (-1 on the boundary),
 * node statuses vstat and the DoF layout vstt / pstt, the five row-offset
 * arrays (checked against the statuses and the adjacency); fills the
 * columns and values, and stats[4] (interface elements, elements with four
 * fluid-solid vertices, fluid elements, solid elements) when not NULL.
 * Fails on a flat element, and on a fluid element on the free surface when
 * g is NULL. */
int nmg_assemble_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double *nodes, const int32_t *elems,
                    const int32_t *vslot, const double *lam, const double *mu, const double *rho, const double *g,
                    const int8_t *efluid, const int32_t *neigh, int purefluid, const double *refM, const double *refD,
                    const double *massF, const int32_t *fmask, const int64_t *adj_ptr, const int32_t *adj,
                    const int8_t *vstat, const int64_t *vstt, const int64_t *pstt, const int64_t *ad_ptr,
                    int32_t *ad_col, double *ad_val, const int64_t *b_ptr, int32_t *b_col, double *b_val,
                    const int64_t *e_ptr, int32_t *e_col, double *e_val, const int64_t *et_ptr, int32_t *et_col,
                    double *et_val, const int64_t *ap_ptr, int32_t *ap_col, double *ap_val, int64_t *stats);

/* Adjoint of a fluid-solid body (the fluid-solid counterpart of
 * nmg_backward; the element terms are CGFSE3D_ISO's,
 * src/mod_cg_create_matrix.f90:103-977): for k eigenpairs (eig[j],
 * U[j*n_dof ...]) of (Ad - E Ap^-1 ET) u = lambda B u, u physical and
 * B-normalised, and P[j*n_p ...] = -Ap^-1 ET u, writes
 * d lambda_j / d m_e = w^T (dK_e/dm - lambda_j dM_e/dm) w, w = (u, p), for m
 * in {rho, lam (Lame), mu} changed uniformly at the nodes of element e, into
 * d_rho, d_lam, d_mu [j*n_elem + e].  Fluid elements: the derivatives of the
 * fluid element matrix and the interface gravity block by forward-mode dual
 * numbers (one thread per element and parameter), d_mu = 0; solid elements:
 * the element stiffness at unit parameters (gravity linear in rho).  The
 * body is described as for nmg_assemble_fs (mu is not needed); n_dof and n_p
 * follow from vstat. */
int nmg_backward_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double *nodes, const int32_t *elems,
                    const int32_t *vslot, const double *lam, const double *rho, const double *g,
                    const int8_t *efluid, const int32_t *neigh, int purefluid, const double *refM, const double *refD,
                    const double *massF, const int32_t *fmask, const int8_t *vstat, const int64_t *vstt,
                    const int64_t *pstt, int k, const double *eig, const double *U, const double *P, double *d_rho,
                    double *d_lam, double *d_mu);

/* Mixed-precision refinement of the Chebyshev solve (no reference
 * counterpart): with passes >= 1 every solve runs `passes` inner Chebyshev
 * solves of degree inner_deg on FP32 values and FP32 iterates (FP64 row
 * accumulation), pass p on the FP64 residual b - B~x of the sum of the
 * previous passes, the sum kept in FP64.  Each pass divides the error by
 * about max(2 rho^inner_deg, the FP32 floor ~1e-6), so two passes reach
 * ~1e-12 with two FP32 solves and one FP64 product.  passes = 0 restores the
 * plain FP64 solve of degree deg.  Applies to the solver's B-solve and to
 * the Ap~ solve of the fluid-solid operator alike (nmg_chebiter_noalloc
 * keeps the plain form). */
int nmg_chebiter_set_refine(nmg_handle cheb, int passes, int inner_deg);
/* persistent: 1 if the persistent solve is planned and selected, 2 if it
 * could be selected (mode 0 now), 0 if it does not apply; grid, LDS bytes
 * (no reference counterpart) */
int nmg_chebiter_info(nmg_handle cheb, int fp32, int *persistent, int *grid, int *lds_bytes);

/* --- eigensolver ------------------------------------------------------- */
/* pEVSL_LANBOUNDS_F90(h, mlan, maxit, tol, lmin, lmax)
 * src/mod_pevsl.f90:84; src/mod_matvec.f90:85, 162 */
int nmg_lanbounds(nmg_handle h, int mlan, int maxit, double tol, double *lmin, double *lmax);
/* pEVSL_FINDPOL_F90(xintv, thr_int, thr_ext, pol)  src/mod_pevsl.f90:115 */
int nmg_findpol(const double *xintv, double thr_int, double thr_ext, nmg_handle *pol);
/* polynomial parameters of a FindPol result: degree, centre/half-width of
 * the spectrum map, delta centre gam, acceptance bar; mu: f64[deg+1] or
 * NULL (no reference counterpart: POL is opaque to the reference) */
int nmg_pol_info(nmg_handle pol, int *deg, double *cc, double *dd, double *gam, double *bar,
                 double *mu);
/* pEVSL_FREEPOL_F90(pol)  src/mod_pevsl.f90:215 */
int nmg_freepol(nmg_handle pol);
/* pEVSL_CHEBLANNR_F90(h, xintv, maxit, tol, pol)  src/mod_pevsl.f90:122 */
int nmg_cheblannr(nmg_handle h, const double *xintv, int maxit, double tol, nmg_handle pol);
/* pEVSL_GET_NEV_F90(h, nev)  src/mod_pevsl.f90:124 */
int nmg_get_nev(nmg_handle h, int *nev);
/* pEVSL_COPY_RESULT_F90(h, val, vec, n_loc)  src/mod_pevsl.f90:130:
 * eigenvalues ascending, vec column j at j*n_loc (scaled coordinates x~,
 * B~-normalised).  vec may be NULL. */
int nmg_copy_result(nmg_handle h, double *val, double *vec, int64_t n_loc);
/* residual check of src/mod_pevsl.f90:137-162 on the device-resident result:
 * res_ref[j] = sqrt(sum_i r_i^2 / N)/|lam_j| (reference definition) and
 * res2[j] = ||r||_2/|lam_j| with x~ B~-normalised, r = A~x~ - lam B~x~. */
int nmg_residuals(nmg_handle h, double *res_ref, double *res2);

/* Adjoint backward (no reference counterpart; the element formulas are those
 * of src/mod_cg_create_matrix.f90:1018-1219): for every eigenpair j of the
 * last nmg_cheblannr and element e, the derivatives of lambda_j with respect
 * to a uniform change of the nodal density, first Lame parameter and shear
 * modulus of element e, x_j^T (dA_e - lambda_j dB_e) x_j with x_j = D x~_j
 * (physical, B-normalised).  elems: int32[n_elem*np] node ids (DoF 3*v+p),
 * detJ[n_elem], invJ[n_elem*9] (row-major, invJ = 2 B^-1 of the affine map),
 * refM[np*np], refD[3*np*np] reference-element mass and derivative matrices.
 * Outputs d_rho, d_lam, d_mu: f64[nev*n_elem], eigenpair-major.  job must be
 * 1 (no reference gravity); JOB=2 runs go through nmg_backward_grav. */
int nmg_backward(nmg_handle h, int64_t n_elem, int np, const int32_t *elems, const double *detJ,
                 const double *invJ, const double *refM, const double *refD, int job, double *d_rho,
                 double *d_lam, double *d_mu);
/* nmg_backward for JOB=2 (no reference counterpart; the gravity terms are
 * OPrho of src/mod_cg_create_matrix.f90:1131-1180, linear in the nodal
 * density, with the gravity field held fixed): d_rho also carries
 * x_j^T (dA_e/drho_e) x_j.  gnode: f64[n_elem*np*3], the reference gravity
 * at each element node (the g0 the stiffness was assembled with); gsym:
 * f64[n_elem*9], the symmetric part (dg + dg^T)/2 of each element's
 * least-squares gravity gradient, dg[r*3+c] = d g_c / d x_r
 * (src/mod_cg_create_matrix.f90:1054-1080). */
int nmg_backward_grav(nmg_handle h, int64_t n_elem, int np, const int32_t *elems, const double *detJ,
                      const double *invJ, const double *refM, const double *refD, const double *gnode,
                      const double *gsym, double *d_rho, double *d_lam, double *d_mu);

/* Density of states by the kernel polynomial method, as pEVSL's kpmdos
 * (the reference fixes NSLICES = 1, src/mod_pevsl.f90:47, 104-105, and never
 * calls it; pEVSL's slicing drivers do).  intv = {a, b, lmin, lmax}.
 * mu[k], k = 0..mdeg: Hutchinson estimates (nvec Rademacher vectors; nvec in
 * {1, 2, 4, 8} or a multiple of 16) of tr T_k((A~B~^-1 - cc)/dd),
 * cc = (lmax+lmin)/2, dd = (lmax-lmin)/2.  ecnt: Jackson-damped estimate of
 * the number of eigenvalues in [a, b]. */
int nmg_kpmdos(nmg_handle h, int mdeg, int nvec, const double *intv, double *mu, double *ecnt);
/* Jackson-damped eigenvalue count in [intv[0], intv[1]] from the moments
 * (pEVSL's count behind spslicer; not called by the reference). */
int nmg_dos_count(const double *mu, int mdeg, const double *intv, double *count);
/* pEVSL spslicer: boundaries sli[0] = a < ... < sli[n_slices] = b with equal
 * estimated counts (bisection of the closed-form count). */
int nmg_spslicer(double *sli, const double *mu, int mdeg, const double *intv, int n_slices);
/* Batched spectrum slices (no reference counterpart: NSLICES = 1 at
 * src/mod_pevsl.f90:47): ChebLanNr on every [sli[i], sli[i+1]] (i <
 * n_slices <= 16), each with its own FindPol(thr_int, thr_ext) filter on
 * [lmin, lmax], its own recurrence and tests, advanced in lockstep so that
 * every sparse product reads A~ and B~ once for all slices.  The union of
 * the pairs (a pair found by two neighbours on their shared boundary kept
 * once) is then what nmg_get_nev / nmg_copy_result / nmg_residuals /
 * nmg_backward return, slice by slice. */
int nmg_solve_slices(nmg_handle h, int n_slices, const double *sli, double lmin, double lmax, double thr_int,
                     double thr_ext, int maxit, double tol);
/* pairs per slice of the last nmg_solve_slices (int[n_slices]; no
 * reference counterpart). */
int nmg_slice_counts(nmg_handle h, int *counts);

/* Tuning/diagnostics of the last nmg_cheblannr / nmg_lanbounds call:
 * Lanczos steps, filter-operator applications, candidate count, seconds
 * (no reference counterpart). */
int nmg_solve_stats(nmg_handle h, int *lan_steps, int64_t *filter_applies, int *ncand,
                    double *seconds);
/* Outcome of the residual stop of the last nmg_cheblannr / nmg_solve_slices
 * (no reference counterpart; the reference stops on its trace test alone,
 * src/mod_pevsl.f90:122): target_met = 1 when every returned pair met the
 * residual target (or the interval held no pair), 0 when the iteration
 * stopped because the residual no longer improved; resid_max is the largest
 * ||A~x - lam B~x||_2/|lam| of the returned pairs measured at that stop (0
 * when the target is 0 and no residual was measured). */
int nmg_solve_quality(nmg_handle h, int *target_met, double *resid_max);
/* Convergence controls of nmg_cheblannr (defaults: ntest 30, cycle 20,
 * rtol 1e-9 on the filtered Ritz residual).  No reference counterpart: the
 * reference passes only TOL = 1e-5 (src/mod_pevsl.f90:52, 122). */
int nmg_set_lanczos_controls(nmg_handle h, int ntest, int cycle, double rtol);

/* Residual target of nmg_cheblannr (no reference counterpart; default
 * 5e-11; 0 = stop on the reference's trace test alone).  Once the trace test and the filtered Ritz
 * residuals pass, the pairs are extracted and the iteration stops when every
 * ||A~x - lam B~x||_2/|lam| (x B~-normalised) is at or below the target, or
 * when that residual stops improving (its floor is set by the B-solve
 * accuracy). */
int nmg_set_residual_target(nmg_handle h, double target);

/* With nmg_set_split_streams(h, p > 1) (no reference counterpart): on = 1
 * (default) runs each block filter application as p independent chains of
 * k/p columns, one per stream (B-solve, stiffness product and recurrence
 * per degree), joined once per application; on = 0 splits only the
 * B-solves and joins after each. */
int nmg_set_split_filter(nmg_handle h, int on);
/* Replay every B-solve (all Chebyshev steps, and the split parts of a
 * k-column solve with their fork and join) from a captured HIP graph, one
 * executable graph per distinct set of buffers, so that a solve costs one
 * host launch instead of one per step (no reference counterpart; on = 0 or
 * 1, default 0).  A change of the B matrix's kernel layout (autotuning,
 * tiling, value order, …) retires the graphs captured before it. */
int nmg_set_graphs(nmg_handle h, int on);
/* Progress of long solves (no reference counterpart): a line on stderr
 * every `every` Lanczos steps of nmg_lanbounds, nmg_cheblannr and
 * nmg_solve_slices (0, the default: silent). */
int nmg_set_verbose(nmg_handle h, int every);
/* Lanczos block width (no reference counterpart; pEVSL's ChebLanNr is the
 * block width 1, the default).  With block > 1, nmg_cheblannr runs the block
 * filtered Lanczos: the same filter, B-solve and extraction, applied to
 * `block` vectors at once so that each sparse product streams the matrices
 * once per block (1, 2, 4, 8 or 16). */
int nmg_set_block_size(nmg_handle h, int block);
/* Split each k-column B-solve into `parts` (2 or 4; 0 or 1: off) solves of
 * k/parts >= 2 columns, run at once on as many streams of the handle (no
 * reference counterpart). */
int nmg_set_split_streams(nmg_handle h, int parts);

/* MVFunc-shaped callbacks, void (*)(double *x, double *y, void *data) as
 * pEVSL's PEVSL_SETAMV_F90 / PEVSL_SETBMV_F90 take them
 * (src/mod_pevsl.f90:69, 77-80), standing for the reference's sparseAV,
 * sparseBV and solveBV (src/mod_matvec.f90:445-483): y = A~x (the DAD or
 * fluid-solid operator set on the solver), y = B~x, y ~ B~^-1 x (the
 * solver's Chebyshev B-solve).  x and y are host vectors of the solver's n;
 * data points at the solver's nmg_handle.  An unmodified pEVSL caller can
 * hand these to its own Lanczos driver.  They return nothing, as MVFunc
 * does: on failure y is filled with NaN and nmg_last_error() holds the
 * message. */
void nmg_amv(double *x, double *y, void *data);
void nmg_bmv(double *x, double *y, void *data);
void nmg_bsol(double *x, double *y, void *data);

/* Kernel timing probe (no reference counterpart): with every > 0, in every
 * `every`-th B-solve the middle fused Chebyshev step launch (which = 0; a
 * persistent or graph-replayed solve is bracketed whole) and every
 * `every`-th fused filter step launch (which = 1) are bracketed by HIP
 * events on the stream they run on.  nmg_kernel_probe returns the mean
 * duration of one launch (bracket time / launches in it), the sample count
 * and the algorithmic bytes of one launch (matrix + vector streams). */
int nmg_set_kernel_probe(nmg_handle h, int every);
int nmg_kernel_probe(nmg_handle h, int which, double *mean_seconds, int64_t *samples,
                     double *bytes_per_launch);

/* --- device-vector access (for tests and benchmarks; none of the entries
 * below has a reference counterpart) ------------------------------------ */
/* One application of the filter operator of the handle:
 * y = p((A B^-1 - cc)/dd) z (z-space of ChebLanNr) with pol; host vectors. */
int nmg_filter_apply(nmg_handle h, nmg_handle pol, const double *z, double *y);

/* Mean seconds of one application of the filter to K device-resident
 * columns (K = 1, 2, 4, .., 64), with the solver's current precision, stream
 * split and operator; timed over `reps` applications after one warm-up.
 * No counterpart in the reference: a measurement entry (per-degree cost of
 * the filter at a given block width). */
int nmg_bench_filter(nmg_handle h, nmg_handle pol, int K, int reps, double *seconds);
/* Operators of the handle on host vectors: which = 0 A, 1 B, 2 B-solve. */
int nmg_apply(nmg_handle h, int which, const double *x, double *y);
/* The same on K host columns at once (X, Y: K blocks of n, column c at c*n)
 * through the k-column kernels of the block filter; K in {1, 2, 4, 8, 16}. */
int nmg_apply_block(nmg_handle h, int which, int K, const double *X, double *Y);
/* Kernel timing probe: runs `reps` products of mat on device-resident
 * vectors and returns the mean kernel time in seconds (HIP events on the
 * library's own stream) and the algorithmic bytes of one product. */
int nmg_bench_spmv(nmg_handle mat, int reps, int fp32, double *seconds, double *bytes);
/* Same for one fused Chebyshev B-solve step (one product + vector update). */
int nmg_bench_chebstep(nmg_handle cheb, int reps, int fp32, double *seconds, double *bytes);
/* The tall-skinny product of the block re-orthogonalisation on host
 * arrays (for checks): C (m x kc) = A^T B, A: n x m, B: n x kc, all
 * column-major. */
int nmg_gemm_tn(int64_t n, int m, int kc, const double *A, const double *B, double *C);
/* Timing of that product on device arrays (no reference counterpart): mean
 * seconds per C = A^T B with A n x m and B n x kc filled on the device;
 * bytes: its algorithmic HBM bytes 8 (n m + n kc + m kc). */
int nmg_bench_gemm_tn(int64_t n, int m, int kc, int reps, double *seconds, double *bytes);
/* Dense products of the block recurrence and the Rayleigh-Ritz extraction
 * (no reference counterpart; pEVSL's ChebLanNr does them with BLAS inside
 * pEVSL_CHEBLANNR_F90, src/mod_pevsl.f90:122), on host arrays (column-major,
 * leading dimensions n, m and n): op 0 C (n x K) = alpha A (n x m) B (m x K)
 * + beta C; op 1 C (m x K) = A^T B with A n x m, B n x K.  kernel 1 runs the
 * hand-written FP64 MFMA kernels (v_mfma_f64_16x16x4f64) the solver uses,
 * kernel 0 rocBLAS DGEMM (op 0) or the FMA split reduction (op 1).
 * nmg_bench_dense times one product on device-resident data and returns its
 * algorithmic bytes and flops. */
int nmg_dense_product(int op, int kernel, int64_t n, int m, int K, double alpha, const double *A,
                      const double *B, double beta, double *C);
int nmg_bench_dense(int op, int kernel, int64_t n, int m, int K, int reps, double *seconds, double *bytes,
                    double *flops);
/* Which kernels a solver's dense products run on (no reference
 * counterpart): 2 (default) the MFMA kernels, except A^T B with fewer than 16
 * columns (the block recurrence's V^T W at block widths 1-8), which runs on
 * the FMA split reduction; 1 the MFMA kernels throughout; 0 rocBLAS DGEMM
 * and the FMA split reduction.  nmg_dense_product / nmg_bench_dense take
 * the same kernel codes. */
int nmg_set_dense_kernels(nmg_handle h, int mfma);
/* A plain copy kernel of n doubles, `reps` times (calibration of HBM byte
 * counters: it moves 16 n bytes per launch); mean seconds per launch. */
int nmg_bench_copy(int64_t n, int reps, double *seconds);
/* Timing of one whole Chebyshev solve (deg steps) on device vectors, in the
 * cheb's current mode; bytes: deg x the algorithmic bytes of one step. */
int nmg_bench_bsolve(nmg_handle cheb, int reps, int fp32, double *seconds, double *bytes);
/* One K-column fused Chebyshev step, `reps` times on device vectors: fp32 =
 * 0 FP64, 1 FP32 values, 2 FP32 values and vectors; bytes: its algorithmic
 * bytes (matrix once, b, x_k, x_(k-1) read and x_(k+1) written). */
int nmg_bench_block_chebstep(nmg_handle cheb, int K, int reps, int fp32, double *seconds, double *bytes);
/* Workgroup order of the sparse-product kernels of mat (no reference
 * counterpart): 1 (default) gives each of the 8 XCDs one contiguous range
 * of rows, so rows that gather the same x share an L2; 0 keeps the plain
 * round-robin order. */
int nmg_set_xcd_order(nmg_handle mat, int on);
/* Matrix values and indices of mat read with non-temporal (streaming)
 * loads (1) or ordinary loads (0, default): a product streams them once, so
 * streaming loads leave the caches to the gathered x. */
int nmg_set_stream_loads(nmg_handle mat, int on);
/* Value order of a 3x3-block matrix (no reference counterpart): 0 planes
 * (the nine values of each block in nine arrays), 1 tiles of 64 blocks
 * (nine runs of 64 values per tile), so that one run of blocks reads
 * contiguous memory.  Products, scaling and diagonals follow the order. */
int nmg_set_value_order(nmg_handle mat, int tiled);
/* Kernel mapping of a matrix (tuning): lanes per row of the single-column
 * products (0 = by row length, else 8/16/32/64) and, for K > 0, rows per
 * wavefront of its K-column products (0 = default, 1, 2, 4 or 8). */
int nmg_set_kernel_map(nmg_handle mat, int group, int K, int rows_per_wave);
/* K-column scalar-mass products through LDS tiles of at most max_nodes
 * nodes (as nmg_set_tiling, 1024-thread workgroups) formed in the wide
 * manner (no reference counterpart): one entry per lane, `lanes` lanes per
 * node row (4, 8 or 16), the K columns in registers, x read from the tile's
 * LDS copy.  max_nodes <= 0 drops the tiling.  The autotuner (kind 5) sets
 * it when fastest. */
int nmg_set_tiled_wide(nmg_handle mat, int K, int max_nodes, int lanes);
/* K-column products of a scalar-mass matrix (MASS3) through the wide layout
 * (no reference counterpart): one matrix entry per lane with the K columns
 * in registers and each neighbour's 3K values read by 16-byte loads,
 * `lanes` lanes per node row (4, 8, 16 or 32; 0 drops it; K = 2, 4, 8, 16;
 * FP32 vectors use it for K >= 4).  Takes precedence over the tiled and
 * row-per-wavefront layouts; the autotuner (kind 4) sets it when fastest. */
int nmg_set_wide(nmg_handle mat, int K, int lanes);
/* Single-column products of a blocked matrix (BSR3 / MASS3) through row
 * groups (no reference counterpart): consecutive node rows holding at most
 * 64 entries (or one longer row) per group, one entry per lane, per_wave
 * groups per wavefront with the next group's loads issued while the
 * current one is reduced (per_wave = 1, 2, 4 or 8; 0 drops the groups).
 * Takes precedence over the other single-column layouts; the autotuner
 * (kind 3) sets it when it is the fastest. */
int nmg_set_row_groups(nmg_handle mat, int per_wave);
/* Single-column products of a 3x3-block matrix through pipelined lane
 * groups (no reference counterpart): `lanes` lanes (16, 32 or 64) walk
 * `rows` consecutive node rows (2, 4 or 8) one entry per lane, loading the
 * columns and values of row j+1 before row j's x gathers, so each row waits
 * on one memory latency instead of three in sequence.  rows = 0 drops it;
 * nmg_set_kernel_map clears it.  The autotuner (kind 6) sets it when
 * fastest. */
int nmg_set_row_pipeline(nmg_handle mat, int lanes, int rows);
/* LDS-tiled K-column products: rows cut into tiles whose referenced nodes
 * (at most max_nodes, <= 64 KB of x in LDS) are staged once per tile by a
 * workgroup of `threads` (256 or 1024; 0 = 1024); max_nodes = 0 removes the
 * tiling. */
int nmg_set_tiling(nmg_handle mat, int K, int max_nodes, int rows_per_wave, int threads);
/* The tiling of a node pattern computed on the host only (no device use):
 * call once with null arrays for ntiles and the list length (tcp[ntiles]),
 * then with trp[ntiles+1], tcp[ntiles+1], tcols[tcp[ntiles]], lcol[nnz]. */
int nmg_tiling_host(int64_t nb, const int64_t *rowptr, const int32_t *col, int max_nodes, int64_t *ntiles,
                    int *max_used, int64_t *trp, int64_t *tcp, int32_t *tcols, uint16_t *lcol);
/* Autotuner (no reference counterpart): times the product layouts of mat for
 * K columns on the device (lanes per row; rows per wavefront; LDS tiles with
 * capacities on a geometric ladder up to 64 KB, workgroups of 1024 or 256
 * threads; for K = 1 also row groups) and keeps the fastest.  kind: 0
 * lanes-per-row, 1 rows per wavefront, 2 tiled, 3 row groups, 4 wide,
 * 5 tiled wide (4 and 5: scalar mass, K >= 2), 6 pipelined lanes-per-row
 * (3x3 blocks, K = 1); rows: lanes per row (kinds 0, 4, 5), rows per
 * wavefront (1, 2), groups per wavefront (3) or rows per lane group (6);
 * cap: tile node capacity (kinds 2, 5) or lanes per group (6); threads:
 * tile workgroup size (kind 2). */
int nmg_csr_autotune(nmg_handle mat, int K, int reps, int fp32, double *seconds, int *kind, int *rows,
                     int *cap, int *threads);

#ifdef __cplusplus
}
#endif
#endif /* NMGPU_H */
