// Internal declarations shared by the nmgpu HIP library sources.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "nmgpu.h"

namespace nmg {

// ---------------------------------------------------------------------------
// errors
struct Error {
  int code;
  std::string msg;
};
void set_error(int code, const std::string& msg);
#define NMG_STR2(x) #x
#define NMG_STR(x) NMG_STR2(x)
#define NMG_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::nmg::Error{NMG_EHIP, std::string(#call) + " at " __FILE__ ":" NMG_STR(__LINE__) \
                                       ": " + hipGetErrorString(e_)};                     \
  } while (0)
#define NMG_CHECK(cond, code, msg)                    \
  do {                                                \
    if (!(cond)) throw ::nmg::Error{(code), (msg)};   \
  } while (0)

// ---------------------------------------------------------------------------
// matrices in HBM
enum Layout { L_CSR = 0, L_BSR3 = 1, L_MASS3 = 2 };

// Row tiles of a BSR3 / MASS3 matrix for the LDS-staged k-column products:
// tile t owns node rows [trp[t], trp[t+1]) and references the nodes
// tcols[tcp[t] .. tcp[t+1]); lcol[k] is entry k's index within its tile's list.
struct Tiling {
  int64_t ntiles = 0;
  int max_nodes = 0;      // largest node list (sets the LDS size)
  int cap = 0;            // node capacity the tiles were cut for
  int rows_per_wave = 1;  // R of the lane layout
  int block = 1024;       // threads of the workgroup sharing one tile (1024 or 256)
  int wide = 0;           // scalar mass, K >= 2: G lanes per row, one entry per lane (0: rows layout)
  int64_t* trp = nullptr;
  int64_t* tcp = nullptr;
  int32_t* tcols = nullptr;
  uint16_t* lcol = nullptr;
  double staged_nodes = 0;  // sum of list lengths (x blocks staged per product)
  ~Tiling();
};

// Row groups of the single-column products (BSR3 / MASS3): group g owns node
// rows [grow[g], grow[g+1]) and entries [gk[g], gk[g+1]); a group holds at
// most 64 entries or is one longer row.  per_wave groups per wavefront.
struct RowGroups {
  int64_t ngroups = 0;
  int per_wave = 8;
  int64_t* grow = nullptr;
  int64_t* gk = nullptr;
  ~RowGroups();
};

// Process-wide source of layout generations (unique across matrices, so a
// matrix freed and re-created at the same address never repeats one).
inline uint64_t next_layout_gen() {
  static std::atomic<uint64_t> g{0};
  return ++g;
}

struct DevMat {
  int64_t nrow = 0, ncol = 0, nnz = 0;  // scalar (DoF-level) sizes
  // changes with every change of the device buffers or of the kernel
  // mapping below; captured HIP graphs of products on this matrix are keyed
  // on it, so a retuned or re-tiled matrix is never replayed with stale
  // kernels or freed buffers
  uint64_t gen = next_layout_gen();
  void touch() { gen = next_layout_gen(); }
  int layout = L_CSR;
  // CSR (any matrix): rowptr[nrow+1], col[nnz], val[nnz]
  int64_t* rowptr = nullptr;
  int32_t* col = nullptr;
  double* val = nullptr;
  float* val32 = nullptr;
  // 3x3-block rows (BSR3) and scalar-mass-times-I3 (MASS3) share the node
  // pattern: nb node rows, nbnz node entries, browptr[nb+1], bcol[nbnz]
  int64_t nb = 0, nbnz = 0;
  int64_t* browptr = nullptr;
  int32_t* bcol = nullptr;
  // BSR3: 9 value planes bval[p*nbnz + k], p = 3*r + c;  MASS3: 1 plane
  double* bval = nullptr;
  float* bval32 = nullptr;
  // kernel mapping knobs (0 = default): lanes per row of the single-column
  // products; rows per wavefront of the k-column products for K = 1,2,4,8,16
  int group = 0;
  int block_map[7] = {0, 0, 0, 0, 0, 0, 0};  // per K = 1,2,4,8,16,32,64
  Tiling* tiling[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // per K = 1..64
  // lanes per node row of the wide k-column scalar-mass kernel (one entry
  // per lane, K columns in registers), per K = 1,2,4,8,16; 0 = not used
  int wide[7] = {0, 0, 0, 0, 0, 0, 0};
  bool single_blocked = false;  // single-column products through the K = 1 block kernels
  RowGroups* rgroups = nullptr; // single-column products through the row-group kernels (first choice)
  int pipe = 0;                 // BSR3 single-column rows per pipelined lane group (2, 4, 8; 0 = off)
  int xcd = 1;                  // XCD-contiguous workgroup order of the row kernels (0: plain order)
  int nt = 0;                   // non-temporal matrix loads in the row kernels
  // BSR3 value order: 0 planes (value q of entry k at q*nbnz + k); 1 tiles of
  // 64 entries (at (k/64)*576 + q*64 + k%64), so the nine values of a run of
  // entries share DRAM pages while each plane load stays contiguous
  int tiled_vals = 0;
  int64_t value_count() const {
    if (layout == L_CSR) return nnz;
    if (layout == L_MASS3) return nbnz;
    return tiled_vals ? 9 * ((nbnz + 63) / 64 * 64) : 9 * nbnz;
  }
  std::vector<int64_t> h_browptr;  // host copy of the node pattern (tiling)
  std::vector<int32_t> h_bcol;
  ~DevMat();
  void release();
  double matrix_bytes(bool fp32) const;       // values + indices + row offsets
  double algorithmic_bytes(bool fp32) const;  // one product: matrix + x + y
};

// Plan of the persistent Chebyshev solve (scalar mass x I3, one column):
// cooperative grid, nnz-balanced row ranges per workgroup, LDS sizes.
constexpr size_t kPersistLdsMax = 64 * 1024;
struct PersistPlan {
  bool ok = false, fp32 = false;
  int grid = 0, rows_cap = 0, ent_cap = 0;
  size_t lds_bytes = 0;
  int64_t* part = nullptr;
  void release() {
    if (part) (void)hipFree(part);
    part = nullptr;
    ok = false;
  }
  ~PersistPlan() { release(); }
};
// plans the persistent solve for M (ok stays false when M is not MASS3, its
// rows do not fit the LDS budget, or cooperative launch is unavailable)
void plan_cheb_persistent(const DevMat& M, bool fp32, PersistPlan& P);
// x ~ M^-1 b by `deg` Chebyshev steps on [lmin, lmax] in one launch, for K
// interleaved columns (K = 1, 2, 4, 8, 16)
void k_cheb_persistent(int K, const DevMat& M, const PersistPlan& P, const double* b, double* out, double* tmp,
                       int deg, double lmin, double lmax, hipStream_t s);

// kernel flags of M's sparse products (bit 0 XCD order, bit 1 streaming
// loads, bit 2 tiled BSR3 value order)
inline int kflags(const DevMat& M) { return (M.xcd ? 1 : 0) | (M.nt ? 2 : 0) | (M.tiled_vals ? 4 : 0); }
// converts the BSR3 values of M (and their FP32 copy) to the tiled order (1)
// or back to planes (0)
void set_value_order(DevMat& M, int tiled, hipStream_t s);

// builds (or, with per_wave <= 0, drops) the row groups of M (blocked layouts)
void build_row_groups(DevMat& M, int per_wave);
// LDS stride (doubles) of one node's x block in the tiled k-column kernels
// (a stride padded by one double, spreading the blocks over LDS banks, was
// measured no faster on MI355X)
constexpr int tile_stride(int K) { return 3 * K; }
// largest tile node capacity within 64 KB of LDS
inline int tile_cap_limit(int K) { return 64 * 1024 / (8 * tile_stride(K)); }
// builds (or, with max_nodes <= 0, drops) the row tiling of M for K columns
void build_tiling(DevMat& M, int K, int max_nodes, int rows_per_wave, int block = 1024, int wide = 0);
// creates a DevMat from host CSR (0-based); picks the layout
DevMat* make_devmat(int64_t nrow, int64_t ncol, const int64_t* rowptr, const int32_t* col,
                    const double* val, bool allow_blocked = true);
// deep copy on the device
DevMat* clone_devmat(const DevMat& M, hipStream_t s);
// M_ij <- M_ij * dl[i] * dr[j] (device vectors, DoF level)
void scale_devmat(DevMat& M, const double* dl, const double* dr, double sign, hipStream_t s);
// diag(M) (DoF level) into d (device)
void diag_devmat(const DevMat& M, double* d, hipStream_t s);
// refresh float copies of the values
void make_fp32(DevMat& M, hipStream_t s);

// ---------------------------------------------------------------------------
// kernel launchers (nmgpu_kernels.hip); all asynchronous on stream s
void k_spmv(const DevMat& M, const double* x, double* y, hipStream_t s, bool fp32 = false);
// xm <- xm + w (xk - xm + (b - M xk) * ith)
void k_cheb_step(const DevMat& M, const double* b, const double* xk, double* xm, double w,
                 double ith, hipStream_t s, bool fp32 = false);
// vkm <- t (A x - cc vk) - (first ? 0 : vkm);  y += mu * vkm
void k_filter_step(const DevMat& A, const double* x, const double* vk, double* vkm, double* y,
                   double t, double cc, double mu, bool first, hipStream_t s, bool fp32 = false);
void k_fill(int64_t n, double v, double* y, hipStream_t s);
void k_copy(int64_t n, const double* x, double* y, hipStream_t s);
void k_copy_kernel(int64_t n, const double* x, double* y, hipStream_t s);  // a copy kernel (not a DMA)
void k_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s);  // y = a x
void k_mul(int64_t n, const double* d, double* x, hipStream_t s);
// interleaved K-column row scalings (K = 1: plain vectors)
void k_rowmul(int64_t n, int K, const double* d, const double* x, double* y, hipStream_t s);        // y = d o x
void k_add_rowmul(int64_t n, int K, const double* d, const double* a, const double* b, double* y,
                  hipStream_t s);                                                                    // y = d o (a + b)
// unfused Chebyshev filter update from a product w = A in: v = t (w - cc vk) - vkm,
// vkm <- v, y += mu v (mu: K per-column coefficients, or null for the scalar mu0)
void k_filter_update(int64_t n, int K, const double* w, const double* vk, double* vkm, double* y,
                     const double* mu, double mu0, double t, double cc, bool first, hipStream_t s);                  // x *= d
// y -= (*a) x  (a device scalar), optionally y -= (*b) z
void k_axpy_dev(int64_t n, const double* a, const double* x, double* y, hipStream_t s);
// x *= 1 / *beta ; z *= 1 / *beta (z may be null)
void k_invscale_dev(int64_t n, const double* beta, double* x, double* z, hipStream_t s);
// deterministic dot into *out (device); sqrt_out: *out = sqrt(dot)
void k_dot(int64_t n, const double* x, const double* y, double* out, bool sqrt_out, double* work,
           hipStream_t s);
// h[0:k] = V(:,0:k)^T z ; V column-major, ld
void k_gemv_t(int64_t n, int k, const double* V, int64_t ld, const double* z, double* h,
              double* work, hipStream_t s);
// z -= Z(:,0:k) h
void k_gemv_n_sub(int64_t n, int k, const double* Z, int64_t ld, const double* h, double* z,
                  hipStream_t s);
// residual pieces for one vector: out[0] = sum r^2, out[1] = x^T bx, r = ax - lam bx
void k_resid(int64_t n, const double* ax, const double* bx, const double* x, double lam,
             double* out, double* work, hipStream_t s);
// k-column products on column-interleaved block vectors (entry (i, c) at i*K + c),
// K in {1, 2, 4, 8, 16}; mu: K per-column filter coefficients of this degree
void k_block_spmv(int K, const DevMat& M, const double* x, double* y, hipStream_t s, bool fp32 = false);
// the same step with FP32 vectors (FP64 product and update, one rounding on store)
void k_block_cheb_step32(int K, const DevMat& M, const float* b, const float* xk, float* xm, double w,
                         double ith, hipStream_t s, bool fp32);
void k_cheb32_start(int64_t n, double ith, const double* b, float* b32, float* x1, float* x0, hipStream_t s);
void k_f32_to_f64(int64_t n, const float* x, double* y, hipStream_t s);
void k_sub_from(int64_t n, const double* b, double* y, hipStream_t s);   // y = b - y
// device producer of the solid-region A and B values on a given CSR pattern
// (nmgpu_assemble.hip); 0, or -(1 + e) for a flat element e
int assemble_elastic_device(int64_t n_nodes, int64_t n_elem, int np, const double* nodes, const int32_t* elems,
                            const int32_t* vslot, const double* lam, const double* mu, const double* rho,
                            const double* g, const double* refM, const double* refD, const int64_t* adj_ptr,
                            const int32_t* adj, const int64_t* a_rowptr, double* a_val, const int64_t* b_rowptr,
                            double* b_val, hipStream_t s);
void assemble_fs_device(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                        const int32_t* vslot, const double* lam, const double* mu, const double* rho, const double* g,
                        const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM,
                        const double* refD, const double* massF, const int32_t* fmask, const int64_t* adj_ptr,
                        const int32_t* adj, const int8_t* vstat, const int64_t* vstt, const int64_t* pstt,
                        const int64_t* ad_ptr, int32_t* ad_col, double* ad_val, const int64_t* b_ptr, int32_t* b_col,
                        double* b_val, const int64_t* e_ptr, int32_t* e_col, double* e_val, const int64_t* et_ptr,
                        int32_t* et_col, double* et_val, const int64_t* ap_ptr, int32_t* ap_col, double* ap_val,
                        int64_t* stats, hipStream_t s);
void backward_fs_device(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                        const int32_t* vslot, const double* lam, const double* rho, const double* g,
                        const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM,
                        const double* refD, const double* massF, const int32_t* fmask, const int8_t* vstat,
                        const int64_t* vstt, const int64_t* pstt, int k, const double* eig, const double* U,
                        const double* P, double* d_rho, double* d_lam, double* d_mu, hipStream_t s);
void k_add_f32(int64_t n, const float* x, double* y, hipStream_t s);     // y += x
void k_block_cheb_step(int K, const DevMat& M, const double* b, const double* xk, double* xm, double w,
                       double ith, hipStream_t s, bool fp32 = false);
void k_block_filter_step(int K, const DevMat& A, const double* x, const double* vk, double* vkm, double* y,
                         const double* mu, double t, double cc, bool first, hipStream_t s, bool fp32 = false);
// out[c] = sum_i x[i*K+c] y[i*K+c] for K in {1,2,4,8,16}; work: block_dot_work_size(K)
size_t block_dot_work_size(int K);
void k_block_dots(int64_t n, int K, const double* x, const double* y, double* out, double* work, hipStream_t s);
void k_block_colscale(int64_t n, int K, const double* mu, const double* x, double* y, hipStream_t s);
// interleaved n x K <-> P (<= 4) interleaved parts of K/P columns
void k_split_parts(int64_t n, int K, int P, const double* x, double* const* parts, hipStream_t s);
void k_merge_parts(int64_t n, int K, int P, double* const* parts, double* y, hipStream_t s);
void k_to_interleaved(int64_t n, int K, const double* cm, int64_t ld, double* il, hipStream_t s);
void k_from_interleaved(int64_t n, int K, const double* il, double* cm, int64_t ld, hipStream_t s);
// per-element eigenvalue sensitivities; gnode (ne x np x 3) and gsym (ne x 9)
// null without reference gravity (see nmgpu_kernels.hip)
void k_adjoint(int np, int64_t ne, int nev, const int32_t* elems, const double* detJ, const double* invJ,
               const double* refM, const double* refD, const double* X, int64_t ldx, const double* d,
               const double* lam, const double* gnode, const double* gsym, double* g_rho, double* g_lam,
               double* g_mu, hipStream_t s);
size_t dot_work_size();
// tall-skinny C (m x K, ldc) = A^T B, A: n x m (lda), B: n x K (ldb), K in
// {1, 2, 4, 8, 16}; work: gemm_tn_work_size doubles
size_t gemm_tn_work_size(int64_t n, int m, int K);
void k_gemm_tn(int64_t n, int m, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double* work, hipStream_t s);
// FP64 MFMA products (nmgpu_mfma.hip), column-major operands:
//   k_gemm_nn_mfma: C (n x K) = alpha A (n x m) B (m x K) + beta C
//   k_gemm_tn_mfma: C (m x K) = A^T B with A n x m, B n x K (K <= 64), split
//                   over chunks of rows; work: gemm_tn_mfma_work_size doubles
void k_gemm_nn_mfma(int64_t n, int K, int m, double alpha, const double* A, int64_t lda, const double* B,
                    int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);
size_t gemm_tn_mfma_work_size(int64_t n, int m, int K);
void k_gemm_tn_mfma(int64_t n, int m, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                    int64_t ldc, double* work, hipStream_t s);
size_t gemv_work_size(int64_t n, int k);

// ---------------------------------------------------------------------------
// host math (nmgpu_host.cpp)
struct PolyFilter {
  int deg = 0;
  int type = 0;  // 0 interior, 1 left end, 2 right end
  double cc = 0, dd = 1, gam = 0, bar = 0, aa = 0, bb = 0;
  std::vector<double> mu;
};
int find_pol(const double* xintv, double thr_int, double thr_ext, int min_deg, int max_deg,
             int damping, PolyFilter& pol);
// eigenvalues of a symmetric tridiagonal (ascending), implicit QL
int tridiag_eigvals(int m, const double* d, const double* e, std::vector<double>& w);
// eigenvector of the tridiagonal for an (isolated) eigenvalue: last component
// and full vector by inverse iteration
void tridiag_eigvec(int m, const double* d, const double* e, double theta, std::vector<double>& s);
// Lanczos start vector (splitmix64 of seed and index, uniform [-1,1))
void start_vector(int64_t n, uint64_t seed, double* v);
// kernel polynomial (KPM) density of states from trace moments
// tau_k ~ tr T_k((M - cc)/dd), k = 0..mdeg, with Jackson damping:
// estimated number of eigenvalues in [a, b]
double dos_count(int mdeg, const double* tau, double lmin, double lmax, double a, double b);
// slice boundaries a = sli[0] < ... < sli[ns] = b holding equal estimated counts
int dos_slicer(int mdeg, const double* tau, double lmin, double lmax, double a, double b, int ns, double* sli);

}  // namespace nmg
