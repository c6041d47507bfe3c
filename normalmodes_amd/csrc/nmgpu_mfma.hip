// FP64 matrix-core (MFMA) kernels for the dense contractions of the block
// filtered Lanczos: the re-orthogonalisation W -= Z H and W -= Z_j A_j, the
// Ritz vectors Y = V S and X = Y W, and the Gram / projection products
// V^T W, Y^T (A~Y), Y^T (B~Y).  All operands are column-major FP64.
//
// The building block is v_mfma_f64_16x16x4f64: one wavefront multiplies a
// 16x4 tile of A by a 4x16 tile of B into a 16x16 accumulator.  Operand and
// result layout per lane l (r = l mod 16, q = l / 16):
//   A operand: A[r][q]     (row r of the 16, depth q of the 4)
//   B operand: B[q][r]     (depth q, column r)
//   result:    D[q + 4 i][r] in element i of the lane's four accumulators
//              (measured on gfx950: the four lane groups interleave by row).
//
// gemm_nn: C (n x K) = alpha A (n x m) B (m x K) + beta C.  One wavefront
// owns RT row tiles of 16 rows and KT column tiles of 16 columns; it streams
// its rows of A once (each MFMA step reads four 128-byte column segments per
// row tile) and reads the small B through the caches.
//
// gemm_tn: C (m x K) = A^T (m x n) B (n x K) with n long.  A workgroup
// reduces a chunk of R rows: it stages the chunk's R x K block of B in LDS
// once (coalesced), then its four wavefronts take 16-column tiles of A in
// turn, stage each R x 16 tile in LDS (coalesced) and run R/4 MFMA steps
// reading both operands from LDS.  Every element of A and B is read from
// HBM once; the chunk partials are summed by a second kernel.
#include <hip/hip_runtime.h>

#include "nmgpu_internal.h"

namespace nmg {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// C = alpha A B + beta C
template <int RT, int KT>
__global__ __launch_bounds__(256) void gemm_nn_mfma_k(int64_t n, int K, int m, double alpha,
                                                      const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb, double beta,
                                                      double* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int64_t row0 = ((int64_t)blockIdx.x * 4 + wave) * (16 * RT);
  const int col0 = (int)blockIdx.y * (16 * KT);
  if (row0 >= n) return;  // uniform per wavefront
  d4 acc[RT][KT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int u = 0; u < KT; ++u) acc[t][u] = d4{0.0, 0.0, 0.0, 0.0};
  bool rowok[RT];
  const double* arow[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int64_t row = row0 + 16 * t + r;
    rowok[t] = row < n;
    arow[t] = A + (rowok[t] ? row : 0);
  }
  bool colok[KT];
  const double* bcol[KT];
#pragma unroll
  for (int u = 0; u < KT; ++u) {
    const int col = col0 + 16 * u + r;
    colok[u] = col < K;
    bcol[u] = B + (int64_t)(colok[u] ? col : 0) * ldb;
  }
  const int mfull = m & ~3;
  for (int j = 0; j < mfull; j += 4) {
    const int64_t jj = j + q;
    double a[RT], b[KT];
#pragma unroll
    for (int t = 0; t < RT; ++t) a[t] = rowok[t] ? arow[t][jj * lda] : 0.0;
#pragma unroll
    for (int u = 0; u < KT; ++u) b[u] = colok[u] ? bcol[u][jj] : 0.0;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int u = 0; u < KT; ++u) acc[t][u] = mfma16(a[t], b[u], acc[t][u]);
  }
  if (mfull < m) {  // the last 1-3 depth slices
    const int64_t jj = mfull + q;
    const bool in = jj < m;
    double a[RT], b[KT];
#pragma unroll
    for (int t = 0; t < RT; ++t) a[t] = (in && rowok[t]) ? arow[t][jj * lda] : 0.0;
#pragma unroll
    for (int u = 0; u < KT; ++u) b[u] = (in && colok[u]) ? bcol[u][jj] : 0.0;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int u = 0; u < KT; ++u) acc[t][u] = mfma16(a[t], b[u], acc[t][u]);
  }
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int u = 0; u < KT; ++u) {
      const int col = col0 + 16 * u + r;
      if (col >= K) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = row0 + 16 * t + q + 4 * i;
        if (row >= n) continue;
        double* c = C + row + (int64_t)col * ldc;
        const double v = alpha * acc[t][u][i];
        *c = (beta == 0.0) ? v : v + beta * *c;
      }
    }
}

void k_gemm_nn_mfma(int64_t n, int K, int m, double alpha, const double* A, int64_t lda, const double* B,
                    int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  if (n <= 0 || K <= 0) return;
  NMG_CHECK(lda >= n && ldb >= m && ldc >= n, NMG_EINVAL, "gemm_nn: leading dimensions too small");
  if (m <= 0) {  // C = beta C
    NMG_CHECK(beta == 0.0 || beta == 1.0, NMG_EINVAL, "gemm_nn: empty product needs beta 0 or 1");
    if (beta == 0.0)
      for (int c = 0; c < K; ++c) k_fill(n, 0.0, C + (int64_t)c * ldc, s);
    return;
  }
  // one row tile per wavefront when the rows alone would leave the CUs idle
  const bool small = n < (int64_t)256 * 4 * 64;
  if (K <= 16) {
    if (small) {
      const dim3 grid((unsigned)((n + 63) / 64), 1);
      hipLaunchKernelGGL((gemm_nn_mfma_k<1, 1>), grid, dim3(256), 0, s, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc);
    } else {
      const dim3 grid((unsigned)((n + 255) / 256), 1);
      hipLaunchKernelGGL((gemm_nn_mfma_k<4, 1>), grid, dim3(256), 0, s, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc);
    }
  } else if (K <= 32) {
    const dim3 grid((unsigned)((n + (small ? 63 : 127)) / (small ? 64 : 128)), 1);
    if (small)
      hipLaunchKernelGGL((gemm_nn_mfma_k<1, 2>), grid, dim3(256), 0, s, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      hipLaunchKernelGGL((gemm_nn_mfma_k<2, 2>), grid, dim3(256), 0, s, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc);
  } else {
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((K + 63) / 64));
    hipLaunchKernelGGL((gemm_nn_mfma_k<1, 4>), grid, dim3(256), 0, s, n, K, m, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  NMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// C = A^T B, split over chunks of kTnRows rows
constexpr int kTnRows = 64;

template <int KT>
__global__ __launch_bounds__(256) void gemm_tn_mfma_part_k(int64_t n, int m, int K, const double* __restrict__ A,
                                                           int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                           double* __restrict__ part, int nch) {
  // +1 column of padding against LDS bank conflicts on the transposed reads
  __shared__ double sB[kTnRows][16 * KT + 1];
  __shared__ double sA[4][kTnRows][17];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int64_t per = ((n + nch - 1) / nch + kTnRows - 1) / kTnRows * kTnRows;
  const int64_t start = (int64_t)blockIdx.x * per;
  const int64_t end = start + per < n ? start + per : n;
  const int mt_count = (m + 15) / 16;
  // accumulators of the wavefront's column tiles of A: at most 4 per wave
  // per pass over the chunk list (grid.y spreads larger m)
  const int mt_base = (int)blockIdx.y * 16;  // 16 tiles = 4 per wave per block row
  d4 acc[4][KT];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int u = 0; u < KT; ++u) acc[v][u] = d4{0.0, 0.0, 0.0, 0.0};
  for (int64_t r0 = start; r0 < end; r0 += kTnRows) {
    __syncthreads();
    // stage B[r0 : r0+R, 0 : 16 KT] (zero outside)
    for (int idx = threadIdx.x; idx < kTnRows * 16 * KT; idx += 256) {
      const int c = idx / kTnRows, rr = idx % kTnRows;
      const int64_t row = r0 + rr;
      sB[rr][c] = (c < K && row < end) ? B[row + (int64_t)c * ldb] : 0.0;
    }
    __syncthreads();
    // every wavefront runs the same iterations (work predicated), so the
    // barriers are reached by the whole workgroup
    for (int v = 0; v < 4; ++v) {
      const int mt = mt_base + 4 * v + wave;
      const bool active = mt < mt_count;
      if (active) {
        // stage A[r0 : r0+R, 16 mt : 16 mt + 16] for this wavefront
        for (int idx = lane; idx < kTnRows * 16; idx += 64) {
          const int c = idx / kTnRows, rr = idx % kTnRows;
          const int64_t row = r0 + rr;
          const int col = 16 * mt + c;
          sA[wave][rr][c] = (col < m && row < end) ? A[row + (int64_t)col * lda] : 0.0;
        }
      }
      __syncthreads();
      if (active) {
#pragma unroll 4
        for (int k = 0; k < kTnRows; k += 4) {
          const double a = sA[wave][k + q][r];
#pragma unroll
          for (int u = 0; u < KT; ++u) acc[v][u] = mfma16(a, sB[k + q][16 * u + r], acc[v][u]);
        }
      }
      __syncthreads();
    }
  }
  // partial of this chunk: part[((col) * m + row) * nch + chunk]
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int mt = mt_base + 4 * v + wave;
    if (mt >= mt_count) continue;
#pragma unroll
    for (int u = 0; u < KT; ++u) {
      const int col = 16 * u + r;
      if (col >= K) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * mt + q + 4 * i;
        if (row < m) part[((int64_t)col * m + row) * nch + blockIdx.x] = acc[v][u][i];
      }
    }
  }
}

__global__ void gemm_tn_mfma_final_k(int m, int K, const double* __restrict__ part, int nch, double* __restrict__ C,
                                     int64_t ldc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * K) return;
  const int k = t / m, j = t % m;
  const double* p = part + (int64_t)t * nch;
  double v = 0.0;
  for (int c = 0; c < nch; ++c) v += p[c];
  C[(int64_t)k * ldc + j] = v;
}

static int tn_chunks(int64_t n, int m) {
  // about two workgroups per CU over all block rows of 16 column tiles
  const int rows_of_tiles = ((m + 15) / 16 + 15) / 16;
  int64_t ch = 512 / (rows_of_tiles > 0 ? rows_of_tiles : 1);
  const int64_t maxch = (n + kTnRows - 1) / kTnRows;
  if (ch > maxch) ch = maxch;
  if (ch < 1) ch = 1;
  return (int)ch;
}

size_t gemm_tn_mfma_work_size(int64_t n, int m, int K) { return (size_t)tn_chunks(n, m) * m * K + 16; }

void k_gemm_tn_mfma(int64_t n, int m, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                    int64_t ldc, double* work, hipStream_t s) {
  if (m <= 0 || K <= 0) return;
  NMG_CHECK(K <= 64, NMG_EINVAL, "gemm_tn: K must be at most 64");
  NMG_CHECK(lda >= n && ldb >= n && ldc >= m, NMG_EINVAL, "gemm_tn: leading dimensions too small");
  const int nch = tn_chunks(n, m);
  const dim3 grid((unsigned)nch, (unsigned)(((m + 15) / 16 + 15) / 16));
  if (K <= 16)
    hipLaunchKernelGGL((gemm_tn_mfma_part_k<1>), grid, dim3(256), 0, s, n, m, K, A, lda, B, ldb, work, nch);
  else if (K <= 32)
    hipLaunchKernelGGL((gemm_tn_mfma_part_k<2>), grid, dim3(256), 0, s, n, m, K, A, lda, B, ldb, work, nch);
  else
    hipLaunchKernelGGL((gemm_tn_mfma_part_k<4>), grid, dim3(256), 0, s, n, m, K, A, lda, B, ldb, work, nch);
  hipLaunchKernelGGL(gemm_tn_mfma_final_k, dim3((m * K + 255) / 256), dim3(256), 0, s, m, K, work, nch, C, ldc);
  NMG_HIP(hipGetLastError());
}

}  // namespace nmg
