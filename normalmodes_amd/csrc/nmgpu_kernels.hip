// gfx950 kernels of the filtered-Lanczos path.
//
// Sparse products run G lanes of a 64-lane wavefront per (node) row: the
// lanes stride through the row's entries, read value planes with unit stride
// (coalesced), gather x, and reduce with cross-lane shuffles.  Three storage
// layouts are used (chosen from the matrix structure at upload):
//   CSR    generic scalar rows;
//   BSR3   3x3 blocks per node pair (the stiffness A of the elastic
//          operator, src/mod_cg_create_matrix.f90:1376-1387): one column
//          index per 9 values, 9 value planes;
//   MASS3  scalar mass matrix times I3 (B, src/mod_cg_create_matrix.f90:
//          1417-1435): one scalar entry serves the three components.
// Each product carries an epilogue functor so that the vector updates of the
// Chebyshev B-solve and of the polynomial filter are fused into the same pass.
#include <hip/hip_cooperative_groups.h>

#include "nmgpu_internal.h"

namespace nmg {

constexpr int kBlock = 256;
constexpr int kDotBlocks = 512;
constexpr int kXcd = 8;  // accelerator complex dies (XCDs) of an MI355X

// Kernel flags of the sparse products: bit 0 XCD-contiguous workgroup order,
// bit 1 non-temporal (streaming) loads of the matrix values and indices.
// Workgroup b of a launch runs on XCD b % 8 (round-robin dispatch).  With
// bit 0 set the logical block index gives each XCD one contiguous range,
// so neighbouring node rows -- which gather the same x entries -- share one
// XCD's L2 instead of being spread over eight.  A bijection of [0, grid).
__device__ __forceinline__ int64_t xcd_block(int flags) {
  const int64_t b = blockIdx.x;
  if (!(flags & 1)) return b;
  const int64_t n = gridDim.x, q = n / kXcd, r = n % kXcd;
  const int64_t x = b % kXcd, i = b / kXcd;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}
// position of value q of BSR3 entry k (planes, or tiles of 64 entries)
__device__ __forceinline__ int64_t bidx(int64_t k, int q, int64_t nbnz, int flags) {
  return (flags & 4) ? ((k >> 6) * 576 + (int64_t)q * 64 + (k & 63)) : ((int64_t)q * nbnz + k);
}
// matrix loads: streamed once per product, so with bit 1 they are issued as
// non-temporal loads that do not displace the gathered x from the caches
template <class T>
__device__ __forceinline__ T ldm(const T* p, int flags) {
  return (flags & 2) ? __builtin_nontemporal_load(p) : *p;
}

// --------------------------------------------------------------------------
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = group_sum<64>(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  }
  __syncthreads();
  return r;  // valid in thread 0
}

// --------------------------------------------------------------------------
// epilogues
// Epilogues load their own operands first (load) and combine them with the
// product once it is reduced (store), so that the loads overlap the product.
struct EpiStore {
  double* y;
  struct Pre {};
  __device__ __forceinline__ Pre load(int64_t) const { return Pre{}; }
  __device__ __forceinline__ void store(int64_t i, double v, const Pre&) const { y[i] = v; }
};
struct EpiCheb {
  const double* b;
  const double* xk;
  double* xm;
  double w, ith;
  struct Pre {
    double b, xk, xm;
  };
  __device__ __forceinline__ Pre load(int64_t i) const { return Pre{b[i], xk[i], xm[i]}; }
  __device__ __forceinline__ void store(int64_t i, double v, const Pre& p) const {
    xm[i] = p.xm + w * (p.xk - p.xm + (p.b - v) * ith);
  }
};
struct EpiFilter {
  const double* vk;
  double* vkm;
  double* y;
  double t, cc, mu;
  int first;
  struct Pre {
    double vk, vkm, y;
  };
  __device__ __forceinline__ Pre load(int64_t i) const { return Pre{vk[i], first ? 0.0 : vkm[i], y[i]}; }
  __device__ __forceinline__ void store(int64_t i, double v, const Pre& p) const {
    const double vn = t * (v - cc * p.vk) - p.vkm;
    vkm[i] = vn;
    y[i] = p.y + mu * vn;
  }
};

// --------------------------------------------------------------------------
// row kernels
template <int G, class VT, class Epi>
__global__ __launch_bounds__(kBlock) void csr_rows(int64_t nrow, const int64_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci,
                                                   const VT* __restrict__ va,
                                                   const double* __restrict__ x, Epi epi, int flags) {
  const int64_t t = xcd_block(flags) * kBlock + threadIdx.x;
  const int64_t row = t / G;
  const int lane = threadIdx.x & (G - 1);
  typename Epi::Pre pre{};
  const bool wr = row < nrow && lane == 0;
  if (wr) pre = epi.load(row);
  double s = 0.0;
  if (row < nrow) {
    const int64_t e = rp[row + 1];
    for (int64_t k = rp[row] + lane; k < e; k += G) s += (double)ldm(va + k, flags) * x[ldm(ci + k, flags)];
  }
  s = group_sum<G>(s);
  if (wr) epi.store(row, s, pre);
}

template <int G, class VT, class Epi>
__global__ __launch_bounds__(kBlock) void bsr3_rows(int64_t nb, int64_t nbnz,
                                                    const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci,
                                                    const VT* __restrict__ bv,
                                                    const double* __restrict__ x, Epi epi, int flags) {
  const int64_t t = xcd_block(flags) * kBlock + threadIdx.x;
  const int64_t row = t / G;
  const int lane = threadIdx.x & (G - 1);
  typename Epi::Pre pre{};
  const bool wr = row < nb && lane < 3;
  if (wr) pre = epi.load(3 * row + lane);
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  if (row < nb) {
    const int64_t e = rp[row + 1];
    for (int64_t k = rp[row] + lane; k < e; k += G) {
      const int64_t c3 = 3 * (int64_t)ldm(ci + k, flags);
      double v[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) v[q] = (double)ldm(bv + bidx(k, q, nbnz, flags), flags);
      const double x0 = x[c3], x1 = x[c3 + 1], x2 = x[c3 + 2];
      y0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
      y1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
      y2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
  }
  y0 = group_sum<G>(y0);
  y1 = group_sum<G>(y1);
  y2 = group_sum<G>(y2);
  if (wr) epi.store(3 * row + lane, lane == 0 ? y0 : (lane == 1 ? y1 : y2), pre);
}

// Single-column BSR3 rows walked by pipelined lane groups: G lanes take NR
// consecutive node rows in order, one entry per lane.  The columns and
// values of row j+1 are loaded before row j's x gathers are issued, so a row
// waits on one memory latency (its gathers, with the next row's matrix loads
// in flight beside them) instead of three in sequence (row offsets, then
// columns and values, then x).  The NR+1 row offsets of the group are read
// once, one per lane.  Entries of a row beyond the first G are finished
// unpipelined.
template <int G, int NR, class VT, class Epi>
__global__ __launch_bounds__(kBlock) void bsr3_rows_pipe(int64_t nb, int64_t nbnz,
                                                         const int64_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const VT* __restrict__ bv,
                                                         const double* __restrict__ x, Epi epi, int flags) {
  static_assert(G > NR, "one lane per row offset of the group");
  const int64_t t = xcd_block(flags) * kBlock + threadIdx.x;
  const int64_t row0 = (t / G) * NR;
  const int lane = threadIdx.x & (G - 1);
  if (row0 >= nb) return;  // the whole lane group leaves together
  const int nr = (int)(nb - row0 < NR ? nb - row0 : NR);
  const int64_t off = lane <= nr ? rp[row0 + lane] : 0;  // rp[row0 .. row0 + nr], nb + 1 entries
  int64_t e = __shfl(off, 1, G);
  int64_t k = __shfl(off, 0, G) + lane;
  int32_t c = 0;
  double v[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) v[q] = 0.0;
  if (k < e) {
    c = ldm(ci + k, flags);
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = (double)ldm(bv + bidx(k, q, nbnz, flags), flags);
  }
  for (int j = 0; j < nr; ++j) {
    const int64_t row = row0 + j;
    typename Epi::Pre pre{};
    if (lane < 3) pre = epi.load(3 * row + lane);
    // row j+1 spans [e, e1); past the last row lane nr holds e, so e1 = e
    const int64_t e1 = __shfl(off, j + 2 < nr ? j + 2 : nr, G);
    const int64_t k1 = e + lane;
    int32_t c1 = 0;
    double v1[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) v1[q] = 0.0;
    if (k1 < e1) {
      c1 = ldm(ci + k1, flags);
#pragma unroll
      for (int q = 0; q < 9; ++q) v1[q] = (double)ldm(bv + bidx(k1, q, nbnz, flags), flags);
    }
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (k < e) {
      const int64_t c3 = 3 * (int64_t)c;
      const double x0 = x[c3], x1 = x[c3 + 1], x2 = x[c3 + 2];
      y0 = v[0] * x0 + v[1] * x1 + v[2] * x2;
      y1 = v[3] * x0 + v[4] * x1 + v[5] * x2;
      y2 = v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    for (int64_t kk = k + G; kk < e; kk += G) {
      const int64_t c3 = 3 * (int64_t)ldm(ci + kk, flags);
      double w[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) w[q] = (double)ldm(bv + bidx(kk, q, nbnz, flags), flags);
      const double x0 = x[c3], x1 = x[c3 + 1], x2 = x[c3 + 2];
      y0 += w[0] * x0 + w[1] * x1 + w[2] * x2;
      y1 += w[3] * x0 + w[4] * x1 + w[5] * x2;
      y2 += w[6] * x0 + w[7] * x1 + w[8] * x2;
    }
    y0 = group_sum<G>(y0);
    y1 = group_sum<G>(y1);
    y2 = group_sum<G>(y2);
    if (lane < 3) epi.store(3 * row + lane, lane == 0 ? y0 : (lane == 1 ? y1 : y2), pre);
    k = k1;
    e = e1;
    c = c1;
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = v1[q];
  }
}

template <int G, class VT, class Epi>
__global__ __launch_bounds__(kBlock) void mass3_rows(int64_t nb, const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci,
                                                     const VT* __restrict__ mv,
                                                     const double* __restrict__ x, Epi epi, int flags) {
  const int64_t t = xcd_block(flags) * kBlock + threadIdx.x;
  const int64_t row = t / G;
  const int lane = threadIdx.x & (G - 1);
  typename Epi::Pre pre{};
  const bool wr = row < nb && lane < 3;
  if (wr) pre = epi.load(3 * row + lane);
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  if (row < nb) {
    const int64_t e = rp[row + 1];
    for (int64_t k = rp[row] + lane; k < e; k += G) {
      const int64_t c3 = 3 * (int64_t)ldm(ci + k, flags);
      const double m = (double)ldm(mv + k, flags);
      y0 += m * x[c3];
      y1 += m * x[c3 + 1];
      y2 += m * x[c3 + 2];
    }
  }
  y0 = group_sum<G>(y0);
  y1 = group_sum<G>(y1);
  y2 = group_sum<G>(y2);
  if (wr) epi.store(3 * row + lane, lane == 0 ? y0 : (lane == 1 ? y1 : y2), pre);
}

// --------------------------------------------------------------------------
// Row-group kernels (single column, BSR3 and MASS3).  The node rows are cut
// on the host into groups of consecutive rows holding at most 64 entries
// (or one longer row); group g owns rows [grow[g], grow[g+1]) and entries
// [gk[g], gk[g+1]).  A wavefront takes NG consecutive groups: lane l holds
// entry gk + l of the current group, so every lane carries an entry whatever
// the row lengths, and the entries' columns and values need no row offset
// first (only the group descriptor, loaded for all NG groups at once).  The
// loads of group i+1 (columns, values, row starts, epilogue operands) are
// issued right after the x gathers of group i, so they are in flight while
// group i is multiplied, reduced and stored.  Lanes find their row by a
// binary search over the group's row starts held one per lane, sum their
// products per row by a segmented butterfly, and lane t < 3 * rows writes
// DoF 3 * grow + t.  Longer rows (one per group) loop over the rest.
__device__ __forceinline__ int64_t rl64(int64_t v, int i) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), i);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), i);
  return ((int64_t)hi << 32) | (uint32_t)lo;
}

template <bool BSR, class VT, class Epi>
struct GroupStage {
  int c;                   // column node of this lane's entry (0 when none)
  double v[BSR ? 9 : 1];   // its value(s) (0 when none)
  int b;                   // lane j < rows: start of row j within the group
  typename Epi::Pre pre;   // lane t < 3 * rows: epilogue operands of DoF 3 * r0 + t
};

template <bool BSR, int NG, class VT, class Epi>
__global__ __launch_bounds__(kBlock) void group_rows(int64_t ngroups, const int64_t* __restrict__ grow,
                                                     const int64_t* __restrict__ gk,
                                                     const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ ci,
                                                     const VT* __restrict__ val, int64_t nbnz,
                                                     const double* __restrict__ x, Epi epi, int flags) {
  constexpr int NV = BSR ? 9 : 1;
  const int lane = threadIdx.x & 63;
  const int64_t g0 = (xcd_block(flags) * (kBlock / 64) + (threadIdx.x >> 6)) * NG;
  if (g0 >= ngroups) return;
  const int ng = (int)((ngroups - g0) < NG ? (ngroups - g0) : NG);
  // descriptors of groups g0 .. g0 + ng (the last one closes group ng - 1)
  int64_t dr = 0, dk = 0;
  if (lane <= ng) {
    dr = grow[g0 + lane];
    dk = gk[g0 + lane];
  }
  using Stage = GroupStage<BSR, VT, Epi>;
  auto load = [&](int i, Stage& st) {
    const int64_t r0 = rl64(dr, i), r1 = rl64(dr, i + 1);
    const int64_t k0 = rl64(dk, i), k1 = rl64(dk, i + 1);
    const int nrows = (int)(r1 - r0);
    const int64_t k = k0 + lane;
    const bool ok = k < k1;
    st.c = ok ? (int)ldm(ci + k, flags) : 0;
#pragma unroll
    for (int q = 0; q < NV; ++q)
      st.v[q] = ok ? (double)ldm(val + (BSR ? bidx(k, q, nbnz, flags) : k), flags) : 0.0;
    st.b = lane < nrows ? (int)(rp[r0 + lane] - k0) : 0x7fffffff;
    if (lane < 3 * nrows) st.pre = epi.load(3 * r0 + lane);
  };
  Stage cur, nxt;
  load(0, cur);
  for (int i = 0; i < ng; ++i) {
    const int64_t r0 = rl64(dr, i), r1 = rl64(dr, i + 1);
    const int64_t k0 = rl64(dk, i), k1 = rl64(dk, i + 1);
    const int nrows = (int)(r1 - r0);
    const int len = (int)(k1 - k0);
    // gathers of this group, then the next group's loads
    const int64_t c3 = 3 * (int64_t)cur.c;
    const double xa = x[c3], xb = x[c3 + 1], xc = x[c3 + 2];
    if (i + 1 < ng) load(i + 1, nxt);
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (lane < len) {
      if (BSR) {
        y0 = cur.v[0] * xa + cur.v[1 % NV] * xb + cur.v[2 % NV] * xc;
        y1 = cur.v[3 % NV] * xa + cur.v[4 % NV] * xb + cur.v[5 % NV] * xc;
        y2 = cur.v[6 % NV] * xa + cur.v[7 % NV] * xb + cur.v[8 % NV] * xc;
      } else {
        y0 = cur.v[0] * xa;
        y1 = cur.v[0] * xb;
        y2 = cur.v[0] * xc;
      }
    }
    if (len > 64) {
      // one long row: the remaining entries, 64 at a time
      for (int64_t k = k0 + 64 + lane; k < k1; k += 64) {
        const int64_t d3 = 3 * (int64_t)ldm(ci + k, flags);
        double v[NV];
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] = (double)ldm(val + (BSR ? bidx(k, q, nbnz, flags) : k), flags);
        const double ua = x[d3], ub = x[d3 + 1], uc = x[d3 + 2];
        if (BSR) {
          y0 += v[0] * ua + v[1 % NV] * ub + v[2 % NV] * uc;
          y1 += v[3 % NV] * ua + v[4 % NV] * ub + v[5 % NV] * uc;
          y2 += v[6 % NV] * ua + v[7 % NV] * ub + v[8 % NV] * uc;
        } else {
          y0 += v[0] * ua;
          y1 += v[0] * ub;
          y2 += v[0] * uc;
        }
      }
    }
    // row of this lane's entry: the largest j < nrows with b_j <= lane
    int rid = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const int cand = rid + step;
      const int bc = __shfl(cur.b, cand & 63, 64);
      if (cand < nrows && bc <= lane) rid = cand;
    }
    // segmented sums: afterwards the first lane of each row holds the row's sum
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int orid = __shfl_down(rid, d, 64);
      const double o0 = __shfl_down(y0, d, 64), o1 = __shfl_down(y1, d, 64), o2 = __shfl_down(y2, d, 64);
      if (lane + d < 64 && orid == rid) {
        y0 += o0;
        y1 += o1;
        y2 += o2;
      }
    }
    // DoF 3 * r0 + t of row j = t / 3, component t % 3
    for (int t0 = 0; t0 < 3 * nrows; t0 += 64) {
      const int t = t0 + lane;
      const int j = t / 3 < 64 ? t / 3 : 63;
      const int bj = __shfl(cur.b, j, 64);
      const int bn = __shfl(cur.b, (j + 1) & 63, 64);
      const int end = (j + 1 < nrows) ? bn : len;
      const int src = bj & 63;
      const double s0 = __shfl(y0, src, 64), s1 = __shfl(y1, src, 64), s2 = __shfl(y2, src, 64);
      if (t < 3 * nrows) {
        const int p = t % 3;
        const double sum = (end > bj) ? (p == 0 ? s0 : (p == 1 ? s1 : s2)) : 0.0;
        if (t0 == 0) epi.store(3 * r0 + t, sum, cur.pre);
        else epi.store(3 * r0 + t, sum, epi.load(3 * r0 + t));
      }
    }
    cur = nxt;
  }
}

template <bool BSR, int NG, class VT, class Epi>
static void launch_groups_NG(const DevMat& M, const VT* v, const double* x, Epi epi, hipStream_t s) {
  const RowGroups& G = *M.rgroups;
  const int64_t waves = (G.ngroups + NG - 1) / NG;
  const int64_t grid = (waves + kBlock / 64 - 1) / (kBlock / 64);
  if (grid <= 0) return;
  hipLaunchKernelGGL((group_rows<BSR, NG, VT, Epi>), dim3((unsigned)grid), dim3(kBlock), 0, s, G.ngroups, G.grow,
                     G.gk, M.browptr, M.bcol, v, M.nbnz, x, epi, kflags(M));
}

template <class VT, class Epi>
static void launch_groups(const DevMat& M, const VT* v, const double* x, Epi epi, hipStream_t s) {
  const bool bsr = M.layout == L_BSR3;
  switch (M.rgroups->per_wave) {
    case 1: bsr ? launch_groups_NG<true, 1>(M, v, x, epi, s) : launch_groups_NG<false, 1>(M, v, x, epi, s); break;
    case 2: bsr ? launch_groups_NG<true, 2>(M, v, x, epi, s) : launch_groups_NG<false, 2>(M, v, x, epi, s); break;
    case 4: bsr ? launch_groups_NG<true, 4>(M, v, x, epi, s) : launch_groups_NG<false, 4>(M, v, x, epi, s); break;
    default: bsr ? launch_groups_NG<true, 8>(M, v, x, epi, s) : launch_groups_NG<false, 8>(M, v, x, epi, s); break;
  }
  NMG_HIP(hipGetLastError());
}

static int pick_group(double avg) {
  if (avg > 40.0) return 64;
  if (avg > 20.0) return 32;
  if (avg > 10.0) return 16;
  return 8;
}

// pipelined BSR3 rows: ceil(nb / NR) lane groups of G lanes
template <int G, int NR, class VT, class Epi>
static void launch_pipe(const DevMat& M, const VT* v, const double* x, Epi epi, hipStream_t s) {
  const int64_t groups = (M.nb + NR - 1) / NR;
  const int64_t grid = (groups * G + kBlock - 1) / kBlock;
  if (grid <= 0) return;
  hipLaunchKernelGGL((bsr3_rows_pipe<G, NR, VT, Epi>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb, M.nbnz,
                     M.browptr, M.bcol, v, x, epi, kflags(M));
}

template <int G, class VT, class Epi>
static void launch_G(const DevMat& M, const VT* v, const double* x, Epi epi, hipStream_t s) {
  if (M.layout == L_CSR) {
    const int64_t thr = M.nrow * G;
    const int64_t grid = (thr + kBlock - 1) / kBlock;
    if (grid > 0)
      hipLaunchKernelGGL((csr_rows<G, VT, Epi>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nrow,
                         M.rowptr, M.col, v, x, epi, kflags(M));
  } else {
    if constexpr (G >= 16) {
      if (M.layout == L_BSR3 && M.pipe > 0) {
        switch (M.pipe) {
          case 2: launch_pipe<G, 2>(M, v, x, epi, s); break;
          case 4: launch_pipe<G, 4>(M, v, x, epi, s); break;
          default: launch_pipe<G, 8>(M, v, x, epi, s); break;
        }
        return;
      }
    }
    const int64_t thr = M.nb * G;
    const int64_t grid = (thr + kBlock - 1) / kBlock;
    if (grid <= 0) return;
    if (M.layout == L_BSR3)
      hipLaunchKernelGGL((bsr3_rows<G, VT, Epi>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb,
                         M.nbnz, M.browptr, M.bcol, v, x, epi, kflags(M));
    else
      hipLaunchKernelGGL((mass3_rows<G, VT, Epi>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb,
                         M.browptr, M.bcol, v, x, epi, kflags(M));
  }
}

template <class VT, class Epi>
static void launch_rows(const DevMat& M, const VT* v, const double* x, Epi epi, hipStream_t s) {
  double avg = M.layout == L_CSR ? (double)M.nnz / (double)(M.nrow ? M.nrow : 1)
                                 : (double)M.nbnz / (double)(M.nb ? M.nb : 1);
  const int g = M.group > 0 ? M.group : pick_group(avg);
  switch (g) {
    case 64: launch_G<64>(M, v, x, epi, s); break;
    case 32: launch_G<32>(M, v, x, epi, s); break;
    case 16: launch_G<16>(M, v, x, epi, s); break;
    default: launch_G<8>(M, v, x, epi, s); break;
  }
  NMG_HIP(hipGetLastError());
}

template <class Epi>
static void dispatch(const DevMat& M, const double* x, Epi epi, hipStream_t s, bool fp32) {
  if (M.rgroups && M.layout != L_CSR) {
    if (fp32 && M.bval32) launch_groups(M, M.bval32, x, epi, s);
    else launch_groups(M, M.bval, x, epi, s);
    return;
  }
  if (M.layout == L_CSR) {
    if (fp32 && M.val32) launch_rows(M, M.val32, x, epi, s);
    else launch_rows(M, M.val, x, epi, s);
  } else {
    if (fp32 && M.bval32) launch_rows(M, M.bval32, x, epi, s);
    else launch_rows(M, M.bval, x, epi, s);
  }
}

void k_spmv(const DevMat& M, const double* x, double* y, hipStream_t s, bool fp32) {
  if (M.single_blocked && !M.rgroups) return k_block_spmv(1, M, x, y, s, fp32);
  dispatch(M, x, EpiStore{y}, s, fp32);
}

void k_cheb_step(const DevMat& M, const double* b, const double* xk, double* xm, double w,
                 double ith, hipStream_t s, bool fp32) {
  if (M.single_blocked && !M.rgroups) return k_block_cheb_step(1, M, b, xk, xm, w, ith, s, fp32);
  dispatch(M, xk, EpiCheb{b, xk, xm, w, ith}, s, fp32);
}

void k_filter_step(const DevMat& A, const double* x, const double* vk, double* vkm, double* y,
                   double t, double cc, double mu, bool first, hipStream_t s, bool fp32);

// --------------------------------------------------------------------------
// vector kernels
__global__ void fill_k(int64_t n, double v, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = v;
}
__global__ void scale_copy_k(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = a * x[i];
}
__global__ void mul_k(int64_t n, const double* __restrict__ d, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    x[i] *= d[i];
}
__global__ void axpy_dev_k(int64_t n, const double* __restrict__ a, const double* __restrict__ x,
                           double* __restrict__ y) {
  const double av = a[0];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] -= av * x[i];
}
__global__ void invscale_dev_k(int64_t n, const double* __restrict__ beta, double* __restrict__ x,
                               double* __restrict__ z) {
  const double ib = 1.0 / beta[0];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    x[i] *= ib;
    if (z) z[i] *= ib;
  }
}

static unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void k_fill(int64_t n, double v, double* y, hipStream_t s) {
  hipLaunchKernelGGL(fill_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, v, y);
  NMG_HIP(hipGetLastError());
}
void k_copy(int64_t n, const double* x, double* y, hipStream_t s) {
  NMG_HIP(hipMemcpyAsync(y, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
}
void k_scale_copy(int64_t n, double a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(scale_copy_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, y);
  NMG_HIP(hipGetLastError());
}
void k_mul(int64_t n, const double* d, double* x, hipStream_t s) {
  hipLaunchKernelGGL(mul_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, x);
  NMG_HIP(hipGetLastError());
}
void k_axpy_dev(int64_t n, const double* a, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(axpy_dev_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, y);
  NMG_HIP(hipGetLastError());
}
void k_invscale_dev(int64_t n, const double* beta, double* x, double* z, hipStream_t s) {
  hipLaunchKernelGGL(invscale_dev_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, beta, x, z);
  NMG_HIP(hipGetLastError());
}

// row scalings and the unfused filter update (operators that are not one
// sparse product, e.g. the fluid-solid Schur operator) ----------------------
// y[i*K + c] = d[i] * x[i*K + c]   (K = 1: y = d o x)
__global__ void rowmul_k(int64_t n, int K, const double* __restrict__ d, const double* __restrict__ x,
                         double* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock)
    y[t] = d[t / K] * x[t];
}
// y[i*K + c] = d[i] * (a[i*K + c] + b[i*K + c])
__global__ void add_rowmul_k(int64_t n, int K, const double* __restrict__ d, const double* __restrict__ a,
                             const double* __restrict__ b, double* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock)
    y[t] = d[t / K] * (a[t] + b[t]);
}
// v = t (w - cc vk) - [!first] vkm;  vkm <- v;  y += mu[c] v
__global__ void filter_update_k(int64_t n, int K, const double* __restrict__ w, const double* __restrict__ vk,
                                double* __restrict__ vkm, double* __restrict__ y, const double* __restrict__ mu,
                                double mu0, double t, double cc, int first) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n * K; i += (int64_t)gridDim.x * kBlock) {
    const double v = t * (w[i] - cc * vk[i]) - (first ? 0.0 : vkm[i]);
    vkm[i] = v;
    y[i] += (mu ? mu[i % K] : mu0) * v;
  }
}
void k_rowmul(int64_t n, int K, const double* d, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(rowmul_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, d, x, y);
  NMG_HIP(hipGetLastError());
}
void k_add_rowmul(int64_t n, int K, const double* d, const double* a, const double* b, double* y,
                  hipStream_t s) {
  hipLaunchKernelGGL(add_rowmul_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, d, a, b, y);
  NMG_HIP(hipGetLastError());
}
void k_filter_update(int64_t n, int K, const double* w, const double* vk, double* vkm, double* y,
                     const double* mu, double mu0, double t, double cc, bool first, hipStream_t s) {
  hipLaunchKernelGGL(filter_update_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, w, vk, vkm, y, mu, mu0, t,
                     cc, first ? 1 : 0);
  NMG_HIP(hipGetLastError());
}

// plain streaming copy (calibration of the HBM byte counters)
__global__ void copy_k(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) y[i] = x[i];
}
void k_copy_kernel(int64_t n, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(copy_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, y);
  NMG_HIP(hipGetLastError());
}

// deterministic two-stage reductions -----------------------------------------
__global__ __launch_bounds__(kBlock) void dot_part_k(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* part) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += x[i] * y[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(kBlock) void sum_final_k(const double* part, int np, double* out,
                                                      int sqrt_out) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kBlock) s += part[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = sqrt_out ? sqrt(s) : s;
}

size_t dot_work_size() { return 2 * kDotBlocks; }

void k_dot(int64_t n, const double* x, const double* y, double* out, bool sqrt_out, double* work,
           hipStream_t s) {
  hipLaunchKernelGGL(dot_part_k, dim3(kDotBlocks), dim3(kBlock), 0, s, n, x, y, work);
  hipLaunchKernelGGL(sum_final_k, dim3(1), dim3(kBlock), 0, s, work, kDotBlocks, out,
                     sqrt_out ? 1 : 0);
  NMG_HIP(hipGetLastError());
}

// K column dot products of two interleaved blocks (i*K + c):
// out[c] = sum_i x[i*K + c] y[i*K + c].  The grid stride is a multiple of K,
// so every thread stays on the column threadIdx.x % K.
__global__ __launch_bounds__(kBlock) void block_dot_part_k(int64_t n, int K, const double* __restrict__ x,
                                                           const double* __restrict__ y, double* part) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  const int64_t nk = n * K;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nk; t += (int64_t)gridDim.x * kBlock)
    s += x[t] * y[t];
  sh[threadIdx.x] = s;
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double q = 0.0;
    for (int j = threadIdx.x; j < kBlock; j += K) q += sh[j];
    part[(int64_t)threadIdx.x * kDotBlocks + blockIdx.x] = q;
  }
}
__global__ __launch_bounds__(kBlock) void block_dot_final_k(const double* part, double* out) {
  __shared__ double sh[kBlock / 64];
  const double* p = part + (int64_t)blockIdx.x * kDotBlocks;
  double s = 0.0;
  for (int i = threadIdx.x; i < kDotBlocks; i += kBlock) s += p[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}
size_t block_dot_work_size(int K) { return (size_t)K * kDotBlocks; }
void k_block_dots(int64_t n, int K, const double* x, const double* y, double* out, double* work, hipStream_t s) {
  if (K < 1 || K > 64 || (kBlock % K) != 0) throw Error{NMG_EINVAL, "block dots: K must divide the block size"};
  hipLaunchKernelGGL(block_dot_part_k, dim3(kDotBlocks), dim3(kBlock), 0, s, n, K, x, y, work);
  hipLaunchKernelGGL(block_dot_final_k, dim3(K), dim3(kBlock), 0, s, work, out);
  NMG_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void resid_part_k(int64_t n, const double* __restrict__ ax,
                                                       const double* __restrict__ bx,
                                                       const double* __restrict__ x, double lam,
                                                       double* part) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0, q = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double r = ax[i] - lam * bx[i];
    s += r * r;
    q += x[i] * bx[i];
  }
  s = block_sum(s, sh);
  q = block_sum(q, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[kDotBlocks + blockIdx.x] = q;
  }
}

void k_resid(int64_t n, const double* ax, const double* bx, const double* x, double lam, double* out,
             double* work, hipStream_t s) {
  hipLaunchKernelGGL(resid_part_k, dim3(kDotBlocks), dim3(kBlock), 0, s, n, ax, bx, x, lam, work);
  hipLaunchKernelGGL(sum_final_k, dim3(1), dim3(kBlock), 0, s, work, kDotBlocks, out, 0);
  hipLaunchKernelGGL(sum_final_k, dim3(1), dim3(kBlock), 0, s, work + kDotBlocks, kDotBlocks,
                     out + 1, 0);
  NMG_HIP(hipGetLastError());
}

// tall-skinny GEMV for the re-orthogonalisation ------------------------------
constexpr int kCG = 8;  // columns per block in V^T z

static int gemv_chunks(int64_t n, int k) {
  const int cgroups = (k + kCG - 1) / kCG;
  int64_t ch = 2048 / (cgroups > 0 ? cgroups : 1);
  int64_t maxch = (n + 4 * kBlock - 1) / (4 * kBlock);
  if (ch > maxch) ch = maxch;
  if (ch < 1) ch = 1;
  if (ch > 1024) ch = 1024;
  return (int)ch;
}

// partial sums of V^T z: at most (k + kCG) columns x 1024 chunks
size_t gemv_work_size(int64_t n, int k) {
  (void)n;
  return (size_t)(k + kCG) * 1024 + 16;
}

__global__ __launch_bounds__(kBlock) void gemv_t_part_k(int64_t n, int k, const double* __restrict__ V,
                                                        int64_t ld, const double* __restrict__ z,
                                                        double* part, int nch) {
  __shared__ double sh[kCG][kBlock / 64];
  const int j0 = blockIdx.y * kCG;
  const int64_t per = ((n + nch - 1) / nch + kBlock - 1) / kBlock * kBlock;
  const int64_t start = (int64_t)blockIdx.x * per;
  int64_t end = start + per;
  if (end > n) end = n;
  double acc[kCG];
#pragma unroll
  for (int c = 0; c < kCG; ++c) acc[c] = 0.0;
  const int ncol = (k - j0) < kCG ? (k - j0) : kCG;
  for (int64_t i = start + threadIdx.x; i < end; i += kBlock) {
    const double zi = z[i];
#pragma unroll
    for (int c = 0; c < kCG; ++c)
      if (c < ncol) acc[c] += V[(int64_t)(j0 + c) * ld + i] * zi;
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < kCG; ++c) {
    double v = group_sum<64>(acc[c]);
    if (l == 0) sh[c][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < ncol) {
    double s = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) s += sh[threadIdx.x][i];
    part[(int64_t)(j0 + threadIdx.x) * nch + blockIdx.x] = s;
  }
}

__global__ void gemv_t_final_k(int k, const double* __restrict__ part, int nch, double* h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[(int64_t)j * nch + c];
  h[j] = s;
}

void k_gemv_t(int64_t n, int k, const double* V, int64_t ld, const double* z, double* h, double* work,
              hipStream_t s) {
  if (k <= 0) return;
  const int nch = gemv_chunks(n, k);
  dim3 grid(nch, (k + kCG - 1) / kCG);
  hipLaunchKernelGGL(gemv_t_part_k, grid, dim3(kBlock), 0, s, n, k, V, ld, z, work, nch);
  hipLaunchKernelGGL(gemv_t_final_k, dim3((k + 127) / 128), dim3(128), 0, s, k, work, nch, h);
  NMG_HIP(hipGetLastError());
}

// Tall-skinny C = A^T B (A: n x m, B: n x K, both column-major; C: m x K):
// the reduction runs over the long dimension n, so it is split into chunks
// (grid.x) and groups of CB columns of A (grid.y); each thread accumulates
// CB x K products over its rows and the workgroup reduces them into one
// partial per chunk; a second kernel sums the chunks.  Memory-bound
// (2K/8 flop per byte of A), so plain FP64 FMAs stream A at full rate.

template <int K, int CB>
__global__ __launch_bounds__(kBlock) void gemm_tn_part_k(int64_t n, int m, const double* __restrict__ A,
                                                         int64_t lda, const double* __restrict__ B,
                                                         int64_t ldb, double* __restrict__ part, int nch) {
  __shared__ double sh[kBlock / 64][CB * K];
  const int j0 = blockIdx.y * CB;
  const int64_t per = ((n + nch - 1) / nch + kBlock - 1) / kBlock * kBlock;
  const int64_t start = (int64_t)blockIdx.x * per;
  const int64_t end = start + per < n ? start + per : n;
  const int ncol = (m - j0) < CB ? (m - j0) : CB;
  double acc[CB][K];
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[c][k] = 0.0;
  for (int64_t i = start + threadIdx.x; i < end; i += kBlock) {
    double b[K];
#pragma unroll
    for (int k = 0; k < K; ++k) b[k] = B[(int64_t)k * ldb + i];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      if (c < ncol) {
        const double a = A[(int64_t)(j0 + c) * lda + i];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[c][k] += a * b[k];
      }
    }
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < CB; ++c)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double v = group_sum<64>(acc[c][k]);
      if (l == 0) sh[w][c * K + k] = v;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < CB * K; t += kBlock) {
    const int c = t / K, k = t % K;
    if (c < ncol) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < kBlock / 64; ++q) v += sh[q][t];
      part[((int64_t)k * m + j0 + c) * nch + blockIdx.x] = v;
    }
  }
}

__global__ void gemm_tn_final_k(int m, int K, const double* __restrict__ part, int nch, double* __restrict__ C,
                                int64_t ldc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * K) return;
  const int k = t / m, j = t % m;
  const double* p = part + (int64_t)t * nch;
  double v = 0.0;
  for (int q = 0; q < nch; ++q) v += p[q];
  C[(int64_t)k * ldc + j] = v;
}

static int gemm_tn_cb(int K) { return K >= 16 ? 4 : (K >= 8 ? 8 : (K >= 4 ? 16 : 32)); }

static int gemm_tn_chunks(int64_t n, int m, int K) {
  const int groups = (m + gemm_tn_cb(K) - 1) / gemm_tn_cb(K);
  int64_t ch = 1024 / (groups > 0 ? groups : 1);
  const int64_t maxch = (n + 4 * kBlock - 1) / (4 * kBlock);
  if (ch > maxch) ch = maxch;
  if (ch < 1) ch = 1;
  return (int)ch;
}

size_t gemm_tn_work_size(int64_t n, int m, int K) { return (size_t)gemm_tn_chunks(n, m, K) * m * K + 16; }

void k_gemm_tn(int64_t n, int m, int K, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
               int64_t ldc, double* work, hipStream_t s) {
  if (m <= 0) return;
  const int nch = gemm_tn_chunks(n, m, K);
  const int cb = gemm_tn_cb(K);
  dim3 grid(nch, (m + cb - 1) / cb);
  switch (K) {
    case 1: hipLaunchKernelGGL((gemm_tn_part_k<1, 32>), grid, dim3(kBlock), 0, s, n, m, A, lda, B, ldb, work, nch); break;
    case 2: hipLaunchKernelGGL((gemm_tn_part_k<2, 32>), grid, dim3(kBlock), 0, s, n, m, A, lda, B, ldb, work, nch); break;
    case 4: hipLaunchKernelGGL((gemm_tn_part_k<4, 16>), grid, dim3(kBlock), 0, s, n, m, A, lda, B, ldb, work, nch); break;
    case 8: hipLaunchKernelGGL((gemm_tn_part_k<8, 8>), grid, dim3(kBlock), 0, s, n, m, A, lda, B, ldb, work, nch); break;
    case 16: hipLaunchKernelGGL((gemm_tn_part_k<16, 4>), grid, dim3(kBlock), 0, s, n, m, A, lda, B, ldb, work, nch); break;
    default: throw Error{NMG_EINVAL, "tall-skinny product: K must be 1, 2, 4, 8 or 16"};
  }
  hipLaunchKernelGGL(gemm_tn_final_k, dim3((m * K + 255) / 256), dim3(256), 0, s, m, K, work, nch, C, ldc);
  NMG_HIP(hipGetLastError());
}

constexpr int kHT = 256;  // h tile staged in LDS

__global__ __launch_bounds__(kBlock) void gemv_n_sub_k(int64_t n, int k, const double* __restrict__ Z,
                                                       int64_t ld, const double* __restrict__ h,
                                                       double* __restrict__ z) {
  __shared__ double hs[kHT];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double s = 0.0;
  for (int j0 = 0; j0 < k; j0 += kHT) {
    const int nt = (k - j0) < kHT ? (k - j0) : kHT;
    __syncthreads();
    if ((int)threadIdx.x < nt) hs[threadIdx.x] = h[j0 + threadIdx.x];
    __syncthreads();
    if (i < n)
      for (int j = 0; j < nt; ++j) s += Z[(int64_t)(j0 + j) * ld + i] * hs[j];
  }
  if (i < n) z[i] -= s;
}

void k_gemv_n_sub(int64_t n, int k, const double* Z, int64_t ld, const double* h, double* z,
                  hipStream_t s) {
  if (k <= 0) return;
  const int64_t grid = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(gemv_n_sub_k, dim3((unsigned)grid), dim3(kBlock), 0, s, n, k, Z, ld, h, z);
  NMG_HIP(hipGetLastError());
}

// --------------------------------------------------------------------------
// matrix maintenance kernels
__global__ void csr_scale_k(int64_t nrow, const int64_t* rp, const int32_t* ci, double* v,
                            const double* dl, const double* dr, double sign) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrow) return;
  const double a = sign * dl[r];
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k) v[k] = a * v[k] * dr[ci[k]];
}
__global__ void bsr3_scale_k(int64_t nb, int64_t nbnz, const int64_t* rp, const int32_t* ci, double* bv,
                             const double* dl, const double* dr, double sign, int flags) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nb) return;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
    const int64_t c = ci[k];
    for (int p = 0; p < 9; ++p) {
      const int rr = p / 3, cc = p % 3;
      const int64_t i = bidx(k, p, nbnz, flags);
      bv[i] = sign * dl[3 * r + rr] * bv[i] * dr[3 * c + cc];
    }
  }
}
__global__ void mass3_scale_k(int64_t nb, const int64_t* rp, const int32_t* ci, double* mv,
                              const double* dl, const double* dr, double sign) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nb) return;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k) mv[k] = sign * dl[3 * r] * mv[k] * dr[3 * (int64_t)ci[k]];
}

void scale_devmat(DevMat& M, const double* dl, const double* dr, double sign, hipStream_t s) {
  if (M.layout == L_CSR) {
    hipLaunchKernelGGL(csr_scale_k, dim3((unsigned)((M.nrow + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nrow, M.rowptr, M.col, M.val, dl, dr, sign);
  } else if (M.layout == L_BSR3) {
    hipLaunchKernelGGL(bsr3_scale_k, dim3((unsigned)((M.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nb, M.nbnz, M.browptr, M.bcol, M.bval, dl, dr, sign, kflags(M));
  } else {
    hipLaunchKernelGGL(mass3_scale_k, dim3((unsigned)((M.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nb, M.browptr, M.bcol, M.bval, dl, dr, sign);
  }
  NMG_HIP(hipGetLastError());
  if (M.val32 || M.bval32) make_fp32(M, s);
}

__global__ void csr_diag_k(int64_t nrow, const int64_t* rp, const int32_t* ci, const double* v, double* d) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nrow) return;
  double x = 0.0;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
    if (ci[k] == r) x = v[k];
  d[r] = x;
}
__global__ void bsr3_diag_k(int64_t nb, int64_t nbnz, const int64_t* rp, const int32_t* ci,
                            const double* bv, double* d, int flags) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nb) return;
  double x0 = 0, x1 = 0, x2 = 0;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
    if (ci[k] == r) {
      x0 = bv[bidx(k, 0, nbnz, flags)];
      x1 = bv[bidx(k, 4, nbnz, flags)];
      x2 = bv[bidx(k, 8, nbnz, flags)];
    }
  d[3 * r] = x0;
  d[3 * r + 1] = x1;
  d[3 * r + 2] = x2;
}
__global__ void mass3_diag_k(int64_t nb, const int64_t* rp, const int32_t* ci, const double* mv, double* d) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= nb) return;
  double x = 0;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
    if (ci[k] == r) x = mv[k];
  d[3 * r] = d[3 * r + 1] = d[3 * r + 2] = x;
}

void diag_devmat(const DevMat& M, double* d, hipStream_t s) {
  if (M.layout == L_CSR)
    hipLaunchKernelGGL(csr_diag_k, dim3((unsigned)((M.nrow + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nrow, M.rowptr, M.col, M.val, d);
  else if (M.layout == L_BSR3)
    hipLaunchKernelGGL(bsr3_diag_k, dim3((unsigned)((M.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nb, M.nbnz, M.browptr, M.bcol, M.bval, d, kflags(M));
  else
    hipLaunchKernelGGL(mass3_diag_k, dim3((unsigned)((M.nb + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       M.nb, M.browptr, M.bcol, M.bval, d);
  NMG_HIP(hipGetLastError());
}

__global__ void to_f32_k(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = (float)x[i];
}

void make_fp32(DevMat& M, hipStream_t s) {
  M.touch();
  if (M.layout == L_CSR) {
    if (!M.val32) NMG_HIP(hipMalloc(&M.val32, sizeof(float) * (M.nnz > 0 ? M.nnz : 1)));
    hipLaunchKernelGGL(to_f32_k, dim3(grid_for(M.nnz)), dim3(kBlock), 0, s, M.nnz, M.val, M.val32);
  } else {
    const int64_t nv = M.value_count();
    if (!M.bval32) NMG_HIP(hipMalloc(&M.bval32, sizeof(float) * (nv > 0 ? nv : 1)));
    hipLaunchKernelGGL(to_f32_k, dim3(grid_for(nv)), dim3(kBlock), 0, s, nv, M.bval, M.bval32);
  }
  NMG_HIP(hipGetLastError());
}

}  // namespace nmg

// ==========================================================================
// k-column ("block") products for the block filter and the batched slices.
// Block vectors are node-major and column-interleaved: entry (dof i, column c)
// at i*K + c, so the three components of a neighbour node for all K columns
// are 3K contiguous doubles.  One wavefront per node row: lane = e*K + c with
// E = 64/K entry slots; the slots stride the row, the K columns share the
// matrix entry; a butterfly over the slots leaves every lane with the sums,
// and slot p < 3 writes component p of its column.
namespace nmg {

template <int K>
struct BEpiStore {
  double* y;
  struct Pre {};
  __device__ __forceinline__ Pre load(int64_t, int) const { return Pre{}; }
  __device__ __forceinline__ void store(int64_t i, int c, double v, const Pre&) const { y[i * K + c] = v; }
};
template <int K>
struct BEpiCheb {
  const double* b;
  const double* xk;
  double* xm;
  double w, ith;
  struct Pre {
    double b, xk, xm;
  };
  __device__ __forceinline__ Pre load(int64_t i, int c) const {
    const int64_t o = i * K + c;
    return Pre{b[o], xk[o], xm[o]};
  }
  __device__ __forceinline__ void store(int64_t i, int c, double v, const Pre& p) const {
    xm[i * K + c] = p.xm + w * (p.xk - p.xm + (p.b - v) * ith);
  }
};
// the Chebyshev step with FP32 vectors (b, x_k, x_{k-1} stored in FP32; the
// product and the update are formed in FP64 and rounded once on store)
template <int K>
struct BEpiCheb32 {
  const float* b;
  const float* xk;
  float* xm;
  double w, ith;
  struct Pre {
    double b, xk, xm;
  };
  __device__ __forceinline__ Pre load(int64_t i, int c) const {
    const int64_t o = i * K + c;
    return Pre{(double)b[o], (double)xk[o], (double)xm[o]};
  }
  __device__ __forceinline__ void store(int64_t i, int c, double v, const Pre& p) const {
    xm[i * K + c] = (float)(p.xm + w * (p.xk - p.xm + (p.b - v) * ith));
  }
};
template <int K>
struct BEpiFilter {
  const double* vk;
  double* vkm;
  double* y;
  const double* mu;  // K coefficients of this degree (one per column)
  double t, cc;
  int first;
  struct Pre {
    double vk, vkm, y, mu;
  };
  __device__ __forceinline__ Pre load(int64_t i, int c) const {
    const int64_t o = i * K + c;
    return Pre{vk[o], first ? 0.0 : vkm[o], y[o], mu[c]};
  }
  __device__ __forceinline__ void store(int64_t i, int c, double v, const Pre& p) const {
    const int64_t o = i * K + c;
    const double vn = t * (v - cc * p.vk) - p.vkm;
    vkm[o] = vn;
    y[o] = p.y + p.mu * vn;
  }
};

template <int K>
__device__ __forceinline__ double slot_sum(double v) {
#pragma unroll
  for (int off = K; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Lane layout of the k-column kernels: a wavefront holds R node rows; each
// row has K columns x E entry slots (R*K*E = 64).  Lane l: row r = l/(K*E),
// column c = l%K, slot e = (l/K)%E.  Each lane covers U entries per pass:
// all U matrix loads are issued, then all gathers, then the FMAs, so a pass
// costs two dependent memory round trips.  Entries past the row end use
// column c of node 0 with weight 0.  A butterfly over the slots leaves every
// slot with the sums; slot e writes components e, e+E, ... (< 3).
template <int K, int R>
struct BlockMap {
  static_assert(K * R <= 64 && (64 % (K * R)) == 0, "bad block map");
  static constexpr int E = 64 / (K * R);
  static constexpr int U = (32 / E) > 8 ? 8 : ((32 / E) > 0 ? (32 / E) : 1);
};

template <int K, int E>
__device__ __forceinline__ double slots_sum(double v) {
#pragma unroll
  for (int off = K; off < K * E; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K, int R, class VT, class Epi, class XT>
__global__ __launch_bounds__(kBlock) void mass3_block_rows(int64_t nb, const int64_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci,
                                                           const VT* __restrict__ mv,
                                                           const XT* __restrict__ x, Epi epi, int flags) {
  constexpr int E = BlockMap<K, R>::E, U = BlockMap<K, R>::U;
  const int lane = threadIdx.x & 63;
  const int64_t row = ((xcd_block(flags) * kBlock + threadIdx.x) >> 6) * R + lane / (K * E);
  const int c = lane % K, e = (lane / K) % E;
  const bool live = row < nb;
  typename Epi::Pre pre[(3 + E - 1) / E];
#pragma unroll
  for (int q = 0; q < (3 + E - 1) / E; ++q)
    if (live && e + q * E < 3) pre[q] = epi.load(3 * row + e + q * E, c);
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  if (live) {
    const int64_t beg = rp[row], end = rp[row + 1];
    for (int64_t k0 = beg + e; k0 < end; k0 += (int64_t)U * E) {
      int64_t base[U];
      double m[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t k = k0 + (int64_t)u * E;
        const bool ok = k < end;
        base[u] = ok ? 3 * (int64_t)ldm(ci + k, flags) * K + c : (int64_t)c;
        m[u] = ok ? (double)ldm(mv + k, flags) : 0.0;
      }
      double xa[U], xb[U], xc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xa[u] = (double)x[base[u]];
        xb[u] = (double)x[base[u] + K];
        xc[u] = (double)x[base[u] + 2 * K];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        y0 += m[u] * xa[u];
        y1 += m[u] * xb[u];
        y2 += m[u] * xc[u];
      }
    }
  }
  y0 = slots_sum<K, E>(y0);
  y1 = slots_sum<K, E>(y1);
  y2 = slots_sum<K, E>(y2);
#pragma unroll
  for (int q = 0; q < (3 + E - 1) / E; ++q) {
    const int p = e + q * E;
    if (live && p < 3) epi.store(3 * row + p, c, p == 0 ? y0 : (p == 1 ? y1 : y2), pre[q]);
  }
}

template <int K, int R, class VT, class Epi, class XT>
__global__ __launch_bounds__(kBlock) void bsr3_block_rows(int64_t nb, int64_t nbnz,
                                                          const int64_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci,
                                                          const VT* __restrict__ bv,
                                                          const XT* __restrict__ x, Epi epi, int flags) {
  constexpr int E = BlockMap<K, R>::E;
  constexpr int U = BlockMap<K, R>::U > 2 ? 2 : BlockMap<K, R>::U;  // 9 values per entry
  const int lane = threadIdx.x & 63;
  const int64_t row = ((xcd_block(flags) * kBlock + threadIdx.x) >> 6) * R + lane / (K * E);
  const int c = lane % K, e = (lane / K) % E;
  const bool live = row < nb;
  typename Epi::Pre pre[(3 + E - 1) / E];
#pragma unroll
  for (int q = 0; q < (3 + E - 1) / E; ++q)
    if (live && e + q * E < 3) pre[q] = epi.load(3 * row + e + q * E, c);
  double y0 = 0.0, y1 = 0.0, y2 = 0.0;
  if (live) {
    const int64_t beg = rp[row], end = rp[row + 1];
    for (int64_t k0 = beg + e; k0 < end; k0 += (int64_t)U * E) {
      int64_t base[U], kk[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t k = k0 + (int64_t)u * E;
        ok[u] = k < end;
        kk[u] = ok[u] ? k : beg;
        base[u] = ok[u] ? 3 * (int64_t)ldm(ci + k, flags) * K + c : (int64_t)c;
      }
      double a[U][9], xa[U], xb[U], xc[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int p = 0; p < 9; ++p) a[u][p] = ok[u] ? (double)ldm(bv + bidx(kk[u], p, nbnz, flags), flags) : 0.0;
        xa[u] = (double)x[base[u]];
        xb[u] = (double)x[base[u] + K];
        xc[u] = (double)x[base[u] + 2 * K];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        y0 += a[u][0] * xa[u] + a[u][1] * xb[u] + a[u][2] * xc[u];
        y1 += a[u][3] * xa[u] + a[u][4] * xb[u] + a[u][5] * xc[u];
        y2 += a[u][6] * xa[u] + a[u][7] * xb[u] + a[u][8] * xc[u];
      }
    }
  }
  y0 = slots_sum<K, E>(y0);
  y1 = slots_sum<K, E>(y1);
  y2 = slots_sum<K, E>(y2);
#pragma unroll
  for (int q = 0; q < (3 + E - 1) / E; ++q) {
    const int p = e + q * E;
    if (live && p < 3) epi.store(3 * row + p, c, p == 0 ? y0 : (p == 1 ? y1 : y2), pre[q]);
  }
}

// generic CSR with K columns: one wavefront per scalar row, lanes (e, c)
template <int K, class VT, class Epi, class XT>
__global__ __launch_bounds__(kBlock) void csr_block_rows(int64_t nrow, const int64_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const VT* __restrict__ va,
                                                         const XT* __restrict__ x, Epi epi, int flags) {
  constexpr int E = 64 / K;
  const int64_t row = (xcd_block(flags) * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int c = lane % K, e = lane / K;
  const bool wr = row < nrow && e == 0;
  typename Epi::Pre pre{};
  if (wr) pre = epi.load(row, c);
  double s = 0.0;
  if (row < nrow) {
    const int64_t end = rp[row + 1];
    for (int64_t k = rp[row] + e; k < end; k += E)
      s += (double)ldm(va + k, flags) * (double)x[(int64_t)ldm(ci + k, flags) * K + c];
  }
  s = slot_sum<K>(s);
  if (wr) epi.store(row, c, s, pre);
}

// "Wide" k-column scalar-mass product: one lane per matrix entry, the K
// columns in registers.  G lanes share a node row (64/G rows per
// wavefront); a lane reads the entry's column and value once and the
// neighbour's 3K contiguous x values with 16-byte loads, so the loads per
// entry do not grow with K.  The 3K row sums are reduced over the G lanes by
// a butterfly; lane g then stores outputs g, g + G, ... of its row.
template <class XT>
__device__ __forceinline__ void load16(const XT* p, double* out);
template <>
__device__ __forceinline__ void load16<double>(const double* p, double* out) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  out[0] = v.x;
  out[1] = v.y;
}
template <>
__device__ __forceinline__ void load16<float>(const float* p, double* out) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}

template <int K, int G, class VT, class Epi, class XT>
__global__ __launch_bounds__(kBlock) void mass3_wide_rows(int64_t nb, const int64_t* __restrict__ rp,
                                                          const int32_t* __restrict__ ci,
                                                          const VT* __restrict__ mv,
                                                          const XT* __restrict__ x, Epi epi, int flags) {
  constexpr int W3 = 3 * K;
  constexpr int PER = 16 / sizeof(XT);  // values per 16-byte load
  static_assert(W3 % PER == 0, "3K must fill 16-byte loads");
  constexpr int Q = (W3 + G - 1) / G;   // outputs per lane
  const int lane = threadIdx.x & 63;
  const int64_t row = ((xcd_block(flags) * kBlock + threadIdx.x) >> 6) * (64 / G) + lane / G;
  const int g = lane % G;
  const bool live = row < nb;
  typename Epi::Pre pre[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = g + q * G;
    if (live && j < W3) pre[q] = epi.load(3 * row + j / K, j % K);
  }
  double acc[W3];
#pragma unroll
  for (int j = 0; j < W3; ++j) acc[j] = 0.0;
  if (live) {
    const int64_t end = rp[row + 1];
    for (int64_t k = rp[row] + g; k < end; k += G) {
      const XT* xb = x + 3 * (int64_t)ldm(ci + k, flags) * K;
      const double m = (double)ldm(mv + k, flags);
      double xv[W3];
#pragma unroll
      for (int j = 0; j < W3; j += PER) load16<XT>(xb + j, xv + j);
#pragma unroll
      for (int j = 0; j < W3; ++j) acc[j] += m * xv[j];
    }
  }
#pragma unroll
  for (int j = 0; j < W3; ++j) acc[j] = group_sum<G>(acc[j]);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = g + q * G;
    double v = 0.0;
#pragma unroll
    for (int jj = 0; jj < W3; ++jj)
      if (jj == j) v = acc[jj];
    if (live && j < W3) epi.store(3 * row + j / K, j % K, v, pre[q]);
  }
}

template <int K, int G, class VT, class Epi, class XT>
static void launch_wide_G(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s) {
  const int64_t waves = (M.nb + (64 / G) - 1) / (64 / G);
  const int64_t grid = (waves * 64 + kBlock - 1) / kBlock;
  if (grid <= 0) return;
  hipLaunchKernelGGL((mass3_wide_rows<K, G, VT, Epi, XT>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb,
                     M.browptr, M.bcol, v, x, epi, kflags(M));
}

// the wide layout applies to scalar-mass matrices with 3K a multiple of the
// 16-byte load (K >= 2 for FP64 vectors, K >= 4 for FP32)
template <int K, class XT>
constexpr bool wide_ok() {
  return K >= 2 && K <= 16 && (3 * K) % (16 / (int)sizeof(XT)) == 0;
}

template <int K, class VT, class Epi, class XT>
static bool launch_wide(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s) {
  if constexpr (!wide_ok<K, XT>()) {
    return false;
  } else {
    int idx = 0;
    while ((1 << idx) != K) ++idx;
    const int G = M.wide[idx];
    if (G <= 0 || M.layout != L_MASS3) return false;
    switch (G) {
      case 4: launch_wide_G<K, 4>(M, v, x, epi, s); break;
      case 8: launch_wide_G<K, 8>(M, v, x, epi, s); break;
      case 16: launch_wide_G<K, 16>(M, v, x, epi, s); break;
      default: launch_wide_G<K, 32>(M, v, x, epi, s); break;
    }
    NMG_HIP(hipGetLastError());
    return true;
  }
}

// rows per wavefront of the k-column kernels: chosen per (matrix, K) by the
// autotuner (tune_block_map); 0 means "not tuned" and uses R = 1
static int block_rows(const DevMat& M, int K) {
  for (int i = 0; i < 7; ++i)
    if ((1 << i) == K) return M.block_map[i] > 0 ? M.block_map[i] : 1;
  return 1;
}

template <int K, int R, class VT, class Epi, class XT>
static void launch_block_R(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s) {
  const int64_t waves = (M.nb + R - 1) / R;
  const int64_t grid = (waves * 64 + kBlock - 1) / kBlock;
  if (grid <= 0) return;
  if (M.layout == L_BSR3)
    hipLaunchKernelGGL((bsr3_block_rows<K, R, VT, Epi, XT>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb,
                       M.nbnz, M.browptr, M.bcol, v, x, epi, kflags(M));
  else
    hipLaunchKernelGGL((mass3_block_rows<K, R, VT, Epi, XT>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nb,
                       M.browptr, M.bcol, v, x, epi, kflags(M));
}

template <int K, class VT, class Epi, class XT>
static void launch_block(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s) {
  if (M.layout == L_CSR) {
    const int64_t grid = (M.nrow * 64 + kBlock - 1) / kBlock;
    if (grid > 0)
      hipLaunchKernelGGL((csr_block_rows<K, VT, Epi, XT>), dim3((unsigned)grid), dim3(kBlock), 0, s, M.nrow,
                         M.rowptr, M.col, v, x, epi, kflags(M));
  } else {
    const int R = block_rows(M, K);
    if (R == 1) launch_block_R<K, 1>(M, v, x, epi, s);
    else if (R == 2 && K * 2 <= 64) launch_block_R<K, (K * 2 <= 64 ? 2 : 1)>(M, v, x, epi, s);
    else if (R == 4 && K * 4 <= 64) launch_block_R<K, (K * 4 <= 64 ? 4 : 1)>(M, v, x, epi, s);
    else if (R == 8 && K * 8 <= 64) launch_block_R<K, (K * 8 <= 64 ? 8 : 1)>(M, v, x, epi, s);
    else launch_block_R<K, 1>(M, v, x, epi, s);
  }
  NMG_HIP(hipGetLastError());
}

template <int K, class VT, class Epi, class XT>
bool launch_tiled(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s);

template <int K, class Epi, class XT>
static void dispatch_block(const DevMat& M, const XT* x, Epi epi, hipStream_t s, bool fp32) {
  if (M.layout == L_MASS3) {
    const bool f = fp32 && M.bval32;
    if (f ? launch_wide<K>(M, M.bval32, x, epi, s) : launch_wide<K>(M, M.bval, x, epi, s)) return;
  }
  if (M.layout != L_CSR) {
    const bool f = fp32 && M.bval32;
    if (f ? launch_tiled<K>(M, M.bval32, x, epi, s) : launch_tiled<K>(M, M.bval, x, epi, s)) return;
  }
  if (M.layout == L_CSR) {
    if (fp32 && M.val32) launch_block<K>(M, M.val32, x, epi, s);
    else launch_block<K>(M, M.val, x, epi, s);
  } else {
    if (fp32 && M.bval32) launch_block<K>(M, M.bval32, x, epi, s);
    else launch_block<K>(M, M.bval, x, epi, s);
  }
}

#define NMG_BLOCK_SWITCH(K_, CALL)                                              \
  switch (K_) {                                                                 \
    case 1: { constexpr int KK = 1; CALL; break; }                              \
    case 2: { constexpr int KK = 2; CALL; break; }                              \
    case 4: { constexpr int KK = 4; CALL; break; }                              \
    case 8: { constexpr int KK = 8; CALL; break; }                              \
    case 16: { constexpr int KK = 16; CALL; break; }                            \
    case 32: { constexpr int KK = 32; CALL; break; }                            \
    case 64: { constexpr int KK = 64; CALL; break; }                            \
    default: throw Error{NMG_EINVAL, "block width must be 1, 2, 4, 8, 16, 32 or 64"};  \
  }

void k_filter_step(const DevMat& A, const double* x, const double* vk, double* vkm, double* y,
                   double t, double cc, double mu, bool first, hipStream_t s, bool fp32) {
  dispatch(A, x, EpiFilter{vk, vkm, y, t, cc, mu, first ? 1 : 0}, s, fp32);
}

void k_block_spmv(int K, const DevMat& M, const double* x, double* y, hipStream_t s, bool fp32) {
  NMG_BLOCK_SWITCH(K, (dispatch_block<KK>(M, x, BEpiStore<KK>{y}, s, fp32)));
}

void k_block_cheb_step(int K, const DevMat& M, const double* b, const double* xk, double* xm, double w,
                       double ith, hipStream_t s, bool fp32) {
  NMG_BLOCK_SWITCH(K, (dispatch_block<KK>(M, xk, BEpiCheb<KK>{b, xk, xm, w, ith}, s, fp32)));
}

void k_block_cheb_step32(int K, const DevMat& M, const float* b, const float* xk, float* xm, double w,
                         double ith, hipStream_t s, bool fp32) {
  NMG_BLOCK_SWITCH(K, (dispatch_block<KK>(M, xk, BEpiCheb32<KK>{b, xk, xm, w, ith}, s, fp32)));
}

// start of the FP32-vector Chebyshev solve: b32 = fl(b), x1 = fl(b/th), x0 = 0
__global__ void cheb32_start_k(int64_t n, double ith, const double* __restrict__ b, float* __restrict__ b32,
                               float* __restrict__ x1, float* __restrict__ x0) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const double v = b[i];
    b32[i] = (float)v;
    x1[i] = (float)(v * ith);
    x0[i] = 0.0f;
  }
}
void k_cheb32_start(int64_t n, double ith, const double* b, float* b32, float* x1, float* x0, hipStream_t s) {
  hipLaunchKernelGGL(cheb32_start_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, ith, b, b32, x1, x0);
  NMG_HIP(hipGetLastError());
}
// y = b - y (the residual of a refinement pass, after y = B~x)
__global__ void sub_from_k(int64_t n, const double* __restrict__ b, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = b[i] - y[i];
}
void k_sub_from(int64_t n, const double* b, double* y, hipStream_t s) {
  hipLaunchKernelGGL(sub_from_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, b, y);
  NMG_HIP(hipGetLastError());
}
// y += x (an FP32 correction added to an FP64 solution)
__global__ void add_f32_k(int64_t n, const float* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] += (double)x[i];
}
void k_add_f32(int64_t n, const float* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(add_f32_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, y);
  NMG_HIP(hipGetLastError());
}
__global__ void f32_to_f64_k(int64_t n, const float* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    y[i] = (double)x[i];
}
void k_f32_to_f64(int64_t n, const float* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(f32_to_f64_k, dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, y);
  NMG_HIP(hipGetLastError());
}

void k_block_filter_step(int K, const DevMat& A, const double* x, const double* vk, double* vkm, double* y,
                         const double* mu, double t, double cc, bool first, hipStream_t s, bool fp32) {
  NMG_BLOCK_SWITCH(K, (dispatch_block<KK>(A, x, BEpiFilter<KK>{vk, vkm, y, mu, t, cc, first ? 1 : 0}, s, fp32)));
}

// y[i*K + c] = mu[c] * x[i*K + c]
__global__ void block_colscale_k(int64_t n, int K, const double* __restrict__ mu, const double* __restrict__ x,
                                 double* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock)
    y[t] = mu[t % K] * x[t];
}
void k_block_colscale(int64_t n, int K, const double* mu, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(block_colscale_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, mu, x, y);
  NMG_HIP(hipGetLastError());
}

// interleaved n x K <-> P interleaved n x K/P parts (part p holds columns
// p*K/P .. (p+1)*K/P - 1)
struct Parts {
  double* p[4];
};
__global__ void split_parts_k(int64_t n, int K, int P, const double* __restrict__ x, Parts out) {
  const int H = K / P;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock) {
    const int64_t i = t / K;
    const int c = (int)(t % K);
    out.p[c / H][i * H + c % H] = x[t];
  }
}
__global__ void merge_parts_k(int64_t n, int K, int P, Parts in, double* __restrict__ y) {
  const int H = K / P;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock) {
    const int64_t i = t / K;
    const int c = (int)(t % K);
    y[t] = in.p[c / H][i * H + c % H];
  }
}
void k_split_parts(int64_t n, int K, int P, const double* x, double* const* parts, hipStream_t s) {
  NMG_CHECK(P >= 1 && P <= 4 && K % P == 0, NMG_EINVAL, "split into 1 to 4 equal parts");
  Parts o{};
  for (int q = 0; q < P; ++q) o.p[q] = parts[q];
  hipLaunchKernelGGL(split_parts_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, P, x, o);
  NMG_HIP(hipGetLastError());
}
void k_merge_parts(int64_t n, int K, int P, double* const* parts, double* y, hipStream_t s) {
  NMG_CHECK(P >= 1 && P <= 4 && K % P == 0, NMG_EINVAL, "merge of 1 to 4 equal parts");
  Parts o{};
  for (int q = 0; q < P; ++q) o.p[q] = parts[q];
  hipLaunchKernelGGL(merge_parts_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, P, o, y);
  NMG_HIP(hipGetLastError());
}

// column-major (n x K, ld) <-> interleaved (i*K + c)
__global__ void to_interleaved_k(int64_t n, int K, const double* __restrict__ cm, int64_t ld,
                                 double* __restrict__ il) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock) {
    const int64_t i = t / K;
    const int c = (int)(t % K);
    il[t] = cm[(int64_t)c * ld + i];
  }
}
__global__ void from_interleaved_k(int64_t n, int K, const double* __restrict__ il, double* __restrict__ cm,
                                   int64_t ld) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n * K; t += (int64_t)gridDim.x * kBlock) {
    const int64_t i = t / K;
    const int c = (int)(t % K);
    cm[(int64_t)c * ld + i] = il[t];
  }
}

void k_to_interleaved(int64_t n, int K, const double* cm, int64_t ld, double* il, hipStream_t s) {
  hipLaunchKernelGGL(to_interleaved_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, cm, ld, il);
  NMG_HIP(hipGetLastError());
}
void k_from_interleaved(int64_t n, int K, const double* il, double* cm, int64_t ld, hipStream_t s) {
  hipLaunchKernelGGL(from_interleaved_k, dim3(grid_for(n * K)), dim3(kBlock), 0, s, n, K, il, cm, ld);
  NMG_HIP(hipGetLastError());
}

}  // namespace nmg

// ==========================================================================
// LDS-tiled k-column products.  The node rows are cut into tiles of
// consecutive rows (the mesh numbering is spatially local); a workgroup owns
// one tile, first copies the x blocks (3K doubles) of every node its rows
// reference into LDS -- each node once per tile instead of once per
// referencing row -- and then forms the products from LDS through 16-bit
// tile-local column indices.  Epilogues work on global memory as before.
namespace nmg {

// TB threads (TB/64 wavefronts) share one tile's LDS: 1024 or 256
template <int K, int R, class VT, class Epi, bool BSR, int TB, class XT>
__global__ __launch_bounds__(TB) void tiled_block_rows(const int64_t* __restrict__ trp,
                                                           const int64_t* __restrict__ tcp,
                                                           const int32_t* __restrict__ tcols,
                                                           const int64_t* __restrict__ rp,
                                                           const uint16_t* __restrict__ lcol,
                                                           const VT* __restrict__ val, int64_t nbnz,
                                                           const XT* __restrict__ x, Epi epi, int flags) {
  extern __shared__ double xs[];
  constexpr int W3 = 3 * K;
  constexpr int WS = tile_stride(K);  // LDS stride of one node's block
  constexpr int E = BlockMap<K, R>::E;
  constexpr int U = BSR ? (BlockMap<K, R>::U > 2 ? 2 : BlockMap<K, R>::U) : BlockMap<K, R>::U;
  const int64_t tile = xcd_block(flags);
  const int64_t r0 = trp[tile], r1 = trp[tile + 1];
  const int64_t c0 = tcp[tile];
  const int nu = (int)(tcp[tile + 1] - c0);
  // stage the tile's x blocks
  for (int t = threadIdx.x; t < nu * W3; t += TB) {
    const int node = t / W3, j = t - node * W3;
    xs[node * WS + j] = (double)x[(int64_t)tcols[c0 + node] * W3 + j];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int c = lane % K, e = (lane / K) % E;
  const int rsub = lane / (K * E);
  for (int64_t rb = r0 + (int64_t)wave * R; rb < r1; rb += (int64_t)(TB / 64) * R) {
    const int64_t row = rb + rsub;
    const bool live = row < r1;
    typename Epi::Pre pre[(3 + E - 1) / E];
#pragma unroll
    for (int q = 0; q < (3 + E - 1) / E; ++q)
      if (live && e + q * E < 3) pre[q] = epi.load(3 * row + e + q * E, c);
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    if (live) {
      const int64_t beg = rp[row], end = rp[row + 1];
      for (int64_t k0 = beg + e; k0 < end; k0 += (int64_t)U * E) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int64_t k = k0 + (int64_t)u * E;
          if (k < end) {
            const double* xn = xs + (int)ldm(lcol + k, flags) * WS + c;
            const double xa = xn[0], xb = xn[K], xc = xn[2 * K];
            if (BSR) {
              double v[9];
#pragma unroll
              for (int q = 0; q < 9; ++q) v[q] = (double)ldm(val + bidx(k, q, nbnz, flags), flags);
              y0 += v[0] * xa + v[1] * xb + v[2] * xc;
              y1 += v[3] * xa + v[4] * xb + v[5] * xc;
              y2 += v[6] * xa + v[7] * xb + v[8] * xc;
            } else {
              const double m = (double)ldm(val + k, flags);
              y0 += m * xa;
              y1 += m * xb;
              y2 += m * xc;
            }
          }
        }
      }
    }
    y0 = slots_sum<K, E>(y0);
    y1 = slots_sum<K, E>(y1);
    y2 = slots_sum<K, E>(y2);
#pragma unroll
    for (int q = 0; q < (3 + E - 1) / E; ++q) {
      const int p = e + q * E;
      if (live && p < 3) epi.store(3 * row + p, c, p == 0 ? y0 : (p == 1 ? y1 : y2), pre[q]);
    }
  }
}

// Tiled wide form (scalar mass, K >= 2): the tile's x blocks are staged in
// LDS as in tiled_block_rows (each node once per tile, contiguous 3K-value
// reads), and the rows are formed as in mass3_wide_rows: one entry per lane,
// G lanes per row, the K columns in registers, 16-byte LDS reads.  The x
// gathers then come from LDS instead of L2.  TB = 1024 threads for K <= 4;
// 512 for K >= 8, whose 3K accumulators per lane need more registers.
template <int K, int G, class VT, class Epi, class XT, int TB>
__global__ __launch_bounds__(TB) void tiled_wide_rows(const int64_t* __restrict__ trp,
                                                        const int64_t* __restrict__ tcp,
                                                        const int32_t* __restrict__ tcols,
                                                        const int64_t* __restrict__ rp,
                                                        const uint16_t* __restrict__ lcol,
                                                        const VT* __restrict__ val, const XT* __restrict__ x,
                                                        Epi epi, int flags) {
  extern __shared__ double2 xs2[];  // 16-byte aligned node blocks
  double* xs = reinterpret_cast<double*>(xs2);
  constexpr int W3 = 3 * K;
  static_assert(W3 % 2 == 0, "node blocks of whole double2");
  constexpr int Q = (W3 + G - 1) / G;
  const int64_t tile = xcd_block(flags);
  const int64_t r0 = trp[tile], r1 = trp[tile + 1];
  const int64_t c0 = tcp[tile];
  const int nu = (int)(tcp[tile + 1] - c0);
  for (int t = threadIdx.x; t < nu * W3; t += TB) {
    const int node = t / W3, j = t - node * W3;
    xs[t] = (double)x[(int64_t)tcols[c0 + node] * W3 + j];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int g = lane % G;
  for (int64_t rb = r0 + (int64_t)wave * (64 / G); rb < r1; rb += (int64_t)(TB / 64) * (64 / G)) {
    const int64_t row = rb + lane / G;
    const bool live = row < r1;
    typename Epi::Pre pre[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = g + q * G;
      if (live && j < W3) pre[q] = epi.load(3 * row + j / K, j % K);
    }
    double acc[W3];
#pragma unroll
    for (int j = 0; j < W3; ++j) acc[j] = 0.0;
    if (live) {
      const int64_t end = rp[row + 1];
      for (int64_t k = rp[row] + g; k < end; k += G) {
        const double2* xb = xs2 + (int)ldm(lcol + k, flags) * (W3 / 2);
        const double m = (double)ldm(val + k, flags);
#pragma unroll
        for (int j = 0; j < W3 / 2; ++j) {
          const double2 v = xb[j];
          acc[2 * j] += m * v.x;
          acc[2 * j + 1] += m * v.y;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < W3; ++j) acc[j] = group_sum<G>(acc[j]);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = g + q * G;
      double v = 0.0;
#pragma unroll
      for (int jj = 0; jj < W3; ++jj)
        if (jj == j) v = acc[jj];
      if (live && j < W3) epi.store(3 * row + j / K, j % K, v, pre[q]);
    }
  }
}

template <int K, class VT, class Epi, class XT>
static bool launch_tiled_wide(const DevMat& M, const Tiling& T, const VT* v, const XT* x, Epi epi,
                              hipStream_t s) {
  if constexpr (K < 2) {
    return false;
  } else {
    if (M.layout != L_MASS3 || T.wide <= 0) return false;
    constexpr int TB = K >= 8 ? 512 : 1024;
    const size_t shmem = (size_t)T.max_nodes * 3 * K * sizeof(double);
    switch (T.wide) {
      case 4:
        hipLaunchKernelGGL((tiled_wide_rows<K, 4, VT, Epi, XT, TB>), dim3((unsigned)T.ntiles), dim3(TB), shmem, s,
                           T.trp, T.tcp, T.tcols, M.browptr, T.lcol, v, x, epi, kflags(M));
        break;
      case 8:
        hipLaunchKernelGGL((tiled_wide_rows<K, 8, VT, Epi, XT, TB>), dim3((unsigned)T.ntiles), dim3(TB), shmem, s,
                           T.trp, T.tcp, T.tcols, M.browptr, T.lcol, v, x, epi, kflags(M));
        break;
      default:
        hipLaunchKernelGGL((tiled_wide_rows<K, 16, VT, Epi, XT, TB>), dim3((unsigned)T.ntiles), dim3(TB), shmem, s,
                           T.trp, T.tcp, T.tcols, M.browptr, T.lcol, v, x, epi, kflags(M));
        break;
    }
    return true;
  }
}

template <int K, int R, class VT, class Epi, int TB, class XT>
static void launch_tiled_RB(const DevMat& M, const Tiling& T, const VT* v, const XT* x, Epi epi,
                            hipStream_t s) {
  const size_t shmem = (size_t)T.max_nodes * tile_stride(K) * sizeof(double);
  if (M.layout == L_BSR3)
    hipLaunchKernelGGL((tiled_block_rows<K, R, VT, Epi, true, TB, XT>), dim3((unsigned)T.ntiles), dim3(TB), shmem,
                       s, T.trp, T.tcp, T.tcols, M.browptr, T.lcol, v, M.nbnz, x, epi, kflags(M));
  else
    hipLaunchKernelGGL((tiled_block_rows<K, R, VT, Epi, false, TB, XT>), dim3((unsigned)T.ntiles), dim3(TB), shmem,
                       s, T.trp, T.tcp, T.tcols, M.browptr, T.lcol, v, M.nbnz, x, epi, kflags(M));
}

template <int K, int R, class VT, class Epi, class XT>
static void launch_tiled_R(const DevMat& M, const Tiling& T, const VT* v, const XT* x, Epi epi,
                           hipStream_t s) {
  if (T.block == 256) launch_tiled_RB<K, R, VT, Epi, 256>(M, T, v, x, epi, s);
  else launch_tiled_RB<K, R, VT, Epi, 1024>(M, T, v, x, epi, s);
}

template <int K, class VT, class Epi, class XT>
bool launch_tiled(const DevMat& M, const VT* v, const XT* x, Epi epi, hipStream_t s) {
  // LDS tiles hold 3K values per node: built for K <= 16 only
  if constexpr (K > 16) {
    return false;
  } else {
  int idx = 0;
  while ((1 << idx) != K) ++idx;
  const Tiling* T = M.tiling[idx];
  if (!T || T->ntiles <= 0 || M.layout == L_CSR) return false;
  if (launch_tiled_wide<K>(M, *T, v, x, epi, s)) {
    NMG_HIP(hipGetLastError());
    return true;
  }
  const int R = T->rows_per_wave;
  if (R == 2 && K * 2 <= 64) launch_tiled_R<K, (K * 2 <= 64 ? 2 : 1)>(M, *T, v, x, epi, s);
  else if (R == 4 && K * 4 <= 64) launch_tiled_R<K, (K * 4 <= 64 ? 4 : 1)>(M, *T, v, x, epi, s);
  else if (R == 8 && K * 8 <= 64) launch_tiled_R<K, (K * 8 <= 64 ? 8 : 1)>(M, *T, v, x, epi, s);
  else launch_tiled_R<K, 1>(M, *T, v, x, epi, s);
  NMG_HIP(hipGetLastError());
  return true;
  }
}

}  // namespace nmg

// ==========================================================================
// Adjoint: per element e and eigenpair j (x B-normalised, physical
// coordinates), the derivatives of lambda_j with respect to a uniform
// change of the nodal Lame parameters and density of element e, from the
// element formulas of the elastic operator:
//   d lambda / d lam_e = detJ d^T M d,           d = sum_i D_i u_i  (divergence)
//   d lambda / d mu_e  = 2 detJ sum_ij e_ij^T M e_ij,  e_ij = (D_j u_i + D_i u_j)/2
//   d lambda / d rho_e = -lambda_j detJ sum_i u_i^T M u_i
// u_i: component i of x at the element's nodes, D_i = sum_r invJ(i,r) Dr_r.
// One thread per (element, eigenpair); the reference-element matrices sit in
// LDS.  NP = 10 (P2) or 4 (P1).  With GRAV (JOB=2) the density derivative
// also carries the gravity part of the stiffness, which is linear in the
// nodal density: for a uniform change of rho over the element it adds
//   detJ * sum_ij u_i^T OPrho_ij u_j
//   = detJ * ( div^T M (g.u) - sum_ij gs_ij u_i^T M u_j
//              - sum_a sum_ij g_i[a] (D_j u_i)[a] (M u_j)[a] )
// (OPrho_ij of src/mod_cg_create_matrix.f90:1131-1180; div = sum_i D_i u_i,
// (g.u)[a] = sum_i g_i[a] u_i[a], gs the symmetric least-squares gravity
// gradient of the element; the two D^T G M terms of OPrho give equal forms).
namespace nmg {

template <int NP, bool GRAV>
__global__ __launch_bounds__(kBlock) void adjoint_k(int64_t ne, int nev, const int32_t* __restrict__ elems,
                                                    const double* __restrict__ detJ,
                                                    const double* __restrict__ invJ,
                                                    const double* __restrict__ refM,
                                                    const double* __restrict__ refD,
                                                    const double* __restrict__ X, int64_t ldx,
                                                    const double* __restrict__ d,
                                                    const double* __restrict__ lam,
                                                    const double* __restrict__ gnode,
                                                    const double* __restrict__ gsym,
                                                    double* __restrict__ g_rho, double* __restrict__ g_lam,
                                                    double* __restrict__ g_mu) {
  __shared__ double Ms[NP * NP];
  __shared__ double Ds[3 * NP * NP];
  for (int t = threadIdx.x; t < NP * NP; t += kBlock) Ms[t] = refM[t];
  for (int t = threadIdx.x; t < 3 * NP * NP; t += kBlock) Ds[t] = refD[t];
  __syncthreads();
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= ne * nev) return;
  const int j = (int)(gid / ne);
  const int64_t e = gid - (int64_t)j * ne;
  const double* xj = X + (int64_t)j * ldx;
  double u[3][NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) {
    const int64_t v = elems[e * NP + a];
#pragma unroll
    for (int i = 0; i < 3; ++i) u[i][a] = xj[3 * v + i] * d[3 * v + i];
  }
  double J[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) J[t] = invJ[e * 9 + t];
  const double dJ = detJ[e];
  // grad[i][k][a] = (D_k u_i)_a
  double gr[3][3][NP];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double ru[3][NP];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int a = 0; a < NP; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) s += Ds[r * NP * NP + a * NP + b] * u[i][b];
        ru[r][a] = s;
      }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int a = 0; a < NP; ++a) gr[i][k][a] = J[k * 3 + 0] * ru[0][a] + J[k * 3 + 1] * ru[1][a] + J[k * 3 + 2] * ru[2][a];
  }
  auto quad = [&](const double* p, const double* q) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < NP; ++b) t += Ms[a * NP + b] * q[b];
      s += p[a] * t;
    }
    return s;
  };
  double dv[NP];
#pragma unroll
  for (int a = 0; a < NP; ++a) dv[a] = gr[0][0][a] + gr[1][1][a] + gr[2][2][a];
  const double glam = dJ * quad(dv, dv);
  double gmu = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double ek[NP];
#pragma unroll
      for (int a = 0; a < NP; ++a) ek[a] = 0.5 * (gr[i][k][a] + gr[k][i][a]);
      gmu += quad(ek, ek);
    }
  gmu *= 2.0 * dJ;
  double grho = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) grho += quad(u[i], u[i]);
  grho *= -lam[j] * dJ;
  if constexpr (GRAV) {
    double Mu[3][NP], gu[NP];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int a = 0; a < NP; ++a) {
        double t = 0.0;
#pragma unroll
        for (int b = 0; b < NP; ++b) t += Ms[a * NP + b] * u[i][b];
        Mu[i][a] = t;
      }
    double t4 = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
      const double g0 = gnode[(e * NP + a) * 3 + 0], g1 = gnode[(e * NP + a) * 3 + 1],
                   g2 = gnode[(e * NP + a) * 3 + 2];
      gu[a] = g0 * u[0][a] + g1 * u[1][a] + g2 * u[2][a];
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        t4 += (g0 * gr[0][jj][a] + g1 * gr[1][jj][a] + g2 * gr[2][jj][a]) * Mu[jj][a];
    }
    const double t1 = quad(dv, gu);
    double t3 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NP; ++a) s += u[i][a] * Mu[jj][a];
        t3 += gsym[e * 9 + i * 3 + jj] * s;
      }
    grho += dJ * (t1 - t3 - t4);
  }
  const int64_t o = (int64_t)j * ne + e;
  g_rho[o] = grho;
  g_lam[o] = glam;
  g_mu[o] = gmu;
}

void k_adjoint(int np, int64_t ne, int nev, const int32_t* elems, const double* detJ, const double* invJ,
               const double* refM, const double* refD, const double* X, int64_t ldx, const double* d,
               const double* lam, const double* gnode, const double* gsym, double* g_rho, double* g_lam,
               double* g_mu, hipStream_t s) {
  const int64_t total = ne * (int64_t)nev;
  const int64_t grid = (total + kBlock - 1) / kBlock;
  if (grid <= 0) return;
  const bool grav = gnode != nullptr;
  if (grav && !gsym) throw Error{NMG_EINVAL, "gravity without its element gradients"};
#define NMG_ADJ(NP, G)                                                                                          \
  hipLaunchKernelGGL((adjoint_k<NP, G>), dim3((unsigned)grid), dim3(kBlock), 0, s, ne, nev, elems, detJ, invJ, \
                     refM, refD, X, ldx, d, lam, gnode, gsym, g_rho, g_lam, g_mu)
  if (np == 10 && grav)
    NMG_ADJ(10, true);
  else if (np == 10)
    NMG_ADJ(10, false);
  else if (np == 4 && grav)
    NMG_ADJ(4, true);
  else if (np == 4)
    NMG_ADJ(4, false);
  else
    throw Error{NMG_EINVAL, "adjoint supports P1 (4 nodes) and P2 (10 nodes) elements"};
#undef NMG_ADJ
  NMG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Persistent Chebyshev B-solve (scalar mass x I3 layout, K interleaved
// columns): all `deg` steps of the semi-iteration in one cooperative launch.
// Each workgroup owns a contiguous, nnz-balanced range of node rows; it
// copies its rows' column indices and values into LDS once, then every step
// only gathers x_k (3K contiguous doubles per neighbour node) and updates
// its own rows of x_{k+1}.  Steps are separated by a grid barrier
// (cooperative_groups grid sync: device-scope fences, so x_{k+1} written by
// one XCD is what every XCD reads in the next step).  Every workgroup runs
// the same number of barriers; there is no early exit.
// Lanes: L = max(K, 8) per node row; lane = slot * K + column, the S = L/K
// slots split the row's entries and are summed by shuffles.
constexpr int kPBlock = 512;  // threads per workgroup

template <int K, class VT>
__global__ __launch_bounds__(kPBlock) void cheb_persistent_k(const int64_t* __restrict__ rp,
                                                             const int32_t* __restrict__ ci,
                                                             const VT* __restrict__ va,
                                                             const int64_t* __restrict__ part, int rows_cap,
                                                             int ent_cap, const double* __restrict__ b,
                                                             double* xa, double* xb, int deg, double ith,
                                                             double sg) {
  constexpr int L = K > 8 ? K : 8;
  constexpr int S = L / K;
  constexpr int RPR = kPBlock / L;  // rows per round
  extern __shared__ __align__(16) unsigned char smem[];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int64_t r0 = part[blockIdx.x], r1 = part[blockIdx.x + 1];
  const int64_t e0 = rp[r0];
  const int nr = (int)(r1 - r0);
  VT* lv = (VT*)smem;                                             // ent_cap values
  int32_t* lc = (int32_t*)(smem + sizeof(VT) * (size_t)ent_cap);  // ent_cap columns
  int32_t* lr = lc + ent_cap;                                     // rows_cap + 1 offsets
  const int ne = (int)(rp[r1] - e0);
  for (int t = threadIdx.x; t < ne; t += kPBlock) {
    lv[t] = va[e0 + t];
    lc[t] = ci[e0 + t];
  }
  for (int t = threadIdx.x; t <= nr; t += kPBlock) lr[t] = (int)(rp[r0 + t] - e0);
  // x_1 = b / th (in xa), x_0 = 0 (in xb), own rows
  for (int64_t i = 3 * K * r0 + threadIdx.x; i < 3 * K * r1; i += kPBlock) {
    xa[i] = b[i] * ith;
    xb[i] = 0.0;
  }
  __syncthreads();
  grid.sync();
  __threadfence();
  const double* xk = xa;
  double* xm = xb;
  const int lane = threadIdx.x % L;
  const int c = lane % K, slot = lane / K;
  double w = 2.0;
  for (int k = 1; k <= deg; ++k) {
    w = 1.0 / (1.0 - w / (4.0 * sg * sg));
    for (int rb = 0; rb < nr; rb += RPR) {
      const int r = rb + threadIdx.x / L;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
      if (r < nr) {
        const int end = lr[r + 1];
        for (int j = lr[r] + slot; j < end; j += S) {
          const double v = (double)lv[j];
          const double* xp = xk + (int64_t)3 * K * lc[j] + c;
          s0 += v * xp[0];
          s1 += v * xp[K];
          s2 += v * xp[2 * K];
        }
      }
#pragma unroll
      for (int off = L / 2; off >= K; off >>= 1) {
        s0 += __shfl_xor(s0, off, L);
        s1 += __shfl_xor(s1, off, L);
        s2 += __shfl_xor(s2, off, L);
      }
      if (r < nr && slot == 0) {
        const int64_t i0 = (int64_t)3 * K * (r0 + r) + c;
        const double sv[3] = {s0, s1, s2};
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const int64_t i = i0 + p * K;
          const double xmi = xm[i], xki = xk[i];
          xm[i] = xmi + w * (xki - xmi + (b[i] - sv[p]) * ith);
        }
      }
    }
    __syncthreads();
    grid.sync();
    __threadfence();
    double* t = const_cast<double*>(xk);
    xk = xm;
    xm = t;
  }
}

template <class VT>
static void* persistent_fn(int K) {
  switch (K) {
    case 1: return reinterpret_cast<void*>(&cheb_persistent_k<1, VT>);
    case 2: return reinterpret_cast<void*>(&cheb_persistent_k<2, VT>);
    case 4: return reinterpret_cast<void*>(&cheb_persistent_k<4, VT>);
    case 8: return reinterpret_cast<void*>(&cheb_persistent_k<8, VT>);
    case 16: return reinterpret_cast<void*>(&cheb_persistent_k<16, VT>);
    default: throw Error{NMG_EINVAL, "block width must be 1, 2, 4, 8 or 16"};
  }
}

void plan_cheb_persistent(const DevMat& M, bool fp32, PersistPlan& P) {
  P.release();
  if (M.layout != L_MASS3 || M.h_browptr.empty() || (fp32 && !M.bval32)) return;
  int dev = 0, ncu = 0;
  NMG_HIP(hipGetDevice(&dev));
  NMG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int coop = 0;
  NMG_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop || ncu <= 0) return;
  const int64_t nb = M.nb;
  const std::vector<int64_t>& rp = M.h_browptr;
  // one workgroup per CU, fewer if the rows are few; nnz-balanced ranges
  int G = (int)std::min<int64_t>(ncu, std::max<int64_t>(1, nb / 8));
  std::vector<int64_t> part(G + 1, 0);
  const double tot = (double)rp[nb] + 4.0 * (double)nb;  // entries plus a per-row cost
  int64_t r = 0;
  for (int g = 1; g < G; ++g) {
    const double target = tot * g / G;
    while (r < nb && (double)rp[r] + 4.0 * (double)r < target) ++r;
    part[g] = r;
  }
  part[G] = nb;
  int64_t max_rows = 0, max_ent = 0;
  for (int g = 0; g < G; ++g) {
    max_rows = std::max(max_rows, part[g + 1] - part[g]);
    max_ent = std::max(max_ent, rp[part[g + 1]] - rp[part[g]]);
  }
  const size_t vsz = fp32 ? sizeof(float) : sizeof(double);
  const size_t bytes = vsz * (size_t)max_ent + 4 * (size_t)max_ent + 4 * (size_t)(max_rows + 1) + 16;
  if (bytes > kPersistLdsMax) return;
  int occ = 0;
  for (int K : {1, 2, 4, 8, 16}) {  // every width must fit the same grid
    NMG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fp32 ? persistent_fn<float>(K) : persistent_fn<double>(K),
                                                         kPBlock, bytes));
    if (occ < 1 || (int64_t)occ * ncu < G) return;
  }
  P.part = nullptr;
  NMG_HIP(hipMalloc(&P.part, sizeof(int64_t) * (G + 1)));
  NMG_HIP(hipMemcpy(P.part, part.data(), sizeof(int64_t) * (G + 1), hipMemcpyHostToDevice));
  P.grid = G;
  P.rows_cap = (int)max_rows;
  P.ent_cap = (int)max_ent;
  P.lds_bytes = bytes;
  P.fp32 = fp32;
  P.ok = true;
}

void k_cheb_persistent(int K, const DevMat& M, const PersistPlan& P, const double* b, double* out, double* tmp,
                       int deg, double lmin, double lmax, hipStream_t s) {
  NMG_CHECK(P.ok && deg >= 1, NMG_ESTATE, "persistent Chebyshev solve not planned");
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin);
  double ith = 1.0 / th, sg = th / de;
  // the result x_{deg+1} ends in xa for even deg, xb for odd
  double* xa = (deg % 2 == 0) ? out : tmp;
  double* xb = (deg % 2 == 0) ? tmp : out;
  const int64_t* rp = M.browptr;
  const int32_t* ci = M.bcol;
  const int64_t* part = P.part;
  int rows_cap = P.rows_cap, ent_cap = P.ent_cap;
  void* args[] = {(void*)&rp, (void*)&ci, nullptr, (void*)&part, (void*)&rows_cap, (void*)&ent_cap,
                  (void*)&b, (void*)&xa, (void*)&xb, (void*)&deg, (void*)&ith, (void*)&sg};
  if (P.fp32) {
    const float* va = M.bval32;
    args[2] = (void*)&va;
    NMG_HIP(hipLaunchCooperativeKernel(persistent_fn<float>(K), dim3(P.grid), dim3(kPBlock), args, P.lds_bytes, s));
  } else {
    const double* va = M.bval;
    args[2] = (void*)&va;
    NMG_HIP(hipLaunchCooperativeKernel(persistent_fn<double>(K), dim3(P.grid), dim3(kPBlock), args, P.lds_bytes, s));
  }
}

// BSR3 values between the plane order and the 64-entry tile order
__global__ void bsr3_reorder_k(int64_t nbnz, int64_t ntot, const double* __restrict__ in, double* __restrict__ out,
                               int to_tiles) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < ntot; t += (int64_t)gridDim.x * kBlock) {
    const int64_t k = (t / 576) * 64 + (t % 64);
    const int q = (int)((t % 576) / 64);
    if (to_tiles) out[t] = k < nbnz ? in[(int64_t)q * nbnz + k] : 0.0;
    else if (k < nbnz) out[(int64_t)q * nbnz + k] = in[t];
  }
}

void set_value_order(DevMat& M, int tiled, hipStream_t s) {
  NMG_CHECK(M.layout == L_BSR3, NMG_EINVAL, "the value order applies to 3x3-block matrices");
  tiled = tiled ? 1 : 0;
  if (tiled == M.tiled_vals) return;
  M.touch();
  const int64_t ntiles = 9 * ((M.nbnz + 63) / 64 * 64);
  const int64_t nplanes = 9 * M.nbnz;
  double* out = nullptr;
  NMG_HIP(hipMalloc(&out, sizeof(double) * (tiled ? ntiles : nplanes)));
  hipLaunchKernelGGL(bsr3_reorder_k, dim3(grid_for(ntiles)), dim3(kBlock), 0, s, M.nbnz, ntiles, M.bval, out, tiled);
  NMG_HIP(hipGetLastError());
  NMG_HIP(hipStreamSynchronize(s));
  NMG_HIP(hipFree(M.bval));
  M.bval = out;
  M.tiled_vals = tiled;
  if (M.bval32) {  // rebuild the FP32 copy in the new order and size
    NMG_HIP(hipFree(M.bval32));
    M.bval32 = nullptr;
    make_fp32(M, s);
  }
}

}  // namespace nmg
