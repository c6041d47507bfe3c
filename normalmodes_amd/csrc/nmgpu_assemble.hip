// Device producer of the solid-region matrices A (stiffness, with the
// gravity terms for JOB=2) and B (mass) of the isotropic elastic
// continuous-Galerkin element, CGE3D_ISO (src/mod_cg_create_matrix.f90:980-1266),
// on the CSR pattern of matrixstruct (:1269-1455).
//
// Two kernels:
//  - elem_k<NP>: one wavefront per element forms the 3NP x 3NP element
//    stiffness and the NP x NP scalar mass in LDS and writes them to an
//    element-major buffer.  Element formulas (invJ = 2 B^{-1}, detJ = det B/8,
//    D_k = sum_r invJ(k,r) Dr_r, ML/MU/M_e as symmetrised nodal products,
//    K_ii = detJ sym(sum_k D_k^T MU D_k + D_i^T MU D_i + D_i^T ML D_i),
//    K_ij = detJ (D_i^T ML D_j + D_j^T MU D_i) evaluated as the average of
//    its two equal forms, and with gravity detJ sym_rho(OPrho_ij) added and
//    the element matrix made exactly symmetric) are those of the CPU
//    producer, term by term, so the two agree to rounding.
//  - gather_k: one thread per DoF row sums the element contributions of its
//    node in ascending element order (the CPU producer's order), so the
//    result is deterministic and A is exactly symmetric.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "nmgpu_fs_elem.h"
#include "nmgpu_internal.h"

namespace nmg {

namespace {

constexpr int kWaveElems = 1;  // elements per 64-thread workgroup

template <int NP>
struct ElemLds {
  double D[3][NP * NP];
  double ML[NP * NP], MU[NP * NP], M[NP * NP];
  double P[3][NP * NP], Q[3][NP * NP], Ot[3][NP * NP];
  double K[9 * NP * NP];
  double geo[12];  // detJ, invJ[9], (unused)
  double dg[9];
  double G[3][NP];
  double lam[NP], mu[NP], rho[NP];
};

__device__ inline bool inv3_dev(const double B[9], double Bi[9], double* det) {
  const double d = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                   B[2] * (B[3] * B[7] - B[4] * B[6]);
  *det = d;
  if (d == 0.0) return false;
  const double id = 1.0 / d;
  Bi[0] = (B[4] * B[8] - B[5] * B[7]) * id;
  Bi[1] = (B[2] * B[7] - B[1] * B[8]) * id;
  Bi[2] = (B[1] * B[5] - B[2] * B[4]) * id;
  Bi[3] = (B[5] * B[6] - B[3] * B[8]) * id;
  Bi[4] = (B[0] * B[8] - B[2] * B[6]) * id;
  Bi[5] = (B[2] * B[3] - B[0] * B[5]) * id;
  Bi[6] = (B[3] * B[7] - B[4] * B[6]) * id;
  Bi[7] = (B[1] * B[6] - B[0] * B[7]) * id;
  Bi[8] = (B[0] * B[4] - B[1] * B[3]) * id;
  return true;
}

// sum_c A[c][a] B[c][b] (A^T B entry) and sum_c A[a][c] B[c][b] (A B entry)
template <int NP>
__device__ inline double tn(const double* A, const double* B, int a, int b) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < NP; ++c) s += A[c * NP + a] * B[c * NP + b];
  return s;
}
template <int NP>
__device__ inline double nn(const double* A, const double* B, int a, int b) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < NP; ++c) s += A[a * NP + c] * B[c * NP + b];
  return s;
}

template <int NP>
__global__ __launch_bounds__(64) void elem_k(int64_t ne, const double* __restrict__ nodes,
                                             const int32_t* __restrict__ elems, const int32_t* __restrict__ vslot,
                                             const double* __restrict__ lam, const double* __restrict__ mu,
                                             const double* __restrict__ rho, const double* __restrict__ g,
                                             const double* __restrict__ refM, const double* __restrict__ refD,
                                             double* __restrict__ Ke, double* __restrict__ Me, int* __restrict__ bad) {
  constexpr int N2 = NP * NP, L = 3 * NP;
  __shared__ ElemLds<NP> S;
  const int lane = threadIdx.x;
  const int64_t e = blockIdx.x;
  if (e >= ne) return;  // uniform per workgroup (one element each)
  const int32_t* el = elems + e * NP;
  // geometry: every lane forms the 3x3 edge matrix and its inverse
  double X[12];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) X[k * 3 + c] = nodes[(int64_t)el[vslot[k]] * 3 + c];
  double B[9], Bi[9], detB;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) B[r * 3 + c] = X[(r + 1) * 3 + c] - X[c];
  const bool ok = inv3_dev(B, Bi, &detB);
  if (!ok) {
    if (lane == 0) atomicMin(bad, (int)e);
    return;
  }
  const double detJ = detB / 8.0;
  double invJ[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) invJ[k] = 2.0 * Bi[k];
  for (int a = lane; a < NP; a += 64) {
    S.lam[a] = lam[e * NP + a];
    S.mu[a] = mu[e * NP + a];
    S.rho[a] = rho[e * NP + a];
  }
  const bool grav = g != nullptr;
  if (grav)
    for (int t = lane; t < 3 * NP; t += 64) S.G[t % 3][t / 3] = g[(e * NP + t / 3) * 3 + t % 3];
  for (int t = lane; t < 3 * N2; t += 64) {
    const int k = t / N2, a = t % N2;
    S.D[k][a] = invJ[k * 3 + 0] * refD[0 * N2 + a] + invJ[k * 3 + 1] * refD[1 * N2 + a] +
                invJ[k * 3 + 2] * refD[2 * N2 + a];
  }
  __syncthreads();
  for (int t = lane; t < N2; t += 64) {
    const int a = t / NP, b = t % NP;
    const double m = refM[t];
    S.M[t] = m;
    S.ML[t] = (m * S.lam[b] + S.lam[a] * m) * 0.5;
    S.MU[t] = (m * S.mu[b] + S.mu[a] * m) * 0.5;
    Me[e * N2 + t] = detJ * ((S.rho[a] * m + m * S.rho[b]) * 0.5);
  }
  if (grav && lane == 0) {
    // least-squares gradient of the nodal gravity over the element
    // (src/mod_cg_create_matrix.f90:1054-1080)
    double xm[3] = {0, 0, 0}, gm[3] = {0, 0, 0};
    for (int a = 0; a < NP; ++a)
      for (int c = 0; c < 3; ++c) {
        xm[c] += nodes[(int64_t)el[a] * 3 + c];
        gm[c] += S.G[c][a];
      }
    for (int c = 0; c < 3; ++c) {
      xm[c] /= NP;
      gm[c] /= NP;
    }
    double Nm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < NP; ++a)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const double xr = nodes[(int64_t)el[a] * 3 + r] - xm[r];
          Nm[r * 3 + c] += xr * (nodes[(int64_t)el[a] * 3 + c] - xm[c]);
          R[r * 3 + c] += xr * (S.G[c][a] - gm[c]);
        }
    double Ni[9], dN;
    inv3_dev(Nm, Ni, &dN);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += Ni[r * 3 + k] * R[k * 3 + c];
        S.dg[r * 3 + c] = s;
      }
  }
  __syncthreads();
  for (int t = lane; t < 3 * N2; t += 64) {
    const int k = t / N2, a = (t % N2) / NP, b = t % NP;
    S.P[k][a * NP + b] = nn<NP>(S.ML, S.D[k], a, b);  // ML D_k
    S.Q[k][a * NP + b] = nn<NP>(S.MU, S.D[k], a, b);  // MU D_k
  }
  __syncthreads();
  for (int t = lane; t < 3 * N2; t += 64) {
    const int k = t / N2, a = (t % N2) / NP, b = t % NP;
    S.Ot[k][a * NP + b] = tn<NP>(S.D[k], S.Q[k], a, b);  // D_k^T MU D_k
  }
  __syncthreads();
  // blocks (i, j), entry (a, b)
  for (int t = lane; t < 9 * N2; t += 64) {
    const int blk = t / N2, i = blk / 3, j = blk % 3, a = (t % N2) / NP, b = t % NP;
    double op;
    if (i == j) {
      const double os_ab = S.Ot[0][a * NP + b] + S.Ot[1][a * NP + b] + S.Ot[2][a * NP + b];
      const double os_ba = S.Ot[0][b * NP + a] + S.Ot[1][b * NP + a] + S.Ot[2][b * NP + a];
      const double t2_ab = os_ab + S.Ot[i][a * NP + b] + tn<NP>(S.D[i], S.P[i], a, b);
      const double t2_ba = os_ba + S.Ot[i][b * NP + a] + tn<NP>(S.D[i], S.P[i], b, a);
      op = (t2_ab + t2_ba) * 0.5;
    } else {
      const double a1 = tn<NP>(S.D[i], S.P[j], a, b);   // D_i^T ML D_j
      const double a2 = tn<NP>(S.D[j], S.Q[i], a, b);   // D_j^T MU D_i
      const double a3 = tn<NP>(S.D[j], S.P[i], b, a);   // (D_j^T ML D_i)(b, a)
      const double a4 = tn<NP>(S.D[i], S.Q[j], b, a);   // (D_i^T MU D_j)(b, a)
      op = ((a3 + a4) + (a1 + a2)) * 0.5;
    }
    if (grav) {
      // OPrho_ij(a, b) = (D_i^T M G_j + G_i M D_j)/2 - M (dg_ij + dg_ji)/2
      //                  - (D_j^T G_i M + M G_j D_i)/2, then sym_rho
      auto oprho = [&](int aa, int bb) {
        double r1 = 0.0, r2 = 0.0, r3 = 0.0, r4 = 0.0;
#pragma unroll
        for (int c = 0; c < NP; ++c) {
          r1 += S.D[i][c * NP + aa] * (S.M[c * NP + bb] * S.G[j][bb]);        // D_i^T (M G_j)
          r2 += S.M[aa * NP + c] * S.D[j][c * NP + bb];                          // (M D_j), times G_i(aa)
          r3 += S.D[j][c * NP + aa] * (S.G[i][c] * S.M[c * NP + bb]);          // D_j^T (G_i M)
          r4 += S.M[aa * NP + c] * (S.G[j][c] * S.D[i][c * NP + bb]);          // M (G_j D_i)
        }
        r2 *= S.G[i][aa];
        const double sg = (S.dg[i * 3 + j] + S.dg[j * 3 + i]) * 0.5;
        return (r1 + r2) * 0.5 - S.M[aa * NP + bb] * sg - (r3 + r4) * 0.5;
      };
      const double o = oprho(a, b);
      op += (o * S.rho[b] + S.rho[a] * o) * 0.5;
    }
    S.K[(i * NP + a) * L + (j * NP + b)] = detJ * op;
  }
  __syncthreads();
  double* Kout = Ke + e * (int64_t)L * L;
  for (int t = lane; t < L * L; t += 64) {
    const int r = t / L, c = t % L;
    double v = S.K[t];
    if (grav && r != c) v = r < c ? 0.5 * (S.K[r * L + c] + S.K[c * L + r]) : 0.5 * (S.K[c * L + r] + S.K[r * L + c]);
    Kout[t] = v;
  }
}

// row r = 3 v + p: zero it, then add the contributions of the elements of
// node v in ascending element order
__global__ void gather_k(int64_t n_nodes, int np, const int32_t* __restrict__ elems,
                         const int64_t* __restrict__ n2e_ptr, const int64_t* __restrict__ n2e,
                         const int32_t* __restrict__ n2s, const int64_t* __restrict__ adj_ptr,
                         const int32_t* __restrict__ adj, const int64_t* __restrict__ a_rowptr,
                         const int64_t* __restrict__ b_rowptr, const double* __restrict__ Ke,
                         const double* __restrict__ Me, double* __restrict__ a_val, double* __restrict__ b_val) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= 3 * n_nodes) return;
  const int64_t v = row / 3;
  const int p = (int)(row % 3);
  const int64_t a0 = adj_ptr[v], deg = adj_ptr[v + 1] - a0;
  const int64_t ra = a_rowptr[row], rb = b_rowptr[row];
  for (int64_t k = 0; k < 3 * deg; ++k) a_val[ra + k] = 0.0;
  for (int64_t k = 0; k < deg; ++k) b_val[rb + k] = 0.0;
  const int L = 3 * np;
  for (int64_t t = n2e_ptr[v]; t < n2e_ptr[v + 1]; ++t) {
    const int64_t e = n2e[t];
    const int m = n2s[t];
    const double* K = Ke + e * (int64_t)L * L;
    const double* Mm = Me + e * (int64_t)np * np;
    for (int nn2 = 0; nn2 < np; ++nn2) {
      const int32_t w = elems[e * np + nn2];
      int64_t lo = 0, hi = deg;  // lower_bound of w in the sorted adjacency of v
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (adj[a0 + mid] < w) lo = mid + 1;
        else hi = mid;
      }
      const int64_t k = lo;
      for (int q = 0; q < 3; ++q) a_val[ra + 3 * k + q] += K[(p * np + m) * L + (q * np + nn2)];
      b_val[rb + k] += Mm[m * np + nn2];
    }
  }
}

// fluid-solid body: per-element data (one thread per element); codes of
// the first flat element and the first fluid element on a free surface
// without gravity
__global__ void fs_info_k(fs::Body M, fs::Info* __restrict__ info, int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M.n_elem) return;
  const int rc = fs::element_info(M, e, &info[e]);
  if (rc) atomicMin(&bad[rc - 1], (int)e);
}

// fluid element matrices, one thread per fluid element of the list
__global__ void fs_fluid_k(fs::Body M, int64_t nf, const int64_t* __restrict__ flist,
                           const int64_t* __restrict__ koff, const int64_t* __restrict__ scm, double* W,
                           double* SC) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int64_t e = flist[i];
  fs::fluid_element(M, e, W + koff[e], scm[e] >= 0 ? SC + scm[e] : nullptr);
}

// rows of the fluid-solid pattern, one thread per node
__global__ void fs_gather_k(fs::Body M, fs::Gather G, unsigned long long* __restrict__ miss) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= M.n_nodes) return;
  const int64_t m = fs::gather_node(M, G, v);
  if (m) atomicAdd(miss, (unsigned long long)m);
}

// derivative fluid element matrices, one thread per (fluid element, parameter):
// parameter 0 seeds drho = 1 at the element's nodes, parameter 1 dlam = 1
__global__ void fs_dfluid_k(fs::Body M, int64_t nf, const int64_t* __restrict__ flist,
                            const int64_t* __restrict__ koff, const int64_t* __restrict__ scm, fs::Dual* Fr,
                            fs::Dual* Fl, fs::Dual* Sr, fs::Dual* Sl) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nf) return;
  const int64_t e = flist[t >> 1];
  const int q = (int)(t & 1);
  fs::Dual lD[fs::kMaxNp], rD[fs::kMaxNp];
  for (int a = 0; a < M.np; ++a) {
    lD[a] = {M.lam[e * M.np + a], q == 1 ? 1.0 : 0.0};
    rD[a] = {M.rho[e * M.np + a], q == 0 ? 1.0 : 0.0};
  }
  fs::fluid_element_t<fs::Dual>(M, e, lD, rD, (q == 0 ? Fr : Fl) + koff[e],
                                scm[e] >= 0 ? (q == 0 ? Sr : Sl) + scm[e] : nullptr);
}

// eigenvalue sensitivities, one thread per (element, pair)
__global__ void fs_adjoint_k(fs::Body M, fs::Adj A) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M.n_elem * A.k) return;
  fs::element_adjoint(M, A, t % M.n_elem, (int)(t / M.n_elem));
}

template <typename T>
struct Dev {
  T* p = nullptr;
  explicit Dev(size_t n) {
    if (n) NMG_HIP(hipMalloc(&p, n * sizeof(T)));
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void up(const T* h, size_t n, hipStream_t s) {
    if (n) NMG_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

}  // namespace

int assemble_elastic_device(int64_t n_nodes, int64_t n_elem, int np, const double* nodes, const int32_t* elems,
                            const int32_t* vslot, const double* lam, const double* mu, const double* rho,
                            const double* g, const double* refM, const double* refD, const int64_t* adj_ptr,
                            const int32_t* adj, const int64_t* a_rowptr, double* a_val, const int64_t* b_rowptr,
                            double* b_val, hipStream_t s) {
  NMG_CHECK(np == 4 || np == 10, NMG_EINVAL, "pNp must be 4 (P1) or 10 (P2)");
  NMG_CHECK(n_nodes > 0 && n_elem > 0, NMG_EINVAL, "empty mesh");
  for (int k = 0; k < 4; ++k) NMG_CHECK(vslot[k] >= 0 && vslot[k] < np, NMG_EINVAL, "bad vertex slot");
  for (int64_t i = 0; i < n_elem * np; ++i)
    NMG_CHECK(elems[i] >= 0 && elems[i] < n_nodes, NMG_EINVAL, "element node id out of range");
  const int L = 3 * np;
  // node -> (element, slot) in ascending element order
  std::vector<int64_t> n2e_ptr(n_nodes + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int a = 0; a < np; ++a) n2e_ptr[elems[e * np + a] + 1]++;
  for (int64_t v = 0; v < n_nodes; ++v) n2e_ptr[v + 1] += n2e_ptr[v];
  std::vector<int64_t> n2e(n2e_ptr[n_nodes]);
  std::vector<int32_t> n2s(n2e_ptr[n_nodes]);
  {
    std::vector<int64_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e)
      for (int a = 0; a < np; ++a) {
        const int64_t v = elems[e * np + a];
        n2e[fill[v]] = e;
        n2s[fill[v]] = a;
        fill[v]++;
      }
  }
  const int64_t nadj = adj_ptr[n_nodes];
  NMG_CHECK(a_rowptr[3 * n_nodes] == 9 * nadj && b_rowptr[3 * n_nodes] == 3 * nadj, NMG_EINVAL,
            "CSR row offsets do not match the node adjacency");
  Dev<double> dn(n_nodes * 3), dl(n_elem * np), dm(n_elem * np), dr(n_elem * np), dM(np * np), dD(3 * np * np);
  Dev<double> dg(g ? n_elem * np * 3 : 0);
  Dev<int32_t> de(n_elem * np), dvs(4), dadj(nadj), dn2s(n2s.size());
  Dev<int64_t> dap(n_nodes + 1), dn2ep(n_nodes + 1), dn2e(n2e.size()), darp(3 * n_nodes + 1), dbrp(3 * n_nodes + 1);
  Dev<double> dKe((size_t)n_elem * L * L), dMe((size_t)n_elem * np * np);
  Dev<double> dav(9 * nadj), dbv(3 * nadj);
  Dev<int> dbad(1);
  dn.up(nodes, n_nodes * 3, s);
  dl.up(lam, n_elem * np, s);
  dm.up(mu, n_elem * np, s);
  dr.up(rho, n_elem * np, s);
  dM.up(refM, np * np, s);
  dD.up(refD, 3 * np * np, s);
  if (g) dg.up(g, n_elem * np * 3, s);
  de.up(elems, n_elem * np, s);
  dvs.up(vslot, 4, s);
  dadj.up(adj, nadj, s);
  dn2s.up(n2s.data(), n2s.size(), s);
  dap.up(adj_ptr, n_nodes + 1, s);
  dn2ep.up(n2e_ptr.data(), n_nodes + 1, s);
  dn2e.up(n2e.data(), n2e.size(), s);
  darp.up(a_rowptr, 3 * n_nodes + 1, s);
  dbrp.up(b_rowptr, 3 * n_nodes + 1, s);
  const int big = 0x7fffffff;
  dbad.up(&big, 1, s);
  if (np == 10)
    hipLaunchKernelGGL(elem_k<10>, dim3((unsigned)n_elem), dim3(64), 0, s, n_elem, dn.p, de.p, dvs.p, dl.p, dm.p,
                       dr.p, g ? dg.p : nullptr, dM.p, dD.p, dKe.p, dMe.p, dbad.p);
  else
    hipLaunchKernelGGL(elem_k<4>, dim3((unsigned)n_elem), dim3(64), 0, s, n_elem, dn.p, de.p, dvs.p, dl.p, dm.p,
                       dr.p, g ? dg.p : nullptr, dM.p, dD.p, dKe.p, dMe.p, dbad.p);
  NMG_HIP(hipGetLastError());
  int badh = big;
  NMG_HIP(hipMemcpyAsync(&badh, dbad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  NMG_HIP(hipStreamSynchronize(s));
  if (badh != big) return -(1 + badh);
  const int64_t rows = 3 * n_nodes;
  hipLaunchKernelGGL(gather_k, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, n_nodes, np, de.p, dn2ep.p,
                     dn2e.p, dn2s.p, dap.p, dadj.p, darp.p, dbrp.p, dKe.p, dMe.p, dav.p, dbv.p);
  NMG_HIP(hipGetLastError());
  NMG_HIP(hipMemcpyAsync(a_val, dav.p, sizeof(double) * 9 * nadj, hipMemcpyDeviceToHost, s));
  NMG_HIP(hipMemcpyAsync(b_val, dbv.p, sizeof(double) * 3 * nadj, hipMemcpyDeviceToHost, s));
  NMG_HIP(hipStreamSynchronize(s));
  return 0;
}

void assemble_fs_device(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                        const int32_t* vslot, const double* lam, const double* mu, const double* rho, const double* g,
                        const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM,
                        const double* refD, const double* massF, const int32_t* fmask, const int64_t* adj_ptr,
                        const int32_t* adj, const int8_t* vstat, const int64_t* vstt, const int64_t* pstt,
                        const int64_t* ad_ptr, int32_t* ad_col, double* ad_val, const int64_t* b_ptr, int32_t* b_col,
                        double* b_val, const int64_t* e_ptr, int32_t* e_col, double* e_val, const int64_t* et_ptr,
                        int32_t* et_col, double* et_val, const int64_t* ap_ptr, int32_t* ap_col, double* ap_val,
                        int64_t* stats, hipStream_t s) {
  using fs::hasp;
  using fs::vnum;
  NMG_CHECK(np == 4 || np == 10, NMG_EINVAL, "pNp must be 4 (P1) or 10 (P2)");
  NMG_CHECK(nfp == (np == 4 ? 3 : 6), NMG_EINVAL, "face nodes must be 3 (P1) or 6 (P2)");
  NMG_CHECK(n_nodes > 0 && n_elem > 0 && n_elem < 0x3fffffff, NMG_EINVAL, "empty or oversized mesh");
  for (int k = 0; k < 4; ++k) NMG_CHECK(vslot[k] >= 0 && vslot[k] < np, NMG_EINVAL, "bad vertex slot");
  for (int k = 0; k < 4 * nfp; ++k) NMG_CHECK(fmask[k] >= 0 && fmask[k] < np, NMG_EINVAL, "bad face mask");
  for (int64_t i = 0; i < n_elem * np; ++i)
    NMG_CHECK(elems[i] >= 0 && elems[i] < n_nodes, NMG_EINVAL, "element node id out of range");
  for (int64_t i = 0; i < n_elem * 4; ++i)
    NMG_CHECK(neigh[i] >= -1 && neigh[i] < n_elem, NMG_EINVAL, "face neighbour out of range");
  for (int64_t e = 0; e < n_elem; ++e) NMG_CHECK(efluid[e] == 0 || efluid[e] == 1, NMG_EINVAL, "bad fluid flag");
  NMG_CHECK(adj_ptr[0] == 0, NMG_EINVAL, "adjacency offsets must start at 0");
  for (int64_t v = 0; v < n_nodes; ++v) {
    NMG_CHECK(adj_ptr[v + 1] >= adj_ptr[v], NMG_EINVAL, "adjacency offsets decrease");
    NMG_CHECK(vstat[v] >= 0 && vstat[v] <= 5, NMG_EINVAL, "node status out of range");
  }
  const int64_t nadj = adj_ptr[n_nodes];
  for (int64_t k = 0; k < nadj; ++k) NMG_CHECK(adj[k] >= 0 && adj[k] < n_nodes, NMG_EINVAL, "adjacency out of range");
  // the DoF layout and the row offsets must be those of the statuses and the
  // adjacency: the gather writes each row's columns from them
  int64_t n_dof = 0, n_p = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    NMG_CHECK(vstt[v] == n_dof && pstt[v] == (hasp(vstat[v]) ? n_p : -1), NMG_EINVAL,
              "vstt / pstt do not match the node statuses");
    n_dof += vnum(vstat[v]);
    if (hasp(vstat[v])) n_p++;
  }
  {
    std::vector<int64_t> ad(n_dof + 1, 0), bb(n_dof + 1, 0), ee(n_dof + 1, 0), et(n_p + 1, 0), ap(n_p + 1, 0);
    for (int64_t v = 0; v < n_nodes; ++v) {
      const int64_t a0 = adj_ptr[v], deg = adj_ptr[v + 1] - a0;
      const int8_t sv = vstat[v];
      int64_t nS = 0, nF = 0, nFS = 0;
      for (int64_t k = 0; k < deg; ++k) {
        const int8_t sw = vstat[adj[a0 + k]];
        if (!hasp(sw) || vnum(sw) == 6) nS++;
        if (hasp(sw)) nF++;
        if (vnum(sw) == 6) nFS++;
      }
      auto set3 = [&](int64_t row, int64_t na, int64_t nb, int64_t ne) {
        for (int j = 0; j < 3; ++j) {
          ad[row + j + 1] = ad[row + j] + na;
          bb[row + j + 1] = bb[row + j] + nb;
          ee[row + j + 1] = ee[row + j] + ne;
        }
      };
      const int64_t r = vstt[v];
      if (hasp(sv)) {
        const int64_t p = pstt[v];
        if (vnum(sv) == 3) {
          NMG_CHECK(nF == deg, NMG_EINVAL, "a pure fluid node has a neighbour without a pressure");
          ap[p + 1] = ap[p] + deg;
          et[p + 1] = et[p] + 3 * deg;
          set3(r, 3 * deg, deg, deg);
        } else {
          set3(r, 3 * nS, nS, nFS);
          ap[p + 1] = ap[p] + nF;
          et[p + 1] = et[p] + 3 * (nF + nFS);
          set3(r + 3, 3 * nF, nF, nF);
        }
      } else {
        NMG_CHECK(vnum(sv) == 3, NMG_EINVAL, "a six-DoF node without a pressure");
        set3(r, 3 * deg, deg, 0);
      }
    }
    NMG_CHECK(std::equal(ad.begin(), ad.end(), ad_ptr) && std::equal(bb.begin(), bb.end(), b_ptr) &&
                  std::equal(ee.begin(), ee.end(), e_ptr) && std::equal(et.begin(), et.end(), et_ptr) &&
                  std::equal(ap.begin(), ap.end(), ap_ptr),
              NMG_EINVAL, "row offsets do not match the fluid-solid pattern of the adjacency");
  }
  const int L = 3 * np, LF = L + np, N2 = np * np, Ls = 3 * nfp;
  // element lists, pool offsets, interface blocks, statistics
  std::vector<int64_t> koff(n_elem), scm(n_elem, -1), flist, slist;
  int64_t wpool = 0, spool = 0;
  int64_t st[4] = {0, 0, 0, 0};
  for (int64_t e = 0; e < n_elem; ++e) {
    if (!efluid[e]) {
      koff[e] = (int64_t)slist.size() * L * L;
      slist.push_back(e);
      st[3]++;
      continue;
    }
    koff[e] = wpool;
    wpool += (int64_t)LF * LF;
    flist.push_back(e);
    st[2]++;
    int c = 0;
    for (int k = 0; k < 4; ++k) c += vstat[elems[e * np + vslot[k]]] == 2;
    if (c == 3) {
      st[0]++;
      if (g) {
        scm[e] = spool;
        spool += (int64_t)Ls * Ls;
      }
    }
    if (c == 4) st[1]++;
  }
  if (stats)
    for (int k = 0; k < 4; ++k) stats[k] = st[k];
  const int64_t ns = (int64_t)slist.size(), nf = (int64_t)flist.size();
  // node -> (element, slot) in ascending element order
  std::vector<int64_t> n2e_ptr(n_nodes + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int a = 0; a < np; ++a) n2e_ptr[elems[e * np + a] + 1]++;
  for (int64_t v = 0; v < n_nodes; ++v) n2e_ptr[v + 1] += n2e_ptr[v];
  std::vector<int64_t> n2e(n2e_ptr[n_nodes]);
  std::vector<int32_t> n2s(n2e_ptr[n_nodes]);
  {
    std::vector<int64_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
    for (int64_t e = 0; e < n_elem; ++e)
      for (int a = 0; a < np; ++a) {
        const int64_t v = elems[e * np + a];
        n2e[fill[v]] = e;
        n2s[fill[v]] = a;
        fill[v]++;
      }
  }
  // solid elements' inputs, compacted for the solid element kernel
  std::vector<int32_t> se((size_t)ns * np);
  std::vector<double> sl((size_t)ns * np), sm((size_t)ns * np), sr((size_t)ns * np), sg(g ? (size_t)ns * np * 3 : 0);
  for (int64_t i = 0; i < ns; ++i) {
    const int64_t e = slist[i];
    for (int a = 0; a < np; ++a) {
      se[i * np + a] = elems[e * np + a];
      sl[i * np + a] = lam[e * np + a];
      sm[i * np + a] = mu[e * np + a];
      sr[i * np + a] = rho[e * np + a];
      if (g)
        for (int c = 0; c < 3; ++c) sg[(i * np + a) * 3 + c] = g[(e * np + a) * 3 + c];
    }
  }
  const int64_t nnz_ad = ad_ptr[n_dof], nnz_b = b_ptr[n_dof], nnz_e = e_ptr[n_dof], nnz_et = et_ptr[n_p],
                nnz_ap = ap_ptr[n_p];
  Dev<double> dn(n_nodes * 3), dl(n_elem * np), dr(n_elem * np), dM(N2), dD(3 * N2), dMF(4 * nfp * nfp);
  Dev<double> dg(g ? n_elem * np * 3 : 0);
  Dev<int32_t> de(n_elem * np), dvs(4), dnb(n_elem * 4), dfm(4 * nfp), dadj(nadj), dn2s(n2s.size());
  Dev<int8_t> def(n_elem), dst(n_nodes);
  Dev<int64_t> dap(n_nodes + 1), dvt(n_nodes), dpt(n_nodes), dn2ep(n_nodes + 1), dn2e(n2e.size());
  Dev<int64_t> dko(n_elem), dsc(n_elem), dfl(nf);
  Dev<int64_t> rp_ad(n_dof + 1), rp_b(n_dof + 1), rp_e(n_dof + 1), rp_et(n_p + 1), rp_ap(n_p + 1);
  Dev<int32_t> c_ad(nnz_ad), c_b(nnz_b), c_e(nnz_e), c_et(nnz_et), c_ap(nnz_ap);
  Dev<double> v_ad(nnz_ad), v_b(nnz_b), v_e(nnz_e), v_et(nnz_et), v_ap(nnz_ap);
  Dev<fs::Info> dinfo(n_elem);
  Dev<double> dW(wpool), dSC(spool), dKs((size_t)ns * L * L), dMs((size_t)ns * N2);
  Dev<int32_t> dse(ns * np);
  Dev<double> dsl(ns * np), dsm(ns * np), dsr(ns * np), dsg(g ? ns * np * 3 : 0);
  Dev<int> dbad(3);
  Dev<unsigned long long> dmiss(1);
  dn.up(nodes, n_nodes * 3, s);
  dl.up(lam, n_elem * np, s);
  dr.up(rho, n_elem * np, s);
  dM.up(refM, N2, s);
  dD.up(refD, 3 * N2, s);
  dMF.up(massF, 4 * nfp * nfp, s);
  if (g) dg.up(g, n_elem * np * 3, s);
  de.up(elems, n_elem * np, s);
  dvs.up(vslot, 4, s);
  dnb.up(neigh, n_elem * 4, s);
  dfm.up(fmask, 4 * nfp, s);
  dadj.up(adj, nadj, s);
  dn2s.up(n2s.data(), n2s.size(), s);
  def.up(efluid, n_elem, s);
  dst.up(vstat, n_nodes, s);
  dap.up(adj_ptr, n_nodes + 1, s);
  dvt.up(vstt, n_nodes, s);
  dpt.up(pstt, n_nodes, s);
  dn2ep.up(n2e_ptr.data(), n_nodes + 1, s);
  dn2e.up(n2e.data(), n2e.size(), s);
  dko.up(koff.data(), n_elem, s);
  dsc.up(scm.data(), n_elem, s);
  dfl.up(flist.data(), nf, s);
  rp_ad.up(ad_ptr, n_dof + 1, s);
  rp_b.up(b_ptr, n_dof + 1, s);
  rp_e.up(e_ptr, n_dof + 1, s);
  rp_et.up(et_ptr, n_p + 1, s);
  rp_ap.up(ap_ptr, n_p + 1, s);
  dse.up(se.data(), se.size(), s);
  dsl.up(sl.data(), sl.size(), s);
  dsm.up(sm.data(), sm.size(), s);
  dsr.up(sr.data(), sr.size(), s);
  if (g) dsg.up(sg.data(), sg.size(), s);
  const int big[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  dbad.up(big, 3, s);
  const unsigned long long zero = 0;
  dmiss.up(&zero, 1, s);

  fs::Body B;
  B.n_nodes = n_nodes;
  B.n_elem = n_elem;
  B.np = np;
  B.nfp = nfp;
  B.purefluid = purefluid ? 1 : 0;
  B.nodes = dn.p;
  B.elems = de.p;
  B.vslot = dvs.p;
  B.lam = dl.p;
  B.rho = dr.p;
  B.g = g ? dg.p : nullptr;
  B.efluid = def.p;
  B.neigh = dnb.p;
  B.vstat = dst.p;
  B.refM = dM.p;
  B.refD = dD.p;
  B.massF = dMF.p;
  B.fmask = dfm.p;
  hipLaunchKernelGGL(fs_info_k, dim3((unsigned)((n_elem + 255) / 256)), dim3(256), 0, s, B, dinfo.p, dbad.p);
  NMG_HIP(hipGetLastError());
  if (ns > 0) {
    // the solid elements' stiffness (CGE3D_ISO with gravity), the kernel of
    // the solid producer on the compacted element list
    if (np == 10)
      hipLaunchKernelGGL(elem_k<10>, dim3((unsigned)ns), dim3(64), 0, s, ns, dn.p, dse.p, dvs.p, dsl.p, dsm.p, dsr.p,
                         g ? dsg.p : nullptr, dM.p, dD.p, dKs.p, dMs.p, dbad.p + 2);
    else
      hipLaunchKernelGGL(elem_k<4>, dim3((unsigned)ns), dim3(64), 0, s, ns, dn.p, dse.p, dvs.p, dsl.p, dsm.p, dsr.p,
                         g ? dsg.p : nullptr, dM.p, dD.p, dKs.p, dMs.p, dbad.p + 2);
    NMG_HIP(hipGetLastError());
  }
  int badh[3];
  NMG_HIP(hipMemcpyAsync(badh, dbad.p, sizeof(badh), hipMemcpyDeviceToHost, s));
  NMG_HIP(hipStreamSynchronize(s));
  NMG_CHECK(badh[0] == big[0], NMG_EINVAL, "flat element " + std::to_string(badh[0]));
  NMG_CHECK(badh[2] == big[0], NMG_EINVAL, "flat element " + std::to_string(slist[badh[2] < ns ? badh[2] : 0]));
  NMG_CHECK(badh[1] == big[0], NMG_EINVAL,
            "fluid element " + std::to_string(badh[1]) +
                " lies on the free surface: its surface term divides by the mean |g.n|, which needs gravity (JOB=2)");
  if (nf > 0) {
    hipLaunchKernelGGL(fs_fluid_k, dim3((unsigned)((nf + 63) / 64)), dim3(64), 0, s, B, nf, dfl.p, dko.p, dsc.p,
                       dW.p, dSC.p);
    NMG_HIP(hipGetLastError());
  }
  fs::Gather G;
  G.adj_ptr = dap.p;
  G.adj = dadj.p;
  G.vstt = dvt.p;
  G.pstt = dpt.p;
  G.n2e_ptr = dn2ep.p;
  G.n2e = dn2e.p;
  G.n2s = dn2s.p;
  G.info = dinfo.p;
  G.koff = dko.p;
  G.scm = dsc.p;
  G.Ks = dKs.p;
  G.W = dW.p;
  G.SC = dSC.p;
  G.ad_ptr = rp_ad.p;
  G.b_ptr = rp_b.p;
  G.e_ptr = rp_e.p;
  G.et_ptr = rp_et.p;
  G.ap_ptr = rp_ap.p;
  G.ad_col = c_ad.p;
  G.b_col = c_b.p;
  G.e_col = c_e.p;
  G.et_col = c_et.p;
  G.ap_col = c_ap.p;
  G.ad_val = v_ad.p;
  G.b_val = v_b.p;
  G.e_val = v_e.p;
  G.et_val = v_et.p;
  G.ap_val = v_ap.p;
  hipLaunchKernelGGL(fs_gather_k, dim3((unsigned)((n_nodes + 127) / 128)), dim3(128), 0, s, B, G, dmiss.p);
  NMG_HIP(hipGetLastError());
  unsigned long long missh = 0;
  NMG_HIP(hipMemcpyAsync(&missh, dmiss.p, sizeof(missh), hipMemcpyDeviceToHost, s));
  auto down = [&](auto* h, auto* d, int64_t n) {
    if (n) NMG_HIP(hipMemcpyAsync(h, d, n * sizeof(*h), hipMemcpyDeviceToHost, s));
  };
  down(ad_col, c_ad.p, nnz_ad);
  down(ad_val, v_ad.p, nnz_ad);
  down(b_col, c_b.p, nnz_b);
  down(b_val, v_b.p, nnz_b);
  down(e_col, c_e.p, nnz_e);
  down(e_val, v_e.p, nnz_e);
  down(et_col, c_et.p, nnz_et);
  down(et_val, v_et.p, nnz_et);
  down(ap_col, c_ap.p, nnz_ap);
  down(ap_val, v_ap.p, nnz_ap);
  NMG_HIP(hipStreamSynchronize(s));
  NMG_CHECK(missh == 0, NMG_EINVAL, "fluid-solid assembly: an element entry has no slot in the pattern");
}

void backward_fs_device(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                        const int32_t* vslot, const double* lam, const double* rho, const double* g,
                        const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM,
                        const double* refD, const double* massF, const int32_t* fmask, const int8_t* vstat,
                        const int64_t* vstt, const int64_t* pstt, int k, const double* eig, const double* U,
                        const double* P, double* d_rho, double* d_lam, double* d_mu, hipStream_t s) {
  using fs::hasp;
  using fs::vnum;
  NMG_CHECK(np == 4 || np == 10, NMG_EINVAL, "pNp must be 4 (P1) or 10 (P2)");
  NMG_CHECK(nfp == (np == 4 ? 3 : 6), NMG_EINVAL, "face nodes must be 3 (P1) or 6 (P2)");
  NMG_CHECK(n_nodes > 0 && n_elem > 0 && n_elem < 0x3fffffff, NMG_EINVAL, "empty or oversized mesh");
  NMG_CHECK(k >= 1, NMG_EINVAL, "no eigenpairs");
  for (int q = 0; q < 4; ++q) NMG_CHECK(vslot[q] >= 0 && vslot[q] < np, NMG_EINVAL, "bad vertex slot");
  for (int q = 0; q < 4 * nfp; ++q) NMG_CHECK(fmask[q] >= 0 && fmask[q] < np, NMG_EINVAL, "bad face mask");
  for (int64_t i = 0; i < n_elem * np; ++i)
    NMG_CHECK(elems[i] >= 0 && elems[i] < n_nodes, NMG_EINVAL, "element node id out of range");
  for (int64_t i = 0; i < n_elem * 4; ++i)
    NMG_CHECK(neigh[i] >= -1 && neigh[i] < n_elem, NMG_EINVAL, "face neighbour out of range");
  for (int64_t e = 0; e < n_elem; ++e) NMG_CHECK(efluid[e] == 0 || efluid[e] == 1, NMG_EINVAL, "bad fluid flag");
  int64_t n_dof = 0, n_p = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    NMG_CHECK(vstat[v] >= 0 && vstat[v] <= 5, NMG_EINVAL, "node status out of range");
    NMG_CHECK(vstt[v] == n_dof && pstt[v] == (hasp(vstat[v]) ? n_p : -1), NMG_EINVAL,
              "vstt / pstt do not match the node statuses");
    n_dof += vnum(vstat[v]);
    if (hasp(vstat[v])) n_p++;
  }
  // a fluid element's nodes all carry a pressure (read by the adjoint)
  for (int64_t e = 0; e < n_elem; ++e)
    if (efluid[e])
      for (int a = 0; a < np; ++a)
        NMG_CHECK(hasp(vstat[elems[e * np + a]]), NMG_EINVAL, "a fluid element has a node without a pressure");
  const int L = 3 * np, LF = L + np, N2 = np * np, Ls = 3 * nfp;
  std::vector<int64_t> koff(n_elem), scm(n_elem, -1), flist, slist;
  int64_t wpool = 0, spool = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    if (!efluid[e]) {
      koff[e] = (int64_t)slist.size() * L * L;
      slist.push_back(e);
      continue;
    }
    koff[e] = wpool;
    wpool += (int64_t)LF * LF;
    flist.push_back(e);
    int c = 0;
    for (int q = 0; q < 4; ++q) c += vstat[elems[e * np + vslot[q]]] == 2;
    if (c == 3 && g) {
      scm[e] = spool;
      spool += (int64_t)Ls * Ls;
    }
  }
  const int64_t ns = (int64_t)slist.size(), nf = (int64_t)flist.size();
  std::vector<int32_t> se((size_t)ns * np);
  std::vector<double> sg(g ? (size_t)ns * np * 3 : 0), one((size_t)ns * np, 1.0), zero((size_t)ns * np, 0.0);
  for (int64_t i = 0; i < ns; ++i) {
    const int64_t e = slist[i];
    for (int a = 0; a < np; ++a) {
      se[i * np + a] = elems[e * np + a];
      if (g)
        for (int c = 0; c < 3; ++c) sg[(i * np + a) * 3 + c] = g[(e * np + a) * 3 + c];
    }
  }
  Dev<double> dn(n_nodes * 3), dl(n_elem * np), dr(n_elem * np), dM(N2), dD(3 * N2), dMF(4 * nfp * nfp);
  Dev<double> dg(g ? n_elem * np * 3 : 0);
  Dev<int32_t> de(n_elem * np), dvs(4), dnb(n_elem * 4), dfm(4 * nfp), dse(ns * np);
  Dev<int8_t> def(n_elem), dst(n_nodes);
  Dev<int64_t> dvt(n_nodes), dpt(n_nodes), dko(n_elem), dsc(n_elem), dfl(nf);
  Dev<fs::Info> dinfo(n_elem);
  Dev<fs::Dual> dFr(wpool), dFl(wpool), dSr(spool), dSl(spool);
  Dev<double> dKl((size_t)ns * L * L), dKm((size_t)ns * L * L), dKr((size_t)ns * L * L), dMs((size_t)ns * N2);
  Dev<double> done(ns * np), dzero(ns * np), dsg(g ? ns * np * 3 : 0);
  Dev<double> deig(k), dU((size_t)k * n_dof), dP((size_t)k * n_p), o_r((size_t)k * n_elem), o_l((size_t)k * n_elem),
      o_m((size_t)k * n_elem);
  Dev<int> dbad(3);
  dn.up(nodes, n_nodes * 3, s);
  dl.up(lam, n_elem * np, s);
  dr.up(rho, n_elem * np, s);
  dM.up(refM, N2, s);
  dD.up(refD, 3 * N2, s);
  dMF.up(massF, 4 * nfp * nfp, s);
  if (g) dg.up(g, n_elem * np * 3, s);
  de.up(elems, n_elem * np, s);
  dvs.up(vslot, 4, s);
  dnb.up(neigh, n_elem * 4, s);
  dfm.up(fmask, 4 * nfp, s);
  dse.up(se.data(), se.size(), s);
  def.up(efluid, n_elem, s);
  dst.up(vstat, n_nodes, s);
  dvt.up(vstt, n_nodes, s);
  dpt.up(pstt, n_nodes, s);
  dko.up(koff.data(), n_elem, s);
  dsc.up(scm.data(), n_elem, s);
  dfl.up(flist.data(), nf, s);
  done.up(one.data(), one.size(), s);
  dzero.up(zero.data(), zero.size(), s);
  if (g) dsg.up(sg.data(), sg.size(), s);
  deig.up(eig, k, s);
  dU.up(U, (size_t)k * n_dof, s);
  dP.up(P, (size_t)k * n_p, s);
  const int big[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  dbad.up(big, 3, s);

  fs::Body B;
  B.n_nodes = n_nodes;
  B.n_elem = n_elem;
  B.np = np;
  B.nfp = nfp;
  B.purefluid = purefluid ? 1 : 0;
  B.nodes = dn.p;
  B.elems = de.p;
  B.vslot = dvs.p;
  B.lam = dl.p;
  B.rho = dr.p;
  B.g = g ? dg.p : nullptr;
  B.efluid = def.p;
  B.neigh = dnb.p;
  B.vstat = dst.p;
  B.refM = dM.p;
  B.refD = dD.p;
  B.massF = dMF.p;
  B.fmask = dfm.p;
  hipLaunchKernelGGL(fs_info_k, dim3((unsigned)((n_elem + 255) / 256)), dim3(256), 0, s, B, dinfo.p, dbad.p);
  NMG_HIP(hipGetLastError());
  if (ns > 0) {
    // the solid elements' stiffness at unit lam, mu and rho: its derivatives
    const double* unit[3][3] = {{done.p, dzero.p, dzero.p}, {dzero.p, done.p, dzero.p}, {dzero.p, dzero.p, done.p}};
    double* outK[3] = {dKl.p, dKm.p, dKr.p};
    for (int q = 0; q < 3; ++q) {
      if (np == 10)
        hipLaunchKernelGGL(elem_k<10>, dim3((unsigned)ns), dim3(64), 0, s, ns, dn.p, dse.p, dvs.p, unit[q][0],
                           unit[q][1], unit[q][2], g ? dsg.p : nullptr, dM.p, dD.p, outK[q], dMs.p, dbad.p + 2);
      else
        hipLaunchKernelGGL(elem_k<4>, dim3((unsigned)ns), dim3(64), 0, s, ns, dn.p, dse.p, dvs.p, unit[q][0],
                           unit[q][1], unit[q][2], g ? dsg.p : nullptr, dM.p, dD.p, outK[q], dMs.p, dbad.p + 2);
      NMG_HIP(hipGetLastError());
    }
  }
  int badh[3];
  NMG_HIP(hipMemcpyAsync(badh, dbad.p, sizeof(badh), hipMemcpyDeviceToHost, s));
  NMG_HIP(hipStreamSynchronize(s));
  NMG_CHECK(badh[0] == big[0], NMG_EINVAL, "flat element " + std::to_string(badh[0]));
  NMG_CHECK(badh[2] == big[0], NMG_EINVAL, "flat element " + std::to_string(slist[badh[2] < ns ? badh[2] : 0]));
  NMG_CHECK(badh[1] == big[0], NMG_EINVAL,
            "fluid element " + std::to_string(badh[1]) +
                " lies on the free surface: its surface term divides by the mean |g.n|, which needs gravity (JOB=2)");
  if (nf > 0) {
    hipLaunchKernelGGL(fs_dfluid_k, dim3((unsigned)((2 * nf + 63) / 64)), dim3(64), 0, s, B, nf, dfl.p, dko.p,
                       dsc.p, dFr.p, dFl.p, dSr.p, dSl.p);
    NMG_HIP(hipGetLastError());
  }
  fs::Adj A;
  A.vstt = dvt.p;
  A.pstt = dpt.p;
  A.info = dinfo.p;
  A.koff = dko.p;
  A.scm = dsc.p;
  A.dFr = dFr.p;
  A.dFl = dFl.p;
  A.dSr = dSr.p;
  A.dKl = dKl.p;
  A.dKm = dKm.p;
  A.dKr = dKr.p;
  A.n_dof = n_dof;
  A.n_p = n_p;
  A.k = k;
  A.eig = deig.p;
  A.U = dU.p;
  A.P = dP.p;
  A.d_rho = o_r.p;
  A.d_lam = o_l.p;
  A.d_mu = o_m.p;
  const int64_t nt = n_elem * (int64_t)k;
  hipLaunchKernelGGL(fs_adjoint_k, dim3((unsigned)((nt + 127) / 128)), dim3(128), 0, s, B, A);
  NMG_HIP(hipGetLastError());
  NMG_HIP(hipMemcpyAsync(d_rho, o_r.p, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
  NMG_HIP(hipMemcpyAsync(d_lam, o_l.p, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
  NMG_HIP(hipMemcpyAsync(d_mu, o_m.p, sizeof(double) * nt, hipMemcpyDeviceToHost, s));
  NMG_HIP(hipStreamSynchronize(s));
}

}  // namespace nmg
