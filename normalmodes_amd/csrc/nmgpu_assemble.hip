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

}  // namespace nmg
