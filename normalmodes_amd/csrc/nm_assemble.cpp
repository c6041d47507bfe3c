// CPU producer of the stiffness matrix A and mass matrix B for the isotropic
// elastic continuous-Galerkin element (solid regions).
//
// Element formulas follow CGE3D_ISO (src/mod_cg_create_matrix.f90:980-1266):
//   lam = rho (vp^2 - 2 vs^2), mu = rho vs^2 per element node        (:1018-1020)
//   D_i = sum_r invJ(i,r) Drst_r, invJ = 2 B^{-1}, detJ = det(B)/8    (:1040-1052,
//                                      src/mod_geometry.f90:1404-1406, 2216-2240)
//   ML = (M diag(lam) + diag(lam) M)/2, MU likewise with mu           (:1109-1124)
//   K_ii = detJ * sym( sum_k D_k^T MU D_k + D_i^T MU D_i + D_i^T ML D_i ) (:1183-1195)
//   K_ij = detJ * ( D_i^T ML D_j + D_j^T MU D_i ), i != j              (:1201-1212)
//   M_e  = detJ * (diag(rho) M + M diag(rho))/2, the same for all three
//          components (B = M (x) I_3)                                  (:1089-1091, 1193)
// With gravity (JOB=2) the block (i,j) adds detJ * sym_rho(OPrho_ij) with
//   OPrho_ij = (D_i^T M G_j + G_i M D_j)/2 - M (dg_ij + dg_ji)/2
//              - (D_j^T G_i M + M G_j D_i)/2,
//   G_j = diag(g_j at the element nodes), dg = least-squares gradient of g
//   (:1054-1080, 1131-1180), sym_rho(X) = (X diag(rho) + diag(rho) X)/2.
//
// Global layout matches the reference's CSR (matrixstruct,
// src/mod_cg_create_matrix.f90:1269-1455): node-major DoFs 3*v+p; the row of
// DoF (v,p) of A holds columns 3*w+q for every node w sharing an element with
// v, sorted ascending; the row of B holds columns 3*w+p.  Row offsets are
// 0-based.  Element matrices are computed in parallel and gathered per row in
// ascending element order, so the result is deterministic and A is exactly
// symmetric.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr int kMaxNp = 10;

struct Elem {
  int np;
  const double* M;    // np x np reference mass
  const double* Dr;   // 3 x np x np reference derivatives
};

inline void matmul(int n, const double* A, const double* B, double* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
      C[i * n + j] = s;
    }
}
// C = A^T B
inline void matmul_tn(int n, const double* A, const double* B, double* C) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += A[k * n + i] * B[k * n + j];
      C[i * n + j] = s;
    }
}

inline bool inv3(const double B[9], double Bi[9], double* det) {
  double d = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
             B[2] * (B[3] * B[7] - B[4] * B[6]);
  *det = d;
  if (d == 0.0) return false;
  double id = 1.0 / d;
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

// Least-squares gradient of the nodal gravity over one element
// (src/mod_cg_create_matrix.f90:1054-1080): dg = (X^T X)^{-1} X^T (g - mean g)
// with X the node coordinates minus their mean; dg[r*3+c] = d g_c / d x_r.
void gravity_gradient(int n, const double* xn, const double* g, double dg[9]) {
  double xm[3] = {0, 0, 0}, gm[3] = {0, 0, 0};
  for (int a = 0; a < n; ++a)
    for (int c = 0; c < 3; ++c) { xm[c] += xn[a * 3 + c]; gm[c] += g[a * 3 + c]; }
  for (int c = 0; c < 3; ++c) { xm[c] /= n; gm[c] /= n; }
  double N[9] = {0}, R[9] = {0};
  for (int a = 0; a < n; ++a)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        N[r * 3 + c] += (xn[a * 3 + r] - xm[r]) * (xn[a * 3 + c] - xm[c]);
        R[r * 3 + c] += (xn[a * 3 + r] - xm[r]) * (g[a * 3 + c] - gm[c]);
      }
  double Ni[9], dN;
  inv3(N, Ni, &dN);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Ni[r * 3 + k] * R[k * 3 + c];
      dg[r * 3 + c] = s;
    }
}

// Element stiffness (3np x 3np, block (i,j) at rows i*np.., cols j*np..) and
// scalar mass (np x np).  Returns false for a flat element.
bool element_matrices(const Elem& E, const double* X /*4x3 vertices*/, const double* lam,
                      const double* mu, const double* rho, const double* g /*np x 3 or null*/,
                      const double* xn /*np x 3 node coords, used with g*/, double* K,
                      double* Me, double* detJ_out, double* invJ_out) {
  const int n = E.np;
  double B[9], Bi[9], detB;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[r * 3 + c] = X[(r + 1) * 3 + c] - X[c];
  if (!inv3(B, Bi, &detB)) return false;
  const double detJ = detB / 8.0;
  double invJ[9];
  for (int k = 0; k < 9; ++k) invJ[k] = 2.0 * Bi[k];
  if (detJ_out) *detJ_out = detJ;
  if (invJ_out) std::memcpy(invJ_out, invJ, sizeof(invJ));

  double D[3][kMaxNp * kMaxNp];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < n * n; ++a)
      D[i][a] = invJ[i * 3 + 0] * E.Dr[0 * n * n + a] + invJ[i * 3 + 1] * E.Dr[1 * n * n + a] +
                invJ[i * 3 + 2] * E.Dr[2 * n * n + a];

  double ML[kMaxNp * kMaxNp], MU[kMaxNp * kMaxNp];
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const double m = E.M[a * n + b];
      ML[a * n + b] = (m * lam[b] + lam[a] * m) * 0.5;
      MU[a * n + b] = (m * mu[b] + mu[a] * m) * 0.5;
      Me[a * n + b] = detJ * ((rho[a] * m + m * rho[b]) * 0.5);
    }
  double P[3][kMaxNp * kMaxNp], Q[3][kMaxNp * kMaxNp];  // ML D_k, MU D_k
  double Ot[3][kMaxNp * kMaxNp];                          // D_k^T MU D_k
  double Os[kMaxNp * kMaxNp];
  for (int k = 0; k < 3; ++k) {
    matmul(n, ML, D[k], P[k]);
    matmul(n, MU, D[k], Q[k]);
    matmul_tn(n, D[k], Q[k], Ot[k]);
  }
  for (int a = 0; a < n * n; ++a) Os[a] = Ot[0][a] + Ot[1][a] + Ot[2][a];

  // gravity operator pieces
  double G[3][kMaxNp];
  double dg[9] = {0};
  const bool grav = (g != nullptr);
  if (grav) {
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < 3; ++c) G[c][a] = g[a * 3 + c];
    gravity_gradient(n, xn, g, dg);
  }

  const int L = 3 * n;
  double T1[kMaxNp * kMaxNp], T2[kMaxNp * kMaxNp];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double Op[kMaxNp * kMaxNp];
      if (i == j) {
        matmul_tn(n, D[i], P[i], T1);  // D_i^T ML D_i
        for (int a = 0; a < n * n; ++a) T2[a] = Os[a] + Ot[i][a] + T1[a];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) Op[a * n + b] = (T2[a * n + b] + T2[b * n + a]) * 0.5;
      } else {
        // OPtmp = D_i^T ML D_j + D_j^T MU D_i ; OPtrs = D_j^T ML D_i + D_i^T MU D_j
        double A1[kMaxNp * kMaxNp], A2[kMaxNp * kMaxNp], A3[kMaxNp * kMaxNp], A4[kMaxNp * kMaxNp];
        matmul_tn(n, D[i], P[j], A1);
        matmul_tn(n, D[j], Q[i], A2);
        matmul_tn(n, D[j], P[i], A3);
        matmul_tn(n, D[i], Q[j], A4);
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            Op[a * n + b] = ((A3[b * n + a] + A4[b * n + a]) + (A1[a * n + b] + A2[a * n + b])) * 0.5;
      }
      if (grav) {
        // OPrho for block (i,j)
        double R1[kMaxNp * kMaxNp], R2[kMaxNp * kMaxNp], OPr[kMaxNp * kMaxNp];
        // D_i^T M G_j
        double MG[kMaxNp * kMaxNp];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) MG[a * n + b] = E.M[a * n + b] * G[j][b];
        matmul_tn(n, D[i], MG, R1);
        // G_i M D_j
        double MD[kMaxNp * kMaxNp];
        matmul(n, E.M, D[j], MD);
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) R2[a * n + b] = G[i][a] * MD[a * n + b];
        const double sg = (dg[i * 3 + j] + dg[j * 3 + i]) * 0.5;
        for (int a = 0; a < n * n; ++a) OPr[a] = (R1[a] + R2[a]) * 0.5 - E.M[a] * sg;
        // D_j^T G_i M
        double GM[kMaxNp * kMaxNp];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) GM[a * n + b] = G[i][a] * E.M[a * n + b];
        matmul_tn(n, D[j], GM, R1);
        // M G_j D_i
        double GD[kMaxNp * kMaxNp];
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) GD[a * n + b] = G[j][a] * D[i][a * n + b];
        matmul(n, E.M, GD, R2);
        for (int a = 0; a < n * n; ++a) OPr[a] -= (R1[a] + R2[a]) * 0.5;
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) Op[a * n + b] += (OPr[a * n + b] * rho[b] + rho[a] * OPr[a * n + b]) * 0.5;
      }
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) K[(i * n + a) * L + (j * n + b)] = detJ * Op[a * n + b];
    }
  if (grav) {
    // exact block symmetry: K_ji := K_ij^T (the two evaluations agree to rounding)
    for (int r = 0; r < L; ++r)
      for (int c = r + 1; c < L; ++c) {
        double s = 0.5 * (K[r * L + c] + K[c * L + r]);
        K[r * L + c] = s;
        K[c * L + r] = s;
      }
  }
  return true;
}

}  // namespace

extern "C" {

// Node adjacency (sorted, includes the node itself) from element connectivity.
// Call with adj == nullptr to get the counts in adj_ptr (n_nodes+1 entries,
// int64); then call again with adj sized adj_ptr[n_nodes].
int nm_node_adjacency(int64_t n_nodes, int64_t n_elem, int np, const int32_t* elems,
                      int64_t* adj_ptr, int32_t* adj) {
  std::vector<int64_t> n2e_ptr(n_nodes + 1, 0);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int a = 0; a < np; ++a) n2e_ptr[elems[e * np + a] + 1]++;
  for (int64_t v = 0; v < n_nodes; ++v) n2e_ptr[v + 1] += n2e_ptr[v];
  std::vector<int64_t> n2e(n2e_ptr[n_nodes]);
  std::vector<int64_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int a = 0; a < np; ++a) n2e[fill[elems[e * np + a]]++] = e;
  if (adj == nullptr) {
    adj_ptr[0] = 0;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t v = 0; v < n_nodes; ++v) {
      std::vector<int32_t> nb;
      for (int64_t k = n2e_ptr[v]; k < n2e_ptr[v + 1]; ++k)
        for (int a = 0; a < np; ++a) nb.push_back(elems[n2e[k] * np + a]);
      std::sort(nb.begin(), nb.end());
      adj_ptr[v + 1] = std::unique(nb.begin(), nb.end()) - nb.begin();
    }
    for (int64_t v = 0; v < n_nodes; ++v) adj_ptr[v + 1] += adj_ptr[v];
    return 0;
  }
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t v = 0; v < n_nodes; ++v) {
    std::vector<int32_t> nb;
    for (int64_t k = n2e_ptr[v]; k < n2e_ptr[v + 1]; ++k)
      for (int a = 0; a < np; ++a) nb.push_back(elems[n2e[k] * np + a]);
    std::sort(nb.begin(), nb.end());
    int64_t m = std::unique(nb.begin(), nb.end()) - nb.begin();
    if (m != adj_ptr[v + 1] - adj_ptr[v]) continue;
    std::memcpy(adj + adj_ptr[v], nb.data(), m * sizeof(int32_t));
  }
  return 0;
}

// Assemble A (3x3-blocked CSR over DoFs) and B (M (x) I3 as DoF CSR).
//   nodes   [n_nodes*3]          node coordinates
//   elems   [n_elem*np]          0-based node ids, reference local order
//   vslot   [4]                  local slots of the 4 vertices
//   lam,mu,rho [n_elem*np]       per element-node Lame parameters and density
//   g       [n_elem*np*3] or null gravity vector per element node (JOB=2)
//   refM [np*np], refD [3*np*np] reference element matrices
//   adj_ptr/adj                  node adjacency from nm_node_adjacency
// Outputs (caller-allocated):
//   a_rowptr [3n+1] int64, a_col [9*nnz_adj] int32, a_val [9*nnz_adj]
//   b_rowptr [3n+1] int64, b_col [3*nnz_adj] int32, b_val [3*nnz_adj]
// Returns 0, or -(1+e) for a flat element e.
int nm_assemble_elastic(int64_t n_nodes, int64_t n_elem, int np, const double* nodes,
                        const int32_t* elems, const int32_t* vslot, const double* lam,
                        const double* mu, const double* rho, const double* g, const double* refM,
                        const double* refD, const int64_t* adj_ptr, const int32_t* adj,
                        int64_t* a_rowptr, int32_t* a_col, double* a_val, int64_t* b_rowptr,
                        int32_t* b_col, double* b_val) {
  const int L = 3 * np;
  Elem E{np, refM, refD};
  std::vector<double> Ke((size_t)n_elem * L * L);
  std::vector<double> Mes((size_t)n_elem * np * np);
  int64_t bad = -1;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elem; ++e) {
    double X[12], xn[kMaxNp * 3];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) X[k * 3 + c] = nodes[(int64_t)elems[e * np + vslot[k]] * 3 + c];
    for (int a = 0; a < np; ++a)
      for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
    bool ok = element_matrices(E, X, lam + e * np, mu + e * np, rho + e * np,
                               g ? g + e * np * 3 : nullptr, xn, &Ke[(size_t)e * L * L],
                               &Mes[(size_t)e * np * np], nullptr, nullptr);
    if (!ok) {
#pragma omp critical
      bad = e;
    }
  }
  if (bad >= 0) return (int)(-(1 + bad));

  // node -> (element, local slot), ascending element order
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
        int64_t v = elems[e * np + a];
        n2e[fill[v]] = e;
        n2s[fill[v]] = a;
        fill[v]++;
      }
  }
  // row pointers
  a_rowptr[0] = 0;
  b_rowptr[0] = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    int64_t deg = adj_ptr[v + 1] - adj_ptr[v];
    for (int p = 0; p < 3; ++p) {
      a_rowptr[3 * v + p + 1] = a_rowptr[3 * v + p] + 3 * deg;
      b_rowptr[3 * v + p + 1] = b_rowptr[3 * v + p] + deg;
    }
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t v = 0; v < n_nodes; ++v) {
    const int64_t a0 = adj_ptr[v], deg = adj_ptr[v + 1] - a0;
    for (int p = 0; p < 3; ++p) {
      int64_t ra = a_rowptr[3 * v + p], rb = b_rowptr[3 * v + p];
      for (int64_t k = 0; k < deg; ++k) {
        int32_t w = adj[a0 + k];
        for (int q = 0; q < 3; ++q) {
          a_col[ra + 3 * k + q] = 3 * w + q;
          a_val[ra + 3 * k + q] = 0.0;
        }
        b_col[rb + k] = 3 * w + p;
        b_val[rb + k] = 0.0;
      }
    }
    for (int64_t t = n2e_ptr[v]; t < n2e_ptr[v + 1]; ++t) {
      const int64_t e = n2e[t];
      const int m = n2s[t];
      const double* K = &Ke[(size_t)e * L * L];
      const double* Me = &Mes[(size_t)e * np * np];
      for (int nn = 0; nn < np; ++nn) {
        int32_t w = elems[e * np + nn];
        const int32_t* lo = adj + a0;
        int64_t k = std::lower_bound(lo, lo + deg, w) - lo;
        for (int p = 0; p < 3; ++p) {
          int64_t ra = a_rowptr[3 * v + p] + 3 * k;
          for (int q = 0; q < 3; ++q) a_val[ra + q] += K[(p * np + m) * L + (q * np + nn)];
          b_val[b_rowptr[3 * v + p] + k] += Me[m * np + nn];
        }
      }
    }
  }
  return 0;
}

// Per-element geometry: detJ and invJ (row-major 3x3), as used by the element
// formulas above (for the adjoint and for tests).
int nm_element_geometry(int64_t n_elem, int np, const double* nodes, const int32_t* elems,
                        const int32_t* vslot, double* detJ, double* invJ) {
  int64_t bad = -1;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elem; ++e) {
    double B[9], Bi[9], d;
    const double* x0 = nodes + (int64_t)elems[e * np + vslot[0]] * 3;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[r * 3 + c] = nodes[(int64_t)elems[e * np + vslot[r + 1]] * 3 + c] - x0[c];
    if (!inv3(B, Bi, &d)) {
#pragma omp critical
      bad = e;
      continue;
    }
    detJ[e] = d / 8.0;
    for (int k = 0; k < 9; ++k) invJ[e * 9 + k] = 2.0 * Bi[k];
  }
  return bad >= 0 ? (int)(-(1 + bad)) : 0;
}

}  // extern "C"

extern "C" {

// Symmetric part of the least-squares gravity gradient of every element,
// gsym[e*9 + i*3 + j] = (dg_ij + dg_ji)/2 with dg from the element's nodal
// gravity g (n_elem x np x 3) and node coordinates: the coefficient of M in
// the gravity block (i,j) of the stiffness (used by the adjoint with JOB=2).
int nm_element_gravity(int64_t n_elem, int np, const double* nodes, const int32_t* elems, const double* g,
                       double* gsym) {
  if (np > kMaxNp) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elem; ++e) {
    double xn[kMaxNp * 3], dg[9];
    for (int a = 0; a < np; ++a)
      for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
    gravity_gradient(np, xn, g + e * np * 3, dg);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) gsym[e * 9 + i * 3 + j] = 0.5 * (dg[i * 3 + j] + dg[j * 3 + i]);
  }
  return 0;
}

// Element matrices for a list of elements (used to check the adjoint):
// K_e (3np x 3np) and M_e (np x np) for the given per-element-node lam, mu,
// rho (n_sel x np each) and optional gravity g (n_sel x np x 3).
int nm_element_matrices(int64_t n_sel, int np, const double* nodes, const int32_t* elems,
                        const int64_t* sel, const int32_t* vslot, const double* lam, const double* mu,
                        const double* rho, const double* g, const double* refM, const double* refD,
                        double* K, double* Me) {
  Elem E{np, refM, refD};
  const int L = 3 * np;
  for (int64_t q = 0; q < n_sel; ++q) {
    const int64_t e = sel[q];
    double X[12], xn[kMaxNp * 3];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) X[k * 3 + c] = nodes[(int64_t)elems[e * np + vslot[k]] * 3 + c];
    for (int a = 0; a < np; ++a)
      for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
    if (!element_matrices(E, X, lam + q * np, mu + q * np, rho + q * np, g ? g + q * np * 3 : nullptr, xn,
                          K + q * L * L, Me + q * np * np, nullptr, nullptr))
      return (int)(-(1 + e));
  }
  return 0;
}

}  // extern "C"
