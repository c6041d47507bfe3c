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

// ---------------------------------------------------------------------------
// Fluid-solid producer: the blocks Ad, B, E, ET, Ap of CGFSE3D_ISO
// (src/mod_cg_create_matrix.f90:103-977) on the DoF layout of
// matrixstruct_general (:1458-2033).
//
// Node status (src/mod_geometry.f90:283-312, 672-680): 0 solid, 1 fluid,
// 2 fluid-solid vertex; P2 edge nodes 3/4/5 from the smaller status of their
// two vertices.  Statuses 2 and 5 carry six displacement DoFs, the solid
// triple first and the fluid triple second (:1467-1473); every status other
// than 0 and 3 carries one pressure DoF (:1581-1613).  DoFs are numbered node
// by node (vstt), pressures in node order (pstt).
//
// Rows of node v (adjacency = nodes sharing an element with v, sorted):
//   pure fluid  Ad/B/E: fluid triples / component / pressure of every neighbour;
//               Ap/ET: pressure / fluid triple of every neighbour.
//   fluid-solid solid rows: solid triples of solid and fluid-solid neighbours,
//               E: pressures of fluid-solid neighbours; fluid rows: fluid
//               triples / pressures of neighbours with a pressure; Ap:
//               pressures of those; ET: every displacement DoF of those.
//   pure solid  Ad/B: solid triples of every neighbour; E: empty.
// Because vstt and pstt increase with the node number, these lists come out
// sorted when the adjacency is.
//
// Element terms (detJ, invJ, D_i and the least-squares gradients as in the
// solid element above; rho_bar the element-average density):
//   solid element: the CGE3D_ISO stiffness (gravity included) and the mass
//     M rho_bar detJ on each component (:287-422).
//   fluid element (max mu < TOL, :279): mass M rho_bar detJ (:468-469);
//     coupling C_i = (M D_i - M diag(rho g_i/lam)) detJ on (u_i, p) and its
//     transpose on (p, u_i) (:488-505, 612-633 without gravity); pressure
//     block -detJ diag(lam^-1/2) M diag(lam^-1/2) (:638-639); with gravity
//     the displacement blocks detJ (Gh_i M Gm_j + Gm_i M Gh_j)/2 with
//     Gh = diag(g/|g|), Gm_j = diag(g_j/|g| N2 rho) and
//     N2 = grad(rho).g/rho - |g|^2 rho/lam (0 where |g| < 1e-15) (:200-249,
//     :472-528).
//   faces of a fluid element (:643-706), with face mass MassF, outward unit
//     normal n and sJac = |cross|/4 (src/mod_geometry.f90:1410-1426,
//     2280-2303): a free-surface face subtracts MassF/(mean|g.n|)/(mean rho)
//     sJac from the pressure block; an interior face with fewer than three
//     fluid-solid vertices adds MassF (mean g.n)(mean rho) sJac n_j^2 to each
//     displacement block j (these cancel between the two fluid elements of a
//     face up to rounding).
//   interface (a fluid element with exactly three fluid-solid vertices; the
//     face opposite the fourth, :786-953): with gravity Ad's solid triples
//     lose rho_f sJac (n_i MassF G_j + n_j G_i MassF)/2; ET(p_m, solid u_q of
//     n) and E(solid u_p of m, p_n) lose n_q sJac MassF(n,m) and
//     n_p sJac MassF(m,n).
// Each row is gathered over its node's elements in ascending element order
// (volume terms of an element before its interface terms), so the result is
// deterministic and E = ET^T exactly.
namespace {

inline int fs_vnum(int8_t st) { return (st == 2 || st == 5) ? 6 : 3; }
inline bool fs_hasp(int8_t st) { return st != 0 && st != 3; }

struct FsElem {
  int kind = 0;      // 0 solid, 1 fluid
  double detJ = 0, rbar = 0;
  int64_t off = 0;   // element matrix in the pool
  int fcid = -1;     // interface face of a fluid element
  double nrm[3] = {0, 0, 0}, sj = 0;
  int64_t scm = -1;  // interface gravity block in the pool
};

// outward unit normals and sJac of the four faces (face f opposite vertex f)
void face_normals(const double* X /*4x3*/, double n[4][3], double sj[4]) {
  for (int f = 0; f < 4; ++f) {
    const double* v[3];
    int k = 0;
    for (int j = 0; j < 4; ++j)
      if (j != f) v[k++] = X + 3 * j;
    double a[3], b[3];
    for (int c = 0; c < 3; ++c) {
      a[c] = v[1][c] - v[0][c];
      b[c] = v[2][c] - v[0][c];
    }
    double m[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    sj[f] = len / 4.0;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      m[c] /= len;
      s += (X[3 * f + c] - v[0][c]) * m[c];
    }
    for (int c = 0; c < 3; ++c) n[f][c] = s > 0 ? -m[c] : m[c];
  }
}

// least-squares gradient of a nodal scalar over one element
void scalar_gradient(int n, const double* xn, const double* f, double grad[3]) {
  double xm[3] = {0, 0, 0}, fm = 0.0;
  for (int a = 0; a < n; ++a) {
    for (int c = 0; c < 3; ++c) xm[c] += xn[a * 3 + c];
    fm += f[a];
  }
  for (int c = 0; c < 3; ++c) xm[c] /= n;
  fm /= n;
  double N[9] = {0}, R[3] = {0, 0, 0};
  for (int a = 0; a < n; ++a)
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) N[r * 3 + c] += (xn[a * 3 + r] - xm[r]) * (xn[a * 3 + c] - xm[c]);
      R[r] += (xn[a * 3 + r] - xm[r]) * (f[a] - fm);
    }
  double Ni[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, d;
  inv3(N, Ni, &d);
  for (int r = 0; r < 3; ++r) grad[r] = Ni[r * 3 + 0] * R[0] + Ni[r * 3 + 1] * R[1] + Ni[r * 3 + 2] * R[2];
}

// Fluid element matrix FT ((3np+np)^2, displacement blocks then pressure).
// Returns 0, 1 for a flat element, 2 for a free surface without gravity.
int fluid_element(const Elem& E, const double* X, const double* lam, const double* rho, const double* g,
                  const double* xn, int purefluid, int nfp, const double* massF, const int32_t* fmask,
                  const int32_t* neigh, const int8_t* vst /*4 vertex statuses*/, double* FT, FsElem& out,
                  double* SCM) {
  const int n = E.np, L = 3 * n, LF = L + n;
  double B[9], Bi[9], detB;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[r * 3 + c] = X[(r + 1) * 3 + c] - X[c];
  if (!inv3(B, Bi, &detB)) return 1;
  const double detJ = detB / 8.0;
  double invJ[9];
  for (int k = 0; k < 9; ++k) invJ[k] = 2.0 * Bi[k];
  double D[3][kMaxNp * kMaxNp];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < n * n; ++a)
      D[i][a] = invJ[i * 3 + 0] * E.Dr[0 * n * n + a] + invJ[i * 3 + 1] * E.Dr[1 * n * n + a] +
                invJ[i * 3 + 2] * E.Dr[2 * n * n + a];
  double rs = 0.0;
  for (int a = 0; a < n; ++a) rs += rho[a];
  out.kind = 1;
  out.detJ = detJ;
  out.rbar = rs / n;
  for (int a = 0; a < LF * LF; ++a) FT[a] = 0.0;
  const bool grav = g != nullptr;
  // gravity quantities per node
  double gh[3][kMaxNp], gm[3][kMaxNp];
  if (grav) {
    double gr[3];
    scalar_gradient(n, xn, rho, gr);
    for (int a = 0; a < n; ++a) {
      const double* ga = g + a * 3;
      const double ng = std::sqrt(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2]);
      double N2 = (gr[0] * ga[0] + gr[1] * ga[1] + gr[2] * ga[2]) / rho[a] - ng * ng / lam[a] * rho[a];
      if (ng < 1e-15 || purefluid) N2 = 0.0;
      for (int i = 0; i < 3; ++i) {
        gh[i][a] = ga[i] / std::max(ng, 1e-15);
        gm[i][a] = gh[i][a] * N2 * rho[a];
      }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) {
            const double m = E.M[a * n + b];
            FT[(i * n + a) * LF + j * n + b] = detJ * ((gh[i][a] * m * gm[j][b] + gm[i][a] * m * gh[j][b]) * 0.5);
          }
  }
  // coupling C_i = M D_i (- M diag(rho g_i / lam))
  for (int i = 0; i < 3; ++i) {
    double C[kMaxNp * kMaxNp];
    matmul(n, E.M, D[i], C);
    if (grav)
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) C[a * n + b] -= E.M[a * n + b] * (rho[b] / lam[b] * g[b * 3 + i]);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        FT[(i * n + a) * LF + L + b] = C[a * n + b] * detJ;
        FT[(L + b) * LF + i * n + a] = C[a * n + b] * detJ;
      }
  }
  // pressure block
  double li[kMaxNp];
  for (int a = 0; a < n; ++a) li[a] = 1.0 / std::sqrt(lam[a]);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) FT[(L + a) * LF + L + b] = -detJ * (li[a] * E.M[a * n + b] * li[b]);
  // faces
  double nf[4][3], sjf[4];
  face_normals(X, nf, sjf);
  for (int f = 0; f < 4; ++f) {
    const int32_t* fm = fmask + f * nfp;
    const double* MF = massF + f * nfp * nfp;
    double srho = 0.0;
    for (int a = 0; a < nfp; ++a) srho += rho[fm[a]];
    srho /= nfp;
    if (neigh[f] < 0) {
      if (!grav) return 2;
      double sg = 0.0;
      for (int a = 0; a < nfp; ++a) {
        double gn = 0.0;
        for (int c = 0; c < 3; ++c) gn -= g[fm[a] * 3 + c] * nf[f][c];
        sg += std::sqrt(gn * gn);
      }
      sg /= nfp;
      for (int a = 0; a < nfp; ++a)
        for (int b = 0; b < nfp; ++b) FT[(L + fm[a]) * LF + L + fm[b]] -= MF[a * nfp + b] / sg / srho * sjf[f];
    } else {
      int cnt = 0;
      for (int j = 0; j < 4; ++j)
        if (j != f && vst[j] == 2) cnt++;
      if (cnt < 3 && grav) {
        double sg = 0.0;
        for (int a = 0; a < nfp; ++a)
          for (int c = 0; c < 3; ++c) sg += g[fm[a] * 3 + c] * nf[f][c];
        sg /= nfp;
        for (int j = 0; j < 3; ++j)
          for (int a = 0; a < nfp; ++a)
            for (int b = 0; b < nfp; ++b)
              FT[(j * n + fm[a]) * LF + j * n + fm[b]] += MF[a * nfp + b] * sg * srho * sjf[f] * (nf[f][j] * nf[f][j]);
      }
    }
  }
  // interface face
  int cnt = 0, other = -1;
  for (int j = 0; j < 4; ++j) {
    if (vst[j] == 2)
      cnt++;
    else
      other = j;
  }
  out.fcid = -1;
  if (cnt == 3) {
    const int f = other;
    out.fcid = f;
    for (int c = 0; c < 3; ++c) out.nrm[c] = nf[f][c];
    out.sj = sjf[f];
    if (grav && SCM) {
      const int32_t* fm = fmask + f * nfp;
      const double* MF = massF + f * nfp * nfp;
      double rf = 0.0;
      for (int a = 0; a < nfp; ++a) rf += rho[fm[a]];
      rf /= nfp;
      const int Ls = 3 * nfp;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int a = 0; a < nfp; ++a)
            for (int b = 0; b < nfp; ++b) {
              const double m = MF[a * nfp + b];
              const double t = (sjf[f] * nf[f][i] * (m * g[fm[b] * 3 + j]) +
                                sjf[f] * nf[f][j] * (g[fm[a] * 3 + i] * m)) * 0.5;
              SCM[(i * nfp + a) * Ls + j * nfp + b] = t * rf;
            }
    }
  }
  return 0;
}

// position of column c in the sorted row [lo, hi) of col, or -1
inline int64_t find_col(const int32_t* col, int64_t lo, int64_t hi, int64_t c) {
  const int32_t* p = std::lower_bound(col + lo, col + hi, (int32_t)c);
  return (p != col + hi && *p == c) ? (p - col) : -1;
}

}  // namespace

extern "C" {

// DoF layout from node statuses: vstt (first displacement DoF of each node),
// pstt (pressure DoF or -1), and the totals.
int nm_fs_layout(int64_t n_nodes, const int8_t* vstat, int64_t* vstt, int64_t* pstt, int64_t* n_dof,
                 int64_t* n_p) {
  int64_t d = 0, p = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    if (vstat[v] < 0 || vstat[v] > 5) return -(int)(1 + v);
    vstt[v] = d;
    d += fs_vnum(vstat[v]);
    pstt[v] = fs_hasp(vstat[v]) ? p++ : -1;
  }
  *n_dof = d;
  *n_p = p;
  return 0;
}

// Row pointers of Ad, B, E (n_dof + 1 each), ET and Ap (n_p + 1 each).
// Returns 0, or -(1+v) for a pure fluid node with a neighbour without a
// pressure (the reference stops there, src/mod_cg_create_matrix.f90:1690-1694),
// or -(1+n_nodes+v) for a six-DoF node without a pressure.
int nm_fs_rowptr(int64_t n_nodes, const int64_t* adj_ptr, const int32_t* adj, const int8_t* vstat,
                 const int64_t* vstt, const int64_t* pstt, int64_t* ad_ptr, int64_t* b_ptr, int64_t* e_ptr,
                 int64_t* et_ptr, int64_t* ap_ptr) {
  ad_ptr[0] = b_ptr[0] = e_ptr[0] = et_ptr[0] = ap_ptr[0] = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    const int64_t a0 = adj_ptr[v], deg = adj_ptr[v + 1] - a0;
    const int8_t sv = vstat[v];
    const int64_t r = vstt[v];
    int64_t nS = 0, nF = 0, nFS = 0;
    for (int64_t k = 0; k < deg; ++k) {
      const int8_t sw = vstat[adj[a0 + k]];
      if (!fs_hasp(sw) || fs_vnum(sw) == 6) nS++;
      if (fs_hasp(sw)) nF++;
      if (fs_vnum(sw) == 6) nFS++;
    }
    auto set3 = [&](int64_t row, int64_t na, int64_t nb, int64_t ne) {
      for (int j = 0; j < 3; ++j) {
        ad_ptr[row + j + 1] = ad_ptr[row + j] + na;
        b_ptr[row + j + 1] = b_ptr[row + j] + nb;
        e_ptr[row + j + 1] = e_ptr[row + j] + ne;
      }
    };
    if (fs_hasp(sv)) {
      const int64_t p = pstt[v];
      if (fs_vnum(sv) == 3) {
        if (nF != deg) return -(int)(1 + v);
        ap_ptr[p + 1] = ap_ptr[p] + deg;
        et_ptr[p + 1] = et_ptr[p] + 3 * deg;
        set3(r, 3 * deg, deg, deg);
      } else {
        set3(r, 3 * nS, nS, nFS);
        ap_ptr[p + 1] = ap_ptr[p] + nF;
        et_ptr[p + 1] = et_ptr[p] + 3 * (nF + nFS);
        set3(r + 3, 3 * nF, nF, nF);
      }
    } else {
      if (fs_vnum(sv) == 6) return -(int)(1 + n_nodes + v);
      set3(r, 3 * deg, deg, 0);
    }
  }
  return 0;
}

// Values (and columns) of Ad, B, E, ET, Ap.  Inputs as nm_assemble_elastic,
// plus: efluid[n_elem] (1 = fluid element), neigh[n_elem*4] (element across
// face f, -1 on the boundary), purefluid, nfp, massF[4*nfp*nfp] and
// fmask[4*nfp] of the reference faces, vstat/vstt/pstt and the row pointers
// from nm_fs_layout / nm_fs_rowptr.  g may be null (no gravity).
// stats[4]: fluid elements with an interface face, fluid elements with four
// fluid-solid vertices (no interface term, as in the reference), fluid
// elements, solid elements.
// Returns 0; -(1+e) for a flat element e; -(1+n_elem+e) for a fluid element
// on the free surface without gravity (the reference divides by mean|g.n|).
int nm_assemble_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes, const int32_t* elems,
                   const int32_t* vslot, const double* lam, const double* mu, const double* rho, const double* g,
                   const int8_t* efluid, const int32_t* neigh, int purefluid, const double* refM,
                   const double* refD, const double* massF, const int32_t* fmask, const int64_t* adj_ptr,
                   const int32_t* adj, const int8_t* vstat, const int64_t* vstt, const int64_t* pstt,
                   const int64_t* ad_ptr, int32_t* ad_col, double* ad_val, const int64_t* b_ptr, int32_t* b_col,
                   double* b_val, const int64_t* e_ptr, int32_t* e_col, double* e_val, const int64_t* et_ptr,
                   int32_t* et_col, double* et_val, const int64_t* ap_ptr, int32_t* ap_col, double* ap_val,
                   int64_t* stats) {
  if (np > kMaxNp || nfp > kMaxNp) return -1;
  Elem E{np, refM, refD};
  const int L = 3 * np, LF = L + np, Ls = 3 * nfp;
  std::vector<FsElem> info(n_elem);
  int64_t pool = 0, spool = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    info[e].kind = efluid[e] ? 1 : 0;
    info[e].off = pool;
    pool += efluid[e] ? (int64_t)LF * LF : (int64_t)L * L;
  }
  std::vector<double> W((size_t)pool);
  std::vector<int8_t> vs4((size_t)n_elem * 4);
  for (int64_t e = 0; e < n_elem; ++e)
    for (int k = 0; k < 4; ++k) vs4[e * 4 + k] = vstat[elems[e * np + vslot[k]]];
  // interface blocks with gravity: count first
  if (g)
    for (int64_t e = 0; e < n_elem; ++e) {
      if (!efluid[e]) continue;
      int c = 0;
      for (int k = 0; k < 4; ++k) c += vs4[e * 4 + k] == 2;
      if (c == 3) {
        info[e].scm = spool;
        spool += (int64_t)Ls * Ls;
      }
    }
  std::vector<double> SC((size_t)std::max<int64_t>(spool, 1));
  int64_t bad = -1;
  int bad_kind = 0;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elem; ++e) {
    double X[12], xn[kMaxNp * 3];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) X[k * 3 + c] = nodes[(int64_t)elems[e * np + vslot[k]] * 3 + c];
    for (int a = 0; a < np; ++a)
      for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
    const double* ge = g ? g + e * np * 3 : nullptr;
    int rc = 0;
    if (efluid[e]) {
      rc = fluid_element(E, X, lam + e * np, rho + e * np, ge, xn, purefluid, nfp, massF, fmask, neigh + e * 4,
                         &vs4[e * 4], &W[info[e].off], info[e], info[e].scm >= 0 ? &SC[info[e].scm] : nullptr);
    } else {
      double Me[kMaxNp * kMaxNp], detJ;
      if (!element_matrices(E, X, lam + e * np, mu + e * np, rho + e * np, ge, xn, &W[info[e].off], Me, &detJ,
                            nullptr))
        rc = 1;
      double rs = 0.0;
      for (int a = 0; a < np; ++a) rs += rho[e * np + a];
      info[e].detJ = detJ;
      info[e].rbar = rs / np;
    }
    if (rc) {
#pragma omp critical
      {
        bad = e;
        bad_kind = rc;
      }
    }
  }
  if (bad >= 0) return bad_kind == 1 ? (int)(-(1 + bad)) : (int)(-(1 + n_elem + bad));
  if (stats) {
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
      if (!efluid[e]) {
        stats[3]++;
        continue;
      }
      stats[2]++;
      int c = 0;
      for (int k = 0; k < 4; ++k) c += vs4[e * 4 + k] == 2;
      if (c == 3) stats[0]++;
      if (c == 4) stats[1]++;
    }
  }
  // node -> (element, slot)
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
  auto fluid3 = [&](int64_t w) { return vstt[w] + fs_vnum(vstat[w]) - 3; };
  int64_t miss = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : miss)
  for (int64_t v = 0; v < n_nodes; ++v) {
    const int64_t a0 = adj_ptr[v], deg = adj_ptr[v + 1] - a0;
    const int8_t sv = vstat[v];
    const bool hp = fs_hasp(sv);
    const int64_t r0 = vstt[v], rf = fluid3(v), pv = pstt[v];
    // columns
    {
      // solid rows (pure solid or the first triple of a fluid-solid node)
      if (!hp || fs_vnum(sv) == 6) {
        for (int j = 0; j < 3; ++j) {
          int64_t ia = ad_ptr[r0 + j], ib = b_ptr[r0 + j], ie = e_ptr[r0 + j];
          for (int64_t k = 0; k < deg; ++k) {
            const int32_t w = adj[a0 + k];
            const int8_t sw = vstat[w];
            if (hp && !(!fs_hasp(sw) || fs_vnum(sw) == 6)) continue;
            for (int q = 0; q < 3; ++q) {
              ad_col[ia] = (int32_t)(vstt[w] + q);
              ad_val[ia++] = 0.0;
            }
            b_col[ib] = (int32_t)(vstt[w] + j);
            b_val[ib++] = 0.0;
            if (hp && fs_vnum(sw) == 6) {
              e_col[ie] = (int32_t)pstt[w];
              e_val[ie++] = 0.0;
            }
          }
        }
      }
      if (hp) {
        // fluid rows (the only triple of a pure fluid node)
        for (int j = 0; j < 3; ++j) {
          int64_t ia = ad_ptr[rf + j], ib = b_ptr[rf + j], ie = e_ptr[rf + j];
          for (int64_t k = 0; k < deg; ++k) {
            const int32_t w = adj[a0 + k];
            if (!fs_hasp(vstat[w])) continue;
            for (int q = 0; q < 3; ++q) {
              ad_col[ia] = (int32_t)(fluid3(w) + q);
              ad_val[ia++] = 0.0;
            }
            b_col[ib] = (int32_t)(fluid3(w) + j);
            b_val[ib++] = 0.0;
            e_col[ie] = (int32_t)pstt[w];
            e_val[ie++] = 0.0;
          }
        }
        int64_t ip = ap_ptr[pv], it = et_ptr[pv];
        for (int64_t k = 0; k < deg; ++k) {
          const int32_t w = adj[a0 + k];
          if (!fs_hasp(vstat[w])) continue;
          ap_col[ip] = (int32_t)pstt[w];
          ap_val[ip++] = 0.0;
          // pure fluid rows see the fluid triple only; fluid-solid rows every DoF
          const int64_t first = fs_vnum(sv) == 6 ? vstt[w] : fluid3(w);
          const int64_t last = vstt[w] + fs_vnum(vstat[w]);
          for (int64_t c = first; c < last; ++c) {
            et_col[it] = (int32_t)c;
            et_val[it++] = 0.0;
          }
        }
      }
    }
    // values, element by element in ascending order
    for (int64_t t = n2e_ptr[v]; t < n2e_ptr[v + 1]; ++t) {
      const int64_t e = n2e[t];
      const int m = n2s[t];
      const FsElem& I = info[e];
      const double* K = &W[I.off];
      const int32_t* el = elems + e * np;
      if (I.kind == 0) {
        for (int nn = 0; nn < np; ++nn) {
          const int64_t w = el[nn];
          for (int p = 0; p < 3; ++p) {
            const int64_t row = r0 + p;
            for (int q = 0; q < 3; ++q) {
              const int64_t pos = find_col(ad_col, ad_ptr[row], ad_ptr[row + 1], vstt[w] + q);
              if (pos < 0) { miss++; continue; }
              ad_val[pos] += K[(p * np + m) * L + q * np + nn];
            }
            const int64_t pb = find_col(b_col, b_ptr[row], b_ptr[row + 1], vstt[w] + p);
            if (pb < 0) { miss++; continue; }
            b_val[pb] += refM[m * np + nn] * I.rbar * I.detJ;
          }
        }
        continue;
      }
      // fluid element
      for (int nn = 0; nn < np; ++nn) {
        const int64_t w = el[nn];
        const int64_t wf = fluid3(w);
        for (int p = 0; p < 3; ++p) {
          const int64_t row = rf + p;
          for (int q = 0; q < 3; ++q) {
            const int64_t pos = find_col(ad_col, ad_ptr[row], ad_ptr[row + 1], wf + q);
            if (pos < 0) { miss++; continue; }
            ad_val[pos] += K[(p * np + m) * LF + q * np + nn];
          }
          const int64_t pb = find_col(b_col, b_ptr[row], b_ptr[row + 1], wf + p);
          const int64_t pe = find_col(e_col, e_ptr[row], e_ptr[row + 1], pstt[w]);
          if (pb < 0 || pe < 0) { miss++; continue; }
          b_val[pb] += refM[m * np + nn] * I.rbar * I.detJ;
          e_val[pe] += K[(p * np + m) * LF + L + nn];
        }
        const int64_t pa = find_col(ap_col, ap_ptr[pv], ap_ptr[pv + 1], pstt[w]);
        if (pa < 0) { miss++; continue; }
        ap_val[pa] += K[(L + m) * LF + L + nn];
        for (int q = 0; q < 3; ++q) {
          const int64_t pt = find_col(et_col, et_ptr[pv], et_ptr[pv + 1], wf + q);
          if (pt < 0) { miss++; continue; }
          et_val[pt] += K[(L + m) * LF + q * np + nn];
        }
      }
      // interface face of this fluid element
      if (I.fcid >= 0) {
        const int32_t* fm = fmask + I.fcid * nfp;
        const double* MF = massF + I.fcid * nfp * nfp;
        int mi = -1;
        for (int a = 0; a < nfp; ++a)
          if (fm[a] == m) mi = a;
        if (mi < 0) continue;
        for (int ni = 0; ni < nfp; ++ni) {
          const int64_t w = el[fm[ni]];
          if (I.scm >= 0) {
            const double* S = &SC[I.scm];
            for (int p = 0; p < 3; ++p)
              for (int q = 0; q < 3; ++q) {
                const int64_t pos = find_col(ad_col, ad_ptr[r0 + p], ad_ptr[r0 + p + 1], vstt[w] + q);
                if (pos < 0) { miss++; continue; }
                ad_val[pos] -= S[(p * nfp + mi) * Ls + q * nfp + ni];
              }
          }
          for (int q = 0; q < 3; ++q) {
            const int64_t pt = find_col(et_col, et_ptr[pv], et_ptr[pv + 1], vstt[w] + q);
            if (pt < 0) { miss++; continue; }
            et_val[pt] -= I.nrm[q] * I.sj * MF[ni * nfp + mi];
          }
          for (int p = 0; p < 3; ++p) {
            const int64_t pe = find_col(e_col, e_ptr[r0 + p], e_ptr[r0 + p + 1], pstt[w]);
            if (pe < 0) { miss++; continue; }
            e_val[pe] -= I.nrm[p] * I.sj * MF[mi * nfp + ni];
          }
        }
      }
    }
  }
  if (miss) return -2;
  return 0;
}

}  // extern "C"
