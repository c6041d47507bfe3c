// Fluid-solid terms of the device FE producer: per-element geometry and
// face data, the fluid element matrix of CGFSE3D_ISO
// (src/mod_cg_create_matrix.f90:103-977) and the per-node gather into the
// layout of matrixstruct_general (:1458-2033).  Every function here is
// __host__ __device__ and works on one element or one node, so each kernel
// runs one thread per element or per node and the same code can be checked
// on the host.
//
// Node status 0 solid, 1 fluid, 2 fluid-solid vertex, 3/4/5 P2 edge nodes;
// statuses 2 and 5 carry six displacement DoFs (the solid triple first),
// every status other than 0 and 3 one pressure DoF.
//
// Fluid element matrix FT ((3np+np)^2: the three displacement components,
// then the pressure), with rho_bar the element-average density:
//   gravity blocks   detJ (Gh_i M Gm_j + Gm_i M Gh_j)/2, Gh = diag(g/|g|),
//                    Gm_j = diag(g_j/|g| N2 rho), N2 = grad(rho).g/rho -
//                    |g|^2 rho/lam (0 where |g| < 1e-15 or for a pure fluid)
//   coupling         C_i = (M D_i - M diag(rho g_i/lam)) detJ on (u_i, p),
//                    its transpose on (p, u_i)
//   pressure         -detJ diag(lam^-1/2) M diag(lam^-1/2)
//   boundary face    pressure block -= MassF/(mean|g.n|)/(mean rho) sJac
//   interior face    (fewer than three fluid-solid vertices, gravity)
//                    block (j,j) += MassF (mean g.n)(mean rho) sJac n_j^2
// and, for a fluid element with three fluid-solid vertices, the interface
// face opposite the fourth: with gravity the solid triples of Ad lose
// rho_f sJac (n_i MassF G_j + n_j G_i MassF)/2 (the SCM block), and E / ET
// lose n sJac MassF.  Rows sum their node's element contributions in
// ascending element order (volume terms before the interface terms), the
// order of the CPU producer, so E = ET^T exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define NMG_HD __host__ __device__

namespace nmg {
namespace fs {

constexpr int kMaxNp = 10;

// inputs shared by the element and gather functions (host or device pointers)
struct Body {
  int64_t n_nodes = 0, n_elem = 0;
  int np = 0, nfp = 0, purefluid = 0;
  const double* nodes = nullptr;    // [n_nodes*3]
  const int32_t* elems = nullptr;   // [n_elem*np], 0-based
  const int32_t* vslot = nullptr;   // [4] local slots of the vertices
  const double* lam = nullptr;      // [n_elem*np]
  const double* rho = nullptr;      // [n_elem*np]
  const double* g = nullptr;        // [n_elem*np*3] or null (no gravity)
  const int8_t* efluid = nullptr;   // [n_elem]
  const int32_t* neigh = nullptr;   // [n_elem*4], -1 on the boundary
  const int8_t* vstat = nullptr;    // [n_nodes]
  const double* refM = nullptr;     // [np*np]
  const double* refD = nullptr;     // [3*np*np]
  const double* massF = nullptr;    // [4*nfp*nfp]
  const int32_t* fmask = nullptr;   // [4*nfp]
};

// per-element data the gather needs
struct Info {
  double detJ, rbar, nrm[3], sj;
  int32_t fcid;   // interface face of a fluid element, or -1
  int32_t pad;
};

NMG_HD inline int vnum(int8_t st) { return (st == 2 || st == 5) ? 6 : 3; }
NMG_HD inline bool hasp(int8_t st) { return st != 0 && st != 3; }

NMG_HD inline bool inv3(const double B[9], double Bi[9], double* det) {
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

// vertex coordinates X (4x3), invJ = 2 B^{-1} (B the edge matrix), detJ = det B / 8
NMG_HD inline bool geometry(const Body& M, int64_t e, double X[12], double invJ[9], double* detJ) {
  for (int k = 0; k < 4; ++k)
    for (int c = 0; c < 3; ++c) X[k * 3 + c] = M.nodes[(int64_t)M.elems[e * M.np + M.vslot[k]] * 3 + c];
  double B[9], Bi[9], detB;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) B[r * 3 + c] = X[(r + 1) * 3 + c] - X[c];
  if (!inv3(B, Bi, &detB)) return false;
  *detJ = detB / 8.0;
  for (int k = 0; k < 9; ++k) invJ[k] = 2.0 * Bi[k];
  return true;
}

// outward unit normal and sJac = |cross|/4 of face f (opposite vertex f)
NMG_HD inline void face_normal(const double X[12], int f, double n[3], double* sj) {
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
  const double len = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  *sj = len / 4.0;
  double s = 0.0;
  for (int c = 0; c < 3; ++c) {
    m[c] /= len;
    s += (X[3 * f + c] - v[0][c]) * m[c];
  }
  for (int c = 0; c < 3; ++c) n[c] = s > 0 ? -m[c] : m[c];
}

// Per-element data for the gather.  Returns 0, 1 for a flat element, 2 for
// a fluid element on the free surface without gravity.
NMG_HD inline int element_info(const Body& M, int64_t e, Info* out) {
  double X[12], invJ[9], detJ;
  out->fcid = -1;
  out->pad = 0;
  out->nrm[0] = out->nrm[1] = out->nrm[2] = 0.0;
  out->sj = 0.0;
  if (!geometry(M, e, X, invJ, &detJ)) return 1;
  double rs = 0.0;
  for (int a = 0; a < M.np; ++a) rs += M.rho[e * M.np + a];
  out->detJ = detJ;
  out->rbar = rs / M.np;
  if (!M.efluid[e]) return 0;
  if (!M.g)
    for (int f = 0; f < 4; ++f)
      if (M.neigh[e * 4 + f] < 0) return 2;
  int cnt = 0, other = -1;
  for (int k = 0; k < 4; ++k) {
    if (M.vstat[M.elems[e * M.np + M.vslot[k]]] == 2)
      cnt++;
    else
      other = k;
  }
  if (cnt == 3) {
    out->fcid = other;
    face_normal(X, other, out->nrm, &out->sj);
  }
  return 0;
}

// Forward-mode dual number: value and derivative along one parameter.  The
// fluid element below is written once over its scalar type; with Dual it
// yields the exact derivative of every entry (the adjoint's dFT/dm).
struct Dual {
  double v, d;
};
NMG_HD inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
NMG_HD inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
NMG_HD inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
NMG_HD inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
NMG_HD inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
NMG_HD inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
NMG_HD inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
NMG_HD inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
NMG_HD inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
NMG_HD inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
NMG_HD inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
NMG_HD inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
NMG_HD inline Dual operator/(double a, Dual b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
NMG_HD inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
NMG_HD inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
NMG_HD inline double root(double a) { return ::sqrt(a); }
NMG_HD inline Dual root(Dual a) {
  const double s = ::sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
NMG_HD inline double value(double a) { return a; }
NMG_HD inline double value(Dual a) { return a.v; }

// least-squares gradient of the nodal density f over element e
template <class T>
NMG_HD inline void rho_gradient(const Body& M, int64_t e, const T* f, T grad[3]) {
  const int n = M.np;
  const int32_t* el = M.elems + e * n;
  double xm[3] = {0, 0, 0};
  T fm = T{0.0};
  for (int a = 0; a < n; ++a) {
    for (int c = 0; c < 3; ++c) xm[c] += M.nodes[(int64_t)el[a] * 3 + c];
    fm += f[a];
  }
  for (int c = 0; c < 3; ++c) xm[c] /= n;
  fm = fm / (double)n;
  double N[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  T R[3] = {T{0.0}, T{0.0}, T{0.0}};
  for (int a = 0; a < n; ++a)
    for (int r = 0; r < 3; ++r) {
      const double xr = M.nodes[(int64_t)el[a] * 3 + r] - xm[r];
      for (int c = 0; c < 3; ++c) N[r * 3 + c] += xr * (M.nodes[(int64_t)el[a] * 3 + c] - xm[c]);
      R[r] += xr * (f[a] - fm);
    }
  double Ni[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, d;
  inv3(N, Ni, &d);
  for (int r = 0; r < 3; ++r) grad[r] = Ni[r * 3 + 0] * R[0] + Ni[r * 3 + 1] * R[1] + Ni[r * 3 + 2] * R[2];
}

// Fluid element matrix FT ((4np)^2, row-major) and, for an interface
// element with gravity, the SCM block ((3nfp)^2) into SCM (may be null), for
// the element's nodal lam and rho (of scalar type T).  D_i = sum_r invJ(i,r)
// Dr_r is formed entry by entry where it is used.
template <class T>
NMG_HD inline void fluid_element_t(const Body& M, int64_t e, const T* lam, const T* rho, T* FT, T* SCM) {
  const int n = M.np, N2 = n * n, L = 3 * n, LF = L + n, nfp = M.nfp;
  double X[12], invJ[9], detJ;
  geometry(M, e, X, invJ, &detJ);
  const double* g = M.g ? M.g + e * n * 3 : nullptr;
  const double* Mr = M.refM;
  for (int t = 0; t < LF * LF; ++t) FT[t] = T{0.0};
  if (g) {
    T gr[3];
    rho_gradient(M, e, rho, gr);
    double gh[3][kMaxNp];
    T gm[3][kMaxNp];
    for (int a = 0; a < n; ++a) {
      const double* ga = g + a * 3;
      const double ng = ::sqrt(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2]);
      T n2 = (gr[0] * ga[0] + gr[1] * ga[1] + gr[2] * ga[2]) / rho[a] - ng * ng / lam[a] * rho[a];
      if (ng < 1e-15 || M.purefluid) n2 = T{0.0};
      for (int i = 0; i < 3; ++i) {
        gh[i][a] = ga[i] / (ng > 1e-15 ? ng : 1e-15);
        gm[i][a] = gh[i][a] * n2 * rho[a];
      }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b) {
            const double m = Mr[a * n + b];
            FT[(i * n + a) * LF + j * n + b] = detJ * ((gh[i][a] * m * gm[j][b] + gm[i][a] * m * gh[j][b]) * 0.5);
          }
  }
  // coupling C_i = M D_i (- M diag(rho g_i / lam))
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        double s0 = 0.0;
        for (int c = 0; c < n; ++c) {
          const int cb = c * n + b;
          const double d = invJ[i * 3 + 0] * M.refD[0 * N2 + cb] + invJ[i * 3 + 1] * M.refD[1 * N2 + cb] +
                           invJ[i * 3 + 2] * M.refD[2 * N2 + cb];
          s0 += Mr[a * n + c] * d;
        }
        T s = T{s0};
        if (g) s = s - Mr[a * n + b] * (rho[b] / lam[b] * g[b * 3 + i]);
        FT[(i * n + a) * LF + L + b] = s * detJ;
        FT[(L + b) * LF + i * n + a] = s * detJ;
      }
  // pressure block
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const T la = 1.0 / root(lam[a]), lb = 1.0 / root(lam[b]);
      FT[(L + a) * LF + L + b] = -detJ * (la * Mr[a * n + b] * lb);
    }
  // vertex statuses
  int8_t vs[4];
  for (int k = 0; k < 4; ++k) vs[k] = M.vstat[M.elems[e * n + M.vslot[k]]];
  // faces
  for (int f = 0; f < 4; ++f) {
    const int32_t* fm = M.fmask + f * nfp;
    const double* MF = M.massF + f * nfp * nfp;
    double nf[3], sjf;
    face_normal(X, f, nf, &sjf);
    T srho = T{0.0};
    for (int a = 0; a < nfp; ++a) srho += rho[fm[a]];
    srho = srho / (double)nfp;
    if (M.neigh[e * 4 + f] < 0) {
      if (!g) continue;   // rejected by element_info before this runs
      double sg = 0.0;
      for (int a = 0; a < nfp; ++a) {
        double gn = 0.0;
        for (int c = 0; c < 3; ++c) gn -= g[fm[a] * 3 + c] * nf[c];
        sg += ::sqrt(gn * gn);
      }
      sg /= nfp;
      for (int a = 0; a < nfp; ++a)
        for (int b = 0; b < nfp; ++b) FT[(L + fm[a]) * LF + L + fm[b]] -= MF[a * nfp + b] / sg / srho * sjf;
    } else {
      int cnt = 0;
      for (int j = 0; j < 4; ++j)
        if (j != f && vs[j] == 2) cnt++;
      if (cnt < 3 && g) {
        double sg = 0.0;
        for (int a = 0; a < nfp; ++a)
          for (int c = 0; c < 3; ++c) sg += g[fm[a] * 3 + c] * nf[c];
        sg /= nfp;
        for (int j = 0; j < 3; ++j)
          for (int a = 0; a < nfp; ++a)
            for (int b = 0; b < nfp; ++b)
              FT[(j * n + fm[a]) * LF + j * n + fm[b]] += MF[a * nfp + b] * sg * srho * sjf * (nf[j] * nf[j]);
      }
    }
  }
  // interface block with gravity
  int cnt = 0, other = -1;
  for (int j = 0; j < 4; ++j) {
    if (vs[j] == 2)
      cnt++;
    else
      other = j;
  }
  if (cnt == 3 && g && SCM) {
    const int f = other;
    const int32_t* fm = M.fmask + f * nfp;
    const double* MF = M.massF + f * nfp * nfp;
    double nf[3], sjf;
    face_normal(X, f, nf, &sjf);
    T rf = T{0.0};
    for (int a = 0; a < nfp; ++a) rf += rho[fm[a]];
    rf = rf / (double)nfp;
    const int Ls = 3 * nfp;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int a = 0; a < nfp; ++a)
          for (int b = 0; b < nfp; ++b) {
            const double m = MF[a * nfp + b];
            const double t = (sjf * nf[i] * (m * g[fm[b] * 3 + j]) + sjf * nf[j] * (g[fm[a] * 3 + i] * m)) * 0.5;
            SCM[(i * nfp + a) * Ls + j * nfp + b] = t * rf;
          }
  }
}

NMG_HD inline void fluid_element(const Body& M, int64_t e, double* FT, double* SCM) {
  fluid_element_t<double>(M, e, M.lam + e * M.np, M.rho + e * M.np, FT, SCM);
}

// Adjoint of a fluid-solid body: d lambda_j / d m_e = w^T (dK_e/dm - lambda_j
// dM_e/dm) w for m in {rho, lam, mu} changed uniformly at the nodes of
// element e, with w = (u, p) the element's displacement and pressure values
// of pair j, u B-normalised and p = -Ap^-1 ET u (the pressures the Schur
// complement eliminates).  Fluid elements: dFT/dm by dual numbers through
// fluid_element_t (dFr, dFl: value and derivative pairs), minus the
// interface block's dSCM/drho on the solid triples, and dB/drho = M detJ
// (element-average density) on the fluid triples; d/dmu is 0 there.  Solid
// elements: dK/dm is the element stiffness at unit lam, mu or rho (the
// gravity term is linear in rho) on the solid triples, dB/drho = M detJ.
struct Adj {
  const int64_t* vstt;
  const int64_t* pstt;
  const Info* info;
  const int64_t* koff;   // solid: into the unit-parameter stiffness pools; fluid: into the dual pools
  const int64_t* scm;    // interface block offset into dSr, or -1
  const Dual* dFr;       // fluid dFT/drho and dFT/dlam, (4np)^2 each at koff
  const Dual* dFl;
  const Dual* dSr;       // interface dSCM/drho, (3nfp)^2 at scm
  const double* dKl;     // solid stiffness at lam = 1, mu = 1, rho = 1 (the others 0), (3np)^2 at koff
  const double* dKm;
  const double* dKr;
  int64_t n_dof, n_p;
  int k;
  const double* eig;     // [k]
  const double* U;       // [k * n_dof], pair-major
  const double* P;       // [k * n_p]
  double *d_rho, *d_lam, *d_mu;   // [k * n_elem]
};

NMG_HD inline void element_adjoint(const Body& M, const Adj& A, int64_t e, int j) {
  const int n = M.np, L = 3 * n, LF = L + n, nfp = M.nfp, Ls = 3 * nfp;
  const int32_t* el = M.elems + e * n;
  const double* u = A.U + (int64_t)j * A.n_dof;
  const double* pr = A.P + (int64_t)j * A.n_p;
  const double lam_j = A.eig[j], detJ = A.info[e].detJ;
  const int64_t out = (int64_t)j * M.n_elem + e;
  if (!M.efluid[e]) {
    // local displacement i*n + a: the solid triple of node a
    double ul[3 * kMaxNp];
    for (int r = 0; r < L; ++r) ul[r] = u[A.vstt[el[r % n]] + r / n];
    double ql = 0.0, qm = 0.0, qr = 0.0, mb = 0.0;
    const double* Kl = A.dKl + A.koff[e];
    const double* Km = A.dKm + A.koff[e];
    const double* Kr = A.dKr + A.koff[e];
    for (int r = 0; r < L; ++r) {
      double sl = 0.0, sm = 0.0, sr = 0.0;
      for (int c = 0; c < L; ++c) {
        sl += Kl[r * L + c] * ul[c];
        sm += Km[r * L + c] * ul[c];
        sr += Kr[r * L + c] * ul[c];
      }
      ql += ul[r] * sl;
      qm += ul[r] * sm;
      qr += ul[r] * sr;
    }
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < n; ++a) {
        double s = 0.0;
        for (int b = 0; b < n; ++b) s += M.refM[a * n + b] * ul[i * n + b];
        mb += ul[i * n + a] * s;
      }
    A.d_lam[out] = ql;
    A.d_mu[out] = qm;
    A.d_rho[out] = qr - lam_j * detJ * mb;
    return;
  }
  // local row r < 3n: fluid triple component r / n of node r % n; r >= 3n: pressure of node r - 3n
  double w[4 * kMaxNp];
  for (int r = 0; r < L; ++r) {
    const int64_t v = el[r % n];
    w[r] = u[A.vstt[v] + vnum(M.vstat[v]) - 3 + r / n];
  }
  for (int a = 0; a < n; ++a) w[L + a] = pr[A.pstt[el[a]]];
  const Dual* Fr = A.dFr + A.koff[e];
  const Dual* Fl = A.dFl + A.koff[e];
  double qr = 0.0, ql = 0.0, mb = 0.0;
  for (int r = 0; r < LF; ++r) {
    double sr = 0.0, sl = 0.0;
    for (int c = 0; c < LF; ++c) {
      sr += Fr[r * LF + c].d * w[c];
      sl += Fl[r * LF + c].d * w[c];
    }
    qr += w[r] * sr;
    ql += w[r] * sl;
  }
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < n; ++a) {
      double s = 0.0;
      for (int b = 0; b < n; ++b) s += M.refM[a * n + b] * w[i * n + b];
      mb += w[i * n + a] * s;
    }
  if (A.scm[e] >= 0) {
    // the interface block enters the solid triples of the face's nodes with a minus sign
    const int32_t* fm = M.fmask + A.info[e].fcid * nfp;
    const Dual* S = A.dSr + A.scm[e];
    double q = 0.0;
    for (int p = 0; p < 3; ++p)
      for (int mi = 0; mi < nfp; ++mi) {
        const double up = u[A.vstt[el[fm[mi]]] + p];
        double s = 0.0;
        for (int qq = 0; qq < 3; ++qq)
          for (int ni = 0; ni < nfp; ++ni) s += S[(p * nfp + mi) * Ls + qq * nfp + ni].d * u[A.vstt[el[fm[ni]]] + qq];
        q += up * s;
      }
    qr -= q;
  }
  A.d_rho[out] = qr - lam_j * detJ * mb;
  A.d_lam[out] = ql;
  A.d_mu[out] = 0.0;
}

// the fluid-solid pattern and the element pools the gather reads
struct Gather {
  const int64_t* adj_ptr;   // [n_nodes+1]
  const int32_t* adj;       // sorted node adjacency
  const int64_t* vstt;      // first displacement DoF of each node
  const int64_t* pstt;      // pressure DoF of each node or -1
  const int64_t* n2e_ptr;   // node -> (element, slot), ascending element order
  const int64_t* n2e;
  const int32_t* n2s;
  const Info* info;         // [n_elem]
  const int64_t* koff;      // element matrix offset: into Ks (solid, (3np)^2) or W (fluid, (4np)^2)
  const int64_t* scm;       // interface block offset into SC, or -1
  const double* Ks;
  const double* W;
  const double* SC;
  const int64_t *ad_ptr, *b_ptr, *e_ptr, *et_ptr, *ap_ptr;
  int32_t *ad_col, *b_col, *e_col, *et_col, *ap_col;
  double *ad_val, *b_val, *e_val, *et_val, *ap_val;
};

// position of column c in the sorted row [lo, hi) of col, or -1
NMG_HD inline int64_t find_col(const int32_t* col, int64_t lo, int64_t hi, int64_t c) {
  int64_t a = lo, b = hi;
  while (a < b) {
    const int64_t mid = a + ((b - a) >> 1);
    if (col[mid] < c)
      a = mid + 1;
    else
      b = mid;
  }
  return (a < hi && col[a] == c) ? a : -1;
}

// Columns (zero values) of every row of node v, then its values summed over
// its elements in ascending order.  Returns the number of element entries
// that found no slot (0 for a consistent pattern).
NMG_HD inline int64_t gather_node(const Body& M, const Gather& G, int64_t v) {
  const int np = M.np, nfp = M.nfp, L = 3 * np, LF = L + np, Ls = 3 * nfp;
  const int64_t a0 = G.adj_ptr[v], deg = G.adj_ptr[v + 1] - a0;
  const int8_t sv = M.vstat[v];
  const bool hp = hasp(sv);
  const int64_t r0 = G.vstt[v], pv = G.pstt[v];
  const int64_t rf = G.vstt[v] + vnum(sv) - 3;
  int64_t miss = 0;
  // solid rows (pure solid, or the first triple of a fluid-solid node)
  if (!hp || vnum(sv) == 6) {
    for (int j = 0; j < 3; ++j) {
      int64_t ia = G.ad_ptr[r0 + j], ib = G.b_ptr[r0 + j], ie = G.e_ptr[r0 + j];
      for (int64_t k = 0; k < deg; ++k) {
        const int32_t w = G.adj[a0 + k];
        const int8_t sw = M.vstat[w];
        if (hp && !(!hasp(sw) || vnum(sw) == 6)) continue;
        for (int q = 0; q < 3; ++q) {
          G.ad_col[ia] = (int32_t)(G.vstt[w] + q);
          G.ad_val[ia++] = 0.0;
        }
        G.b_col[ib] = (int32_t)(G.vstt[w] + j);
        G.b_val[ib++] = 0.0;
        if (hp && vnum(sw) == 6) {
          G.e_col[ie] = (int32_t)G.pstt[w];
          G.e_val[ie++] = 0.0;
        }
      }
    }
  }
  if (hp) {
    // fluid rows (the only triple of a pure fluid node)
    for (int j = 0; j < 3; ++j) {
      int64_t ia = G.ad_ptr[rf + j], ib = G.b_ptr[rf + j], ie = G.e_ptr[rf + j];
      for (int64_t k = 0; k < deg; ++k) {
        const int32_t w = G.adj[a0 + k];
        if (!hasp(M.vstat[w])) continue;
        const int64_t wf = G.vstt[w] + vnum(M.vstat[w]) - 3;
        for (int q = 0; q < 3; ++q) {
          G.ad_col[ia] = (int32_t)(wf + q);
          G.ad_val[ia++] = 0.0;
        }
        G.b_col[ib] = (int32_t)(wf + j);
        G.b_val[ib++] = 0.0;
        G.e_col[ie] = (int32_t)G.pstt[w];
        G.e_val[ie++] = 0.0;
      }
    }
    int64_t ip = G.ap_ptr[pv], it = G.et_ptr[pv];
    for (int64_t k = 0; k < deg; ++k) {
      const int32_t w = G.adj[a0 + k];
      if (!hasp(M.vstat[w])) continue;
      G.ap_col[ip] = (int32_t)G.pstt[w];
      G.ap_val[ip++] = 0.0;
      // pure fluid rows see the fluid triple only; fluid-solid rows every DoF
      const int64_t first = vnum(sv) == 6 ? G.vstt[w] : G.vstt[w] + vnum(M.vstat[w]) - 3;
      const int64_t last = G.vstt[w] + vnum(M.vstat[w]);
      for (int64_t c = first; c < last; ++c) {
        G.et_col[it] = (int32_t)c;
        G.et_val[it++] = 0.0;
      }
    }
  }
  // values, element by element in ascending order
  for (int64_t t = G.n2e_ptr[v]; t < G.n2e_ptr[v + 1]; ++t) {
    const int64_t e = G.n2e[t];
    const int m = G.n2s[t];
    const Info& I = G.info[e];
    const int32_t* el = M.elems + e * np;
    if (!M.efluid[e]) {
      const double* K = G.Ks + G.koff[e];
      for (int nn = 0; nn < np; ++nn) {
        const int64_t w = el[nn];
        for (int p = 0; p < 3; ++p) {
          const int64_t row = r0 + p;
          for (int q = 0; q < 3; ++q) {
            const int64_t pos = find_col(G.ad_col, G.ad_ptr[row], G.ad_ptr[row + 1], G.vstt[w] + q);
            if (pos < 0) {
              miss++;
              continue;
            }
            G.ad_val[pos] += K[(p * np + m) * L + q * np + nn];
          }
          const int64_t pb = find_col(G.b_col, G.b_ptr[row], G.b_ptr[row + 1], G.vstt[w] + p);
          if (pb < 0) {
            miss++;
            continue;
          }
          G.b_val[pb] += M.refM[m * np + nn] * I.rbar * I.detJ;
        }
      }
      continue;
    }
    // fluid element
    const double* K = G.W + G.koff[e];
    for (int nn = 0; nn < np; ++nn) {
      const int64_t w = el[nn];
      const int64_t wf = G.vstt[w] + vnum(M.vstat[w]) - 3;
      for (int p = 0; p < 3; ++p) {
        const int64_t row = rf + p;
        for (int q = 0; q < 3; ++q) {
          const int64_t pos = find_col(G.ad_col, G.ad_ptr[row], G.ad_ptr[row + 1], wf + q);
          if (pos < 0) {
            miss++;
            continue;
          }
          G.ad_val[pos] += K[(p * np + m) * LF + q * np + nn];
        }
        const int64_t pb = find_col(G.b_col, G.b_ptr[row], G.b_ptr[row + 1], wf + p);
        const int64_t pe = find_col(G.e_col, G.e_ptr[row], G.e_ptr[row + 1], G.pstt[w]);
        if (pb < 0 || pe < 0) {
          miss++;
          continue;
        }
        G.b_val[pb] += M.refM[m * np + nn] * I.rbar * I.detJ;
        G.e_val[pe] += K[(p * np + m) * LF + L + nn];
      }
      const int64_t pa = find_col(G.ap_col, G.ap_ptr[pv], G.ap_ptr[pv + 1], G.pstt[w]);
      if (pa < 0) {
        miss++;
        continue;
      }
      G.ap_val[pa] += K[(L + m) * LF + L + nn];
      for (int q = 0; q < 3; ++q) {
        const int64_t pt = find_col(G.et_col, G.et_ptr[pv], G.et_ptr[pv + 1], wf + q);
        if (pt < 0) {
          miss++;
          continue;
        }
        G.et_val[pt] += K[(L + m) * LF + q * np + nn];
      }
    }
    // interface face of this fluid element
    if (I.fcid >= 0) {
      const int32_t* fm = M.fmask + I.fcid * nfp;
      const double* MF = M.massF + I.fcid * nfp * nfp;
      int mi = -1;
      for (int a = 0; a < nfp; ++a)
        if (fm[a] == m) mi = a;
      if (mi < 0) continue;
      for (int ni = 0; ni < nfp; ++ni) {
        const int64_t w = el[fm[ni]];
        if (G.scm[e] >= 0) {
          const double* S = G.SC + G.scm[e];
          for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) {
              const int64_t pos = find_col(G.ad_col, G.ad_ptr[r0 + p], G.ad_ptr[r0 + p + 1], G.vstt[w] + q);
              if (pos < 0) {
                miss++;
                continue;
              }
              G.ad_val[pos] -= S[(p * nfp + mi) * Ls + q * nfp + ni];
            }
        }
        for (int q = 0; q < 3; ++q) {
          const int64_t pt = find_col(G.et_col, G.et_ptr[pv], G.et_ptr[pv + 1], G.vstt[w] + q);
          if (pt < 0) {
            miss++;
            continue;
          }
          G.et_val[pt] -= I.nrm[q] * I.sj * MF[ni * nfp + mi];
        }
        for (int p = 0; p < 3; ++p) {
          const int64_t pe = find_col(G.e_col, G.e_ptr[r0 + p], G.e_ptr[r0 + p + 1], G.pstt[w]);
          if (pe < 0) {
            miss++;
            continue;
          }
          G.e_val[pe] -= I.nrm[p] * I.sj * MF[mi * nfp + ni];
        }
      }
    }
  }
  return miss;
}

}  // namespace fs
}  // namespace nmg
