// Host build of the device fluid-solid functions (nmgpu_fs_elem.h, shared
// host/device code) for the CPU tests: the producer and the adjoint with the
// per-element and per-node functions the kernels run, called in host loops,
// with the entry points' argument lists (nm_assemble_fs / nmg_backward_fs).
// The CPU producer's element routines come from nm_assemble.cpp, included
// here so that its solid element is the one the device path replaces.
#include "nm_assemble.cpp"
#include "nmgpu_fs_elem.h"

extern "C" int host_assemble_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes,
                               const int32_t* elems, const int32_t* vslot, const double* lam, const double* mu,
                               const double* rho, const double* g, const int8_t* efluid, const int32_t* neigh,
                               int purefluid, const double* refM, const double* refD, const double* massF,
                               const int32_t* fmask, const int64_t* adj_ptr, const int32_t* adj, const int8_t* vstat,
                               const int64_t* vstt, const int64_t* pstt, const int64_t* ad_ptr, int32_t* ad_col,
                               double* ad_val, const int64_t* b_ptr, int32_t* b_col, double* b_val,
                               const int64_t* e_ptr, int32_t* e_col, double* e_val, const int64_t* et_ptr,
                               int32_t* et_col, double* et_val, const int64_t* ap_ptr, int32_t* ap_col,
                               double* ap_val, int64_t* stats) {
  using namespace nmg;
  const int L = 3 * np, LF = L + np, Ls = 3 * nfp;
  std::vector<int64_t> koff(n_elem), scm(n_elem, -1);
  int64_t wpool = 0, spool = 0, ns = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    if (!efluid[e]) { koff[e] = ns * L * L; ns++; continue; }
    koff[e] = wpool; wpool += (int64_t)LF * LF;
    int c = 0;
    for (int k = 0; k < 4; ++k) c += vstat[elems[e * np + vslot[k]]] == 2;
    if (c == 3 && g) { scm[e] = spool; spool += (int64_t)Ls * Ls; }
  }
  fs::Body B;
  B.n_nodes = n_nodes; B.n_elem = n_elem; B.np = np; B.nfp = nfp; B.purefluid = purefluid;
  B.nodes = nodes; B.elems = elems; B.vslot = vslot; B.lam = lam; B.rho = rho; B.g = g; B.efluid = efluid;
  B.neigh = neigh; B.vstat = vstat; B.refM = refM; B.refD = refD; B.massF = massF; B.fmask = fmask;
  std::vector<fs::Info> info(n_elem);
  std::vector<double> W(wpool + 1), SC(spool + 1), Ks(ns * L * L + 1);
  Elem E{np, refM, refD};
  for (int64_t e = 0; e < n_elem; ++e) {
    int rc = fs::element_info(B, e, &info[e]);
    if (rc == 1) return -(int)(1 + e);
    if (rc == 2) return -(int)(1 + n_elem + e);
  }
  for (int64_t e = 0; e < n_elem; ++e) {
    if (efluid[e]) {
      fs::fluid_element(B, e, &W[koff[e]], scm[e] >= 0 ? &SC[scm[e]] : nullptr);
    } else {
      double X[12], xn[kMaxNp * 3], Me[kMaxNp * kMaxNp], detJ;
      for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) X[k * 3 + c] = nodes[(int64_t)elems[e * np + vslot[k]] * 3 + c];
      for (int a = 0; a < np; ++a)
        for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
      const double* ge = g ? g + e * np * 3 : nullptr;
      if (!element_matrices(E, X, lam + e * np, mu + e * np, rho + e * np, ge, xn, &Ks[koff[e]], Me, &detJ, nullptr))
        return -(int)(1 + e);
    }
  }
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
        n2e[fill[v]] = e; n2s[fill[v]] = a; fill[v]++;
      }
  }
  fs::Gather G;
  G.adj_ptr = adj_ptr; G.adj = adj; G.vstt = vstt; G.pstt = pstt; G.n2e_ptr = n2e_ptr.data(); G.n2e = n2e.data();
  G.n2s = n2s.data(); G.info = info.data(); G.koff = koff.data(); G.scm = scm.data(); G.Ks = Ks.data();
  G.W = W.data(); G.SC = SC.data();
  G.ad_ptr = ad_ptr; G.b_ptr = b_ptr; G.e_ptr = e_ptr; G.et_ptr = et_ptr; G.ap_ptr = ap_ptr;
  G.ad_col = ad_col; G.b_col = b_col; G.e_col = e_col; G.et_col = et_col; G.ap_col = ap_col;
  G.ad_val = ad_val; G.b_val = b_val; G.e_val = e_val; G.et_val = et_val; G.ap_val = ap_val;
  int64_t miss = 0;
  for (int64_t v = 0; v < n_nodes; ++v) miss += fs::gather_node(B, G, v);
  if (stats) {
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    for (int64_t e = 0; e < n_elem; ++e) {
      if (!efluid[e]) { stats[3]++; continue; }
      stats[2]++;
      int c = 0;
      for (int k = 0; k < 4; ++k) c += vstat[elems[e * np + vslot[k]]] == 2;
      if (c == 3) stats[0]++;
      if (c == 4) stats[1]++;
    }
  }
  return miss ? -2 : 0;
}

extern "C" int host_backward_fs(int64_t n_nodes, int64_t n_elem, int np, int nfp, const double* nodes,
                               const int32_t* elems, const int32_t* vslot, const double* lam, const double* rho,
                               const double* g, const int8_t* efluid, const int32_t* neigh, int purefluid,
                               const double* refM, const double* refD, const double* massF, const int32_t* fmask,
                               const int8_t* vstat, const int64_t* vstt, const int64_t* pstt, int k,
                               const double* eig, const double* U, const double* P, double* d_rho, double* d_lam,
                               double* d_mu) {
  using namespace nmg;
  int64_t n_dof = 0, n_p = 0;
  for (int64_t v = 0; v < n_nodes; ++v) {
    n_dof += fs::vnum(vstat[v]);
    if (fs::hasp(vstat[v])) n_p++;
  }
  const int L = 3 * np, LF = L + np, Ls = 3 * nfp;
  std::vector<int64_t> koff(n_elem), scm(n_elem, -1);
  int64_t wpool = 0, spool = 0, ns = 0;
  for (int64_t e = 0; e < n_elem; ++e) {
    if (!efluid[e]) { koff[e] = ns * L * L; ns++; continue; }
    koff[e] = wpool; wpool += (int64_t)LF * LF;
    int c = 0;
    for (int q = 0; q < 4; ++q) c += vstat[elems[e * np + vslot[q]]] == 2;
    if (c == 3 && g) { scm[e] = spool; spool += (int64_t)Ls * Ls; }
  }
  fs::Body B;
  B.n_nodes = n_nodes; B.n_elem = n_elem; B.np = np; B.nfp = nfp; B.purefluid = purefluid;
  B.nodes = nodes; B.elems = elems; B.vslot = vslot; B.lam = lam; B.rho = rho; B.g = g; B.efluid = efluid;
  B.neigh = neigh; B.vstat = vstat; B.refM = refM; B.refD = refD; B.massF = massF; B.fmask = fmask;
  std::vector<fs::Info> info(n_elem);
  for (int64_t e = 0; e < n_elem; ++e)
    if (fs::element_info(B, e, &info[e])) return -1;
  std::vector<fs::Dual> Fr(wpool + 1), Fl(wpool + 1), Sr(spool + 1), Sl(spool + 1);
  std::vector<double> Kl(ns * L * L + 1), Km(ns * L * L + 1), Kr(ns * L * L + 1);
  Elem E{np, refM, refD};
  std::vector<double> one(np, 1.0), zero(np, 0.0);
  for (int64_t e = 0; e < n_elem; ++e) {
    if (efluid[e]) {
      for (int q = 0; q < 2; ++q) {
        fs::Dual lD[fs::kMaxNp], rD[fs::kMaxNp];
        for (int a = 0; a < np; ++a) {
          lD[a] = {lam[e * np + a], q == 1 ? 1.0 : 0.0};
          rD[a] = {rho[e * np + a], q == 0 ? 1.0 : 0.0};
        }
        fs::fluid_element_t<fs::Dual>(B, e, lD, rD, (q == 0 ? Fr : Fl).data() + koff[e],
                                      scm[e] >= 0 ? (q == 0 ? Sr : Sl).data() + scm[e] : nullptr);
      }
    } else {
      double X[12], xn[kMaxNp * 3], Me[kMaxNp * kMaxNp], detJ;
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) X[q * 3 + c] = nodes[(int64_t)elems[e * np + vslot[q]] * 3 + c];
      for (int a = 0; a < np; ++a)
        for (int c = 0; c < 3; ++c) xn[a * 3 + c] = nodes[(int64_t)elems[e * np + a] * 3 + c];
      const double* ge = g ? g + e * np * 3 : nullptr;
      element_matrices(E, X, one.data(), zero.data(), zero.data(), ge, xn, &Kl[koff[e]], Me, &detJ, nullptr);
      element_matrices(E, X, zero.data(), one.data(), zero.data(), ge, xn, &Km[koff[e]], Me, &detJ, nullptr);
      element_matrices(E, X, zero.data(), zero.data(), one.data(), ge, xn, &Kr[koff[e]], Me, &detJ, nullptr);
    }
  }
  fs::Adj A;
  A.vstt = vstt; A.pstt = pstt; A.info = info.data(); A.koff = koff.data(); A.scm = scm.data();
  A.dFr = Fr.data(); A.dFl = Fl.data(); A.dSr = Sr.data(); A.dKl = Kl.data(); A.dKm = Km.data(); A.dKr = Kr.data();
  A.n_dof = n_dof; A.n_p = n_p; A.k = k; A.eig = eig; A.U = U; A.P = P; A.d_rho = d_rho; A.d_lam = d_lam; A.d_mu = d_mu;
  for (int j = 0; j < k; ++j)
    for (int64_t e = 0; e < n_elem; ++e) fs::element_adjoint(B, A, e, j);
  return 0;
}
