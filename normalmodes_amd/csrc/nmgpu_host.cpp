// Host-side math of the filtered-Lanczos path: polynomial filter design,
// tridiagonal eigenvalues and the Lanczos start vector.
#include <algorithm>
#include <cmath>
#include <vector>

#include "nmgpu_internal.h"

namespace nmg {

// ---------------------------------------------------------------------------
// Chebyshev-delta filter (restates EVSL/pEVSL find_pol; called by the
// reference as pEVSL_FINDPOL_F90(XINTV, 0.8, 0.7, POL), src/mod_pevsl.f90:108-115).
// The spectrum [lmin, lmax] = xintv[2..3] is mapped to [-1, 1]; the filter is
// p(t) = sum_{k<=m} mu_k T_k(t), mu_k = g_k cos(k th_c) / p(cos th_c), with
// damping factors g_k (g_0 = 1/2; 1 = Jackson, 2 = Lanczos sigma).  An
// interval touching an end of the spectrum centres the delta there and uses
// thr_ext; an interior one balances p(a) = p(b) and uses thr_int.  The degree
// is the first m >= min_deg whose endpoint values do not exceed the threshold.
static void damping(int m, int type, std::vector<double>& g) {
  g.assign(m + 1, 1.0);
  g[0] = 0.5;
  const double dm = m;
  for (int k = 1; k <= m; ++k) {
    if (type == 1) {
      const double th = M_PI / (dm + 2.0), a1 = 1.0 / (dm + 2.0);
      g[k] = a1 * std::sin((k + 1) * th) / std::sin(th) + (1.0 - (k + 1) * a1) * std::cos(k * th);
    } else if (type == 2) {
      const double th = M_PI / (dm + 1.0);
      g[k] = std::sin(k * th) / (k * th);
    }
  }
}

static double cheb_series(const std::vector<double>& c, int m, double t) {
  // Clenshaw summation of sum_{k<=m} c_k T_k(t)
  double b1 = 0.0, b2 = 0.0;
  for (int k = m; k >= 1; --k) {
    const double b0 = 2.0 * t * b1 - b2 + c[k];
    b2 = b1;
    b1 = b0;
  }
  return c[0] + t * b1 - b2;
}

static void delta_series(int m, const std::vector<double>& g, double thc, std::vector<double>& mu) {
  mu.resize(m + 1);
  for (int k = 0; k <= m; ++k) mu[k] = g[k] * std::cos(k * thc);
  const double peak = cheb_series(mu, m, std::cos(thc));
  for (int k = 0; k <= m; ++k) mu[k] /= peak;
}

int find_pol(const double* xintv, double thr_int, double thr_ext, int min_deg, int max_deg,
             int damp, PolyFilter& pol) {
  const double lmin = xintv[2], lmax = xintv[3];
  const double a = std::max(xintv[0], lmin), b = std::min(xintv[1], lmax);
  if (!(b > a) || !(lmax > lmin)) return NMG_EINVAL;
  const double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  const double aa = std::max((a - cc) / dd, -1.0), bb = std::min((b - cc) / dd, 1.0);
  const double kEndTol = 0.0005;
  int type = 0;
  double thr = thr_int;
  if (aa - kEndTol <= -1.0) {
    type = 1;
    thr = thr_ext;
  } else if (bb + kEndTol >= 1.0) {
    type = 2;
    thr = thr_ext;
  }
  std::vector<double> g, mu;
  const double tha = std::acos(aa), thb = std::acos(bb);
  for (int m = min_deg; m <= max_deg; ++m) {
    damping(m, damp, g);
    double thc;
    if (type == 1) thc = M_PI;
    else if (type == 2) thc = 0.0;
    else {
      // balance the two endpoint values: bisection on the centre angle
      double lo = thb, hi = tha;
      for (int it = 0; it < 100 && hi - lo > 1e-15; ++it) {
        const double mid = 0.5 * (lo + hi);
        delta_series(m, g, mid, mu);
        const double diff = cheb_series(mu, m, aa) - cheb_series(mu, m, bb);
        if (diff > 0.0) hi = mid; else lo = mid;
      }
      thc = 0.5 * (lo + hi);
    }
    delta_series(m, g, thc, mu);
    const double va = cheb_series(mu, m, aa), vb = cheb_series(mu, m, bb);
    const double worst = type == 1 ? vb : (type == 2 ? va : std::max(va, vb));
    if (worst <= thr) {
      pol.deg = m;
      pol.type = type;
      pol.cc = cc;
      pol.dd = dd;
      pol.gam = std::cos(thc);
      pol.bar = type == 1 ? vb : (type == 2 ? va : std::min(va, vb));
      pol.aa = aa;
      pol.bb = bb;
      pol.mu = mu;
      return NMG_OK;
    }
  }
  return NMG_ENUM;
}

// ---------------------------------------------------------------------------
// Eigenvalues of the symmetric tridiagonal T (diagonal d[0..m-1], off-diagonal
// e[0..m-2]) by implicit QL iterations with a Wilkinson-type shift taken from
// the leading 2x2 of each unreduced block (the mirror image of the implicit
// QR sweep), ascending on return.
int tridiag_eigvals(int m, const double* din, const double* ein, std::vector<double>& w) {
  std::vector<double> d(din, din + m), e(m, 0.0);
  for (int i = 0; i + 1 < m; ++i) e[i] = ein[i];
  const double eps = 2.2e-16;
  int guard = 0;
  for (int l = 0; l < m; ++l) {
    for (;;) {
      int r = l;
      while (r < m - 1 && std::fabs(e[r]) > eps * (std::fabs(d[r]) + std::fabs(d[r + 1]))) ++r;
      if (r == l) break;
      if (++guard > 60 * m + 100) return NMG_ENUM;
      // shift from the leading 2x2 [d_l e_l; e_l d_{l+1}] closest to d_l
      const double tt = 0.5 * (d[l + 1] - d[l]);
      const double bb = e[l];
      const double den = tt + (tt >= 0 ? 1.0 : -1.0) * std::hypot(tt, bb);
      const double mu = d[l] - (den != 0.0 ? bb * bb / den : 0.0);
      // chase the bulge upward from the bottom of the block [l, r]
      double x = d[r] - mu, z = e[r - 1];
      for (int k = r; k > l; --k) {
        const double rr = std::hypot(x, z);
        const double c = rr == 0.0 ? 1.0 : x / rr, s = rr == 0.0 ? 0.0 : -z / rr;
        if (k < r) e[k] = rr;
        // rotate rows/cols (k, k-1)
        const double a11 = d[k], a12 = e[k - 1], a22 = d[k - 1];
        d[k] = c * c * a11 - 2 * c * s * a12 + s * s * a22;
        d[k - 1] = s * s * a11 + 2 * c * s * a12 + c * c * a22;
        e[k - 1] = c * s * (a11 - a22) + (c * c - s * s) * a12;
        if (k > l + 1) {
          x = e[k - 1];
          z = -s * e[k - 2];
          e[k - 2] = c * e[k - 2];
        }
      }
    }
  }
  std::sort(d.begin(), d.end());
  w.swap(d);
  return NMG_OK;
}

// Eigenvector of T for eigenvalue theta by two steps of inverse iteration
// (Gaussian elimination with partial pivoting on T - theta' I).
void tridiag_eigvec(int m, const double* d, const double* e, double theta, std::vector<double>& s) {
  s.assign(m, 1.0);
  if (m == 1) return;
  double scale = 0.0;
  for (int i = 0; i < m; ++i) scale = std::max(scale, std::fabs(d[i]) + (i + 1 < m ? std::fabs(e[i]) : 0));
  const double th = theta + 1e-13 * (scale > 0 ? scale : 1.0);
  for (int it = 0; it < 3; ++it) {
    // band LU of (T - th I) with row pivoting: upper has 2 super-diagonals
    std::vector<double> dg(m), u1(m, 0.0), u2(m, 0.0), lo(m, 0.0);
    std::vector<int> piv(m, 0);
    for (int i = 0; i < m; ++i) dg[i] = d[i] - th;
    for (int i = 0; i + 1 < m; ++i) u1[i] = e[i];
    std::vector<double> sub(m, 0.0);
    for (int i = 0; i + 1 < m; ++i) sub[i] = e[i];
    std::vector<double> rhs = s;
    for (int i = 0; i + 1 < m; ++i) {
      // rows i and i+1: [dg_i u1_i u2_i ; sub_i dg_{i+1} u1_{i+1}]
      if (std::fabs(sub[i]) > std::fabs(dg[i])) {
        std::swap(dg[i], sub[i]);
        std::swap(u1[i], dg[i + 1]);
        const double t2 = u2[i];
        u2[i] = u1[i + 1];
        u1[i + 1] = t2;
        std::swap(rhs[i], rhs[i + 1]);
        piv[i] = 1;
      }
      if (dg[i] == 0.0) dg[i] = 1e-300;
      const double f = sub[i] / dg[i];
      lo[i] = f;
      dg[i + 1] -= f * u1[i];
      u1[i + 1] -= f * u2[i];
      rhs[i + 1] -= f * rhs[i];
    }
    if (dg[m - 1] == 0.0) dg[m - 1] = 1e-300;
    std::vector<double> x(m);
    for (int i = m - 1; i >= 0; --i) {
      double t = rhs[i];
      if (i + 1 < m) t -= u1[i] * x[i + 1];
      if (i + 2 < m) t -= u2[i] * x[i + 2];
      x[i] = t / dg[i];
    }
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) s[i] = x[i] / nrm;
  }
}

void start_vector(int64_t n, uint64_t seed, double* v) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)i + 0x632BE59BD9B4E019ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    v[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// Kernel polynomial method.  With t = (x - cc)/dd on [-1, 1] and
//   sum_i delta(t - t_i) = 1/(pi sqrt(1-t^2)) sum_k (2 - [k=0]) T_k(t) tau_k,
// tau_k = sum_i T_k(t_i), the Jackson-damped count in [a, b] integrates in
// closed form through t = cos(theta):
//   count = 1/pi sum_k (2 - [k=0]) g_k tau_k I_k,
//   I_0 = theta_a - theta_b,  I_k = (sin k theta_a - sin k theta_b)/k,
// g_k = ((M+2-k) cos(k al) + sin(k al) cot(al))/(M+2), al = pi/(M+2)
// (Jackson kernel for M+1 moments; Weisse et al., Rev. Mod. Phys. 78, 2006).
double dos_count(int mdeg, const double* tau, double lmin, double lmax, double a, double b) {
  const double cc = 0.5 * (lmax + lmin), dd = 0.5 * (lmax - lmin);
  auto theta = [&](double x) { return std::acos(std::min(1.0, std::max(-1.0, (x - cc) / dd))); };
  const double ta = theta(a), tb = theta(b);
  const double al = M_PI / (mdeg + 2);
  double sum = tau[0] * (ta - tb);  // k = 0: g_0 = 1
  for (int k = 1; k <= mdeg; ++k) {
    const double g = ((mdeg + 2 - k) * std::cos(k * al) + std::sin(k * al) / std::tan(al)) / (mdeg + 2);
    sum += 2.0 * g * tau[k] * (std::sin(k * ta) - std::sin(k * tb)) / k;
  }
  return sum / M_PI;
}

int dos_slicer(int mdeg, const double* tau, double lmin, double lmax, double a, double b, int ns, double* sli) {
  if (ns < 1 || !(b > a)) return NMG_EINVAL;
  const double total = dos_count(mdeg, tau, lmin, lmax, a, b);
  sli[0] = a;
  sli[ns] = b;
  for (int i = 1; i < ns; ++i) {
    const double target = total * i / ns;
    double lo = sli[i - 1], hi = b;
    for (int it = 0; it < 100; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (dos_count(mdeg, tau, lmin, lmax, a, mid) < target) lo = mid;
      else hi = mid;
    }
    sli[i] = 0.5 * (lo + hi);
    if (!(sli[i] > sli[i - 1])) return NMG_ENUM;
  }
  return NMG_OK;
}

}  // namespace nmg
