// Host solvers for the small projected problems of the Krylov loops:
// symmetric tridiagonal eigenvalues by the implicit QL method with Wilkinson
// shifts (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293,
// the "tql1/tql2" procedures), eigenvectors by accumulating the rotations, and
// single eigenvectors by inverse iteration on the shifted tridiagonal.
// Sizes are the Krylov window (m <= a few hundred), so this is O(m^2)-O(m^3)
// host work beside the O(m·N) device work.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "qmbed.hpp"

namespace qmbed {

namespace {

// QL with implicit shifts. d: diagonal (in/out -> eigenvalues, unsorted),
// e: subdiagonal e[0..n-2] (destroyed). z (optional, n×n row-major) receives
// eigenvectors as columns when non-null (initialised to identity by caller).
void tql(int n, std::vector<double>& d, std::vector<double> e, double* z) {
  if (n <= 1) return;
  e.resize(n);
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 60) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z) {
            for (int k = 0; k < n; ++k) {
              double zk1 = z[size_t(k) * n + i + 1];
              double zk = z[size_t(k) * n + i];
              z[size_t(k) * n + i + 1] = s * zk + c * zk1;
              z[size_t(k) * n + i] = c * zk - s * zk1;
            }
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

}  // namespace

void tridiagonal_eigen(const std::vector<double>& alpha, const std::vector<double>& beta, std::vector<double>& values,
                       std::vector<double>* vecs) {
  const int n = int(alpha.size());
  std::vector<double> d = alpha;
  std::vector<double> e(beta.begin(), beta.begin() + std::max(0, n - 1));
  std::vector<double> z;
  if (vecs) {
    z.assign(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) z[size_t(i) * n + i] = 1.0;
  }
  tql(n, d, e, vecs ? z.data() : nullptr);
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  values.resize(n);
  for (int j = 0; j < n; ++j) values[j] = d[order[j]];
  if (vecs) {
    vecs->assign(size_t(n) * n, 0.0);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (*vecs)[size_t(i) * n + j] = z[size_t(i) * n + order[j]];
  }
}

namespace {

// number of eigenvalues below x: negative pivots of the LDL^T factorization
// of T - xI (Sturm sequence), a pivot within pivmin of zero taken negative
int count_below(const std::vector<double>& a, const std::vector<double>& b2, double x, double pivmin) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < a.size(); ++i) {
    q = (a[i] - x) - (i ? b2[i - 1] / q : 0.0);
    if (std::fabs(q) <= pivmin) q = -pivmin;
    if (q < 0) ++cnt;
  }
  return cnt;
}

}  // namespace

// Eigenvalues lo..hi-1 (ascending order, 0 = smallest) of the symmetric
// tridiagonal (alpha, beta) by bisection on Sturm counts (Barth, Martin &
// Wilkinson, Numer. Math. 9 (1967) 386), each to a few ulps of the
// spectrum's scale: O(n) per bisection step instead of the O(n^2) of a full
// QL sweep when only the extreme Ritz values are wanted.
std::vector<double> tridiagonal_eigenvalue_range(const std::vector<double>& alpha, const std::vector<double>& beta,
                                                 int lo, int hi) {
  const int n = int(alpha.size());
  lo = std::max(lo, 0);
  hi = std::min(hi, n);
  std::vector<double> out;
  if (lo >= hi) return out;
  std::vector<double> b2(size_t(std::max(n - 1, 0)));
  double bmax = 0;
  for (int i = 0; i + 1 < n; ++i) {
    b2[i] = beta[i] * beta[i];
    bmax = std::max(bmax, b2[i]);
  }
  const double pivmin = 2.2250738585072014e-308 * std::max(1.0, bmax);
  double gl = INFINITY, gu = -INFINITY;
  for (int i = 0; i < n; ++i) {
    const double r = (i ? std::fabs(beta[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(beta[i]) : 0.0);
    gl = std::min(gl, alpha[i] - r);
    gu = std::max(gu, alpha[i] + r);
  }
  const double eps = 2.220446049250313e-16;
  const double tnorm = std::max(std::fabs(gl), std::fabs(gu));
  gl -= 2.0 * eps * tnorm * n + 2.0 * pivmin;
  gu += 2.0 * eps * tnorm * n + 2.0 * pivmin;
  double prev = gl;
  for (int j = lo; j < hi; ++j) {
    double l = prev, u = gu;  // eigenvalue j is not below eigenvalue j - 1
    for (int it = 0; it < 200; ++it) {
      if (u - l <= 2.0 * eps * std::max(std::fabs(l), std::fabs(u)) + 4.0 * pivmin) break;
      const double mid = 0.5 * (l + u);
      if (mid <= l || mid >= u) break;
      if (count_below(alpha, b2, mid, pivmin) > j) u = mid;
      else l = mid;
    }
    const double v = 0.5 * (l + u);
    out.push_back(v);
    prev = l;
  }
  return out;
}

// Eigenvector of the symmetric tridiagonal (alpha, beta) for an eigenvalue
// lam, by inverse iteration with partial pivoting (LAPACK dstein-style).
std::vector<double> tridiagonal_eigvec(const std::vector<double>& alpha, const std::vector<double>& beta, double lam) {
  const int n = int(alpha.size());
  std::vector<double> x(n, 1.0 / std::sqrt(double(std::max(n, 1))));
  if (n == 1) return {1.0};
  double scale = 0;
  for (int i = 0; i < n; ++i) scale = std::max(scale, std::fabs(alpha[i]));
  for (int i = 0; i + 1 < n; ++i) scale = std::max(scale, std::fabs(beta[i]));
  const double pert = std::max(scale, 1.0) * 1e-14;
  // LU of (T - lam I) with partial pivoting: rows i, i+1 can swap; store
  // U with up to two superdiagonals.
  std::vector<double> dl(n), dd(n), du(n), du2(n);
  std::vector<int> piv(n);
  for (int i = 0; i < n; ++i) dd[i] = alpha[i] - lam;
  for (int i = 0; i + 1 < n; ++i) {
    dl[i] = beta[i];
    du[i] = beta[i];
  }
  for (int i = 0; i < n; ++i) du2[i] = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      piv[i] = 0;
      if (dd[i] == 0.0) dd[i] = pert;
      double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
      du2[i] = 0.0;
    } else {
      piv[i] = 1;
      double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  if (dd[n - 1] == 0.0) dd[n - 1] = pert;
  for (int i = 0; i < n; ++i)
    if (std::fabs(dd[i]) < pert) dd[i] = dd[i] < 0 ? -pert : pert;
  for (int it = 0; it < 3; ++it) {
    // forward (L)
    for (int i = 0; i + 1 < n; ++i) {
      if (piv[i] == 0) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        double t = x[i];
        x[i] = x[i + 1];
        x[i + 1] = t - dl[i] * x[i + 1];
      }
    }
    // backward (U)
    x[n - 1] /= dd[n - 1];
    if (n >= 2) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / dd[n - 2];
    for (int i = n - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
    double nrm = 0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (double& v : x) v /= nrm;
  }
  return x;
}

// Dense Hermitian eigen-decomposition for the projected matrices of the
// thick-restarted Lanczos (orders up to a few hundred): Householder
// reduction to tridiagonal form, a diagonal phase that makes the
// off-diagonal real and non-negative, the implicit QL iteration above, and
// back-transformation. a: row-major n×n (only its Hermitian part is used);
// values ascending; vectors row-major n×n with column j the unit eigenvector
// of values[j].
void hermitian_eigen(int n, std::vector<cplx> a, std::vector<double>& values, std::vector<cplx>& vectors) {
  auto at = [&](int i, int j) -> cplx& { return a[size_t(i) * n + j]; };
  for (int i = 0; i < n; ++i)  // symmetrize: the Hermitian part of a
    for (int j = i; j < n; ++j) {
      const cplx h = 0.5 * (at(i, j) + std::conj(at(j, i)));
      at(i, j) = h;
      at(j, i) = std::conj(h);
    }
  std::vector<cplx> q(size_t(n) * n, cplx(0, 0));
  for (int i = 0; i < n; ++i) q[size_t(i) * n + i] = 1.0;
  std::vector<cplx> u(n), p(n), qu(n);
  for (int k = 0; k + 2 < n; ++k) {
    // reflector H = I - 2 u u^H (u on rows k+1..n-1) with H x = alpha e_{k+1}
    const int m0 = k + 1;
    double sigma2 = 0;
    for (int i = m0; i < n; ++i) sigma2 += std::norm(at(i, k));
    const double sigma = std::sqrt(sigma2);
    double below = 0;
    for (int i = m0 + 1; i < n; ++i) below += std::norm(at(i, k));
    if (below == 0.0) continue;  // already tridiagonal in this column
    const cplx x0 = at(m0, k);
    const double ax0 = std::abs(x0);
    const cplx ph = ax0 > 0 ? x0 / ax0 : cplx(1, 0);
    const cplx alpha = -ph * sigma;
    for (int i = m0; i < n; ++i) u[i] = at(i, k);
    u[m0] -= alpha;
    double un = 0;
    for (int i = m0; i < n; ++i) un += std::norm(u[i]);
    un = std::sqrt(un);
    for (int i = m0; i < n; ++i) u[i] /= un;
    // trailing block: A <- A - 2 u q^H - 2 q u^H, p = A u, q = p - (u^H p) u
    for (int i = m0; i < n; ++i) {
      cplx acc(0, 0);
      for (int j = m0; j < n; ++j) acc += at(i, j) * u[j];
      p[i] = acc;
    }
    cplx mu(0, 0);
    for (int i = m0; i < n; ++i) mu += std::conj(u[i]) * p[i];
    const double mur = mu.real();
    for (int i = m0; i < n; ++i) p[i] -= mur * u[i];
    for (int i = m0; i < n; ++i)
      for (int j = m0; j < n; ++j) at(i, j) -= 2.0 * (u[i] * std::conj(p[j]) + p[i] * std::conj(u[j]));
    at(m0, k) = alpha;
    at(k, m0) = std::conj(alpha);
    for (int i = m0 + 1; i < n; ++i) {
      at(i, k) = 0;
      at(k, i) = 0;
    }
    // Q <- Q H on columns m0..n-1
    for (int r = 0; r < n; ++r) {
      cplx acc(0, 0);
      for (int j = m0; j < n; ++j) acc += q[size_t(r) * n + j] * u[j];
      qu[r] = acc;
    }
    for (int r = 0; r < n; ++r)
      for (int j = m0; j < n; ++j) q[size_t(r) * n + j] -= 2.0 * qu[r] * std::conj(u[j]);
  }
  // T = Q^H A Q: real diagonal, complex subdiagonal t_i = T[i+1][i]; with
  // D_0 = 1, D_{i+1} = D_i t_i / |t_i|, D^H T D is real symmetric with
  // off-diagonal |t_i|
  std::vector<double> d(n), e(std::max(n - 1, 0));
  std::vector<cplx> ph(n, cplx(1, 0));
  for (int i = 0; i < n; ++i) d[i] = at(i, i).real();
  for (int i = 0; i + 1 < n; ++i) {
    const cplx t = at(i + 1, i);
    const double at_ = std::abs(t);
    e[i] = at_;
    ph[i + 1] = at_ > 0 ? ph[i] * (t / at_) : ph[i];
  }
  std::vector<double> z(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) z[size_t(i) * n + i] = 1.0;
  tql(n, d, e, z.data());
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  values.resize(n);
  vectors.assign(size_t(n) * n, cplx(0, 0));
  for (int jj = 0; jj < n; ++jj) {
    const int j = order[jj];
    values[jj] = d[j];
    // eigenvector Q D z_j
    for (int r = 0; r < n; ++r) {
      cplx acc(0, 0);
      for (int i = 0; i < n; ++i) acc += q[size_t(r) * n + i] * ph[i] * z[size_t(i) * n + j];
      vectors[size_t(r) * n + jj] = acc;
    }
  }
}

}  // namespace qmbed
