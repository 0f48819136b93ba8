// Small dense complex linear algebra on the host, for the k x k matrices of
// the block eigensolver (k <= 4*NEV): Cholesky, triangular inverse, Hermitian
// eigendecomposition and the projected harmonic-Ritz pencil.  Column-major
// storage throughout.  The reference calls LAPACK (ZPOTRF, ZTRSM, ZGGEV); these
// are this library's own implementations of the textbook algorithms named at
// each function.
#include "dex_internal.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace dex {

using cd = std::complex<double>;

// Upper Cholesky factor A = R^H R (ZPOTRF 'U' semantics, lamod.f90:167).
// Returns 0 or the 1-based index of the first non-positive pivot.  Only the
// upper triangle of `a` is referenced; the strict lower triangle of the
// result is zeroed.
int zpotrf_upper(int n, cd *a, int lda) {
  for (int j = 0; j < n; ++j) {
    double djj = std::real(a[j + size_t(lda) * j]);
    for (int k = 0; k < j; ++k) djj -= std::norm(a[k + size_t(lda) * j]);
    if (!(djj > 0.0) || !std::isfinite(djj)) return j + 1;
    const double rjj = std::sqrt(djj);
    a[j + size_t(lda) * j] = rjj;
    for (int i = j + 1; i < n; ++i) {
      cd s = a[j + size_t(lda) * i];
      for (int k = 0; k < j; ++k) s -= std::conj(a[k + size_t(lda) * j]) * a[k + size_t(lda) * i];
      a[j + size_t(lda) * i] = s / rjj;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) a[i + size_t(lda) * j] = 0.0;
  return 0;
}

// In-place inverse of an upper triangular matrix (back substitution by columns).
void ztrtri_upper(int n, cd *r, int ldr) {
  std::vector<cd> inv(size_t(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    inv[j + size_t(n) * j] = 1.0 / r[j + size_t(ldr) * j];
    for (int i = j - 1; i >= 0; --i) {
      cd s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += r[i + size_t(ldr) * k] * inv[k + size_t(n) * j];
      inv[i + size_t(n) * j] = -s / r[i + size_t(ldr) * i];
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) r[i + size_t(ldr) * j] = inv[i + size_t(n) * j];
}

// Hermitian eigendecomposition by cyclic complex Jacobi rotations: each
// off-diagonal pair is first made real by a diagonal phase, then annihilated by
// a real plane rotation.  Eigenvalues ascending in w, eigenvectors in v.
void zheev_jacobi(int n, const cd *a_in, int lda, std::vector<double> &w, std::vector<cd> &v) {
  std::vector<cd> a(size_t(n) * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[i + size_t(n) * j] = a_in[i + size_t(lda) * j];
  v.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + size_t(n) * i] = 1.0;
  auto A = [&](int i, int j) -> cd & { return a[i + size_t(n) * j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += std::norm(A(i, j));
        if (i != j) off += std::norm(A(i, j));
      }
    if (off <= 1e-30 * tot || tot == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cd g = A(p, q);
        const double ag = std::abs(g);
        if (ag < 1e-300) continue;
        const cd ph = g / ag;  // e^{i phi}
        const double app = std::real(A(p, p)), aqq = std::real(A(q, q));
        const double tau = (aqq - app) / (2.0 * ag);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        // V = [[c, s*ph], [-s*conj(ph)... ]] restricted to (p, q): columns
        // v_p = c e_p - s conj(ph) e_q,  v_q = s ph e_p + c e_q
        const cd vpp = c, vqp = -s * std::conj(ph), vpq = s * ph, vqq = c;
        for (int k = 0; k < n; ++k) {  // A <- A V
          const cd akp = A(k, p), akq = A(k, q);
          A(k, p) = akp * vpp + akq * vqp;
          A(k, q) = akp * vpq + akq * vqq;
        }
        for (int k = 0; k < n; ++k) {  // A <- V^H A
          const cd apk = A(p, k), aqk = A(q, k);
          A(p, k) = std::conj(vpp) * apk + std::conj(vqp) * aqk;
          A(q, k) = std::conj(vpq) * apk + std::conj(vqq) * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const cd xkp = v[k + size_t(n) * p], xkq = v[k + size_t(n) * q];
          v[k + size_t(n) * p] = xkp * vpp + xkq * vqp;
          v[k + size_t(n) * q] = xkp * vpq + xkq * vqq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return std::real(A(x, x)) < std::real(A(y, y)); });
  w.resize(n);
  std::vector<cd> vs(size_t(n) * n);
  for (int j = 0; j < n; ++j) {
    w[j] = std::real(A(idx[j], idx[j]));
    for (int i = 0; i < n; ++i) vs[i + size_t(n) * j] = v[i + size_t(n) * idx[j]];
  }
  v.swap(vs);
}

// Eigenvalues and right eigenvectors of a general complex n x n matrix:
// Householder reduction to upper Hessenberg form, complex single-shift QR with
// Wilkinson shifts and deflation to Schur form T = Z^H A Z (Golub & Van Loan,
// Matrix Computations, 4th ed., Alg. 7.4.2 and 7.5), then back substitution
// on T and back-transformation by Z.  Returns 0, or 1 if QR did not converge.
int zgeev_right(int n, const cd *a_in, int lda, std::vector<cd> &lam, std::vector<cd> &vec) {
  std::vector<cd> h(size_t(n) * n), z(size_t(n) * n, 0.0);
  auto H = [&](int i, int j) -> cd & { return h[i + size_t(n) * j]; };
  auto Z = [&](int i, int j) -> cd & { return z[i + size_t(n) * j]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) H(i, j) = a_in[i + size_t(lda) * j];
  for (int i = 0; i < n; ++i) Z(i, i) = 1.0;
  double anorm = 0.0;
  for (size_t i = 0; i < h.size(); ++i) anorm = std::max(anorm, std::abs(h[i]));
  // Hessenberg reduction
  std::vector<cd> u(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha2 = 0.0;
    for (int i = k + 1; i < n; ++i) alpha2 += std::norm(H(i, k));
    const double alpha = std::sqrt(alpha2);
    if (alpha == 0.0) continue;
    const cd x0 = H(k + 1, k);
    const cd ph = std::abs(x0) > 0 ? x0 / std::abs(x0) : cd(1.0);
    for (int i = 0; i < n; ++i) u[i] = 0.0;
    for (int i = k + 1; i < n; ++i) u[i] = H(i, k);
    u[k + 1] += ph * alpha;
    double un2 = 0.0;
    for (int i = k + 1; i < n; ++i) un2 += std::norm(u[i]);
    if (un2 == 0.0) continue;
    const double beta = 2.0 / un2;
    for (int j = 0; j < n; ++j) {  // H <- (I - beta u u^H) H
      cd s = 0.0;
      for (int i = k + 1; i < n; ++i) s += std::conj(u[i]) * H(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) H(i, j) -= u[i] * s;
    }
    for (int i = 0; i < n; ++i) {  // H <- H (I - beta u u^H), Z <- Z (I - beta u u^H)
      cd s = 0.0, t = 0.0;
      for (int j = k + 1; j < n; ++j) {
        s += H(i, j) * u[j];
        t += Z(i, j) * u[j];
      }
      s *= beta;
      t *= beta;
      for (int j = k + 1; j < n; ++j) {
        H(i, j) -= s * std::conj(u[j]);
        Z(i, j) -= t * std::conj(u[j]);
      }
    }
    for (int i = k + 2; i < n; ++i) H(i, k) = 0.0;
  }
  // Shifted QR on the Hessenberg matrix
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    while (l > 0) {
      const double s = std::abs(H(l - 1, l - 1)) + std::abs(H(l, l));
      const double tol = eps * (s > 0 ? s : anorm);
      if (std::abs(H(l, l - 1)) <= tol) {
        H(l, l - 1) = 0.0;
        break;
      }
      --l;
    }
    if (l == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 100 * n) return 1;
    ++iter;
    cd mu;
    if (iter % 11 == 10) {
      mu = H(hi, hi) + cd(std::abs(std::real(H(hi, hi - 1))) + std::abs(H(hi - 1, hi - 2 >= 0 ? hi - 2 : 0)), 0.0);
    } else {
      const cd a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const cd tr = a + d, det = a * d - b * c;
      const cd disc = std::sqrt(tr * tr * 0.25 - det);
      const cd e1 = tr * 0.5 + disc, e2 = tr * 0.5 - disc;
      mu = std::abs(e1 - d) < std::abs(e2 - d) ? e1 : e2;
    }
    // QR step on rows/cols l..hi via Givens rotations, applied to the full matrix
    std::vector<double> cs(hi - l);
    std::vector<cd> sn(hi - l);
    for (int k = l; k <= hi; ++k) H(k, k) -= mu;
    for (int k = l; k < hi; ++k) {
      const cd x = H(k, k), y = H(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      double c;
      cd s;
      if (r == 0.0) {
        c = 1.0;
        s = 0.0;
      } else {
        c = std::abs(x) / r;
        const cd xph = std::abs(x) > 0 ? x / std::abs(x) : cd(1.0);
        s = xph * std::conj(y) / r;
      }
      cs[k - l] = c;
      sn[k - l] = s;
      // rows k, k+1: [c s; -conj(s) c]
      for (int j = k; j < n; ++j) {
        const cd hk = H(k, j), hk1 = H(k + 1, j);
        H(k, j) = c * hk + s * hk1;
        H(k + 1, j) = -std::conj(s) * hk + c * hk1;
      }
    }
    for (int k = l; k < hi; ++k) {
      const double c = cs[k - l];
      const cd s = sn[k - l];
      // columns k, k+1 times G^H
      const int rmax = std::min(k + 2, hi);
      for (int i = 0; i <= rmax; ++i) {
        const cd hk = H(i, k), hk1 = H(i, k + 1);
        H(i, k) = c * hk + std::conj(s) * hk1;
        H(i, k + 1) = -s * hk + c * hk1;
      }
      for (int i = 0; i < n; ++i) {
        const cd zk = Z(i, k), zk1 = Z(i, k + 1);
        Z(i, k) = c * zk + std::conj(s) * zk1;
        Z(i, k + 1) = -s * zk + c * zk1;
      }
    }
    for (int k = l; k <= hi; ++k) H(k, k) += mu;
  }
  lam.resize(n);
  for (int i = 0; i < n; ++i) lam[i] = H(i, i);
  // eigenvectors of the triangular T, then back-transform
  vec.assign(size_t(n) * n, 0.0);
  std::vector<cd> y(n);
  const double smin = std::max(eps * anorm, 1e-300);
  for (int k = 0; k < n; ++k) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    y[k] = 1.0;
    for (int i = k - 1; i >= 0; --i) {
      cd s = 0.0;
      for (int j = i + 1; j <= k; ++j) s += H(i, j) * y[j];
      cd den = H(i, i) - lam[k];
      if (std::abs(den) < smin) den = smin;
      y[i] = -s / den;
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      cd s = 0.0;
      for (int j = 0; j <= k; ++j) s += Z(i, j) * y[j];
      vec[i + size_t(n) * k] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    if (nrm > 0)
      for (int i = 0; i < n; ++i) vec[i + size_t(n) * k] /= nrm;
  }
  return 0;
}

// LU with partial pivoting, solve A X = B in place (B n x m).  Returns 0 or
// the 1-based index of a zero pivot.
static int zgesv(int n, std::vector<cd> a, int m, cd *b, int ldb) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(a[k + size_t(n) * k]);
    for (int i = k + 1; i < n; ++i)
      if (std::abs(a[i + size_t(n) * k]) > best) {
        best = std::abs(a[i + size_t(n) * k]);
        p = i;
      }
    if (best == 0.0) return k + 1;
    piv[k] = p;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k + size_t(n) * j], a[p + size_t(n) * j]);
      for (int j = 0; j < m; ++j) std::swap(b[k + size_t(ldb) * j], b[p + size_t(ldb) * j]);
    }
    for (int i = k + 1; i < n; ++i) {
      const cd f = a[i + size_t(n) * k] / a[k + size_t(n) * k];
      a[i + size_t(n) * k] = f;
      for (int j = k + 1; j < n; ++j) a[i + size_t(n) * j] -= f * a[k + size_t(n) * j];
      for (int j = 0; j < m; ++j) b[i + size_t(ldb) * j] -= f * b[k + size_t(ldb) * j];
    }
  }
  for (int j = 0; j < m; ++j)
    for (int i = n - 1; i >= 0; --i) {
      cd s = b[i + size_t(ldb) * j];
      for (int k = i + 1; k < n; ++k) s -= a[i + size_t(n) * k] * b[k + size_t(ldb) * j];
      b[i + size_t(ldb) * j] = s / a[i + size_t(n) * i];
    }
  return 0;
}

// Projected harmonic-Ritz pencil K y = mu M y of order k (leading block of
// leading dimension ld), the NEV pairs of smallest |mu| (SOLVE_GEVP + EIGSORT,
// gplmr.F90:395-439,640-660).  K = B^H B is Hermitian positive definite when
// the projected basis is independent, so with K = R^H R the pencil is
// equivalent to the standard problem C z = nu z, C = R^{-H} M R^{-1},
// nu = 1/mu, y = R^{-1} z; the smallest |mu| are the largest |nu|.  If K is not
// numerically definite the pencil is reduced through M^{-1} K instead.
// Ritz vectors are scaled to unit 2-norm (their scale is immaterial: the
// solver renormalises X and re-orthonormalises P).  Returns 0 or DEX_EGEVP.
int harmonic_pencil(int k, const cd *kp, const cd *mp, int ld, int nev, std::vector<cd> &coords,
                    std::vector<cd> &mu) {
  if (nev > k) return DEX_EGEVP;
  std::vector<cd> r(size_t(k) * k), c(size_t(k) * k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) r[i + size_t(k) * j] = kp[i + size_t(ld) * j];
  std::vector<cd> lam, vec;
  std::vector<double> key(k);
  bool chol = zpotrf_upper(k, r.data(), k) == 0;
  if (chol) {
    std::vector<cd> rinv = r;
    ztrtri_upper(k, rinv.data(), k);
    // C = R^{-H} M R^{-1}
    std::vector<cd> t(size_t(k) * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        cd s = 0.0;
        for (int l = 0; l <= j; ++l) s += mp[i + size_t(ld) * l] * rinv[l + size_t(k) * j];
        t[i + size_t(k) * j] = s;
      }
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        cd s = 0.0;
        for (int l = 0; l <= i; ++l) s += std::conj(rinv[l + size_t(k) * i]) * t[l + size_t(k) * j];
        c[i + size_t(k) * j] = s;
      }
    if (zgeev_right(k, c.data(), k, lam, vec) != 0) return DEX_EGEVP;
    for (int i = 0; i < k; ++i) key[i] = std::abs(lam[i]) > 0 ? 1.0 / std::abs(lam[i]) : HUGE_VAL;
    // y = R^{-1} z
    std::vector<cd> y(size_t(k) * k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        cd s = 0.0;
        for (int l = i; l < k; ++l) s += rinv[i + size_t(k) * l] * vec[l + size_t(k) * j];
        y[i + size_t(k) * j] = s;
      }
    vec.swap(y);
    for (int i = 0; i < k; ++i) lam[i] = std::abs(lam[i]) > 0 ? 1.0 / lam[i] : cd(HUGE_VAL);
  } else {
    std::vector<cd> m(size_t(k) * k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        m[i + size_t(k) * j] = mp[i + size_t(ld) * j];
        c[i + size_t(k) * j] = kp[i + size_t(ld) * j];
      }
    if (zgesv(k, m, k, c.data(), k) != 0) return DEX_EGEVP;
    if (zgeev_right(k, c.data(), k, lam, vec) != 0) return DEX_EGEVP;
    for (int i = 0; i < k; ++i) key[i] = std::abs(lam[i]);
  }
  for (int i = 0; i < k; ++i)
    if (!std::isfinite(key[i])) key[i] = HUGE_VAL;
  std::vector<int> idx(k);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  coords.assign(size_t(k) * nev, 0.0);
  mu.resize(nev);
  for (int j = 0; j < nev; ++j) {
    const int s = idx[j];
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm += std::norm(vec[i + size_t(k) * s]);
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) return DEX_EGEVP;
    for (int i = 0; i < k; ++i) coords[i + size_t(k) * j] = vec[i + size_t(k) * s] / nrm;
    mu[j] = lam[s];
  }
  return 0;
}

}  // namespace dex
