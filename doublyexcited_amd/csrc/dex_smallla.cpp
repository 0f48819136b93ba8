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

// Complex products written out (a b and conj(a) b) for the O(n^3) inner
// loops: std::complex's operator* adds an inf/NaN recovery branch per
// product, which the finite data here never needs.
static inline cd mul(const cd &a, const cd &b) {
  return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
static inline cd mulcj(const cd &a, const cd &b) {
  return cd(a.real() * b.real() + a.imag() * b.imag(), a.real() * b.imag() - a.imag() * b.real());
}

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

// Complex Schur form A = Z T Z^H of a general n x n matrix: Householder
// reduction to upper Hessenberg form, then complex single-shift QR with
// Wilkinson shifts and deflation (Golub & Van Loan, Matrix Computations, 4th
// ed., Alg. 7.4.2 and 7.5).  T is returned in h (upper triangular), Z in z.
// Returns 0, or 1 if QR did not converge.
static int zschur(int n, const cd *a_in, int lda, std::vector<cd> &h, std::vector<cd> &z, double &anorm) {
  h.assign(size_t(n) * n, 0.0);
  z.assign(size_t(n) * n, 0.0);
  auto H = [&](int i, int j) -> cd & { return h[i + size_t(n) * j]; };
  auto Z = [&](int i, int j) -> cd & { return z[i + size_t(n) * j]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) H(i, j) = a_in[i + size_t(lda) * j];
  for (int i = 0; i < n; ++i) Z(i, i) = 1.0;
  anorm = 0.0;
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
      for (int i = k + 1; i < n; ++i) s += mulcj(u[i], H(i, j));
      s *= beta;
      for (int i = k + 1; i < n; ++i) H(i, j) -= mul(u[i], s);
    }
    for (int i = 0; i < n; ++i) {  // H <- H (I - beta u u^H), Z <- Z (I - beta u u^H)
      cd s = 0.0, t = 0.0;
      for (int j = k + 1; j < n; ++j) {
        s += mul(H(i, j), u[j]);
        t += mul(Z(i, j), u[j]);
      }
      s *= beta;
      t *= beta;
      for (int j = k + 1; j < n; ++j) {
        H(i, j) -= mul(s, std::conj(u[j]));
        Z(i, j) -= mul(t, std::conj(u[j]));
      }
    }
    for (int i = k + 2; i < n; ++i) H(i, k) = 0.0;
  }
  // Shifted QR on the Hessenberg matrix.  Deflation uses the Ahues-Tisseur
  // criterion (SIAM J. Matrix Anal. Appl. 1997; as in LAPACK's ZLAHQR), which
  // also deflates multiple eigenvalues promptly; every 10th step without
  // deflation takes an exceptional shift.
  const double ulp = 2.220446049250313e-16;
  const double smlnum = 1e-300 * (double(n) / ulp);
  auto cabs1 = [](const cd &z) { return std::fabs(z.real()) + std::fabs(z.imag()); };
  const int itmax = 30 * std::max(10, n);
  const int kexsh = 10;
  int kdefl = 0;
  std::vector<double> cs(n);
  std::vector<cd> sn(n);
  int i = n - 1;
  while (i >= 0) {
    int l = 0;
    bool converged = false;
    for (int its = 0; its <= itmax; ++its) {
      int k;
      for (k = i; k > l; --k) {
        if (cabs1(H(k, k - 1)) <= smlnum) break;
        double tst = cabs1(H(k - 1, k - 1)) + cabs1(H(k, k));
        if (tst == 0.0) {
          if (k - 2 >= l) tst += std::fabs(H(k - 1, k - 2).real());
          if (k + 1 <= i) tst += std::fabs(H(k + 1, k).real());
        }
        if (std::fabs(H(k, k - 1).real()) <= ulp * tst) {
          const double ab = std::max(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
          const double ba = std::min(cabs1(H(k, k - 1)), cabs1(H(k - 1, k)));
          const double aa = std::max(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
          const double bb = std::min(cabs1(H(k, k)), cabs1(H(k - 1, k - 1) - H(k, k)));
          const double sden = aa + ab;
          if (ba * (ab / sden) <= std::max(smlnum, ulp * (bb * (aa / sden)))) break;
        }
      }
      l = k;
      if (l > 0) H(l, l - 1) = 0.0;
      if (l >= i) {
        converged = true;
        break;
      }
      ++kdefl;
      cd mu;
      if (kdefl % (2 * kexsh) == 0) {
        mu = H(i, i) + 0.75 * std::fabs(H(i, i - 1).real());
      } else if (kdefl % kexsh == 0) {
        mu = H(l, l) + 0.75 * std::fabs(H(l + 1, l).real());
      } else {
        // Eigenvalue of the trailing 2 x 2 block [[a, b], [c, d]] nearer d,
        // written as d - u^2 / (x + y) with x = (a - d)/2, u^2 = b c and
        // y = sqrt(x^2 + u^2) (no cancellation between nearly equal terms).
        mu = H(i, i);
        const cd u = std::sqrt(H(i - 1, i)) * std::sqrt(H(i, i - 1));
        double sc = cabs1(u);
        if (sc != 0.0) {
          const cd x = 0.5 * (H(i - 1, i - 1) - mu);
          const double sx = cabs1(x);
          sc = std::max(sc, sx);
          const cd xs = x / sc, us = u / sc;
          cd y = sc * std::sqrt(xs * xs + us * us);
          if (sx > 0.0) {
            const cd xn = x / sx;
            if (xn.real() * y.real() + xn.imag() * y.imag() < 0.0) y = -y;
          }
          mu -= u * (u / (x + y));
        }
      }
      // explicit shifted QR step on rows/cols l..i, applied to the full matrix
      for (int kk = l; kk <= i; ++kk) H(kk, kk) -= mu;
      for (int kk = l; kk < i; ++kk) {
        const cd x = H(kk, kk), y = H(kk + 1, kk);
        const double r = std::sqrt(std::norm(x) + std::norm(y));
        double c;
        cd sv;
        if (r == 0.0) {
          c = 1.0;
          sv = 0.0;
        } else {
          c = std::abs(x) / r;
          const cd xph = std::abs(x) > 0 ? x / std::abs(x) : cd(1.0);
          sv = xph * std::conj(y) / r;
        }
        cs[kk] = c;
        sn[kk] = sv;
        for (int j = kk; j < n; ++j) {
          const cd hk = H(kk, j), hk1 = H(kk + 1, j);
          H(kk, j) = c * hk + mul(sv, hk1);
          H(kk + 1, j) = c * hk1 - mulcj(sv, hk);
        }
      }
      for (int kk = l; kk < i; ++kk) {
        const double c = cs[kk];
        const cd sv = sn[kk];
        const int rmax = std::min(kk + 2, i);
        for (int r = 0; r <= rmax; ++r) {
          const cd hk = H(r, kk), hk1 = H(r, kk + 1);
          H(r, kk) = c * hk + mulcj(sv, hk1);
          H(r, kk + 1) = c * hk1 - mul(sv, hk);
        }
        for (int r = 0; r < n; ++r) {
          const cd zk = Z(r, kk), zk1 = Z(r, kk + 1);
          Z(r, kk) = c * zk + mulcj(sv, zk1);
          Z(r, kk + 1) = c * zk1 - mul(sv, zk);
        }
      }
      for (int kk = l; kk <= i; ++kk) H(kk, kk) += mu;
    }
    if (!converged) return 1;
    kdefl = 0;
    i = l - 1;
  }
  for (int j = 0; j < n; ++j)
    for (int r = j + 1; r < n; ++r) H(r, j) = 0.0;
  return 0;
}

// Exchanges the diagonal entries k and k + 1 of the Schur form (T, Z) with one
// plane rotation whose first column is the eigenvector of the 2 x 2 block for
// T(k+1, k+1) (Golub & Van Loan, Sec. 7.6.2).
static void schur_swap(int n, std::vector<cd> &t, std::vector<cd> &z, int k) {
  auto T = [&](int i, int j) -> cd & { return t[i + size_t(n) * j]; };
  const cd t11 = T(k, k), t22 = T(k + 1, k + 1), t12 = T(k, k + 1);
  const cd d = t22 - t11;
  const double r = std::sqrt(std::norm(t12) + std::norm(d));
  if (r == 0.0) return;  // equal entries, nothing to exchange
  const cd g1 = t12 / r, g2 = d / r;
  // G = [[g1, -conj(g2)], [g2, conj(g1)]];  T <- G^H T G,  Z <- Z G
  for (int j = 0; j < n; ++j) {
    const cd a = T(k, j), b = T(k + 1, j);
    T(k, j) = std::conj(g1) * a + std::conj(g2) * b;
    T(k + 1, j) = -g2 * a + g1 * b;
  }
  for (int i = 0; i < n; ++i) {
    const cd a = T(i, k), b = T(i, k + 1);
    T(i, k) = a * g1 + b * g2;
    T(i, k + 1) = -a * std::conj(g2) + b * std::conj(g1);
    const cd za = z[i + size_t(n) * k], zb = z[i + size_t(n) * (k + 1)];
    z[i + size_t(n) * k] = za * g1 + zb * g2;
    z[i + size_t(n) * (k + 1)] = -za * std::conj(g2) + zb * std::conj(g1);
  }
  T(k + 1, k) = 0.0;
  T(k, k) = t22;
  T(k + 1, k + 1) = t11;
}

// Right eigenvectors (unit 2-norm) for the leading p diagonal entries of the
// Schur form, by back substitution on T and multiplication by Z.  Entries of
// one cluster (contiguous, equal to within tolc relative) are treated as a
// multiple eigenvalue: each member keeps its own Schur direction instead of
// dividing by a vanishing difference, so the cluster's vectors span its
// invariant subspace with an orthonormal basis.
static void leading_eigvecs(int n, int p, const std::vector<cd> &t, const std::vector<cd> &z, double anorm,
                            std::vector<cd> &vec) {
  const double eps = 2.220446049250313e-16, tolc = 1e-9;
  auto T = [&](int i, int j) -> const cd & { return t[i + size_t(n) * j]; };
  std::vector<int> cl(p, 0);
  for (int i = 1; i < p; ++i) {
    const double sc = std::max(std::abs(T(i, i)), std::abs(T(i - 1, i - 1)));
    cl[i] = std::abs(T(i, i) - T(i - 1, i - 1)) <= tolc * sc ? cl[i - 1] : cl[i - 1] + 1;
  }
  const double smin = std::max(eps * anorm, 1e-300);
  vec.assign(size_t(n) * p, 0.0);
  std::vector<cd> y(p);
  for (int k = 0; k < p; ++k) {
    std::fill(y.begin(), y.end(), cd(0.0));
    y[k] = 1.0;
    for (int i = k - 1; i >= 0; --i) {
      if (cl[i] == cl[k]) continue;  // cluster mate: y_i = 0
      cd s = 0.0;
      for (int j = i + 1; j <= k; ++j) s += T(i, j) * y[j];
      cd den = T(i, i) - T(k, k);
      if (std::abs(den) < smin) den = smin;
      y[i] = -s / den;
    }
    double nrm = 0.0;
    for (int r = 0; r < n; ++r) {
      cd s = 0.0;
      for (int j = 0; j <= k; ++j) s += z[r + size_t(n) * j] * y[j];
      vec[r + size_t(n) * k] = s;
      nrm += std::norm(s);
    }
    nrm = std::sqrt(nrm);
    if (nrm > 0)
      for (int r = 0; r < n; ++r) vec[r + size_t(n) * k] /= nrm;
  }
}

// Eigenvalues and right eigenvectors of a general complex n x n matrix.
int zgeev_right(int n, const cd *a_in, int lda, std::vector<cd> &lam, std::vector<cd> &vec) {
  std::vector<cd> t, z;
  double anorm = 0.0;
  if (zschur(n, a_in, lda, t, z, anorm) != 0) return 1;
  lam.resize(n);
  for (int i = 0; i < n; ++i) lam[i] = t[i + size_t(n) * i];
  leading_eigvecs(n, n, t, z, anorm, vec);
  return 0;
}

// Selects the p eigenvalues of smallest key (key(lambda) from `key`), moves
// them, in that order, to the leading positions of the Schur form, and returns
// them with their eigenvectors.  Returns 0, or 1 if QR did not converge.
template <typename Key>
static int select_eigen(int n, const cd *a, int lda, int p, Key key, std::vector<cd> &lam,
                        std::vector<cd> &vec) {
  std::vector<cd> t, z;
  double anorm = 0.0;
  if (zschur(n, a, lda, t, z, anorm) != 0) return 1;
  for (int pos = 0; pos < p; ++pos) {
    int best = pos;
    double bk = key(t[pos + size_t(n) * pos]);
    for (int i = pos + 1; i < n; ++i) {
      const double ki = key(t[i + size_t(n) * i]);
      if (ki < bk) {
        bk = ki;
        best = i;
      }
    }
    for (int i = best; i > pos; --i) schur_swap(n, t, z, i - 1);
  }
  lam.resize(p);
  for (int i = 0; i < p; ++i) lam[i] = t[i + size_t(n) * i];
  leading_eigvecs(n, p, t, z, anorm, vec);
  return 0;
}
// Least squares min ||A Y - B||_F for a full-column-rank A (rows >= cols),
// by Householder QR (Golub & Van Loan 5.2.1, 5.3.3).  With lower >= 0, A has
// at most `lower` nonzero subdiagonals (a block Hessenberg matrix with upper
// triangular subdiagonal blocks of order k has lower = k) and every
// reflector spans lower + 1 rows; lower < 0: A is dense.  A (rows x cols, lda)
// and B (rows x nrhs, ldb) are overwritten; Y is returned in the leading cols
// rows of B.  Returns 0, or j + 1 when column j of the triangular factor is
// numerically zero.
int zlstsq(int rows, int cols, cd *a, int lda, cd *b, int ldb, int nrhs, int lower) {
  if (rows < cols) return -1;
  double amax = 0.0;
  for (int j = 0; j < cols; ++j)
    for (int i = 0; i < rows; ++i) amax = std::max(amax, std::abs(a[i + size_t(lda) * j]));
  std::vector<cd> v(rows);
  for (int j = 0; j < cols; ++j) {
    // rows j .. iend-1 of column j may be nonzero; reflections keep that band
    // (they combine rows j .. j + lower, which later columns already span)
    const int iend = lower < 0 ? rows : std::min(rows, j + lower + 1);
    double nrm2 = 0.0;
    for (int i = j; i < iend; ++i) nrm2 += std::norm(a[i + size_t(lda) * j]);
    const double nrm = std::sqrt(nrm2);
    if (nrm == 0.0) return j + 1;
    const cd x0 = a[j + size_t(lda) * j];
    const double ax0 = std::abs(x0);
    const cd ph = ax0 > 0.0 ? x0 / ax0 : cd(1.0, 0.0);
    const cd alpha = -ph * nrm;  // reflect x onto alpha e_1, |alpha| = ||x||
    for (int i = j; i < iend; ++i) v[i] = a[i + size_t(lda) * j];
    v[j] -= alpha;
    double vn2 = 0.0;
    for (int i = j; i < iend; ++i) vn2 += std::norm(v[i]);
    if (vn2 > 0.0) {
      // apply I - 2 v v^H / (v^H v) to the trailing columns of A and to B
      auto reflect = [&](cd *col) {
        cd dot = 0.0;
        for (int i = j; i < iend; ++i) dot += std::conj(v[i]) * col[i];
        const cd f = 2.0 * dot / vn2;
        for (int i = j; i < iend; ++i) col[i] -= f * v[i];
      };
      for (int jj = j + 1; jj < cols; ++jj) reflect(a + size_t(lda) * jj);
      for (int r = 0; r < nrhs; ++r) reflect(b + size_t(ldb) * r);
    }
    a[j + size_t(lda) * j] = alpha;
    for (int i = j + 1; i < iend; ++i) a[i + size_t(lda) * j] = 0.0;
  }
  for (int j = 0; j < cols; ++j)
    if (!(std::abs(a[j + size_t(lda) * j]) > 1e-14 * amax)) return j + 1;
  for (int r = 0; r < nrhs; ++r) {
    cd *y = b + size_t(ldb) * r;
    for (int i = cols - 1; i >= 0; --i) {
      cd acc = y[i];
      for (int jj = i + 1; jj < cols; ++jj) acc -= a[i + size_t(lda) * jj] * y[jj];
      y[i] = acc / a[i + size_t(lda) * i];
    }
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
  std::vector<cd> lam, vec;
  bool ok = false;
  // Route 1: K = R^H R, C = R^{-H} M R^{-1}, nu = 1/mu, y = R^{-1} z; the
  // wanted pairs (smallest |mu|) are those of largest |nu|.
  {
    std::vector<cd> r(size_t(k) * k), c(size_t(k) * k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) r[i + size_t(k) * j] = kp[i + size_t(ld) * j];
    if (zpotrf_upper(k, r.data(), k) == 0) {
      std::vector<cd> rinv = r;
      ztrtri_upper(k, rinv.data(), k);
      std::vector<cd> t(size_t(k) * k, 0.0);
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
          cd s = 0.0;
          for (int l = 0; l <= j; ++l) s += mul(mp[i + size_t(ld) * l], rinv[l + size_t(k) * j]);
          t[i + size_t(k) * j] = s;
        }
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
          cd s = 0.0;
          for (int l = 0; l <= i; ++l) s += mulcj(rinv[l + size_t(k) * i], t[l + size_t(k) * j]);
          c[i + size_t(k) * j] = s;
        }
      auto key = [](const cd &nu) { return std::abs(nu) > 0 ? 1.0 / std::abs(nu) : HUGE_VAL; };
      if (select_eigen(k, c.data(), k, nev, key, lam, vec) == 0) {
        std::vector<cd> y(size_t(k) * nev);
        for (int j = 0; j < nev; ++j)
          for (int i = 0; i < k; ++i) {
            cd s = 0.0;
            for (int l = i; l < k; ++l) s += rinv[i + size_t(k) * l] * vec[l + size_t(k) * j];
            y[i + size_t(k) * j] = s;
          }
        vec.swap(y);
        for (int j = 0; j < nev; ++j) lam[j] = std::abs(lam[j]) > 0 ? 1.0 / lam[j] : cd(HUGE_VAL);
        ok = true;
      }
    }
  }
  // Route 2: M^{-1} K by LU with partial pivoting.
  if (!ok) {
    std::vector<cd> m(size_t(k) * k), c(size_t(k) * k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        m[i + size_t(k) * j] = mp[i + size_t(ld) * j];
        c[i + size_t(k) * j] = kp[i + size_t(ld) * j];
      }
    if (zgesv(k, m, k, c.data(), k) != 0) return DEX_EGEVP;
    auto key = [](const cd &x) { return std::isfinite(std::abs(x)) ? std::abs(x) : HUGE_VAL; };
    if (select_eigen(k, c.data(), k, nev, key, lam, vec) != 0) return DEX_EGEVP;
  }
  coords.assign(size_t(k) * nev, 0.0);
  mu.resize(nev);
  for (int j = 0; j < nev; ++j) {
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm += std::norm(vec[i + size_t(k) * j]);
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) return DEX_EGEVP;
    for (int i = 0; i < k; ++i) coords[i + size_t(k) * j] = vec[i + size_t(k) * j] / nrm;
    mu[j] = lam[j];
  }
  return 0;
}

}  // namespace dex
