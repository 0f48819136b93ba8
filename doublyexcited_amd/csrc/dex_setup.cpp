// Host-side construction of the discretised Hamiltonian pieces (run once per
// context): the 1-D sinc-DVR kinetic matrix and the regularised Coulomb
// table Tinv.  These are set-up costs (milliseconds), not the hot path.
#include "dex_internal.h"

#include <cmath>
#include <vector>

namespace dex {

// Dimensionless sinc-DVR kinetic symbol: T_ij = tau(|i-j|) / delta^2 with
// tau(0) = pi^2/6 and tau(m) = (-1)^m / m^2 (hammod.F90:248-258).
double kinetic_tau(int m) {
  if (m < 0) m = -m;
  if (m == 0) return M_PI * M_PI / 6.0;
  const double s = (m & 1) ? -1.0 : 1.0;
  return s / (double(m) * double(m));
}

// KEMATRIX (hammod.F90:235-260): column-major NN x NN, real.
void kematrix(int nn, double delta, double *t) {
  const double inv = 1.0 / (delta * delta);
  for (int j = 0; j < nn; ++j)
    for (int i = 0; i < nn; ++i) t[i + nn * j] = kinetic_tau(i - j) * inv;
}

// Cyclic Jacobi eigendecomposition of a real symmetric n x n matrix
// (column-major).  On return a holds garbage, w the eigenvalues and v the
// eigenvectors as columns.  Used for the 45 x 45 kinetic matrix of the
// Coulomb solve (the reference calls LAPACK DSYEV at tooth.f90:70).
static void jacobi_eig(int n, std::vector<double> &a, std::vector<double> &w,
                       std::vector<double> &v) {
  v.assign(size_t(n) * n, 0.0);
  for (int i = 0; i < n; ++i) v[i + n * i] = 1.0;
  auto A = [&](int i, int j) -> double & { return a[i + size_t(n) * j]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += A(i, j) * A(i, j);
        if (i != j) off += A(i, j) * A(i, j);
      }
    if (off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (std::fabs(apq) < 1e-300) continue;
        const double app = A(p, p), aqq = A(q, q);
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k + size_t(n) * p], vkq = v[k + size_t(n) * q];
          v[k + size_t(n) * p] = c * vkp - s * vkq;
          v[k + size_t(n) * q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = A(i, i);
}

// GETINVERSE (tooth.f90:14-101).  Output Tinv(-nout:nout)^3 with the first
// index fastest, i.e. tinv[(i+nout) + L*((j+nout) + L*(k+nout))], L = 2 nout + 1.
//   R(n) = delta^3 / |n delta| (0 at the origin) on the big grid [-nbig, nbig]^3,
//   nbig = nsmall * tfactor (RVECTOR, tooth.f90:104-173);
//   TR = -(T R)|small + 2 pi delta_0, with the kinetic sum over the big grid;
//   Q solves T3D Q = TR on the small grid through the 1-D eigenbasis of T;
//   Tinv = R + Q on the small grid, R elsewhere.
int getinverse(int nout, double delta, int nsmall, int tfactor, std::vector<double> &tinv) {
  const int nbig = nsmall * tfactor;
  if (nbig < nout || nsmall < 1) return -1;  // tooth.f90:38-40
  const int ns = 2 * nsmall + 1;
  const double d2 = delta * delta;
  auto rval = [&](long i, long j, long k) -> double {
    const long r2 = i * i + j * j + k * k;
    if (r2 == 0) return 0.0;
    return d2 / std::sqrt(double(r2));  // delta^3 / (delta |n|)
  };
  // T between a small-grid label and a big-grid label: tau(i-p)/delta^2.
  std::vector<double> tcol(size_t(2 * (nsmall + nbig) + 1));
  const int toff = nsmall + nbig;
  for (int m = -toff; m <= toff; ++m) tcol[m + toff] = kinetic_tau(m) / d2;
  // R(p, j, k) for p on the big grid and j, k on the small grid; R is symmetric
  // under permutations of its arguments, which covers all three sums below.
  const int nb = 2 * nbig + 1;
  std::vector<double> rs(size_t(nb) * ns * ns);
  for (int k = -nsmall; k <= nsmall; ++k)
    for (int j = -nsmall; j <= nsmall; ++j)
      for (int p = -nbig; p <= nbig; ++p)
        rs[(p + nbig) + size_t(nb) * ((j + nsmall) + size_t(ns) * (k + nsmall))] = rval(p, j, k);
  auto RS = [&](int p, int j, int k) -> double {
    return rs[(p + nbig) + size_t(nb) * ((j + nsmall) + size_t(ns) * (k + nsmall))];
  };
  std::vector<double> tr(size_t(ns) * ns * ns);
  auto TR = [&](int i, int j, int k) -> double & {
    return tr[(i + nsmall) + size_t(ns) * ((j + nsmall) + size_t(ns) * (k + nsmall))];
  };
  for (int k = -nsmall; k <= nsmall; ++k)
    for (int j = -nsmall; j <= nsmall; ++j)
      for (int i = -nsmall; i <= nsmall; ++i) {
        double acc = 0.0;
        for (int p = -nbig; p <= nbig; ++p) {
          acc += tcol[i - p + toff] * RS(p, j, k);
          acc += tcol[j - p + toff] * RS(p, i, k);
          acc += tcol[k - p + toff] * RS(p, i, j);
        }
        TR(i, j, k) = -acc;
      }
  TR(0, 0, 0) += 2.0 * M_PI;
  // 1-D eigenbasis of the small kinetic matrix.
  std::vector<double> a(size_t(ns) * ns), w, v;
  for (int q = 0; q < ns; ++q)
    for (int p = 0; p < ns; ++p) a[p + size_t(ns) * q] = kinetic_tau(p - q) / d2;
  jacobi_eig(ns, a, w, v);
  // Q = (V x V x V) [ (V^T x V^T x V^T) TR / (w_a + w_b + w_c) ]
  auto mode_mult = [&](const std::vector<double> &in, std::vector<double> &out, int axis,
                       bool transpose) {
    out.assign(in.size(), 0.0);
    const size_t sa = axis == 0 ? 1 : (axis == 1 ? size_t(ns) : size_t(ns) * ns);
    for (int k = 0; k < ns; ++k)
      for (int j = 0; j < ns; ++j)
        for (int i = 0; i < ns; ++i) {
          const int idx[3] = {i, j, k};
          const int o = idx[axis];
          const size_t base = size_t(i) + size_t(ns) * (j + size_t(ns) * k) - sa * o;
          double acc = 0.0;
          for (int l = 0; l < ns; ++l) {
            // out_o = sum_l M(o,l) in_l with M = V^T (transpose) or V
            const double mval = transpose ? v[l + size_t(ns) * o] : v[o + size_t(ns) * l];
            acc += mval * in[base + sa * l];
          }
          out[size_t(i) + size_t(ns) * (j + size_t(ns) * k)] = acc;
        }
  };
  std::vector<double> t1, t2;
  mode_mult(tr, t1, 0, true);
  mode_mult(t1, t2, 1, true);
  mode_mult(t2, t1, 2, true);
  for (int k = 0; k < ns; ++k)
    for (int j = 0; j < ns; ++j)
      for (int i = 0; i < ns; ++i) t1[i + size_t(ns) * (j + size_t(ns) * k)] /= (w[i] + w[j] + w[k]);
  mode_mult(t1, t2, 0, false);
  mode_mult(t2, t1, 1, false);
  mode_mult(t1, t2, 2, false);  // t2 = Q
  const int L = 2 * nout + 1;
  tinv.assign(size_t(L) * L * L, 0.0);
  for (int k = -nout; k <= nout; ++k)
    for (int j = -nout; j <= nout; ++j)
      for (int i = -nout; i <= nout; ++i) {
        double val = rval(i, j, k);
        if (std::abs(i) <= nsmall && std::abs(j) <= nsmall && std::abs(k) <= nsmall)
          val += t2[(i + nsmall) + size_t(ns) * ((j + nsmall) + size_t(ns) * (k + nsmall))];
        tinv[(i + nout) + size_t(L) * ((j + nout) + size_t(L) * (k + nout))] = val;
      }
  return 0;
}

}  // namespace dex
