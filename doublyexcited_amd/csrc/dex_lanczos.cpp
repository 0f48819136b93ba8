// One-electron (He+) block Lanczos on the device: BLOCKLANCZOS
// (lanmod.F90:18-219) over KEMULT1 / POTENTIAL1 (hammod.F90:71-134,432-498),
// the orbitals of the reference's product start vectors.
//
// The operator H1 = T3D / SCALE^2 + Vn / SCALE is complex symmetric and the
// Lanczos is the reference's complex-symmetric one: every inner product of the
// recurrence is the bilinear form x^T y (MPIDOT), the Ritz vectors are
// normalised and their residuals measured in the Hermitian norm (MPIHERMDOT),
// as lanmod.F90:163-175 does.  Per step J (lanmod.F90:89-130):
//   W = H1 V_J;  A_J = V_J^T W;
//   VNEW = W made bilinear-orthonormal to V_1..V_J and within itself, twice
//          (GSORTH: the reference's column Gram-Schmidt against VMAT, VOLD and
//          the earlier new columns, then V / sqrt(V^T V), repeated);
//   V_{J+1} = VNEW;  B_{J+1} = V_{J+1}^T W;
//   T(J, J) = A_J, T(J, J+1) = B_{J+1}^T, T(J, J-1) = B_J.
// Every CHECK steps the block tridiagonal T (J BB square) is diagonalised, the
// NUMEIGS eigenvalues of smallest real part are taken (EIGSORT on the real
// part), their Ritz vectors formed from the basis and their residuals
// ||H1 u - lambda u|| / ||u|| measured; all below TOL ends the run
// (lanmod.F90:132-195).
//
// Differences, each giving the same vectors in exact arithmetic: the
// projection against V_1..V_J is one block product per pass instead of a
// column loop (VOLD is V_J, already among them); the orthonormalisation
// within the new block is the complex-symmetric Cholesky factor of its
// bilinear Gram matrix, VNEW R^{-1} with R^T R = VNEW^T VNEW, which is what
// column Gram-Schmidt with bilinear normalisation computes.  The start block
// is seeded uniform [0, 1) real parts plus 1e-4 i uniform [0, 1) imaginary
// parts, as RANDOM_NUMBER fills it (lanmod.F90:76-79), from std::mt19937_64
// (the reference leaves its generator unseeded).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dex_internal.h"

namespace dex {
namespace {

using cd = std::complex<double>;

// Complex-symmetric Cholesky G = R^T R (R upper triangular, no conjugation)
// of the k x k column-major G, then R^{-1} (upper) in place of R.  Returns
// false on a zero pivot (a breakdown of the bilinear form).
bool csym_chol_inverse(int k, std::vector<cd> &g) {
  std::vector<cd> r(size_t(k) * k, 0.0);
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i <= j; ++i) {
      cd s = g[i + size_t(k) * j];
      for (int l = 0; l < i; ++l) s -= r[l + size_t(k) * i] * r[l + size_t(k) * j];
      if (i == j) {
        const cd d = std::sqrt(s);
        if (!(std::abs(d) > 0.0) || !std::isfinite(std::abs(d))) return false;
        r[i + size_t(k) * j] = d;
      } else {
        r[i + size_t(k) * j] = s / r[i + size_t(k) * i];
      }
    }
  }
  // inverse of the upper triangular r
  std::vector<cd> inv(size_t(k) * k, 0.0);
  for (int j = 0; j < k; ++j) {
    inv[j + size_t(k) * j] = 1.0 / r[j + size_t(k) * j];
    for (int i = j - 1; i >= 0; --i) {
      cd s = 0.0;
      for (int l = i + 1; l <= j; ++l) s += r[i + size_t(k) * l] * inv[l + size_t(k) * j];
      inv[i + size_t(k) * j] = -s / r[i + size_t(k) * i];
    }
  }
  g = inv;
  return true;
}

struct LanczosRun {
  hipStream_t s;
  int nn;
  int64_t m;
  const double *vn;
  MatvecCoef coef;
  Dbl2 *d_small = nullptr;  // p x q results and coefficient uploads

  int hip(hipError_t e, const char *where) {
    if (e == hipSuccess) return 0;
    return set_error(DEX_EHIP, (std::string(where) + ": " + hipGetErrorString(e)).c_str());
  }

  // out (p x q, host) = op(A)^T B
  int dots(const Dbl2 *a, int p, const Dbl2 *b, int q, int herm, std::vector<cd> &out) {
    out.assign(size_t(p) * q, 0.0);
    if (int st = hip(launch_small_dots(a, p, b, q, m, herm, d_small, s), "small dots")) return st;
    if (int st = hip(hipMemcpyAsync(out.data(), d_small, out.size() * sizeof(cd), hipMemcpyDeviceToHost, s),
                     "small dots copy"))
      return st;
    return hip(hipStreamSynchronize(s), "small dots sync");
  }

  // Y = beta Y + A C with C (p x q) from the host (the upload completes before
  // the call returns, so the caller's C may go; the launch stays queued)
  int lincomb(const Dbl2 *a, int p, const std::vector<cd> &cm, int q, int beta, Dbl2 *y) {
    if (int st = hip(hipMemcpyAsync(d_small, cm.data(), size_t(p) * q * sizeof(cd), hipMemcpyHostToDevice, s),
                     "coefficient upload"))
      return st;
    if (int st = hip(hipStreamSynchronize(s), "coefficient upload sync")) return st;
    return hip(launch_small_lincomb(a, p, d_small, q, m, beta, y, s), "small lincomb");
  }
};

}  // namespace

int heplus_lanczos(hipStream_t stream, int nn, const double *d_vn, const MatvecCoef &coef, int bb, int kdim,
                   int check, double tol, int numeigs, uint64_t seed, std::vector<cd> &eig,
                   std::vector<cd> &vecs, std::vector<double> &resid, int &iters) {
  const int64_t m = int64_t(nn) * nn * nn;
  if (bb < 1 || kdim < 2 || check < 1 || numeigs < 1 || numeigs > bb * (kdim - 1) || !(tol > 0))
    return set_error(DEX_EINVAL, "block Lanczos: need BB >= 1, KDIM >= 2, CHECK >= 1, "
                                 "1 <= NUMEIGS <= BB (KDIM - 1), TOL > 0");
  if (int64_t(bb) * kdim > m) return set_error(DEX_EINVAL, "block Lanczos: BB * KDIM exceeds the grid size M");
  LanczosRun L{stream, nn, m, d_vn, coef};
  const int64_t pmax = std::max<int64_t>(int64_t(bb) * kdim, numeigs);
  const size_t small_elems = size_t(pmax) * size_t(std::max(bb, numeigs));
  Dbl2 *vmat = nullptr, *w = nullptr, *vnew = nullptr, *tmpb = nullptr, *u = nullptr, *hu = nullptr;
  auto release = [&]() {
    for (Dbl2 *p : {vmat, w, vnew, tmpb, u, hu, L.d_small})
      if (p) (void)hipFree(p);
  };
  hipError_t e = hipMalloc(&vmat, size_t(m) * bb * kdim * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&w, size_t(m) * bb * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&vnew, size_t(m) * bb * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&tmpb, size_t(m) * bb * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&u, size_t(m) * numeigs * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&hu, size_t(m) * numeigs * sizeof(Dbl2));
  if (e == hipSuccess) e = hipMalloc(&L.d_small, small_elems * sizeof(Dbl2));
  if (e != hipSuccess) {
    release();
    return set_error(DEX_EOOM, "block Lanczos: device allocation failed");
  }
  int st = 0;
  auto col = [&](Dbl2 *base, int64_t j) { return base + j * m; };
  // bilinear-orthonormalise the bb columns at y against the first `nprev`
  // basis columns and within themselves, twice
  auto orthonormalise = [&](Dbl2 *y, int nprev) -> int {
    for (int pass = 0; pass < 2; ++pass) {
      if (nprev > 0) {
        std::vector<cd> c;
        if (int r = L.dots(vmat, nprev, y, bb, 0, c)) return r;
        for (cd &v : c) v = -v;
        if (int r = L.lincomb(vmat, nprev, c, bb, 1, y)) return r;
      }
      std::vector<cd> g;
      if (int r = L.dots(y, bb, y, bb, 0, g)) return r;
      if (!csym_chol_inverse(bb, g))
        return set_error(DEX_ECHOLESKY, "block Lanczos: the bilinear form broke down (zero pivot)");
      // y <- y R^{-1} through the scratch block
      if (int r = L.lincomb(y, bb, g, bb, 0, tmpb)) return r;
      if (int r = L.hip(hipMemcpyAsync(y, tmpb, size_t(m) * bb * sizeof(Dbl2), hipMemcpyDeviceToDevice, stream),
                        "block copy"))
        return r;
    }
    return 0;
  };
  // start block (lanmod.F90:76-86)
  {
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::vector<cd> v0(size_t(m) * bb);
    for (cd &z : v0) z = cd(uni(gen), 0.0);
    for (cd &z : v0) z += cd(0.0, 1e-4 * uni(gen));
    st = L.hip(hipMemcpyAsync(vmat, v0.data(), v0.size() * sizeof(cd), hipMemcpyHostToDevice, stream),
               "start block upload");
    if (st == 0) st = L.hip(hipStreamSynchronize(stream), "start block sync");
    if (st == 0) st = orthonormalise(vmat, 0);
  }
  const int ntot = bb * kdim;
  std::vector<cd> tdm(size_t(ntot) * ntot, 0.0);
  auto T = [&](int i, int j) -> cd & { return tdm[i + size_t(ntot) * j]; };
  std::vector<cd> bprev;  // B_J
  bool converged = false;
  iters = 0;
  for (int j = 1; st == 0 && j <= kdim - 1; ++j) {
    Dbl2 *vj = col(vmat, int64_t(j - 1) * bb);
    // W = H1 V_J; A_J = V_J^T W (lanmod.F90:89-96)
    st = L.hip(launch_h1(vj, w, nn, bb, d_vn, coef, stream), "H1 apply");
    std::vector<cd> a;
    if (st == 0) st = L.dots(vj, bb, w, bb, 0, a);
    // VNEW = W, orthonormalised (lanmod.F90:101-116)
    if (st == 0)
      st = L.hip(hipMemcpyAsync(vnew, w, size_t(m) * bb * sizeof(Dbl2), hipMemcpyDeviceToDevice, stream),
                 "W copy");
    if (st == 0) st = orthonormalise(vnew, j * bb);
    if (st) break;
    // V_{J+1} = VNEW (the basis keeps it; vmat has room for KDIM blocks)
    Dbl2 *vj1 = col(vmat, int64_t(j) * bb);
    st = L.hip(hipMemcpyAsync(vj1, vnew, size_t(m) * bb * sizeof(Dbl2), hipMemcpyDeviceToDevice, stream),
               "basis store");
    // B_{J+1} = V_{J+1}^T W (lanmod.F90:120-124)
    std::vector<cd> b;
    if (st == 0) st = L.dots(vj1, bb, w, bb, 0, b);
    if (st) break;
    const int o = (j - 1) * bb;
    for (int jb = 0; jb < bb; ++jb)
      for (int ib = 0; ib < bb; ++ib) {
        T(o + ib, o + jb) = a[ib + size_t(bb) * jb];                    // TDM(:,J,:,J) = A
        if (j < kdim) T(o + ib, o + bb + jb) = b[jb + size_t(bb) * ib];  // TDM(:,J,:,J+1) = B^T
        if (j > 1) T(o + ib, o - bb + jb) = bprev[ib + size_t(bb) * jb];  // TDM(:,J,:,J-1) = B_J
      }
    bprev = b;
    iters = j;
    if (j % check != 0) continue;
    // diagonalise the leading J BB block (ZGEEV, lanmod.F90:139-147)
    const int n = j * bb;
    std::vector<cd> t(size_t(n) * n);
    for (int cc = 0; cc < n; ++cc)
      for (int r = 0; r < n; ++r) t[r + size_t(n) * cc] = T(r, cc);
    std::vector<cd> lam, z;
    if (zgeev_right(n, t.data(), n, lam, z) != 0) {
      st = set_error(DEX_EGEVP, "block Lanczos: eigenproblem of the block tridiagonal matrix failed");
      break;
    }
    // EIGSORT on the real part (lanmod.F90:148, 262-284): first minimum on ties
    std::vector<int> ind(n);
    for (int i = 0; i < n; ++i) ind[i] = i;
    std::stable_sort(ind.begin(), ind.end(), [&](int x, int y) { return lam[x].real() < lam[y].real(); });
    ind.resize(numeigs);
    // Ritz vectors U_k = VMAT Z(:, k), Hermitian-normalised; residuals (lanmod.F90:155-175)
    std::vector<cd> zc(size_t(n) * numeigs);
    for (int k = 0; k < numeigs; ++k)
      for (int r = 0; r < n; ++r) zc[r + size_t(n) * k] = z[r + size_t(n) * ind[k]];
    st = L.lincomb(vmat, n, zc, numeigs, 0, u);
    std::vector<cd> un;
    if (st == 0) st = L.dots(u, numeigs, u, numeigs, 1, un);
    if (st) break;
    std::vector<cd> sc(size_t(numeigs) * numeigs, 0.0);
    for (int k = 0; k < numeigs; ++k) sc[k + size_t(numeigs) * k] = 1.0 / std::sqrt(un[k + size_t(numeigs) * k].real());
    st = L.lincomb(u, numeigs, sc, numeigs, 0, hu);  // hu <- normalised u
    if (st == 0)
      st = L.hip(hipMemcpyAsync(u, hu, size_t(m) * numeigs * sizeof(Dbl2), hipMemcpyDeviceToDevice, stream),
                 "Ritz copy");
    if (st == 0) st = L.hip(launch_h1(u, hu, nn, numeigs, d_vn, coef, stream), "H1 apply");
    // hu <- H1 u - lambda u
    std::vector<cd> dl(size_t(numeigs) * numeigs, 0.0);
    for (int k = 0; k < numeigs; ++k) dl[k + size_t(numeigs) * k] = -lam[ind[k]];
    if (st == 0) st = L.lincomb(u, numeigs, dl, numeigs, 1, hu);
    std::vector<cd> rr;
    if (st == 0) st = L.dots(hu, numeigs, hu, numeigs, 1, rr);
    if (st) break;
    eig.assign(numeigs, 0.0);
    resid.assign(numeigs, 0.0);
    double maxerr = 0.0;
    for (int k = 0; k < numeigs; ++k) {
      eig[k] = lam[ind[k]];
      resid[k] = std::sqrt(std::max(rr[k + size_t(numeigs) * k].real(), 0.0));
      maxerr = std::max(maxerr, resid[k]);
    }
    if (maxerr < tol) {
      converged = true;
      break;
    }
  }
  if (st == 0 && !eig.empty()) {
    vecs.resize(size_t(m) * numeigs);
    st = L.hip(hipMemcpyAsync(vecs.data(), u, vecs.size() * sizeof(cd), hipMemcpyDeviceToHost, stream),
               "Ritz vectors copy");
    if (st == 0) st = L.hip(hipStreamSynchronize(stream), "Ritz vectors sync");
  }
  release();
  if (st) return st;
  if (eig.empty()) return set_error(DEX_ENOTCONVERGED, "LANCZOS FAILED TO CONVERGE (no check step reached)");
  return converged ? DEX_OK : set_error(DEX_ENOTCONVERGED, "LANCZOS FAILED TO CONVERGE.");
}

}  // namespace dex
