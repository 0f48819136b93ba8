// Internal declarations shared by the host sources and the HIP kernels.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../../include/dex.h"

namespace dex {

// Records msg as the calling thread's last error and returns status.
int set_error(int status, const char *msg);

// ---- host set-up (dex_setup.cpp) -----------------------------------------
double kinetic_tau(int m);
void kematrix(int nn, double delta, double *t);
int getinverse(int nout, double delta, int nsmall, int tfactor, std::vector<double> &tinv);

// ---- small dense LA (dex_smallla.cpp) -------------------------------------
int zpotrf_upper(int n, std::complex<double> *a, int lda);
void ztrtri_upper(int n, std::complex<double> *r, int ldr);
void zheev_jacobi(int n, const std::complex<double> *a, int lda, std::vector<double> &w,
                  std::vector<std::complex<double>> &v);
int zgeev_right(int n, const std::complex<double> *a, int lda, std::vector<std::complex<double>> &lam,
                std::vector<std::complex<double>> &vec);
int zlstsq(int rows, int cols, std::complex<double> *a, int lda, std::complex<double> *b, int ldb,
           int nrhs, int lower = -1);
int harmonic_pencil(int k, const std::complex<double> *kp, const std::complex<double> *mp, int ld,
                    int nev, std::vector<std::complex<double>> &coords,
                    std::vector<std::complex<double>> &mu);

// ---- device kernels (dex_kernels.hip) ---------------------------------------
struct Dbl2 {
  double x, y;
};

// One column of a tall block as seen by the block kernels: the value at row r
// is a[r] - sig * b[r] (b may be null).  Used to form A Q - sigma Q on the fly.
// In the Gram kernel, bslot >= 0 also stores the raw b[r] in that extra
// column slot (so Q is read once when both B and Q are operands), and a
// column with a == nullptr is one whose values another column supplies.
struct ColDesc {
  const Dbl2 *a;
  const Dbl2 *b;
  double sig_re, sig_im;
  int32_t bslot;
  int32_t fmt;  // storage of a and b (and of an output column): 0 complex128, 1 complex64
};

struct MatvecCoef {
  double s2_re, s2_im;   // e^{-2 i theta} / delta^2 multiplying the kinetic sum
  double c1_re, c1_im;   // e^{-i theta} multiplying the real potential
  double inv_d3;         // 1 / delta^3 (electron-electron term)
};

// Sets every kernel's function attributes and caches the occupancy figures the
// launchers use (once per process, before the first launch; thread-safe).
hipError_t prepare_kernels();

// Returns false if no fast (LDS two-pass) kernel exists for (nn, single).
bool fast_matvec_supported(int nn, bool single);
bool generic_matvec_supported(int nn);

// Arithmetic / storage of a matvec: FP64 throughout; FP32 with complex64 X
// and Y; FP32 arithmetic reading complex128 X and writing complex128 Y (the
// intermediate U is complex64 in both FP32 forms).
enum { kIoF64 = 0, kIoF32 = 1, kIoMixed = 2 };

// Up to kBatchMax columns of one matvec launch, by value in the kernel arguments.
constexpr int kBatchMax = 16;
struct ColPtrs {
  const void *p[kBatchMax];
};
struct ColPtrsW {
  void *p[kBatchMax];
};

// Fast path: U = K1 X (electron 1), then Y = s2 (U + K2 X) + c1 Vr X.
hipError_t launch_pass1(int nn, int io, const void *x, void *u, hipStream_t s);
// pass 1 on ncols (1..kBatchMax) columns in one launch: U_j = K1 X_j
hipError_t launch_pass1_batch(int nn, int io, const void *const *x, void *const *u, int ncols, hipStream_t s);
hipError_t launch_pass2(int nn, int io, const void *x, const void *u, void *y,
                        const double *vn, const double *tinv, const MatvecCoef &coef,
                        hipStream_t s);
// Generic FP64 path for any odd nn <= 27: six axis passes plus the potential.
hipError_t launch_matvec_generic(int nn, const Dbl2 *x, Dbl2 *y, const double *vn,
                                 const double *tinv, const MatvecCoef &coef, hipStream_t s);

// G(kl x kr, column-major) = L^H R summed over d rows, kl <= 64, kr <= 128.
// The first `share` columns of R are the L columns themselves (share <= kl);
// `descs` holds kl L descriptors then kr - share R descriptors.  With
// share > 0 the leading share x share block is Hermitian and only its upper
// 16 x 16 blocks are computed (the rest of that block is left unset).
// `partial` must hold gram_partial_elems(kl, kr) Dbl2.
int gram_blocks();
size_t gram_partial_elems(int kl, int kr);
hipError_t launch_gram(const ColDesc *descs, int kl, int kr, int share, int64_t d, Dbl2 *partial,
                       Dbl2 *out, hipStream_t s);

// out_j = beta_j out_j + sum_i in_i C(i, j) row by row; in-place aliasing of
// an output with an input is allowed.  `descs`: nin input descriptors then
// nout output pointers (as ColDesc.a); C (nin x nout) and beta (nout, may be
// null) are device arrays.  nin <= 64, nout <= 48.
hipError_t launch_combine(const ColDesc *descs, int nin, int nout, const Dbl2 *c,
                          const Dbl2 *beta, int64_t d, hipStream_t s);

// Per column j: sum |a_j|^2, sum conj(a_j) b_j, sum |b_j|^2 (b from descs[k + j]).
size_t coldots_partial_elems(int k);
hipError_t launch_coldots(const ColDesc *descs, int k, int64_t d, double *partial, double *out,
                          hipStream_t s);

// The GPLMR residual step per column j: X_j <- s_j X_j, AX_j <- s_j AX_j,
// W_j = AX_j - E_j X_j (scaled columns); out[j] = ||W_j||^2.  descs: X, AX, W
// (nev each); sc (nev reals) and ev (nev complex E_j) device arrays; `partial`
// holds coldots_partial_elems(nev) doubles.
hipError_t launch_resid(const ColDesc *descs, int nev, const double *sc, const Dbl2 *ev, int64_t d,
                        double *partial, double *out, hipStream_t s);

// Philox4x32-10 complex Gaussian fill: element i of column c gets
// (g0 + i g1)/sqrt(2) with g from Box-Muller on the counter (i, c, stream, 0).
hipError_t launch_fill_gaussian(Dbl2 *x, int64_t n, uint64_t seed, uint32_t col, uint32_t stream_id,
                                hipStream_t s);
// Exchange projection X <- (X + sign X^T)/2 in the M x M view (in place).
hipError_t launch_sym_project(Dbl2 *x, int64_t m, int sign, hipStream_t s);
// out(p, q) = sum_b T(p, b) X(q, b) (M x M from two M x K blocks)
hipError_t launch_lowrank(const Dbl2 *t, const Dbl2 *x, Dbl2 *out, int64_t m, int k, hipStream_t s);
// W = Y X (conj_x: Y conj(X)) for M x M Y and M x K X
hipError_t launch_ywx(const Dbl2 *y, const Dbl2 *x, Dbl2 *w, int64_t m, int k, int conj_x, hipStream_t s);
// Second matvec pass for X = sign X^T: Y = s2 (U + sign U^T) + c1 Vr o (X + sign X^T)/2.
hipError_t launch_symadd(int nn, int io, const void *x, const void *u, void *y,
                         const double *vn, const double *tinv, const MatvecCoef &coef, int sign,
                         hipStream_t s);
// the same on ncols (1..kBatchMax) columns in one launch
hipError_t launch_symadd_batch(int nn, int io, const void *const *x, const void *const *u, void *const *y,
                               int ncols, const double *vn, const double *tinv, const MatvecCoef &coef,
                               int sign, hipStream_t s);
// One-electron operator Y = s2 tau3D X + c1 vn o X on ncols columns of NN^3 entries.
hipError_t launch_h1(const Dbl2 *x, Dbl2 *y, int nn, int ncols, const double *vn, const MatvecCoef &coef,
                     hipStream_t s);
// out (p x q) = op(A)^T B over m rows, op = identity (herm 0) or conj (herm 1).
hipError_t launch_small_dots(const Dbl2 *a, int p, const Dbl2 *b, int q, int64_t m, int herm, Dbl2 *out,
                             hipStream_t s);
// Y = beta Y + A C (A m x p, C p x q, Y m x q; beta 0 or 1).
hipError_t launch_small_lincomb(const Dbl2 *a, int p, const Dbl2 *c, int q, int64_t m, int beta, Dbl2 *y,
                                hipStream_t s);
hipError_t launch_to_single(const Dbl2 *x, void *xs, int64_t n, hipStream_t s);

// ---- one-electron block Lanczos (dex_lanczos.cpp) --------------------------
// BLOCKLANCZOS (lanmod.F90:18-219) of H1 = s2 tau3D + c1 vn on the device:
// block size bb, at most kdim blocks, a convergence check every `check` steps,
// numeigs eigenvalues of smallest real part, Hermitian residual tolerance tol.
// Returns DEX_OK, or DEX_ENOTCONVERGED with the last check's values.
int heplus_lanczos(hipStream_t stream, int nn, const double *d_vn, const MatvecCoef &coef, int bb, int kdim,
                   int check, double tol, int numeigs, uint64_t seed, std::vector<std::complex<double>> &eig,
                   std::vector<std::complex<double>> &vecs, std::vector<double> &resid, int &iters);
hipError_t launch_to_double(const void *xs, Dbl2 *x, int64_t n, hipStream_t s);

}  // namespace dex
