/*
 * CPU oracle (test infrastructure only): a C restatement of the reference's
 * two-electron HAMMULT on one process, used by tests/ as a checker and by
 * bench.py's cpu_baseline leg as the timed CPU path.  The product never links
 * or calls it.
 *
 * It follows the reference's algorithm step by step
 * (complexscale-djhnow/code/hammod.F90):
 *   HAMMULT  (:19-67):   Y = KEMULT2(X) + V * X
 *   KEMULT2  (:138-187): Y(:,J) = KEMULTS(X(:,J)) for every column J
 *                        (OpenMP over columns, :164-168); W = X^T
 *                        (MPITRANSPOSE with NP = 1, :898-899); Z(:,J) =
 *                        KEMULTS(W(:,J)); Y = Y + Z^T (:182-185)
 *   KEMULTS  (:264-286): three complex SUMs over the 1-D kinetic matrix T
 * with T(NN,NN) = KEMATRIX / SCALE^2 and V = POTENTIAL2 / SCALE given by the
 * caller (built by oracle/hamiltonian.py, which restates KEMATRIX,
 * GETINVERSE and POTENTIAL2).  Complex numbers are interleaved doubles.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

/* KEMULTS: Y(I,J,K) = SUM(T(:,I)*X(:,J,K)) + SUM(T(:,J)*X(I,:,K)) + SUM(T(:,K)*X(I,J,:)) */
static void kemults(const zc *t, const zc *x, zc *y, int nn) {
  const int n2 = nn * nn;
  for (int k = 0; k < nn; ++k)
    for (int j = 0; j < nn; ++j)
      for (int i = 0; i < nn; ++i) {
        zc s1 = 0, s2 = 0, s3 = 0;
        for (int l = 0; l < nn; ++l) s1 += t[l + nn * i] * x[l + nn * j + n2 * k];
        for (int l = 0; l < nn; ++l) s2 += t[l + nn * j] * x[i + nn * l + n2 * k];
        for (int l = 0; l < nn; ++l) s3 += t[l + nn * k] * x[i + nn * j + n2 * l];
        y[i + nn * j + n2 * k] = s1 + s2 + s3;
      }
}

static void transpose(const zc *x, zc *w, int64_t m) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < m; ++j)
    for (int64_t i = 0; i < m; ++i) w[j + m * i] = x[i + m * j];
}

/* Y = H X for `howmany` vectors (MYMATMUL, doublyexcited.F90:24-44).
 * t: NN x NN complex (column-major), v: D complex, x/y: howmany * D complex.
 * work: 2 * D complex scratch (the reference's SAVEd W, Z of KEMULT2).
 * nthreads <= 0 keeps the OpenMP default.  Returns 0. */
int oracle_hammult(int nn, const double *t_ri, const double *v_ri, const double *x_ri, double *y_ri,
                   int howmany, double *work_ri, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int64_t m = (int64_t)nn * nn * nn;
  const int64_t d = m * m;
  const zc *t = (const zc *)t_ri;
  const zc *v = (const zc *)v_ri;
  zc *w = (zc *)work_ri;
  zc *z = w + d;
  for (int h = 0; h < howmany; ++h) {
    const zc *x = (const zc *)x_ri + h * d;
    zc *y = (zc *)y_ri + h * d;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) kemults(t, x + m * j, y + m * j, nn);
    transpose(x, w, m);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < m; ++j) kemults(t, w + m * j, z + m * j, nn);
    transpose(z, w, m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d; ++i) y[i] = y[i] + w[i] + v[i] * x[i];
  }
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
