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

/* POTENTIAL2 (hammod.F90:502-582) with the 1/SCALE of initmod.F90:224, one
 * nucleus of charge z at the origin grid point:
 *   V(P,Q) = [A Tinv(r_Q) + A Tinv(r_P) + Tinv(r_Q - r_P) / delta^3] / SCALE,
 * A = -z / delta^3 (hammod.F90:554-555,571-573), r = (i, j, k) - N with i
 * fastest (hammod.F90:560-562).  tinv: (2 nout + 1)^3 reals indexed
 * [i + nout] fastest (the reference's TINV(-NOUT:NOUT)^3, nout = NN here,
 * initmod.F90:39); v: D complex, V(P,Q) at P + M Q.  scale = (cos, sin)(theta). */
int oracle_potential2(int nn, int nout, const double *tinv, double z, double delta, double scale_re,
                      double scale_im, double *v_ri) {
  const int n = (nn - 1) / 2;
  const int64_t m = (int64_t)nn * nn * nn;
  const int64_t L = 2 * (int64_t)nout + 1;
  const double a = -z / (delta * delta * delta);
  const double inv_d3 = 1.0 / (delta * delta * delta);
  const zc inv_scale = 1.0 / (scale_re + I * scale_im);
  zc *v = (zc *)v_ri;
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < m; ++q) {
    const int qi = (int)(q % nn) - n, qj = (int)((q / nn) % nn) - n, qk = (int)(q / ((int64_t)nn * nn)) - n;
    const double tq = tinv[(qi + nout) + L * ((qj + nout) + L * (qk + nout))];
    for (int64_t p = 0; p < m; ++p) {
      const int pi = (int)(p % nn) - n, pj = (int)((p / nn) % nn) - n, pk = (int)(p / ((int64_t)nn * nn)) - n;
      const double tp = tinv[(pi + nout) + L * ((pj + nout) + L * (pk + nout))];
      const double tee = tinv[(qi - pi + nout) + L * ((qj - pj + nout) + L * (qk - pk + nout))];
      v[p + m * q] = (a * tq + a * tp + tee * inv_d3) * inv_scale;
    }
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
