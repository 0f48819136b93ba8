/*
 * dex.h - C-ABI of the MI355X implementation of the helium two-electron
 * complex-scaled sinc-DVR matvec and the GPLMR / harmonic-Ritz block
 * eigensolver (one GPU, one HIP stream per context).
 *
 * Conventions (those of the reference's Fortran boundary,
 * complexscale-djhnow/code/doublyexcited.F90:24-59 and gplmr.F90:9-36):
 *   - complex numbers are interleaved (re, im) doubles, i.e. COMPLEX*16;
 *   - a two-electron vector is X(P, Q), P = electron 1 flattened with i1
 *     fastest, Q = electron 2 (hammod.F90:149, initmod.F90:86-98); its length
 *     is D = NN^6 with NN = 2N+1 points per axis;
 *   - a block of k vectors is vector-major: vector j starts at j*D;
 *   - every call returns 0 on success or a DEX_E* status; nothing calls exit().
 *
 * Threading: a context is used by one host thread at a time; calls on one
 * context are serialised on its stream.  Contexts are independent.
 */
#ifndef DEX_H
#define DEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
  DEX_OK = 0,
  DEX_ESIZE = 1,          /* checksize / M*MB mismatch: "SIZE ERRORRRR 66", doublyexcited.F90:33-36 */
  DEX_ECHOLESKY = 2,      /* cholQR failed (strict mode): "Cholesky failed!!!", lamod.f90:169-175 */
  DEX_EGEVP = 3,          /* projected pencil solve failed, gplmr.F90:430-433 */
  DEX_ENOTCONVERGED = 4,  /* MAXIT reached with max residual > TOL (silent in the reference, gplmr.F90:86) */
  DEX_EHIP = 5,           /* a HIP runtime call failed */
  DEX_EOOM = 6,           /* device allocation failed */
  DEX_EINVAL = 7,         /* invalid argument */
  DEX_EADMISSION = 8,     /* the solve does not fit the context's HBM cap */
  DEX_EUNSUPPORTED = 9,   /* grid size / precision combination not built */
  DEX_EZEROBLOCK = 10     /* all-zero block in cholQR ("Warning rsum=0"), lamod.f90:157-166 */
};

enum { DEX_FP64 = 0, DEX_FP32 = 1 };

/* Grid and Hamiltonian parameters: the fields of PARAMETERS / MOLECULE the path
 * uses (structs.F90:28-60, input.F90:27-44).  One nucleus of charge `charge`
 * at the origin grid point (NUC_ON_GRID, RNUC = 0: input.F90:244-250). */
typedef struct dex_params {
  int32_t n;          /* half-width N; NN = 2N+1 (initmod.F90:31) */
  double delta;       /* DELTA, grid spacing (bohr) */
  double theta;       /* THETA, complex scaling angle (rad) */
  int32_t charge;     /* CHARGE (default 2, input.F90:40) */
  int32_t nsmall;     /* NSMALL (default 22, input.F90:31) */
  int32_t tfactor;    /* TFACTOR (default 4, input.F90:33) */
  int32_t precision;  /* DEX_FP64, or DEX_FP32: the mixed form (FP32 kinetic contraction on
                         complex64 vectors for the solver's search blocks; see dex_gplmr_eig) */
  int32_t device;     /* HIP device ordinal */
} dex_params;

typedef struct dex_ctx dex_ctx;

typedef struct dex_info {
  int32_t n, nn;
  int64_t m;          /* NN^3 */
  int64_t d;          /* NN^6 = vector length */
  double delta, theta;
  int32_t precision;
  int32_t fast_path;  /* 1 if the LDS two-pass matvec is used for this NN/precision */
} dex_info;

/* Fills params with the reference defaults (input.F90:27-44): N=2, DELTA=0.5,
 * THETA=0, CHARGE=2, NSMALL=22, TFACTOR=4, FP64, device 0. */
void dex_params_default(dex_params *p);

/* INIT_STRUCTS(NELEC=2) + INIT_HAMILTONIAN (input.F90:365-417,
 * initmod.F90:174-228): builds T (KEMATRIX), Tinv (GETINVERSE) and the
 * potential tables, uploads them to the device and creates the stream. */
int dex_ctx_create(const dex_params *p, dex_ctx **out);
int dex_ctx_destroy(dex_ctx *ctx);
int dex_ctx_info(const dex_ctx *ctx, dex_info *info);

/* stop = 1 asks a solve running on this context (dex_gplmr_eig_device and
 * its restarts) to return at its next iteration boundary, with its current
 * X and E and DEX_ENOTCONVERGED unless it had converged; later solves on the
 * context return at their first boundary until stop = 0 clears the request.
 * Safe to call from another thread while the solve runs.  No counterpart in
 * the reference (one run per process, ended by MAXIT): it is for servers that
 * end work at a deadline. */
int dex_ctx_stop(dex_ctx *ctx, int32_t stop);

/* Copies of the host-side tables (for checking): T(NN,NN) real before the
 * 1/SCALE^2 factor, Tinv(-NN:NN)^3 real (the reference's HAM%TINV). */
int dex_get_kematrix(const dex_ctx *ctx, double *t /* NN*NN */);
int dex_get_tinv(const dex_ctx *ctx, double *tinv /* (2NN+1)^3 */);

/* MYMATMUL(IN,OUT,HOWMANY,CHECKSIZE), doublyexcited.F90:24-44, i.e. HOWMANY
 * calls of HAMMULT (hammod.F90:19-67).  Host pointers; checksize must equal D. */
int dex_matvec(dex_ctx *ctx, const double *in, double *out, int32_t howmany, int64_t checksize);
/* Same on device pointers (complex128, or complex64 when the context is
 * DEX_FP32 and `single` is 1), ordered on the context's stream. */
int dex_matvec_device(dex_ctx *ctx, const void *in, void *out, int32_t howmany,
                      int64_t checksize, int32_t single);

/* HAMMULT for exchange-(anti)symmetric vectors, X(P,Q) = symmetry * X(Q,P)
 * (symmetry = +1 singlet, -1 triplet; 0 is dex_matvec_device).  Electron 2's
 * kinetic term is then the transpose of electron 1's, so the second pass is a
 * coalesced transpose-add; the result is Y = H Xs with Xs = (X + symmetry X^T)/2,
 * exactly (anti)symmetric.  This is how the reference's start vectors enter
 * (HELIUM_ORBITAL with SIGN = +-1, hammod.F90:1059-1063); HAMMULT itself has
 * no such mode.  Needs the LDS kernels for the grid (dex_info.fast_path).
 * With howmany >= 2 on complex128 columns (and kernel timing off) the columns
 * go through the passes in batches of up to 16 per launch, where the grid is
 * small enough for their intermediates to fit in 512 MiB (NN <= 15). */
int dex_matvec_sym_device(dex_ctx *ctx, const void *in, void *out, int32_t howmany,
                          int64_t checksize, int32_t single, int32_t symmetry);

/* dex_matvec_sym_device without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops that forbid it).  The intermediate U lives in a static 64 MiB
 * device array that the library holds on every GPU it runs on until the
 * process ends, so the LDS kernels' grids are limited to NN <= 11 in an FP64
 * context and NN <= 13 in an FP32 one (larger: DEX_ESIZE); FP64 grids without
 * the LDS kernels (NN >= 23) use the six-pass kernels, which need no scratch.
 * Columns go one at a time (no batched launches) and no kernel timing is
 * recorded.  Calls that overlap on one GPU share the array and race: order them
 * on one stream or serialise them.  U needs no clearing (each column's first
 * pass writes all of it). */
int dex_matvec_noalloc(dex_ctx *ctx, const void *in, void *out, int32_t howmany,
                       int64_t checksize, int32_t single, int32_t symmetry);

/* A HIP graph of dex_matvec_noalloc(ctx, in, out, howmany, checksize, single,
 * symmetry), captured on the context's stream: each dex_graph_launch replays
 * the same kernels on the same device pointers (reading what `in` holds at
 * replay time) as one launch on that stream; dex_sync orders the host after
 * it.  No counterpart in the reference (one matvec call per MYMATMUL); it is
 * for callers that repeat a block matvec in a loop, such as scans of many
 * small solves, and want one host launch per block.  Same limits and errors as
 * dex_matvec_noalloc (the capture records its launches; nothing runs at
 * creation).  Replays share that function's static scratch array: replays and
 * other allocation-free calls on one GPU must be ordered on one stream.
 * Destroy a graph before its context. */
typedef struct dex_graph dex_graph;
int dex_matvec_graph_create(dex_ctx *ctx, const void *in, void *out, int32_t howmany,
                            int64_t checksize, int32_t single, int32_t symmetry, dex_graph **out_graph);
int dex_graph_launch(dex_graph *graph);
int dex_graph_destroy(dex_graph *graph);

/* MYDOTSUB(BRAS,KETS,SIZE,NN,MM,OUTDOTS), doublyexcited.F90:47-59:
 * OUT(NN,MM) = BRAS^H KETS, column-major.  Host or device pointers. */
int dex_dots(dex_ctx *ctx, const double *bras, const double *kets, int64_t size,
             int32_t nn, int32_t mm, double *out);
int dex_dots_device(dex_ctx *ctx, const void *bras, const void *kets, int64_t size,
                    int32_t nn, int32_t mm, double *out_host);

typedef struct dex_result {
  int32_t iterations;     /* ITER - 1 at exit */
  int32_t converged;      /* 1 if max residual <= TOL */
  int64_t matvecs;        /* HAMMULT applications */
  double max_residual;
  double t_total;         /* TOT_TIME (s) */
  double t_matvec;        /* TSPMV (s), gplmr.F90:71-74: host time issuing the matvecs (they run
                             asynchronously; device time per kernel class: dex_get_kernel_stats) */
  double t_prec;          /* TPREC (s) */
  double t_gram;          /* projection + orthogonalisation Grams (s; host wall time, which includes
                             waiting for the matvecs queued before each Gram) */
  double t_update;        /* block recombinations (s) */
  double t_smalleig;      /* host projected pencil (s) */
  int32_t fallbacks;      /* orthonormalisations that left the plain cholQR path */
  int32_t fp64_restarts_used; /* FP32 contexts: recomputations of AX = H X in FP64 (refreshes) */
} dex_result;

typedef struct dex_solve_opts {
  int32_t strict_cholesky; /* 1: fail with DEX_ECHOLESKY where the reference STOPs */
  int32_t verbose;         /* 1: print "EIG j = ..." every OUTMODULUS iterations like gplmr.F90:343-352;
                              2: one progress line to stderr every OUTMODULUS iterations instead */
  int32_t refresh_every;   /* >0: recompute AX = H X every this many iterations; 0: never in FP64
                              contexts (as the reference), every 10 in FP32 contexts */
  int32_t fp64_restarts;   /* FP32 contexts: at most this many restarts (fresh FP64 AX, no P block) after
                              a run that ends at MAXIT unconverged (default 0) */
  double fp32_tol;         /* FP32 contexts: AX is also recomputed in FP64 when the residual first falls
                              below this (default 1e-5), where the FP32 drift starts to matter */
  int32_t symmetry;        /* 0: general HAMMULT; +1 / -1: project the start block to
                              singlet / triplet and use dex_matvec_sym_device */
  int32_t lock;            /* 1: build W, S, P only for columns whose residual is above
                              TOL (all columns stay in X and the projection); 0 (default):
                              every column, as the reference */
  int32_t inner_block;     /* inner solve (inner_iter >= 2): >0 solves column groups of
                              this many at a time, each in its own block Krylov space
                              (an (m + 1) inner_block column basis instead of (m + 1) NEV);
                              0 (default): the whole block at once, as the reference */
  int32_t guard;           /* extra block columns (seeded Gaussians) iterated beside the
                              caller's NEV: convergence is tested, and values returned, on
                              the NEV nearest; a block that would end inside a degenerate
                              level then holds the whole level.  0 (default): the
                              reference's block of exactly NEV.  Needs NS(NEV+guard) <= 48 and
                              (3+NS)(NEV+guard) <= 96 */
  int32_t inner_fp32;      /* inner solve: 1 keeps its Arnoldi basis in complex64 and applies H
                              in FP32 (half the traffic of the preconditioner; the outer
                              iteration, its residuals and the returned pairs stay FP64);
                              0 (default): FP64, as the reference */
} dex_solve_opts;

void dex_solve_opts_default(dex_solve_opts *o);

/* GPLMR_EIG(HAM,PARAMS,MPIDATA,GMSOLVER,X,E,M,MB,SIGMAIN,NEV,NS,MAXIT,TOL,
 *           PRECONDITIONER,OUTMODULUS), gplmr.F90:9.
 * X (M*MB*NEV complex, host): start block in, eigenvectors out.  E (NEV
 * complex, host): Rayleigh quotients out.  M*MB must equal D (one process
 * holds the whole vector).  inner_iter is GM_MAX_INNER_ITER (PRECBLOCK,
 * gplmr.F90:466-502): 0 no inner solve, 1 W <- H W, m >= 2 the in-house
 * order-m block minimal-residual solve of (H - sigma) x = b
 * (PRECBLOCK_BASIC, gplmr.F90:674-770) applied to every column of the
 * block.  res/opts may be NULL. */
int dex_gplmr_eig(dex_ctx *ctx, double *x, double *e, int32_t m, int32_t mb,
                  double sigma_re, double sigma_im, int32_t nev, int32_t ns,
                  int32_t maxit, double tol, int32_t inner_iter, int32_t outmodulus,
                  const dex_solve_opts *opts, dex_result *res);
/* Same with X resident on the device (complex128, M*MB*NEV).
 *
 * In an FP32 context (the mixed form) X and AX stay complex128 and AX = H X is
 * applied in FP64, while the search blocks W, S, P, their images and the inner
 * solve's Arnoldi basis are stored in complex64 and multiplied by H in FP32
 * (FP32 accumulation); every Gram, combination, Cholesky factor and the
 * projected pencil are FP64.  AX, carried by the recurrence over FP32 images,
 * is recomputed as H X in FP64 every refresh_every iterations (10 by default),
 * when the residual first falls below fp32_tol, and whenever it falls below
 * TOL: the solve ends only on a residual formed from an FP64 AX.  The
 * reference is FP64 throughout (definitions.INC:1-2). */
int dex_gplmr_eig_device(dex_ctx *ctx, void *x_dev, double *e, int32_t m, int32_t mb,
                         double sigma_re, double sigma_im, int32_t nev, int32_t ns,
                         int32_t maxit, double tol, int32_t inner_iter, int32_t outmodulus,
                         const dex_solve_opts *opts, dex_result *res);

/* Seeded Philox4x32-10 complex Gaussian start block on the device: entry i of
 * column c is (g1 + i g2)/sqrt(2) from counter (i, c, stream_id, 0) and key
 * seed; symmetry +1 / -1 then projects each column, X <- (X +- X^T)/2. */
int dex_fill_gaussian_device(dex_ctx *ctx, void *x_dev, int32_t ncols, uint64_t seed,
                             uint32_t stream_id, int32_t symmetry);
int dex_sync(dex_ctx *ctx);

/* Residual norms ||H x_j - E_j x_j|| / ||x_j|| recomputed in FP64 from scratch. */
int dex_residuals_device(dex_ctx *ctx, const void *x_dev, const double *e, int32_t nev,
                         double *res_out);

/* PRECBLOCK(HAM,PARAMS,MPIDATA,GMSOLVER,X,PX,SIGMA,M,MB,NEV,MAX_INNER_ITER,...),
 * gplmr.F90:466-502, on device blocks of NEV columns (X is read, PX written;
 * they must not alias).  inner_iter = 1: PX = H X (MATBLOCK); inner_iter =
 * m >= 2: PX = argmin over the block Krylov space K_m(H, X) of
 * ||(H - sigma) PX - X|| (PRECBLOCK_BASIC -> basicblocklansolve,
 * gplmr.F90:674-770, harmonic_ritz_small.F90:258-324), every column
 * multiplied.  inner_iter = 0 is an error, as in the reference (:487-489).
 * symmetry: 0, or +-1 for exchange-(anti)symmetric X. */
int dex_precblock_device(dex_ctx *ctx, const void *x_dev, void *px_dev, int32_t m, int32_t mb,
                         int32_t nev, double sigma_re, double sigma_im, int32_t inner_iter,
                         int32_t symmetry);

/* Physically targeted start vectors from one-electron orbitals:
 * HELIUM_ORBITALGATHER's Gram-Schmidt of the orbitals (bilinear products, or
 * Hermitian for hermopt != 0; hammod.F90:1080-1113), DIAGSUBSPACE over the
 * (anti)symmetrised orbital products (singlet: a <= b, triplet: a < b;
 * hammod.F90:1125-1256) with its eigenvalues ordered by |E0 - etarget| as
 * mysort (hammod.F90:1658-1690), and GET_EXPECTS with the Hermitian product
 * (cshiftopt false, hammod.F90:1261-1301).  orbitals: M x numorbs complex,
 * column-major, host (the He+ eigenvectors the reference's Lanczos returns).
 * psi_dev: numout device vectors (D complex each, Hermitian-normalised); e0,
 * expects: numout complex each.  The grid needs the LDS matvec kernels. */
int dex_diagsubspace_device(dex_ctx *ctx, const double *orbitals, int32_t numorbs, int32_t singlet,
                            int32_t hermopt, double etarget_re, double etarget_im, int32_t numout,
                            void *psi_dev, double *e0, double *expects);

/* BLOCKLANCZOS(PARAMS,LANPAR,HAM,MPIDATA,BB), lanmod.F90:18-219, on the
 * device: the complex-symmetric block Lanczos of the one-electron (He+)
 * operator H1 = T3D / SCALE^2 + Vn / SCALE of this context's grid (KEMULT1 and
 * POTENTIAL1, hammod.F90:71-134,432-498), whose eigenvectors are the orbitals
 * of the product start vectors (dex_diagsubspace_device).  blocksize is
 * LANBLOCKSIZE, kdim KDIM (blocks kept; blocksize * kdim <= NN^3), check CHECK
 * (a convergence test every this many steps), tol LANTOL on the Hermitian
 * residual norms, numeigs NUMORBS (the eigenvalues of smallest real part).
 * The start block is seeded (uniform [0,1) real parts, 1e-4 i uniform
 * imaginary parts, as RANDOM_NUMBER fills it).  Out: eigenvalues (numeigs
 * complex), orbitals (M x numeigs complex, column-major, Hermitian-normalised;
 * may be NULL), residuals (numeigs; may be NULL), iterations (may be NULL).
 * DEX_ENOTCONVERGED ("LANCZOS FAILED TO CONVERGE") returns the last check's
 * values. */
int dex_heplus_lanczos_device(dex_ctx *ctx, int32_t blocksize, int32_t kdim, int32_t check, double tol,
                              int32_t numeigs, uint64_t seed, double *eigenvalues, double *orbitals,
                              double *residuals, int32_t *iterations);

/* Bytes of device memory a dex_gplmr_eig_device solve of (nev, ns, inner_iter,
 * opts: inner_block, guard and inner_fp32; NULL for the defaults) holds on this
 * context: the caller's block X, the solver's work vectors (complex64 search
 * blocks in an FP32 context) and the context's own fixed buffers (tables,
 * block-kernel scratch, matvec intermediates).  No reference counterpart (the
 * reference allocates its work arrays inside GPLMR_EIG, gplmr.F90:31-36). */
int64_t dex_gplmr_footprint(const dex_ctx *ctx, int32_t nev, int32_t ns, int32_t inner_iter,
                            const dex_solve_opts *opts);
/* The same figure from the parameters alone (p->n and p->precision), before
 * any context exists and without a device: a request server's admission test. */
int64_t dex_solve_footprint(const dex_params *p, int32_t nev, int32_t ns, int32_t inner_iter,
                            const dex_solve_opts *opts);

/* Eigenvector files in the reference's format, READWRITE_EIGENVECTORS
 * (utils.F90:82-181): a Fortran unformatted sequential file (gfortran record
 * framing, subrecords beyond 2147483639 bytes) with a header record M, NUM,
 * SIZEOF(vector) and one M x M complex128 record per vector.  The size field
 * is written as int64 and read as int64 or, from a 12-byte header, as its low
 * 32 bits (the reference reads it into a 4-byte integer, utils.F90:95, which
 * wraps at 2 GiB).  Reading into a different odd-cubed M applies SIXD_FIX
 * (utils.F90:184-198): the common central block is kept, the rest is zero.
 * psi: host, num consecutive vectors of M*M complex (interleaved doubles). */
int dex_write_eigvecs(const char *path, const double *psi, int32_t m, int32_t num);
int dex_read_eigvecs_info(const char *path, int32_t *m, int32_t *num);
int dex_read_eigvecs(const char *path, int32_t m, int32_t num, double *psi);
/* SIXD_FIX alone: one vector from the (2 n_in + 1)^6 grid to the (2 n_out + 1)^6 grid. */
int dex_host_sixd_fix(const double *in, int32_t n_in, double *out, int32_t n_out);

/* Host-only set-up and small-matrix routines (no device use), exported so the
 * host arithmetic can be checked on machines without a GPU:
 *   KEMATRIX (hammod.F90:235-260): t = NN x NN column-major, real;
 *   GETINVERSE (tooth.f90:14-101): tinv = (2 nout + 1)^3 reals, first index fastest;
 *   SOLVE_GEVP (gplmr.F90:395-439): the NEV pairs of K y = mu M y with smallest
 *   |mu|; kp, mp: k x k complex (leading dimension ld); coords: k x nev complex
 *   (unit 2-norm columns), mu: nev complex. */
int dex_host_kematrix(int32_t nn, double delta, double *t);
int dex_host_getinverse(int32_t nout, double delta, int32_t nsmall, int32_t tfactor, double *tinv);
int dex_host_solve_gevp(int32_t k, const double *kp, const double *mp, int32_t ld, int32_t nev,
                        double *coords, double *mu);

/* Human-readable text of a status code, and the last error message on this thread. */
const char *dex_status_string(int status);
const char *dex_last_error(void);

/* Device event timing of the dominant matvec kernels of the most recent
 * dex_matvec_device call with timing enabled (ms per launch, -1 if none).
 * enable: 0 off, 1 every launch, n > 1 a random sample of about 1 in n
 * launches for the kernel-class statistics below (each launch drawn
 * independently; events cost about 10% of a solve's time when every launch is
 * timed). */
int dex_set_kernel_timing(dex_ctx *ctx, int32_t enable);
int dex_get_kernel_timing(const dex_ctx *ctx, double *pass1_ms, double *pass2_ms);

/* With timing enabled, every timed launch of the library's kernels is bracketed by
 * HIP events on the context's stream (resolved lazily, so no synchronisation
 * is added to the work); per kernel class the summed device time, the number
 * of timed launches and their algorithmic HBM bytes and flops (counted on
 * the host from the launch shapes) accumulate until dex_reset_kernel_stats.  No reference
 * counterpart (the reference times with MPI_WTIME only, gplmr.F90:64-380). */
enum {
  DEX_K_PASS1 = 0,   /* electron-1 three-axis kinetic pass */
  DEX_K_PASS2 = 1,   /* electron-2 pass with the potential (general vectors) */
  DEX_K_SYMADD = 2,  /* exchange transpose-add pass with the potential */
  DEX_K_GRAM = 3,    /* block Gram (with its partial-sum reduction) */
  DEX_K_COMBINE = 4, /* block linear combination */
  DEX_K_COLDOTS = 5, /* per-column dots and norms, and the residual step (scaled X, AX and W) */
  DEX_K_GENERIC = 6, /* generic-grid matvec */
  DEX_KCLASSES = 7
};
int dex_get_kernel_stats(dex_ctx *ctx, int32_t kclass, double *total_ms, int64_t *launches,
                         double *alg_bytes, double *alg_flops);
/* Per kernel class, every launch since the context was created or the stats
 * were last reset, timed or not: the launch count and the algorithmic HBM
 * bytes and flops (host-side counts from the launch shapes; no device call). */
int dex_get_kernel_totals(const dex_ctx *ctx, int32_t kclass, int64_t *launches, double *alg_bytes,
                          double *alg_flops);
int dex_reset_kernel_stats(dex_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* DEX_H */
