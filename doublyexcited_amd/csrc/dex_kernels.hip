// HIP kernels for gfx950 (MI355X): the two-electron sinc-DVR matvec and the
// tall-skinny block kernels of the GPLMR eigensolver.
//
// Matvec.  A two-electron vector is the M x M matrix X[p + M q] (p: electron 1,
// q: electron 2, each p = i + NN j + NN^2 k).  With T = tau / delta^2 the real
// Toeplitz sinc-DVR matrix (tau(0) = pi^2/6, tau(m) = (-1)^m / m^2) the
// reference's HAMMULT (hammod.F90:19-67) is
//     Y = e^{-2 i theta} (T3D x I + I x T3D) X + e^{-i theta} Vr o X,
// Vr(p, q) = -Z/delta^3 (Tinv(r_p) + Tinv(r_q)) + Tinv(r_q - r_p)/delta^3.
// tau is a compile-time table, so the kinetic contraction is a stream of
// FMAs with scalar constants.  Two passes, each over an NN^3 cube held in LDS:
//   pass 1 (electron 1): U[:, q] = tau3D X[:, q]  (contiguous cubes),
//   pass 2 (electron 2): Y[p, :] = s2 (U[p, :] + tau3D X[p, :]) + c1 Vr X[p, :]
// (cubes at stride M; neighbouring p share cache lines, so the workgroup ->
// p map keeps groups of 8 consecutive p on one XCD).  The kinetic operator is
// real, so the real and imaginary parts are contracted in turn through two
// LDS buffers of NN^3 reals each (148 KB at NN = 21 in FP64).
#include <atomic>
#include <cstdlib>
#include <mutex>

#include <hip/hip_runtime.h>

#include "dex_internal.h"

namespace dex {

namespace {

constexpr double kPi2over6 = 1.6449340668482264;  // pi^2 / 6

template <int M>
__device__ __forceinline__ constexpr double tau_c() {
  if constexpr (M == 0) {
    return kPi2over6;
  } else {
    return ((M & 1) ? -1.0 : 1.0) / double(M * M);
  }
}

__device__ __forceinline__ constexpr double tau_rt(int m) {
  m = m < 0 ? -m : m;
  return m == 0 ? kPi2over6 : ((m & 1) ? -1.0 : 1.0) / double(m * m);
}

template <typename R>
struct Cplx;
template <>
struct Cplx<double> {
  using type = double2;
};
template <>
struct Cplx<float> {
  using type = float2;
};

constexpr int round_up(int a, int b) { return (a + b - 1) / b * b; }

template <int NN>
constexpr int cube_threads() {
  return round_up(NN * NN, 64) > 1024 ? 1024 : round_up(NN * NN, 64);
}

// b = tau3D a on one NN^3 cube in LDS (a, b: NN^3 reals, i fastest).
// One thread per line: it reads the NN inputs of its line once and forms all
// NN outputs with NN^2 FMAs against compile-time coefficients.
template <int NN, int NT, typename R>
__device__ __forceinline__ void cube_apply(const R *__restrict__ a, R *__restrict__ b, int tid) {
  constexpr int N2 = NN * NN;
  // axis i (stride 1): line c = (j, k), base c * NN
  for (int c = tid; c < N2; c += NT) {
    R acc[NN];
#pragma unroll
    for (int i = 0; i < NN; ++i) acc[i] = R(0);
#pragma unroll
    for (int l = 0; l < NN; ++l) {
      const R xv = a[c * NN + l];
#pragma unroll
      for (int i = 0; i < NN; ++i) acc[i] += R(tau_rt(l - i)) * xv;
    }
#pragma unroll
    for (int i = 0; i < NN; ++i) b[c * NN + i] = acc[i];
  }
  __syncthreads();
  // axis j (stride NN): line c = (i, k)
  for (int c = tid; c < N2; c += NT) {
    const int i = c % NN, k = c / NN;
    const int base = i + N2 * k;
    R acc[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) acc[j] = R(0);
#pragma unroll
    for (int l = 0; l < NN; ++l) {
      const R xv = a[base + NN * l];
#pragma unroll
      for (int j = 0; j < NN; ++j) acc[j] += R(tau_rt(l - j)) * xv;
    }
#pragma unroll
    for (int j = 0; j < NN; ++j) b[base + NN * j] += acc[j];
  }
  __syncthreads();
  // axis k (stride NN^2): line c = (i, j)
  for (int c = tid; c < N2; c += NT) {
    R acc[NN];
#pragma unroll
    for (int k = 0; k < NN; ++k) acc[k] = R(0);
#pragma unroll
    for (int l = 0; l < NN; ++l) {
      const R xv = a[c + N2 * l];
#pragma unroll
      for (int k = 0; k < NN; ++k) acc[k] += R(tau_rt(l - k)) * xv;
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) b[c + N2 * k] += acc[k];
  }
  __syncthreads();
}

// Pass 1: electron-1 cubes (contiguous, one per electron-2 index q).
// U = tau3D X on each cube (no phase, no potential).  Workgroups are
// persistent: while a workgroup contracts cube q in LDS, the loads of its next
// cube (q + gridDim.x) are already in flight into registers.
template <int NN, typename R, typename XT>
__global__ __launch_bounds__(cube_threads<NN>()) void k_pass1(ColPtrs xs, ColPtrsW us, int ncols) {
  using C = typename Cplx<R>::type;
  constexpr int NT = cube_threads<NN>();
  constexpr int N3 = NN * NN * NN;
  constexpr int PER = (N3 + NT - 1) / NT;
  extern __shared__ __align__(16) unsigned char smem[];
  R *a = reinterpret_cast<R *>(smem);
  R *b = a + N3;
  const int tid = threadIdx.x;
  // work item q: cube q % N3 of column q / N3
  const int64_t total = int64_t(ncols) * N3;
  auto xcube = [&](int64_t qq) {
    const int cc = int(qq / N3);
    return static_cast<const XT *>(xs.p[cc]) + (qq - int64_t(cc) * N3) * N3;
  };
  XT pre[PER];
  int64_t q = blockIdx.x;
  if (q < total) {
    const XT *xq = xcube(q);
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = tid + t * NT;
      if (e < N3) pre[t] = xq[e];
    }
  }
  for (; q < total; q += gridDim.x) {
    R keep[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = tid + t * NT;
      if (e < N3) {
        a[e] = R(pre[t].x);
        keep[t] = R(pre[t].y);
      }
    }
    __syncthreads();
    const int64_t qn = q + gridDim.x;
    if (qn < total) {
      const XT *xq = xcube(qn);
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int e = tid + t * NT;
        if (e < N3) pre[t] = xq[e];
      }
    }
    cube_apply<NN, NT, R>(a, b, tid);
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = tid + t * NT;
      if (e < N3) {
        const R re = b[e];
        a[e] = keep[t];
        keep[t] = re;
      }
    }
    __syncthreads();
    cube_apply<NN, NT, R>(a, b, tid);
    const int cq = int(q / N3);
    C *uq = static_cast<C *>(us.p[cq]) + (q - int64_t(cq) * N3) * N3;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = tid + t * NT;
      if (e < N3) {
        C o;
        o.x = keep[t];
        o.y = b[e];
        uq[e] = o;
      }
    }
    __syncthreads();
  }
}

// Workgroup -> electron-1 index p for pass 2: consecutive workgroups are
// dealt round-robin to the 8 XCDs, so the s-th workgroup on XCD x takes
// p = 8 (x + 8 (s / 8)) + s % 8, which puts 8 consecutive p (one 128-byte
// line of complex doubles per q) on the same XCD at nearly the same time.
__device__ __forceinline__ int64_t xcd_group_map(int64_t b, int64_t m) {
  const int64_t full = (m / 64) * 64;
  if (b >= full) return b;
  const int64_t x = b % 8, s = b / 8;
  return 8 * (x + 8 * (s / 8)) + (s % 8);
}

template <int NN, typename R, typename XT, typename YT>
__global__ __launch_bounds__(cube_threads<NN>()) void k_pass2(
    const XT *__restrict__ x, const typename Cplx<R>::type *__restrict__ u, YT *__restrict__ y,
    const double *__restrict__ vn, const double *__restrict__ tinv, MatvecCoef cf) {
  using C = typename Cplx<R>::type;
  constexpr int NT = cube_threads<NN>();
  constexpr int N3 = NN * NN * NN;
  constexpr int PER = (N3 + NT - 1) / NT;
  constexpr int L = 2 * NN + 1;
  extern __shared__ __align__(16) unsigned char smem[];
  R *a = reinterpret_cast<R *>(smem);
  R *b = a + N3;
  const int tid = threadIdx.x;
  const int64_t p = xcd_group_map(blockIdx.x, N3);
  const int64_t M = N3;
  R keep[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int e = tid + t * NT;
    if (e < N3) {
      const XT v = x[p + M * e];
      a[e] = R(v.x);
      keep[t] = R(v.y);
    }
  }
  __syncthreads();
  cube_apply<NN, NT, R>(a, b, tid);
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int e = tid + t * NT;
    if (e < N3) {
      const R re = b[e];
      a[e] = keep[t];
      keep[t] = re;
    }
  }
  __syncthreads();
  cube_apply<NN, NT, R>(a, b, tid);
  // epilogue: Y = s2 (U + K) + c1 Vr X
  const int ip = int(p % NN), jp = int((p / NN) % NN), kp = int(p / (NN * NN));
  const double vnp = vn[p];
  const R s2r = R(cf.s2_re), s2i = R(cf.s2_im), c1r = R(cf.c1_re), c1i = R(cf.c1_im);
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int e = tid + t * NT;
    if (e < N3) {
      const int iq = e % NN, jq = (e / NN) % NN, kq = e / (NN * NN);
      const int ti = (iq - ip + NN) + L * ((jq - jp + NN) + L * (kq - kp + NN));
      const R vr = R(vnp + vn[e] + tinv[ti] * cf.inv_d3);
      const C uv = u[p + M * e];
      const XT xv = x[p + M * e];
      const R sr = uv.x + keep[t], si = uv.y + b[e];
      const R vxr = vr * R(xv.x), vxi = vr * R(xv.y);
      YT o;
      o.x = s2r * sr - s2i * si + c1r * vxr - c1i * vxi;
      o.y = s2r * si + s2i * sr + c1r * vxi + c1i * vxr;
      y[p + M * e] = o;
    }
  }
}

// ---- exchange-symmetric second pass ------------------------------------------
// For X = sign X^T (sign = +1 singlet, -1 triplet) electron 2's kinetic term is
// the transpose of electron 1's: (I x T3D) X = sign (T3D X)^T.  So after pass 1
// (U = tau3D X) the result is Y = s2 (U + sign U^T) + c1 Vr o Xs with
// Xs = (X + sign X^T) / 2, formed tile pair by tile pair: a workgroup owns the
// 32 x 32 tile (ta, tb), ta <= tb, and its mirror (tb, ta); all global
// accesses run along p (coalesced) and the transposes go through LDS.  Each
// unordered pair {(p, q), (q, p)} is computed once and written with its sign,
// so the output is exactly (anti)symmetric.
constexpr int kSymT = 32;

template <typename R, typename XT, typename YT>
__global__ __launch_bounds__(256) void k_symadd(ColPtrs xs, ColPtrs us, ColPtrsW ys, int64_t m, int nn,
                                                const double *__restrict__ vn,
                                                const double *__restrict__ tinv, MatvecCoef cf,
                                                int sign) {
  using C = typename Cplx<R>::type;
  const int ta = blockIdx.x, tb = blockIdx.y;
  if (ta > tb) return;
  // column blockIdx.z of the launch
  const XT *__restrict__ x = static_cast<const XT *>(xs.p[blockIdx.z]);
  const C *__restrict__ u = static_cast<const C *>(us.p[blockIdx.z]);
  YT *__restrict__ y = static_cast<YT *>(ys.p[blockIdx.z]);
  __shared__ C su[kSymT][kSymT + 1], sx[kSymT][kSymT + 1];
  __shared__ YT sy[kSymT][kSymT + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int64_t pa = int64_t(ta) * kSymT, qb = int64_t(tb) * kSymT;
  // mirror tile: su[r][c] = U[q + m p] with q = qb + c, p = pa + r
  for (int r = ty; r < kSymT; r += 8) {
    const int64_t q = qb + tx, p = pa + r;
    if (q < m && p < m) {
      su[r][tx] = u[q + m * p];
      const XT xv = x[q + m * p];
      C xc;
      xc.x = R(xv.x);
      xc.y = R(xv.y);
      sx[r][tx] = xc;
    }
  }
  __syncthreads();
  const int L = 2 * nn + 1;
  const R s2r = R(cf.s2_re), s2i = R(cf.s2_im), c1r = R(cf.c1_re), c1i = R(cf.c1_im);
  const R sg = R(sign);
  const bool diag = ta == tb;
  for (int r = ty; r < kSymT; r += 8) {
    const int64_t p = pa + tx, q = qb + r;  // main element (p, q) at p + m q
    if (p < m && q < m) {
      const C uv = u[p + m * q];
      const XT xv = x[p + m * q];
      const C um = su[tx][r], xm = sx[tx][r];  // U[q + m p], X[q + m p]
      const R sr = uv.x + sg * um.x, si = uv.y + sg * um.y;
      const R xr = R(0.5) * (R(xv.x) + sg * xm.x), xi = R(0.5) * (R(xv.y) + sg * xm.y);
      const int ip = int(p % nn), jp = int((p / nn) % nn), kp = int(p / (nn * nn));
      const int iq = int(q % nn), jq = int((q / nn) % nn), kq = int(q / (nn * nn));
      const int ti = (iq - ip + nn) + L * ((jq - jp + nn) + L * (kq - kp + nn));
      const R vr = R(vn[p] + vn[q] + tinv[ti] * cf.inv_d3);
      const R vxr = vr * xr, vxi = vr * xi;
      YT o;
      o.x = s2r * sr - s2i * si + c1r * vxr - c1i * vxi;
      o.y = s2r * si + s2i * sr + c1r * vxi + c1i * vxr;
      sy[tx][r] = o;
      if (!diag || tx <= r) y[p + m * q] = o;
    }
  }
  __syncthreads();
  // mirror write: y[q + m p] = sign * Y(p, q), q = qb + tx, p = pa + r
  for (int r = ty; r < kSymT; r += 8) {
    const int64_t q = qb + tx, p = pa + r;
    if (q < m && p < m && (!diag || r < tx)) {
      const YT o = sy[r][tx];
      YT w;
      w.x = sg * o.x;
      w.y = sg * o.y;
      y[q + m * p] = w;
    }
  }
}

// ---- generic FP64 path ------------------------------------------------------
// One thread per line of the given axis (stride s): y (+)= tau * x along it.
template <int NN>
__global__ __launch_bounds__(256) void k_axis(const double2 *__restrict__ x, double2 *__restrict__ y,
                                              int64_t stride, int64_t nlines, int accumulate) {
  const int64_t line = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (line >= nlines) return;
  const int64_t off = (line % stride) + (line / stride) * stride * NN;
  double2 xv[NN];
#pragma unroll
  for (int l = 0; l < NN; ++l) xv[l] = x[off + stride * l];
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    double ar = 0.0, ai = 0.0;
#pragma unroll
    for (int l = 0; l < NN; ++l) {
      const double t = tau_rt(l - i);
      ar += t * xv[l].x;
      ai += t * xv[l].y;
    }
    double2 o = accumulate ? y[off + stride * i] : make_double2(0.0, 0.0);
    o.x += ar;
    o.y += ai;
    y[off + stride * i] = o;
  }
}

__global__ void k_potential(const double2 *__restrict__ x, double2 *__restrict__ y,
                            const double *__restrict__ vn, const double *__restrict__ tinv, int nn,
                            int64_t m, MatvecCoef cf) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= m * m) return;
  const int64_t p = idx % m, q = idx / m;
  const int L = 2 * nn + 1;
  const int ip = int(p % nn), jp = int((p / nn) % nn), kp = int(p / (int64_t(nn) * nn));
  const int iq = int(q % nn), jq = int((q / nn) % nn), kq = int(q / (int64_t(nn) * nn));
  const int64_t ti = (iq - ip + nn) + int64_t(L) * ((jq - jp + nn) + int64_t(L) * (kq - kp + nn));
  const double vr = vn[p] + vn[q] + tinv[ti] * cf.inv_d3;
  const double2 xv = x[idx];
  const double2 kv = y[idx];
  double2 o;
  o.x = cf.s2_re * kv.x - cf.s2_im * kv.y + vr * (cf.c1_re * xv.x - cf.c1_im * xv.y);
  o.y = cf.s2_re * kv.y + cf.s2_im * kv.x + vr * (cf.c1_re * xv.y + cf.c1_im * xv.x);
  y[idx] = o;
}

template <int NN>
hipError_t matvec_generic_nn(const double2 *x, double2 *y, const double *vn, const double *tinv,
                             const MatvecCoef &cf, hipStream_t s) {
  const int64_t m = int64_t(NN) * NN * NN;
  const int64_t d = m * m;
  const int64_t nlines = d / NN;
  const int bs = 256;
  const int64_t nb = (nlines + bs - 1) / bs;
  int64_t stride = 1;
  for (int axis = 0; axis < 6; ++axis) {
    hipLaunchKernelGGL(k_axis<NN>, dim3(unsigned(nb)), dim3(bs), 0, s, x, y, stride, nlines,
                       axis == 0 ? 0 : 1);
    stride *= NN;
  }
  const int64_t nb2 = (d + bs - 1) / bs;
  hipLaunchKernelGGL(k_potential, dim3(unsigned(nb2)), dim3(bs), 0, s, x, y, vn, tinv, NN, m, cf);
  return hipGetLastError();
}

// Function attributes and occupancy figures are set once per kernel, under
// g_setup_mu: prepare_kernels() does all of them when the first context is
// created, before any launch, so no hipFuncSetAttribute call ever runs while
// another thread launches (the launchers below only read the cached values,
// taking the lock themselves only if prepare_kernels() was never called).
std::mutex g_setup_mu;

// Workgroups of k_pass1<NN, R, XT> resident on the device at once (its
// dynamic LDS attribute set on first use).  0 on a HIP error.
template <int NN, typename R, typename XT>
int pass1_setup() {
  static std::atomic<int> resident_cached{0};
  int resident = resident_cached.load(std::memory_order_acquire);
  if (resident != 0) return resident;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  resident = resident_cached.load(std::memory_order_acquire);
  if (resident != 0) return resident;
  constexpr int N3 = NN * NN * NN;
  const size_t lds = size_t(2) * N3 * sizeof(R);
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pass1<NN, R, XT>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
    return 0;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(&k_pass1<NN, R, XT>),
                                                   cube_threads<NN>(), lds) != hipSuccess || per < 1)
    per = 1;
  resident = cus * per;
  resident_cached.store(resident, std::memory_order_release);
  return resident;
}

template <int NN, typename R, typename XT, typename YT>
bool pass2_setup() {
  static std::atomic<bool> attr{false};
  if (attr.load(std::memory_order_acquire)) return true;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (attr.load(std::memory_order_acquire)) return true;
  constexpr int N3 = NN * NN * NN;
  const size_t lds = size_t(2) * N3 * sizeof(R);
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pass2<NN, R, XT, YT>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
    return false;
  attr.store(true, std::memory_order_release);
  return true;
}

template <int NN, typename R, typename XT>
hipError_t pass1_nn(const ColPtrs &xs, const ColPtrsW &us, int ncols, hipStream_t s) {
  constexpr int N3 = NN * NN * NN;
  const size_t lds = size_t(2) * N3 * sizeof(R);
  // persistent grid: as many workgroups as fit on the device at once
  const int resident = pass1_setup<NN, R, XT>();
  if (resident == 0) return hipErrorInvalidDeviceFunction;
  const int64_t total = int64_t(ncols) * N3;
  const int grid = int(resident < total ? resident : total);
  hipLaunchKernelGGL((k_pass1<NN, R, XT>), dim3(grid), dim3(cube_threads<NN>()), lds, s, xs, us, ncols);
  return hipGetLastError();
}

template <int NN, typename R, typename XT, typename YT>
hipError_t pass2_nn(const void *x, const void *u, void *y, const double *vn, const double *tinv,
                    const MatvecCoef &cf, hipStream_t s) {
  using C = typename Cplx<R>::type;
  constexpr int N3 = NN * NN * NN;
  const size_t lds = size_t(2) * N3 * sizeof(R);
  if (!pass2_setup<NN, R, XT, YT>()) return hipErrorInvalidDeviceFunction;
  hipLaunchKernelGGL((k_pass2<NN, R, XT, YT>), dim3(N3), dim3(cube_threads<NN>()), lds, s,
                     static_cast<const XT *>(x), static_cast<const C *>(u), static_cast<YT *>(y), vn,
                     tinv, cf);
  return hipGetLastError();
}

// LDS budget per workgroup on gfx950.
constexpr size_t kMaxLds = 160 * 1024;

constexpr bool fast_fits(int nn, size_t rsize) {
  return size_t(2) * nn * nn * nn * rsize <= kMaxLds;
}

}  // namespace

#define DEX_NN_LIST(X) X(3) X(5) X(7) X(9) X(11) X(13) X(15) X(17) X(19) X(21) X(23) X(25)

bool fast_matvec_supported(int nn, bool single) {
  switch (nn) {
#define CASE(N) \
  case N:       \
    return fast_fits(N, single ? sizeof(float) : sizeof(double));
    DEX_NN_LIST(CASE)
#undef CASE
    default:
      return false;
  }
}

bool generic_matvec_supported(int nn) {
  switch (nn) {
#define CASE(N) \
  case N:       \
    return true;
    DEX_NN_LIST(CASE)
#undef CASE
    default:
      return false;
  }
}

hipError_t launch_pass1_batch(int nn, int io, const void *const *x, void *const *u, int ncols, hipStream_t s) {
  if (ncols < 1 || ncols > kBatchMax) return hipErrorInvalidValue;
  ColPtrs xs{};
  ColPtrsW us{};
  for (int j = 0; j < ncols; ++j) {
    xs.p[j] = x[j];
    us.p[j] = u[j];
  }
  switch (nn) {
#define CASE(N)                                                                             \
  case N:                                                                                   \
    if constexpr (fast_fits(N, sizeof(double))) {                                           \
      if (io == kIoF64) return pass1_nn<N, double, double2>(xs, us, ncols, s);              \
    }                                                                                       \
    if constexpr (fast_fits(N, sizeof(float))) {                                            \
      if (io == kIoF32) return pass1_nn<N, float, float2>(xs, us, ncols, s);                \
      if (io == kIoMixed) return pass1_nn<N, float, double2>(xs, us, ncols, s);             \
    }                                                                                       \
    return hipErrorInvalidValue;
    DEX_NN_LIST(CASE)
#undef CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_pass1(int nn, int io, const void *x, void *u, hipStream_t s) {
  return launch_pass1_batch(nn, io, &x, &u, 1, s);
}

hipError_t launch_pass2(int nn, int io, const void *x, const void *u, void *y,
                        const double *vn, const double *tinv, const MatvecCoef &coef,
                        hipStream_t s) {
  switch (nn) {
#define CASE(N)                                                                                   \
  case N:                                                                                         \
    if constexpr (fast_fits(N, sizeof(double))) {                                                 \
      if (io == kIoF64) return pass2_nn<N, double, double2, double2>(x, u, y, vn, tinv, coef, s); \
    }                                                                                             \
    if constexpr (fast_fits(N, sizeof(float))) {                                                  \
      if (io == kIoF32) return pass2_nn<N, float, float2, float2>(x, u, y, vn, tinv, coef, s);    \
      if (io == kIoMixed) return pass2_nn<N, float, double2, double2>(x, u, y, vn, tinv, coef, s); \
    }                                                                                             \
    return hipErrorInvalidValue;
    DEX_NN_LIST(CASE)
#undef CASE
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_symadd_batch(int nn, int io, const void *const *x, const void *const *u, void *const *y,
                               int ncols, const double *vn, const double *tinv, const MatvecCoef &coef,
                               int sign, hipStream_t s) {
  if (ncols < 1 || ncols > kBatchMax) return hipErrorInvalidValue;
  ColPtrs xs{}, us{};
  ColPtrsW ys{};
  for (int j = 0; j < ncols; ++j) {
    xs.p[j] = x[j];
    us.p[j] = u[j];
    ys.p[j] = y[j];
  }
  const int64_t m = int64_t(nn) * nn * nn;
  const unsigned nt = unsigned((m + kSymT - 1) / kSymT);
  const dim3 g(nt, nt, unsigned(ncols)), b(kSymT, 8);
  if (io == kIoF32) {
    hipLaunchKernelGGL((k_symadd<float, float2, float2>), g, b, 0, s, xs, us, ys, m, nn, vn, tinv, coef, sign);
  } else if (io == kIoMixed) {
    hipLaunchKernelGGL((k_symadd<float, double2, double2>), g, b, 0, s, xs, us, ys, m, nn, vn, tinv, coef, sign);
  } else {
    hipLaunchKernelGGL((k_symadd<double, double2, double2>), g, b, 0, s, xs, us, ys, m, nn, vn, tinv, coef,
                       sign);
  }
  return hipGetLastError();
}

hipError_t launch_symadd(int nn, int io, const void *x, const void *u, void *y,
                         const double *vn, const double *tinv, const MatvecCoef &coef, int sign,
                         hipStream_t s) {
  return launch_symadd_batch(nn, io, &x, &u, &y, 1, vn, tinv, coef, sign, s);
}

hipError_t launch_matvec_generic(int nn, const Dbl2 *x, Dbl2 *y, const double *vn,
                                 const double *tinv, const MatvecCoef &coef, hipStream_t s) {
  const double2 *xx = reinterpret_cast<const double2 *>(x);
  double2 *yy = reinterpret_cast<double2 *>(y);
  switch (nn) {
#define CASE(N) \
  case N:       \
    return matvec_generic_nn<N>(xx, yy, vn, tinv, coef, s);
    DEX_NN_LIST(CASE)
#undef CASE
    default:
      return hipErrorInvalidValue;
  }
}

// ---- block kernels ----------------------------------------------------------
namespace {

constexpr int kGramThreads = 256;
constexpr int kGramMaxL = 64;
constexpr int kGramMaxR = 128;
constexpr int kGramBlocks = 1024;

// Element r of a complex128 (fmt 0) or complex64 (fmt 1) column, in FP64.
__device__ __forceinline__ double2 ld_col(const void *p, int fmt, int64_t r) {
  if (fmt != 0) {
    const float2 v = reinterpret_cast<const float2 *>(p)[r];
    return make_double2(double(v.x), double(v.y));
  }
  return reinterpret_cast<const double2 *>(p)[r];
}

__device__ __forceinline__ void st_col(void *p, int fmt, int64_t r, double2 v) {
  if (fmt != 0)
    reinterpret_cast<float2 *>(p)[r] = make_float2(float(v.x), float(v.y));
  else
    reinterpret_cast<double2 *>(p)[r] = v;
}

// Two adjacent rows r, r + 1 of a complex128 (fmt 0) or complex64 (fmt 1)
// column in FP64; `two` false reads only row r (the last row of an odd-length
// column).  A complex64 pair is one 16-byte load, a complex128 pair two.
__device__ __forceinline__ void ld_col2(const void *p, int fmt, int64_t r, bool two, double2 &v0, double2 &v1) {
  if (fmt != 0) {
    const float2 *q = reinterpret_cast<const float2 *>(p) + r;
    if (two) {
      const float4 w = *reinterpret_cast<const float4 *>(q);
      v0 = make_double2(double(w.x), double(w.y));
      v1 = make_double2(double(w.z), double(w.w));
    } else {
      const float2 w = *q;
      v0 = make_double2(double(w.x), double(w.y));
      v1 = make_double2(0.0, 0.0);
    }
    return;
  }
  const double2 *q = reinterpret_cast<const double2 *>(p) + r;
  v0 = q[0];
  v1 = two ? q[1] : make_double2(0.0, 0.0);
}

__device__ __forceinline__ void st_col2(void *p, int fmt, int64_t r, bool two, double2 v0, double2 v1) {
  if (fmt != 0) {
    float2 *q = reinterpret_cast<float2 *>(p) + r;
    if (two)
      *reinterpret_cast<float4 *>(q) = make_float4(float(v0.x), float(v0.y), float(v1.x), float(v1.y));
    else
      *q = make_float2(float(v0.x), float(v0.y));
    return;
  }
  double2 *q = reinterpret_cast<double2 *>(p) + r;
  q[0] = v0;
  if (two) q[1] = v1;
}

__device__ __forceinline__ double2 col_value(const ColDesc &c, int64_t r) {
  const double2 va = ld_col(c.a, c.fmt, r);
  if (c.b == nullptr) return va;
  const double2 vb = ld_col(c.b, c.fmt, r);
  double2 o;
  o.x = va.x - (c.sig_re * vb.x - c.sig_im * vb.y);
  o.y = va.y - (c.sig_re * vb.y + c.sig_im * vb.x);
  return o;
}

// Partial Gram G = L^H R over a contiguous row range per workgroup, on the FP64
// matrix cores.  G = L^H R is a GEMM whose
// inner dimension is the vector length: per 4 rows, v_mfma_f64_16x16x4f64
// multiplies a 16 x 4 slice of L^T by a 4 x 16 slice of R into a 16 x 16
// output tile (lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] and
// holds D[4 v + l / 16][l % 16], v < 4).  A complex tile takes four real
// products: Re += Lr Rr + Li Ri, Im += Lr Ri - Li Rr.  The rows are staged
// through LDS (one coalesced pass over every column); each LDS value then
// feeds 16 MACs of a matrix instruction instead of one vector FMA (a vector-FMA
// version of this kernel was bound by the LDS bandwidth at ~2.5 TB/s).  The 4 waves of a
// workgroup share the output tiles (MT x NT of them) round-robin.
typedef double dex_v4d __attribute__((ext_vector_type(4)));

// Rows per staged tile: 16 (TALL = 32 for the narrowest shapes), halved while
// the tile exceeds 48 KB.  16-row tiles (8-33 KB) leave room for 4+ resident
// workgroups per CU; 32-row tiles measured 10-23% slower at 54-64 x 10 (2
// workgroups per CU, VGPR-bound).  A Gram of one 16 x 16 output tile (the
// inner solve's projections of 4-column groups) keeps few bytes in flight per
// workgroup with 16-row tiles: 32-row tiles measured 5-6% faster there at
// NN=21 (8 x 4, 16 x 4, 12 x 10) and 5-7% slower at 28 x 4 and 32 x 16.
constexpr int gram_mfma_rows(int mt, int nt, int tall = 16) {
  const int cap = tall >= 64 ? 64 * 1024 : 48 * 1024;  // tiles of up to 64 rows may take 64 KB of LDS
  int r = tall;
  while (r > 4 && r * (16 * mt + 16 * nt + 1) * 16 > cap) r /= 2;
  return r;
}

template <int MT, int NT, int TALL = 16>
__global__ __launch_bounds__(kGramThreads) void k_gram_mfma(const ColDesc *__restrict__ descs, int kl,
                                                            int kr, int share, int64_t d,
                                                            int64_t rows_per_block,
                                                            double2 *__restrict__ partial) {
  constexpr int LD = 16 * MT + 16 * NT + 1;
  constexpr int ROWS = gram_mfma_rows(MT, NT, TALL);
  constexpr int QMAX = (MT * NT + 3) / 4;
  // With fewer 16 x 16 output tiles than waves, KS waves share each tile and
  // split its rows (k steps) between them, their sums added through LDS at the
  // end: with one wave per tile, a one-tile Gram ran all its MFMAs on one SIMD
  // of four and every workgroup waited on them
  constexpr int KS = MT * NT == 1 ? 4 : (MT * NT == 2 ? 2 : 1);
  static_assert(gram_mfma_rows(MT, NT, TALL) % (4 * KS) == 0, "every wave takes whole row steps");
  // loads per thread per tile, each an adjacent row pair of one column (16-byte
  // accesses for complex64 and complex128 alike): every thread issues all of
  // them before any wait, and the next tile's while the matrix cores work on
  // this one
  constexpr int HALF = ROWS / 2;
  constexpr int PMAX = (HALF * (16 * MT + 16 * NT) + kGramThreads - 1) / kGramThreads;
  // the row tile, reused after the loop for the k-split sums (4 KB per wave)
  constexpr int TILE2 = ROWS * LD > MT * NT * (KS - 1) * 256 ? ROWS * LD : MT * NT * (KS - 1) * 256;
  __shared__ double2 tile[TILE2];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int ks = KS > 1 ? wave % KS : 0;  // this wave's share of the k steps
  const int nload = kl + (kr - share);
  const int mt_valid = (kl + 15) / 16, nt_valid = (kr + 15) / 16;
  int tm[QMAX], tn[QMAX];
  bool tv[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    const int t = KS > 1 ? wave / KS : wave + 4 * q;
    tm[q] = t % MT;
    tn[q] = t / MT;
    tv[q] = t < MT * NT && tm[q] < mt_valid && tn[q] < nt_valid &&
            !(tm[q] > tn[q] && 16 * (tm[q] + 1) <= share);
  }
  dex_v4d accr[QMAX], acci[QMAX];
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    accr[q] = dex_v4d{0.0, 0.0, 0.0, 0.0};
    acci[q] = dex_v4d{0.0, 0.0, 0.0, 0.0};
  }
  // Element p of this thread is (row pair e % HALF, column e / HALF) of every
  // tile, e = tid + 256 p: its column pointers are fixed for the whole loop.
  const void *pa[PMAX];
  const void *pb[PMAX];
  int pf[PMAX];
#pragma unroll
  for (int p = 0; p < PMAX; ++p) {
    const int e = tid + kGramThreads * p;
    pa[p] = nullptr;
    pb[p] = nullptr;
    pf[p] = 0;
    if (e < HALF * nload) {
      const ColDesc &dc = descs[e / HALF];
      pa[p] = dc.a;
      pb[p] = dc.b;
      pf[p] = dc.fmt;
    }
  }
  // rows_per_block > 0: this workgroup's rows are one contiguous range; 0: the
  // workgroups take tiles round robin (tile t to workgroup t % gridDim.x), so
  // the grid sweeps every column together through a narrow window of rows
  const bool strided = rows_per_block == 0;
  const int64_t r0 = strided ? int64_t(blockIdx.x) * ROWS : int64_t(blockIdx.x) * rows_per_block;
  const int64_t r1 = strided ? d : (r0 + rows_per_block < d ? r0 + rows_per_block : d);
  const int64_t rstep = strided ? int64_t(gridDim.x) * ROWS : ROWS;
  // One tile in flight per workgroup while the matrix cores work on the
  // previous one: three register stages (more bytes in flight per workgroup)
  // measured 4-8% slower on the narrow shapes (fewer resident workgroups)
  double2 va[PMAX][2], vb[PMAX][2];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      const int64_t row = rb + 2 * ((tid + kGramThreads * p) % HALF);
      va[p][0] = va[p][1] = vb[p][0] = vb[p][1] = make_double2(0.0, 0.0);
      if (pa[p] != nullptr && row < r1) {
        const bool two = row + 1 < r1;
        ld_col2(pa[p], pf[p], row, two, va[p][0], va[p][1]);
        if (pb[p] != nullptr) ld_col2(pb[p], pf[p], row, two, vb[p][0], vb[p][1]);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      if (pa[p] == nullptr) continue;  // beyond the loads, or supplied through a bslot
      const int e = tid + kGramThreads * p;
      const int r = 2 * (e % HALF), c = e / HALF;
      const int slot = c < kl ? c : 16 * MT + (c - kl);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (pb[p] == nullptr) {
          tile[(r + h) * LD + slot] = va[p][h];
        } else {
          const ColDesc &dc = descs[c];
          double2 o;
          o.x = va[p][h].x - (dc.sig_re * vb[p][h].x - dc.sig_im * vb[p][h].y);
          o.y = va[p][h].y - (dc.sig_re * vb[p][h].y + dc.sig_im * vb[p][h].x);
          tile[(r + h) * LD + slot] = o;
          if (dc.bslot >= 0) tile[(r + h) * LD + dc.bslot] = vb[p][h];
        }
      }
    }
  };
  if (r0 < r1) fetch(r0);
  for (int64_t rb = r0; rb < r1; rb += rstep) {
    store();
    __syncthreads();
    if (rb + rstep < r1) fetch(rb + rstep);
    // this wave's row steps: k0 = 4 (ks + KS s)
#pragma unroll
    for (int st = 0; st < ROWS / (4 * KS); ++st) {
      const int k0 = 4 * (ks + KS * st);
      const double2 *trow = tile + (k0 + lk) * LD;
#pragma unroll
      for (int q = 0; q < QMAX; ++q) {
        if (!tv[q]) continue;
        const double2 av = trow[16 * tm[q] + li];
        const int jj = 16 * tn[q] + li;
        const double2 bv = trow[jj < share ? jj : 16 * MT + (jj - share)];
        accr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.x, accr[q], 0, 0, 0);
        accr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, bv.y, accr[q], 0, 0, 0);
        acci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, bv.y, acci[q], 0, 0, 0);
        acci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.y, bv.x, acci[q], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  if constexpr (KS > 1) {
    // the waves with ks > 0 hand their sums to the tile's ks = 0 wave (the last
    // __syncthreads of the loop has freed the row tile)
    double *red = reinterpret_cast<double *>(tile);
    const int t = wave / KS;
    if (ks > 0 && tv[0]) {
      double *slot = red + size_t(t * (KS - 1) + ks - 1) * 512;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        slot[v * 64 + lane] = accr[0][v];
        slot[(4 + v) * 64 + lane] = acci[0][v];
      }
    }
    __syncthreads();
    if (ks == 0 && tv[0]) {
#pragma unroll
      for (int k = 1; k < KS; ++k) {
        const double *slot = red + size_t(t * (KS - 1) + k - 1) * 512;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          accr[0][v] += slot[v * 64 + lane];
          acci[0][v] += slot[(4 + v) * 64 + lane];
        }
      }
    }
  }
  double2 *out = partial + size_t(blockIdx.x) * kl * kr;
#pragma unroll
  for (int q = 0; q < QMAX; ++q) {
    if (!tv[q] || ks != 0) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = 16 * tm[q] + 4 * v + lk, j = 16 * tn[q] + li;
      if (i < kl && j < kr) out[i + size_t(kl) * j] = make_double2(accr[q][v], acci[q][v]);
    }
  }
}

// Partial Gram G = L^H R for narrow shapes (kl <= 8 M8 <= 32, kr <= 8: the
// inner solve's projections of 4-column groups) with the real and imaginary
// parts as separate MFMA rows and columns: A = [Re L_c; Im L_c] (16 rows for 8
// left columns), B = [Re R_r, Im R_r] (16 columns for 8 right columns), so one
// v_mfma_f64_16x16x4f64 per 4 rows forms all four real products of an 8 x 8
// complex block, O = [Lr Rr, Lr Ri; Li Rr, Li Ri], and G(c, r) =
// O(c, r) + O(c+8, r+8) + i (O(c, r+8) - O(c+8, r)) at the end.  The 16 x 16
// complex tiles of k_gram_mfma take 4 MFMAs for any kl <= 16, kr <= 16; this
// layout takes ceil(kl / 8).  Rows are staged in LDS as in k_gram_mfma (one
// row tile per workgroup, the next prefetched, tiles round robin over the
// grid); KS waves share each 8 x 8 block and split its row steps.  Right
// columns are unshared (share 0) or all shared with the first left columns
// (share == kr).  Layout of the partials as k_gram_mfma's.
template <int M8, int ROWS>
__global__ __launch_bounds__(kGramThreads) void k_gram_pk(const ColDesc *__restrict__ descs, int kl, int kr,
                                                          int share, int64_t d, double2 *__restrict__ partial) {
  constexpr int LD = 8 * M8 + 8 + 1;
  constexpr int KS = M8 == 1 ? 4 : (M8 == 2 ? 2 : 1);
  static_assert(ROWS % (4 * KS) == 0, "every wave takes whole row steps");
  constexpr int HALF = ROWS / 2;
  constexpr int PMAX = (HALF * (8 * M8 + 8) + kGramThreads - 1) / kGramThreads;
  // the row tile, reused after the loop for the waves' 16 x 16 sums (256 doubles each)
  constexpr int TILE2 = ROWS * LD > M8 * KS * 128 ? ROWS * LD : M8 * KS * 128;
  __shared__ double2 tile[TILE2];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int t = KS > 1 ? wave / KS : wave;  // this wave's 8-column block of L
  const int ks = KS > 1 ? wave % KS : 0;    // and its share of the row steps
  const bool tv = t < M8 && 8 * t < kl;
  const int nload = kl + (kr - share);
  dex_v4d acc = dex_v4d{0.0, 0.0, 0.0, 0.0};
  const void *pa[PMAX];
  const void *pb[PMAX];
  int pf[PMAX];
#pragma unroll
  for (int p = 0; p < PMAX; ++p) {
    const int e = tid + kGramThreads * p;
    pa[p] = nullptr;
    pb[p] = nullptr;
    pf[p] = 0;
    if (e < HALF * nload) {
      const ColDesc &dc = descs[e / HALF];
      pa[p] = dc.a;
      pb[p] = dc.b;
      pf[p] = dc.fmt;
    }
  }
  const int64_t r0 = int64_t(blockIdx.x) * ROWS;
  const int64_t rstep = int64_t(gridDim.x) * ROWS;
  double2 va[PMAX][2], vb[PMAX][2];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      const int64_t row = rb + 2 * ((tid + kGramThreads * p) % HALF);
      va[p][0] = va[p][1] = vb[p][0] = vb[p][1] = make_double2(0.0, 0.0);
      if (pa[p] != nullptr && row < d) {
        const bool two = row + 1 < d;
        ld_col2(pa[p], pf[p], row, two, va[p][0], va[p][1]);
        if (pb[p] != nullptr) ld_col2(pb[p], pf[p], row, two, vb[p][0], vb[p][1]);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      if (pa[p] == nullptr) continue;
      const int e = tid + kGramThreads * p;
      const int r = 2 * (e % HALF), c = e / HALF;
      const int slot = c < kl ? c : 8 * M8 + (c - kl);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double2 o = va[p][h];
        if (pb[p] != nullptr) {
          const ColDesc &dc = descs[c];
          o.x -= dc.sig_re * vb[p][h].x - dc.sig_im * vb[p][h].y;
          o.y -= dc.sig_re * vb[p][h].y + dc.sig_im * vb[p][h].x;
        }
        tile[(r + h) * LD + slot] = o;
      }
    }
  };
  // this lane's A row (a part of left column 8 t + (li & 7)) and B column (a
  // part of right column li & 7; shared right columns are the first left ones)
  const int acol = 8 * t + (li & 7);
  const int rj = li & 7;
  const int bcol = rj < share ? rj : 8 * M8 + (rj - share);
  const bool im = li >= 8;
  if (r0 < d) fetch(r0);
  for (int64_t rb = r0; rb < d; rb += rstep) {
    store();
    __syncthreads();
    if (rb + rstep < d) fetch(rb + rstep);
    if (tv) {
#pragma unroll
      for (int st = 0; st < ROWS / (4 * KS); ++st) {
        const double2 *trow = tile + (4 * (ks + KS * st) + lk) * LD;
        const double2 av = trow[acol];
        const double2 bv = trow[bcol];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(im ? av.y : av.x, im ? bv.y : bv.x, acc, 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // every wave's 16 x 16 sums to LDS (the loop's last __syncthreads freed the
  // tile); element (i, j) of the block, i = 4 v + lk, j = li, sits at
  // [v * 64 + lane] of the wave's 256 doubles
  double *red = reinterpret_cast<double *>(tile);
  if (tv) {
#pragma unroll
    for (int v = 0; v < 4; ++v) red[size_t(t * KS + ks) * 256 + v * 64 + lane] = acc[v];
  }
  __syncthreads();
  double2 *out = partial + size_t(blockIdx.x) * kl * kr;
  for (int e = tid; e < M8 * 64; e += kGramThreads) {
    const int tb = e / 64, c = (e % 64) / 8, r = e % 8;
    const int i = 8 * tb + c;
    if (i >= kl || r >= kr) continue;
    auto O = [&](int k, int ii, int jj) {
      return red[size_t(tb * KS + k) * 256 + (ii / 4) * 64 + (jj + 16 * (ii % 4))];
    };
    double gr = 0.0, gi = 0.0;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      gr += O(k, c, r) + O(k, c + 8, r + 8);
      gi += O(k, c, r + 8) - O(k, c + 8, r);
    }
    out[i + size_t(kl) * r] = make_double2(gr, gi);
  }
}

// Sum of nparts partial results per element: one wavefront per element, each
// lane summing a strided subset, then a shuffle tree (fixed order).
__global__ __launch_bounds__(256) void k_reduce_partials(const double2 *__restrict__ partial,
                                                         int nparts, int nelem,
                                                         double2 *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= nelem) return;
  double sx = 0.0, sy = 0.0;
  for (int b = lane; b < nparts; b += 64) {
    const double2 v = partial[size_t(b) * nelem + e];
    sx += v.x;
    sy += v.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sx += __shfl_down(sx, off);
    sy += __shfl_down(sy, off);
  }
  if (lane == 0) out[e] = make_double2(sx, sy);
}

constexpr int kCombRows = 32;
constexpr int kCombThreads = 256;
constexpr int kCombMaxIn = 64;
constexpr int kCombGroups = 8;  // output groups per workgroup (kCombThreads / kCombRows)
constexpr int kCombMaxOut = 48;

// out_j = sum_i in_i C(i, j) (+ beta_j out_j), row by row.  A workgroup stages
// C (nin x nout, padded to kCombGroups * JPT columns) once in LDS, then walks
// tiles of kCombRows rows: the tile of every input column goes to LDS, and
// thread t forms, for row t % kCombRows, the JPT outputs of group
// t / kCombRows.  All loads of a tile finish before any of its outputs is
// written, so an output may alias an input.
template <int JPT>
__global__ __launch_bounds__(kCombThreads) void k_combine(const ColDesc *__restrict__ descs, int nin,
                                                          int nout, const double2 *__restrict__ cm,
                                                          const double2 *__restrict__ beta, int64_t d) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NP = kCombGroups * JPT;
  // input elements per thread per tile: all issued before any wait, the next
  // tile's while this one is combined; each an adjacent row pair of one
  // column (16-byte accesses for complex64 and complex128 alike)
  constexpr int HALF = kCombRows / 2;
  constexpr int PMAX = HALF * kCombMaxIn / kCombThreads;
  double2 *sc = reinterpret_cast<double2 *>(smem);  // [nin][NP]
  double2 *tin = sc + size_t(nin) * NP;              // [nin][kCombRows]
  const int tid = threadIdx.x;
  const int row = tid % kCombRows;
  const int grp = tid / kCombRows;
  for (int e = tid; e < nin * NP; e += kCombThreads) {
    const int i = e / NP, j = e % NP;
    sc[e] = j < nout ? cm[i + size_t(nin) * j] : make_double2(0.0, 0.0);
  }
  // element p of this thread: row pair e % HALF of input column e / HALF
  const void *pa[PMAX];
  const void *pb[PMAX];
  int pf[PMAX];
#pragma unroll
  for (int p = 0; p < PMAX; ++p) {
    const int e = tid + kCombThreads * p;
    pa[p] = nullptr;
    pb[p] = nullptr;
    pf[p] = 0;
    if (e < HALF * nin) {
      const ColDesc &dc = descs[e / HALF];
      pa[p] = dc.a;
      pb[p] = dc.b;
      pf[p] = dc.fmt;
    }
  }
  double2 va[PMAX][2], vb[PMAX][2];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      const int64_t r = rb + 2 * ((tid + kCombThreads * p) % HALF);
      if (pa[p] != nullptr && r < d) {
        ld_col2(pa[p], pf[p], r, r + 1 < d, va[p][0], va[p][1]);
        if (pb[p] != nullptr) ld_col2(pb[p], pf[p], r, r + 1 < d, vb[p][0], vb[p][1]);
      }
    }
  };
  const int64_t ntile = (d + kCombRows - 1) / kCombRows;
  if (int64_t(blockIdx.x) < ntile) fetch(int64_t(blockIdx.x) * kCombRows);
  for (int64_t tb = blockIdx.x; tb < ntile; tb += gridDim.x) {
    const int64_t rb = tb * kCombRows;
    const int nr = int(d - rb < kCombRows ? d - rb : kCombRows);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PMAX; ++p) {
      if (pa[p] == nullptr) continue;
      const int e = tid + kCombThreads * p;
      const int r = 2 * (e % HALF), c = e / HALF;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (r + h >= nr) continue;
        double2 v = va[p][h];
        if (pb[p] != nullptr) {
          const ColDesc &dc = descs[c];
          v.x -= dc.sig_re * vb[p][h].x - dc.sig_im * vb[p][h].y;
          v.y -= dc.sig_re * vb[p][h].y + dc.sig_im * vb[p][h].x;
        }
        tin[c * kCombRows + r + h] = v;
      }
    }
    __syncthreads();
    if (tb + gridDim.x < ntile) fetch((tb + gridDim.x) * kCombRows);
    if (row < nr) {
      double ax[JPT], ay[JPT];
#pragma unroll
      for (int q = 0; q < JPT; ++q) ax[q] = ay[q] = 0.0;
      const double2 *cg = sc + grp * JPT;
#pragma unroll 4
      for (int i = 0; i < nin; ++i) {
        const double2 a = tin[i * kCombRows + row];
#pragma unroll
        for (int q = 0; q < JPT; ++q) {
          const double2 c = cg[i * NP + q];
          ax[q] = fma(a.x, c.x, ax[q]);
          ax[q] = fma(-a.y, c.y, ax[q]);
          ay[q] = fma(a.x, c.y, ay[q]);
          ay[q] = fma(a.y, c.x, ay[q]);
        }
      }
#pragma unroll
      for (int q = 0; q < JPT; ++q) {
        const int j = grp * JPT + q;
        if (j < nout) {
          const ColDesc &od = descs[nin + j];
          void *o = const_cast<Dbl2 *>(od.a);
          const int64_t r = rb + row;
          double ox = ax[q], oy = ay[q];
          if (beta != nullptr) {
            const double2 bj = beta[j];
            const double2 ov = ld_col(o, od.fmt, r);
            ox += bj.x * ov.x - bj.y * ov.y;
            oy += bj.x * ov.y + bj.y * ov.x;
          }
          st_col(o, od.fmt, r, make_double2(ox, oy));
        }
      }
    }
  }
}

// The combine for narrow outputs (nout <= 16; also 33..48, where the LDS
// kernel's staging would exceed 64 KB per workgroup), without LDS: a wavefront
// owns 64 * RPL consecutive rows and every output column of them, so the
// coefficients C(i, j) are wavefront-uniform (scalar loads, no LDS traffic)
// and each input element is read once, by one lane.  With RPL = 2 a lane owns
// two adjacent rows, so every load and store is 16 bytes wide whatever the
// storage (one complex64 pair, or two complex128 halves of a 32-byte run) and a
// wavefront reads 2 KB (complex128) or 1 KB (complex64) contiguous per column:
// with one complex64 row per lane the loads were 8 bytes wide and the kernel
// moved half the bytes per instruction.  With RPL = 1 a lane owns one row.
// JN, the output width the registers hold, is the smallest of 4, 8, 12, 16, 48
// that covers nout: spare accumulators cost occupancy.  UNR inputs are
// unrolled.  All inputs of a row are read before any output of it is written,
// so outputs may alias inputs.  Columns start 256-byte aligned and row pairs
// start at even rows, so the 16-byte accesses are aligned.
template <int JN, int RPL, int UNR>
__global__ __launch_bounds__(256) void k_combine_w(const ColDesc *__restrict__ descs, int nin, int nout,
                                                   const double2 *__restrict__ cm,
                                                   const double2 *__restrict__ beta, int64_t d) {
  static_assert(RPL == 1 || RPL == 2, "one row or an adjacent row pair per lane");
  const int lane = threadIdx.x & 63;
  const int64_t wave = int64_t(blockIdx.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t stride = int64_t(gridDim.x) * 4 * 64 * RPL;
  for (int64_t rb = wave * 64 * RPL; rb < d; rb += stride) {
    const int64_t r0 = rb + RPL * lane;  // this lane's first row
    const bool ok0 = r0 < d, ok1 = RPL == 2 && r0 + 1 < d;
    double ax[RPL][JN], ay[RPL][JN];
#pragma unroll
    for (int k = 0; k < RPL; ++k)
#pragma unroll
      for (int q = 0; q < JN; ++q) ax[k][q] = ay[k][q] = 0.0;
#pragma unroll UNR
    for (int i = 0; i < nin; ++i) {
      const ColDesc dc = descs[i];
      double2 v[RPL];
#pragma unroll
      for (int k = 0; k < RPL; ++k) v[k] = make_double2(0.0, 0.0);
      if (ok0) {
        if constexpr (RPL == 2) {
          ld_col2(dc.a, dc.fmt, r0, ok1, v[0], v[RPL - 1]);
          if (dc.b != nullptr) {
            double2 b0, b1;
            ld_col2(dc.b, dc.fmt, r0, ok1, b0, b1);
            v[0].x -= dc.sig_re * b0.x - dc.sig_im * b0.y;
            v[0].y -= dc.sig_re * b0.y + dc.sig_im * b0.x;
            v[RPL - 1].x -= dc.sig_re * b1.x - dc.sig_im * b1.y;
            v[RPL - 1].y -= dc.sig_re * b1.y + dc.sig_im * b1.x;
          }
        } else {
          v[0] = ld_col(dc.a, dc.fmt, r0);
          if (dc.b != nullptr) {
            const double2 vb = ld_col(dc.b, dc.fmt, r0);
            v[0].x -= dc.sig_re * vb.x - dc.sig_im * vb.y;
            v[0].y -= dc.sig_re * vb.y + dc.sig_im * vb.x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < JN; ++q) {
        if (q < nout) {
          const double2 c = cm[i + size_t(nin) * q];
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            ax[k][q] = fma(v[k].x, c.x, ax[k][q]);
            ax[k][q] = fma(-v[k].y, c.y, ax[k][q]);
            ay[k][q] = fma(v[k].x, c.y, ay[k][q]);
            ay[k][q] = fma(v[k].y, c.x, ay[k][q]);
          }
        }
      }
    }
    if (!ok0) continue;
#pragma unroll
    for (int q = 0; q < JN; ++q) {
      if (q < nout) {
        const ColDesc &od = descs[nin + q];
        void *o = const_cast<Dbl2 *>(od.a);
        double2 out[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) out[k] = make_double2(ax[k][q], ay[k][q]);
        if (beta != nullptr) {
          const double2 bq = beta[q];
          double2 ov[RPL];
          if constexpr (RPL == 2)
            ld_col2(o, od.fmt, r0, ok1, ov[0], ov[RPL - 1]);
          else
            ov[0] = ld_col(o, od.fmt, r0);
#pragma unroll
          for (int k = 0; k < RPL; ++k) {
            out[k].x += bq.x * ov[k].x - bq.y * ov[k].y;
            out[k].y += bq.x * ov[k].y + bq.y * ov[k].x;
          }
        }
        if constexpr (RPL == 2)
          st_col2(o, od.fmt, r0, ok1, out[0], out[RPL - 1]);
        else
          st_col(o, od.fmt, r0, out[0]);
      }
    }
  }
}

constexpr int kDotThreads = 256;
constexpr int kDotBlocks = 512;
constexpr int kDotMaxCols = 64;

// Per column: |a|^2, conj(a) b (re, im), |b|^2 -> 4 doubles per column per block.
__global__ __launch_bounds__(kDotThreads) void k_coldots(const ColDesc *__restrict__ descs, int k,
                                                         int64_t d, double *__restrict__ partial) {
  __shared__ double red[kDotThreads / 64][4];
  const int tid = threadIdx.x;
  for (int j = 0; j < k; ++j) {
    const ColDesc ca = descs[j];
    const ColDesc cb = descs[k + j];
    double aa = 0.0, abr = 0.0, abi = 0.0, bb = 0.0;
    for (int64_t r = int64_t(blockIdx.x) * kDotThreads + tid; r < d; r += int64_t(gridDim.x) * kDotThreads) {
      const double2 a = col_value(ca, r);
      const double2 b = cb.a != nullptr ? col_value(cb, r) : make_double2(0.0, 0.0);
      aa += a.x * a.x + a.y * a.y;
      abr += a.x * b.x + a.y * b.y;
      abi += a.x * b.y - a.y * b.x;
      bb += b.x * b.x + b.y * b.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
      aa += __shfl_down(aa, off);
      abr += __shfl_down(abr, off);
      abi += __shfl_down(abi, off);
      bb += __shfl_down(bb, off);
    }
    if ((tid & 63) == 0) {
      red[tid >> 6][0] = aa;
      red[tid >> 6][1] = abr;
      red[tid >> 6][2] = abi;
      red[tid >> 6][3] = bb;
    }
    __syncthreads();
    if (tid < 4) {
      double s = 0.0;
      for (int w = 0; w < kDotThreads / 64; ++w) s += red[w][tid];
      partial[(size_t(blockIdx.x) * k + j) * 4 + tid] = s;
    }
    __syncthreads();
  }
}

// The residual step of GPLMR (gplmr.F90:324-341) for column j = blockIdx.y
// after the scaling s_j = 1 / ||X_j|| and the Rayleigh quotient E_j are known:
// X_j <- s_j X_j, AX_j <- s_j AX_j, W_j = AX_j - E_j X_j (scaled), one pass over
// X and AX (row pairs, 16-byte accesses), with each workgroup's partial sum of
// |W_j|^2 (in FP64, before W is rounded to its storage) in
// partial[blockIdx.x * nev + j].  descs: X (nev), AX (nev), W (nev).
__global__ __launch_bounds__(256) void k_resid(const ColDesc *__restrict__ descs, int nev,
                                               const double *__restrict__ sc, const double2 *__restrict__ ev,
                                               int64_t d, int64_t rows_per_block, double *__restrict__ partial) {
  const int j = blockIdx.y;
  const ColDesc cx = descs[j], cax = descs[nev + j], cw = descs[2 * nev + j];
  const double s = sc[j];
  const double2 f = ev[j];
  const int64_t r0 = int64_t(blockIdx.x) * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < d ? r0 + rows_per_block : d;
  double acc = 0.0;
  for (int64_t r = r0 + 2 * int64_t(threadIdx.x); r < r1; r += 2 * int64_t(blockDim.x)) {
    const bool two = r + 1 < r1;
    double2 x[2], a[2], w[2];
    ld_col2(cx.a, cx.fmt, r, two, x[0], x[1]);
    ld_col2(cax.a, cax.fmt, r, two, a[0], a[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      x[h].x *= s;
      x[h].y *= s;
      a[h].x *= s;
      a[h].y *= s;
      w[h].x = a[h].x - (f.x * x[h].x - f.y * x[h].y);
      w[h].y = a[h].y - (f.x * x[h].y + f.y * x[h].x);
    }
    st_col2(const_cast<Dbl2 *>(cx.a), cx.fmt, r, two, x[0], x[1]);
    st_col2(const_cast<Dbl2 *>(cax.a), cax.fmt, r, two, a[0], a[1]);
    st_col2(const_cast<Dbl2 *>(cw.a), cw.fmt, r, two, w[0], w[1]);
    acc += w[0].x * w[0].x + w[0].y * w[0].y;
    if (two) acc += w[1].x * w[1].x + w[1].y * w[1].y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[size_t(blockIdx.x) * nev + j] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_reduce_real(const double *__restrict__ partial, int nparts,
                                                     int nelem, double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= nelem) return;
  double s = 0.0;
  for (int b = lane; b < nparts; b += 64) s += partial[size_t(b) * nelem + e];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) out[e] = s;
}

// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(M0) * c[0];
    const uint64_t p1 = uint64_t(M1) * c[2];
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

__global__ void k_fill_gaussian(double2 *__restrict__ x, int64_t n, uint64_t seed, uint32_t col,
                                uint32_t sid) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c[4] = {uint32_t(i), uint32_t(uint64_t(i) >> 32), col, sid};
  philox10(c, uint32_t(seed), uint32_t(seed >> 32));
  // 53-bit uniforms in (0, 1]
  const double u1 = (double((uint64_t(c[0]) << 21) ^ (c[1] >> 11)) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = double((uint64_t(c[2]) << 21) ^ (c[3] >> 11)) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586 * u2;
  double sn, cs;
  sincos(ang, &sn, &cs);
  const double s = 0.7071067811865476;
  x[i] = make_double2(rad * cs * s, rad * sn * s);
}

// X <- (X + sign X^T)/2 on the M x M matrix, tiles of 32 x 32 above the diagonal.
__global__ void k_singlet(double2 *__restrict__ x, int64_t m, double sign) {
  __shared__ double2 ta[32][33], tb[32][33];
  const int64_t bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = bi * 32 + tx, j = bj * 32 + r;  // element (i, j): x[i + m j]
    ta[r][tx] = (i < m && j < m) ? x[i + m * j] : make_double2(0, 0);
    const int64_t i2 = bj * 32 + tx, j2 = bi * 32 + r;
    tb[r][tx] = (i2 < m && j2 < m) ? x[i2 + m * j2] : make_double2(0, 0);
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int64_t i = bi * 32 + tx, j = bj * 32 + r;
    const double2 a = ta[r][tx], b = tb[tx][r];  // b = x[j + m i]
    const double2 s = make_double2(0.5 * (a.x + sign * b.x), 0.5 * (a.y + sign * b.y));
    if (i < m && j < m) x[i + m * j] = s;
    const int64_t i2 = bj * 32 + tx, j2 = bi * 32 + r;
    const double2 a2 = tb[r][tx], b2 = ta[tx][r];
    const double2 s2 = make_double2(0.5 * (a2.x + sign * b2.x), 0.5 * (a2.y + sign * b2.y));
    if (i2 < m && j2 < m) x[i2 + m * j2] = s2;
  }
}

__global__ void k_to_single(const double2 *__restrict__ x, float2 *__restrict__ y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = make_float2(float(x[i].x), float(x[i].y));
}

__global__ void k_to_double(const float2 *__restrict__ x, double2 *__restrict__ y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = make_double2(double(x[i].x), double(x[i].y));
}

// ---- one-electron (He+) operator and the small block products of its Lanczos --
// H1 x = s2 tau3D x + c1 vn o x on vectors of M = NN^3 entries (i fastest):
// KEMULT1 (hammod.F90:71-134) with the 1/SCALE^2 of the kinetic matrix folded
// into s2 and POTENTIAL1 (hammod.F90:432-498) / SCALE into c1 vn.  One thread
// per entry of one column; the column (at most 315 KB at NN = 27) stays in L2.
__global__ __launch_bounds__(256) void k_h1(const double2 *__restrict__ x, double2 *__restrict__ y, int nn,
                                            int64_t m, int ncols, const double *__restrict__ vn,
                                            MatvecCoef cf) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * ncols) return;
  const int64_t c = e / m, p = e % m;
  const int i = int(p % nn), j = int((p / nn) % nn), k = int(p / (int64_t(nn) * nn));
  const double2 *xc = x + c * m;
  const int64_t base_i = p - i, base_j = p - int64_t(nn) * j, base_k = p - int64_t(nn) * nn * k;
  double sr = 0.0, si = 0.0;
  for (int l = 0; l < nn; ++l) {
    const double ti = tau_rt(l - i), tj = tau_rt(l - j), tk = tau_rt(l - k);
    const double2 a = xc[base_i + l], b = xc[base_j + int64_t(nn) * l], d = xc[base_k + int64_t(nn) * nn * l];
    sr += ti * a.x + tj * b.x + tk * d.x;
    si += ti * a.y + tj * b.y + tk * d.y;
  }
  const double2 xv = xc[p];
  const double v = vn[p];
  double2 o;
  o.x = cf.s2_re * sr - cf.s2_im * si + v * (cf.c1_re * xv.x - cf.c1_im * xv.y);
  o.y = cf.s2_re * si + cf.s2_im * sr + v * (cf.c1_re * xv.y + cf.c1_im * xv.x);
  y[e] = o;
}

// out(i, j) = sum_r op(A_i[r]) B_j[r] over m rows, op = identity (the
// bilinear product of the complex-symmetric Lanczos, MPIDOT) or conj (the
// Hermitian one, MPIHERMDOT); A (m x p) and B (m x q) column-major with
// leading dimension m, out (p x q) column-major.  One workgroup per output
// element, a fixed-order reduction.
__global__ __launch_bounds__(256) void k_small_dots(const double2 *__restrict__ a, int p,
                                                    const double2 *__restrict__ b, int q, int64_t m,
                                                    int herm, double2 *__restrict__ out) {
  const int i = blockIdx.x % p, j = blockIdx.x / p;
  const double2 *ai = a + int64_t(i) * m;
  const double2 *bj = b + int64_t(j) * m;
  const double s = herm ? -1.0 : 1.0;
  double sr = 0.0, si = 0.0;
  for (int64_t r = threadIdx.x; r < m; r += blockDim.x) {
    const double2 u = ai[r], v = bj[r];
    const double uy = s * u.y;
    sr += u.x * v.x - uy * v.y;
    si += u.x * v.y + uy * v.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) {
      red[0][threadIdx.x] += red[0][threadIdx.x + off];
      red[1][threadIdx.x] += red[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i + int64_t(p) * j] = make_double2(red[0][0], red[1][0]);
}

// Y = beta Y + A C for A (m x p, leading dimension m), C (p x q, leading
// dimension p), Y (m x q, leading dimension m); beta 0 or 1.  One thread per
// element of Y.
__global__ __launch_bounds__(256) void k_small_lincomb(const double2 *__restrict__ a, int p,
                                                       const double2 *__restrict__ cm, int q, int64_t m,
                                                       int beta, double2 *__restrict__ y) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * q) return;
  const int64_t r = e % m;
  const int j = int(e / m);
  double sr = 0.0, si = 0.0;
  for (int i = 0; i < p; ++i) {
    const double2 u = a[r + m * i], c = cm[i + int64_t(p) * j];
    sr += u.x * c.x - u.y * c.y;
    si += u.x * c.y + u.y * c.x;
  }
  if (beta) {
    const double2 o = y[e];
    sr += o.x;
    si += o.y;
  }
  y[e] = make_double2(sr, si);
}

// Two-electron vector of rank <= K over one-electron orbitals:
// out(p, q) = sum_b T(p, b) X(q, b), p electron 1 (fastest), q electron 2;
// T and X are M x K column-major.  One thread per element.
__global__ __launch_bounds__(256) void k_lowrank(const double2 *__restrict__ t, const double2 *__restrict__ x,
                                                 double2 *__restrict__ out, int64_t m, int k) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m * m) return;
  const int64_t p = e % m, q = e / m;
  double re = 0.0, im = 0.0;
  for (int b = 0; b < k; ++b) {
    const double2 tv = t[p + m * b];
    const double2 xv = x[q + m * b];
    re += tv.x * xv.x - tv.y * xv.y;
    im += tv.x * xv.y + tv.y * xv.x;
  }
  out[e] = make_double2(re, im);
}

// W = Y X (or Y conj(X)) for the M x M matrix Y (a two-electron vector,
// column q = electron-2 index contiguous) and the M x K orbital block X:
// each thread owns one row p and all K outputs; X(q, .) is the same for the
// whole workgroup (L1-resident), Y is read once, coalesced along p.
template <int KM>
__global__ __launch_bounds__(256) void k_ywx(const double2 *__restrict__ y, const double2 *__restrict__ x,
                                             double2 *__restrict__ w, int64_t m, int k, int conj_x) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= m) return;
  double re[KM], im[KM];
#pragma unroll
  for (int b = 0; b < KM; ++b) re[b] = im[b] = 0.0;
  const double sx = conj_x ? -1.0 : 1.0;
  for (int64_t q = 0; q < m; ++q) {
    const double2 yv = y[p + m * q];
#pragma unroll
    for (int b = 0; b < KM; ++b) {
      if (b < k) {
        const double2 xv = x[q + m * b];
        const double xi = sx * xv.y;
        re[b] += yv.x * xv.x - yv.y * xi;
        im[b] += yv.x * xi + yv.y * xv.x;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < KM; ++b)
    if (b < k) w[p + m * b] = make_double2(re[b], im[b]);
}

}  // namespace

hipError_t launch_lowrank(const Dbl2 *t, const Dbl2 *x, Dbl2 *out, int64_t m, int k, hipStream_t s) {
  if (k < 1 || m < 1) return hipErrorInvalidValue;
  const int64_t n = m * m;
  hipLaunchKernelGGL(k_lowrank, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2 *>(t), reinterpret_cast<const double2 *>(x),
                     reinterpret_cast<double2 *>(out), m, k);
  return hipGetLastError();
}

hipError_t launch_ywx(const Dbl2 *y, const Dbl2 *x, Dbl2 *w, int64_t m, int k, int conj_x, hipStream_t s) {
  if (k < 1 || k > 32 || m < 1) return hipErrorInvalidValue;
  const dim3 grid(unsigned((m + 255) / 256)), block(256);
  const double2 *yy = reinterpret_cast<const double2 *>(y), *xx = reinterpret_cast<const double2 *>(x);
  double2 *ww = reinterpret_cast<double2 *>(w);
  if (k <= 16)
    hipLaunchKernelGGL(k_ywx<16>, grid, block, 0, s, yy, xx, ww, m, k, conj_x);
  else
    hipLaunchKernelGGL(k_ywx<32>, grid, block, 0, s, yy, xx, ww, m, k, conj_x);
  return hipGetLastError();
}

int gram_blocks() { return kGramBlocks; }

size_t gram_partial_elems(int kl, int kr) { return size_t(kGramBlocks) * kl * kr; }

namespace {
// Workgroups of k_gram_mfma<MA, NB> that fit on the device at once (LDS sets
// it: 3 or 4 per CU), capped at kGramBlocks.  A grid of exactly that many runs
// in one wave of workgroups; 1024 workgroups at 3 per CU would leave a second
// round on a third of the chip.
template <int MA, int NB, int TALL>
int gram_grid() {
  static std::atomic<int> cached{0};
  int grid = cached.load(std::memory_order_acquire);
  if (grid != 0) return grid;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  grid = cached.load(std::memory_order_acquire);
  if (grid != 0) return grid;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(&k_gram_mfma<MA, NB, TALL>), kGramThreads, 0) != hipSuccess ||
      cus < 1 || per_cu < 1)
    return kGramBlocks;
  grid = cus * per_cu < kGramBlocks ? cus * per_cu : kGramBlocks;
  cached.store(grid, std::memory_order_release);
  return grid;
}

// Tile height choice (read once): DEX_GRAM_ROWS=16 keeps 16-row tiles
// everywhere, 64 takes up-to-64-row tiles (64 KB of LDS) for every shape;
// otherwise 32-row tiles for single-tile Grams and 16 elsewhere.
int gram_rows_choice() {
  static const int rows = [] {
    const char *v = std::getenv("DEX_GRAM_ROWS");
    return v ? std::atoi(v) : 32;
  }();
  return rows;
}

// The packed narrow Gram (on unless DEX_GRAM_PK=0, read once).
bool gram_pk_on() {
  static const bool on = [] {
    const char *v = std::getenv("DEX_GRAM_PK");
    return !(v && std::atoi(v) == 0);
  }();
  return on;
}

// Row-tile height of the packed Gram: 64, or DEX_PK_ROWS = 32 or 128 (read once).
// Measured on whole NN=21 block-16 solves: 64 rows took 5-7% less Gram time
// than 32, 128 rows 1-3% less.
int pk_rows_choice() {
  static const int rows = [] {
    const char *v = std::getenv("DEX_PK_ROWS");
    const int r = v ? std::atoi(v) : 64;
    return r == 32 || r == 128 ? r : 64;
  }();
  return rows;
}

template <int M8, int ROWS>
int gram_pk_grid() {
  static std::atomic<int> cached{0};
  int grid = cached.load(std::memory_order_acquire);
  if (grid != 0) return grid;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  grid = cached.load(std::memory_order_acquire);
  if (grid != 0) return grid;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(&k_gram_pk<M8, ROWS>), kGramThreads, 0) != hipSuccess ||
      cus < 1 || per_cu < 1)
    return kGramBlocks;
  grid = cus * per_cu < kGramBlocks ? cus * per_cu : kGramBlocks;
  cached.store(grid, std::memory_order_release);
  return grid;
}

template <int M8, int ROWS>
int gram_pk_launch_r(const ColDesc *descs, int kl, int kr, int share, int64_t d, double2 *partial,
                     hipStream_t s) {
  const int grid = gram_pk_grid<M8, ROWS>();
  hipLaunchKernelGGL((k_gram_pk<M8, ROWS>), dim3(grid), dim3(kGramThreads), 0, s, descs, kl, kr, share, d,
                     partial);
  return grid;
}

template <int M8>
int gram_pk_launch(const ColDesc *descs, int kl, int kr, int share, int64_t d, double2 *partial, hipStream_t s) {
  switch (pk_rows_choice()) {
    case 128: return gram_pk_launch_r<M8, 128>(descs, kl, kr, share, d, partial, s);
    case 64: return gram_pk_launch_r<M8, 64>(descs, kl, kr, share, d, partial, s);
    default: return gram_pk_launch_r<M8, 32>(descs, kl, kr, share, d, partial, s);
  }
}

template <int M8>
void gram_pk_grids() {
  const int r = pk_rows_choice();
  if (r == 128) (void)gram_pk_grid<M8, 128>();
  else if (r == 64) (void)gram_pk_grid<M8, 64>();
  else (void)gram_pk_grid<M8, 32>();
}

template <int MA>
void gram_grids_nb() {
  if (gram_rows_choice() == 64) {
    (void)gram_grid<MA, 1, 64>();
    (void)gram_grid<MA, 2, 64>();
    (void)gram_grid<MA, 3, 64>();
    (void)gram_grid<MA, 4, 64>();
    (void)gram_grid<MA, 5, 64>();
    (void)gram_grid<MA, 6, 64>();
    (void)gram_grid<MA, 7, 64>();
    (void)gram_grid<MA, 8, 64>();
  }
  (void)gram_grid<MA, 1, 16>();
  (void)gram_grid<MA, 2, 16>();
  (void)gram_grid<MA, 3, 16>();
  (void)gram_grid<MA, 4, 16>();
  (void)gram_grid<MA, 5, 16>();
  (void)gram_grid<MA, 6, 16>();
  (void)gram_grid<MA, 7, 16>();
  (void)gram_grid<MA, 8, 16>();
  if constexpr (MA == 1) (void)gram_grid<MA, 1, 32>();
}

// Round-robin tiles (the default) or, with DEX_GRAM_STRIDE=0 (read once), one
// contiguous row range per workgroup.  Measured on whole NN=21 block-16 solves:
// Gram time within 1% either way (complex128 and complex64 columns).
bool gram_stride_on() {
  static const bool on = [] {
    const char *v = std::getenv("DEX_GRAM_STRIDE");
    return !(v && std::atoi(v) == 0);
  }();
  return on;
}

template <int MA, int NB, int TALL>
int gram_launch_t(const ColDesc *descs, int kl, int kr, int share, int64_t d, double2 *partial,
                  hipStream_t s) {
  const int grid = gram_grid<MA, NB, TALL>();
  int64_t rpb = (d + grid - 1) / grid;
  rpb = (rpb + 63) / 64 * 64;  // a whole number of tiles for every tile height
  if (gram_stride_on()) rpb = 0;
  hipLaunchKernelGGL((k_gram_mfma<MA, NB, TALL>), dim3(grid), dim3(kGramThreads), 0, s, descs, kl, kr, share,
                     d, rpb, partial);
  return grid;
}

template <int MA, int NB>
int gram_launch(const ColDesc *descs, int kl, int kr, int share, int64_t d, double2 *partial,
                hipStream_t s) {
  const int rows = gram_rows_choice();
  if (rows == 64) return gram_launch_t<MA, NB, 64>(descs, kl, kr, share, d, partial, s);
  if constexpr (MA == 1 && NB == 1) {
    if (rows != 16) return gram_launch_t<MA, NB, 32>(descs, kl, kr, share, d, partial, s);
  }
  return gram_launch_t<MA, NB, 16>(descs, kl, kr, share, d, partial, s);
}

template <int MA>
int gram_launch_nb(int nb, const ColDesc *descs, int kl, int kr, int share, int64_t d, double2 *partial,
                   hipStream_t s) {
  switch (nb) {
    case 1: return gram_launch<MA, 1>(descs, kl, kr, share, d, partial, s);
    case 2: return gram_launch<MA, 2>(descs, kl, kr, share, d, partial, s);
    case 3: return gram_launch<MA, 3>(descs, kl, kr, share, d, partial, s);
    case 4: return gram_launch<MA, 4>(descs, kl, kr, share, d, partial, s);
    case 5: return gram_launch<MA, 5>(descs, kl, kr, share, d, partial, s);
    case 6: return gram_launch<MA, 6>(descs, kl, kr, share, d, partial, s);
    case 7: return gram_launch<MA, 7>(descs, kl, kr, share, d, partial, s);
    default: return gram_launch<MA, 8>(descs, kl, kr, share, d, partial, s);
  }
}
}  // namespace

hipError_t launch_gram(const ColDesc *descs, int kl, int kr, int share, int64_t d, Dbl2 *partial,
                       Dbl2 *out, hipStream_t s) {
  if (kl < 1 || kr < 1 || kl > kGramMaxL || kr > kGramMaxR) return hipErrorInvalidValue;
  if (share < 0 || share > kl || share > kr) return hipErrorInvalidValue;
  const int ma = (kl + 15) / 16, nb = (kr + 15) / 16;
  double2 *pp = reinterpret_cast<double2 *>(partial);
  int nparts = 0;
  if (kr <= 8 && kl <= 32 && (share == 0 || share == kr) && gram_pk_on()) {
    switch ((kl + 7) / 8) {
      case 1: nparts = gram_pk_launch<1>(descs, kl, kr, share, d, pp, s); break;
      case 2: nparts = gram_pk_launch<2>(descs, kl, kr, share, d, pp, s); break;
      case 3: nparts = gram_pk_launch<3>(descs, kl, kr, share, d, pp, s); break;
      default: nparts = gram_pk_launch<4>(descs, kl, kr, share, d, pp, s); break;
    }
    const int nelem = kl * kr;
    hipLaunchKernelGGL(k_reduce_partials, dim3((nelem + 3) / 4), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(partial), nparts, nelem,
                       reinterpret_cast<double2 *>(out));
    return hipGetLastError();
  }
  switch (ma) {
    case 1: nparts = gram_launch_nb<1>(nb, descs, kl, kr, share, d, pp, s); break;
    case 2: nparts = gram_launch_nb<2>(nb, descs, kl, kr, share, d, pp, s); break;
    case 3: nparts = gram_launch_nb<3>(nb, descs, kl, kr, share, d, pp, s); break;
    default: nparts = gram_launch_nb<4>(nb, descs, kl, kr, share, d, pp, s); break;
  }
  const int nelem = kl * kr;
  hipLaunchKernelGGL(k_reduce_partials, dim3((nelem + 3) / 4), dim3(256), 0, s,
                     reinterpret_cast<const double2 *>(partial), nparts, nelem,
                     reinterpret_cast<double2 *>(out));
  return hipGetLastError();
}

namespace {
template <int JPT>
bool combine_setup() {
  static std::atomic<bool> attr{false};
  if (attr.load(std::memory_order_acquire)) return true;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (attr.load(std::memory_order_acquire)) return true;
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_combine<JPT>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
    return false;
  attr.store(true, std::memory_order_release);
  return true;
}

template <int JPT>
hipError_t combine_launch(const ColDesc *descs, int nin, int nout, const Dbl2 *c, const Dbl2 *beta,
                          int64_t d, hipStream_t s) {
  const size_t lds = (size_t(nin) * kCombGroups * JPT + size_t(kCombRows) * nin) * sizeof(double2);
  if (!combine_setup<JPT>()) return hipErrorInvalidDeviceFunction;
  const int64_t ntile = (d + kCombRows - 1) / kCombRows;
  const unsigned grid = unsigned(ntile < 2048 ? ntile : 2048);
  hipLaunchKernelGGL((k_combine<JPT>), dim3(grid), dim3(kCombThreads), lds, s, descs, nin, nout,
                     reinterpret_cast<const double2 *>(c), reinterpret_cast<const double2 *>(beta), d);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_combine(const ColDesc *descs, int nin, int nout, const Dbl2 *c, const Dbl2 *beta,
                          int64_t d, hipStream_t s) {
  if (nin < 1 || nin > kCombMaxIn || nout < 1 || nout > kCombMaxOut) return hipErrorInvalidValue;
  if (nout <= 16 || nout > kCombGroups * 4) {
    const double2 *cc = reinterpret_cast<const double2 *>(c);
    const double2 *bb = reinterpret_cast<const double2 *>(beta);
    // one wavefront per 64 * RPL rows, at most 2048 workgroups of 4 (grid-stride beyond)
    auto grid = [d](int rpl) {
      const int64_t waves = (d + 64 * rpl - 1) / (64 * rpl);
      return dim3(unsigned((waves + 3) / 4 < 2048 ? (waves + 3) / 4 : 2048));
    };
    if (nout <= 4)
      hipLaunchKernelGGL((k_combine_w<4, 2, 8>), grid(2), dim3(256), 0, s, descs, nin, nout, cc, bb, d);
    else if (nout <= 8)
      hipLaunchKernelGGL((k_combine_w<8, 2, 8>), grid(2), dim3(256), 0, s, descs, nin, nout, cc, bb, d);
    else if (nout <= 12)
      hipLaunchKernelGGL((k_combine_w<12, 2, 8>), grid(2), dim3(256), 0, s, descs, nin, nout, cc, bb, d);
    else if (nout <= 16)
      hipLaunchKernelGGL((k_combine_w<16, 2, 16>), grid(2), dim3(256), 0, s, descs, nin, nout, cc, bb, d);
    else
      hipLaunchKernelGGL((k_combine_w<kCombMaxOut, 1, 4>), grid(1), dim3(256), 0, s, descs, nin, nout, cc, bb,
                         d);
    return hipGetLastError();
  }
  // 17..32 outputs: C and one row tile in LDS, (64 * 8 * 4 + 32 * 64) * 16 B = 64 KB at most
  switch ((nout + kCombGroups - 1) / kCombGroups) {
    case 3: return combine_launch<3>(descs, nin, nout, c, beta, d, s);
    default: return combine_launch<4>(descs, nin, nout, c, beta, d, s);
  }
}

size_t coldots_partial_elems(int k) { return size_t(kDotBlocks) * k * 4; }

hipError_t launch_resid(const ColDesc *descs, int nev, const double *sc, const Dbl2 *ev, int64_t d,
                        double *partial, double *out, hipStream_t s) {
  if (nev < 1 || nev > kDotMaxCols) return hipErrorInvalidValue;
  // kDotBlocks * 4 partial sums per column at most: the coldots partial buffer holds them
  const int nb = kDotBlocks;
  int64_t rpb = (d + nb - 1) / nb;
  rpb = (rpb + 1) / 2 * 2;  // row pairs start at even rows
  hipLaunchKernelGGL(k_resid, dim3(unsigned(nb), unsigned(nev)), dim3(256), 0, s, descs, nev, sc,
                     reinterpret_cast<const double2 *>(ev), d, rpb, partial);
  hipLaunchKernelGGL(k_reduce_real, dim3((nev + 3) / 4), dim3(256), 0, s, partial, nb, nev, out);
  return hipGetLastError();
}

hipError_t launch_coldots(const ColDesc *descs, int k, int64_t d, double *partial, double *out,
                          hipStream_t s) {
  if (k < 1 || k > kDotMaxCols) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_coldots, dim3(kDotBlocks), dim3(kDotThreads), 0, s, descs, k, d, partial);
  const int nelem = k * 4;
  hipLaunchKernelGGL(k_reduce_real, dim3((nelem + 3) / 4), dim3(256), 0, s, partial, kDotBlocks,
                     nelem, out);
  return hipGetLastError();
}

hipError_t launch_fill_gaussian(Dbl2 *x, int64_t n, uint64_t seed, uint32_t col, uint32_t stream_id,
                                hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(k_fill_gaussian, dim3(unsigned((n + bs - 1) / bs)), dim3(bs), 0, s,
                     reinterpret_cast<double2 *>(x), n, seed, col, stream_id);
  return hipGetLastError();
}

hipError_t launch_sym_project(Dbl2 *x, int64_t m, int sign, hipStream_t s) {
  const unsigned nb = unsigned((m + 31) / 32);
  hipLaunchKernelGGL(k_singlet, dim3(nb, nb), dim3(32, 8), 0, s, reinterpret_cast<double2 *>(x), m,
                     double(sign));
  return hipGetLastError();
}

hipError_t launch_to_single(const Dbl2 *x, void *xs, int64_t n, hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(k_to_single, dim3(unsigned((n + bs - 1) / bs)), dim3(bs), 0, s,
                     reinterpret_cast<const double2 *>(x), static_cast<float2 *>(xs), n);
  return hipGetLastError();
}

hipError_t launch_h1(const Dbl2 *x, Dbl2 *y, int nn, int ncols, const double *vn, const MatvecCoef &coef,
                     hipStream_t s) {
  const int64_t m = int64_t(nn) * nn * nn;
  const int64_t n = m * ncols;
  if (ncols < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_h1, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, reinterpret_cast<const double2 *>(x),
                     reinterpret_cast<double2 *>(y), nn, m, ncols, vn, coef);
  return hipGetLastError();
}

hipError_t launch_small_dots(const Dbl2 *a, int p, const Dbl2 *b, int q, int64_t m, int herm, Dbl2 *out,
                             hipStream_t s) {
  if (p < 1 || q < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_dots, dim3(unsigned(int64_t(p) * q)), dim3(256), 0, s,
                     reinterpret_cast<const double2 *>(a), p, reinterpret_cast<const double2 *>(b), q, m, herm,
                     reinterpret_cast<double2 *>(out));
  return hipGetLastError();
}

hipError_t launch_small_lincomb(const Dbl2 *a, int p, const Dbl2 *c, int q, int64_t m, int beta, Dbl2 *y,
                                hipStream_t s) {
  if (p < 1 || q < 1) return hipErrorInvalidValue;
  const int64_t n = m * q;
  hipLaunchKernelGGL(k_small_lincomb, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const double2 *>(a), p, reinterpret_cast<const double2 *>(c), q, m, beta,
                     reinterpret_cast<double2 *>(y));
  return hipGetLastError();
}

hipError_t launch_to_double(const void *xs, Dbl2 *x, int64_t n, hipStream_t s) {
  const int bs = 256;
  hipLaunchKernelGGL(k_to_double, dim3(unsigned((n + bs - 1) / bs)), dim3(bs), 0, s,
                     static_cast<const float2 *>(xs), reinterpret_cast<double2 *>(x), n);
  return hipGetLastError();
}

namespace {
template <int NN>
bool prepare_nn() {
  bool ok = true;
  if constexpr (fast_fits(NN, sizeof(double))) {
    ok = ok && pass1_setup<NN, double, double2>() != 0;
    ok = ok && pass2_setup<NN, double, double2, double2>();
  }
  if constexpr (fast_fits(NN, sizeof(float))) {
    ok = ok && pass1_setup<NN, float, float2>() != 0 && pass1_setup<NN, float, double2>() != 0;
    ok = ok && pass2_setup<NN, float, float2, float2>() && pass2_setup<NN, float, double2, double2>();
  }
  return ok;
}
}  // namespace

hipError_t prepare_kernels() {
  bool ok = true;
#define CASE(N) ok = prepare_nn<N>() && ok;
  DEX_NN_LIST(CASE)
#undef CASE
  ok = combine_setup<3>() && ok;
  ok = combine_setup<4>() && ok;
  gram_grids_nb<1>();
  gram_grids_nb<2>();
  gram_grids_nb<3>();
  gram_grids_nb<4>();
  gram_pk_grids<1>();
  gram_pk_grids<2>();
  gram_pk_grids<3>();
  gram_pk_grids<4>();
  return ok ? hipSuccess : hipErrorInvalidDeviceFunction;
}

}  // namespace dex
