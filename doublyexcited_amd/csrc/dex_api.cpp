// C-ABI entry points: context set-up, the matvec (MYMATMUL / HAMMULT),
// Hermitian dots (MYDOTSUB) and the GPLMR / harmonic-Ritz block eigensolver
// (GPLMR_EIG, gplmr.F90:9-386) driven from the host over device kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "dex_internal.h"

using dex::ColDesc;
using dex::Dbl2;
using cd = std::complex<double>;

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

int hip_fail(hipError_t e, const char *where) {
  return fail(DEX_EHIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define DEX_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return hip_fail(e_, #call);   \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int kMaxDesc = 512;
constexpr int kMaxSmall = 96 * 64;

}  // namespace

namespace dex {
int set_error(int status, const char *msg) { return fail(status, msg); }
}  // namespace dex

// ---- device scratch for the allocation-free matvec ---------------------------
// dex_matvec_noalloc keeps the matvec intermediate U in this file-scope device
// array instead of a library allocation.  Every GPU the library runs on holds
// its own copy (64 MiB) until the process ends.  Host code never passes the
// array's name to a kernel: it is a host-side stand-in; the device address of
// the current device's copy comes from hipGetSymbolAddress, once per device.
namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
hipError_t scratch_address(int device, void **ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

struct dex_ctx {
  dex_params p{};
  int nn = 0;
  int64_t m = 0, d = 0;
  std::vector<double> t_host, tinv_host, vn_host;
  double *d_tinv = nullptr, *d_vn = nullptr;
  dex::MatvecCoef coef{};
  hipStream_t stream = nullptr;
  bool fast64 = false, fast32 = false;
  void *d_u = nullptr;  // matvec intermediate, D complex128 (also used as D complex64)
  void *d_ub = nullptr;  // intermediates of a batched matvec launch: batch_cols(c) x D complex128
  void *d_scratch = nullptr;  // this device's copy of `scratch` (dex_matvec_noalloc)
  ColDesc *d_desc = nullptr;
  unsigned char *d_args = nullptr;  // a combine's descriptors, C and beta, packed (one upload)
  Dbl2 *d_gram_out = nullptr;
  void *d_partial = nullptr;
  size_t partial_bytes = 0;
  double *d_dots = nullptr;
  int timing = 0;                 // 0 off, 1 every launch, n > 1 a random 1 in n of them
  std::atomic<int> stop{0};       // dex_ctx_stop: solves return at their next iteration boundary
  uint64_t sample_state = 0x9e3779b97f4a7c15ull;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  double t_pass1 = -1.0, t_pass2 = -1.0;
  // per-kernel-class launch timing: event pairs recorded on the stream and
  // resolved lazily (no synchronisation added to the timed work)
  struct Pending {
    hipEvent_t a, b;
    int cls;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> free_events;
  double k_ms[DEX_KCLASSES] = {};
  int64_t k_n[DEX_KCLASSES] = {};
  double k_bytes[DEX_KCLASSES] = {};  // algorithmic HBM bytes of the timed launches
  double k_flops[DEX_KCLASSES] = {};  // algorithmic flops of the timed launches
  // every launch, timed or not (counted on the host from the launch shapes)
  int64_t k_all_n[DEX_KCLASSES] = {};
  double k_all_bytes[DEX_KCLASSES] = {};
  double k_all_flops[DEX_KCLASSES] = {};
  // pinned staging ring for the small host-to-device uploads (descriptors,
  // coefficients): copies from pageable memory would wait for the stream
  char *h_stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
};

namespace {

hipEvent_t take_event(dex_ctx *c) {
  if (!c->free_events.empty()) {
    hipEvent_t e = c->free_events.back();
    c->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Accumulates every recorded pair (waits for the last one).
hipError_t resolve_timing(dex_ctx *c) {
  if (c->pending.empty()) return hipSuccess;
  hipError_t e = hipEventSynchronize(c->pending.back().b);
  for (const auto &pd : c->pending) {
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, pd.a, pd.b);
    c->k_ms[pd.cls] += ms;
    c->k_n[pd.cls] += 1;
    c->free_events.push_back(pd.a);
    c->free_events.push_back(pd.b);
  }
  c->pending.clear();
  return e;
}

// Brackets the launches of one kernel class with events when timing is on.
// `bytes` and `flops` are the launch's algorithmic HBM traffic and arithmetic.
// Whether this launch is timed: all with timing 1; with timing n > 1 each
// launch independently with probability 1/n (xorshift64), so a periodic
// launch pattern cannot alias with the sample.
bool sampled(dex_ctx *c) {
  if (c->timing <= 1) return c->timing == 1;
  uint64_t x = c->sample_state;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  c->sample_state = x;
  return (x >> 11) % uint64_t(c->timing) == 0;
}

struct KTimer {
  dex_ctx *c;
  int cls;
  hipEvent_t a = nullptr;
  KTimer(dex_ctx *c_, int cls_, double bytes, double flops) : c(c_), cls(cls_) {
    c->k_all_n[cls] += 1;
    c->k_all_bytes[cls] += bytes;
    c->k_all_flops[cls] += flops;
    if (c->timing && sampled(c) && (a = take_event(c)) != nullptr) {
      (void)hipEventRecord(a, c->stream);
      c->k_bytes[cls] += bytes;
      c->k_flops[cls] += flops;
    }
  }
  ~KTimer() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) {
      c->free_events.push_back(a);
      return;
    }
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({a, b, cls});
    if (c->pending.size() >= 8192) (void)resolve_timing(c);
  }
};

// Uploads `bytes` from host memory to `dev` on the context's stream through the
// pinned ring (asynchronous); the ring wraps after a stream synchronisation,
// so no staged copy is overwritten before it has run.
hipError_t stage_upload(dex_ctx *c, void *dev, const void *src, size_t bytes) {
  if (!c->h_stage || bytes > c->stage_cap)
    return hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, c->stream);
  const size_t need = (bytes + 255) / 256 * 256;
  if (c->stage_off + need > c->stage_cap) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    c->stage_off = 0;
  }
  char *h = c->h_stage + c->stage_off;
  c->stage_off += need;
  std::memcpy(h, src, bytes);
  return hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, c->stream);
}

int ensure_u(dex_ctx *c) {
  if (c->d_u) return 0;
  DEX_HIP(hipMalloc(&c->d_u, size_t(c->d) * sizeof(Dbl2)));
  return 0;
}

// Columns per batched matvec launch: as many intermediates as fit in 512 MiB,
// at most kBatchMax (16 at NN = 11, 6 at NN = 13, 2 at NN = 15, none from
// NN = 17 on, where one column's passes already fill the device).
int64_t batch_cols(const dex_ctx *c) {
  const int64_t b = (int64_t(512) << 20) / (c->d * 16);
  return b >= 2 ? std::min<int64_t>(b, dex::kBatchMax) : 1;
}

int ensure_ub(dex_ctx *c) {
  if (c->d_ub || batch_cols(c) < 2) return 0;
  DEX_HIP(hipMalloc(&c->d_ub, size_t(batch_cols(c)) * size_t(c->d) * sizeof(Dbl2)));
  return 0;
}

// H on ncols columns with X = sym X^T (sym = +-1) in two launches (batched
// electron-1 pass, batched transpose-add pass).  io: dex::kIoF64 (complex128,
// FP64), kIoMixed (complex128, FP32 arithmetic) or kIoF32 (complex64 columns).
int apply_h_batch(dex_ctx *c, const Dbl2 *const *x, Dbl2 *const *y, int ncols, int io, int sym) {
  const bool single = io != dex::kIoF64;
  if (ncols < 1 || ncols > batch_cols(c) || sym == 0 || !(single ? c->fast32 : c->fast64))
    return fail(DEX_EINVAL, "batched matvec: bad column count or form");
  if (int st = ensure_ub(c)) return st;
  const double dd = double(c->d) * ncols, nn = double(c->nn);
  const double ub = single ? 8.0 : 16.0;     // intermediate U entry bytes
  const double xb = io == dex::kIoF32 ? 8.0 : 16.0;  // X and Y entry bytes
  const void *xs[dex::kBatchMax];
  void *us[dex::kBatchMax];
  void *ys[dex::kBatchMax];
  for (int j = 0; j < ncols; ++j) {
    xs[j] = x[j];
    us[j] = static_cast<char *>(c->d_ub) + size_t(j) * size_t(c->d) * sizeof(Dbl2);
    ys[j] = y[j];
  }
  {
    KTimer kt(c, DEX_K_PASS1, (xb + ub) * dd, 12.0 * nn * dd);
    DEX_HIP(dex::launch_pass1_batch(c->nn, io, xs, us, ncols, c->stream));
  }
  {
    KTimer kt(c, DEX_K_SYMADD, (2.0 * xb + ub) * dd, 24.0 * dd);
    DEX_HIP(dex::launch_symadd_batch(c->nn, io, xs, us, ys, ncols, c->d_vn, c->d_tinv, c->coef, sym,
                                     c->stream));
  }
  return 0;
}

// Second pass: the general electron-2 cube pass, or for X = sym X^T (sym = +-1)
// the transpose-add pass.
hipError_t second_pass(dex_ctx *c, int io, const void *x, void *y, int sym) {
  if (sym != 0)
    return dex::launch_symadd(c->nn, io, x, c->d_u, y, c->d_vn, c->d_tinv, c->coef, sym, c->stream);
  return dex::launch_pass2(c->nn, io, x, c->d_u, y, c->d_vn, c->d_tinv, c->coef, c->stream);
}

// One application of H on complex128 device vectors; `single` selects FP32
// arithmetic (complex64 intermediate, complex128 in and out).
// pass_events: also record c->ev[0..2] around the two passes (the matvec
// entry point's per-pass timing).
int apply_h(dex_ctx *c, const Dbl2 *x, Dbl2 *y, bool single, int sym = 0, bool pass_events = false) {
  if (int st = ensure_u(c)) return st;
  hipStream_t s = c->stream;
  const bool fast = single ? c->fast32 : c->fast64;
  if (single && !fast) return fail(DEX_EUNSUPPORTED, "no FP32 matvec for this grid");
  if (fast) {
    const int io = single ? dex::kIoMixed : dex::kIoF64;
    if (c->timing && pass_events) DEX_HIP(hipEventRecord(c->ev[0], s));
    const double dd = double(c->d), nn = double(c->nn);
    const double ub = single ? 8.0 : 16.0;  // intermediate U entry bytes
    {
      // reads X, writes U; three axes of NN real-by-complex MACs per entry
      KTimer kt(c, DEX_K_PASS1, (16.0 + ub) * dd, 12.0 * nn * dd);
      DEX_HIP(dex::launch_pass1(c->nn, io, x, c->d_u, s));
    }
    if (c->timing && pass_events) DEX_HIP(hipEventRecord(c->ev[1], s));
    {
      // reads U and X, writes Y; the general pass also applies the three
      // electron-2 axes, the transpose-add pass the exchange image
      const double fl = sym != 0 ? 24.0 * dd : (12.0 * nn + 14.0) * dd;
      KTimer kt(c, sym != 0 ? DEX_K_SYMADD : DEX_K_PASS2, (32.0 + ub) * dd, fl);
      DEX_HIP(second_pass(c, io, x, y, sym));
    }
    if (c->timing && pass_events) DEX_HIP(hipEventRecord(c->ev[2], s));
    return 0;
  }
  KTimer kt(c, DEX_K_GENERIC, 32.0 * double(c->d), (24.0 * c->nn + 14.0) * double(c->d));
  DEX_HIP(dex::launch_matvec_generic(c->nn, x, y, c->d_vn, c->d_tinv, c->coef, s));
  // The general operator carries the rounding-level exchange-odd (even) part of
  // an exchange-even (odd) x into y, where the transpose-add kernels, which the
  // mixed form applies to its complex64 blocks, drop it: in a solve that mixes
  // the two, AX then disagrees with the images of the search blocks in that
  // part, and the iteration drifts away from the wanted levels after tens of
  // iterations (seen at NN=25, where FP64 has only these kernels).  Projecting y
  // keeps both paths on the exchange-symmetric subspace.
  if (sym != 0) DEX_HIP(dex::launch_sym_project(y, c->m, sym > 0 ? 1 : -1, s));
  return 0;
}

// H on complex64 device vectors in FP32 arithmetic (x, y hold D complex64).
int apply_h32(dex_ctx *c, const Dbl2 *x, Dbl2 *y, int sym) {
  if (int st = ensure_u(c)) return st;
  if (!c->fast32) return fail(DEX_EUNSUPPORTED, "no FP32 matvec for this grid");
  const double dd = double(c->d), nn = double(c->nn);
  {
    KTimer kt(c, DEX_K_PASS1, 16.0 * dd, 12.0 * nn * dd);
    DEX_HIP(dex::launch_pass1(c->nn, dex::kIoF32, x, c->d_u, c->stream));
  }
  {
    const double fl = sym != 0 ? 24.0 * dd : (12.0 * nn + 14.0) * dd;
    KTimer kt(c, sym != 0 ? DEX_K_SYMADD : DEX_K_PASS2, 24.0 * dd, fl);
    DEX_HIP(second_pass(c, dex::kIoF32, x, y, sym));
  }
  return 0;
}

// ---- block helpers on one context ------------------------------------------
// The solver's work vectors are complex128 (D x 16 B) or complex64 (D x 8 B)
// buffers.  A complex64 buffer's pointer carries a tag in bit 0 (device
// allocations are 256-byte aligned, so the bit is otherwise zero): blocks can
// then mix storage formats column by column, and every descriptor built from a
// pointer (plain, shifted) takes the untagged address and the format from it.
inline bool is32(const void *p) { return (reinterpret_cast<uintptr_t>(p) & uintptr_t(1)) != 0; }
inline Dbl2 *untag(const Dbl2 *p) {
  return reinterpret_cast<Dbl2 *>(reinterpret_cast<uintptr_t>(p) & ~uintptr_t(1));
}
inline Dbl2 *tag32(Dbl2 *p) { return reinterpret_cast<Dbl2 *>(reinterpret_cast<uintptr_t>(p) | uintptr_t(1)); }

// Columns of a block (device vectors of D entries).  `fmt` is the storage of
// the columns grow() adds: 0 complex128, 1 complex64 (tagged pointers).
struct Block {
  std::vector<Dbl2 *> cols;
  int fmt = 0;
  int size() const { return int(cols.size()); }
};

ColDesc plain(const Dbl2 *a) { return ColDesc{untag(a), nullptr, 0.0, 0.0, -1, is32(a) ? 1 : 0}; }
// a - sig b; a and b share one storage format
ColDesc shifted(const Dbl2 *a, const Dbl2 *b, cd sig) {
  return ColDesc{untag(a), untag(b), sig.real(), sig.imag(), -1, is32(a) ? 1 : 0};
}

int upload_descs(dex_ctx *c, const std::vector<ColDesc> &ds) {
  if (int(ds.size()) > kMaxDesc) return fail(DEX_EINVAL, "too many block columns");
  DEX_HIP(stage_upload(c, c->d_desc, ds.data(), ds.size() * sizeof(ColDesc)));
  return 0;
}

// Bytes per entry summed over the distinct vectors the descriptors read (a, and
// b where shifted): 16 for a complex128 column, 8 for a complex64 one.
double distinct_entry_bytes(const std::vector<ColDesc> &ds, size_t n) {
  std::vector<std::pair<const void *, int>> v;
  for (size_t i = 0; i < n && i < ds.size(); ++i) {
    if (ds[i].a) v.push_back({ds[i].a, ds[i].fmt});
    if (ds[i].b) v.push_back({ds[i].b, ds[i].fmt});
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  double b = 0.0;
  for (const auto &pv : v) b += pv.second ? 8.0 : 16.0;
  return b;
}

// Runs one Gram launch (kl <= 64, kr <= 128) and copies the kl x kr result
// into g (leading dimension ldg) at (i0, j0).  With share > 0 the skipped lower
// blocks of the Hermitian share x share part are filled by conjugate symmetry.
int gram_launch_copy(dex_ctx *c, const std::vector<ColDesc> &ds, int kl, int kr, int share,
                     std::vector<cd> &g, int ldg, int i0, int j0) {
  if (int st = upload_descs(c, ds)) return st;
  {
    const double hskip = share > 0 ? 0.5 * double(share) * double(share - 1) : 0.0;
    KTimer kt(c, DEX_K_GRAM, double(c->d) * distinct_entry_bytes(ds, ds.size()),
              8.0 * double(c->d) * (double(kl) * kr - hskip));
    DEX_HIP(dex::launch_gram(c->d_desc, kl, kr, share, c->d, reinterpret_cast<Dbl2 *>(c->d_partial),
                             c->d_gram_out, c->stream));
  }
  std::vector<cd> part(size_t(kl) * kr);
  DEX_HIP(hipMemcpyAsync(part.data(), c->d_gram_out, part.size() * sizeof(Dbl2), hipMemcpyDeviceToHost,
                         c->stream));
  DEX_HIP(hipStreamSynchronize(c->stream));
  for (int j = 0; j < kr; ++j)
    for (int i = 0; i < kl; ++i) {
      cd v = part[i + size_t(kl) * j];
      if (share > 0 && (i / 16) > (j / 16) && 16 * (i / 16 + 1) <= share) v = std::conj(part[j + size_t(kl) * i]);
      g[(i0 + i) + size_t(ldg) * (j0 + j)] = v;
    }
  return 0;
}

// G = L^H R (kl x kr, column-major) on the host; any sizes (chunked launches).
int gram(dex_ctx *c, const std::vector<ColDesc> &l, const std::vector<ColDesc> &r,
         std::vector<cd> &g) {
  const int kl = int(l.size()), kr = int(r.size());
  g.assign(size_t(kl) * kr, 0.0);
  for (int i0 = 0; i0 < kl; i0 += 64)
    for (int j0 = 0; j0 < kr; j0 += 128) {
      const int bl = std::min(64, kl - i0), br = std::min(128, kr - j0);
      std::vector<ColDesc> ds(l.begin() + i0, l.begin() + i0 + bl);
      ds.insert(ds.end(), r.begin() + j0, r.begin() + j0 + br);
      if (int st = gram_launch_copy(c, ds, bl, br, 0, g, kl, i0, j0)) return st;
    }
  return 0;
}

// G = X^H X for k <= 64 columns: every column is read once.
int gram_self(dex_ctx *c, const std::vector<ColDesc> &x, std::vector<cd> &g) {
  const int k = int(x.size());
  if (k > 64) return gram(c, x, x, g);
  g.assign(size_t(k) * k, 0.0);
  return gram_launch_copy(c, x, k, k, k, g, k, 0, 0);
}

// [K | M] = B^H [B | Q] with B = AQ - sigma Q: for nq <= 64 one launch that reads
// AQ and Q once; wider bases (up to 96) go through the chunked general Gram.
int gram(dex_ctx *c, const std::vector<ColDesc> &l, const std::vector<ColDesc> &r, std::vector<cd> &g);

int gram_harmonic(dex_ctx *c, const std::vector<Dbl2 *> &q, const std::vector<Dbl2 *> &aq, cd sigma,
                  std::vector<cd> &g) {
  const int nq = int(q.size());
  if (nq < 1 || nq > 96) return fail(DEX_EINVAL, "projected basis must have 1..96 columns");
  if (nq > 64) {
    // beyond one launch: the general chunked Gram of B = AQ - sigma Q against [B | Q]
    std::vector<ColDesc> l, r;
    for (int j = 0; j < nq; ++j) l.push_back(shifted(aq[j], q[j], sigma));
    r = l;
    for (int j = 0; j < nq; ++j) r.push_back(plain(q[j]));
    return gram(c, l, r, g);
  }
  const int ma = (nq + 15) / 16;
  std::vector<ColDesc> ds;
  for (int j = 0; j < nq; ++j) {
    ColDesc dsc = shifted(aq[j], q[j], sigma);
    dsc.bslot = 16 * ma + j;  // raw Q_j lands in R column nq + j
    ds.push_back(dsc);
  }
  for (int j = 0; j < nq; ++j) ds.push_back(ColDesc{nullptr, nullptr, 0, 0, -1, 0});
  g.assign(size_t(nq) * 2 * nq, 0.0);
  return gram_launch_copy(c, ds, nq, 2 * nq, nq, g, nq, 0, 0);
}

// The iteration's X/P update as two launches of up to 16 outputs (on unless
// DEX_UPDATE_SPLIT=0, read once): 6-7% less combine time at NN=21, block 16.
bool split_update() {
  static const bool on = [] {
    const char *v = std::getenv("DEX_UPDATE_SPLIT");
    return !(v && std::atoi(v) == 0);
  }();
  return on;
}

// out_j = sum_i in_i C(i, j) (+ beta_j out_j), row-local, in-place safe; each
// output column is stored in its own pointer's format.
int combine(dex_ctx *c, const std::vector<ColDesc> &in, const std::vector<Dbl2 *> &out,
            const std::vector<cd> &cm, const std::vector<cd> *beta = nullptr) {
  const int nin = int(in.size()), nout = int(out.size());
  if (nout == 0) return 0;
  if (nin * nout > kMaxSmall || cm.size() < size_t(nin) * nout || cm.size() > size_t(kMaxSmall) ||
      (beta && (beta->size() < size_t(nout) || beta->size() > size_t(kMaxSmall))))
    return fail(DEX_EINVAL, "combine matrix size");
  std::vector<ColDesc> ds(in);
  double out_bytes = 0.0;
  for (Dbl2 *o : out) {
    ds.push_back(plain(o));
    out_bytes += is32(o) ? 8.0 : 16.0;
  }
  if (int(ds.size()) > kMaxDesc) return fail(DEX_EINVAL, "too many block columns");
  // descriptors | C | beta in one image, one copy
  const size_t bd = (ds.size() * sizeof(ColDesc) + 255) / 256 * 256;
  const size_t bc = (cm.size() * sizeof(Dbl2) + 255) / 256 * 256;
  const size_t bb = beta ? beta->size() * sizeof(Dbl2) : 0;
  std::vector<unsigned char> img(bd + bc + bb);
  std::memcpy(img.data(), ds.data(), ds.size() * sizeof(ColDesc));
  std::memcpy(img.data() + bd, cm.data(), cm.size() * sizeof(Dbl2));
  if (beta) std::memcpy(img.data() + bd + bc, beta->data(), bb);
  DEX_HIP(stage_upload(c, c->d_args, img.data(), img.size()));
  KTimer kt(c, DEX_K_COMBINE,
            double(c->d) * (distinct_entry_bytes(in, in.size()) + out_bytes * (beta ? 2.0 : 1.0)),
            8.0 * double(c->d) * nin * nout);
  const Dbl2 *dc = reinterpret_cast<const Dbl2 *>(c->d_args + bd);
  const Dbl2 *dbeta = beta ? reinterpret_cast<const Dbl2 *>(c->d_args + bd + bc) : nullptr;
  DEX_HIP(dex::launch_combine(reinterpret_cast<const ColDesc *>(c->d_args), nin, nout, dc, dbeta, c->d,
                              c->stream));
  return 0;
}

// Per column j: aa_j = |a_j|^2, ab_j = a_j^H b_j, bb_j = |b_j|^2.
int coldots(dex_ctx *c, const std::vector<ColDesc> &a, const std::vector<ColDesc> &b,
            std::vector<double> &aa, std::vector<cd> &ab, std::vector<double> &bb) {
  const int k = int(a.size());
  std::vector<ColDesc> ds(a);
  if (b.empty()) {
    for (int j = 0; j < k; ++j) ds.push_back(ColDesc{nullptr, nullptr, 0, 0, -1, 0});
  } else {
    ds.insert(ds.end(), b.begin(), b.end());
  }
  if (int st = upload_descs(c, ds)) return st;
  {
    KTimer kt(c, DEX_K_COLDOTS, double(c->d) * distinct_entry_bytes(ds, ds.size()),
              (b.empty() ? 4.0 : 16.0) * double(c->d) * k);
    DEX_HIP(dex::launch_coldots(c->d_desc, k, c->d, reinterpret_cast<double *>(c->d_partial), c->d_dots,
                                c->stream));
  }
  std::vector<double> h(size_t(k) * 4);
  DEX_HIP(hipMemcpyAsync(h.data(), c->d_dots, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DEX_HIP(hipStreamSynchronize(c->stream));
  aa.resize(k);
  ab.resize(k);
  bb.resize(k);
  for (int j = 0; j < k; ++j) {
    aa[j] = h[4 * j];
    ab[j] = cd(h[4 * j + 1], h[4 * j + 2]);
    bb[j] = h[4 * j + 3];
  }
  return 0;
}

// The residual step per column j: X_j <- s_j X_j, AX_j <- s_j AX_j,
// W_j = AX_j - E_j X_j (scaled), and rn_j = ||W_j||, in one pass over X and AX.
int residual_cols(dex_ctx *c, const std::vector<Dbl2 *> &x, const std::vector<Dbl2 *> &ax,
                  const std::vector<Dbl2 *> &w, const std::vector<double> &sc, const std::vector<cd> &ev,
                  std::vector<double> &rn) {
  const int nev = int(x.size());
  if (nev < 1 || int(ax.size()) != nev || int(w.size()) < nev || nev > 64)
    return fail(DEX_EINVAL, "residual step: bad block widths");
  std::vector<ColDesc> ds;
  for (Dbl2 *p : x) ds.push_back(plain(p));
  for (Dbl2 *p : ax) ds.push_back(plain(p));
  for (int j = 0; j < nev; ++j) ds.push_back(plain(w[j]));
  // descriptors | scalings | E in one image, one copy
  const size_t bd = (ds.size() * sizeof(ColDesc) + 255) / 256 * 256;
  const size_t bs = (sc.size() * sizeof(double) + 255) / 256 * 256;
  std::vector<unsigned char> img(bd + bs + ev.size() * sizeof(cd));
  std::memcpy(img.data(), ds.data(), ds.size() * sizeof(ColDesc));
  std::memcpy(img.data() + bd, sc.data(), sc.size() * sizeof(double));
  std::memcpy(img.data() + bd + bs, ev.data(), ev.size() * sizeof(cd));
  DEX_HIP(stage_upload(c, c->d_args, img.data(), img.size()));
  {
    double bytes = 0.0;  // read X and AX, write X, AX and W
    for (int j = 0; j < nev; ++j)
      bytes += 2.0 * (is32(x[j]) ? 8.0 : 16.0) + 2.0 * (is32(ax[j]) ? 8.0 : 16.0) + (is32(w[j]) ? 8.0 : 16.0);
    KTimer kt(c, DEX_K_COLDOTS, double(c->d) * bytes, 12.0 * double(c->d) * nev);
    DEX_HIP(dex::launch_resid(reinterpret_cast<const ColDesc *>(c->d_args), nev,
                              reinterpret_cast<const double *>(c->d_args + bd),
                              reinterpret_cast<const Dbl2 *>(c->d_args + bd + bs), c->d,
                              reinterpret_cast<double *>(c->d_partial), c->d_dots, c->stream));
  }
  std::vector<double> h(nev);
  DEX_HIP(hipMemcpyAsync(h.data(), c->d_dots, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DEX_HIP(hipStreamSynchronize(c->stream));
  rn.resize(nev);
  for (int j = 0; j < nev; ++j) rn[j] = std::sqrt(h[j]);
  return 0;
}

std::vector<ColDesc> descs_of(const Block &b) {
  std::vector<ColDesc> v;
  for (Dbl2 *p : b.cols) v.push_back(plain(p));
  return v;
}

}  // namespace

// ============================================================================
extern "C" {

void dex_params_default(dex_params *p) {
  if (!p) return;
  p->n = 2;
  p->delta = 0.5;
  p->theta = 0.0;
  p->charge = 2;
  p->nsmall = 22;
  p->tfactor = 4;
  p->precision = DEX_FP64;
  p->device = 0;
}

void dex_solve_opts_default(dex_solve_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->fp64_restarts = 0;
  o->fp32_tol = 1e-5;
}

const char *dex_status_string(int status) {
  switch (status) {
    case DEX_OK: return "ok";
    case DEX_ESIZE: return "size mismatch";
    case DEX_ECHOLESKY: return "Cholesky failed";
    case DEX_EGEVP: return "projected eigenvalue problem failed";
    case DEX_ENOTCONVERGED: return "not converged";
    case DEX_EHIP: return "HIP runtime error";
    case DEX_EOOM: return "device allocation failed";
    case DEX_EINVAL: return "invalid argument";
    case DEX_EADMISSION: return "does not fit the HBM cap";
    case DEX_EUNSUPPORTED: return "unsupported grid/precision";
    case DEX_EZEROBLOCK: return "all-zero block in cholQR";
    default: return "unknown status";
  }
}

const char *dex_last_error(void) { return g_last_error.c_str(); }

int dex_ctx_create(const dex_params *p, dex_ctx **out) {
  if (!p || !out) return fail(DEX_EINVAL, "null argument");
  *out = nullptr;
  if (p->n < 1 || !(p->delta > 0) || p->nsmall < 1 || p->tfactor < 1 ||
      (p->precision != DEX_FP64 && p->precision != DEX_FP32))
    return fail(DEX_EINVAL, "invalid parameters");
  const int nn = 2 * p->n + 1;
  if (!dex::generic_matvec_supported(nn)) return fail(DEX_EUNSUPPORTED, "NN not built (odd 3..25)");
  if (p->nsmall * p->tfactor < nn) return fail(DEX_EINVAL, "NSMALL*TFACTOR < NN (tooth.f90:38)");
  DEX_HIP(hipSetDevice(p->device));
  // kernel attributes and occupancy figures for every grid size, once per
  // process and before any launch (launches may come from several threads)
  static std::once_flag prepared;
  static hipError_t prep_err = hipSuccess;
  std::call_once(prepared, [] { prep_err = dex::prepare_kernels(); });
  if (prep_err != hipSuccess) return hip_fail(prep_err, "kernel attribute set-up");
  auto *c = new dex_ctx();
  c->p = *p;
  c->nn = nn;
  c->m = int64_t(nn) * nn * nn;
  c->d = c->m * c->m;
  c->t_host.resize(size_t(nn) * nn);
  dex::kematrix(nn, p->delta, c->t_host.data());
  if (dex::getinverse(nn, p->delta, p->nsmall, p->tfactor, c->tinv_host) != 0) {
    delete c;
    return fail(DEX_EINVAL, "GETINVERSE failed");
  }
  // A Tinv(r_p), A = -Z/delta^3 (hammod.F90:554,571), nucleus at the origin
  const int L = 2 * nn + 1;
  const int n = p->n;
  c->vn_host.resize(size_t(c->m));
  const double a = -double(p->charge) / (p->delta * p->delta * p->delta);
  for (int k = 0; k < nn; ++k)
    for (int j = 0; j < nn; ++j)
      for (int i = 0; i < nn; ++i) {
        const int ti = (i - n + nn) + L * ((j - n + nn) + L * (k - n + nn));
        c->vn_host[i + nn * (j + nn * k)] = a * c->tinv_host[ti];
      }
  // T / SCALE^2 and V / SCALE with SCALE = e^{i theta} (initmod.F90:48,207,224)
  const cd scale(std::cos(p->theta), std::sin(p->theta));
  const cd s2 = 1.0 / (scale * scale) / (p->delta * p->delta);
  const cd c1 = 1.0 / scale;
  c->coef.s2_re = s2.real();
  c->coef.s2_im = s2.imag();
  c->coef.c1_re = c1.real();
  c->coef.c1_im = c1.imag();
  c->coef.inv_d3 = 1.0 / (p->delta * p->delta * p->delta);
  c->fast64 = dex::fast_matvec_supported(nn, false);
  c->fast32 = dex::fast_matvec_supported(nn, true);
  // DEX_MATVEC=generic selects the six-pass FP64 kernels even where the LDS
  // two-pass kernels exist (used to cross-check the two implementations).
  if (const char *mv = std::getenv("DEX_MATVEC"))
    if (std::strcmp(mv, "generic") == 0) c->fast64 = false;
  // A blocking stream: work on it is ordered after work already queued on the
  // default stream (where callers such as torch place their kernels) and vice versa.
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);
  if (e == hipSuccess) e = hipMalloc(&c->d_tinv, c->tinv_host.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_vn, c->vn_host.size() * sizeof(double));
  if (e == hipSuccess)
    e = hipMemcpy(c->d_tinv, c->tinv_host.data(), c->tinv_host.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->d_vn, c->vn_host.data(), c->vn_host.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_desc, kMaxDesc * sizeof(ColDesc));
  if (e == hipSuccess) e = hipMalloc(&c->d_args, kMaxDesc * sizeof(ColDesc) + 2 * kMaxSmall * sizeof(Dbl2) + 512);
  if (e == hipSuccess) e = hipMalloc(&c->d_gram_out, 192 * 192 * sizeof(Dbl2));
  c->partial_bytes = std::max(dex::gram_partial_elems(64, 128) * sizeof(Dbl2),
                              dex::coldots_partial_elems(64) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_partial, c->partial_bytes);
  if (e == hipSuccess) e = hipMalloc(&c->d_dots, 64 * 4 * sizeof(double));
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
  if (e == hipSuccess) {
    c->stage_cap = size_t(4) << 20;
    e = hipHostMalloc(reinterpret_cast<void **>(&c->h_stage), c->stage_cap, hipHostMallocDefault);
  }
  // the allocation-free matvec's scratch on this device (the device is current)
  if (e == hipSuccess) e = scratch_address(p->device, &c->d_scratch);
  if (e != hipSuccess) {
    dex_ctx_destroy(c);
    return hip_fail(e, "dex_ctx_create");
  }
  *out = c;
  return DEX_OK;
}

int dex_ctx_destroy(dex_ctx *c) {
  if (!c) return DEX_OK;
  (void)hipSetDevice(c->p.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void *ptr : {(void *)c->d_tinv, (void *)c->d_vn, c->d_u, c->d_ub, (void *)c->d_desc,
                    (void *)c->d_args, (void *)c->d_gram_out, c->d_partial,
                    (void *)c->d_dots})
    if (ptr) (void)hipFree(ptr);
  for (auto &ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (const auto &pd : c->pending) {
    (void)hipEventDestroy(pd.a);
    (void)hipEventDestroy(pd.b);
  }
  for (hipEvent_t ev : c->free_events) (void)hipEventDestroy(ev);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return DEX_OK;
}

int dex_ctx_info(const dex_ctx *c, dex_info *info) {
  if (!c || !info) return fail(DEX_EINVAL, "null argument");
  info->n = c->p.n;
  info->nn = c->nn;
  info->m = c->m;
  info->d = c->d;
  info->delta = c->p.delta;
  info->theta = c->p.theta;
  info->precision = c->p.precision;
  info->fast_path = c->p.precision == DEX_FP32 ? int(c->fast32) : int(c->fast64);
  return DEX_OK;
}

int dex_get_kematrix(const dex_ctx *c, double *t) {
  if (!c || !t) return fail(DEX_EINVAL, "null argument");
  std::memcpy(t, c->t_host.data(), c->t_host.size() * sizeof(double));
  return DEX_OK;
}

int dex_get_tinv(const dex_ctx *c, double *tinv) {
  if (!c || !tinv) return fail(DEX_EINVAL, "null argument");
  std::memcpy(tinv, c->tinv_host.data(), c->tinv_host.size() * sizeof(double));
  return DEX_OK;
}

int dex_ctx_stop(dex_ctx *c, int32_t stop) {
  if (!c) return fail(DEX_EINVAL, "null context");
  c->stop.store(stop != 0 ? 1 : 0, std::memory_order_relaxed);
  return 0;
}

int dex_set_kernel_timing(dex_ctx *c, int32_t enable) {
  if (!c) return fail(DEX_EINVAL, "null argument");
  if (enable < 0) return fail(DEX_EINVAL, "timing period must be >= 0");
  c->timing = enable;
  return DEX_OK;
}

int dex_get_kernel_stats(dex_ctx *c, int32_t kclass, double *total_ms, int64_t *launches,
                         double *alg_bytes, double *alg_flops) {
  if (!c || !total_ms || !launches || !alg_bytes || !alg_flops) return fail(DEX_EINVAL, "null argument");
  if (kclass < 0 || kclass >= DEX_KCLASSES) return fail(DEX_EINVAL, "unknown kernel class");
  DEX_HIP(hipSetDevice(c->p.device));
  DEX_HIP(resolve_timing(c));
  *total_ms = c->k_ms[kclass];
  *launches = c->k_n[kclass];
  *alg_bytes = c->k_bytes[kclass];
  *alg_flops = c->k_flops[kclass];
  return DEX_OK;
}

int dex_get_kernel_totals(const dex_ctx *c, int32_t kclass, int64_t *launches, double *alg_bytes,
                          double *alg_flops) {
  if (!c || !launches || !alg_bytes || !alg_flops) return fail(DEX_EINVAL, "null argument");
  if (kclass < 0 || kclass >= DEX_KCLASSES) return fail(DEX_EINVAL, "unknown kernel class");
  *launches = c->k_all_n[kclass];
  *alg_bytes = c->k_all_bytes[kclass];
  *alg_flops = c->k_all_flops[kclass];
  return DEX_OK;
}

int dex_reset_kernel_stats(dex_ctx *c) {
  if (!c) return fail(DEX_EINVAL, "null argument");
  DEX_HIP(hipSetDevice(c->p.device));
  DEX_HIP(resolve_timing(c));
  for (int i = 0; i < DEX_KCLASSES; ++i) {
    c->k_ms[i] = c->k_bytes[i] = c->k_flops[i] = 0.0;
    c->k_n[i] = 0;
    c->k_all_bytes[i] = c->k_all_flops[i] = 0.0;
    c->k_all_n[i] = 0;
  }
  return DEX_OK;
}

int dex_get_kernel_timing(const dex_ctx *c, double *p1, double *p2) {
  if (!c || !p1 || !p2) return fail(DEX_EINVAL, "null argument");
  *p1 = c->t_pass1;
  *p2 = c->t_pass2;
  return DEX_OK;
}

int dex_matvec_device(dex_ctx *c, const void *in, void *out, int32_t howmany, int64_t checksize,
                      int32_t single) {
  return dex_matvec_sym_device(c, in, out, howmany, checksize, single, 0);
}

int dex_matvec_sym_device(dex_ctx *c, const void *in, void *out, int32_t howmany, int64_t checksize,
                          int32_t single, int32_t symmetry) {
  if (symmetry < -1 || symmetry > 1) return fail(DEX_EINVAL, "symmetry must be -1, 0 or 1");
  if (!c || !in || !out) return fail(DEX_EINVAL, "null argument");
  // an FP32 context applies the FP32 kernels to complex128 columns as well
  if (symmetry != 0 && !((single || c->p.precision == DEX_FP32) ? c->fast32 : c->fast64))
    return fail(DEX_EUNSUPPORTED, "exchange-symmetric matvec needs the LDS kernels for this grid");
  if (checksize != c->d) return fail(DEX_ESIZE, "SIZE ERROR: checksize != M*MBLK");
  if (howmany < 0) return fail(DEX_EINVAL, "howmany < 0");
  DEX_HIP(hipSetDevice(c->p.device));
  const bool sp = single != 0;
  if (sp && c->p.precision != DEX_FP32) return fail(DEX_EINVAL, "complex64 I/O needs an FP32 context");
  if (sp && !c->fast32) return fail(DEX_EUNSUPPORTED, "no FP32 matvec for this grid");
  if (int st = ensure_u(c)) return st;
  const int nb = int(batch_cols(c));
  if (!sp && symmetry != 0 && !c->timing && howmany >= 2 && nb >= 2 &&
      (c->p.precision == DEX_FP32 ? c->fast32 : c->fast64)) {
    // exchange-symmetric complex128 columns: nb per pair of launches
    std::vector<const Dbl2 *> xs(howmany);
    std::vector<Dbl2 *> ys(howmany);
    for (int j = 0; j < howmany; ++j) {
      xs[j] = static_cast<const Dbl2 *>(in) + size_t(j) * c->d;
      ys[j] = static_cast<Dbl2 *>(out) + size_t(j) * c->d;
    }
    for (int j0 = 0; j0 < howmany; j0 += nb) {
      const int k = std::min(nb, howmany - j0);
      if (int st = apply_h_batch(c, xs.data() + j0, ys.data() + j0, k,
                                 c->p.precision == DEX_FP32 ? dex::kIoMixed : dex::kIoF64, symmetry))
        return st;
    }
    return DEX_OK;
  }
  for (int j = 0; j < howmany; ++j) {
    if (sp) {
      const char *xi = static_cast<const char *>(in) + size_t(j) * c->d * 8;
      char *yo = static_cast<char *>(out) + size_t(j) * c->d * 8;
      if (c->timing) DEX_HIP(hipEventRecord(c->ev[0], c->stream));
      DEX_HIP(dex::launch_pass1(c->nn, dex::kIoF32, xi, c->d_u, c->stream));
      if (c->timing) DEX_HIP(hipEventRecord(c->ev[1], c->stream));
      DEX_HIP(second_pass(c, dex::kIoF32, xi, yo, symmetry));
      if (c->timing) DEX_HIP(hipEventRecord(c->ev[2], c->stream));
    } else {
      const Dbl2 *xi = static_cast<const Dbl2 *>(in) + size_t(j) * c->d;
      Dbl2 *yo = static_cast<Dbl2 *>(out) + size_t(j) * c->d;
      if (int st = apply_h(c, xi, yo, c->p.precision == DEX_FP32, symmetry, true)) return st;
    }
  }
  const bool timed_fast = c->p.precision == DEX_FP32 ? c->fast32 : c->fast64;
  if (c->timing && howmany > 0 && timed_fast) {
    DEX_HIP(hipEventSynchronize(c->ev[2]));
    float a = 0.f, b = 0.f;
    DEX_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    DEX_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    c->t_pass1 = a;
    c->t_pass2 = b;
  }
  return DEX_OK;
}

int dex_matvec_noalloc(dex_ctx *c, const void *in, void *out, int32_t howmany, int64_t checksize,
                       int32_t single, int32_t symmetry) {
  if (symmetry < -1 || symmetry > 1) return fail(DEX_EINVAL, "symmetry must be -1, 0 or 1");
  if (!c || !in || !out) return fail(DEX_EINVAL, "null argument");
  if (checksize != c->d) return fail(DEX_ESIZE, "SIZE ERROR: checksize != M*MBLK");
  if (howmany < 0) return fail(DEX_EINVAL, "howmany < 0");
  const bool sp = single != 0;
  if (sp && c->p.precision != DEX_FP32) return fail(DEX_EINVAL, "complex64 I/O needs an FP32 context");
  const bool f32 = c->p.precision == DEX_FP32;  // FP32 arithmetic (complex64 U)
  const bool fast = f32 ? c->fast32 : c->fast64;
  if (!fast && (f32 || symmetry != 0))
    return fail(DEX_EUNSUPPORTED, "this grid and precision have no LDS matvec kernels");
  if (fast && size_t(c->d) * (f32 ? 8 : 16) > kScratchBytes)
    return fail(DEX_ESIZE, "dex_matvec_noalloc: the intermediate (NN^6 complex128, or complex64 in an FP32 "
                           "context) exceeds the 64 MiB device scratch: NN <= 11 in FP64, NN <= 13 in FP32");
  if (!c->d_scratch) return fail(DEX_EHIP, "no device scratch address");
  DEX_HIP(hipSetDevice(c->p.device));
  const int io = sp ? dex::kIoF32 : (f32 ? dex::kIoMixed : dex::kIoF64);
  const size_t xb = sp ? 8 : 16;  // bytes per entry of X and Y
  for (int j = 0; j < howmany; ++j) {
    const char *x = static_cast<const char *>(in) + size_t(j) * size_t(c->d) * xb;
    char *y = static_cast<char *>(out) + size_t(j) * size_t(c->d) * xb;
    if (!fast) {
      DEX_HIP(dex::launch_matvec_generic(c->nn, reinterpret_cast<const Dbl2 *>(x), reinterpret_cast<Dbl2 *>(y),
                                         c->d_vn, c->d_tinv, c->coef, c->stream));
      continue;
    }
    // U (fully written by the electron-1 pass before the second pass reads it)
    DEX_HIP(dex::launch_pass1(c->nn, io, x, c->d_scratch, c->stream));
    if (symmetry != 0)
      DEX_HIP(dex::launch_symadd(c->nn, io, x, c->d_scratch, y, c->d_vn, c->d_tinv, c->coef, symmetry, c->stream));
    else
      DEX_HIP(dex::launch_pass2(c->nn, io, x, c->d_scratch, y, c->d_vn, c->d_tinv, c->coef, c->stream));
  }
  return DEX_OK;
}

struct dex_graph {
  dex_ctx *ctx;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int dex_matvec_graph_create(dex_ctx *c, const void *in, void *out, int32_t howmany, int64_t checksize,
                            int32_t single, int32_t symmetry, dex_graph **out_graph) {
  if (!c || !out_graph) return fail(DEX_EINVAL, "null argument");
  *out_graph = nullptr;
  if (howmany < 1) return fail(DEX_EINVAL, "howmany < 1");
  DEX_HIP(hipSetDevice(c->p.device));
  // The kernels' attributes were set when the context was created, so the capture
  // records kernel launches only: no allocation, copy, memset or synchronisation.
  DEX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  const int st = dex_matvec_noalloc(c, in, out, howmany, checksize, single, symmetry);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
  if (st != DEX_OK) {  // its own message stands
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return hip_fail(ec, "hipStreamEndCapture");
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(ei, "hipGraphInstantiate");
  }
  auto *g = new (std::nothrow) dex_graph{c, graph, exec};
  if (!g) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return fail(DEX_EOOM, "graph handle");
  }
  *out_graph = g;
  return DEX_OK;
}

int dex_graph_launch(dex_graph *g) {
  if (!g) return fail(DEX_EINVAL, "null graph");
  DEX_HIP(hipSetDevice(g->ctx->p.device));
  DEX_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
  return DEX_OK;
}

int dex_graph_destroy(dex_graph *g) {
  if (!g) return DEX_OK;
  DEX_HIP(hipSetDevice(g->ctx->p.device));
  const hipError_t e1 = hipGraphExecDestroy(g->exec);
  const hipError_t e2 = hipGraphDestroy(g->graph);
  delete g;
  if (e1 != hipSuccess) return hip_fail(e1, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return hip_fail(e2, "hipGraphDestroy");
  return DEX_OK;
}

int dex_matvec(dex_ctx *c, const double *in, double *out, int32_t howmany, int64_t checksize) {
  if (!c || !in || !out) return fail(DEX_EINVAL, "null argument");
  if (checksize != c->d) return fail(DEX_ESIZE, "SIZE ERROR: checksize != M*MBLK");
  DEX_HIP(hipSetDevice(c->p.device));
  Dbl2 *dx = nullptr, *dy = nullptr;
  const size_t bytes = size_t(c->d) * sizeof(Dbl2);
  DEX_HIP(hipMalloc(&dx, bytes));
  hipError_t e = hipMalloc(&dy, bytes);
  if (e != hipSuccess) {
    (void)hipFree(dx);
    return hip_fail(e, "hipMalloc");
  }
  int st = DEX_OK;
  for (int j = 0; j < howmany && st == DEX_OK; ++j) {
    e = hipMemcpyAsync(dx, in + size_t(j) * c->d * 2, bytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      st = hip_fail(e, "copy in");
      break;
    }
    st = dex_matvec_device(c, dx, dy, 1, checksize, 0);
    if (st) break;
    e = hipMemcpyAsync(out + size_t(j) * c->d * 2, dy, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = hip_fail(e, "copy out");
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  return st;
}

int dex_dots_device(dex_ctx *c, const void *bras, const void *kets, int64_t size, int32_t nn,
                    int32_t mm, double *out) {
  if (!c || !bras || !kets || !out) return fail(DEX_EINVAL, "null argument");
  if (size != c->d) return fail(DEX_ESIZE, "SIZE ERROR: size != M*MBLK");
  if (nn < 1 || mm < 1) return fail(DEX_EINVAL, "nn, mm out of range");
  DEX_HIP(hipSetDevice(c->p.device));
  std::vector<ColDesc> l, r;
  for (int i = 0; i < nn; ++i) l.push_back(plain(static_cast<const Dbl2 *>(bras) + size_t(i) * size));
  for (int j = 0; j < mm; ++j) r.push_back(plain(static_cast<const Dbl2 *>(kets) + size_t(j) * size));
  std::vector<cd> g;
  if (int st = gram(c, l, r, g)) return st;
  std::memcpy(out, g.data(), g.size() * sizeof(cd));
  return DEX_OK;
}

int dex_dots(dex_ctx *c, const double *bras, const double *kets, int64_t size, int32_t nn,
             int32_t mm, double *out) {
  if (!c || !bras || !kets || !out) return fail(DEX_EINVAL, "null argument");
  if (size != c->d) return fail(DEX_ESIZE, "SIZE ERROR: size != M*MBLK");
  DEX_HIP(hipSetDevice(c->p.device));
  void *db = nullptr, *dk = nullptr;
  DEX_HIP(hipMalloc(&db, size_t(nn) * size * 16));
  hipError_t e = hipMalloc(&dk, size_t(mm) * size * 16);
  if (e != hipSuccess) {
    (void)hipFree(db);
    return hip_fail(e, "hipMalloc");
  }
  int st = DEX_OK;
  e = hipMemcpy(db, bras, size_t(nn) * size * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dk, kets, size_t(mm) * size * 16, hipMemcpyHostToDevice);
  if (e != hipSuccess)
    st = hip_fail(e, "copy in");
  else
    st = dex_dots_device(c, db, dk, size, nn, mm, out);
  (void)hipFree(db);
  (void)hipFree(dk);
  return st;
}

}  // extern "C"

// ============================================================================
// GPLMR
namespace {

struct Solver {
  dex_ctx *c;
  int nev, ns;
  bool strict;
  int sym;      // 0, or +-1 for exchange-(anti)symmetric blocks
  int fallbacks = 0;
  int64_t matvecs = 0;
  double t_matvec = 0, t_gram = 0, t_update = 0, t_small = 0, t_prec = 0;
  bool inner32 = false;  // inner solve: complex64 Arnoldi basis and FP32 matvecs
  std::vector<Dbl2 *> pool;    // free complex128 work vectors
  std::vector<Dbl2 *> pool32;  // free complex64 work vectors (tagged pointers)

  // H on every column of `in` into `out`.  complex128 columns get the FP64
  // matvec, complex64 columns (both sides) the FP32 one.
  int matblock(const Block &in, const Block &out) {
    const double t0 = now_s();
    if (out.size() < in.size()) return fail(DEX_EINVAL, "matblock: output block too small");
    const int nb = int(batch_cols(c));
    bool same = true;
    for (int j = 0; j < in.size(); ++j)
      if (is32(in.cols[j]) != is32(in.cols[0]) || is32(out.cols[j]) != is32(in.cols[j])) same = false;
    const bool lo = in.size() > 0 && is32(in.cols[0]);
    if (same && sym != 0 && nb >= 2 && in.size() >= 2 && (lo ? c->fast32 : c->fast64)) {
      // exchange-symmetric columns: nb columns per pair of launches
      std::vector<Dbl2 *> xi, yo;
      for (int j = 0; j < in.size(); ++j) {
        xi.push_back(untag(in.cols[j]));
        yo.push_back(untag(out.cols[j]));
      }
      for (int j0 = 0; j0 < in.size(); j0 += nb) {
        const int k = std::min(nb, in.size() - j0);
        if (int st = apply_h_batch(c, xi.data() + j0, yo.data() + j0, k, lo ? dex::kIoF32 : dex::kIoF64, sym))
          return st;
        matvecs += k;
      }
      t_matvec += now_s() - t0;
      return 0;
    }
    for (int j = 0; j < in.size(); ++j) {
      if (is32(in.cols[j]) != is32(out.cols[j])) return fail(DEX_EINVAL, "matblock: mixed storage formats");
      if (is32(in.cols[j])) {
        if (int st = apply_h32(c, untag(in.cols[j]), untag(out.cols[j]), sym)) return st;
      } else if (int st = apply_h(c, in.cols[j], out.cols[j], false, sym)) {
        return st;
      }
      ++matvecs;
    }
    // no synchronisation: the next block operation is ordered after these
    // launches on the same stream (t_matvec is the host time to issue them)
    t_matvec += now_s() - t0;
    return 0;
  }

  // cholQR (lamod.f90:128-180) with the rank-revealing fallback.  Returns in
  // `t` (k x k', column-major) the transform with xq = x t; k' < k only when
  // `drop` removed dependent directions.
  int orth_transform(const Block &x, bool drop, std::vector<cd> &t, int &kout) {
    const int k = x.size();
    std::vector<cd> g;
    double t0 = now_s();
    if (int st = gram_self(c, descs_of(x), g)) return st;
    t_gram += now_s() - t0;
    t0 = now_s();
    double asum = 0.0;
    for (const cd &v : g) asum += std::abs(v);
    if (asum == 0.0) {
      if (drop) {
        kout = 0;
        t.clear();
        return 0;
      }
      return fail(DEX_EZEROBLOCK, "all-zero block in cholQR");
    }
    std::vector<double> dg(k);
    bool ok = true;
    for (int i = 0; i < k; ++i) {
      dg[i] = std::sqrt(std::max(g[i + size_t(k) * i].real(), 0.0));
      if (!(dg[i] > 0)) ok = false;
    }
    std::vector<cd> r = g;
    if (ok && dex::zpotrf_upper(k, r.data(), k) == 0) {
      for (int i = 0; i < k && ok; ++i) {
        const double piv = r[i + size_t(k) * i].real() / dg[i];
        if (!(piv * piv > 1e-12)) ok = false;
      }
    } else {
      ok = false;
    }
    if (ok) {
      dex::ztrtri_upper(k, r.data(), k);
      t = r;
      kout = k;
      t_small += now_s() - t0;
      return 0;
    }
    if (strict) return fail(DEX_ECHOLESKY, "Cholesky failed");
    ++fallbacks;
    // SVQB on the unit-diagonal Gram
    std::vector<cd> gs(size_t(k) * k);
    std::vector<double> dinv(k);
    for (int i = 0; i < k; ++i) dinv[i] = dg[i] > 0 ? 1.0 / dg[i] : 0.0;
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) gs[i + size_t(k) * j] = dinv[i] * g[i + size_t(k) * j] * dinv[j];
    std::vector<double> w;
    std::vector<cd> v;
    dex::zheev_jacobi(k, gs.data(), k, w, v);
    const double lmax = *std::max_element(w.begin(), w.end());
    std::vector<int> keep;
    for (int i = 0; i < k; ++i)
      if (!drop || w[i] > 1e-12 * lmax) keep.push_back(i);
    kout = int(keep.size());
    t.assign(size_t(k) * kout, 0.0);
    for (int jj = 0; jj < kout; ++jj) {
      const int j = keep[jj];
      const double lam = std::max(w[j], 1e-12 * lmax);
      for (int i = 0; i < k; ++i) t[i + size_t(k) * jj] = dinv[i] * v[i + size_t(k) * j] / std::sqrt(lam);
    }
    t_small += now_s() - t0;
    return 0;
  }

  // Orthonormalise block x in place (and its image ax if given) with a second
  // cholQR pass after a fallback.  Returns the new column count.
  int orth(Block &x, Block *ax, bool drop) {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<cd> t;
      int kout = 0;
      const int k = x.size();
      const int before = fallbacks;
      if (int st = orth_transform(x, drop, t, kout)) return -st;
      const double t0 = now_s();
      if (kout > 0) {
        std::vector<ColDesc> in = descs_of(x);
        std::vector<Dbl2 *> out(x.cols.begin(), x.cols.begin() + kout);
        if (int st = combine(c, in, out, t, nullptr)) return -st;
        if (ax) {
          std::vector<ColDesc> ain = descs_of(*ax);
          std::vector<Dbl2 *> aout(ax->cols.begin(), ax->cols.begin() + kout);
          if (int st = combine(c, ain, aout, t, nullptr)) return -st;
        }
      }
      t_update += now_s() - t0;
      for (int j = kout; j < k; ++j) give(x.cols[j]);
      x.cols.resize(kout);
      if (ax) {
        for (int j = kout; j < k; ++j) give(ax->cols[j]);
        ax->cols.resize(kout);
      }
      if (fallbacks == before || kout == 0) break;  // plain cholQR path: one pass
    }
    return 0;
  }

  // b <- b - q (q^H b); the image (ab <- ab - aq (q^H b)) when given.
  int project_out(const Block &q, const Block *aq, Block &b, Block *ab) {
    if (q.size() == 0 || b.size() == 0) return 0;
    std::vector<cd> g;
    double t0 = now_s();
    if (int st = gram(c, descs_of(q), descs_of(b), g)) return st;
    t_gram += now_s() - t0;
    t0 = now_s();
    const int kq = q.size(), kb = b.size();
    // b <- [b q] [I; -G]: q in chunks that fit one launch, the first chunk
    // carrying b itself (in place), later ones accumulating with beta = 1.
    std::vector<cd> one(kb, 1.0);
    for (int q0 = 0; q0 < kq;) {
      const bool first = q0 == 0;
      const int nq = std::min(first ? 64 - kb : 64, kq - q0);
      const int nin = (first ? kb : 0) + nq;
      const int off = first ? kb : 0;
      std::vector<cd> cm(size_t(nin) * kb, 0.0);
      for (int j = 0; j < kb; ++j) {
        if (first) cm[j + size_t(nin) * j] = 1.0;
        for (int i = 0; i < nq; ++i) cm[off + i + size_t(nin) * j] = -g[q0 + i + size_t(kq) * j];
      }
      std::vector<ColDesc> in;
      if (first) in = descs_of(b);
      for (int i = 0; i < nq; ++i) in.push_back(plain(q.cols[q0 + i]));
      if (int st = combine(c, in, b.cols, cm, first ? nullptr : &one)) return st;
      if (ab) {
        std::vector<ColDesc> ain;
        if (first) ain = descs_of(*ab);
        for (int i = 0; i < nq; ++i) ain.push_back(plain(aq->cols[q0 + i]));
        if (int st = combine(c, ain, ab->cols, cm, first ? nullptr : &one)) return st;
      }
      q0 += nq;
    }
    t_update += now_s() - t0;
    return 0;
  }

  // G = [w; q]^H w ((k + kq) x k, rows w then q): one pass over w and q
  // (q in chunks of up to 64 columns per launch beyond the first).
  int gram_wq(const std::vector<ColDesc> &w, const Block &q, std::vector<cd> &g) {
    const int k = int(w.size()), kq = q.size(), rows = k + kq;
    if (k < 1 || k > 48) return fail(DEX_EINVAL, "block width must be 1..48");
    g.assign(size_t(rows) * k, 0.0);
    const double t0 = now_s();
    const int n1 = std::min(kq, 64 - k);
    std::vector<ColDesc> ds(w);
    for (int i = 0; i < n1; ++i) ds.push_back(plain(q.cols[i]));
    if (int st = gram_launch_copy(c, ds, k + n1, k, k, g, rows, 0, 0)) return st;
    for (int i0 = n1; i0 < kq; i0 += 64) {
      const int bl = std::min(64, kq - i0);
      std::vector<ColDesc> d2;
      for (int i = 0; i < bl; ++i) d2.push_back(plain(q.cols[i0 + i]));
      d2.insert(d2.end(), w.begin(), w.end());
      if (int st = gram_launch_copy(c, d2, bl, k, 0, g, rows, k + i0, 0)) return st;
    }
    t_gram += now_s() - t0;
    return 0;
  }

  // out = [w q] [T; -Cq T] (k x k T, kq x k Cq), written over the columns of
  // `out` (which may be w's own storage: the combine is row-local and w is read
  // in the first launch only).
  int apply_wq(const std::vector<ColDesc> &w, const Block &q, const std::vector<cd> &t,
               const std::vector<cd> &cq, Block &out) {
    const int k = int(w.size()), kq = q.size();
    const double t0 = now_s();
    std::vector<cd> one(k, 1.0);
    // -Cq T
    std::vector<cd> ct(size_t(kq) * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int l = 0; l < k; ++l) {
        const cd tlj = t[l + size_t(k) * j];
        if (tlj == cd(0.0))
          continue;
        for (int i = 0; i < kq; ++i) ct[i + size_t(kq) * j] -= cq[i + size_t(kq) * l] * tlj;
      }
    for (int q0 = 0; q0 == 0 || q0 < kq;) {
      const bool first = q0 == 0;
      const int nq = std::min(first ? 64 - k : 64, kq - q0);
      const int nin = (first ? k : 0) + nq;
      const int off = first ? k : 0;
      std::vector<cd> cm(size_t(nin) * k, 0.0);
      for (int j = 0; j < k; ++j) {
        if (first)
          for (int i = 0; i < k; ++i) cm[i + size_t(nin) * j] = t[i + size_t(k) * j];
        for (int i = 0; i < nq; ++i) cm[off + i + size_t(nin) * j] = ct[q0 + i + size_t(kq) * j];
      }
      std::vector<ColDesc> in;
      if (first) in = w;
      for (int i = 0; i < nq; ++i) in.push_back(plain(q.cols[q0 + i]));
      if (int st = combine(c, in, out.cols, cm, first ? nullptr : &one)) return st;
      q0 += nq;
      if (nq == 0) break;
    }
    t_update += now_s() - t0;
    return 0;
  }

  // Projected Gram P = W^H W - Cq^H Cq of (I - q q^H) w from g = [w; q]^H w,
  // and the smallest ratio P_jj / (W^H W)_jj (the fraction of a column's norm
  // squared that survives the projection).
  static void projected_gram(const std::vector<cd> &g, int k, int kq, std::vector<cd> &p,
                             double &min_keep) {
    const int rows = k + kq;
    p.assign(size_t(k) * k, 0.0);
    min_keep = 1.0;
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        cd v = g[i + size_t(rows) * j];
        for (int l = 0; l < kq; ++l) v -= std::conj(g[k + l + size_t(rows) * i]) * g[k + l + size_t(rows) * j];
        p[i + size_t(k) * j] = v;
      }
    for (int j = 0; j < k; ++j) {
      const double w2 = g[j + size_t(rows) * j].real();
      const double keep = w2 > 0.0 ? p[j + size_t(k) * j].real() / w2 : 0.0;
      min_keep = std::min(min_keep, keep);
    }
  }

  // Upper Cholesky factor R of the Hermitian p and its inverse, when every
  // pivot keeps more than 1e-12 of the corresponding column's norm squared
  // (w2: those norms); false otherwise.
  static bool chol_inverse(const std::vector<cd> &p, const std::vector<double> &w2, int k, std::vector<cd> &r,
                           std::vector<cd> &rinv) {
    r = p;
    if (dex::zpotrf_upper(k, r.data(), k) != 0) return false;
    for (int i = 0; i < k; ++i) {
      const double piv = r[i + size_t(k) * i].real();
      if (!(piv > 0.0) || !(piv * piv > 1e-12 * w2[i])) return false;
    }
    rinv = r;
    dex::ztrtri_upper(k, rinv.data(), k);
    for (int j = 0; j < k; ++j)
      for (int i = j + 1; i < k; ++i) r[i + size_t(k) * j] = 0.0;
    return true;
  }

  // Block classical Gram-Schmidt with cholQR in one pass: w <- (w - q C) R^{-1}
  // with C = q^H w and R^H R = w^H w - C^H C, i.e. one Gram [w; q]^H w and one
  // combine over [w q].  A second pass runs when the projection removed more
  // than the fraction 1 - keep_min of some column's norm squared: keep_min =
  // 0.5 is the criterion of Daniel, Gragg, Kaufman and Stewart (Math. Comp. 30
  // (1976) 772-795); the subtraction loses about eps / keep of relative
  // accuracy in R, so a smaller keep_min trades orthogonality for a pass.  q
  // must be orthonormal.  The input is `wd` (descriptors: plain columns or a - sigma b),
  // the result goes to `w` (which may be the storage wd reads).  On success
  // `cq` and `r` satisfy w_in = q cq + w_out r; `ok` is false when the
  // projected block is numerically rank deficient (then w holds the projected,
  // unnormalised block and the caller decides).  `materialize` false skips the
  // final combine (the Arnoldi step that only needs C and R).
  int cgs_cholqr(const Block &q, const std::vector<ColDesc> &wd, Block &w, std::vector<cd> &cq,
                 std::vector<cd> &r, bool &ok, bool materialize = true, double keep_min = 0.5) {
    const int k = int(wd.size()), kq = q.size();
    std::vector<cd> g, p, rinv;
    double keep = 1.0;
    if (int st = gram_wq(wd, q, g)) return st;
    double t0 = now_s();
    projected_gram(g, k, kq, p, keep);
    std::vector<double> w2(k);
    for (int j = 0; j < k; ++j) w2[j] = g[j + size_t(k + kq) * j].real();
    cq.assign(size_t(kq) * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < kq; ++i) cq[i + size_t(kq) * j] = g[k + i + size_t(k + kq) * j];
    bool good = keep > keep_min && chol_inverse(p, w2, k, r, rinv);
    t_small += now_s() - t0;
    if (good) {
      ok = true;
      if (!materialize) return 0;
      return apply_wq(wd, q, rinv, cq, w);
    }
    // second pass on the projected block
    std::vector<cd> id(size_t(k) * k, 0.0);
    for (int j = 0; j < k; ++j) id[j + size_t(k) * j] = 1.0;
    if (int st = apply_wq(wd, q, id, cq, w)) return st;
    if (kq == 0) {
      ok = false;
      return 0;
    }
    std::vector<cd> g2, cq2;
    if (int st = gram_wq(descs_of(w), q, g2)) return st;
    t0 = now_s();
    projected_gram(g2, k, kq, p, keep);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < kq; ++i) cq[i + size_t(kq) * j] += g2[k + i + size_t(k + kq) * j];
    cq2.assign(size_t(kq) * k, 0.0);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < kq; ++i) cq2[i + size_t(kq) * j] = g2[k + i + size_t(k + kq) * j];
    good = keep > 0.5 && chol_inverse(p, w2, k, r, rinv);
    t_small += now_s() - t0;
    if (!good) {
      ok = false;
      if (int st = apply_wq(descs_of(w), q, id, cq2, w)) return st;
      return 0;
    }
    ok = true;
    if (!materialize) return 0;
    return apply_wq(descs_of(w), q, rinv, cq2, w);
  }

  // w <- orthonormal basis of (I - q q^H) w (q orthonormal), in place, with
  // the optional image block aw transformed alike (aq being q's image).  The
  // fused pass above; where it reports rank deficiency, the projection is
  // completed and the rank-revealing cholQR (orth) takes over, dropping
  // dependent directions when `drop`.
  int orth_against(const Block &q, Block &w, const Block *aq, Block *aw, bool drop) {
    if (w.size() == 0) return 0;
    if (aw == nullptr) {
      std::vector<cd> cq, r;
      bool ok = false;
      if (int st = cgs_cholqr(q, descs_of(w), w, cq, r, ok)) return st;
      if (ok) return 0;
      const int kk = orth(w, nullptr, drop);
      return kk < 0 ? -kk : 0;
    }
    // with an image block: the same coefficients for both
    const int k = w.size(), kq = q.size();
    std::vector<cd> g, p, r, rinv;
    double keep = 1.0;
    if (int st = gram_wq(descs_of(w), q, g)) return st;
    const double t0 = now_s();
    projected_gram(g, k, kq, p, keep);
    std::vector<double> w2(k);
    for (int j = 0; j < k; ++j) w2[j] = g[j + size_t(k + kq) * j].real();
    std::vector<cd> cq(size_t(kq) * k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < kq; ++i) cq[i + size_t(kq) * j] = g[k + i + size_t(k + kq) * j];
    const bool good = keep > 0.5 && chol_inverse(p, w2, k, r, rinv);
    t_small += now_s() - t0;
    if (good) {
      if (int st = apply_wq(descs_of(*aw), *aq, rinv, cq, *aw)) return st;
      return apply_wq(descs_of(w), q, rinv, cq, w);
    }
    if (kq > 0) {
      if (int st = project_out(q, aq, w, aw)) return st;
      if (int st = project_out(q, aq, w, aw)) return st;
    }
    const int kk = orth(w, aw, drop);
    return kk < 0 ? -kk : 0;
  }

  // PRECBLOCK_BASIC -> basicblocklansolve (gplmr.F90:674-770;
  // harmonic_ritz_small.F90:258-324): b <- argmin over x in K_m(H, b) of
  // ||(H - sigma) x - b||, the fixed-order unrestarted block minimal-residual
  // approximation to (H - sigma)^{-1} b, for every column of the block (the
  // reference's allhmultsub multiplies only the first, gplmr.F90:714).
  //
  // Same Krylov space (K_m(H, b) = K_m(H - sigma, b)) and the same
  // least-squares minimiser as the reference, computed as block GMRES: an
  // orthonormal Arnoldi basis Q of K_m(H - sigma, b) with
  // (H - sigma) Q_m = Q_{m+1} Hbar, one matvec per basis column, one fused
  // projection/cholQR pass per block, x = Q_m y with
  // y = argmin ||[R0; 0] - Hbar y|| (b = Q_1 R0).  The reference instead keeps
  // both V and B = (H - sigma) V, two matvecs per column and two Gram-Schmidt
  // passes over both (harmonic_ritz_small.F90:100-155).
  int inner_solve(Block &b, int m, cd sigma) {
    const int k = b.size();
    if (k == 0 || m < 1) return 0;
    const double t0 = now_s();
    const int bf = inner32 ? 1 : 0;  // storage of the Arnoldi basis
    Block qall;
    qall.fmt = bf;
    // Q_1 R0 = b (into fresh storage: b is overwritten only by the solution)
    Block q1;
    q1.fmt = bf;
    if (int st = grow(q1, k)) return st;
    std::vector<cd> cq, r0;
    bool ok = false;
    if (int st = cgs_cholqr(Block{}, descs_of(b), q1, cq, r0, ok)) return st;
    int kb = k;  // Arnoldi block width
    if (!ok) {
      // rank-deficient right-hand side (q1 holds a copy of b): an orthonormal
      // basis of its range, and R0 = Q1^H b
      const int kk = orth(q1, nullptr, true);
      if (kk < 0) return -kk;
      kb = q1.size();
      if (kb == 0) {
        t_prec += now_s() - t0;
        return 0;  // b = 0: x = 0, which b already holds
      }
      const double tg = now_s();
      if (int st = gram(c, descs_of(q1), descs_of(b), r0)) return st;
      t_gram += now_s() - tg;
    }
    qall = q1;
    // Hbar by block columns: column block j holds C (rows of all earlier
    // blocks) and R_j (the new block's rows)
    std::vector<std::vector<cd>> hc;
    std::vector<std::vector<cd>> hr;
    int nblk = 0;
    for (int j = 0; j < m; ++j) {
      Block qj;
      qj.fmt = bf;
      qj.cols.assign(qall.cols.end() - kb, qall.cols.end());
      Block hq;
      hq.fmt = bf;
      if (int st = grow(hq, kb)) return st;
      if (int st = matblock(qj, hq)) return st;
      std::vector<ColDesc> wd;
      for (int i = 0; i < kb; ++i) wd.push_back(shifted(hq.cols[i], qj.cols[i], sigma));
      std::vector<cd> cj, rj;
      const bool last = j == m - 1;
      // a preconditioner's basis: one pass unless the projection cancels
      // more than 99.9% of a column (R keeps ~1e-13 relative accuracy)
      if (int st = cgs_cholqr(qall, wd, hq, cj, rj, ok, !last, 1e-3)) return st;
      if (!ok) rj.assign(size_t(kb) * kb, 0.0);  // (H - sigma) Q_j lies in span(Q): the space is exhausted
      hc.push_back(cj);
      hr.push_back(rj);
      ++nblk;
      if (!ok || last) {
        for (Dbl2 *p : hq.cols) give(p);
        break;
      }
      qall.cols.insert(qall.cols.end(), hq.cols.begin(), hq.cols.end());
    }
    // least squares on the (nblk + 1) k x nblk k block Hessenberg matrix
    const double ts = now_s();
    const int rows = (nblk + 1) * kb, cols = nblk * kb;
    std::vector<cd> hb(size_t(rows) * cols, 0.0), rhs(size_t(rows) * k, 0.0);
    for (int j = 0; j < nblk; ++j) {
      const int kq = (j + 1) * kb;  // rows of C for block column j
      for (int cc = 0; cc < kb; ++cc) {
        for (int i = 0; i < kq; ++i) hb[i + size_t(rows) * (j * kb + cc)] = hc[j][i + size_t(kq) * cc];
        for (int i = 0; i < kb; ++i) hb[kq + i + size_t(rows) * (j * kb + cc)] = hr[j][i + size_t(kb) * cc];
      }
    }
    for (int cc = 0; cc < k; ++cc)
      for (int i = 0; i < kb; ++i) rhs[i + size_t(rows) * cc] = r0[i + size_t(kb) * cc];
    // block Hessenberg with upper triangular subdiagonal blocks: kb subdiagonals
    const int lst = dex::zlstsq(rows, cols, hb.data(), rows, rhs.data(), rows, k, kb);
    t_small += now_s() - ts;
    if (lst != 0) {
      for (Dbl2 *p : qall.cols) give(p);
      return fail(DEX_EGEVP, "inner solve: projected least-squares problem is singular");
    }
    std::vector<cd> y(size_t(cols) * k);
    for (int cc = 0; cc < k; ++cc)
      for (int i = 0; i < cols; ++i) y[i + size_t(cols) * cc] = rhs[i + size_t(rows) * cc];
    Block qm;
    qm.fmt = bf;
    qm.cols.assign(qall.cols.begin(), qall.cols.begin() + cols);
    const double tu = now_s();
    if (int st = combine_chunked(qm, y, b)) return st;
    t_update += now_s() - tu;
    for (Dbl2 *p : qall.cols) give(p);
    t_prec += now_s() - t0;
    return 0;
  }

  // out = in C for an input block of any width (64 columns per launch,
  // accumulated with beta = 1); `out` must not alias `in`.
  int combine_chunked(const Block &in, const std::vector<cd> &cm, Block &out) {
    const int nin = in.size(), nout = out.size();
    for (int i0 = 0; i0 < nin; i0 += 64) {
      const int bi = std::min(64, nin - i0);
      std::vector<ColDesc> ds;
      for (int i = 0; i < bi; ++i) ds.push_back(plain(in.cols[i0 + i]));
      std::vector<cd> sub(size_t(bi) * nout);
      for (int j = 0; j < nout; ++j)
        for (int i = 0; i < bi; ++i) sub[i + size_t(bi) * j] = cm[i0 + i + size_t(nin) * j];
      std::vector<cd> one(nout, 1.0);
      if (int st = combine(c, ds, out.cols, sub, i0 > 0 ? &one : nullptr)) return st;
    }
    return 0;
  }

  std::vector<Dbl2 *> *store = nullptr;  // every buffer ever handed out (freed by the caller)

  // A free work vector of format fmt (0 complex128, 1 complex64, tagged): from
  // its pool, else a fresh allocation.  nullptr on OOM.
  Dbl2 *take(int fmt) {
    std::vector<Dbl2 *> &pl = fmt ? pool32 : pool;
    if (!pl.empty()) {
      Dbl2 *p = pl.back();
      pl.pop_back();
      return p;
    }
    Dbl2 *p = nullptr;
    if (hipMalloc(&p, size_t(c->d) * (fmt ? 8 : 16)) != hipSuccess) return nullptr;
    if (fmt) p = tag32(p);
    store->push_back(p);
    return p;
  }

  // Returns a work vector to the pool of its format.
  void give(Dbl2 *p) { (is32(p) ? pool32 : pool).push_back(p); }

  // Appends k work vectors of b's format to b, or returns DEX_EOOM.
  int grow(Block &b, int k) {
    for (int j = 0; j < k; ++j) {
      Dbl2 *p = take(b.fmt);
      if (!p) return fail(DEX_EOOM, "device allocation for GPLMR work vectors failed");
      b.cols.push_back(p);
    }
    return 0;
  }
};

// One GPLMR run on the block `xcols` (the caller's columns, then any guard
// columns); convergence and the returned values concern the first `nconv`
// columns, which the harmonic-Ritz selection keeps ordered nearest-first.
// lowp: the search blocks W, S, P, their images and the inner solve's basis are
// complex64 with FP32 matvecs (FP64 accumulation in every Gram and
// combination); X and AX stay complex128 and AX = H X is applied in FP64.
int gplmr_run(dex_ctx *c, const std::vector<Dbl2 *> &xcols, int nconv, cd *e_out, cd sigma, int ns,
              int maxit, double tol, int inner, int outmodulus, const dex_solve_opts &o, dex_result *res,
              bool lowp, std::vector<Dbl2 *> &store) {
  const int nev = int(xcols.size());
  Solver S{c, nev, ns, o.strict_cholesky != 0, o.symmetry};
  S.inner32 = o.inner_fp32 != 0 || lowp;
  const int lf = lowp ? 1 : 0;  // storage of the search blocks
  // columns: X in xcols; the rest from `store` (every buffer in it is free at entry)
  Block X, AX, W, AW, P, AP;
  W.fmt = AW.fmt = P.fmt = AP.fmt = lf;
  X.cols = xcols;
  for (Dbl2 *p : store) S.give(p);
  S.store = &store;
  if (int st = S.grow(AX, nev)) return st;
  if (int st = S.grow(W, nev)) return st;
  const double t_start = now_s();
  if (o.symmetry != 0) {
    // the exchange-symmetric matvec assumes X = sym X^T: project the start block
    for (int j = 0; j < nev; ++j) DEX_HIP(dex::launch_sym_project(X.cols[j], c->m, o.symmetry, c->stream));
  }
  // NORMALIZE(X) (gplmr.F90:65)
  {
    std::vector<double> aa, bb;
    std::vector<cd> ab;
    if (int st = coldots(c, descs_of(X), {}, aa, ab, bb)) return st;
    std::vector<cd> dm(size_t(nev) * nev, 0.0);
    for (int j = 0; j < nev; ++j) {
      if (!(aa[j] > 0)) return fail(DEX_EZEROBLOCK, "zero start vector");
      dm[j + size_t(nev) * j] = 1.0 / std::sqrt(aa[j]);
    }
    if (int st = combine(c, descs_of(X), X.cols, dm)) return st;
  }
  auto precblock = [&](Block &b) -> int {
    if (inner == 0 || b.size() == 0) return 0;
    if (inner == 1) {
      Block hb;
      hb.fmt = b.fmt;
      if (int st = S.grow(hb, b.size())) return st;
      if (int st = S.matblock(b, hb)) return st;
      for (Dbl2 *p : b.cols) S.give(p);
      b = hb;
      return 0;
    }
    const int grp = o.inner_block > 0 ? o.inner_block : b.size();
    for (int j0 = 0; j0 < b.size(); j0 += grp) {
      Block sub;
      sub.cols.assign(b.cols.begin() + j0, b.cols.begin() + std::min(b.size(), j0 + grp));
      if (int st = S.inner_solve(sub, inner, sigma)) return st;
    }
    return 0;
  };
  if (int st = S.matblock(X, AX)) return st;  // :68
  std::vector<cd> e(nev);
  std::vector<double> resn(nev);
  auto residual = [&]() -> int {
    // E = diag(X^H AX) of unit columns and W = AX - X E, ||W_j|| (:324-341):
    // X and AX are first scaled by 1/||X_j||.
    std::vector<double> aa, bb;
    std::vector<cd> ab;
    double t0 = now_s();
    if (int st = coldots(c, descs_of(X), descs_of(AX), aa, ab, bb)) return st;
    S.t_gram += now_s() - t0;
    t0 = now_s();
    // per column: X_j / n, AX_j / n and W_j = AX_j/n - E_j X_j/n with ||W_j||, one pass
    std::vector<double> sc(nev);
    std::vector<cd> ev(nev);
    for (int j = 0; j < nev; ++j) {
      sc[j] = 1.0 / std::sqrt(aa[j]);
      e[j] = ab[j] / aa[j];
      ev[j] = e[j];
    }
    if (int st = residual_cols(c, X.cols, AX.cols, W.cols, sc, ev, resn)) return st;
    S.t_update += now_s() - t0;
    return 0;
  };
  // First E and W: X is already unit, the scaling is 1.
  if (int st = residual()) return st;
  double maxres = *std::max_element(resn.begin(), resn.begin() + nconv);
  int iter = 1;
  Block Sblk, ASblk;
  Sblk.fmt = ASblk.fmt = lf;
  // With complex64 search blocks (lowp) AX is carried by a recurrence over FP32
  // images, which leaves it off H X by the FP32 rounding of the steps since AX
  // was last formed.  AX = H X is recomputed in FP64 every `refresh` iterations
  // (refresh_every, 10 by default in that form), once when the residual first
  // falls below fp32_tol, and whenever the residual falls below TOL: the run
  // ends only on a residual formed from an FP64 AX.
  const int refresh = o.refresh_every > 0 ? o.refresh_every : (lowp ? 10 : 0);
  bool fresh = true;         // AX was formed by a matvec since the last update
  bool crossed = false;      // the residual has fallen below fp32_tol once
  int refreshes = 0;
  auto refresh_ax = [&]() -> int {
    if (int st = S.matblock(X, AX)) return st;
    if (int st = residual()) return st;
    maxres = *std::max_element(resn.begin(), resn.begin() + nconv);
    fresh = true;
    ++refreshes;
    return 0;
  };
  while (iter <= maxit && (maxres > tol || !fresh) && c->stop.load(std::memory_order_relaxed) == 0) {
    if (lowp && maxres <= tol && !fresh) {
      // converged on the recurrence: check against an FP64 AX
      if (int st = refresh_ax()) return st;
      if (maxres <= tol) break;
    }
    if (refresh > 0 && iter > 1 && (iter - 1) % refresh == 0) {
      if (!lowp) {
        if (int st = S.matblock(X, AX)) return st;
      } else if (!fresh) {
        if (int st = refresh_ax()) return st;
      }
    }
    if (lowp && !crossed && maxres <= o.fp32_tol) {
      crossed = true;
      if (!fresh)
        if (int st = refresh_ax()) return st;
    }
    fresh = !lowp;  // in FP64 the recurrence is the reference's: its residual is accepted
    // active columns: all, or with lock only those still above TOL
    std::vector<int> active;
    for (int j = 0; j < nev; ++j)
      if (!o.lock || resn[j] > tol) active.push_back(j);
    {
      Block wa;
      wa.fmt = lf;
      std::vector<bool> keep(nev, false);
      for (int j : active) keep[j] = true;
      for (int j = 0; j < nev; ++j) {
        if (keep[j]) wa.cols.push_back(W.cols[j]);
        else S.give(W.cols[j]);
      }
      W = wa;
    }
    std::vector<cd> eact;
    for (int j : active) eact.push_back(e[j]);
    // CHOLQR(X); AX <- AX R^{-1}  (:88-90)
    {
      Block none, anone;
      if (int st = S.orth_against(none, X, &anone, &AX, false)) return st;
    }
    if (X.size() != nev) return fail(DEX_ECHOLESKY, "X lost rank");
    // PRECBLOCK (:92-99): GM_MAX_INNER_ITER = 1 gives W <- H W (:491-493),
    // larger values the inner minimal-residual solve
    if (int st = precblock(W)) return st;
    // W <- W - X (X^H W); CHOLQR(W)  (:101-105)
    if (int st = S.orth_against(X, W, nullptr, nullptr, true)) return st;
    AW.cols.clear();
    if (int st = S.grow(AW, W.size())) return st;
    if (int st = S.matblock(W, AW)) return st;  // :108
    // S blocks (:112-172)
    for (Dbl2 *p : Sblk.cols) S.give(p);
    for (Dbl2 *p : ASblk.cols) S.give(p);
    Sblk.cols.clear();
    ASblk.cols.clear();
    Block prev = W, aprev = AW;
    for (int l = 1; l <= ns; ++l) {
      Block sl;
      sl.fmt = lf;
      const int k = prev.size();
      if (int st = S.grow(sl, k)) return st;
      if (k > 0) {
        // S_l = A S_{l-1} - S_{l-1} E  (W for l = 1)
        std::vector<cd> cm(size_t(2 * k) * k, 0.0);
        for (int j = 0; j < k; ++j) {
          cm[j + size_t(2 * k) * j] = 1.0;
          cm[k + j + size_t(2 * k) * j] = -eact[j];
        }
        std::vector<ColDesc> in = descs_of(aprev);
        for (Dbl2 *p : prev.cols) in.push_back(plain(p));
        if (int st = combine(c, in, sl.cols, cm)) return st;
      }
      if (int st = precblock(sl)) return st;  // :118-126, :148-156
      {
        // X, W and the earlier S blocks are mutually orthonormal: one projection
        Block xws = X;
        xws.cols.insert(xws.cols.end(), W.cols.begin(), W.cols.end());
        xws.cols.insert(xws.cols.end(), Sblk.cols.begin(), Sblk.cols.end());
        if (int st = S.orth_against(xws, sl, nullptr, nullptr, true)) return st;
      }
      Block asl;
      asl.fmt = lf;
      if (int st = S.grow(asl, sl.size())) return st;
      if (int st = S.matblock(sl, asl)) return st;
      Sblk.cols.insert(Sblk.cols.end(), sl.cols.begin(), sl.cols.end());
      ASblk.cols.insert(ASblk.cols.end(), asl.cols.begin(), asl.cols.end());
      prev = sl;
      aprev = asl;
    }
    // P against X, W, S; CHOLQR(P); AP <- AP R^{-1}  (:177-196)
    if (iter > 1 && P.size() > 0) {
      Block q = X, aq = AX;
      q.cols.insert(q.cols.end(), W.cols.begin(), W.cols.end());
      q.cols.insert(q.cols.end(), Sblk.cols.begin(), Sblk.cols.end());
      aq.cols.insert(aq.cols.end(), AW.cols.begin(), AW.cols.end());
      aq.cols.insert(aq.cols.end(), ASblk.cols.begin(), ASblk.cols.end());
      if (int st = S.orth_against(q, P, &aq, &AP, true)) return st;
    }
    // Harmonic projection K = (AQ - sigma Q)^H (AQ - sigma Q), M = (AQ - sigma Q)^H Q (:201-277)
    Block Q, AQ;
    for (const Block *b : {&X, &W, &Sblk, &P}) Q.cols.insert(Q.cols.end(), b->cols.begin(), b->cols.end());
    for (const Block *b : {&AX, &AW, &ASblk, &AP})
      AQ.cols.insert(AQ.cols.end(), b->cols.begin(), b->cols.end());
    const int nq = Q.size();
    if (nq > 96) return fail(DEX_EINVAL, "projected basis larger than 96");
    std::vector<cd> km;
    double t0 = now_s();
    if (int st = gram_harmonic(c, Q.cols, AQ.cols, sigma, km)) return st;
    S.t_gram += now_s() - t0;
    t0 = now_s();
    std::vector<cd> kp(size_t(nq) * nq), mp(size_t(nq) * nq);
    for (int j = 0; j < nq; ++j)
      for (int i = 0; i < nq; ++i) {
        kp[i + size_t(nq) * j] = km[i + size_t(nq) * j];
        mp[i + size_t(nq) * j] = km[i + size_t(nq) * (nq + j)];
      }
    std::vector<cd> coords, mu;
    if (dex::harmonic_pencil(nq, kp.data(), mp.data(), nq, nev, coords, mu) != 0)
      return fail(DEX_EGEVP, "SOLVE OF THE PROJECTED EIGENVALUE PROBLEM FAILED");
    S.t_small += now_s() - t0;
    t0 = now_s();
    // P = W CW + S CS + P CP;  X = X CX + P  (and images) (:298-320).
    // The new P keeps the columns of the active set.
    const int nw = W.size(), nsb = Sblk.size(), np = P.size();
    const int nact = int(active.size());
    // inputs [X W S P], outputs [X, Pnew] with Pnew in the old P's storage.
    Block Pn, APn;
    Pn.fmt = APn.fmt = lf;
    for (int j = 0; j < np && j < nact; ++j) {
      Pn.cols.push_back(P.cols[j]);
      APn.cols.push_back(AP.cols[j]);
    }
    if (int st = S.grow(Pn, nact - Pn.size())) return st;
    if (int st = S.grow(APn, nact - APn.size())) return st;
    std::vector<cd> cm(size_t(nq) * (nev + nact), 0.0);
    for (int j = 0; j < nev; ++j)
      for (int i = 0; i < nq; ++i) cm[i + size_t(nq) * j] = coords[i + size_t(nq) * j];  // X_new
    for (int jj = 0; jj < nact; ++jj) {
      const int j = active[jj];
      for (int i = nev; i < nq; ++i) cm[i + size_t(nq) * (nev + jj)] = coords[i + size_t(nq) * j];  // P_new
    }
    {
      // One launch takes at most 64 inputs and 48 outputs (nev + nact <=
      // 2 (NEV + guard) <= 48 since (3 + NS)(NEV + guard) <= 96).  Wider Q goes
      // in chunks accumulated with beta = 1, X and P (whose storage the outputs
      // overwrite: 2 (NEV + guard) <= 48 columns) in the first, W and S after.
      std::vector<int> order;
      for (int i = 0; i < nev; ++i) order.push_back(i);
      for (int i = nev + nw + nsb; i < nq; ++i) order.push_back(i);
      for (int i = nev; i < nev + nw + nsb; ++i) order.push_back(i);
      std::vector<Dbl2 *> out = X.cols, aout = AX.cols;
      out.insert(out.end(), Pn.cols.begin(), Pn.cols.end());
      aout.insert(aout.end(), APn.cols.begin(), APn.cols.end());
      const int nout = int(out.size());
      std::vector<cd> one(nout, 1.0);
      for (int i0 = 0; i0 < nq; i0 += 64) {
        const int bi = std::min(64, nq - i0);
        std::vector<cd> sub(size_t(bi) * nout);
        for (int j = 0; j < nout; ++j)
          for (int i = 0; i < bi; ++i) sub[i + size_t(bi) * j] = cm[order[i0 + i] + size_t(nq) * j];
        std::vector<ColDesc> qd, aqd;
        for (int i = 0; i < bi; ++i) {
          qd.push_back(plain(Q.cols[order[i0 + i]]));
          aqd.push_back(plain(AQ.cols[order[i0 + i]]));
        }
        if (split_update() && nout > 16 && nev <= 16 && nout - nev <= 16) {
          // Two launches of at most 16 outputs on the LDS-free kernel in place of one
          // launch of the LDS kernel.  X_new first (in place; it reads the old P),
          // then P_new (in the old P's storage), which reads no X column.
          const std::vector<Dbl2 *> outx(out.begin(), out.begin() + nev), outp(out.begin() + nev, out.end());
          const std::vector<Dbl2 *> aoutx(aout.begin(), aout.begin() + nev), aoutp(aout.begin() + nev, aout.end());
          const std::vector<cd> subx(sub.begin(), sub.begin() + size_t(bi) * nev);
          std::vector<ColDesc> qdp, aqdp;
          std::vector<int> rows_p;
          for (int i = 0; i < bi; ++i)
            if (order[i0 + i] >= nev) {
              rows_p.push_back(i);
              qdp.push_back(qd[i]);
              aqdp.push_back(aqd[i]);
            }
          const int bp = int(rows_p.size()), npo = nout - nev;
          std::vector<cd> subp(size_t(bp) * npo);
          for (int j = 0; j < npo; ++j)
            for (int k = 0; k < bp; ++k) subp[k + size_t(bp) * j] = sub[rows_p[k] + size_t(bi) * (nev + j)];
          std::vector<cd> onex(nev, 1.0), onep(npo, 1.0);
          if (int st = combine(c, qd, outx, subx, i0 > 0 ? &onex : nullptr)) return st;
          if (int st = combine(c, aqd, aoutx, subx, i0 > 0 ? &onex : nullptr)) return st;
          if (bp > 0) {
            if (int st = combine(c, qdp, outp, subp, i0 > 0 ? &onep : nullptr)) return st;
            if (int st = combine(c, aqdp, aoutp, subp, i0 > 0 ? &onep : nullptr)) return st;
          } else if (i0 == 0) {
            return fail(DEX_EINVAL, "update: no input for the new P in the first chunk");
          }
          continue;
        }
        if (int st = combine(c, qd, out, sub, i0 > 0 ? &one : nullptr)) return st;
        if (int st = combine(c, aqd, aout, sub, i0 > 0 ? &one : nullptr)) return st;
      }
    }
    for (int j = nact; j < np; ++j) {
      S.give(P.cols[j]);
      S.give(AP.cols[j]);
    }
    P = Pn;
    AP = APn;
    DEX_HIP(hipStreamSynchronize(c->stream));
    S.t_update += now_s() - t0;
    // free W and AW columns (W is rebuilt as the residual in place of the old W storage)
    for (Dbl2 *p : AW.cols) S.give(p);
    AW.cols.clear();
    // W gets nev columns again
    if (int st = S.grow(W, nev - W.size())) return st;
    if (int st = residual()) return st;
    maxres = *std::max_element(resn.begin(), resn.begin() + nconv);
    if (o.verbose == 2 && (maxres <= tol || (outmodulus > 0 && iter % outmodulus == 0))) {
      int nc = 0;
      for (int j = 0; j < nconv; ++j) nc += resn[j] <= tol;
      std::fprintf(stderr, "GPLMR NN=%d block %d%s: iteration %d, %d/%d below %.1e, max residual %.2e\n", c->nn,
                   nconv, lowp ? " (complex64 search blocks)" : "", iter, nc, nconv, tol, maxres);
      std::fflush(stderr);
    } else if (o.verbose && (maxres <= tol || (outmodulus > 0 && iter % outmodulus == 0))) {
      std::printf("ITERATION %5d\n", iter);
      for (int j = 0; j < nev; ++j)
        std::printf("EIG %4d = %20.12f%20.12f,  ERROR = %10.2E\n", j + 1, e[j].real(), e[j].imag(), resn[j]);
      std::fflush(stdout);
    }
    ++iter;
  }
  if (lowp && !fresh) {
    // the iteration budget ran out on a recurrence residual: report the true one
    if (int st = refresh_ax()) return st;
  }
  for (int j = 0; j < nconv; ++j) e_out[j] = e[j];
  if (res) {
    res->iterations = iter - 1;
    res->converged = maxres <= tol;
    res->fp64_restarts_used += refreshes;
    res->matvecs += S.matvecs;
    res->max_residual = maxres;
    res->t_total += now_s() - t_start;
    res->t_matvec += S.t_matvec;
    res->t_gram += S.t_gram;
    res->t_update += S.t_update;
    res->t_smalleig += S.t_small;
    res->t_prec += S.t_prec;
    res->fallbacks += S.fallbacks;
  }
  return maxres <= tol ? DEX_OK : DEX_ENOTCONVERGED;
}

}  // namespace

// Columns per batched matvec launch for vectors of d entries (batch_cols).
int64_t batch_cols_d(int64_t d) {
  const int64_t b = (int64_t(512) << 20) / (d * 16);
  return b >= 2 ? std::min<int64_t>(b, dex::kBatchMax) : 1;
}

// Device bytes a context of grid nn holds apart from any solve: the tables,
// the block kernels' descriptor, coefficient and partial-sum buffers (as
// dex_ctx_create allocates them), and the matvec intermediates (U, and on small
// grids those of the batched launches), which a solve's first matvec allocates.
int64_t ctx_fixed_bytes(int nn) {
  const int64_t m = int64_t(nn) * nn * nn, d = m * m, vec = d * 16;
  const int64_t L = 2 * nn + 1;
  const int64_t partial = int64_t(std::max(dex::gram_partial_elems(64, 128) * sizeof(Dbl2),
                                           dex::coldots_partial_elems(64) * sizeof(double)));
  const int64_t bc = batch_cols_d(d);
  return (L * L * L + m) * 8 + int64_t(kMaxDesc) * int64_t(sizeof(ColDesc)) +
         (int64_t(kMaxDesc) * int64_t(sizeof(ColDesc)) + 2 * kMaxSmall * 16 + 512) + 192 * 192 * 16 + partial +
         64 * 4 * 8 + vec + (bc >= 2 ? bc * vec : 0);
}

// Bytes of a solve (nev, ns, inner_iter, opts) on a context of grid nn and the
// given precision, the context's fixed buffers included.  The caller's block X
// (NEV) and the guard columns, AX, the solver's (5 + 2 NS) (NEV + guard)
// search-block vectors (W, S, P, their images and the new P), and for an inner
// solve of order m on groups of g columns its Arnoldi basis ((m + 1) g) and the
// next block (g).  X and AX are complex128; the search blocks and the inner
// basis are complex64 in the mixed form (FP32 context), the inner basis also
// with inner_fp32.
int64_t solve_footprint(int nn, int precision, int nev, int ns, int inner_iter, const dex_solve_opts *opts) {
  const int64_t d = int64_t(nn) * nn * nn * nn * nn * nn;
  const int64_t vec = d * 16;
  const bool mixed = precision == DEX_FP32;
  const int64_t lo = mixed ? d * 8 : vec;
  const int64_t guard = opts && opts->guard > 0 ? opts->guard : 0;
  const int64_t nblk = nev + guard;
  const int64_t ib = opts ? opts->inner_block : 0;
  const int64_t grp = ib > 0 && ib < nblk ? ib : nblk;
  const int64_t ivec = (mixed || (opts && opts->inner_fp32 != 0)) ? d * 8 : vec;
  const int64_t inner = inner_iter >= 2 ? int64_t(inner_iter + 2) * grp * ivec : (inner_iter == 1 ? nblk * lo : 0);
  return 2 * nblk * vec + int64_t(5 + 2 * ns) * nblk * lo + inner + ctx_fixed_bytes(nn);
}

extern "C" int64_t dex_gplmr_footprint(const dex_ctx *c, int32_t nev, int32_t ns, int32_t inner_iter,
                                       const dex_solve_opts *opts) {
  if (!c) return -1;
  return solve_footprint(c->nn, c->p.precision, nev, ns, inner_iter, opts);
}

extern "C" int64_t dex_solve_footprint(const dex_params *p, int32_t nev, int32_t ns, int32_t inner_iter,
                                       const dex_solve_opts *opts) {
  if (!p || p->n < 1 || nev < 1 || ns < 1 || inner_iter < 0) return -1;
  return solve_footprint(2 * p->n + 1, p->precision, nev, ns, inner_iter, opts);
}

extern "C" int dex_gplmr_eig_device(dex_ctx *c, void *x_dev, double *e, int32_t m, int32_t mb,
                                    double sigma_re, double sigma_im, int32_t nev, int32_t ns,
                                    int32_t maxit, double tol, int32_t inner_iter, int32_t outmodulus,
                                    const dex_solve_opts *opts, dex_result *res) {
  if (!c || !x_dev || !e) return fail(DEX_EINVAL, "null argument");
  if (int64_t(m) * mb != c->d) return fail(DEX_ESIZE, "SIZE ERROR: M*MB != NN^6");
  if (opts && opts->inner_fp32 != 0 && !c->fast32)
    return fail(DEX_EUNSUPPORTED, "inner_fp32 needs the FP32 matvec kernels for this grid");
  const int guard = opts ? opts->guard : 0;
  if (guard < 0) return fail(DEX_EINVAL, "guard must be >= 0");
  if (nev < 1 || ns < 1 || ns * (nev + guard) > 48 || (3 + ns) * (nev + guard) > 96)
    return fail(DEX_EINVAL, "need NEV >= 1, NS >= 1, NS*(NEV+guard) <= 48, (3+NS)*(NEV+guard) <= 96");
  if (inner_iter < 0) return fail(DEX_EINVAL, "GM_MAX_INNER_ITER must be >= 0");
  if (opts && (opts->symmetry < -1 || opts->symmetry > 1)) return fail(DEX_EINVAL, "symmetry must be -1, 0 or 1");
  if (opts && opts->symmetry != 0 && !c->fast64 && !(c->p.precision == DEX_FP32 && c->fast32))
    return fail(DEX_EUNSUPPORTED, "exchange-symmetric matvec needs the LDS kernels for this grid");
  DEX_HIP(hipSetDevice(c->p.device));
  dex_solve_opts o;
  dex_solve_opts_default(&o);
  if (opts) o = *opts;
  if (res) std::memset(res, 0, sizeof(*res));
  const int nblk = nev + guard;
  const bool mixed = c->p.precision == DEX_FP32;
  // the work pool plus headroom for the new P block: AX (complex128) and
  // (5 + 2 NS) (NEV + guard) search-block vectors, complex64 in the mixed form
  const int nwork = (5 + 2 * ns) * nblk + nblk;
  std::vector<Dbl2 *> store, guards;
  auto free_all = [&]() {
    for (Dbl2 *q : store) (void)hipFree(untag(q));
    for (Dbl2 *q : guards) (void)hipFree(q);
  };
  for (int j = 0; j < nwork + guard; ++j) {
    const bool lo = mixed && j >= nblk && j < nwork;
    Dbl2 *p = nullptr;
    hipError_t err = hipMalloc(&p, size_t(c->d) * (lo ? 8 : 16));
    if (err != hipSuccess) {
      free_all();
      return fail(DEX_EOOM, "device allocation for GPLMR work vectors failed");
    }
    if (j < nwork)
      store.push_back(lo ? tag32(p) : p);
    else
      guards.push_back(p);
  }
  // the block: the caller's NEV columns, then the guard columns (seeded
  // complex Gaussians with the start block's exchange symmetry)
  std::vector<Dbl2 *> xcols;
  for (int j = 0; j < nev; ++j) xcols.push_back(static_cast<Dbl2 *>(x_dev) + size_t(j) * c->d);
  for (int j = 0; j < guard; ++j) {
    DEX_HIP(dex::launch_fill_gaussian(guards[j], c->d, 0x67756172ull, uint32_t(j), 0xFFFFFFFFu, c->stream));
    if (o.symmetry != 0) DEX_HIP(dex::launch_sym_project(guards[j], c->m, o.symmetry, c->stream));
    xcols.push_back(guards[j]);
  }
  const cd sigma(sigma_re, sigma_im);
  std::vector<cd> ev(nev);
  int st;
  // Mixed form (FP32 contexts): complex64 search blocks with FP32 matvecs
  // around the complex128 X and AX, whose FP64 refreshes gplmr_run does; a run
  // that ends at MAXIT unconverged may be followed by up to fp64_restarts
  // restarts from its X (each with a fresh FP64 AX and no P block).
  st = gplmr_run(c, xcols, nev, ev.data(), sigma, ns, maxit, tol, inner_iter, outmodulus, o, res, mixed, store);
  if (mixed && st == DEX_ENOTCONVERGED) {
    for (int r = 0; r < o.fp64_restarts && c->stop.load(std::memory_order_relaxed) == 0; ++r) {
      dex_result sub{};
      st = gplmr_run(c, xcols, nev, ev.data(), sigma, ns, maxit, tol, inner_iter, outmodulus, o,
                     res ? &sub : nullptr, true, store);
      if (res) {
        res->matvecs += sub.matvecs;
        res->iterations += sub.iterations;
        res->t_total += sub.t_total;
        res->t_matvec += sub.t_matvec;
        res->t_gram += sub.t_gram;
        res->t_update += sub.t_update;
        res->t_smalleig += sub.t_smalleig;
        res->t_prec += sub.t_prec;
        res->fallbacks += sub.fallbacks;
        res->max_residual = sub.max_residual;
        res->converged = sub.converged;
        res->fp64_restarts_used += sub.fp64_restarts_used;
      }
      if (st != DEX_ENOTCONVERGED) break;
    }
  }
  free_all();
  for (int j = 0; j < nev; ++j) {
    e[2 * j] = ev[j].real();
    e[2 * j + 1] = ev[j].imag();
  }
  return st;
}

extern "C" int dex_precblock_device(dex_ctx *c, const void *x_dev, void *px_dev, int32_t m, int32_t mb,
                                    int32_t nev, double sigma_re, double sigma_im, int32_t inner_iter,
                                    int32_t symmetry) {
  if (!c || !x_dev || !px_dev) return fail(DEX_EINVAL, "null argument");
  if (int64_t(m) * mb != c->d) return fail(DEX_ESIZE, "SIZE ERROR: M*MB != NN^6");
  if (nev < 1 || nev > 48) return fail(DEX_EINVAL, "need 1 <= NEV <= 48");
  if (inner_iter < 1) return fail(DEX_EINVAL, "PRECBLOCK needs GM_MAX_INNER_ITER >= 1");  // gplmr.F90:487-489
  if (symmetry < -1 || symmetry > 1) return fail(DEX_EINVAL, "symmetry must be -1, 0 or 1");
  if (symmetry != 0 && !c->fast64) return fail(DEX_EUNSUPPORTED, "exchange-symmetric matvec needs the LDS kernels for this grid");
  if (x_dev == px_dev) return fail(DEX_EINVAL, "X and PX must not alias");
  DEX_HIP(hipSetDevice(c->p.device));
  Solver S{c, nev, 1, false, false, symmetry};
  std::vector<Dbl2 *> store;
  S.store = &store;
  Block in, out;
  for (int j = 0; j < nev; ++j) {
    in.cols.push_back(const_cast<Dbl2 *>(static_cast<const Dbl2 *>(x_dev)) + size_t(j) * c->d);
    out.cols.push_back(static_cast<Dbl2 *>(px_dev) + size_t(j) * c->d);
  }
  int st;
  if (inner_iter == 1) {
    st = S.matblock(in, out);
  } else {
    st = 0;
    const size_t bytes = size_t(c->d) * nev * sizeof(Dbl2);
    if (hipMemcpyAsync(px_dev, x_dev, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      st = fail(DEX_EHIP, "device copy failed");
    if (!st) st = S.inner_solve(out, inner_iter, cd(sigma_re, sigma_im));
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = fail(DEX_EHIP, "stream synchronise failed");
  for (Dbl2 *p : store) (void)hipFree(p);
  return st;
}

// HELIUM_ORBITALGATHER (hammod.F90:1080-1113) orthonormalisation, then
// DIAGSUBSPACE (hammod.F90:1125-1256) and GET_EXPECTS (hammod.F90:1261-1301).
extern "C" int dex_diagsubspace_device(dex_ctx *c, const double *orbitals, int32_t numorbs, int32_t singlet,
                                       int32_t hermopt, double etarget_re, double etarget_im, int32_t numout,
                                       void *psi_dev, double *e0, double *expects) {
  if (!c || !orbitals || !psi_dev || !e0 || !expects) return fail(DEX_EINVAL, "null argument");
  if (numorbs < 1 || numorbs > 32) return fail(DEX_EINVAL, "numorbs must be 1..32");
  const int k = numorbs;
  const int64_t m = c->m;
  const int sgn = singlet ? 1 : -1;
  // the orbital-product basis: singlet a <= b, triplet a < b (:1170-1180)
  std::vector<std::pair<int, int>> pairs;
  for (int a = 0; a < k; ++a)
    for (int b = a + (singlet ? 0 : 1); b < k; ++b) pairs.emplace_back(a, b);
  const int nb = int(pairs.size());
  if (numout < 1 || numout > nb) return fail(DEX_EINVAL, "numout must be 1..(size of the orbital-product basis)");
  if (!(c->fast64)) return fail(DEX_EUNSUPPORTED, "needs the LDS matvec kernels for this grid");
  DEX_HIP(hipSetDevice(c->p.device));
  // orthonormalise the orbitals: bilinear (hermopt = 0) or Hermitian products,
  // one Gram-Schmidt pass as HELIUM_ORBITALGATHER
  std::vector<cd> x(size_t(m) * k);
  for (size_t i = 0; i < x.size(); ++i) x[i] = cd(orbitals[2 * i], orbitals[2 * i + 1]);
  auto dot = [&](int a, int b) {
    cd sacc = 0.0;
    for (int64_t p = 0; p < m; ++p) {
      const cd xa = x[p + m * a];
      sacc += (hermopt ? std::conj(xa) : xa) * x[p + m * b];
    }
    return sacc;
  };
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < i; ++j) {
      const cd pr = dot(j, i);
      for (int64_t p = 0; p < m; ++p) x[p + m * i] -= x[p + m * j] * pr;
    }
    const cd nrm = std::sqrt(dot(i, i));
    if (std::abs(nrm) == 0.0) return fail(DEX_EZEROBLOCK, "zero orbital");
    for (int64_t p = 0; p < m; ++p) x[p + m * i] /= nrm;
  }
  // device buffers: X, conj-or-plain X for the Y X products, T, W, one vector
  Dbl2 *dx = nullptr, *dt = nullptr, *dw = nullptr, *dpsi = nullptr, *dy = nullptr;
  auto cleanup = [&]() {
    for (void *ptr : {(void *)dx, (void *)dt, (void *)dw, (void *)dpsi, (void *)dy})
      if (ptr) (void)hipFree(ptr);
  };
  const size_t mk = size_t(m) * k * sizeof(Dbl2);
  if (hipMalloc(&dx, mk) != hipSuccess || hipMalloc(&dt, mk) != hipSuccess || hipMalloc(&dw, mk) != hipSuccess ||
      hipMalloc(&dpsi, size_t(c->d) * sizeof(Dbl2)) != hipSuccess ||
      hipMalloc(&dy, size_t(c->d) * sizeof(Dbl2)) != hipSuccess) {
    cleanup();
    return fail(DEX_EOOM, "device allocation for the orbital subspace failed");
  }
  int st = DEX_OK;
  auto hip = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && st == DEX_OK) st = fail(DEX_EHIP, what);
    return st == DEX_OK;
  };
  hip(hipMemcpyAsync(dx, x.data(), mk, hipMemcpyHostToDevice, c->stream), "upload orbitals");
  // coefficient block S (k x k) of a combination sum_I c_I psi_I -> T = X S
  auto build = [&](const std::vector<cd> &cvec, Dbl2 *out) {
    std::vector<cd> sm(size_t(k) * k, 0.0);
    for (int i = 0; i < nb; ++i) {
      const int a = pairs[i].first, b = pairs[i].second;
      const double n = a == b ? 0.5 : 1.0 / std::sqrt(2.0);
      sm[a + size_t(k) * b] += cvec[i] * n;
      sm[b + size_t(k) * a] += double(sgn) * cvec[i] * n;
    }
    std::vector<cd> t(size_t(m) * k, 0.0);
    for (int b = 0; b < k; ++b)
      for (int a = 0; a < k; ++a) {
        const cd sab = sm[a + size_t(k) * b];
        if (sab == cd(0.0)) continue;
        for (int64_t p = 0; p < m; ++p) t[p + m * b] += x[p + m * a] * sab;
      }
    if (!hip(hipMemcpyAsync(dt, t.data(), mk, hipMemcpyHostToDevice, c->stream), "upload T")) return false;
    if (!hip(dex::launch_lowrank(dt, dx, out, m, k, c->stream), "lowrank")) return false;
    return hip(hipStreamSynchronize(c->stream), "sync");  // t is a host temporary
  };
  // A(I, J) = psi_I . H psi_J (bilinear) or psi_I^H H psi_J (:1196-1220)
  std::vector<cd> amat(size_t(nb) * nb);
  std::vector<cd> wh(size_t(m) * k);
  for (int jj = 0; jj < nb && st == DEX_OK; ++jj) {
    std::vector<cd> cvec(nb, 0.0);
    cvec[jj] = 1.0;
    if (!build(cvec, dpsi)) break;
    if (int e = apply_h(c, dpsi, dy, false, sgn)) {
      st = e;
      break;
    }
    if (!hip(dex::launch_ywx(dy, dx, dw, m, k, hermopt ? 1 : 0, c->stream), "ywx")) break;
    if (!hip(hipMemcpyAsync(wh.data(), dw, mk, hipMemcpyDeviceToHost, c->stream), "download W")) break;
    if (!hip(hipStreamSynchronize(c->stream), "sync")) break;
    // G(a, b) = x_a . W_b (conjugated x_a for the Hermitian product)
    std::vector<cd> g(size_t(k) * k, 0.0);
    for (int b = 0; b < k; ++b)
      for (int a = 0; a < k; ++a) {
        cd sacc = 0.0;
        for (int64_t p = 0; p < m; ++p) {
          const cd xa = x[p + m * a];
          sacc += (hermopt ? std::conj(xa) : xa) * wh[p + m * b];
        }
        g[a + size_t(k) * b] = sacc;
      }
    for (int ii = 0; ii < nb; ++ii) {
      const int a = pairs[ii].first, b = pairs[ii].second;
      const double n = a == b ? 0.5 : 1.0 / std::sqrt(2.0);
      amat[ii + size_t(nb) * jj] = n * (g[a + size_t(k) * b] + double(sgn) * g[b + size_t(k) * a]);
    }
  }
  std::vector<cd> lam, vec;
  if (st == DEX_OK && dex::zgeev_right(nb, amat.data(), nb, lam, vec) != 0)
    st = fail(DEX_EGEVP, "subspace eigenproblem failed");
  if (st == DEX_OK) {
    // mysort (hammod.F90:1658-1690): ascending |E0 - etarget|, the last index on ties
    const cd et(etarget_re, etarget_im);
    std::vector<int> order, taken(nb, 0);
    for (int j = 0; j < nb; ++j) {
      int best = -1;
      double low = 0.0;
      for (int i = 0; i < nb; ++i) {
        if (taken[i]) continue;
        const double dv = std::abs(lam[i] - et);
        if (best < 0 || dv <= low) {
          best = i;
          low = dv;
        }
      }
      taken[best] = 1;
      order.push_back(best);
    }
    for (int o = 0; o < numout && st == DEX_OK; ++o) {
      const int col = order[o];
      std::vector<cd> cvec(vec.begin() + size_t(nb) * col, vec.begin() + size_t(nb) * (col + 1));
      Dbl2 *out = static_cast<Dbl2 *>(psi_dev) + size_t(o) * c->d;
      if (!build(cvec, out)) break;
      // Hermitian normalisation (:1250-1252) and GET_EXPECTS with the Hermitian product
      std::vector<double> aa, bb;
      std::vector<cd> ab;
      if (int e = coldots(c, {plain(out)}, {}, aa, ab, bb)) {
        st = e;
        break;
      }
      std::vector<cd> sc(1, 1.0 / std::sqrt(aa[0]));
      if (int e = combine(c, {plain(out)}, {out}, sc)) {
        st = e;
        break;
      }
      if (int e = apply_h(c, out, dy, false, sgn)) {
        st = e;
        break;
      }
      if (int e = coldots(c, {plain(out)}, {plain(dy)}, aa, ab, bb)) {
        st = e;
        break;
      }
      e0[2 * o] = lam[col].real();
      e0[2 * o + 1] = lam[col].imag();
      const cd ex = ab[0] / aa[0];
      expects[2 * o] = ex.real();
      expects[2 * o + 1] = ex.imag();
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess && st == DEX_OK) st = fail(DEX_EHIP, "sync");
  cleanup();
  return st;
}

extern "C" int dex_gplmr_eig(dex_ctx *c, double *x, double *e, int32_t m, int32_t mb, double sigma_re,
                             double sigma_im, int32_t nev, int32_t ns, int32_t maxit, double tol,
                             int32_t inner_iter, int32_t outmodulus, const dex_solve_opts *opts,
                             dex_result *res) {
  if (!c || !x || !e) return fail(DEX_EINVAL, "null argument");
  if (int64_t(m) * mb != c->d) return fail(DEX_ESIZE, "SIZE ERROR: M*MB != NN^6");
  if (nev < 1) return fail(DEX_EINVAL, "NEV < 1");
  DEX_HIP(hipSetDevice(c->p.device));
  const size_t bytes = size_t(c->d) * nev * sizeof(Dbl2);
  void *dx = nullptr;
  DEX_HIP(hipMalloc(&dx, bytes));
  hipError_t err = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    (void)hipFree(dx);
    return hip_fail(err, "copy in");
  }
  int st = dex_gplmr_eig_device(c, dx, e, m, mb, sigma_re, sigma_im, nev, ns, maxit, tol, inner_iter,
                                outmodulus, opts, res);
  err = hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(dx);
  if (err != hipSuccess && st == DEX_OK) return hip_fail(err, "copy out");
  return st;
}

extern "C" int dex_heplus_lanczos_device(dex_ctx *c, int32_t blocksize, int32_t kdim, int32_t check, double tol,
                                         int32_t numeigs, uint64_t seed, double *eigenvalues, double *orbitals,
                                         double *residuals, int32_t *iterations) {
  if (!c || !eigenvalues) return fail(DEX_EINVAL, "null argument");
  DEX_HIP(hipSetDevice(c->p.device));
  std::vector<cd> eig, vecs;
  std::vector<double> resid;
  int iters = 0;
  const int st = dex::heplus_lanczos(c->stream, c->nn, c->d_vn, c->coef, blocksize, kdim, check, tol, numeigs,
                                     seed, eig, vecs, resid, iters);
  if (st != DEX_OK && st != DEX_ENOTCONVERGED) return st;
  for (int k = 0; k < numeigs && k < int(eig.size()); ++k) {
    eigenvalues[2 * k] = eig[k].real();
    eigenvalues[2 * k + 1] = eig[k].imag();
    if (residuals) residuals[k] = resid[k];
  }
  if (orbitals && !vecs.empty()) std::memcpy(orbitals, vecs.data(), vecs.size() * sizeof(cd));
  if (iterations) *iterations = iters;
  return st;
}

extern "C" int dex_residuals_device(dex_ctx *c, const void *x_dev, const double *e, int32_t nev,
                                    double *res_out) {
  if (!c || !x_dev || !e || !res_out) return fail(DEX_EINVAL, "null argument");
  DEX_HIP(hipSetDevice(c->p.device));
  Dbl2 *ax = nullptr;
  DEX_HIP(hipMalloc(&ax, size_t(c->d) * sizeof(Dbl2)));
  int st = DEX_OK;
  for (int j = 0; j < nev && st == DEX_OK; ++j) {
    const Dbl2 *xj = static_cast<const Dbl2 *>(x_dev) + size_t(j) * c->d;
    st = apply_h(c, xj, ax, false);
    if (st) break;
    std::vector<double> aa, bb;
    std::vector<cd> ab;
    // r = AX - E X: coldots of (AX - E X) via the sigma descriptor
    std::vector<ColDesc> a{shifted(ax, xj, cd(e[2 * j], e[2 * j + 1]))};
    std::vector<ColDesc> b{plain(xj)};
    st = coldots(c, a, b, aa, ab, bb);
    if (st) break;
    res_out[j] = std::sqrt(aa[0]) / std::sqrt(bb[0]);
  }
  (void)hipFree(ax);
  return st;
}

extern "C" int dex_fill_gaussian_device(dex_ctx *c, void *x_dev, int32_t ncols, uint64_t seed,
                                        uint32_t stream_id, int32_t symmetry) {
  if (!c || !x_dev) return fail(DEX_EINVAL, "null argument");
  DEX_HIP(hipSetDevice(c->p.device));
  for (int j = 0; j < ncols; ++j) {
    Dbl2 *xj = static_cast<Dbl2 *>(x_dev) + size_t(j) * c->d;
    DEX_HIP(dex::launch_fill_gaussian(xj, c->d, seed, uint32_t(j), stream_id, c->stream));
    if (symmetry != 0) DEX_HIP(dex::launch_sym_project(xj, c->m, symmetry > 0 ? 1 : -1, c->stream));
  }
  DEX_HIP(hipStreamSynchronize(c->stream));
  return DEX_OK;
}

extern "C" int dex_sync(dex_ctx *c) {
  if (!c) return fail(DEX_EINVAL, "null argument");
  DEX_HIP(hipStreamSynchronize(c->stream));
  return DEX_OK;
}

extern "C" int dex_host_kematrix(int32_t nn, double delta, double *t) {
  if (nn < 1 || !(delta > 0) || !t) return fail(DEX_EINVAL, "invalid argument");
  dex::kematrix(nn, delta, t);
  return DEX_OK;
}

extern "C" int dex_host_getinverse(int32_t nout, double delta, int32_t nsmall, int32_t tfactor,
                                   double *tinv) {
  if (nout < 0 || !(delta > 0) || !tinv) return fail(DEX_EINVAL, "invalid argument");
  std::vector<double> t;
  if (dex::getinverse(nout, delta, nsmall, tfactor, t) != 0) return fail(DEX_EINVAL, "nbig < nout");
  std::memcpy(tinv, t.data(), t.size() * sizeof(double));
  return DEX_OK;
}

extern "C" int dex_host_solve_gevp(int32_t k, const double *kp, const double *mp, int32_t ld,
                                   int32_t nev, double *coords, double *mu) {
  if (k < 1 || nev < 1 || nev > k || ld < k || !kp || !mp || !coords || !mu)
    return fail(DEX_EINVAL, "invalid argument");
  std::vector<cd> cv, mv;
  const int st = dex::harmonic_pencil(k, reinterpret_cast<const cd *>(kp), reinterpret_cast<const cd *>(mp), ld,
                                      nev, cv, mv);
  if (st != 0) return fail(DEX_EGEVP, "SOLVE OF THE PROJECTED EIGENVALUE PROBLEM FAILED");
  std::memcpy(coords, cv.data(), cv.size() * sizeof(cd));
  std::memcpy(mu, mv.data(), mv.size() * sizeof(cd));
  return DEX_OK;
}
