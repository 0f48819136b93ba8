// Eigenvector files in the reference's format: READWRITE_EIGENVECTORS and
// SIXD_FIX (utils.F90:82-198).
//
// A Fortran unformatted sequential file.  Record 1 holds M (int32), NUM (int32)
// and SIZEOF(ALLPSI) (the byte count of one M x M complex128 vector, int64 as
// gfortran writes it; readers that declared it int32 (utils.F90:95) see a
// 12-byte prefix of the same record, which wraps at 2 GiB).  Records 2..NUM+1
// hold one vector each, ALLPSI(M, M) column-major (electron 1 fastest).
// Record framing follows gfortran: a 4-byte little-endian length before and
// after each subrecord; records longer than 2147483639 bytes are split into
// subrecords whose leading marker is negated when the record continues and
// whose trailing marker is negated when it continued from a previous one.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dex.h"
#include "dex_internal.h"

namespace {

constexpr int64_t kMaxSubrecord = 2147483639;

struct File {
  std::FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put32(std::FILE *f, int32_t v) { return std::fwrite(&v, 4, 1, f) == 1; }

bool write_record(std::FILE *f, const void *data, int64_t n) {
  const char *p = static_cast<const char *>(data);
  int64_t left = n;
  bool first = true;
  do {
    const int64_t len = left > kMaxSubrecord ? kMaxSubrecord : left;
    const bool more = left > len;
    if (!put32(f, int32_t(more ? -len : len))) return false;
    if (len > 0 && std::fwrite(p, 1, size_t(len), f) != size_t(len)) return false;
    if (!put32(f, int32_t(first ? len : -len))) return false;
    p += len;
    left -= len;
    first = false;
  } while (left > 0);
  return true;
}

// Reads one whole record into `out` (resized); false on a framing error.
bool read_record(std::FILE *f, std::vector<char> &out) {
  out.clear();
  for (;;) {
    int32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, f) != 1) return false;
    const bool more = head < 0;
    const int64_t len = more ? -int64_t(head) : int64_t(head);
    const size_t at = out.size();
    out.resize(at + size_t(len));
    if (len > 0 && std::fread(out.data() + at, 1, size_t(len), f) != size_t(len)) return false;
    if (std::fread(&tail, 4, 1, f) != 1) return false;
    if ((tail < 0 ? -int64_t(tail) : int64_t(tail)) != len) return false;
    if (!more) return true;
  }
}

int cube_root_odd(int64_t m) {
  const int r = int(std::floor(std::cbrt(double(m)) + 0.5));
  return (int64_t(r) * r * r == m && r % 2 == 1) ? r : -1;
}

// SIXD_FIX (utils.F90:184-198): the central [-nmin, nmin]^6 block of a
// vector on the (2 nin + 1)^6 grid placed on the (2 nout + 1)^6 grid, zero
// elsewhere.
void sixd_fix(const double *in, int nin, double *out, int nout) {
  const int ni = 2 * nin + 1, no = 2 * nout + 1, nmin = nin < nout ? nin : nout;
  const int64_t mo = int64_t(no) * no * no;
  std::memset(out, 0, size_t(mo * mo) * 16);
  int64_t si[6], so[6];
  si[0] = so[0] = 1;
  for (int k = 1; k < 6; ++k) {
    si[k] = si[k - 1] * ni;
    so[k] = so[k - 1] * no;
  }
  const int w = 2 * nmin + 1;
  int idx[6] = {0, 0, 0, 0, 0, 0};
  for (;;) {
    int64_t pi = 0, po = 0;
    for (int k = 0; k < 6; ++k) {
      pi += (idx[k] - nmin + nin) * si[k];
      po += (idx[k] - nmin + nout) * so[k];
    }
    out[2 * po] = in[2 * pi];
    out[2 * po + 1] = in[2 * pi + 1];
    int k = 0;
    while (k < 6 && ++idx[k] == w) idx[k++] = 0;
    if (k == 6) break;
  }
}

}  // namespace

extern "C" int dex_write_eigvecs(const char *path, const double *psi, int32_t m, int32_t num) {
  if (!path || (!psi && num > 0) || m < 1 || num < 0) return dex::set_error(DEX_EINVAL, "bad argument");
  File fh;
  fh.f = std::fopen(path, "wb");
  if (!fh.f) return dex::set_error(DEX_EINVAL, "cannot open the eigenvector file for writing");
  const int64_t bytes = int64_t(m) * m * 16;
  char head[16];
  std::memcpy(head, &m, 4);
  std::memcpy(head + 4, &num, 4);
  std::memcpy(head + 8, &bytes, 8);
  if (!write_record(fh.f, head, 16)) return dex::set_error(DEX_EINVAL, "write failed");
  for (int32_t v = 0; v < num; ++v)
    if (!write_record(fh.f, psi + size_t(v) * size_t(bytes / 8), bytes))
      return dex::set_error(DEX_EINVAL, "write failed");
  return DEX_OK;
}

extern "C" int dex_read_eigvecs_info(const char *path, int32_t *m, int32_t *num) {
  if (!path || !m || !num) return dex::set_error(DEX_EINVAL, "null argument");
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return dex::set_error(DEX_EINVAL, "cannot open the eigenvector file");
  std::vector<char> rec;
  if (!read_record(fh.f, rec) || rec.size() < 12) return dex::set_error(DEX_EINVAL, "bad header record");
  std::memcpy(m, rec.data(), 4);
  std::memcpy(num, rec.data() + 4, 4);
  return DEX_OK;
}

extern "C" int dex_read_eigvecs(const char *path, int32_t m, int32_t num, double *psi) {
  if (!path || !psi || m < 1 || num < 0) return dex::set_error(DEX_EINVAL, "bad argument");
  File fh;
  fh.f = std::fopen(path, "rb");
  if (!fh.f) return dex::set_error(DEX_EINVAL, "cannot open the eigenvector file");
  std::vector<char> rec;
  if (!read_record(fh.f, rec) || rec.size() < 12) return dex::set_error(DEX_EINVAL, "bad header record");
  int32_t m_in = 0, num_in = 0;
  std::memcpy(&m_in, rec.data(), 4);
  std::memcpy(&num_in, rec.data() + 4, 4);
  const int64_t bytes_in = int64_t(m_in) * m_in * 16;
  // the size field: int64 (16-byte header) or its low 32 bits (12-byte header)
  if (rec.size() >= 16) {
    int64_t sz = 0;
    std::memcpy(&sz, rec.data() + 8, 8);
    if (sz != bytes_in) return dex::set_error(DEX_ESIZE, "header size field does not match M");
  } else {
    uint32_t sz = 0;
    std::memcpy(&sz, rec.data() + 8, 4);
    if (sz != uint32_t(uint64_t(bytes_in))) return dex::set_error(DEX_ESIZE, "header size field does not match M");
  }
  if (num > num_in) return dex::set_error(DEX_ESIZE, "the file holds fewer vectors than requested");
  const bool fix = m_in != m;
  int n_in = 0, n_out = 0;
  if (fix) {
    // utils.F90:118-133: both M must be odd cubes
    const int ri = cube_root_odd(m_in), ro = cube_root_odd(m);
    if (ri < 0 || ro < 0) return dex::set_error(DEX_ESIZE, "M is not the cube of an odd grid size");
    n_in = (ri - 1) / 2;
    n_out = (ro - 1) / 2;
  }
  const size_t out_elems = size_t(m) * m;
  for (int32_t v = 0; v < num; ++v) {
    if (!read_record(fh.f, rec) || int64_t(rec.size()) != bytes_in)
      return dex::set_error(DEX_ESIZE, "vector record has the wrong length");
    double *dst = psi + 2 * out_elems * size_t(v);
    if (fix)
      sixd_fix(reinterpret_cast<const double *>(rec.data()), n_in, dst, n_out);
    else
      std::memcpy(dst, rec.data(), rec.size());
  }
  return DEX_OK;
}

extern "C" int dex_host_sixd_fix(const double *in, int32_t n_in, double *out, int32_t n_out) {
  if (!in || !out || n_in < 0 || n_out < 0) return dex::set_error(DEX_EINVAL, "bad argument");
  sixd_fix(in, n_in, out, n_out);
  return DEX_OK;
}
