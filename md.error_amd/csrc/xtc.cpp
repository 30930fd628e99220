// xtc.cpp -- GROMACS xtc trajectory frames (compressed coordinates): write
// and read.  Host only.
//
// The reference writes its trajectory with libxdrfile's write_xtc from
// MDSystem::initWriteXtc / writeHostDataXtc / endWriteXtc
// (H/src/MDSystem_interface.cu:243-290, called every few steps by the driver
// H/app/lj.nhc-xtc.cu:155-158,260,275).  libxdrfile is an external library
// that the reference neither vendors nor pins; this file restates the
// published xtc frame layout so that the files are readable by the usual xtc
// readers.
//
// Frame (XDR: big-endian 4-byte words):
//   int 1995 (magic), int natoms, int step, float time, float box[3][3],
//   int natoms again, then
//   natoms <= 9: 3 natoms floats (uncompressed), or
//   float precision, int min[3], int max[3], int first class index,
//   int byte count, the bit stream (zero-padded to a multiple of 4 bytes).
// Coordinates become integers at the precision (x * precision rounded half
// away from zero: the FP32 product plus 0.5, truncated).  The stream is a
// sequence of "large" atoms, each followed by a run of up to 8 "small"
// atoms:
//   * a large atom: its integer coordinates minus min[], as one mixed-radix
//     number of bit length bitlen(product of the three ranges) when every
//     range is at most 2^24, else as three fields of bitlen(range) bits;
//   * one bit: 1 if a 5-bit code follows (run length in atoms times 3 plus
//     the class step + 1), 0 if the previous run length is reused with no
//     class step;
//   * the run: each small atom as its difference from the previous atom plus
//     s/2 per component (s = class size), packed as one mixed-radix number of
//     `class index` bits.
// The class (index into kClassSize, whose entry i is the largest size s
// with s^3 <= 2^i, up to a few historical entries that the format keeps)
// moves by at most one index per large atom.  When the atom after a large
// atom lies within s/2 of it the two are swapped, so that the second is sent
// as the large atom and the first as the first small one (the format's
// water-molecule optimisation; a reader swaps them back).
//
// Bits are packed most significant first; a mixed-radix number of b bits is
// sent as its bytes from the least significant, 8 bits each, the last with
// the remaining b mod 8 (or 8) bits.
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <climits>
#include <vector>

#include "../../include/sre.h"

namespace sre {
sre_status fail(sre_status s, const char *fmt, ...);
}
using sre::fail;

struct sre_xtc {
  FILE *fp = nullptr;
  bool reading = false;
  std::vector<int> ints;             // integer coordinates of a frame
  std::vector<unsigned char> bytes;  // bit stream of a frame
};

namespace {

constexpr int kMagic = 1995;
constexpr int kFirstClass = 9;
// class sizes of the format (index = bits of three packed small values)
constexpr int kClassSize[] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0,
    8, 10, 12, 16, 20, 25, 32, 40, 50, 64,
    80, 101, 128, 161, 203, 256, 322, 406, 512, 645,
    812, 1024, 1290, 1625, 2048, 2580, 3250, 4096, 5060, 6501,
    8192, 10321, 13003, 16384, 20642, 26007, 32768, 41285, 52015, 65536,
    82570, 104031, 131072, 165140, 208063, 262144, 330280, 416127, 524287, 660561,
    832255, 1048576, 1321122, 1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042,
    8388607, 10568983, 13316085, 16777216};
constexpr int kNumClasses = (int)(sizeof(kClassSize) / sizeof(kClassSize[0]));
static_assert(kNumClasses == 73, "class table");

typedef unsigned __int128 u128;

// number of bits of v (0 for v == 0)
int bit_length(u128 v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// ---- XDR words ---------------------------------------------------------

bool put_u32(FILE *f, uint32_t v) {
  unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16),
                        (unsigned char)(v >> 8), (unsigned char)v};
  return fwrite(b, 1, 4, f) == 4;
}
bool put_i32(FILE *f, int v) { return put_u32(f, (uint32_t)v); }
bool put_f32(FILE *f, float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return put_u32(f, u);
}
// 0 = ok, 1 = clean end of file before the word, -1 = truncated
int get_u32(FILE *f, uint32_t *v) {
  unsigned char b[4];
  const size_t n = fread(b, 1, 4, f);
  if (n == 0) return 1;
  if (n != 4) return -1;
  *v = ((uint32_t)b[0] << 24) | ((uint32_t)b[1] << 16) | ((uint32_t)b[2] << 8) | b[3];
  return 0;
}
bool get_i32(FILE *f, int *v) {
  uint32_t u;
  if (get_u32(f, &u) != 0) return false;
  *v = (int)u;
  return true;
}
bool get_f32(FILE *f, float *v) {
  uint32_t u;
  if (get_u32(f, &u) != 0) return false;
  memcpy(v, &u, 4);
  return true;
}

// ---- bit stream --------------------------------------------------------

struct BitWriter {
  std::vector<unsigned char> &out;
  uint32_t pending = 0;   // low `npend` bits not yet written
  int npend = 0;
  explicit BitWriter(std::vector<unsigned char> &o) : out(o) { out.clear(); }
  void bits(int n, uint32_t v) {   // the low n bits of v, most significant first
    while (n > 0) {
      const int take = n < 8 ? n : 8;
      const uint32_t chunk = (v >> (n - take)) & ((1u << take) - 1u);
      pending = (pending << take) | chunk;
      npend += take;
      n -= take;
      if (npend >= 8) {
        out.push_back((unsigned char)(pending >> (npend - 8)));
        npend -= 8;
        pending &= (1u << npend) - 1u;
      }
    }
  }
  // the number v < 2^nbits as bytes from the least significant
  void number(int nbits, u128 v) {
    while (nbits > 8) {
      bits(8, (uint32_t)(v & 0xff));
      v >>= 8;
      nbits -= 8;
    }
    bits(nbits, (uint32_t)v);
  }
  void finish() {
    if (npend > 0) out.push_back((unsigned char)(pending << (8 - npend)));
    npend = 0;
    pending = 0;
  }
};

struct BitReader {
  const std::vector<unsigned char> &in;
  size_t pos = 0;     // next byte
  uint32_t cur = 0;   // unread low `ncur` bits of the bytes read so far
  int ncur = 0;
  bool overrun = false;
  explicit BitReader(const std::vector<unsigned char> &i) : in(i) {}
  uint32_t bits(int n) {
    uint32_t v = 0;
    while (n > 0) {
      if (ncur == 0) {
        if (pos >= in.size()) {
          overrun = true;
          return 0;
        }
        cur = in[pos++];
        ncur = 8;
      }
      const int take = n < ncur ? n : ncur;
      v = (v << take) | ((cur >> (ncur - take)) & ((1u << take) - 1u));
      ncur -= take;
      n -= take;
    }
    return v;
  }
  u128 number(int nbits) {
    u128 v = 0;
    int shift = 0;
    while (nbits > 8) {
      v |= (u128)bits(8) << shift;
      shift += 8;
      nbits -= 8;
    }
    v |= (u128)bits(nbits) << shift;
    return v;
  }
};

// three values as one mixed-radix number, the first most significant
u128 pack3(const uint32_t v[3], const uint32_t size[3]) {
  return ((u128)v[0] * size[1] + v[1]) * size[2] + v[2];
}
void unpack3(u128 n, const uint32_t size[3], int v[3]) {
  v[2] = (int)(n % size[2]);
  n /= size[2];
  v[1] = (int)(n % size[1]);
  v[0] = (int)(n / size[1]);
}

// FP32 product plus one half, truncated (half away from zero)
int to_int(float x, float prec) {
  const float p = x * prec;
  const float lf = p >= 0.f ? (float)((double)p + 0.5) : (float)((double)p - 0.5);
  return (int)lf;
}

// The small-difference class: index, size s, s/2 ("smallnum") and the half
// size of the class below ("smaller", 0 at the first class).
struct Classes {
  int idx, smallnum, smaller;
  uint32_t size;
  explicit Classes(int i)
      : idx(i), smallnum(kClassSize[i] / 2),
        smaller(kClassSize[i - 1 > kFirstClass ? i - 1 : kFirstClass] / 2),
        size((uint32_t)kClassSize[i]) {}
  // one step down or up after a large atom
  void adapt(int step) {
    if (step == 0) return;
    idx += step;
    if (step < 0) {
      smallnum = smaller;
      smaller = idx > kFirstClass ? kClassSize[idx - 1] / 2 : 0;
    } else {
      smaller = smallnum;
      smallnum = kClassSize[idx] / 2;
    }
    size = (uint32_t)kClassSize[idx];
  }
};

bool close_all(const int *a, const int *b, int lim) {
  return abs(a[0] - b[0]) < lim && abs(a[1] - b[1]) < lim && abs(a[2] - b[2]) < lim;
}

// Encodes n > 9 integer coordinates (in place: the swaps are applied to q).
// Returns false if a value cannot be represented.
bool encode(int n, int *q, const int mn[3], const int mx[3], int mindiff, int *first_class,
            std::vector<unsigned char> &out) {
  uint32_t range[3];
  for (int d = 0; d < 3; ++d) range[d] = (uint32_t)((int64_t)mx[d] - mn[d] + 1);
  const bool wide = (range[0] | range[1] | range[2]) > 0xffffffu;
  int field_bits[3] = {0, 0, 0};
  int large_bits = 0;
  if (wide) {
    for (int d = 0; d < 3; ++d) field_bits[d] = bit_length(range[d]);
  } else {
    large_bits = bit_length((u128)range[0] * range[1] * range[2]);
  }
  int ci = kFirstClass;
  while (ci < kNumClasses && kClassSize[ci] < mindiff) ++ci;
  if (ci >= kNumClasses) ci = kNumClasses - 1;   // atoms far apart: no small atoms
  *first_class = ci;
  const int hi_class = ci + 8 < kNumClasses - 1 ? ci + 8 : kNumClasses - 1;
  const int lo_class = hi_class - 8;
  const int larger = kClassSize[hi_class] / 2;
  Classes c(ci);
  BitWriter w(out);
  int prev[3] = {0, 0, 0};
  int prevrun = -1;
  uint32_t run_vals[8 * 3];
  int i = 0;
  while (i < n) {
    int *cur = q + 3 * i;
    int step;   // class step after this large atom: -1, 0, +1
    if (c.idx < hi_class && i >= 1 && close_all(cur, prev, larger)) step = 1;
    else if (c.idx > lo_class) step = -1;
    else step = 0;
    bool small_next = false;
    if (i + 1 < n && close_all(cur, cur + 3, c.smallnum)) {
      for (int d = 0; d < 3; ++d) {
        const int t = cur[d];
        cur[d] = cur[d + 3];
        cur[d + 3] = t;
      }
      small_next = true;
    }
    uint32_t rel[3];
    for (int d = 0; d < 3; ++d) rel[d] = (uint32_t)((int64_t)cur[d] - mn[d]);
    if (wide) {
      for (int d = 0; d < 3; ++d) w.bits(field_bits[d], rel[d]);
    } else {
      w.number(large_bits, pack3(rel, range));
    }
    for (int d = 0; d < 3; ++d) prev[d] = cur[d];
    ++i;
    int run = 0;
    if (!small_next && step == -1) step = 0;
    while (small_next && run < 8 * 3) {
      const int *nx = q + 3 * i;
      if (step == -1) {
        const int64_t dx = nx[0] - prev[0], dy = nx[1] - prev[1], dz = nx[2] - prev[2];
        if (dx * dx + dy * dy + dz * dz >= (int64_t)c.smaller * c.smaller) step = 0;
      }
      for (int d = 0; d < 3; ++d) run_vals[run++] = (uint32_t)(nx[d] - prev[d] + c.smallnum);
      for (int d = 0; d < 3; ++d) prev[d] = nx[d];
      ++i;
      small_next = i < n && close_all(q + 3 * i, prev, c.smallnum);
    }
    if (run != prevrun || step != 0) {
      prevrun = run;
      w.bits(1, 1);
      w.bits(5, (uint32_t)(run + step + 1));
    } else {
      w.bits(1, 0);
    }
    const uint32_t sz[3] = {c.size, c.size, c.size};
    for (int k = 0; k < run; k += 3) {
      for (int d = 0; d < 3; ++d)
        if (run_vals[k + d] >= c.size) return false;
      w.number(c.idx, pack3(run_vals + k, sz));
    }
    c.adapt(step);
  }
  w.finish();
  return true;
}

bool decode(int n, int *q, const int mn[3], const int mx[3], int first_class,
            const std::vector<unsigned char> &in) {
  uint32_t range[3];
  for (int d = 0; d < 3; ++d) range[d] = (uint32_t)((int64_t)mx[d] - mn[d] + 1);
  const bool wide = (range[0] | range[1] | range[2]) > 0xffffffu;
  int field_bits[3] = {0, 0, 0};
  int large_bits = 0;
  if (wide) {
    for (int d = 0; d < 3; ++d) field_bits[d] = bit_length(range[d]);
  } else {
    large_bits = bit_length((u128)range[0] * range[1] * range[2]);
  }
  if (first_class < kFirstClass || first_class >= kNumClasses) return false;
  Classes c(first_class);
  BitReader r(in);
  int run = 0;
  int i = 0;
  while (i < n) {
    int *cur = q + 3 * i;
    if (wide) {
      for (int d = 0; d < 3; ++d) cur[d] = (int)r.bits(field_bits[d]);
    } else {
      unpack3(r.number(large_bits), range, cur);
    }
    for (int d = 0; d < 3; ++d) cur[d] += mn[d];
    ++i;
    int step = 0;
    if (r.bits(1)) {
      const int code = (int)r.bits(5);
      step = code % 3 - 1;
      run = code - code % 3;
    }
    if (run > 0) {
      if (i + run / 3 > n) return false;
      int *ref = cur;   // the atom the next difference is taken from
      const uint32_t sz[3] = {c.size, c.size, c.size};
      for (int k = 0; k < run; k += 3) {
        int *nx = q + 3 * i;
        unpack3(r.number(c.idx), sz, nx);
        for (int d = 0; d < 3; ++d) nx[d] += ref[d] - c.smallnum;
        if (k == 0) {
          // the large atom was the second of the pair: swap back, and the
          // next difference is taken from the first atom (now at ref)
          for (int d = 0; d < 3; ++d) {
            const int t = nx[d];
            nx[d] = ref[d];
            ref[d] = t;
          }
        } else {
          ref = nx;
        }
        ++i;
      }
    }
    if (step != 0 && (c.idx + step < kFirstClass || c.idx + step >= kNumClasses)) return false;
    c.adapt(step);
    if (r.overrun) return false;
  }
  return !r.overrun;
}

}  // namespace

extern "C" sre_status sre_xtc_open(const char *path, const char *mode, sre_xtc **out) {
  if (!path || !mode || !out) return fail(SRE_ERR_INVALID, "sre_xtc_open: null argument");
  *out = nullptr;
  const char *m = nullptr;
  bool reading = false;
  if (!strcmp(mode, "w")) m = "wb";
  else if (!strcmp(mode, "a")) m = "ab";
  else if (!strcmp(mode, "r")) { m = "rb"; reading = true; }
  else return fail(SRE_ERR_INVALID, "sre_xtc_open: mode \"%s\" (w, a or r)", mode);
  FILE *fp = fopen(path, m);
  if (!fp) return fail(SRE_ERR_IO, "cannot open %s (MDExcptCannotOpenFile)", path);
  sre_xtc *f = new sre_xtc;
  f->fp = fp;
  f->reading = reading;
  *out = f;
  return SRE_OK;
}

extern "C" sre_status sre_xtc_close(sre_xtc *f) {
  if (!f) return SRE_OK;
  const int r = f->fp ? fclose(f->fp) : 0;
  delete f;
  return r == 0 ? SRE_OK : fail(SRE_ERR_IO, "xtc: close failed");
}

extern "C" sre_status sre_xtc_write(sre_xtc *f, int natoms, int step, float time,
                                    const float box[9], const float *xyz, float precision) {
  if (!f || !box || (natoms > 0 && !xyz) || natoms < 0)
    return fail(SRE_ERR_INVALID, "sre_xtc_write: bad argument");
  if (f->reading) return fail(SRE_ERR_INVALID, "sre_xtc_write: file opened for reading");
  if (!(precision > 0.f)) precision = 1000.f;
  FILE *fp = f->fp;
  bool ok = put_i32(fp, kMagic) && put_i32(fp, natoms) && put_i32(fp, step) &&
            put_f32(fp, time);
  for (int k = 0; k < 9 && ok; ++k) ok = put_f32(fp, box[k]);
  ok = ok && put_i32(fp, natoms);
  if (!ok) return fail(SRE_ERR_IO, "xtc: write failed");
  if (natoms <= 9) {
    for (int k = 0; k < 3 * natoms && ok; ++k) ok = put_f32(fp, xyz[k]);
    return ok ? SRE_OK : fail(SRE_ERR_IO, "xtc: write failed");
  }
  std::vector<int> &q = f->ints;
  q.resize((size_t)3 * natoms);
  int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
  int mindiff = INT_MAX;
  const float lim = (float)(INT_MAX - 2);
  for (int i = 0; i < natoms; ++i) {
    for (int d = 0; d < 3; ++d) {
      const float p = xyz[3 * i + d] * precision;
      if (!(fabsf(p) < lim))
        return fail(SRE_ERR_INVALID, "xtc: coordinate %g of atom %d beyond the integer range "
                    "at precision %g", (double)xyz[3 * i + d], i, (double)precision);
      const int v = to_int(xyz[3 * i + d], precision);
      q[3 * i + d] = v;
      if (v < mn[d]) mn[d] = v;
      if (v > mx[d]) mx[d] = v;
    }
    if (i > 0) {
      const int64_t diff = llabs((int64_t)q[3 * i] - q[3 * i - 3]) +
                           llabs((int64_t)q[3 * i + 1] - q[3 * i - 2]) +
                           llabs((int64_t)q[3 * i + 2] - q[3 * i - 1]);
      if (diff < mindiff) mindiff = (int)diff;
    }
  }
  for (int d = 0; d < 3; ++d)
    if ((int64_t)mx[d] - mn[d] >= (int64_t)INT_MAX - 2)
      return fail(SRE_ERR_INVALID, "xtc: coordinate range too large at precision %g",
                  (double)precision);
  int first_class = kFirstClass;
  if (!encode(natoms, q.data(), mn, mx, mindiff, &first_class, f->bytes))
    return fail(SRE_ERR_INVALID, "xtc: a small difference exceeds its class (bug)");
  ok = put_f32(fp, precision);
  for (int d = 0; d < 3 && ok; ++d) ok = put_i32(fp, mn[d]);
  for (int d = 0; d < 3 && ok; ++d) ok = put_i32(fp, mx[d]);
  ok = ok && put_i32(fp, first_class) && put_i32(fp, (int)f->bytes.size());
  if (ok && !f->bytes.empty())
    ok = fwrite(f->bytes.data(), 1, f->bytes.size(), fp) == f->bytes.size();
  static const unsigned char zero[4] = {0, 0, 0, 0};
  const size_t pad = (4 - f->bytes.size() % 4) % 4;
  if (ok && pad) ok = fwrite(zero, 1, pad, fp) == pad;
  return ok ? SRE_OK : fail(SRE_ERR_IO, "xtc: write failed");
}

extern "C" sre_status sre_xtc_read(sre_xtc *f, int natoms, int *step, float *time, float box[9],
                                   float *xyz, float *precision) {
  if (!f || !step || !time || !box || (natoms > 0 && !xyz) || natoms < 0)
    return fail(SRE_ERR_INVALID, "sre_xtc_read: bad argument");
  if (!f->reading) return fail(SRE_ERR_INVALID, "sre_xtc_read: file opened for writing");
  FILE *fp = f->fp;
  uint32_t magic;
  const int e = get_u32(fp, &magic);
  if (e == 1) return SRE_END_OF_FILE;
  if (e < 0 || (int)magic != kMagic) return fail(SRE_ERR_IO, "xtc: bad frame header");
  int n = 0, n2 = 0;
  bool ok = get_i32(fp, &n) && get_i32(fp, step) && get_f32(fp, time);
  for (int k = 0; k < 9 && ok; ++k) ok = get_f32(fp, &box[k]);
  ok = ok && get_i32(fp, &n2);
  if (!ok) return fail(SRE_ERR_IO, "xtc: truncated frame");
  if (n != natoms || n2 != natoms)
    return fail(SRE_ERR_INVALID, "xtc: frame holds %d atoms, %d expected", n, natoms);
  if (precision) *precision = 0.f;
  if (natoms <= 9) {
    for (int k = 0; k < 3 * natoms && ok; ++k) ok = get_f32(fp, &xyz[k]);
    return ok ? SRE_OK : fail(SRE_ERR_IO, "xtc: truncated frame");
  }
  float prec = 0.f;
  int mn[3], mx[3], first_class = 0, nbytes = 0;
  ok = get_f32(fp, &prec);
  for (int d = 0; d < 3 && ok; ++d) ok = get_i32(fp, &mn[d]);
  for (int d = 0; d < 3 && ok; ++d) ok = get_i32(fp, &mx[d]);
  ok = ok && get_i32(fp, &first_class) && get_i32(fp, &nbytes);
  if (!ok || nbytes < 0 || !(prec > 0.f)) return fail(SRE_ERR_IO, "xtc: bad frame header");
  for (int d = 0; d < 3; ++d)
    if (mx[d] < mn[d]) return fail(SRE_ERR_IO, "xtc: bad frame header");
  const size_t padded = ((size_t)nbytes + 3) / 4 * 4;
  f->bytes.resize(padded);
  if (padded && fread(f->bytes.data(), 1, padded, fp) != padded)
    return fail(SRE_ERR_IO, "xtc: truncated frame");
  f->bytes.resize((size_t)nbytes);
  std::vector<int> &q = f->ints;
  q.assign((size_t)3 * natoms, 0);
  if (!decode(natoms, q.data(), mn, mx, first_class, f->bytes))
    return fail(SRE_ERR_IO, "xtc: corrupt coordinate stream");
  const float inv = 1.f / prec;
  for (size_t k = 0; k < q.size(); ++k) xyz[k] = (float)q[k] * inv;
  if (precision) *precision = prec;
  return SRE_OK;
}

extern "C" sre_status sre_xtc_natoms(const char *path, int *natoms) {
  if (!path || !natoms) return fail(SRE_ERR_INVALID, "sre_xtc_natoms: null argument");
  FILE *fp = fopen(path, "rb");
  if (!fp) return fail(SRE_ERR_IO, "cannot open %s", path);
  int magic = 0, n = 0;
  const bool ok = get_i32(fp, &magic) && get_i32(fp, &n);
  fclose(fp);
  if (!ok || magic != kMagic) return fail(SRE_ERR_IO, "%s: not an xtc file", path);
  *natoms = n;
  return SRE_OK;
}
