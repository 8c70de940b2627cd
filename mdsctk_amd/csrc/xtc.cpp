// XTC trajectory codec (GROMACS compressed coordinates), written from the
// published format: XDR (big-endian) frames of
//   int magic (1995), int natoms, int step, float time, float box[3][3],
//   int natoms, then the xdr3dfcoord block:
//     natoms <= 9 : 3*natoms raw floats
//     otherwise   : float precision, int minint[3], int maxint[3],
//                   int smallidx, int nbytes, nbytes of bit-packed data
//                   (padded to 4 bytes).
// The bit stream stores each coordinate as an integer round(x*precision);
// a "large" triple is a mixed-radix number over (maxint-minint+1) per axis,
// and runs of following atoms are stored as small differences whose radix
// comes from the magic-integer table below, adapting smallidx as it goes.
// Reference call sites: knn_rms.cpp:155-178 (open/read_first_xtc),
// :186-206 (read_next_xtc).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mdsctk_io.h"
#include "../../include/mdsctk_knn.h"

namespace {

thread_local std::string g_err;

int io_fail(int code, const char* fmt, const char* a = "", long long b = 0) {
  char buf[512];
  snprintf(buf, sizeof buf, fmt, a, b);
  g_err = buf;
  return code;
}

constexpr int kMagic = 1995;

// Radix table of the small-difference coding ("magic integers"): about
// 2^(i/3), the table of the published xdr3dfcoord format.  Entries below
// kFirstIdx are unused.
constexpr int kMagicInts[] = {
    0,       0,       0,       0,       0,       0,        0,        0,        0,
    8,       10,      12,      16,      20,      25,       32,       40,       50,
    64,      80,      101,     128,     161,     203,      256,      322,      406,
    512,     645,     812,     1024,    1290,    1625,     2048,     2580,     3250,
    4096,    5060,    6501,    8192,    10321,   13003,    16384,    20642,    26007,
    32768,   41285,   52015,   65536,   82570,   104031,   131072,   165140,   208063,
    262144,  330280,  416127,  524287,  660561,  832255,   1048576,  1321122,  1664510,
    2097152, 2642245, 3329021, 4194304, 5284491, 6658042,  8388607,  10568983, 13316085,
    16777216};
constexpr int kFirstIdx = 9;
constexpr int kLastIdx = (int)(sizeof(kMagicInts) / sizeof(kMagicInts[0]));

// ---- big-endian scalar access ------------------------------------------------
inline uint32_t be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
inline int32_t be_i32(const unsigned char* p) { return (int32_t)be32(p); }
inline float be_f32(const unsigned char* p) {
  uint32_t u = be32(p);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
inline void put_be32(std::vector<unsigned char>& o, uint32_t v) {
  o.push_back((unsigned char)(v >> 24));
  o.push_back((unsigned char)(v >> 16));
  o.push_back((unsigned char)(v >> 8));
  o.push_back((unsigned char)v);
}
inline void put_i32(std::vector<unsigned char>& o, int32_t v) { put_be32(o, (uint32_t)v); }
inline void put_f32(std::vector<unsigned char>& o, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  put_be32(o, u);
}

// ---- bit-level helpers ---------------------------------------------------------

// Number of bits needed to hold values 0..size.
int bits_for(unsigned size) {
  unsigned num = 1;
  int nb = 0;
  while (size >= num && nb < 32) {
    ++nb;
    num <<= 1;
  }
  return nb;
}

// Number of bits needed for a mixed-radix number with radices sizes[0..n).
int bits_for_product(int n, const unsigned* sizes) {
  unsigned bytes[32];
  int nbytes = 1;
  bytes[0] = 1;
  for (int i = 0; i < n; ++i) {
    unsigned carry = 0;
    int b = 0;
    for (; b < nbytes; ++b) {
      carry = bytes[b] * sizes[i] + carry;
      bytes[b] = carry & 0xff;
      carry >>= 8;
    }
    while (carry != 0) {
      bytes[b++] = carry & 0xff;
      carry >>= 8;
    }
    nbytes = b;
  }
  int nbits = 0;
  unsigned num = 1;
  --nbytes;
  while (bytes[nbytes] >= num) {
    ++nbits;
    num *= 2;
  }
  return nbits + nbytes * 8;
}

// MSB-first bit reader over the packed block.
struct BitReader {
  const unsigned char* data;
  size_t size;
  size_t pos = 0;
  unsigned lastbits = 0;
  unsigned lastbyte = 0;
  bool overrun = false;

  unsigned char next() {
    if (pos >= size) {
      overrun = true;
      return 0;
    }
    return data[pos++];
  }
  int bits(int nbits) {
    const unsigned mask = nbits >= 32 ? 0xffffffffu : ((1u << nbits) - 1u);
    unsigned num = 0;
    while (nbits >= 8) {
      lastbyte = (lastbyte << 8) | next();
      num |= (lastbyte >> lastbits) << (nbits - 8);
      nbits -= 8;
    }
    if (nbits > 0) {
      if ((int)lastbits < nbits) {
        lastbits += 8;
        lastbyte = (lastbyte << 8) | next();
      }
      lastbits -= nbits;
      num |= (lastbyte >> lastbits) & ((1u << nbits) - 1u);
    }
    return (int)(num & mask);
  }
  // Mixed-radix triple: the packed number is
  //   X = (n0 * s1 + n1) * s2 + n2, stored little-endian in bytes.
  void ints3(int nbits, const unsigned* sizes, int* nums) {
    int bytes[32];
    bytes[1] = bytes[2] = bytes[3] = 0;
    int nb = 0;
    while (nbits > 8) {
      bytes[nb++] = bits(8);
      nbits -= 8;
    }
    if (nbits > 0) bytes[nb++] = bits(nbits);
    for (int i = 2; i > 0; --i) {
      unsigned num = 0;
      for (int j = nb - 1; j >= 0; --j) {
        num = (num << 8) | (unsigned)bytes[j];
        const unsigned q = num / sizes[i];
        bytes[j] = (int)q;
        num = num - q * sizes[i];
      }
      nums[i] = (int)num;
    }
    nums[0] = bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24);
  }
};

// MSB-first bit writer.
struct BitWriter {
  std::vector<unsigned char> out;
  unsigned lastbits = 0;
  unsigned lastbyte = 0;

  void bits(int nbits, unsigned num) {
    while (nbits >= 8) {
      lastbyte = (lastbyte << 8) | ((num >> (nbits - 8)) & 0xffu);
      out.push_back((unsigned char)(lastbyte >> lastbits));
      nbits -= 8;
    }
    if (nbits > 0) {
      lastbyte = (lastbyte << nbits) | (num & ((1u << nbits) - 1u));
      lastbits += nbits;
      if (lastbits >= 8) {
        lastbits -= 8;
        out.push_back((unsigned char)(lastbyte >> lastbits));
      }
    }
  }
  void flush() {
    if (lastbits > 0) out.push_back((unsigned char)(lastbyte << (8 - lastbits)));
    lastbits = 0;
  }
  void ints3(int nbits, const unsigned* sizes, const unsigned* nums) {
    // X = (n0 * s1 + n1) * s2 + n2 as little-endian bytes.
    unsigned bytes[32] = {0};
    int nb = 0;
    unsigned v = nums[0];
    do {
      bytes[nb++] = v & 0xff;
      v >>= 8;
    } while (v != 0);
    for (int i = 1; i < 3; ++i) {
      unsigned carry = nums[i];
      int b = 0;
      for (; b < nb; ++b) {
        const unsigned long long t = (unsigned long long)bytes[b] * sizes[i] + carry;
        bytes[b] = (unsigned)(t & 0xff);
        carry = (unsigned)(t >> 8);
      }
      while (carry != 0) {
        bytes[b++] = carry & 0xff;
        carry >>= 8;
      }
      nb = b;
    }
    int j = 0;
    while (nbits > 8) {
      bits(8, bytes[j++]);
      nbits -= 8;
    }
    if (nbits > 0) bits(nbits, bytes[j]);
  }
};

// ---- frame index ---------------------------------------------------------------

struct FrameRef {
  size_t off;
  size_t len;
};

struct FileData {
  std::vector<unsigned char> bytes;
};

int load_file(const char* path, std::vector<unsigned char>& bytes) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(MDSCTK_ERR_IO, "cannot open %s%lld", path, 0);
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return io_fail(MDSCTK_ERR_IO, "cannot size %s%lld", path, 0);
  }
  bytes.resize((size_t)sz);
  size_t got = sz ? std::fread(bytes.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if (got != (size_t)sz) return io_fail(MDSCTK_ERR_IO, "short read on %s%lld", path, 0);
  return 0;
}

// Walk frame headers; fills `frames` and the atom count.
int index_frames(const std::vector<unsigned char>& b, std::vector<FrameRef>& frames,
                 int* natoms_out, float* prec_out) {
  size_t pos = 0;
  int natoms = -1;
  float prec = 0.f;
  while (pos < b.size()) {
    const size_t start = pos;
    if (b.size() - pos < 56) return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame header%s at frame %lld", "", (long long)frames.size());
    if (be_i32(&b[pos]) != kMagic)
      return io_fail(MDSCTK_ERR_FORMAT, "bad XTC magic%s in frame %lld", "", (long long)frames.size());
    const int na = be_i32(&b[pos + 4]);
    const int na2 = be_i32(&b[pos + 52]);
    if (na <= 0 || na != na2)
      return io_fail(MDSCTK_ERR_FORMAT, "inconsistent XTC atom count%s in frame %lld", "", (long long)frames.size());
    if (natoms < 0) natoms = na;
    if (na != natoms)
      return io_fail(MDSCTK_ERR_FORMAT, "XTC atom count changes%s at frame %lld", "", (long long)frames.size());
    pos += 56;
    if (na <= 9) {
      if (b.size() - pos < (size_t)12 * na) return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame%s %lld", "", (long long)frames.size());
      pos += (size_t)12 * na;
    } else {
      if (b.size() - pos < 36) return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame%s %lld", "", (long long)frames.size());
      if (frames.empty()) prec = be_f32(&b[pos]);
      const int nbytes = be_i32(&b[pos + 32]);
      if (nbytes < 0) return io_fail(MDSCTK_ERR_FORMAT, "negative XTC block size%s in frame %lld", "", (long long)frames.size());
      const size_t padded = ((size_t)nbytes + 3) & ~size_t(3);
      pos += 36;
      if (b.size() - pos < padded) return io_fail(MDSCTK_ERR_FORMAT, "truncated XTC frame%s %lld", "", (long long)frames.size());
      pos += padded;
    }
    frames.push_back({start, pos - start});
  }
  if (natoms_out) *natoms_out = natoms < 0 ? 0 : natoms;
  if (prec_out) *prec_out = prec;
  return 0;
}

// Decode one frame.  Returns 0 or a negative error code.
int decode_frame(const unsigned char* p, size_t len, int natoms, float* xyz, int32_t* step,
                 float* time, float* box) {
  if (step) *step = be_i32(p + 8);
  if (time) *time = be_f32(p + 12);
  if (box)
    for (int i = 0; i < 9; ++i) box[i] = be_f32(p + 16 + 4 * i);
  const unsigned char* q = p + 56;
  if (natoms <= 9) {
    for (int i = 0; i < 3 * natoms; ++i) xyz[i] = be_f32(q + 4 * i);
    return 0;
  }
  const float precision = be_f32(q);
  int minint[3], maxint[3];
  for (int i = 0; i < 3; ++i) {
    minint[i] = be_i32(q + 4 + 4 * i);
    maxint[i] = be_i32(q + 16 + 4 * i);
  }
  int smallidx = be_i32(q + 28);
  const int nbytes = be_i32(q + 32);
  if (!(precision > 0.f) || smallidx < kFirstIdx || smallidx >= kLastIdx)
    return io_fail(MDSCTK_ERR_FORMAT, "bad XTC compression header%s%lld", "", 0);
  if ((size_t)nbytes + 92 > len + 3) return io_fail(MDSCTK_ERR_FORMAT, "bad XTC block size%s%lld", "", 0);
  unsigned sizeint[3];
  int bitsizeint[3] = {0, 0, 0};
  int bitsize = 0;
  for (int i = 0; i < 3; ++i) sizeint[i] = (unsigned)(maxint[i] - minint[i]) + 1u;
  if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffffu) {
    for (int i = 0; i < 3; ++i) bitsizeint[i] = bits_for(sizeint[i]);
    bitsize = 0;
  } else {
    bitsize = bits_for_product(3, sizeint);
  }
  int smaller = kMagicInts[std::max(kFirstIdx, smallidx - 1)] / 2;
  int smallnum = kMagicInts[smallidx] / 2;
  unsigned sizesmall[3];
  sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagicInts[smallidx];

  BitReader br{q + 36, (size_t)nbytes};
  const float inv_precision = 1.0f / precision;
  float* out = xyz;
  int i = 0, run = 0;
  int cur[3], prev[3];
  while (i < natoms) {
    if (bitsize == 0) {
      for (int d = 0; d < 3; ++d) cur[d] = br.bits(bitsizeint[d]);
    } else {
      br.ints3(bitsize, sizeint, cur);
    }
    ++i;
    for (int d = 0; d < 3; ++d) {
      cur[d] += minint[d];
      prev[d] = cur[d];
    }
    const int flag = br.bits(1);
    int is_smaller = 0;
    if (flag == 1) {
      run = br.bits(5);
      is_smaller = run % 3;
      run -= is_smaller;
      --is_smaller;
    }
    if (run > 0) {
      if (i + run / 3 > natoms) return io_fail(MDSCTK_ERR_FORMAT, "XTC run overflows the frame%s%lld", "", 0);
      for (int k = 0; k < run; k += 3) {
        int nxt[3];
        br.ints3(smallidx, sizesmall, nxt);
        ++i;
        for (int d = 0; d < 3; ++d) nxt[d] += prev[d] - smallnum;
        if (k == 0) {
          // The first two atoms of a run are stored swapped (water-friendly
          // ordering); emit the small-coded one first.
          for (int d = 0; d < 3; ++d) std::swap(nxt[d], prev[d]);
          for (int d = 0; d < 3; ++d) *out++ = (float)prev[d] * inv_precision;
        } else {
          for (int d = 0; d < 3; ++d) prev[d] = nxt[d];
        }
        for (int d = 0; d < 3; ++d) *out++ = (float)nxt[d] * inv_precision;
      }
    } else {
      for (int d = 0; d < 3; ++d) *out++ = (float)cur[d] * inv_precision;
    }
    smallidx += is_smaller;
    if (smallidx < kFirstIdx || smallidx >= kLastIdx)
      return io_fail(MDSCTK_ERR_FORMAT, "XTC small index out of range%s%lld", "", 0);
    if (is_smaller < 0) {
      smallnum = smaller;
      smaller = (smallidx > kFirstIdx) ? kMagicInts[smallidx - 1] / 2 : 0;
    } else if (is_smaller > 0) {
      smaller = smallnum;
      smallnum = kMagicInts[smallidx] / 2;
    }
    sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagicInts[smallidx];
  }
  if (br.overrun) return io_fail(MDSCTK_ERR_FORMAT, "XTC bit stream overrun%s%lld", "", 0);
  return 0;
}

// Quantise as the format does: round half away from zero.
int quantise(float x, float precision, int* out) {
  double lf = (double)x * (double)precision;
  lf = lf >= 0.0 ? lf + 0.5 : lf - 0.5;
  if (std::fabs(lf) > 2147483647.0 - 1.0) return -1;
  *out = (int)lf;
  return 0;
}

// Encode one frame.  Every atom is written as a full mixed-radix triple
// followed by a 0 flag bit (no small-difference runs), which every XTC
// reader decodes.
int encode_frame(std::vector<unsigned char>& o, const float* xyz, int natoms, float precision,
                 int32_t step, float time, const float* box) {
  put_i32(o, kMagic);
  put_i32(o, natoms);
  put_i32(o, step);
  put_f32(o, time);
  for (int i = 0; i < 9; ++i) put_f32(o, box ? box[i] : 0.0f);
  put_i32(o, natoms);
  if (natoms <= 9) {
    for (int i = 0; i < 3 * natoms; ++i) put_f32(o, xyz[i]);
    return 0;
  }
  std::vector<int> q((size_t)natoms * 3);
  int minint[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
  int maxint[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int i = 0; i < natoms; ++i)
    for (int d = 0; d < 3; ++d) {
      int v;
      if (quantise(xyz[3 * i + d], precision, &v))
        return io_fail(MDSCTK_ERR_ARG, "coordinate too large for XTC precision%s%lld", "", 0);
      q[3 * i + d] = v;
      minint[d] = std::min(minint[d], v);
      maxint[d] = std::max(maxint[d], v);
    }
  unsigned sizeint[3];
  int bitsizeint[3] = {0, 0, 0};
  int bitsize;
  for (int d = 0; d < 3; ++d) {
    if ((long long)maxint[d] - (long long)minint[d] >= 0x7fffffffLL)
      return io_fail(MDSCTK_ERR_ARG, "coordinate range too large for XTC%s%lld", "", 0);
    sizeint[d] = (unsigned)(maxint[d] - minint[d]) + 1u;
  }
  if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffffu) {
    for (int d = 0; d < 3; ++d) bitsizeint[d] = bits_for(sizeint[d]);
    bitsize = 0;
  } else {
    bitsize = bits_for_product(3, sizeint);
  }
  put_f32(o, precision);
  for (int d = 0; d < 3; ++d) put_i32(o, minint[d]);
  for (int d = 0; d < 3; ++d) put_i32(o, maxint[d]);
  put_i32(o, kFirstIdx);
  BitWriter bw;
  for (int i = 0; i < natoms; ++i) {
    unsigned v[3];
    for (int d = 0; d < 3; ++d) v[d] = (unsigned)(q[3 * i + d] - minint[d]);
    if (bitsize == 0) {
      for (int d = 0; d < 3; ++d) bw.bits(bitsizeint[d], v[d]);
    } else {
      bw.ints3(bitsize, sizeint, v);
    }
    bw.bits(1, 0u);
  }
  bw.flush();
  put_i32(o, (int32_t)bw.out.size());
  o.insert(o.end(), bw.out.begin(), bw.out.end());
  while (o.size() % 4) o.push_back(0);
  return 0;
}

}  // namespace

extern "C" {

const char* mdsctk_io_last_error(void) { return g_err.c_str(); }

void mdsctk_io_set_error(const char* msg) { g_err = msg ? msg : ""; }

int mdsctk_xtc_info(const char* path, int64_t* n_frames, int32_t* natoms, float* precision) {
  if (!path) return io_fail(MDSCTK_ERR_ARG, "null path%s%lld", "", 0);
  std::vector<unsigned char> b;
  int rc = load_file(path, b);
  if (rc) return rc;
  std::vector<FrameRef> fr;
  int na = 0;
  float prec = 0;
  if ((rc = index_frames(b, fr, &na, &prec))) return rc;
  if (n_frames) *n_frames = (int64_t)fr.size();
  if (natoms) *natoms = na;
  if (precision) *precision = prec;
  return 0;
}

int mdsctk_xtc_read(const char* path, float* xyz, int64_t max_frames, int32_t natoms,
                    int64_t* n_read, int32_t* steps, float* times, float* boxes, int nthreads) {
  if (!path || (!xyz && max_frames > 0)) return io_fail(MDSCTK_ERR_ARG, "null argument%s%lld", "", 0);
  std::vector<unsigned char> b;
  int rc = load_file(path, b);
  if (rc) return rc;
  std::vector<FrameRef> fr;
  int na = 0;
  if ((rc = index_frames(b, fr, &na, nullptr))) return rc;
  if (!fr.empty() && na != natoms)
    return io_fail(MDSCTK_ERR_NATOMS, "%s has a different number of atoms (%lld)", path, na);
  const int64_t n = std::min<int64_t>((int64_t)fr.size(), max_frames);
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  nthreads = (int)std::min<int64_t>(nthreads, std::max<int64_t>(1, n));
  std::atomic<int64_t> next{0};
  std::atomic<int> err{0};
  std::vector<std::string> errs(nthreads);
  auto work = [&](int t) {
    for (;;) {
      const int64_t f = next.fetch_add(1);
      if (f >= n || err.load()) break;
      int r = decode_frame(&b[fr[f].off], fr[f].len, natoms, xyz + (size_t)f * natoms * 3,
                           steps ? steps + f : nullptr, times ? times + f : nullptr,
                           boxes ? boxes + (size_t)f * 9 : nullptr);
      if (r) {
        int expected = 0;
        if (err.compare_exchange_strong(expected, r)) errs[t] = g_err + " (frame " + std::to_string(f) + ")";
        break;
      }
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work, t);
    for (auto& t : th) t.join();
  }
  if (err.load()) {
    for (auto& e : errs)
      if (!e.empty()) g_err = e;
    return err.load();
  }
  if (n_read) *n_read = n;
  return 0;
}

int mdsctk_xtc_natoms(const char* path, int32_t* natoms) {
  if (!path || !natoms) return io_fail(MDSCTK_ERR_ARG, "null argument%s%lld", "", 0);
  FILE* f = std::fopen(path, "rb");
  if (!f) return io_fail(MDSCTK_ERR_IO, "cannot open %s%lld", path, 0);
  unsigned char h0[8];
  const size_t got = std::fread(h0, 1, sizeof h0, f);
  std::fclose(f);
  if (got != sizeof h0 || be_i32(h0) != kMagic)
    return io_fail(MDSCTK_ERR_FORMAT, "%s does not start with an XTC frame%lld", path, 0);
  *natoms = be_i32(h0 + 4);
  return 0;
}

int mdsctk_xtc_write(const char* path, const float* xyz, int64_t n_frames, int32_t natoms,
                     float precision, const int32_t* steps, const float* times,
                     const float* boxes) {
  if (!path || (!xyz && n_frames > 0) || natoms <= 0 || n_frames < 0 || !(precision > 0.f))
    return io_fail(MDSCTK_ERR_ARG, "bad arguments to xtc_write%s%lld", "", 0);
  FILE* f = std::fopen(path, "wb");
  if (!f) return io_fail(MDSCTK_ERR_IO, "cannot create %s%lld", path, 0);
  std::vector<unsigned char> o;
  for (int64_t i = 0; i < n_frames; ++i) {
    o.clear();
    int rc = encode_frame(o, xyz + (size_t)i * natoms * 3, natoms, precision,
                          steps ? steps[i] : (int32_t)i, times ? times[i] : 0.0f,
                          boxes ? boxes + (size_t)i * 9 : nullptr);
    if (rc) {
      std::fclose(f);
      return rc;
    }
    if (std::fwrite(o.data(), 1, o.size(), f) != o.size()) {
      std::fclose(f);
      return io_fail(MDSCTK_ERR_IO, "write failed on %s%lld", path, 0);
    }
  }
  if (std::fclose(f) != 0) return io_fail(MDSCTK_ERR_IO, "close failed on %s%lld", path, 0);
  return 0;
}

}  // extern "C"
