// XTC decode on the GPU: one lane per frame.
//
// The compressed coordinates of a frame (xdr3dfcoord, the format described
// in xtc.cpp) are a serial bit stream, but frames are independent: once the
// host has walked the frame headers (a cheap chain of lengths), every frame
// can be decoded at the same time.  The whole file is uploaded as it is
// read; each lane walks its own frame's bit stream through a cached 32-bit
// word and writes the frame's float32 nm coordinates.  The arithmetic
// (integer unpacking, mixed-radix division, float(i) * (1/precision)) is the
// host decoder's, so both produce the same bits.
// Reference call sites: knn_rms.cpp:155-206 (read_first_xtc / read_next_xtc).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mdsctk {
namespace {

constexpr int kXtcMagic = 1995;

// The format's radix table of small differences (about 2^(i/3)); entries
// below kFirstIdx are unused.
__constant__ int kMagicIntsDev[] = {
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
constexpr int kLastIdx = 73;
// Bytes of the largest mixed-radix number (3 x 24 bits, or a small triple
// of at most kLastIdx bits).
constexpr int kRadixBytes = 10;

__device__ __forceinline__ uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// Big-endian 32-bit field at a 4-byte aligned byte offset of the frame.
__device__ __forceinline__ uint32_t be32_at(const uint32_t* w, int byte_off) {
  return bswap32(w[byte_off >> 2]);
}

__device__ __forceinline__ int bitlen64(uint64_t v) { return v ? 64 - __clzll((long long)v) : 0; }

// Bits needed for values 0..size (capped at 32).
__device__ __forceinline__ int bits_for(unsigned size) {
  const int b = bitlen64(size);
  return b < 32 ? b : 32;
}

// Bit length of s0 * s1 * s2 (the packed size of a large triple).
__device__ __forceinline__ int bits_for_product3(unsigned s0, unsigned s1, unsigned s2) {
  const uint64_t p = (uint64_t)s0 * s1;  // < 2^48
  const uint64_t lo = (p & 0xffffffffull) * s2;
  const uint64_t hi = (p >> 32) * s2 + (lo >> 32);
  return hi ? 32 + bitlen64(hi) : bitlen64(lo & 0xffffffffull);
}

// MSB-first bit reader over a frame's packed block (4-byte aligned).
struct DevBits {
  const uint32_t* words;
  unsigned nbytes;
  unsigned pos;
  unsigned lastbits;
  unsigned lastbyte;
  unsigned wi;
  uint32_t w;
  bool overrun;

  __device__ unsigned next() {
    if (pos >= nbytes) {
      overrun = true;
      return 0;
    }
    const unsigned i = pos >> 2;
    if (i != wi) {
      w = words[i];
      wi = i;
    }
    const unsigned b = (w >> (8 * (pos & 3))) & 0xffu;
    ++pos;
    return b;
  }
  __device__ int bits(int nbits) {
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
  // Mixed-radix triple X = (n0 * s1 + n1) * s2 + n2, little-endian bytes;
  // the byte array stays in registers (every index is a constant after
  // unrolling).
  __device__ void ints3(int nbits, const unsigned* sizes, int* nums) {
    unsigned bytes[kRadixBytes];
    int nb = 0;
#pragma unroll
    for (int j = 0; j < kRadixBytes; ++j) {
      bytes[j] = 0;
      if (nbits > 8) {
        bytes[j] = (unsigned)bits(8);
        nbits -= 8;
        nb = j + 1;
      } else if (nbits > 0) {
        bytes[j] = (unsigned)bits(nbits);
        nbits = 0;
        nb = j + 1;
      }
    }
#pragma unroll
    for (int i = 2; i > 0; --i) {
      unsigned num = 0;
#pragma unroll
      for (int j = kRadixBytes - 1; j >= 0; --j) {
        if (j < nb) {
          num = (num << 8) | bytes[j];
          const unsigned q = num / sizes[i];
          bytes[j] = q;
          num -= q * sizes[i];
        }
      }
      nums[i] = (int)num;
    }
    nums[0] = (int)(bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24));
  }
};

// Decode frame f ([off[f], off[f+1]) of the file) into xyz[f].  Returns 0,
// or 1 when the frame is malformed.
__device__ int decode_frame_dev(const unsigned char* __restrict__ file, int64_t off, int64_t len,
                                int natoms, float* __restrict__ out) {
  const uint32_t* fw = reinterpret_cast<const uint32_t*>(file + off);
  if (len < 56 || (int)be32_at(fw, 0) != kXtcMagic || (int)be32_at(fw, 4) != natoms ||
      (int)be32_at(fw, 52) != natoms)
    return 1;
  if (natoms <= 9) {
    if (len < 56 + 12 * (int64_t)natoms) return 1;
    for (int i = 0; i < 3 * natoms; ++i) out[i] = __uint_as_float(be32_at(fw, 56 + 4 * i));
    return 0;
  }
  if (len < 92) return 1;
  const float precision = __uint_as_float(be32_at(fw, 56));
  int minint[3], maxint[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    minint[d] = (int)be32_at(fw, 60 + 4 * d);
    maxint[d] = (int)be32_at(fw, 72 + 4 * d);
  }
  int smallidx = (int)be32_at(fw, 84);
  const int nbytes = (int)be32_at(fw, 88);
  if (!(precision > 0.f) || smallidx < kFirstIdx || smallidx >= kLastIdx || nbytes < 0 ||
      (int64_t)nbytes + 92 > len + 3)
    return 1;
  unsigned sizeint[3];
  int bitsizeint[3] = {0, 0, 0};
  int bitsize = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) sizeint[d] = (unsigned)(maxint[d] - minint[d]) + 1u;
  if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffffu) {
#pragma unroll
    for (int d = 0; d < 3; ++d) bitsizeint[d] = bits_for(sizeint[d]);
  } else {
    bitsize = bits_for_product3(sizeint[0], sizeint[1], sizeint[2]);
  }
  int smaller = kMagicIntsDev[max(kFirstIdx, smallidx - 1)] / 2;
  int smallnum = kMagicIntsDev[smallidx] / 2;
  unsigned sizesmall[3];
  sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagicIntsDev[smallidx];

  DevBits br;
  br.words = fw + 23;  // the packed block starts at byte 92
  br.nbytes = (unsigned)nbytes;
  br.pos = 0;
  br.lastbits = 0;
  br.lastbyte = 0;
  br.wi = 0xffffffffu;
  br.w = 0;
  br.overrun = false;
  const float inv_precision = __frcp_rn(precision);  // the correctly rounded 1 / precision
  int i = 0, run = 0;
  int cur[3], prev[3];
  while (i < natoms) {
    if (bitsize == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) cur[d] = br.bits(bitsizeint[d]);
    } else {
      br.ints3(bitsize, sizeint, cur);
    }
    ++i;
#pragma unroll
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
      if (i + run / 3 > natoms) return 1;
      for (int k = 0; k < run; k += 3) {
        int nxt[3];
        br.ints3(smallidx, sizesmall, nxt);
        ++i;
#pragma unroll
        for (int d = 0; d < 3; ++d) nxt[d] += prev[d] - smallnum;
        if (k == 0) {
          // the first two atoms of a run are stored swapped
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const int t = nxt[d];
            nxt[d] = prev[d];
            prev[d] = t;
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) *out++ = (float)prev[d] * inv_precision;
        } else {
#pragma unroll
          for (int d = 0; d < 3; ++d) prev[d] = nxt[d];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) *out++ = (float)nxt[d] * inv_precision;
      }
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) *out++ = (float)cur[d] * inv_precision;
    }
    smallidx += is_smaller;
    if (smallidx < kFirstIdx || smallidx >= kLastIdx) return 1;
    if (is_smaller < 0) {
      smallnum = smaller;
      smaller = (smallidx > kFirstIdx) ? kMagicIntsDev[smallidx - 1] / 2 : 0;
    } else if (is_smaller > 0) {
      smaller = smallnum;
      smallnum = kMagicIntsDev[smallidx] / 2;
    }
    sizesmall[0] = sizesmall[1] = sizesmall[2] = (unsigned)kMagicIntsDev[smallidx];
  }
  return br.overrun ? 1 : 0;
}

// One lane per frame; *bad receives the smallest index of a malformed frame
// (xtc_bad_init_kernel sets it to ~0 first).
__global__ __launch_bounds__(64) void xtc_decode_kernel(const unsigned char* __restrict__ file,
                                                        const int64_t* __restrict__ off,
                                                        int64_t n, int natoms,
                                                        float* __restrict__ xyz,
                                                        unsigned long long* __restrict__ bad) {
  const int64_t f = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (f >= n) return;
  const int64_t o = off[f];
  if (decode_frame_dev(file, o, off[f + 1] - o, natoms, xyz + f * (int64_t)natoms * 3))
    atomicMin(bad, (unsigned long long)f);
}

__global__ void xtc_bad_init_kernel(unsigned long long* bad) { *bad = ~0ull; }

}  // namespace

// Decode frames [0, n) of an XTC file held in device memory (`file`, with
// at least 4 bytes of slack after its end), whose frame f spans bytes
// [off[f], off[f+1]) (off: n+1 device int64, every offset a multiple of 4),
// into xyz[n][natoms][3].  *bad (device) receives the smallest malformed
// frame index, or ~0.
hipError_t launch_xtc_decode(const unsigned char* file, const int64_t* off, int64_t n, int natoms,
                             float* xyz, unsigned long long* bad, hipStream_t st) {
  hipLaunchKernelGGL(xtc_bad_init_kernel, dim3(1), dim3(1), 0, st, bad);
  if (n > 0)
    hipLaunchKernelGGL(xtc_decode_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, file,
                       off, n, natoms, xyz, bad);
  return hipGetLastError();
}

}  // namespace mdsctk
