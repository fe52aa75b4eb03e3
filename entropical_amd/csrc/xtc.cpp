// xtc.cpp -- GROMACS XTC trajectory reader (host side of the data path).
//
// An XTC frame is XDR (big-endian) data: magic 1995, natoms, step, time,
// a 3x3 box, then the coordinate block.  Blocks of at most 9 atoms are plain
// floats; larger ones use Frans van Hoesel's compression, as the xdrfile
// library that the reference vendors decodes it
// (coords_file/xdrfile/xdrfile.c:761-968, xdrfile_xtc.c:18-73):
//   precision, minint[3], maxint[3], smallidx, byte count, packed bits.
// Each group starts with one atom coded as a mixed-radix integer over the
// ranges maxint - minint + 1 (or three plain bit fields when a range exceeds
// 2^24), followed by a flag bit and, when set, a 5-bit run length whose
// remainder mod 3 moves smallidx down, keeps it or moves it up; a run codes
// run/3 more atoms as small differences (radix magicints[smallidx]) from the
// previous atom, the first two atoms of a run stored swapped.  Coordinates
// are int * (1/precision) in float.  The integer table below is part of the
// format and is the same as the library's.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/entropical_kde.h"

namespace ent {
void set_error(const std::string& msg);  // engine.hip: ent_last_error() message
}

namespace {

using ent::set_error;

constexpr int32_t kMagic = 1995;
// Radices of the small-difference code (format constant).
constexpr int32_t kSmallRadix[] = {
    0,       0,       0,       0,        0,        0,        0,        0,        0,       8,       10,
    12,      16,      20,      25,       32,       40,       50,       64,       80,      101,     128,
    161,     203,     256,     322,      406,      512,      645,      812,      1024,    1290,    1625,
    2048,    2580,    3250,    4096,     5060,     6501,     8192,     10321,    13003,   16384,   20642,
    26007,   32768,   41285,   52015,    65536,    82570,    104031,   131072,   165140,  208063,  262144,
    330280,  416127,  524287,  660561,   832255,   1048576,  1321122,  1664510,  2097152, 2642245, 3329021,
    4194304, 5284491, 6658042, 8388607,  10568983, 13316085, 16777216};
constexpr int kFirstIdx = 9;
constexpr int kLastIdx = (int)(sizeof(kSmallRadix) / sizeof(kSmallRadix[0]));

struct Xdr {
  const uint8_t* p;
  const uint8_t* end;
  bool u32(uint32_t& v) {
    if (end - p < 4) return false;
    v = (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
    p += 4;
    return true;
  }
  bool i32(int32_t& v) {
    uint32_t u;
    if (!u32(u)) return false;
    v = (int32_t)u;
    return true;
  }
  bool f32(float& f) {
    uint32_t u;
    if (!u32(u)) return false;
    std::memcpy(&f, &u, 4);
    return true;
  }
  bool skip(size_t n) {
    if ((size_t)(end - p) < n) return false;
    p += n;
    return true;
  }
};

// Most-significant-bit-first reader over the packed bytes.
struct Bits {
  const uint8_t* b;
  size_t n_bytes;
  size_t pos = 0;  // bit position
  bool ok = true;
  uint32_t get(int k) {  // k <= 32
    if (k == 0) return 0;
    if (pos + (size_t)k > n_bytes * 8) {
      ok = false;
      return 0;
    }
    uint64_t v = 0;
    size_t byte = pos >> 3;
    const int off = (int)(pos & 7);
    const int need = (off + k + 7) >> 3;  // bytes touched, <= 5
    for (int i = 0; i < need; ++i) v = (v << 8) | b[byte + i];
    v >>= need * 8 - off - k;
    pos += (size_t)k;
    return (uint32_t)(v & ((k == 32) ? 0xffffffffull : ((1ull << k) - 1)));
  }
  // Three integers coded as one mixed-radix number of n_bits bits: the bits
  // come as bytes, least significant byte first, the last one possibly short.
  void get3(int n_bits, const uint32_t sizes[3], int32_t out[3]) {
    unsigned __int128 v = 0;
    int shift = 0;
    while (n_bits > 8) {
      v |= (unsigned __int128)get(8) << shift;
      shift += 8;
      n_bits -= 8;
    }
    if (n_bits > 0) v |= (unsigned __int128)get(n_bits) << shift;
    out[2] = (int32_t)(uint32_t)(v % sizes[2]);
    v /= sizes[2];
    out[1] = (int32_t)(uint32_t)(v % sizes[1]);
    v /= sizes[1];
    out[0] = (int32_t)(uint32_t)v;
  }
};

// bits needed for values 0..size-1 of one field: smallest b with size < 2^b
int bits_of(uint32_t size) {
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) <= size) ++b;
  return b;
}

// bits of the mixed-radix code of three fields: bit length of the product
int bits_of_product(const uint32_t sizes[3]) {
  unsigned __int128 prod = (unsigned __int128)sizes[0] * sizes[1] * sizes[2];
  int b = 0;
  while (prod) {
    ++b;
    prod >>= 1;
  }
  return b;
}

// One frame's coordinate block; xyz receives natoms*3 floats when non-null.
bool read_coords(Xdr& x, int32_t natoms, float* xyz) {
  int32_t lsize;
  if (!x.i32(lsize) || lsize != natoms) return false;
  if (natoms <= 9) {
    for (int i = 0; i < 3 * natoms; ++i) {
      float f;
      if (!x.f32(f)) return false;
      if (xyz) xyz[i] = f;
    }
    return true;
  }
  float precision;
  int32_t minint[3], maxint[3], smallidx, n_bytes;
  if (!x.f32(precision)) return false;
  for (int d = 0; d < 3; ++d)
    if (!x.i32(minint[d])) return false;
  for (int d = 0; d < 3; ++d)
    if (!x.i32(maxint[d])) return false;
  if (!x.i32(smallidx) || !x.i32(n_bytes) || n_bytes < 0) return false;
  const uint8_t* packed = x.p;
  if (!x.skip(((size_t)n_bytes + 3) & ~(size_t)3)) return false;
  if (!xyz) return true;
  if (smallidx < kFirstIdx || smallidx >= kLastIdx) return false;

  uint32_t sizeint[3];
  for (int d = 0; d < 3; ++d) sizeint[d] = (uint32_t)(maxint[d] - minint[d]) + 1u;
  int bitsizeint[3] = {0, 0, 0};
  int bitsize = 0;  // 0: three separate fields
  if ((sizeint[0] | sizeint[1] | sizeint[2]) > 0xffffffu) {
    for (int d = 0; d < 3; ++d) bitsizeint[d] = bits_of(sizeint[d]);
  } else {
    bitsize = bits_of_product(sizeint);
  }
  int smaller = kSmallRadix[std::max(kFirstIdx, smallidx - 1)] / 2;
  int smallnum = kSmallRadix[smallidx] / 2;
  uint32_t sizesmall[3];
  sizesmall[0] = sizesmall[1] = sizesmall[2] = (uint32_t)kSmallRadix[smallidx];

  Bits in{packed, (size_t)n_bytes};
  const float inv_precision = (float)(1.0 / (double)precision);
  auto emit = [&](const int32_t c[3], float*& out) {
    for (int d = 0; d < 3; ++d) *out++ = (float)c[d] * inv_precision;
  };
  float* out = xyz;
  int32_t i = 0;
  int run = 0;
  while (i < natoms) {
    int32_t cur[3], prev[3];
    if (bitsize == 0) {
      for (int d = 0; d < 3; ++d) cur[d] = (int32_t)in.get(bitsizeint[d]);
    } else {
      in.get3(bitsize, sizeint, cur);
    }
    ++i;
    for (int d = 0; d < 3; ++d) {
      cur[d] += minint[d];
      prev[d] = cur[d];
    }
    int step = 0;  // change of smallidx after this group
    if (in.get(1)) {
      run = (int)in.get(5);
      step = run % 3;
      run -= step;
      step -= 1;
    }
    if (run > 0) {
      if (i + run / 3 > natoms) return false;
      for (int k = 0; k < run; k += 3) {
        int32_t nxt[3];
        in.get3(smallidx, sizesmall, nxt);
        ++i;
        for (int d = 0; d < 3; ++d) nxt[d] += prev[d] - smallnum;
        if (k == 0) {
          // the group's first two atoms are stored swapped
          emit(nxt, out);
          emit(prev, out);
          for (int d = 0; d < 3; ++d) prev[d] = nxt[d];
        } else {
          for (int d = 0; d < 3; ++d) prev[d] = nxt[d];
          emit(nxt, out);
        }
      }
    } else {
      emit(cur, out);
    }
    smallidx += step;
    if (smallidx < kFirstIdx || smallidx >= kLastIdx) return false;
    if (step < 0) {
      smallnum = smaller;
      smaller = smallidx > kFirstIdx ? kSmallRadix[smallidx - 1] / 2 : 0;
    } else if (step > 0) {
      smaller = smallnum;
      smallnum = kSmallRadix[smallidx] / 2;
    }
    sizesmall[0] = sizesmall[1] = sizesmall[2] = (uint32_t)kSmallRadix[smallidx];
    if (!in.ok) return false;
  }
  return in.ok;
}

// Frame header; false at a clean end of data.
bool read_header(Xdr& x, int32_t& natoms, bool& bad) {
  bad = false;
  if (x.p == x.end) return false;
  int32_t magic, step;
  float time;
  if (!x.i32(magic) || magic != kMagic || !x.i32(natoms) || natoms <= 0 || !x.i32(step) || !x.f32(time) ||
      !x.skip(9 * 4)) {
    bad = true;
    return false;
  }
  return true;
}

}  // namespace

extern "C" int ent_read_xtc(const char* path, float* out, size_t capacity, size_t* n_frames, uint32_t* natoms) {
  if (!path || !n_frames || !natoms) {
    set_error("null argument");
    return ENT_ERR_ARG;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    set_error(std::string("cannot open XTC file '") + path + "'");
    return ENT_ERR_ARG;
  }
  std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  // pass 1: frame count and atom count
  Xdr x{data.data(), data.data() + data.size()};
  size_t frames = 0;
  int32_t na = -1;
  for (;;) {
    int32_t n;
    bool bad;
    if (!read_header(x, n, bad)) {
      if (bad) {
        set_error("malformed XTC frame header after frame " + std::to_string(frames));
        return ENT_ERR_ARG;
      }
      break;
    }
    if (na >= 0 && n != na) {
      set_error("XTC frames with different atom counts");
      return ENT_ERR_ARG;
    }
    na = n;
    if (!read_coords(x, n, nullptr)) {
      set_error("truncated XTC coordinate block in frame " + std::to_string(frames));
      return ENT_ERR_ARG;
    }
    ++frames;
  }
  *n_frames = frames;
  *natoms = na < 0 ? 0u : (uint32_t)na;
  if (!out) return ENT_OK;
  const size_t cols = 3 * (size_t)*natoms;
  if (capacity < frames * cols) {
    set_error("output capacity below frames * 3 * natoms");
    return ENT_ERR_ARG;
  }
  // pass 2: decode; frame f is row f, atom a dim e is column 3a + e
  std::vector<float> xyz(cols);
  x = Xdr{data.data(), data.data() + data.size()};
  for (size_t fr = 0; fr < frames; ++fr) {
    int32_t n;
    bool bad;
    if (!read_header(x, n, bad) || !read_coords(x, n, xyz.data())) {
      set_error("corrupt XTC coordinate block in frame " + std::to_string(fr));
      return ENT_ERR_ARG;
    }
    for (size_t c = 0; c < cols; ++c) out[c * frames + fr] = xyz[c];
  }
  return ENT_OK;
}
