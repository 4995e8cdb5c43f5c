// (key, index) list words of the tiled search's FP32 top-K lists.
//
// Compiled for the device by the search kernel and for the host by the CPU test of the
// encoding (the same source).  A list entry -- an FP32 key c (never NaN) and a point
// index i in [0, 2^31) -- is one 64-bit word W whose value read as an IEEE double is
// finite and increases with (c, i) in lexicographic order, where keys compare as floats
// (-0 == +0).  An insertion into the ascending list w[0..N) is then
//     w[j] = min(w[j], max(w[j-1], x))   (old values, w[-1] = -inf)
// per slot: two FP64 min/max, in place, with no compare masks.
//
//   c >= +0 (bits b, top bit clear):  W = b << 31 | i
//   c <= -0 (bits b, top bit set):    W = 1 << 63 | (b & 0x7fffffff) << 31 | (0x7fffffff - i)
//
// Positive words lie in [0, 0x3fc00000_7fffffff] (b <= 0x7f800000 = +inf): finite, and
// ordered like unsigned integers, which for non-negative doubles is numeric order
// (denormal words arise for keys below 2^-106 and for key 0, so FP64 denormals must not be
// flushed, which is the device default).  Negative words have magnitude >= 2^31 (b & 0x7fffffff >= 1 once -0 is
// folded into +0), so -0.0 never occurs, and a larger magnitude (more negative key, or
// same key and smaller index) is a smaller double.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define AKNN_OHD __host__ __device__
#else
#define AKNN_OHD
#endif

namespace aknn {

// Word of (c, i); c is folded to +0 when it is -0 (c + 0.0f does that in round-to-nearest).
AKNN_OHD inline uint64_t list_word(float c, int32_t i) {
  const uint32_t b = __builtin_bit_cast(uint32_t, c + 0.0f);
  const uint32_t neg = static_cast<uint32_t>(static_cast<int32_t>(b) >> 31);  // all ones for c < 0
  const uint32_t hi = (b >> 1) ^ (neg & 0xc0000000u);
  const uint32_t lo = ((b << 31) | static_cast<uint32_t>(i)) ^ (neg & 0x7fffffffu);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Key of a word (bit-exact, +0 for a folded -0).
AKNN_OHD inline float list_word_key(uint64_t w) {
  const uint32_t hi = static_cast<uint32_t>(w >> 32);
  const uint32_t mag = static_cast<uint32_t>(w >> 31) & 0x7fffffffu;  // bits 62..31
  return __builtin_bit_cast(float, mag | (hi & 0x80000000u));
}

// Index of a word.
AKNN_OHD inline int32_t list_word_index(uint64_t w) {
  const uint32_t lo = static_cast<uint32_t>(w) & 0x7fffffffu;
  return static_cast<int32_t>((w >> 63) ? 0x7fffffffu - lo : lo);
}

}  // namespace aknn
