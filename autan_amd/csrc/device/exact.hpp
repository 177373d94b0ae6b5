// Host side of the order-independent exact sums. The words are the reference's exact-sum format
// (runtime/src/exact_sum.cpp:22-103, kept so that the words mean the same): 68 signed 64-bit limbs,
// limb i weighing 2^(32 i - 1088), then the counts of +inf, -inf and NaN terms. The device kernels
// add into the same words.
//
// Rounding: the limbs are resolved into the sign and the base-2^32 digits of |sum| (carry passes in
// 128-bit arithmetic); the 64 bits from the leading one down are read as one integer
// whose lowest bit is OR-ed with every bit below them (rounding to odd at 64 bits), and that integer
// is converted to a double, which rounds it to nearest even at 53 bits. Rounding to odd at p + 2 or
// more bits then to nearest at p bits gives the correctly rounded result (Boldo and Melquiond 2008,
// "Emulation of FMA and correctly rounded sums: proved algorithms using rounding to odd", IEEE TC
// 57(4)). The power-of-two scaling after it is exact: a sum of doubles is a multiple of 2^-1074, so
// a result below 2^-1021 has fewer than 54 significant bits and was never rounded, and above it the
// scaled value is normal.
//
// exact_value and its helpers are host and device functions when compiled as HIP: the device
// rounds words it summed itself (the implicit solve's inner products) with the same code, so both
// sides give the same double.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#if defined(__HIP__)
#define MXE_HOST_DEVICE __host__ __device__
#else
#define MXE_HOST_DEVICE
#endif

namespace mxe {

inline constexpr int kExactLimbs = 68;
inline constexpr int kExactWords = kExactLimbs + 3;
inline constexpr int kExactBias = 1088;  // limb 0's weight is 2^-kExactBias

// The value of exact words, rounded once to the nearest double (ties to even). Two passes over the
// limbs in one register set (no digit array, so a device thread runs it without private memory):
// the first resolves the carries only to learn the sign; the second resolves them again into the
// base-2^32 digits of |sum| (two's complement on the fly when negative) and keeps the highest
// non-zero digit, the two digits under it and whether any digit below those is non-zero.
MXE_HOST_DEVICE inline double exact_value(const long long* in) {
  constexpr double inf = std::numeric_limits<double>::infinity();
  const long long pos_inf = in[kExactLimbs], neg_inf = in[kExactLimbs + 1], nans = in[kExactLimbs + 2];
  if (nans > 0 || (pos_inf > 0 && neg_inf > 0)) return std::numeric_limits<double>::quiet_NaN();
  if (pos_inf > 0) return inf;
  if (neg_inf > 0) return -inf;
  // Unrolled: the limbs' loads do not wait on the carry chain.
  __int128 carry = 0;
#pragma GCC unroll 68
  for (int i = 0; i < kExactLimbs; ++i) carry = (carry + in[i]) >> 32;  // floor division
  const bool negative = carry < 0;
  carry = 0;
  unsigned long long up = 1;  // the +1 of the two's complement, carried up the digits
  std::uint32_t prev1 = 0, prev2 = 0, prev3 = 0;  // digits i - 1, i - 2, i - 3
  bool low_any = false;                            // any non-zero digit below i - 3
  int top_index = -1;                              // the highest non-zero digit so far
  std::uint32_t top0 = 0, top1 = 0, top2 = 0;      // that digit and the two under it
  bool top_sticky = false;                         // any non-zero digit under those three
#pragma GCC unroll 68
  for (int i = 0; i < kExactLimbs; ++i) {
    const __int128 acc = carry + in[i];
    std::uint32_t digit = static_cast<std::uint32_t>(static_cast<unsigned __int128>(acc) & 0xffffffffu);
    carry = acc >> 32;
    if (negative) {
      const unsigned long long v = static_cast<unsigned long long>(~digit) + up;
      digit = static_cast<std::uint32_t>(v);
      up = v >> 32;
    }
    low_any = low_any || prev3 != 0;
    if (digit != 0) {
      top_index = i;
      top0 = digit;
      top1 = prev1;
      top2 = prev2;
      top_sticky = low_any || prev3 != 0;
    }
    prev3 = prev2;
    prev2 = prev1;
    prev1 = digit;
  }
  // Multiples of 2^(32 * kExactLimbs) = 2^1088 (in units of limb 0) and above: past the largest
  // double.
  const unsigned __int128 over =
      negative ? static_cast<unsigned __int128>(-(carry + 1)) + up : static_cast<unsigned __int128>(carry);
  if (over != 0) return negative ? -inf : inf;
  if (top_index < 0) return 0.0;
  int clz = 0;
  while (((top0 << clz) & 0x80000000u) == 0) ++clz;
  const int lead = 32 * top_index + 31 - clz;  // position of the leading one
  // The 96 bits of the three digits; the 64 from the leading one down, rounded to odd by every bit
  // under them (the rest of the window and the digits below it).
  const int shift = 32 - clz;  // 1 .. 32
  const unsigned __int128 window = (static_cast<unsigned __int128>(top0) << 64) |
                                   (static_cast<unsigned __int128>(top1) << 32) | top2;
  std::uint64_t top = static_cast<std::uint64_t>(window >> shift);
  const bool sticky = top_sticky || (window & ((static_cast<unsigned __int128>(1) << shift) - 1)) != 0;
  if (sticky) top |= 1u;
  const double rounded = static_cast<double>(top);  // to nearest even at 53 bits
  const double magnitude = std::ldexp(rounded, lead - 63 - kExactBias);
  return negative ? -magnitude : magnitude;
}

// A running exact sum on the host, in the same words.
class ExactSum {
 public:
  void add(double term) {
    std::uint64_t bits;
    std::memcpy(&bits, &term, sizeof bits);
    const unsigned field = static_cast<unsigned>(bits >> 52) & 0x7ffu;
    const std::uint64_t frac = bits & ((std::uint64_t{1} << 52) - 1);
    const bool minus = (bits >> 63) != 0;
    if (field == 0x7ffu) {  // inf or NaN: counted, never added
      ++words_[frac != 0 ? kExactLimbs + 2 : kExactLimbs + (minus ? 1 : 0)];
      return;
    }
    // term = significand * 2^(exponent); position in units of limb 0's weight
    const std::uint64_t significand = field == 0 ? frac : frac | (std::uint64_t{1} << 52);
    if (significand == 0) return;
    const int position = (field == 0 ? -1074 : int(field) - 1075) + kExactBias;
    const unsigned __int128 placed = static_cast<unsigned __int128>(significand) << (position & 31);
    const int first = position >> 5;
    for (int k = 0; k < 3; ++k) {
      const long long piece = static_cast<long long>(static_cast<std::uint64_t>(placed >> (32 * k)) & 0xffffffffu);
      words_[first + k] += minus ? -piece : piece;
    }
    // A limb moves by less than 2^32 per term; resolving carries every 2^30 terms keeps every limb
    // below 2^63 in magnitude.
    if (++since_carry_ == (1u << 30)) carry();
  }
  [[nodiscard]] double value() const { return exact_value(words_); }

 private:
  void carry() {
    for (int i = 0; i + 1 < kExactLimbs; ++i) {
      const long long high = words_[i] >> 32;  // floor: the limb keeps its low 32 bits
      words_[i] &= 0xffffffffLL;
      words_[i + 1] += high;
    }
    since_carry_ = 0;
  }
  long long words_[kExactWords] = {};
  unsigned since_carry_ = 0;
};

}  // namespace mxe
