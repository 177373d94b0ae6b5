// Host side of the order-independent exact sums. The words are the reference's exact-sum format
// (runtime/src/exact_sum.cpp:22-103, kept so that the words mean the same): 68 signed 64-bit limbs,
// limb i weighing 2^(32 i - 1088), then the counts of +inf, -inf and NaN terms. The device kernels
// add into the same words.
//
// Rounding: the limbs are resolved once into the sign and the 68 base-2^32 digits of |sum| (one
// carry pass in 128-bit arithmetic); the 64 bits from the leading one down are read as one integer
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

namespace exact_detail {

// |sum| as base-2^32 digits plus whatever lies above the last limb (`over`), and the sign.
struct Magnitude {
  std::uint32_t digit[kExactLimbs];
  unsigned __int128 over;  // multiples of 2^(32 * kExactLimbs) in units of limb 0
  bool negative;
};

MXE_HOST_DEVICE inline Magnitude resolve(const long long* w) {
  Magnitude m{};
  __int128 carry = 0;
  for (int i = 0; i < kExactLimbs; ++i) {
    const __int128 acc = carry + w[i];
    m.digit[i] = static_cast<std::uint32_t>(static_cast<unsigned __int128>(acc) & 0xffffffffu);
    carry = acc >> 32;  // floor division: the digits stay in [0, 2^32)
  }
  m.negative = carry < 0;
  if (!m.negative) {
    m.over = static_cast<unsigned __int128>(carry);
    return m;
  }
  // Two's complement of (carry, digits): invert, add one from the lowest digit up.
  unsigned long long up = 1;
  for (int i = 0; i < kExactLimbs; ++i) {
    const unsigned long long v = static_cast<unsigned long long>(~m.digit[i]) + up;
    m.digit[i] = static_cast<std::uint32_t>(v);
    up = v >> 32;
  }
  m.over = static_cast<unsigned __int128>(-(carry + 1)) + up;
  return m;
}

// Bit `b` (0 = the lowest bit of digit 0) of the digits.
MXE_HOST_DEVICE inline unsigned bit_of(const Magnitude& m, int b) {
  return b < 0 ? 0u : (m.digit[b >> 5] >> (b & 31)) & 1u;
}

// Whether any bit below position `b` is set.
MXE_HOST_DEVICE inline bool any_below(const Magnitude& m, int b) {
  if (b <= 0) return false;
  const int whole = b >> 5;
  for (int i = 0; i < whole; ++i)
    if (m.digit[i] != 0) return true;
  const int part = b & 31;
  return part != 0 && (m.digit[whole] & ((1u << part) - 1u)) != 0;
}

}  // namespace exact_detail

// The value of exact words, rounded once to the nearest double (ties to even).
MXE_HOST_DEVICE inline double exact_value(const long long* in) {
  using namespace exact_detail;
  constexpr double inf = std::numeric_limits<double>::infinity();
  const long long pos_inf = in[kExactLimbs], neg_inf = in[kExactLimbs + 1], nans = in[kExactLimbs + 2];
  if (nans > 0 || (pos_inf > 0 && neg_inf > 0)) return std::numeric_limits<double>::quiet_NaN();
  if (pos_inf > 0) return inf;
  if (neg_inf > 0) return -inf;
  const Magnitude m = resolve(in);
  // 2^(32 * 68 - 1088) = 2^1088 and above is past the largest double.
  if (m.over != 0) return m.negative ? -inf : inf;
  int lead = -1;  // position of the leading one
  for (int i = kExactLimbs - 1; i >= 0 && lead < 0; --i)
    if (m.digit[i] != 0) lead = 32 * i + 31 - __builtin_clz(m.digit[i]);
  if (lead < 0) return 0.0;
  // The 64 bits lead - 63 .. lead (zeros below bit 0), rounded to odd by everything under them.
  std::uint64_t top = 0;
  for (int b = lead; b > lead - 64; --b) top = (top << 1) | bit_of(m, b);
  if (any_below(m, lead - 63)) top |= 1u;
  const double rounded = static_cast<double>(top);  // to nearest even at 53 bits
  const double magnitude = std::ldexp(rounded, lead - 63 - kExactBias);
  return m.negative ? -magnitude : magnitude;
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
