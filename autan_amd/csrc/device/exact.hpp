// Host side of the order-independent exact sum (runtime/src/exact_sum.cpp:22-103): 68 limbs of
// 32 bits, limb i weighing 2^(32 i - 1088), plus counts of +inf, -inf and NaN terms; the sum is
// rounded once, to nearest even. The device kernels add into the same word layout.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

namespace mxe {

inline constexpr int kExactLimbs = 68;
inline constexpr int kExactWords = kExactLimbs + 3;

// The value of exact words, rounded once to the nearest double (ties to even).
inline double exact_value(const long long* in) {
  constexpr double inf = std::numeric_limits<double>::infinity();
  constexpr int L = kExactLimbs;
  if (in[L + 2] > 0 || (in[L] > 0 && in[L + 1] > 0)) return std::numeric_limits<double>::quiet_NaN();
  if (in[L] > 0) return inf;
  if (in[L + 1] > 0) return -inf;
  long long w[L];
  for (int i = 0; i < L; ++i) w[i] = in[i];
  auto normalize = [&w] {
    for (int i = 0; i + 1 < L; ++i) {
      const long long carry = w[i] >> 32;
      w[i] -= carry * (1LL << 32);
      w[i + 1] += carry;
    }
  };
  normalize();
  const bool negative = w[L - 1] < 0;
  if (negative) {
    for (long long& v : w) v = -v;
    normalize();
  }
  int t = L - 1;
  while (t >= 0 && w[t] == 0) --t;
  if (t < 0) return 0.0;
  auto limb = [&w](int k) { return k < 0 ? 0ULL : static_cast<unsigned long long>(w[k]); };
  const unsigned long long top = limb(t);
  if (top > 0xffffffffULL) return negative ? -inf : inf;
  int b = 0;
  while ((top >> b) != 0) ++b;
  const unsigned long long below = (limb(t - 1) << 32) | limb(t - 2);
  const unsigned long long lead = (top << (64 - b)) | (below >> b);
  bool sticky = (below & ((1ULL << b) - 1)) != 0;
  for (int k = t - 3; k >= 0 && !sticky; --k) sticky = limb(k) != 0;
  unsigned long long mantissa = lead >> 11;
  const unsigned long long rest = lead & 0x7ff;
  if (rest > 0x400 || (rest == 0x400 && (sticky || (mantissa & 1) != 0))) ++mantissa;
  int exponent = 32 * (t - 2) - 1088 + b + 11;
  if (mantissa == (1ULL << 53)) {
    mantissa >>= 1;
    ++exponent;
  }
  const double magnitude = std::ldexp(static_cast<double>(mantissa), exponent);
  return negative ? -magnitude : magnitude;
}

// A running exact sum on the host.
class ExactSum {
 public:
  void add(double term) {
    std::uint64_t bits;
    std::memcpy(&bits, &term, 8);
    const int biased = static_cast<int>((bits >> 52) & 0x7ff);
    const std::uint64_t fraction = bits & ((std::uint64_t{1} << 52) - 1);
    const bool negative = (bits >> 63) != 0;
    if (biased == 0x7ff) {
      ++w_[fraction != 0 ? kExactLimbs + 2 : (negative ? kExactLimbs + 1 : kExactLimbs)];
      return;
    }
    if (biased == 0 && fraction == 0) return;
    const std::uint64_t m = biased == 0 ? fraction : (fraction | (std::uint64_t{1} << 52));
    const int p = (biased == 0 ? -1074 : biased - 1075) + 1088;
    const int i = p >> 5, s = p & 31;
    const std::uint64_t lo = (m & 0xffffffffULL) << s, hi = (m >> 32) << s;
    const long long sign = negative ? -1 : 1;
    w_[i] += sign * static_cast<long long>(lo & 0xffffffffULL);
    w_[i + 1] += sign * static_cast<long long>((lo >> 32) + (hi & 0xffffffffULL));
    w_[i + 2] += sign * static_cast<long long>(hi >> 32);
    // Each add moves a limb by less than 2^33: carry well before 2^63.
    if (++pending_ == (1LL << 29)) normalize();
  }
  [[nodiscard]] double value() const { return exact_value(w_); }

 private:
  void normalize() {
    for (int i = 0; i + 1 < kExactLimbs; ++i) {
      const long long carry = w_[i] >> 32;
      w_[i] -= carry * (1LL << 32);
      w_[i + 1] += carry;
    }
    pending_ = 0;
  }
  long long w_[kExactWords] = {};
  long long pending_ = 0;
};

}  // namespace mxe
