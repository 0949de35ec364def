// printf's "%g" (6 significant digits) of a float, written without the
// general-purpose double formatter: the float's exact binary value m * 2^e is
// scaled by a power of ten in 128-bit integer arithmetic, rounded half to even
// on the exact value (as printf rounds), and laid out by %g's rules (fixed
// notation for decimal exponents -4..5, else d.ddddde+XX; trailing zeros and a
// bare point removed). Values outside [1e-20, 1e20), zeros' signs aside, and
// non-finite values go to std::to_chars' general format at precision 6, which
// is specified as printf's %.6g.
#ifndef IHE_HOST_GFMT_H
#define IHE_HOST_GFMT_H

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace ihe_host {

using u128 = unsigned __int128;

// 10^n for n in [0, 19]
constexpr uint64_t kPow10[20] = {1ull,
                                 10ull,
                                 100ull,
                                 1000ull,
                                 10000ull,
                                 100000ull,
                                 1000000ull,
                                 10000000ull,
                                 100000000ull,
                                 1000000000ull,
                                 10000000000ull,
                                 100000000000ull,
                                 1000000000000ull,
                                 10000000000000ull,
                                 100000000000000ull,
                                 1000000000000000ull,
                                 10000000000000000ull,
                                 100000000000000000ull,
                                 1000000000000000000ull,
                                 10000000000000000000ull};

// 10^n, n in [0, 38]
inline u128 pow10_u128(int n) {
  return n < 20 ? (u128)kPow10[n] : (u128)kPow10[19] * kPow10[n - 19];
}

// "000".."999" (and a spare byte), and the trailing zeros of 0..999 (3 for 0)
struct Digits3 {
  char s[3001];
  uint8_t tz[1000];
  constexpr Digits3() : s(), tz() {
    for (int i = 0; i < 1000; ++i) {
      s[3 * i] = (char)('0' + i / 100);
      s[3 * i + 1] = (char)('0' + i / 10 % 10);
      s[3 * i + 2] = (char)('0' + i % 10);
      tz[i] = (uint8_t)(i == 0 ? 3 : i % 100 == 0 ? 2 : i % 10 == 0 ? 1 : 0);
    }
  }
};
constexpr Digits3 kDigitTable{};
constexpr const char *kDigits3 = kDigitTable.s;
constexpr const uint8_t *kTrailingZeros3 = kDigitTable.tz;

// floor(a * 10^n) and whether the rest is <, == or > one half (-1, 0, 1), for
// a = m * 2^e; n is in [-16, 27] and the result below 10^7.
inline uint32_t scaled_floor(uint32_t m, int e, int n, int &half) {
  if (n >= 0 && n < 20 && e < 0) {  // m * 10^n / 2^-e: a shift
    const u128 num = (u128)m * kPow10[n];
    const int sh = -e;
    const u128 q = num >> sh, r = num - (q << sh), hv = (u128)1 << (sh - 1);
    half = r < hv ? -1 : r == hv ? 0 : 1;
    return (uint32_t)q;
  }
  u128 num, den;
  if (n >= 0) {
    num = (u128)m * pow10_u128(n);
    den = 1;
  } else {
    num = m;
    den = pow10_u128(-n);
  }
  if (e >= 0) num <<= e;
  else den <<= -e;
  const u128 q = num / den, r = num - q * den;
  const u128 twice = r << 1;
  half = twice < den ? -1 : twice == den ? 0 : 1;
  return (uint32_t)q;
}

// Writes %g of v at out (at most 16 characters, with up to 8 bytes past them
// scratched); returns the end.
inline char *format_g6(char *out, float v) {
  uint32_t bits;
  std::memcpy(&bits, &v, 4);
  const uint32_t biased = (bits >> 23) & 0xFF, frac = bits & 0x7FFFFF;
  if (biased == 0xFF || (biased == 0 && frac == 0) || !(std::fabs(v) >= 1e-20f && std::fabs(v) < 1e20f)) {
    if (biased == 0 && frac == 0) {  // +-0
      if (bits >> 31) *out++ = '-';
      *out++ = '0';
      return out;
    }
    return std::to_chars(out, out + 32, (double)v, std::chars_format::general, 6).ptr;
  }
  if (bits >> 31) *out++ = '-';
  // a = m * 2^e exactly (normal floats here: the range excludes subnormals)
  const uint32_t m = frac | 0x800000u;
  const int e = (int)biased - 150;
  // decimal exponent X: 10^5 <= a * 10^(5 - X) < 10^6; from the estimate
  // floor(b log10 2) for 2^b <= a < 2^(b+1) (78913 / 2^18 ~ log10 2), moved
  // until the scaled value has six digits
  const int b = (int)biased - 127;
  int X = b >= 0 ? (b * 78913) >> 18 : -((-b * 78913 + (1 << 18) - 1) >> 18);
  int half;
  uint32_t q = scaled_floor(m, e, 5 - X, half);
  while (q >= 1000000) {
    ++X;
    q = scaled_floor(m, e, 5 - X, half);
  }
  while (q < 100000) {
    --X;
    q = scaled_floor(m, e, 5 - X, half);
  }
  if (half > 0 || (half == 0 && (q & 1))) ++q;  // to nearest, ties to even
  if (q == 1000000) {
    q = 100000;
    ++X;
  }
  // the six digits, and how many remain once trailing zeros go
  char d[16];  // 8-byte copies below read past the six digits
  const uint32_t hi = q / 1000, lo = q - hi * 1000;
  std::memcpy(d, kDigits3 + 3 * hi, 4);
  std::memcpy(d + 3, kDigits3 + 3 * lo, 4);
  const int nd = lo ? 6 - kTrailingZeros3[lo] : 3 - kTrailingZeros3[hi];
  if (X >= -4 && X < 6) {
    if (X >= 0) {
      std::memcpy(out, d, 8);
      out += X + 1;
      if (nd > X + 1) {
        *out = '.';
        std::memcpy(out + 1, d + X + 1, 8);
        out += 1 + nd - (X + 1);
      }
    } else {
      std::memcpy(out, "0.000000", 8);
      out += 1 - X;
      std::memcpy(out, d, 8);
      out += nd;
    }
  } else {
    out[0] = d[0];
    out[1] = '.';
    std::memcpy(out + 2, d + 1, 8);
    out += nd > 1 ? 1 + nd : 1;
    out[0] = 'e';
    out[1] = X < 0 ? '-' : '+';
    const int ax = X < 0 ? -X : X;
    out[2] = (char)('0' + ax / 10);
    out[3] = (char)('0' + ax % 10);
    out += 4;
  }
  return out;
}

}  // namespace ihe_host

#endif  // IHE_HOST_GFMT_H
