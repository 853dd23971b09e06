// decimal.hpp - exact decimal text of doubles in the reference's checkpoint
// number format, and the fast exact path of the inverse, for host and device.
//
// The reference writes checkpoint values with fprintf("%.10E ")
// (io/src/writegp.cpp:780-803, io/inc/io.h CHECKPOINT_OUTPUT_ACCURACY 10,
// format "E") and reads them back with fscanf("%lf") (io/src/readgp.cpp).
// Both C library routines are correctly rounded, so the text of a value is a
// function of the value alone: 11 significant digits, round half to even on
// the exact binary value.
//
// format_e<P> produces that text (and the "% 10lE" = "% .6E" text of the
// plot and CSV files) without a big-integer conversion: with E the decimal
// exponent, t = |v| * 10^(P-E) is computed in double arithmetic from
// correctly rounded powers of ten (at most three roundings, so |t - t*| <=
// 3.3e-16 * t < 3.3e-5 for t < 1e11).  The P+1 digits are round(t*); when the
// fraction of t is farther than 1e-4 from 1/2, round(t) = round(t*) and the
// result is exact.  Values closer to a tie, subnormals, infinities and NaNs are
// reported as not handled; the caller formats those few with the C library.
//
// parse_fast is Clinger's fast path: a decimal with at most 15 significant
// digits m and a power-of-ten exponent |e| <= 22 is m * 10^e (or m / 10^-e)
// with both operands exact doubles, so the one IEEE operation rounds
// correctly.  Other tokens are reported as not handled.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sps {
namespace dec {

__host__ __device__ constexpr double pow10_exact_small(int k) {
  // 10^0 .. 10^22 are exact doubles
  constexpr double p[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,
                            1e8,  1e9,  1e10, 1e11, 1e12, 1e13, 1e14, 1e15,
                            1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  return p[k];
}

// 10^k for -298 <= k <= 300 as 10^(16q) * 10^r: a correctly rounded table
// entry (the compiler rounds the literals) times an exact small power, so the
// relative error is at most 2^-52.
__host__ __device__ inline double pow10_approx(int k) {
  constexpr double big[39] = {
      1e-304, 1e-288, 1e-272, 1e-256, 1e-240, 1e-224, 1e-208, 1e-192,
      1e-176, 1e-160, 1e-144, 1e-128, 1e-112, 1e-96,  1e-80,  1e-64,
      1e-48,  1e-32,  1e-16,  1e0,    1e16,   1e32,   1e48,   1e64,
      1e80,   1e96,   1e112,  1e128,  1e144,  1e160,  1e176,  1e192,
      1e208,  1e224,  1e240,  1e256,  1e272,  1e288,  1e304};
  // k = 16*q + r with r in [0, 16)
  int q = k >= 0 ? k / 16 : -((-k + 15) / 16);
  int r = k - 16 * q;
  return big[q + 19] * pow10_exact_small(r);
}

// Text of v as fprintf("%.<P>E", v) writes it (fprintf("% .<P>E") when
// space is set: a blank where a non-negative value has no sign), into out (at
// most P + 8 chars, no terminator).  Returns the length, or 0 when the value
// must be formatted by the C library (near a rounding tie, subnormal,
// infinite or NaN).  1 <= P <= 14.
template <int P>
__host__ __device__ inline int format_e(double v, char* out, bool space = false) {
  static_assert(P >= 1 && P <= 14, "digits after the point");
  constexpr double lo = pow10_exact_small(P) - 0.5 - 1e-4;       // t below: E too big
  constexpr double hi = pow10_exact_small(P + 1) - 0.5 + 1e-4;   // t at or above: E too small
  uint64_t bits;
  __builtin_memcpy(&bits, &v, 8);
  const bool neg = (bits >> 63) != 0;
  const uint64_t mag = bits & 0x7fffffffffffffffull;
  int n = 0;
  if (neg) out[n++] = '-';
  else if (space) out[n++] = ' ';
  if (mag == 0) {
    out[n++] = '0';
    out[n++] = '.';
    for (int i = 0; i < P; ++i) out[n++] = '0';
    out[n++] = 'E';
    out[n++] = '+';
    out[n++] = '0';
    out[n++] = '0';
    return n;
  }
  const int bexp = (int)(mag >> 52);
  if (bexp == 0 || bexp == 0x7ff) return 0;  // subnormal, inf, nan
  double a;
  __builtin_memcpy(&a, &mag, 8);
  // decimal exponent estimate from the binary one, floor(log10(2) * e2): the
  // true exponent is that or one more, and rounding to P+1 digits can carry
  // once more, so two straight-line corrections suffice (no loop whose exit
  // values are used later)
  int E = (int)__builtin_floor((bexp - 1023) * 0.30102999566398120);
  auto scaled = [a](int Ex) {
    const int j = P - Ex;
    return j > 300 ? (a * 1e300) * pow10_approx(j - 300) : a * pow10_approx(j);
  };
  double t = scaled(E);
  if (t < lo) {
    E -= 1;
    t = scaled(E);
  } else if (t >= hi) {
    E += 1;
    t = scaled(E);
  }
  if (t < lo) {
    E -= 1;
    t = scaled(E);
  } else if (t >= hi) {
    E += 1;
    t = scaled(E);
  }
  if (!(t >= lo && t < hi)) return 0;
  const double d0 = __builtin_floor(t);
  const double f = t - d0;  // exact: t < 2^53
  if (__builtin_fabs(f - 0.5) < 1e-4) return 0;  // too close to a tie
  uint64_t d = (uint64_t)d0 + (f > 0.5 ? 1 : 0);
  const uint64_t dlo = (uint64_t)pow10_exact_small(P);
  if (d < dlo || d >= 10 * dlo) return 0;
  // digits
  char dig[P + 1];
  for (int i = P; i >= 0; --i) {
    dig[i] = (char)('0' + (int)(d % 10));
    d /= 10;
  }
  out[n++] = dig[0];
  out[n++] = '.';
  for (int i = 1; i <= P; ++i) out[n++] = dig[i];
  out[n++] = 'E';
  out[n++] = E < 0 ? '-' : '+';
  int ae = E < 0 ? -E : E;
  if (ae >= 100) {
    out[n++] = (char)('0' + ae / 100);
    ae %= 100;
  }
  out[n++] = (char)('0' + ae / 10);
  out[n++] = (char)('0' + ae % 10);
  return n;
}

// The checkpoint number format, "%.10E".
__host__ __device__ inline int format_e10(double v, char* out) {
  return format_e<10>(v, out, false);
}

__host__ __device__ inline bool is_space(char c) {
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' ||
         c == '\f';
}

// Parse the token s[0..len) as strtod would, when it is a plain decimal
// ([+-] digits [. digits] [(e|E) [+-] digits]) with at most 15 significant
// digits and a decimal exponent in [-22, 22] after scaling.  Returns false
// otherwise (the caller uses the C library).
__host__ __device__ inline bool parse_fast(const char* s, int len, double* out) {
  // written without loops that exit early into code reading their index:
  // the exponent marker is found by a full reverse scan, and the mantissa
  // and exponent are parsed over fixed ranges
  if (len < 1 || len > 40) return false;
  const bool neg = s[0] == '-';
  const int i0 = (s[0] == '+' || s[0] == '-') ? 1 : 0;
  int epos = len;  // first 'e' or 'E', or len
  for (int j = len - 1; j >= i0; --j)
    if (s[j] == 'e' || s[j] == 'E') epos = j;
  // mantissa [i0, epos): digits with at most one '.'
  uint64_t m = 0;
  int sig = 0, frac = 0, ndig = 0, points = 0;
  for (int j = i0; j < epos; ++j) {
    const char c = s[j];
    if (c == '.') {
      ++points;
      continue;
    }
    if (c < '0' || c > '9') return false;
    ++ndig;
    if (points) ++frac;
    if (m == 0 && c == '0') continue;  // leading zeros
    if (++sig > 15) return false;
    m = m * 10 + (uint64_t)(c - '0');
  }
  if (ndig == 0 || points > 1) return false;
  // exponent (epos, len): [+-]? and 1..4 digits
  int e = 0;
  if (epos < len) {
    int j = epos + 1;
    int sign = 1;
    if (j < len && (s[j] == '+' || s[j] == '-')) {
      if (s[j] == '-') sign = -1;
      ++j;
    }
    const int nd = len - j;
    if (nd < 1 || nd > 4) return false;
    for (int q = 0; q < 4; ++q) {
      if (q < nd) {
        const int c = (int)s[j + q] - '0';
        if (c < 0 || c > 9) return false;
        e = e * 10 + c;
      }
    }
    e *= sign;
  }
  double r;
  if (m == 0) {
    r = 0.0;
  } else {
    const int k = e - frac;
    if (k < -22 || k > 22) return false;
    r = k >= 0 ? (double)m * pow10_exact_small(k)
               : (double)m / pow10_exact_small(-k);
  }
  *out = neg ? -r : r;
  return true;
}

}  // namespace dec
}  // namespace sps
