// Host check of symphas_amd/csrc/decimal.hpp against the C library:
// format_e10 against snprintf("%.10E"), format_e<6> with the blank flag
// against snprintf("% 10lE") (the plot and CSV number format) and parse_fast
// against strtod.
//   decimal_check <count> <seed>
// Prints "checked N formatted F (fallback B) parsed P" and exits non-zero
// on the first mismatch.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../symphas_amd/csrc/decimal.hpp"

static uint64_t state;
static uint64_t next() {  // SplitMix64
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static long formatted = 0, fallback = 0, parsed = 0, checked = 0;

static bool check(double v) {
  ++checked;
  char want[64], got[32];
  std::snprintf(want, sizeof(want), "%.10E", v);
  const int n = sps::dec::format_e10(v, got);
  if (n == 0) {
    ++fallback;
  } else {
    ++formatted;
    if ((size_t)n != std::strlen(want) || std::memcmp(got, want, n) != 0) {
      std::printf("FORMAT MISMATCH %a: want %s got %.*s\n", v, want, n, got);
      return false;
    }
  }
  char want6[64], got6[32];
  std::snprintf(want6, sizeof(want6), "% 10lE", v);
  const int n6 = sps::dec::format_e<6>(v, got6, true);
  if (n6 != 0 &&
      ((size_t)n6 != std::strlen(want6) || std::memcmp(got6, want6, n6) != 0)) {
    std::printf("FORMAT6 MISMATCH %a: want '%s' got '%.*s'\n", v, want6, n6, got6);
    return false;
  }
  double r;
  if (sps::dec::parse_fast(want, (int)std::strlen(want), &r)) {
    ++parsed;
    const double w = std::strtod(want, nullptr);
    if (std::memcmp(&r, &w, 8) != 0) {
      std::printf("PARSE MISMATCH %s: want %a got %a\n", want, w, r);
      return false;
    }
  }
  return true;
}

static bool check_token(const char* tok) {
  double r;
  if (!sps::dec::parse_fast(tok, (int)std::strlen(tok), &r)) return true;
  char* end;
  const double w = std::strtod(tok, &end);
  if (*end != 0 || std::memcmp(&r, &w, 8) != 0) {
    std::printf("TOKEN MISMATCH %s: want %a got %a\n", tok, w, r);
    return false;
  }
  ++parsed;
  return true;
}

int main(int argc, char** argv) {
  const long count = argc > 1 ? std::atol(argv[1]) : 1000000;
  state = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
  // special and boundary values
  const double special[] = {0.0, -0.0, 1.0, -1.0, 0.1, 9.99999999995e-1,
                            9.999999999949999e-1, 9.99999999995e9, 1e10,
                            99999999999.5, 1.00000000005, 2.5e-16, 1e-300,
                            1e300, 1.7976931348623157e308, 2.2250738585072014e-308,
                            4.9e-324, INFINITY, -INFINITY, NAN, 123456789012345.0,
                            0.5, 0.25, 1.5, 1.25e-5, 6.103515625e-05};
  for (double v : special)
    if (!check(v)) return 1;
  for (long i = 0; i < count; ++i) {
    double v;
    const uint64_t r = next();
    switch (i % 4) {
      case 0: {  // any finite bit pattern
        uint64_t b = r;
        std::memcpy(&v, &b, 8);
        if (!std::isfinite(v)) continue;
        break;
      }
      case 1:  // phase-field magnitudes
        v = ((double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5) * 4.0;
        break;
      case 2: {  // values on an 11-digit decimal grid +- a few ulps (ties)
        const double base = (double)(r % 90000000000ull + 10000000000ull) + 0.5;
        const int e = (int)((r >> 40) % 40) - 20;
        v = base * std::pow(10.0, (double)(e - 10));
        uint64_t b;
        std::memcpy(&b, &v, 8);
        b += (int64_t)((r >> 56) % 5) - 2;
        std::memcpy(&v, &b, 8);
        break;
      }
      default: {  // small integers and dyadic fractions (exact ties exist)
        v = (double)(int64_t)(r % 2000001) / (double)(1ull << ((r >> 32) % 40));
        break;
      }
    }
    if (!check(v)) return 1;
  }
  const char* tokens[] = {"0", "-0", "1", "1.", ".5", "-.5e3", "12345678901234",
                          "1234567890123456", "1e22", "1e23", "1E-22", "1e-23",
                          "3.1415926535E+00", "-2.7182818285E-05", "0.000001",
                          "+7", "1e", "e5", "--1", "1.2.3", "INF", "NAN"};
  for (const char* t : tokens)
    if (!check_token(t)) return 1;
  std::printf("checked %ld formatted %ld (fallback %ld) parsed %ld\n", checked,
              formatted, fallback, parsed);
  return 0;
}
