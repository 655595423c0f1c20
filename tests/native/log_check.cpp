// Accuracy of pf::pf_log against a long-double logarithm (x87 80-bit, 64-bit
// mantissa).  Prints: max_ulp frac_correctly_rounded n
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include "pf_log.hpp"

static double ulp_of(double y) {
  const double a = std::fabs(y);
  if (a == 0.0) return std::ldexp(1.0, -1074);
  int e;
  std::frexp(a, &e);
  return std::ldexp(1.0, e - 53);
}

int main() {
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return st;
  };
  double max_ulp = 0.0;
  long n = 0, exact = 0;
  auto check = [&](double x) {
    const double y = pf::pf_log(x);
    const long double ref = logl((long double)x);
    const double err = (double)std::fabs((long double)y - ref) / ulp_of((double)ref);
    if (err > max_ulp) max_ulp = err;
    if (y == (double)ref) ++exact;
    ++n;
  };
  // whole normal range, log-uniform
  for (int i = 0; i < 2000000; ++i) {
    uint64_t b = next();
    uint64_t ex = 1 + (b >> 11) % 2046;
    uint64_t bits = (ex << 52) | (next() & 0xFFFFFFFFFFFFFull);
    double x; std::memcpy(&x, &bits, 8);
    check(x);
  }
  // the chemical-potential domain: x and 1 - x for x in [1e-12, 1 - 1e-12]
  for (int i = 0; i < 2000000; ++i) {
    double u = (double)(next() >> 11) * 0x1.0p-53;
    double x = 1e-12 + u * (1.0 - 2e-12);
    check(x);
    check(1.0 - x);
  }
  // close to 1 on both sides, and the reduction boundary 1/sqrt2
  for (int k = 1; k <= 52; ++k) {
    for (int j = -64; j <= 64; ++j) {
      check(1.0 + std::ldexp((double)j, -k - 6));
      check(std::nextafter(0.70710678118654752440, 0.0) + std::ldexp((double)j, -60));
    }
  }
  check(1.0); check(2.0); check(0.5); check(1e-12); check(1.0 - 1e-12);
  std::printf("%.4f %.6f %ld\n", max_ulp, (double)exact / (double)n, n);
  // special values route to the library
  const bool sp = std::isnan(pf::pf_log(-1.0)) && std::isinf(pf::pf_log(0.0)) && pf::pf_log(0.0) < 0 &&
                  std::isinf(pf::pf_log(INFINITY)) && std::isnan(pf::pf_log(NAN)) &&
                  pf::pf_log(1.0) == 0.0 && std::fabs(pf::pf_log(4.9406564584124654e-324) + 744.44007192138126) < 1e-12;
  std::printf("%d\n", sp ? 1 : 0);
  return 0;
}
