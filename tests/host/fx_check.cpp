// Host check of the exact fixed-point energy sums (exachem_amd/csrc/xt_fixed.h):
// split values and add them in several orders with the kernels' carry rule,
// also through running sums over runs of partials (fx_add_local, as a
// workgroup sums its boxes of one task); the totals must be bit-identical
// across orders, equal to the exactly
// rounded sum, and non-finite or out-of-range partials must be refused.
// Built and run by tests/test_host_safety.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <vector>

#include "../../exachem_amd/csrc/xt_fixed.h"

static void add(unsigned long long* acc, unsigned long long lo, unsigned long long hi) {
  const unsigned long long old   = acc[0];
  acc[0]                         = old + lo;
  const unsigned long long carry = (acc[0] < old) ? 1ull : 0ull;
  acc[1] += hi + carry;
}

static int fails = 0;
#define CHECK(c)                                                     \
  do {                                                               \
    if(!(c)) {                                                       \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);               \
      fails++;                                                       \
    }                                                                \
  } while(0)

int main() {
  unsigned long long lo, hi;
  // refusals
  CHECK(!xt::fx_split(std::nan(""), lo, hi) && lo == 0 && hi == 0);
  CHECK(!xt::fx_split(std::numeric_limits<double>::infinity(), lo, hi));
  CHECK(!xt::fx_split(-std::numeric_limits<double>::infinity(), lo, hi));
  CHECK(!xt::fx_split(xt::kFxLimit, lo, hi));
  CHECK(!xt::fx_split(-xt::kFxLimit, lo, hi));
  CHECK(!xt::fx_split(1e300, lo, hi));
  CHECK(xt::fx_split(std::nextafter(xt::kFxLimit, 0.0), lo, hi));
  // single values come back as their value truncated toward zero at 2^-64,
  // to a few ulp
  const double vals[] = {0.0, -0.0, 1.0, -1.0, 0.5, -0.75, 3.25, -1234.5, 1e-19, -1e-19,
                         std::ldexp(1.0, -64), -std::ldexp(1.0, -64), std::ldexp(1.0, 61),
                         -std::ldexp(1.0, 61) - 1.0};
  for(double v: vals) {
    unsigned long long acc[2] = {0, 0};
    CHECK(xt::fx_split(v, lo, hi));
    add(acc, lo, hi);
    const double back = xt::fx_value(acc);
    const double want = (double) std::ldexp(std::truncl(std::ldexp((long double) v, 64)), -64);
    if(std::fabs(v) >= 1.0 || v == 0.0) CHECK(back == v);
    else CHECK(std::fabs(back - want) <= std::ldexp(std::fabs(want), -51));
    CHECK((v < 0) == (back < 0) || back == 0.0);
  }
  // order independence and exactness over many signed partials
  std::mt19937_64                        g(7);
  std::normal_distribution<double>       nd(0.0, 1.0);
  std::vector<double>                    p(20000);
  for(size_t i = 0; i < p.size(); i++) p[i] = nd(g) * std::pow(10.0, (double) (i % 9) - 4.0);
  unsigned long long ref[2] = {0, 0};
  long double        exact  = 0.0L;
  for(double v: p) {
    CHECK(xt::fx_split(v, lo, hi));
    add(ref, lo, hi);
    // the split truncates toward zero below 2^-64
    const long double t = std::ldexp(std::truncl(std::ldexp((long double) v, 64)), -64);
    exact += t;
  }
  for(int rep = 0; rep < 5; rep++) {
    std::shuffle(p.begin(), p.end(), g);
    unsigned long long acc[2] = {0, 0};
    for(double v: p) {
      xt::fx_split(v, lo, hi);
      add(acc, lo, hi);
    }
    CHECK(acc[0] == ref[0] && acc[1] == ref[1]);
  }
  // a workgroup's running sum over consecutive partials of one task
  // (fx_add_local), added once per run of partials, gives the same bits
  for(int rep = 0; rep < 5; rep++) {
    std::shuffle(p.begin(), p.end(), g);
    unsigned long long acc[2] = {0, 0};
    size_t             i      = 0;
    while(i < p.size()) {
      const size_t       run  = 1 + g() % 700;
      unsigned long long w[2] = {0, 0};
      for(size_t k = 0; k < run && i < p.size(); k++, i++) {
        xt::fx_split(p[i], lo, hi);
        xt::fx_add_local(w, lo, hi);
      }
      add(acc, w[0], w[1]);
    }
    CHECK(acc[0] == ref[0] && acc[1] == ref[1]);
  }
  const double got = xt::fx_value(ref);
  CHECK(std::fabs((long double) got - exact) <= 1e-15L * std::fabs(exact) + 1e-18L);
  std::printf("fx_check: %s (sum %.17g)\n", fails ? "FAILED" : "ok", got);
  return fails ? 1 : 0;
}
