// The FP64 instance's division, reciprocal and square root on the host -- TEST HARNESS (built and run by
// tests/test_fp64_helpers_cpu.py).
//
// Instantiates p2r-tests_amd/csrc/p2r_lane_math.inc with real = double and P2R_DOUBLE_EST = 1, the helpers the
// library's FP64 instance (p2r::dbl) compiles, with the hardware estimates v_rcp_f64 / v_rsq_f64 replaced by
// the exact value times (1 + d): d is drawn per call from [-2^-E, 2^-E] (E from the command line; the
// hardware's estimates are documented to within 2^29 ulps, about 2^-23, closer than E = 20).  Prints, over
// N random operands of magnitudes 1e-30 to 1e30 and both signs, the largest error of m_rcp, m_div and m_sqrt
// in ulps of the IEEE result; then the special values "sqrt0 sqrtinf sqrtneg" (1 where m_sqrt(0) == 0,
// m_sqrt(inf) == inf, m_sqrt(-1) is NaN); then, over N arguments uniform in [-1e5, 1e5], N in [-4, 4] and
// the doubles nearest to k pi/2 for |k| <= 1000, the largest error of m_sincos's sine and cosine against
// libm's, in units of 2^-52 (absolute error; both lie in [-1, 1]), and "sinnan" (1 where m_sincos of
// NaN and of infinity gives NaN).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define __device__
#define __forceinline__ inline

#include "../p2r-tests_amd/csrc/p2r_types.hpp"

namespace est {

uint64_t& rng() {
  static uint64_t s = 0x9E3779B97F4A7C15ull;
  return s;
}
double uniform() {  // [0, 1), splitmix64
  uint64_t z = (rng() += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53;
}
int& est_bits() {
  static int e = 20;
  return e;
}
double perturb(double v) { return v * (1.0 + std::ldexp(2.0 * uniform() - 1.0, -est_bits())); }
// the estimates the FP64 instance refines
inline double rcp_est(double x) { return perturb(1.0 / x); }
inline double rsq_est(double x) { return perturb(1.0 / std::sqrt(x)); }
inline void opaque(double&) {}
inline void phase_boundary() {}

using real = double;
using Trk = p2r::TrkT<real>;
using Hit = p2r::HitT<real>;
using Jac = p2r::JacT<real>;
using p2r::kNiter;
using p2r::sym;
#define P2R_EXACT 0
#define P2R_DOUBLE 1
#define P2R_DOUBLE_EST 1
#include "../p2r-tests_amd/csrc/p2r_lane_math.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE
#undef P2R_DOUBLE_EST

}  // namespace est

static double ulps(double got, double want) {
  if (got == want) return 0.0;
  if (!std::isfinite(got) || !std::isfinite(want)) return INFINITY;
  const double u = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
  return std::fabs(got - want) / u;
}

int main(int argc, char** argv) {
  using namespace est;
  if (argc < 3) return 1;
  const long n = std::atol(argv[1]);
  est_bits() = std::atoi(argv[2]);
  double e_rcp = 0, e_div = 0, e_sqrt = 0;
  for (long i = 0; i < n; ++i) {
    const double a = std::pow(10.0, -30.0 + 60.0 * uniform()) * (uniform() < 0.5 ? -1.0 : 1.0);
    const double b = std::pow(10.0, -30.0 + 60.0 * uniform()) * (uniform() < 0.5 ? -1.0 : 1.0);
    e_rcp = std::fmax(e_rcp, ulps(m_rcp(b), 1.0 / b));
    e_div = std::fmax(e_div, ulps(m_div(a, b), a / b));
    e_sqrt = std::fmax(e_sqrt, ulps(m_sqrt(std::fabs(a)), std::sqrt(std::fabs(a))));
  }
  std::printf("%.3f %.3f %.3f\n", e_rcp, e_div, e_sqrt);
  std::printf("%d %d %d\n", m_sqrt(0.0) == 0.0 ? 1 : 0, m_sqrt((double)INFINITY) == (double)INFINITY ? 1 : 0,
              std::isnan(m_sqrt(-1.0)) ? 1 : 0);
  double e_sin = 0, e_cos = 0;
  auto check = [&](double x) {
    double sv, cv;
    m_sincos(x, sv, cv);
    e_sin = std::fmax(e_sin, std::fabs(sv - std::sin(x)) * 0x1.0p52);
    e_cos = std::fmax(e_cos, std::fabs(cv - std::cos(x)) * 0x1.0p52);
  };
  for (long i = 0; i < n; ++i) check(-1.0e5 + 2.0e5 * uniform());
  for (long i = 0; i < n; ++i) check(-4.0 + 8.0 * uniform());
  for (int k = -1000; k <= 1000; ++k) check((double)k * 1.5707963267948966);
  double sn, cn, si, ci;
  m_sincos((double)NAN, sn, cn);
  m_sincos((double)INFINITY, si, ci);
  std::printf("%.3f %.3f %d\n", e_sin, e_cos, std::isnan(sn) && std::isnan(cn) && std::isnan(si) && std::isnan(ci));
  return 0;
}
