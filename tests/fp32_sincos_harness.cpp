// The fast FP32 instance's sin/cos on the host -- TEST HARNESS (built and run by tests/test_fp32_sincos_cpu.py).
//
// Instantiates p2r-tests_amd/csrc/p2r_lane_math.inc with real = float as the library's fast instance
// (p2r::fast) does, the GPU-only estimates stood in for by their exact host values (they are not used by
// sin/cos), and compares its m_sincos with libm's double sin/cos over N arguments uniform in [-4, 4], N in
// [-100, 100], N in [-6000, 6000] and the floats nearest to k pi/2 for |k| <= 1000.  Prints the largest
// absolute error of the sine and of the cosine in units of 2^-24, and "nan" (1 where NaN and infinite
// arguments give NaN).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

#define __device__
#define __forceinline__ inline
#define __builtin_amdgcn_sqrtf(x) std::sqrt(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))

#include "../p2r-tests_amd/csrc/p2r_types.hpp"

namespace fst {

inline double rcp_est(double x) { return 1.0 / x; }
inline void opaque(float&) {}
inline void phase_boundary() {}

using real = float;
using Trk = p2r::TrkT<real>;
using Hit = p2r::HitT<real>;
using Jac = p2r::JacT<real>;
using p2r::kNiter;
using p2r::sym;
#define P2R_EXACT 0
#define P2R_DOUBLE 0
#include "../p2r-tests_amd/csrc/p2r_lane_math.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE

}  // namespace fst

int main(int argc, char** argv) {
  using namespace fst;
  if (argc < 2) return 1;
  const long n = std::atol(argv[1]);
  uint64_t st = 0x2545F4914F6CDD1Dull;
  auto uniform = [&] {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * 0x1.0p-53;
  };
  double e_sin = 0, e_cos = 0;
  auto check = [&](float x) {
    float sv, cv;
    m_sincos(x, sv, cv);
    e_sin = std::fmax(e_sin, std::fabs((double)sv - std::sin((double)x)) * 0x1.0p24);
    e_cos = std::fmax(e_cos, std::fabs((double)cv - std::cos((double)x)) * 0x1.0p24);
  };
  for (double span : {4.0, 100.0, 6000.0})
    for (long i = 0; i < n; ++i) check((float)(span * (2.0 * uniform() - 1.0)));
  for (int k = -1000; k <= 1000; ++k) check((float)(k * 1.5707963267948966));
  float sn, cn, si, ci;
  m_sincos(NAN, sn, cn);
  m_sincos(INFINITY, si, ci);
  std::printf("%.3f %.3f %d\n", e_sin, e_cos, std::isnan(sn) && std::isnan(cn) && std::isnan(si) && std::isnan(ci));
  return 0;
}
