// The product's per-lane forward and reverse bodies on the host, in FP64 -- TEST HARNESS (built and run by
// tests/test_adjoint_host_cpu.py).
//
// Instantiates p2r-tests_amd/csrc/p2r_lane_math.inc and p2r_adjoint.inc -- the code the GPU kernels run --
// with real = double, as the library's FP64 instance does (with IEEE division and square root and libm's
// sin/cos where that instance refines hardware estimates), and runs the chained fit of each track read
// from stdin (27 state floats, the charge, kfact, nlayer, then nlayer hits of 9 floats) forward with
// layer_forward_values, then backward with layer_adjoint.  Prints per track one line of the gradient of
// the total chi2: d/d(6 initial parameters), d/d(hit x, y, z) per layer, d/dkfact -- the layout of the
// oracle's forward-mode Dual gradient it is compared with.
#include <cmath>
#include <cstdio>
#include <vector>

#define __device__
#define __forceinline__ inline

#include "../p2r-tests_amd/csrc/p2r_types.hpp"

namespace hd {

// the kernels' optimizer barriers (no code on the GPU either)
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
#include "../p2r-tests_amd/csrc/p2r_lane_math.inc"
#include "../p2r-tests_amd/csrc/p2r_adjoint.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE

}  // namespace hd

int main() {
  using namespace hd;
  int ntrk = 0;
  if (std::scanf("%d", &ntrk) != 1) return 1;
  for (int t = 0; t < ntrk; ++t) {
    float v[27];
    int q, nl;
    double kf;
    for (float& x : v)
      if (std::scanf("%f", &x) != 1) return 1;
    if (std::scanf("%d %lf %d", &q, &kf, &nl) != 3 || nl < 1) return 1;
    std::vector<Hit> h(nl);
    for (int l = 0; l < nl; ++l) {
      float hv[9];
      for (float& x : hv)
        if (std::scanf("%f", &x) != 1) return 1;
      for (int i = 0; i < 3; ++i) h[l].pos[i] = hv[i];
      for (int i = 0; i < 6; ++i) h[l].cov[i] = hv[3 + i];
    }
    const real k = (double)(float)((float)q * (float)kf);  // k = q kfact in FP32, as the kernels form it
    std::vector<Trk> in(nl + 1);
    for (int i = 0; i < 6; ++i) in[0].par[i] = v[i];
    for (int i = 0; i < 21; ++i) in[0].cov[i] = v[6 + i];
    for (int l = 0; l < nl; ++l) {
      LayerFwd f;
      layer_forward_values(in[l], k, h[l], f, in[l + 1]);
    }
    Trk bar;
    for (int i = 0; i < 6; ++i) bar.par[i] = 0.0;
    for (int i = 0; i < 21; ++i) bar.cov[i] = 0.0;
    real kbar = 0.0;
    std::vector<double> ghit(3 * nl);
    for (int l = nl - 1; l >= 0; --l) {
      real hb[3];
      layer_adjoint(in[l], k, h[l], bar, kbar, hb);
      for (int c = 0; c < 3; ++c) ghit[3 * l + c] = hb[c];
    }
    for (int i = 0; i < 6; ++i) std::printf("%.17g ", bar.par[i]);
    for (double g : ghit) std::printf("%.17g ", g);
    std::printf("%.17g\n", (double)q * kbar);
  }
  return 0;
}
