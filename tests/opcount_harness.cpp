// Operation count of the product's per-lane bodies, on the host -- TEST HARNESS (built and run by
// tests/test_adjoint_opcount_cpu.py).
//
// Instantiates p2r-tests_amd/csrc/p2r_lane_math.inc and p2r_adjoint.inc -- the code the GPU kernels run --
// with a scalar type that counts every arithmetic operation, and prints for each track read from stdin
// (27 state floats, the charge, kfact, nlayer, then nlayer hits of 9 floats) and each layer of its chained
// fit one line "fwd values adjoint cond iter upd_out" of operation totals:
//   fwd      propagate_to_r + kalman_update (the forward update; its chi2 included)
//   values   layer_forward_values (the layer's forward values and its updated state)
//   adjoint  layer_adjoint (the forward values it needs plus the reverse sweep proper)
//   cond     fixed-point iterations whose convergence branch the reverse took (0..4)
//   iter     iteration_values of the 5 iterations (what the reverse of the iterations recomputes)
//   upd_out  update_out (the updated state, which layer_adjoint does not compute)
// A total counts add/sub, mul, div, sqrt, sin and cos each, and the 4 FP64 operations of the update's
// determinant and its reciprocal; negation, comparison and selection are not counted.
#include <cmath>
#include <cstdio>

#define __device__
#define __forceinline__ inline

#include "../p2r-tests_amd/csrc/p2r_types.hpp"

namespace cnt {

long& ops() {
  static long n = 0;
  return n;
}

struct Cnt {
  double v = 0.0;
  Cnt() = default;
  Cnt(double x) : v(x) {}
  explicit operator double() const { return v; }
};
inline Cnt operator+(Cnt a, Cnt b) { ++ops(); return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++ops(); return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++ops(); return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++ops(); return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { return a = a + b; }
inline Cnt& operator-=(Cnt& a, Cnt b) { return a = a - b; }
inline Cnt& operator*=(Cnt& a, Cnt b) { return a = a * b; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
// the kernels' optimizer barriers (no code on the GPU either)
inline void opaque(Cnt&) {}
inline void phase_boundary() {}
inline Cnt sqrt(Cnt x) { ++ops(); return Cnt(std::sqrt(x.v)); }
inline Cnt floor(Cnt x) { return Cnt(std::floor(x.v)); }
inline void sincos(Cnt x, Cnt* s, Cnt* c) {
  ops() += 2;
  *s = Cnt(std::sin(x.v));
  *c = Cnt(std::cos(x.v));
}

using real = Cnt;
using Trk = p2r::TrkT<real>;
using Hit = p2r::HitT<real>;
using Jac = p2r::JacT<real>;
using p2r::kNiter;
using p2r::sym;

// The generic instance of the per-lane bodies (the one the kernels build for double).  Its FP64
// determinant is computed on doubles converted from `real`, which the counter does not see: the
// 4 operations are added per update below.
#define P2R_EXACT 0
#define P2R_DOUBLE 1
#include "../p2r-tests_amd/csrc/p2r_lane_math.inc"
#include "../p2r-tests_amd/csrc/p2r_adjoint.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE
constexpr long kDetOps = 4;

}  // namespace cnt

int main() {
  using namespace cnt;
  int ntrk = 0;
  if (std::scanf("%d", &ntrk) != 1) return 1;
  for (int t = 0; t < ntrk; ++t) {
    float v[27];
    int q, nl;
    double kf;
    for (float& x : v)
      if (std::scanf("%f", &x) != 1) return 1;
    if (std::scanf("%d %lf %d", &q, &kf, &nl) != 3) return 1;
    Trk in;
    for (int i = 0; i < 6; ++i) in.par[i] = v[i];
    for (int i = 0; i < 21; ++i) in.cov[i] = v[6 + i];
    const real k = (double)q * kf;
    for (int l = 0; l < nl; ++l) {
      float hv[9];
      for (float& x : hv)
        if (std::scanf("%f", &x) != 1) return 1;
      Hit h;
      for (int i = 0; i < 3; ++i) h.pos[i] = hv[i];
      for (int i = 0; i < 6; ++i) h.cov[i] = hv[3 + i];
      ops() = 0;
      Trk prop;
      propagate_to_r(in, k, h.pos[0], h.pos[1], prop);
      kalman_update(prop, h);
      const long fwd = ops() + kDetOps;
      ops() = 0;
      LayerFwd f;
      Trk out;
      layer_forward_values(in, k, h, f, out);
      const long values = ops() + kDetOps;
      int cond = 0;
      for (int i = 0; i + 1 < kNiter; ++i) cond += f.x.cond[i] ? 1 : 0;
      ops() = 0;
      for (int i = 0; i < kNiter; ++i) {
        IterFwd it;
        iteration_values(f.x.xo[i], f.x.yo[i], f.r, in.par[3], f.x.kinv, i + 1 == kNiter, it);
      }
      const long iter = ops();
      ops() = 0;
      update_out(f.prop, f.u, out);
      const long upd_out = ops();
      ops() = 0;
      Trk bar;
      for (int i = 0; i < 6; ++i) bar.par[i] = 0.0;
      for (int i = 0; i < 21; ++i) bar.cov[i] = 0.0;
      real kbar = 0.0, hbar[3];
      layer_adjoint(in, k, h, bar, kbar, hbar);
      const long adjoint = ops() + kDetOps;
      std::printf("%ld %ld %ld %d %ld %ld\n", fwd, values, adjoint, cond, iter, upd_out);
      in = prop;  // the chained fit: the next layer starts from this layer's update
    }
  }
  return 0;
}
