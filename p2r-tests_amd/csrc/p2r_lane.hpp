// Per-lane p2r math for gfx950: one track per lane, all state in VGPRs.
//
// The arithmetic is the reference's (cerati/p2r-tests,
// src/propagate-tor-test_tbb.cpp): helix propagation to the hit radius with 5
// fixed-point steps and a 4th-order Taylor sin/cos of the turning angle
// (tbb.cpp:656-808), covariance transport P' = J P J^T (tbb.cpp:296-373) and
// the tangent-plane Kalman update with an FP64 2x2 determinant
// (tbb.cpp:433-642).  Formulas and their evaluation order follow those lines
// so results agree with the reference's FP32 CPU path to a few ulps (the
// exact instance rounds every operation as the FP32 source does; the fast
// instance fuses multiply-adds and uses the hardware reciprocal and sqrt).
//
// Differences in form, not in value: the transport skips the Jacobian's
// structural zeros and ones, including row 3 (a unit row) which the reference
// multiplies out generically (tbb.cpp:322-327, 357-360); for finite inputs the
// result is identical.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/p2r.h"
#include "p2r_types.hpp"

namespace p2r {

using Trk = TrkT<float>;
using Hit = HitT<float>;

// Keeps `v` opaque to the optimizer (no code).  Used so values that are
// loop-invariant only because of reference semantics are recomputed per layer
// rather than hoisted: each layer then does the full work it is credited with.
__device__ __forceinline__ void opaque(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque(double& v) { asm volatile("" : "+v"(v)); }
// No instruction is scheduled across this point (code size 0): separates phases whose values should not
// be live at the same time.
__device__ __forceinline__ void phase_boundary() { __builtin_amdgcn_sched_barrier(0); }
__device__ __forceinline__ void opaque(Trk& t) {
#pragma unroll
  for (int i = 0; i < 6; ++i) opaque(t.par[i]);
#pragma unroll
  for (int i = 0; i < 21; ++i) opaque(t.cov[i]);
}

template <class A, class B> __device__ __forceinline__ void convert(const TrkT<A>& a, TrkT<B>& b) {
#pragma unroll
  for (int i = 0; i < 6; ++i) b.par[i] = (B)a.par[i];
#pragma unroll
  for (int i = 0; i < 21; ++i) b.cov[i] = (B)a.cov[i];
}
template <class A, class B> __device__ __forceinline__ void convert(const HitT<A>& a, HitT<B>& b) {
#pragma unroll
  for (int i = 0; i < 3; ++i) b.pos[i] = (B)a.pos[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) b.cov[i] = (B)a.cov[i];
}

// FP64 hardware estimates of 1/x and 1/sqrt(x) (v_rcp_f64, v_rsq_f64), refined by Newton steps in the
// FP64 instance's division and square root and in the fast instance's FP64 reciprocal.
__device__ __forceinline__ double rcp_est(double x) { return __builtin_amdgcn_rcp(x); }
__device__ __forceinline__ double rsq_est(double x) { return __builtin_amdgcn_rsq(x); }

// Instances of the per-lane math (see p2r_lane_math.inc) and of its adjoint
// (p2r_adjoint.inc).
namespace exact {
#pragma clang fp contract(off)
using real = float;
using Trk = TrkT<real>;
using Hit = HitT<real>;
using Jac = JacT<real>;
#define P2R_EXACT 1
#define P2R_DOUBLE 0
#include "p2r_lane_math.inc"
#include "p2r_adjoint.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE
}  // namespace exact

namespace fast {
#pragma clang fp contract(fast)
using real = float;
using Trk = TrkT<real>;
using Hit = HitT<real>;
using Jac = JacT<real>;
#define P2R_EXACT 0
#define P2R_DOUBLE 0
#include "p2r_lane_math.inc"
#include "p2r_adjoint.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE
}  // namespace fast

namespace dbl {
#pragma clang fp contract(fast)
using real = double;
using Trk = TrkT<real>;
using Hit = HitT<real>;
using Jac = JacT<real>;
#define P2R_EXACT 0
#define P2R_DOUBLE 1
#define P2R_DOUBLE_EST 1
#include "p2r_lane_math.inc"
#include "p2r_adjoint.inc"
#undef P2R_EXACT
#undef P2R_DOUBLE
#undef P2R_DOUBLE_EST
}  // namespace dbl

// Compile-time selection of the instance.
template <bool EXACT> struct Lane;
template <> struct Lane<true> {
  static __device__ __forceinline__ void propagate(const Trk& in, float k, float hx, float hy, Trk& out) {
    exact::propagate_to_r(in, k, hx, hy, out);
  }
  template <class PC>
  static __device__ __forceinline__ void propagate(const float (&par)[6], const PC& P, float k, float hx, float hy,
                                                   Trk& out) {
    exact::propagate_to_r_c(par, P, k, hx, hy, out);
  }
  static __device__ __forceinline__ float update(Trk& t, const Hit& h) { return exact::kalman_update(t, h); }
  static __device__ __forceinline__ float wrap_phi(float phi) { return exact::m_wrap_phi(phi); }
};
template <> struct Lane<false> {
  static __device__ __forceinline__ void propagate(const Trk& in, float k, float hx, float hy, Trk& out) {
    fast::propagate_to_r(in, k, hx, hy, out);
  }
  template <class PC>
  static __device__ __forceinline__ void propagate(const float (&par)[6], const PC& P, float k, float hx, float hy,
                                                   Trk& out) {
    fast::propagate_to_r_c(par, P, k, hx, hy, out);
  }
  static __device__ __forceinline__ float update(Trk& t, const Hit& h) { return fast::kalman_update(t, h); }
  static __device__ __forceinline__ float wrap_phi(float phi) { return fast::m_wrap_phi(phi); }
};

// The options of p2r_params applied to a layer's updated state (tbb.cpp:527, :635): phi wrapped; a
// negative 1/pt flips its sign and the charge, and k = q kfact follows the charge.
template <bool EXACT>
__device__ __forceinline__ void apply_options(int opts, Trk& t, int& q, float& k, float kfact) {
  if (opts & P2R_OPT_WRAP_PHI) t.par[4] = Lane<EXACT>::wrap_phi(t.par[4]);
  if ((opts & P2R_OPT_CHARGE_FROM_IPT) && t.par[3] < 0.f) {
    t.par[3] = -t.par[3];
    q = -q;
    k = (float)q * kfact;
  }
}

}  // namespace p2r
