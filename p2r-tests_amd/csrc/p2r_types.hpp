// Per-lane data types of the p2r math: one track's state, one hit, the non-trivial Jacobian entries.
// Free of HIP headers so that host code can instantiate the per-lane bodies (p2r_lane_math.inc,
// p2r_adjoint.inc) with another scalar type.
#pragma once

#if defined(__HIPCC__)
#define P2R_HD __host__ __device__
#else
#define P2R_HD
#endif

namespace p2r {

constexpr int kNiter = 5;  // fixed-point steps of propagateToR, tbb.cpp:655

template <class R> struct TrkT {
  R par[6];   // x, y, z, 1/pt, phi, theta
  R cov[21];  // packed lower-triangular symmetric
};
template <class R> struct HitT {
  R pos[3];
  R cov[6];  // xx, yx, yy, zx, zy, zz
};

// Index of (i, j) in the packed lower-triangular storage (tbb.cpp:42-54).
P2R_HD constexpr int sym(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// Non-trivial entries of the propagation Jacobian (errorProp, tbb.cpp:753-807).
// Rows 3 and 5 are unit rows; J22 = 1; J02 = J05 = J12 = J15 = J42 = J45 = 0.
template <class R> struct JacT {
  R j00, j01, j03, j04;
  R j10, j11, j13, j14;
  R j20, j21, j23, j24, j25;
  R j40, j41, j43, j44;
};

}  // namespace p2r
