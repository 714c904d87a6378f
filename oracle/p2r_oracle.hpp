// p2r oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference p2r algorithm (cerati/p2r-tests,
// src/propagate-tor-test_tbb.cpp) used as the checker for the HIP product
// path.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it; the product library never links, calls or falls back to it.
//
// Every per-lane routine is a template over the scalar type T so that the
// same operation sequence runs as
//   float            bit-level restatement of the reference's FP32 path,
//   double           FP64 replay of the same algorithm (same constants),
//   OpCount          op counter that freezes the FLOP-per-update figure,
//   Dual<ND>         forward-mode derivatives (reference gradient for the
//                    hand-written adjoint of the product).
// The operation order (including C++'s left-to-right evaluation of sums and
// products) follows tbb.cpp line by line; each routine cites the lines.
//
// Pinning (what checks this file): see tests/test_oracle_*.py -- bit-exact
// against the reference's own compiled propagateToR / KalmanUpdate /
// MultHelixProp[Transp] (built from src/propagate-tor-test_acc.v3.cpp, whose
// routines are arithmetically identical to tbb.cpp's apart from the kfact
// constant), bit-exact generator against src/propagate-tor-test_pstl.cpp's
// prepareTracks/prepareHits, the reference program's printed summary, and
// closed forms (textbook Kalman update in FP64, finite-difference Jacobian,
// glibc rand() prefix, packed-symmetric index maps).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <type_traits>

namespace orc {

constexpr int kNiter = 5;  // tbb.cpp:655 (fixed-point iterations, part of the algorithm)

// ---------------------------------------------------------------------------
// Packed lower-triangular symmetric storage: element (i,j), i>=j, lives at
// i*(i+1)/2 + j.  Same map as tbb.cpp:51-54 (SymOffsets66) and :46-49.
inline constexpr int sym6(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
inline constexpr int sym3(int i, int j) { return sym6(i, j); }

// ---------------------------------------------------------------------------
// Scalar helpers, overloaded per scalar type.
inline float o_sqrt(float x) { return std::sqrt(x); }
inline double o_sqrt(double x) { return std::sqrt(x); }
// Trig of the FP32 path: 0 = libm cosf/sinf as the reference calls them
// (tbb.cpp:689, :768-769); 1 = the FP64 value rounded to float (correctly
// rounded), the variant the GPU's exact arithmetic is checked against
// bit-for-bit.  glibc's sinf/cosf differ from the correctly rounded value on
// ~0.1% of arguments; nothing else differs between the two modes.
inline int& trig_mode() { static int m = 0; return m; }
inline float o_cos(float x) { return trig_mode() ? (float)std::cos((double)x) : std::cos(x); }
inline double o_cos(double x) { return std::cos(x); }
inline float o_sin(float x) { return trig_mode() ? (float)std::sin((double)x) : std::sin(x); }
inline double o_sin(double x) { return std::sin(x); }
inline double o_val(float x) { return x; }
inline double o_val(double x) { return x; }

// Wide<T>: the type the reference uses for the 2x2 determinant (tbb.cpp:457
// computes it in double even on the FP32 path).
template <class T> struct Wide { using type = T; };
template <> struct Wide<float> { using type = double; };

// ---------------------------------------------------------------------------
// OpCount: counts floating-point operations of the algorithm.  Values that are
// structural constants of the algorithm (the 0 and 1 entries of the Jacobian)
// are tracked so that multiplications/additions with them can be reported
// separately: `ops` counts every op the reference source performs, `useful`
// excludes ops whose operand is a structural 0 or 1.
struct OpTally {
  long long add = 0, mul = 0, div = 0, sqrt = 0, trig = 0, trivial = 0;
  long long total() const { return add + mul + div + sqrt + trig; }
  long long useful() const { return total() - trivial; }
};
inline OpTally& tally() { static thread_local OpTally t; return t; }

struct OpCount {
  double v = 0;
  signed char structural = -1;  // -1: ordinary value, 0/1: structural constant 0/1
  OpCount() = default;
  OpCount(double x) : v(x) {}
  static OpCount S(double x) { OpCount o(x); o.structural = (x == 0.0 ? 0 : (x == 1.0 ? 1 : -1)); return o; }
};
// Arithmetic propagates structural constants: x*0 -> structural 0, x*1 -> x,
// x+0 -> x; such ops are counted in `trivial` as well as in their class.
inline OpCount operator+(OpCount a, OpCount b) {
  auto& t = tally(); t.add++;
  if (a.structural == 0) { t.trivial++; return b; }
  if (b.structural == 0) { t.trivial++; return a; }
  return OpCount(a.v + b.v);
}
inline OpCount operator-(OpCount a, OpCount b) {
  auto& t = tally(); t.add++;
  if (b.structural == 0) { t.trivial++; return a; }
  return OpCount(a.v - b.v);
}
inline OpCount operator*(OpCount a, OpCount b) {
  auto& t = tally(); t.mul++;
  if (a.structural == 0 || b.structural == 0) { t.trivial++; return OpCount::S(0.0); }
  if (a.structural == 1) { t.trivial++; return b; }
  if (b.structural == 1) { t.trivial++; return a; }
  return OpCount(a.v * b.v);
}
inline OpCount operator/(OpCount a, OpCount b) { tally().div++; return OpCount(a.v / b.v); }
inline OpCount operator-(OpCount a) { OpCount r(-a.v); if (a.structural == 0) r.structural = 0; return r; }
inline OpCount& operator+=(OpCount& a, OpCount b) { a = a + b; return a; }
inline OpCount& operator-=(OpCount& a, OpCount b) { a = a - b; return a; }
inline OpCount& operator*=(OpCount& a, OpCount b) { a = a * b; return a; }
inline OpCount o_sqrt(OpCount x) { tally().sqrt++; return OpCount(std::sqrt(x.v)); }
inline OpCount o_cos(OpCount x) { tally().trig++; return OpCount(std::cos(x.v)); }
inline OpCount o_sin(OpCount x) { tally().trig++; return OpCount(std::sin(x.v)); }
inline double o_val(const OpCount& x) { return x.v; }

// ---------------------------------------------------------------------------
// Dual<ND>: forward-mode AD with ND tangent directions, FP64.
template <int ND> struct Dual {
  double v = 0;
  double d[ND] = {};
  Dual() = default;
  Dual(double x) : v(x) {}
};
template <int ND> inline Dual<ND> operator+(const Dual<ND>& a, const Dual<ND>& b) { Dual<ND> r(a.v + b.v); for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int ND> inline Dual<ND> operator-(const Dual<ND>& a, const Dual<ND>& b) { Dual<ND> r(a.v - b.v); for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int ND> inline Dual<ND> operator*(const Dual<ND>& a, const Dual<ND>& b) { Dual<ND> r(a.v * b.v); for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int ND> inline Dual<ND> operator/(const Dual<ND>& a, const Dual<ND>& b) { Dual<ND> r(a.v / b.v); const double ib = 1.0 / b.v; for (int i = 0; i < ND; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib; return r; }
template <int ND> inline Dual<ND> operator-(const Dual<ND>& a) { Dual<ND> r(-a.v); for (int i = 0; i < ND; ++i) r.d[i] = -a.d[i]; return r; }
template <int ND> inline Dual<ND>& operator+=(Dual<ND>& a, const Dual<ND>& b) { a = a + b; return a; }
template <int ND> inline Dual<ND>& operator-=(Dual<ND>& a, const Dual<ND>& b) { a = a - b; return a; }
template <int ND> inline Dual<ND>& operator*=(Dual<ND>& a, const Dual<ND>& b) { a = a * b; return a; }
template <int ND> inline Dual<ND> o_sqrt(const Dual<ND>& x) { Dual<ND> r(std::sqrt(x.v)); const double g = 0.5 / r.v; for (int i = 0; i < ND; ++i) r.d[i] = x.d[i] * g; return r; }
template <int ND> inline Dual<ND> o_cos(const Dual<ND>& x) { Dual<ND> r(std::cos(x.v)); const double g = -std::sin(x.v); for (int i = 0; i < ND; ++i) r.d[i] = x.d[i] * g; return r; }
template <int ND> inline Dual<ND> o_sin(const Dual<ND>& x) { Dual<ND> r(std::sin(x.v)); const double g = std::cos(x.v); for (int i = 0; i < ND; ++i) r.d[i] = x.d[i] * g; return r; }
template <int ND> inline double o_val(const Dual<ND>& x) { return x.v; }
template <int ND> struct Wide<Dual<ND>> { using type = Dual<ND>; };

// Constants are spelled as the float literals of the reference and converted
// to T, so the FP64 replay uses the same constant values.
template <class T> inline T C(float x) { return T(x); }
template <> inline OpCount C<OpCount>(float x) { return OpCount::S(x); }

// hipo: tbb.cpp:428-431.
template <class T> inline T hipo(const T& x, const T& y) { return o_sqrt(x * x + y * y); }

// sincos4: 4th-order Taylor expansion, tbb.cpp:644-652.
template <class T> inline void sincos4(const T& x, T& s, T& c) {
  const T x2 = x * x;
  c = T(1.f) - T(0.5f) * x2 + T(0.04166667f) * x2 * x2;
  s = x - T(0.16666667f) * x * x2;
}

// ---------------------------------------------------------------------------
// Jacobian sparsity pattern of p2r's errorProp (tbb.cpp:753-807).  For each
// row: the columns that enter the products of MultHelixProp (tbb.cpp:296-341)
// and MultHelixPropTransp (:343-373), in the order the reference sums them.
// kind 'm' multiplies by the entry, kind '1' adds the other operand unscaled
// (the entry is a structural 1 that the reference does not multiply).
struct JTerm { int col; char kind; };
struct JRow { int n; JTerm t[6]; };
inline const JRow& jrow(int i) {
  static const JRow rows[6] = {
      {4, {{0, 'm'}, {1, 'm'}, {3, 'm'}, {4, 'm'}}},                        // row 0
      {4, {{0, 'm'}, {1, 'm'}, {3, 'm'}, {4, 'm'}}},                        // row 1
      {6, {{0, 'm'}, {1, 'm'}, {2, '1'}, {3, 'm'}, {4, 'm'}, {5, 'm'}}},    // row 2
      {4, {{0, 'm'}, {1, 'm'}, {3, 'm'}, {4, 'm'}}},                        // row 3 (generic in the reference)
      {4, {{0, 'm'}, {1, 'm'}, {3, 'm'}, {4, 'm'}}},                        // row 4
      {1, {{5, '1'}}},                                                      // row 5
  };
  return rows[i];
}

// MultHelixProp (tbb.cpp:296-341): A = J * P, P packed symmetric, A full 6x6.
// Row 5 of A is row 5 of P (tbb.cpp:334-339).
template <class T> inline void mult_helix_prop(const T J[36], const T P[21], T A[36]) {
  for (int i = 0; i < 6; ++i) {
    const JRow& row = jrow(i);
    for (int j = 0; j < 6; ++j) {
      auto term = [&](int t) -> T {
        const int k = row.t[t].col;
        return row.t[t].kind == 'm' ? J[i * 6 + k] * P[sym6(k, j)] : P[sym6(k, j)];
      };
      T acc = term(0);
      for (int t = 1; t < row.n; ++t) acc = acc + term(t);
      A[i * 6 + j] = acc;
    }
  }
}

// MultHelixPropTransp (tbb.cpp:343-373): P' = A * J^T, packed output.
// Element (i,j) sums over row j of J:  sum_k A[i][k] * J[j][k].
template <class T> inline void mult_helix_prop_transp(const T J[36], const T A[36], T Pout[21]) {
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      const JRow& row = jrow(j);
      auto term = [&](int t) -> T {
        const int k = row.t[t].col;
        return row.t[t].kind == 'm' ? A[i * 6 + k] * J[j * 6 + k] : A[i * 6 + k];
      };
      T acc = term(0);
      for (int t = 1; t < row.n; ++t) acc = acc + term(t);
      Pout[sym6(i, j)] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Track state of one lane: par = (x, y, z, ipt, phi, theta), cov packed 21.
template <class T> struct Trk { T par[6]; T cov[21]; };
template <class T> struct Hit { T pos[3]; T cov[6]; };

// propagateToR for one lane (tbb.cpp:656-812).  `kf` is the kfact constant
// (tbb.cpp:654; a runtime value here so the same code serves both reference
// lineages and the field-scale derivative).  Writes out = propagated state and
// optionally the Jacobian.
template <class T>
inline void propagate_to_r(const Trk<T>& in, int q, const T hitpos[3], const T& kf, Trk<T>& out, T Jout[36] = nullptr) {
  T J[36];
  for (int i = 0; i < 36; ++i) J[i] = C<T>(0.f);
  for (int i = 0; i < 6; ++i) J[i * 6 + i] = C<T>(1.f);  // tbb.cpp:663

  T r0 = hipo(in.par[0], in.par[1]);                 // :665
  const T k = T(float(q)) * kf;                      // :667
  const T r = hipo(hitpos[0], hitpos[1]);            // :668
  const T xin = in.par[0], yin = in.par[1], iptin = in.par[3], phiin = in.par[4], thetain = in.par[5];
  T xo = xin, yo = yin;                              // :677-682 (outPar initialised to inPar)

  const T kinv = T(1.f) / k;                         // :684
  const T pt = T(1.f) / iptin;                       // :685
  T D = T(0.f), cosa = T(0.f), sina = T(0.f), id = T(0.f);
  T cosPorT = o_cos(phiin), sinPorT = o_sin(phiin);  // :689
  T pxin = cosPorT * pt;
  T pyin = sinPorT * pt;
  // :694-695.  Both quotients are formed and then selected (same values as
  // the reference's conditional expressions; keeps the op count branch-free).
  const T dx0 = -xin / r0, dy0 = -yin / r0;
  T dDdx = o_val(r0) > 0.0 ? dx0 : T(0.f);
  T dDdy = o_val(r0) > 0.0 ? dy0 : T(0.f);
  T dDdipt = T(0.f);
  T dDdphi = T(0.f);

  for (int i = 0; i < kNiter; ++i) {                 // :701-743
    r0 = hipo(xo, yo);
    id = (r - r0);
    D += id;
    sincos4(id * iptin * kinv, sina, cosa);
    if (i + 1 != kNiter) {
      const T xtmp = xo;
      const T ytmp = yo;
      const T rinv = T(1.f) / r0;
      const T oor0 = (o_val(r0) > 0.0 && std::fabs(o_val(id)) < double(0.0001f)) ? rinv : T(0.f);  // id == r - r0
      const T dadipt = id * kinv;
      const T dadx = -xtmp * iptin * kinv * oor0;
      const T dady = -ytmp * iptin * kinv * oor0;
      const T pxca = pxin * cosa;
      const T pxsa = pxin * sina;
      const T pyca = pyin * cosa;
      const T pysa = pyin * sina;
      T tmp = k * dadx;
      dDdx -= (xtmp * (T(1.f) + tmp * (pxca - pysa)) + ytmp * tmp * (pyca + pxsa)) * oor0;
      tmp = k * dady;
      dDdy -= (xtmp * tmp * (pxca - pysa) + ytmp * (T(1.f) + tmp * (pyca + pxsa))) * oor0;
      tmp = dadipt * iptin;
      dDdipt -= k * (xtmp * (pxca * tmp - pysa * tmp - pyca - pxsa + pyin) +
                     ytmp * (pyca * tmp + pxsa * tmp - pysa + pxca - pxin)) * pt * oor0;
      dDdphi += k * (xtmp * (pysa - pxin + pxca) - ytmp * (pxsa - pyin + pyca)) * oor0;
    }
    xo = xo + k * (pxin * sina - pyin * (T(1.f) - cosa));
    yo = yo + k * (pyin * sina + pxin * (T(1.f) - cosa));
    const T pxinold = pxin;
    pxin = pxin * cosa - pyin * sina;
    pyin = pyin * cosa + pxinold * sina;
  }

  const T alpha = D * iptin * kinv;                  // :745-749
  const T dadx = dDdx * iptin * kinv;
  const T dady = dDdy * iptin * kinv;
  const T dadipt = (iptin * dDdipt + D) * kinv;
  const T dadphi = dDdphi * iptin * kinv;
  sincos4(alpha, sina, cosa);                        // :751

  J[0 * 6 + 0] = T(1.f) + k * dadx * (cosPorT * cosa - sinPorT * sina) * pt;   // :753-758
  J[0 * 6 + 1] = k * dady * (cosPorT * cosa - sinPorT * sina) * pt;
  J[0 * 6 + 2] = C<T>(0.f);
  J[0 * 6 + 3] = k * (cosPorT * (iptin * dadipt * cosa - sina) + sinPorT * ((T(1.f) - cosa) - iptin * dadipt * sina)) * pt * pt;
  J[0 * 6 + 4] = k * (cosPorT * dadphi * cosa - sinPorT * dadphi * sina - sinPorT * sina + cosPorT * cosa - cosPorT) * pt;
  J[0 * 6 + 5] = C<T>(0.f);

  J[1 * 6 + 0] = k * dadx * (sinPorT * cosa + cosPorT * sina) * pt;             // :760-765
  J[1 * 6 + 1] = T(1.f) + k * dady * (sinPorT * cosa + cosPorT * sina) * pt;
  J[1 * 6 + 2] = C<T>(0.f);
  J[1 * 6 + 3] = k * (sinPorT * (iptin * dadipt * cosa - sina) + cosPorT * (iptin * dadipt * sina - (T(1.f) - cosa))) * pt * pt;
  J[1 * 6 + 4] = k * (sinPorT * dadphi * cosa + cosPorT * dadphi * sina + sinPorT * cosa + cosPorT * sina - sinPorT) * pt;
  J[1 * 6 + 5] = C<T>(0.f);

  cosPorT = o_cos(thetain);                          // :768-771
  sinPorT = o_sin(thetain);
  sinPorT = T(1.f) / sinPorT;

  out.par[0] = xo;
  out.par[1] = yo;
  out.par[2] = in.par[2] + k * alpha * cosPorT * pt * sinPorT;                  // :773

  J[2 * 6 + 0] = k * cosPorT * dadx * pt * sinPorT;                             // :775-780
  J[2 * 6 + 1] = k * cosPorT * dady * pt * sinPorT;
  J[2 * 6 + 2] = C<T>(1.f);
  J[2 * 6 + 3] = k * cosPorT * (iptin * dadipt - alpha) * pt * pt * sinPorT;
  J[2 * 6 + 4] = k * dadphi * cosPorT * pt * sinPorT;
  J[2 * 6 + 5] = -k * alpha * pt * sinPorT * sinPorT;

  out.par[3] = iptin;                                                           // :782
  for (int c = 0; c < 6; ++c) J[3 * 6 + c] = C<T>(c == 3 ? 1.f : 0.f);         // :784-789

  out.par[4] = phiin + alpha;                                                   // :791
  J[4 * 6 + 0] = dadx;                                                          // :793-798
  J[4 * 6 + 1] = dady;
  J[4 * 6 + 2] = C<T>(0.f);
  J[4 * 6 + 3] = dadipt;
  J[4 * 6 + 4] = T(1.f) + dadphi;
  J[4 * 6 + 5] = C<T>(0.f);

  out.par[5] = thetain;                                                         // :800
  for (int c = 0; c < 6; ++c) J[5 * 6 + c] = C<T>(c == 5 ? 1.f : 0.f);         // :802-807

  T A[36];
  mult_helix_prop(J, in.cov, A);                                                // :810
  mult_helix_prop_transp(J, A, out.cov);                                        // :811
  if (Jout) for (int i = 0; i < 36; ++i) Jout[i] = J[i];
}

// 2x2 inverse of the measurement-residual covariance with the determinant in
// the reference's wide type (tbb.cpp:454-464).  S = {s00, s01, s11} in/out.
template <class T> inline void invert_2x2_sym(T S[3]) {
  using WT = typename Wide<T>::type;
  const WT det = WT(S[0]) * WT(S[2]) - WT(S[1]) * WT(S[1]);
  const T s = T(WT(1.f) / det);
  const T tmp = s * S[2];
  S[1] *= -s;
  S[2] = s * S[0];
  S[0] = tmp;
}
template <> inline void invert_2x2_sym<float>(float S[3]) {
  const double det = (double)S[0] * S[2] - (double)S[1] * S[1];
  const float s = 1.f / det;
  const float tmp = s * S[2];
  S[1] *= -s;
  S[2] = s * S[0];
  S[0] = tmp;
}

// KalmanUpdate for one lane (tbb.cpp:433-642), in place on `t`.
// Optionally returns the local chi2 = r^T S^-1 r built from the residual and
// the inverted residual covariance the update already computes (the
// reference leaves Chi2Similarity commented out, tbb.cpp:631); chi2 is a
// quantity this project adds and is not part of the reference's output.
template <class T>
inline void kalman_update(Trk<T>& t, const Hit<T>& h, T* chi2 = nullptr) {
  const T* P = t.cov;
  const T* V = h.cov;
  const T r = hipo(h.pos[0], h.pos[1]);                                         // :441
  const T rotT00 = -(h.pos[1] + t.par[1]) / (T(2.f) * r);                       // :442
  const T rotT01 = (h.pos[0] + t.par[0]) / (T(2.f) * r);                        // :443

  T S[3];
  S[0] = (rotT00 * (P[0] + V[0]) + rotT01 * (P[1] + V[1])) * rotT00 +          // :445-451
         (rotT00 * (P[1] + V[1]) + rotT01 * (P[2] + V[2])) * rotT01;
  S[1] = (P[3] + V[3]) * rotT00 + (P[4] + V[4]) * rotT01;
  S[2] = (P[5] + V[5]);
  invert_2x2_sym(S);                                                            // :454-464

  // Gain K (6x2), tbb.cpp:466-512.  Row i uses P(i,0), P(i,1), P(i,2).
  T K[6][2];
  for (int i = 0; i < 6; ++i) {
    const T pi0 = P[sym6(i, 0)], pi1 = P[sym6(i, 1)], pi2 = P[sym6(i, 2)];
    K[i][0] = pi0 * (rotT00 * S[0]) + pi1 * (rotT01 * S[0]) + pi2 * S[1];
    K[i][1] = pi0 * (rotT00 * S[1]) + pi1 * (rotT01 * S[1]) + pi2 * S[2];
  }

  // Residual in the local frame and parameter update, tbb.cpp:514-528.
  const T res0 = rotT00 * (h.pos[0] - t.par[0]) + rotT01 * (h.pos[1] - t.par[1]);
  const T res1 = h.pos[2] - t.par[2];
  if (chi2) *chi2 = res0 * (S[0] * res0 + S[1] * res1) + res1 * (S[1] * res0 + S[2] * res1);
  for (int i = 0; i < 6; ++i) t.par[i] = t.par[i] + K[i][0] * res0 + K[i][1] * res1;

  // Covariance update P' = P - K H P, tbb.cpp:530-619 (all 21 entries).
  T KHP[21];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j)
      KHP[sym6(i, j)] = K[i][0] * rotT00 * P[sym6(j, 0)] + K[i][0] * rotT01 * P[sym6(j, 1)] + K[i][1] * P[sym6(j, 2)];
  T Pn[21];
  for (int e = 0; e < 21; ++e) Pn[e] = P[e] - KHP[e];
  for (int e = 0; e < 21; ++e) t.cov[e] = Pn[e];                                 // :641
}

}  // namespace orc
