// p2r oracle C-ABI -- TEST INFRASTRUCTURE ONLY (see p2r_oracle.hpp header).
//
// Matriplex byte layout (identical to the reference's, tbb.cpp:68-125):
//   MPTRK block = float par[6*bs] | float cov[21*bs] | int q[bs]   (28*bs words)
//   MPHIT block = float pos[3*bs] | float cov[6*bs]                (9*bs words)
//   track block (ev, ib)        at index ib + nb*ev                (tbb.cpp:147-153)
//   hit block   (ev, ib, layer) at index layer + nlayer*(ib+nb*ev) (tbb.cpp:215-217)
//   element (field f, lane it)  at data[f*bs + it]                 (tbb.cpp:159-161)
#include "p2r_oracle.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

#include "c0_inputs.inc"

inline const float* trk_block(const float* base, long blk, int bs) { return base + blk * 28L * bs; }
inline float* trk_block(float* base, long blk, int bs) { return base + blk * 28L * bs; }
inline const float* hit_block(const float* base, long hblk, int bs) { return base + hblk * 9L * bs; }

template <class T>
inline void load_trk(const float* blk, int bs, int it, Trk<T>& t, int& q) {
  for (int f = 0; f < 6; ++f) t.par[f] = T(blk[f * bs + it]);
  for (int f = 0; f < 21; ++f) t.cov[f] = T(blk[(6 + f) * bs + it]);
  int qi;
  std::memcpy(&qi, blk + 27 * bs + it, sizeof(int));
  q = qi;
}
template <class T>
inline void load_hit(const float* blk, int bs, int it, Hit<T>& h) {
  for (int f = 0; f < 3; ++f) h.pos[f] = T(blk[f * bs + it]);
  for (int f = 0; f < 6; ++f) h.cov[f] = T(blk[(3 + f) * bs + it]);
}
inline void store_trk(float* blk, int bs, int it, const Trk<float>& t, int q) {
  for (int f = 0; f < 6; ++f) blk[f * bs + it] = t.par[f];
  for (int f = 0; f < 21; ++f) blk[(6 + f) * bs + it] = t.cov[f];
  std::memcpy(blk + 27 * bs + it, &q, sizeof(int));
}

// One Matriplex block through all layers.  semantics 0 = reference (every
// layer restarts from the input track, tbb.cpp:880-911); 1 = chained (layer
// L starts from layer L-1's updated state, propagate-tor-test_hip.cu:839-845).
// mask (nullable): per-lane bitmask, bit L set = layer L present; absent layers
// are skipped (chained semantics only).
// options (bitwise): 1 = wrap phi after each update, phi - 2pi floor((phi + pi) * (1/2pi)) in T
// (the squash the reference leaves commented out, tbb.cpp:635); 2 = an update that makes 1/pt negative
// flips 1/pt's sign and the charge (tbb.cpp:527 note), and later layers use the new charge; 4 = chi2
// receives each update's chi2 (Chi2Similarity, commented out at tbb.cpp:631) at chi2[l * lstride + it]
// instead of the per-track sum, 0 for absent layers.
template <class T>
T wrap_phi(T phi) {
  const T n = std::floor((phi + T(3.14159265358979323846)) * T(0.15915494309189533577));
  return phi - n * T(6.28318530717958647693);
}

template <class T>
void run_block(const float* tb, const float* hb0, int bs, int nlayer, int semantics, T kf,
               const uint32_t* mask, Trk<T>* out, int* qout, T* chi2, int options = 0, long lstride = 0) {
  for (int it = 0; it < bs; ++it) {
    Trk<T> in;
    int q_in;
    load_trk(tb, bs, it, in, q_in);
    int q = q_in;
    Trk<T> cur = in;
    T c2 = T(0.f);
    for (int l = 0; l < nlayer; ++l) {
      if (mask && !((mask[it] >> l) & 1u)) {
        if (chi2 && (options & 4)) chi2[l * lstride + it] = T(0.f);
        continue;
      }
      Hit<T> h;
      load_hit(hb0 + (long)l * 9 * bs, bs, it, h);
      const Trk<T>& src = (semantics == 0) ? in : cur;
      if (semantics == 0) q = q_in;
      Trk<T> prop;
      propagate_to_r(src, q, h.pos, kf, prop);
      T c2l;
      kalman_update(prop, h, &c2l);
      if (options & 1) prop.par[4] = wrap_phi(prop.par[4]);
      if ((options & 2) && prop.par[3] < T(0.f)) {
        prop.par[3] = -prop.par[3];
        q = -q;
      }
      if (chi2 && (options & 4)) chi2[l * lstride + it] = c2l;
      c2 = c2 + c2l;
      cur = prop;
    }
    out[it] = cur;
    qout[it] = q;
    if (chi2 && !(options & 4)) chi2[it] = c2;
  }
}

// glibc-rand()-based Gaussian, the reference's randn (tbb.cpp:127-145).
struct RefRandn {
  float X1 = 0, X2 = 0;
  int call = 0;
  float operator()(float mu, float sigma) {
    float U1, U2, W, mult;
    if (call == 1) {
      call = !call;
      return (mu + sigma * (float)X2);
    }
    do {
      U1 = -1 + ((float)rand() / RAND_MAX) * 2;
      U2 = -1 + ((float)rand() / RAND_MAX) * 2;
      W = std::pow(U1, 2) + std::pow(U2, 2);
    } while (W >= 1 || W == 0);
    mult = std::sqrt((-2 * std::log(W)) / W);
    X1 = U1 * mult;
    X2 = U2 * mult;
    call = !call;
    return (mu + sigma * (float)X1);
  }
};

}  // namespace

extern "C" {

// ---- block-level mirrors of the reference entry points ---------------------

// propagateToR on one Matriplex block (tbb.cpp:656-812).  Pointers address the
// start of the MP6x6SF / MP6F / MP1I / MP3F arrays of that block.
void orc_propagateToR_block(int bs, float kfact, const float* inErr, const float* inPar, const int* inChg,
                            const float* msP, float* outErr, float* outPar) {
  for (int it = 0; it < bs; ++it) {
    Trk<float> in, out;
    for (int f = 0; f < 6; ++f) in.par[f] = inPar[f * bs + it];
    for (int f = 0; f < 21; ++f) in.cov[f] = inErr[f * bs + it];
    const float hp[3] = {msP[it], msP[bs + it], msP[2 * bs + it]};
    propagate_to_r(in, inChg[it], hp, kfact, out);
    for (int f = 0; f < 6; ++f) outPar[f * bs + it] = out.par[f];
    for (int f = 0; f < 21; ++f) outErr[f * bs + it] = out.cov[f];
  }
}

// KalmanUpdate on one Matriplex block, in place (tbb.cpp:433-642).
void orc_KalmanUpdate_block(int bs, float* trkErr, float* inPar, const float* hitErr, const float* msP) {
  for (int it = 0; it < bs; ++it) {
    Trk<float> t;
    Hit<float> h;
    for (int f = 0; f < 6; ++f) t.par[f] = inPar[f * bs + it];
    for (int f = 0; f < 21; ++f) t.cov[f] = trkErr[f * bs + it];
    for (int f = 0; f < 3; ++f) h.pos[f] = msP[f * bs + it];
    for (int f = 0; f < 6; ++f) h.cov[f] = hitErr[f * bs + it];
    kalman_update(t, h);
    for (int f = 0; f < 6; ++f) inPar[f * bs + it] = t.par[f];
    for (int f = 0; f < 21; ++f) trkErr[f * bs + it] = t.cov[f];
  }
}

// Propagate one lane and return the 6x6 Jacobian (row-major) -- used by the
// finite-difference and closed-form tests.
void orc_propagate_lane_f64(const double par[6], const double cov[21], int q, const double hit[3], double kfact,
                            double par_out[6], double cov_out[21], double J_out[36]) {
  Trk<double> in, out;
  for (int f = 0; f < 6; ++f) in.par[f] = par[f];
  for (int f = 0; f < 21; ++f) in.cov[f] = cov[f];
  propagate_to_r(in, q, hit, kfact, out, J_out);
  for (int f = 0; f < 6; ++f) par_out[f] = out.par[f];
  for (int f = 0; f < 21; ++f) cov_out[f] = out.cov[f];
}
void orc_kalman_lane_f64(double par[6], double cov[21], const double hitpos[3], const double hitcov[6], double* chi2) {
  Trk<double> t;
  Hit<double> h;
  for (int f = 0; f < 6; ++f) t.par[f] = par[f];
  for (int f = 0; f < 21; ++f) t.cov[f] = cov[f];
  for (int f = 0; f < 3; ++f) h.pos[f] = hitpos[f];
  for (int f = 0; f < 6; ++f) h.cov[f] = hitcov[f];
  kalman_update(t, h, chi2);
  for (int f = 0; f < 6; ++f) par[f] = t.par[f];
  for (int f = 0; f < 21; ++f) cov[f] = t.cov[f];
}

// ---- full runs --------------------------------------------------------------

// Selects the FP32 path's sin/cos (see trig_mode in p2r_oracle.hpp).
void orc_set_trig_mode(int mode) { trig_mode() = mode ? 1 : 0; }

// FP32 run over nblk Matriplex blocks in the reference layout (the driver loop
// of tbb.cpp:873-925 for one repetition).  `nblk` = nevts*nb.  chi2 (nullable)
// receives per-track sum of local chi2 over the processed layers, indexed by
// block*bs + lane (options & 4: each layer's chi2, [nlayer][nblk*bs]).  mask
// (nullable) = per-track layer bitmask, same indexing.
int orc_run_f32(const float* trk, const float* hits, float* out, long nblk, int nlayer, int bs, int semantics,
                float kfact, int nthreads, float* chi2, const uint32_t* mask, int options) {
  if (bs <= 0 || nlayer <= 0 || nblk < 0) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (long b = 0; b < nblk; ++b) {
    std::vector<Trk<float>> res(bs);
    std::vector<int> q(bs);
    run_block<float>(trk_block(trk, b, bs), hit_block(hits, b * nlayer, bs), bs, nlayer, semantics, kfact,
                     mask ? mask + b * bs : nullptr, res.data(), q.data(), chi2 ? chi2 + b * bs : nullptr, options,
                     nblk * bs);
    float* ob = trk_block(out, b, bs);
    for (int it = 0; it < bs; ++it) store_trk(ob, bs, it, res[it], q[it]);
  }
  return 0;
}

// FP64 replay: the same algorithm in double precision on the same FP32 inputs.
// Outputs per track (index block*bs + lane): par_out[6], cov_out[21], chi2.
int orc_run_f64(const float* trk, const float* hits, double* par_out, double* cov_out, double* chi2, long nblk,
                int nlayer, int bs, int semantics, double kfact, int nthreads, const uint32_t* mask, int options) {
  if (bs <= 0 || nlayer <= 0 || nblk < 0) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (long b = 0; b < nblk; ++b) {
    std::vector<Trk<double>> res(bs);
    std::vector<int> q(bs);
    std::vector<double> c2(bs);
    run_block<double>(trk_block(trk, b, bs), hit_block(hits, b * nlayer, bs), bs, nlayer, semantics, kfact,
                      mask ? mask + b * bs : nullptr, res.data(), q.data(), c2.data(), options & 3);
    for (int it = 0; it < bs; ++it) {
      const long t = b * bs + it;
      for (int f = 0; f < 6; ++f) par_out[t * 6 + f] = res[it].par[f];
      for (int f = 0; f < 21; ++f) cov_out[t * 21 + f] = res[it].cov[f];
      if (chi2) chi2[t] = c2[it];
    }
  }
  return 0;
}

// ---- reference-compatible generator (C0) ------------------------------------

// kfact of the reference's current lineage, tbb.cpp:654 (double expression
// narrowed to float).
float orc_kfact_tbb(void) { return (float)(100 / (-0.299792458 * 3.8112)); }

// prepareTracks + prepareHits of tbb.cpp:243-293 with the input of
// input_track.h (layer l uses inputhit[21-l], tbb.cpp:818-820), glibc rand()
// in its unseeded state (srand(1)) and smear 1e-5.  Writes nevts*nb blocks.
int orc_gen_c0(float* trk, float* hits, int nevts, int ntrks, int nlayer, int bs) {
  if (bs <= 0 || ntrks % bs != 0 || nlayer > 22 || nlayer <= 0) return 1;
  const int nb = ntrks / bs;
  const double smear = 0.00001;
  srand(1);
  RefRandn randn;
  float tpar[6], tcov[21];
  for (int i = 0; i < 6; ++i) tpar[i] = (float)kC0TrackPar[i];
  for (int i = 0; i < 21; ++i) tcov[i] = (float)kC0TrackCov[i];
  for (long ie = 0; ie < nevts; ++ie)
    for (long ib = 0; ib < nb; ++ib) {
      float* blk = trk + (ib + nb * ie) * 28L * bs;
      for (int it = 0; it < bs; ++it) {
        for (int ip = 0; ip < 6; ++ip) blk[ip * bs + it] = (1 + smear * randn(0, 1)) * tpar[ip];
        for (int ip = 0; ip < 21; ++ip) blk[(6 + ip) * bs + it] = (1 + smear * randn(0, 1)) * tcov[ip] * 100;
        const int q = kC0TrackQ;
        std::memcpy(blk + 27 * bs + it, &q, sizeof(int));
      }
    }
  for (int lay = 0; lay < nlayer; ++lay) {
    float hp[3], hc[6];
    for (int i = 0; i < 3; ++i) hp[i] = (float)kC0Hit[21 - lay][i];
    for (int i = 0; i < 6; ++i) hc[i] = (float)kC0Hit[21 - lay][3 + i];
    for (long ie = 0; ie < nevts; ++ie)
      for (long ib = 0; ib < nb; ++ib) {
        float* blk = hits + (lay + nlayer * (ib + nb * ie)) * 9L * bs;
        for (int it = 0; it < bs; ++it) {
          for (int ip = 0; ip < 3; ++ip) blk[ip * bs + it] = (1 + smear * randn(0, 1)) * hp[ip];
          for (int ip = 0; ip < 6; ++ip) blk[(3 + ip) * bs + it] = (1 + smear * randn(0, 1)) * hc[ip];
        }
      }
  }
  return 0;
}

// Unsmeared C0 hit position of layer `lay` (the summary's reference hit).
void orc_c0_hit(int lay, float pos[3]) {
  for (int i = 0; i < 3; ++i) pos[i] = (float)kC0Hit[21 - lay][i];
}

// First n values of glibc rand() after srand(seed).
void orc_rand_prefix(unsigned seed, int n, int* out) {
  srand(seed);
  for (int i = 0; i < n; ++i) out[i] = rand();
}

// ---- summary (tbb.cpp:935-1048) ---------------------------------------------
// stats[0..21]: avgx, stdx/|avgx|, avgy, std/|avgy|, avgz, std/|avgz|,
//   avgr, stdr/|avgz| (the reference divides by avgz, tbb.cpp:1039),
//   avgdx, stddx, avgdy, stddy, avgdz, stddz, avgdr, stddr,
//   avgpt, avgphi, avgtheta, and (additions of this project) rms of pt, phi,
//   theta about their means over the same accepted tracks.
// counts[0] = tracks with nans, counts[1] = failed tracks.
void orc_summary(const float* out, int nevts, int ntrks, int bs, float hx_, float hy_, float hz_, double* stats,
                 int* counts) {
  const int nb = ntrks / bs;
  auto fld = [&](long ie, long tk, int f) -> float {
    const long ib = tk / bs, it = tk % bs;
    return out[(ib + (long)nb * ie) * 28L * bs + f * bs + it];
  };
  int nnans = 0, nfail = 0;
  float avgx = 0, avgy = 0, avgz = 0, avgr = 0;
  float avgpt = 0, avgphi = 0, avgtheta = 0;
  float avgdx = 0, avgdy = 0, avgdz = 0, avgdr = 0;
  const float hr_ = sqrtf(hx_ * hx_ + hy_ * hy_);
  for (long ie = 0; ie < nevts; ++ie)
    for (long it = 0; it < ntrks; ++it) {
      float x_ = fld(ie, it, 0), y_ = fld(ie, it, 1), z_ = fld(ie, it, 2);
      float r_ = sqrtf(x_ * x_ + y_ * y_);
      float pt_ = std::abs(1. / fld(ie, it, 3));
      float phi_ = fld(ie, it, 4), theta_ = fld(ie, it, 5);
      if (!std::isfinite(x_) || !std::isfinite(y_) || !std::isfinite(z_) || !std::isfinite(pt_) ||
          !std::isfinite(phi_) || !std::isfinite(theta_)) {
        nnans++;
        continue;
      }
      if (std::fabs((x_ - hx_) / hx_) > 1. || std::fabs((y_ - hy_) / hy_) > 1. || std::fabs((z_ - hz_) / hz_) > 1.) {
        nfail++;
        continue;
      }
      avgpt += pt_; avgphi += phi_; avgtheta += theta_;
      avgx += x_; avgy += y_; avgz += z_; avgr += r_;
      avgdx += (x_ - hx_) / x_; avgdy += (y_ - hy_) / y_; avgdz += (z_ - hz_) / z_; avgdr += (r_ - hr_) / r_;
    }
  const float n = float(nevts * ntrks);
  avgpt = avgpt / n; avgphi = avgphi / n; avgtheta = avgtheta / n;
  avgx = avgx / n; avgy = avgy / n; avgz = avgz / n; avgr = avgr / n;
  avgdx = avgdx / n; avgdy = avgdy / n; avgdz = avgdz / n; avgdr = avgdr / n;
  float stdx = 0, stdy = 0, stdz = 0, stdr = 0, stddx = 0, stddy = 0, stddz = 0, stddr = 0;
  double spt = 0, sphi = 0, stheta = 0;
  for (long ie = 0; ie < nevts; ++ie)
    for (long it = 0; it < ntrks; ++it) {
      float x_ = fld(ie, it, 0), y_ = fld(ie, it, 1), z_ = fld(ie, it, 2);
      float r_ = sqrtf(x_ * x_ + y_ * y_);
      if (!std::isfinite(x_) || !std::isfinite(y_) || !std::isfinite(z_)) continue;
      if (std::fabs((x_ - hx_) / hx_) > 1. || std::fabs((y_ - hy_) / hy_) > 1. || std::fabs((z_ - hz_) / hz_) > 1.)
        continue;
      stdx += (x_ - avgx) * (x_ - avgx); stdy += (y_ - avgy) * (y_ - avgy);
      stdz += (z_ - avgz) * (z_ - avgz); stdr += (r_ - avgr) * (r_ - avgr);
      stddx += ((x_ - hx_) / x_ - avgdx) * ((x_ - hx_) / x_ - avgdx);
      stddy += ((y_ - hy_) / y_ - avgdy) * ((y_ - hy_) / y_ - avgdy);
      stddz += ((z_ - hz_) / z_ - avgdz) * ((z_ - hz_) / z_ - avgdz);
      stddr += ((r_ - hr_) / r_ - avgdr) * ((r_ - hr_) / r_ - avgdr);
      const float pt_ = std::abs(1. / fld(ie, it, 3));
      if (std::isfinite(pt_) && std::isfinite(fld(ie, it, 4)) && std::isfinite(fld(ie, it, 5))) {
        spt += double(pt_ - avgpt) * double(pt_ - avgpt);
        sphi += double(fld(ie, it, 4) - avgphi) * double(fld(ie, it, 4) - avgphi);
        stheta += double(fld(ie, it, 5) - avgtheta) * double(fld(ie, it, 5) - avgtheta);
      }
    }
  stdx = sqrtf(stdx / n); stdy = sqrtf(stdy / n); stdz = sqrtf(stdz / n); stdr = sqrtf(stdr / n);
  stddx = sqrtf(stddx / n); stddy = sqrtf(stddy / n); stddz = sqrtf(stddz / n); stddr = sqrtf(stddr / n);
  const double v[22] = {avgx, std::fabs(stdx / avgx), avgy, std::fabs(stdy / avgy), avgz, std::fabs(stdz / avgz),
                        avgr, std::fabs(stdr / avgz), avgdx, stddx, avgdy, stddy, avgdz, stddz, avgdr, stddr,
                        avgpt, avgphi, avgtheta, std::sqrt(spt / n), std::sqrt(sphi / n), std::sqrt(stheta / n)};
  for (int i = 0; i < 22; ++i) stats[i] = v[i];
  counts[0] = nnans;
  counts[1] = nfail;
}

// ---- reference gradients of the chained fit's total chi2 --------------------
// Sum over layers of the local chi2 (r^T S^-1 r, the quantity this project
// adds to the reference's KalmanUpdate) of one track in chained semantics, as
// a function of: the 6 initial parameters, the 3 coordinates of every hit
// and kfact.  Input layout per track (AoS): par[6], cov[21], hits[nlayer][9]
// (pos 3, cov 6).  Gradient layout: [0..5] d/dpar, [6 + 3*l + c] d/d hit l
// coordinate c, [6 + 3*nlayer] d/dkfact.

}  // extern "C"

namespace {
template <class T>
T chained_chi2(const T par[6], const float cov[21], int q, const T* hitpos, const float* hits, int nlayer, T kf) {
  Trk<T> cur;
  for (int f = 0; f < 6; ++f) cur.par[f] = par[f];
  for (int f = 0; f < 21; ++f) cur.cov[f] = T(cov[f]);
  T tot = T(0.f);
  for (int l = 0; l < nlayer; ++l) {
    Hit<T> h;
    for (int c = 0; c < 3; ++c) h.pos[c] = hitpos[3 * l + c];
    for (int c = 0; c < 6; ++c) h.cov[c] = T(hits[9 * l + 3 + c]);
    Trk<T> prop;
    propagate_to_r(cur, q, h.pos, kf, prop);
    T c2;
    kalman_update(prop, h, &c2);
    tot = tot + c2;
    cur = prop;
  }
  return tot;
}
}  // namespace

extern "C" {

// FP64 value of the chained total chi2 with the inputs given in double (for
// central finite differences).  par/hitpos are the (perturbed) variables.
double orc_chi2_f64(const double par[6], const float cov[21], int q, const double* hitpos, const float* hits,
                    int nlayer, double kfact) {
  return chained_chi2<double>(par, cov, q, hitpos, hits, nlayer, kfact);
}

// Forward-mode (Dual, FP64) gradient, 16 directions per pass.
void orc_chi2_grad_dual(const float par[6], const float cov[21], int q, const float* hits, int nlayer, double kfact,
                        double* grad, double* chi2) {
  constexpr int ND = 16;
  using D = Dual<ND>;
  const int nvar = 6 + 3 * nlayer + 1;
  std::vector<D> hp(3 * nlayer);
  for (int c0 = 0; c0 < nvar; c0 += ND) {
    D p[6], kf(kfact);
    for (int f = 0; f < 6; ++f) {
      p[f] = D(par[f]);
      if (f >= c0 && f < c0 + ND) p[f].d[f - c0] = 1.0;
    }
    for (int l = 0; l < nlayer; ++l)
      for (int c = 0; c < 3; ++c) {
        const int v = 6 + 3 * l + c;
        hp[3 * l + c] = D(hits[9 * l + c]);
        if (v >= c0 && v < c0 + ND) hp[3 * l + c].d[v - c0] = 1.0;
      }
    if (nvar - 1 >= c0 && nvar - 1 < c0 + ND) kf.d[nvar - 1 - c0] = 1.0;
    const D tot = chained_chi2<D>(p, cov, q, hp.data(), hits, nlayer, kf);
    for (int v = c0; v < c0 + ND && v < nvar; ++v) grad[v] = tot.d[v - c0];
    if (chi2) *chi2 = tot.v;
  }
}

// ---- op count ---------------------------------------------------------------
// FLOPs of one propagateToR + KalmanUpdate (+ the chi2 this project adds) for
// one lane.  out[0..5] = add, mul, div, sqrt, trig, trivial; out[6] = chi2 ops
// (counted separately, not part of the reference algorithm).
void orc_opcount(long long out[7]) {
  tally() = OpTally{};
  Trk<OpCount> in, prop;
  const double par[6] = {-12.8, -7.7, 38.1, 0.237, -2.61, 0.356};
  for (int f = 0; f < 6; ++f) in.par[f] = OpCount(par[f]);
  for (int f = 0; f < 21; ++f) in.cov[f] = OpCount(f == sym6(f % 6, f % 6) ? 1e-4 : 1e-6);
  OpCount hp[3] = {OpCount(-20.0), OpCount(-12.0), OpCount(57.0)};
  Hit<OpCount> h;
  for (int f = 0; f < 3; ++f) h.pos[f] = hp[f];
  for (int f = 0; f < 6; ++f) h.cov[f] = OpCount(f == 0 || f == 2 || f == 5 ? 1e-4 : 0.0);
  propagate_to_r(in, 1, hp, OpCount(-87.5), prop);
  kalman_update(prop, h);
  const OpTally base = tally();
  out[0] = base.add; out[1] = base.mul; out[2] = base.div; out[3] = base.sqrt; out[4] = base.trig;
  out[5] = base.trivial;
  // chi2-only ops = (an update with chi2) - (the same update without chi2)
  tally() = OpTally{};
  Trk<OpCount> t2 = prop;
  OpCount c2;
  kalman_update(t2, h, &c2);
  const OpTally with = tally();
  tally() = OpTally{};
  Trk<OpCount> t3 = prop;
  kalman_update(t3, h);
  const OpTally upd = tally();
  out[6] = with.total() - upd.total();
}

}  // extern "C"
