// Host-side input generators (declarations: include/p2r_gen.h).
#include "../../include/p2r_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace {

#include "p2r_reference_inputs.inc"

// ---- reference-compatible generator ------------------------------------------

// The reference's Gaussian (tbb.cpp:127-145): Marsaglia polar method on glibc
// rand(), returning the two variates of a pair on successive calls.
struct PolarRandn {
  float x1 = 0, x2 = 0;
  bool have = false;
  float operator()(float mu, float sigma) {
    if (have) {
      have = false;
      return mu + sigma * x2;
    }
    float u1, u2, w;
    do {
      u1 = -1 + ((float)rand() / RAND_MAX) * 2;
      u2 = -1 + ((float)rand() / RAND_MAX) * 2;
      w = std::pow(u1, 2) + std::pow(u2, 2);  // promoted to double, as in the reference
    } while (w >= 1 || w == 0);
    const float mult = std::sqrt((-2 * std::log(w)) / w);
    x1 = u1 * mult;
    x2 = u2 * mult;
    have = true;
    return mu + sigma * x1;
  }
};

// ---- counter-based generator -------------------------------------------------

inline uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
enum Stream : uint64_t { kPar = 1, kCov = 2, kCharge = 3, kHitPos = 4, kMask = 5, kCount = 6, kMaskPerm = 7, kSeed = 8 };
inline uint64_t counter(uint64_t seed, uint64_t stream, uint64_t item, uint64_t field) {
  return mix64(mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ item) ^ field);
}
inline double uniform01(uint64_t bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }
// N(0,1) by Box-Muller from one 64-bit counter value.
inline double normal(uint64_t seed, uint64_t stream, uint64_t item, uint64_t field) {
  const uint64_t a = counter(seed, stream, item, field);
  const uint64_t b = mix64(a ^ 0x632BE59BD9B4E019ull);
  const double u1 = 1.0 - uniform01(a);  // (0, 1]
  const double u2 = uniform01(b);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
// N(0,1) truncated to |N| <= 3 by redrawing (field + 64*k for the k-th redraw; 16 redraws never all miss
// in practice, the last one is clamped).
inline double normal_trunc3(uint64_t seed, uint64_t stream, uint64_t item, uint64_t field) {
  double v = 0;
  for (uint64_t k = 0; k < 16; ++k) {
    v = normal(seed, stream, item, field + 64 * k);
    if (std::fabs(v) <= 3.0) return v;
  }
  return std::fmax(-3.0, std::fmin(3.0, v));
}
inline uint64_t track_item(uint64_t ev, uint64_t tk) { return (ev << 32) | tk; }

// Base state of the benchmark workloads (propagate-tor-test_acc.v3.cpp:808-812).
const double kBasePar[6] = {-12.806846618652344, -7.723824977874756, 38.13014221191406,
                            0.23732035065189902, -2.613372802734375, 0.35594117641448975};
const double kBaseCov[21] = {6.290299552347278e-07, 4.1375109560704004e-08, 7.526661534029699e-07,
                             2.0973730840978533e-07, 1.5431574240665213e-07, 9.626245400795597e-08,
                             -2.804026640189443e-06, 6.219111130687595e-06, 2.649119409845118e-07,
                             0.00253512163402557, -2.419662877381737e-07, 4.3124190760040646e-07,
                             3.1068903991780678e-09, 0.000923913115050627, 0.00040678296006807003,
                             -7.755406890332818e-07, 1.68539375883925e-06, 6.676875566525437e-08,
                             0.0008420574605423793, 7.356584799406111e-05, 0.0002306247719158348};

// Point of the exact helix of the base track (charge +1) after turning angle
// a, in the reference's parametrisation (tbb.cpp:738-742, :773).
void helix_point(double kfact, double a, double out[3]) {
  const double k = kfact, pt = 1.0 / kBasePar[3];
  const double px = std::cos(kBasePar[4]) * pt, py = std::sin(kBasePar[4]) * pt;
  out[0] = kBasePar[0] + k * (px * std::sin(a) - py * (1 - std::cos(a)));
  out[1] = kBasePar[1] + k * (py * std::sin(a) + px * (1 - std::cos(a)));
  out[2] = kBasePar[2] + k * a * pt * std::cos(kBasePar[5]) / std::sin(kBasePar[5]);
}
// Crossing of that helix with the cylinder of radius r (FP64 bisection on the
// turning angle, outward from the start point).
void helix_crossing(double kfact, double r, double out[3]) {
  const double dir = (kBasePar[3] / kfact) > 0 ? 1.0 : -1.0;
  auto rad = [&](double t) {
    double p[3];
    helix_point(kfact, dir * t, p);
    return std::hypot(p[0], p[1]);
  };
  double lo = 0.0, hi = 1e-3;
  while (rad(hi) < r && hi < 3.0) hi *= 1.5;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    (rad(mid) < r ? lo : hi) = mid;
  }
  helix_point(kfact, dir * 0.5 * (lo + hi), out);
}

template <class F>
void run_parallel(int64_t n, F fn) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 64) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto& x : th) x.join();
}

struct LayerHits {
  std::vector<double> base;  // 3 per layer: crossing of the base track (literal mode)
};

LayerHits layer_bases(int nlayer, const p2r_gen_config& c) {
  LayerHits h;
  h.base.resize(3 * (size_t)nlayer);
  const double kf = (double)p2r_default_params().kfact / (double)c.b_scale_truth;
  for (int l = 0; l < nlayer; ++l) helix_crossing(kf, (double)c.r0 + l * (double)c.dr, &h.base[3 * l]);
  return h;
}

// Positive-definite correlation of the consistent mode: the base correlations
// (not positive-definite as given) shrunk towards the identity,
// rho_s = lam*rho + (1-lam)*I, with the largest lam in steps of 0.05 whose
// Cholesky pivots are all >= 0.05.  Returns the Cholesky factor (row-major).
struct Corr {
  double L[36] = {};
  double lam = 1.0;
};
bool cholesky6(const double A[36], double L[36], double min_pivot) {
  for (int i = 0; i < 36; ++i) L[i] = 0;
  for (int j = 0; j < 6; ++j) {
    double d = A[j * 6 + j];
    for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k];
    if (!(d >= min_pivot)) return false;
    L[j * 6 + j] = std::sqrt(d);
    for (int i = j + 1; i < 6; ++i) {
      double v = A[i * 6 + j];
      for (int k = 0; k < j; ++k) v -= L[i * 6 + k] * L[j * 6 + k];
      L[i * 6 + j] = v / L[j * 6 + j];
    }
  }
  return true;
}
const Corr& shrunk_correlation() {
  static const Corr corr = [] {
    Corr c;
    double rho[36];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        const int e = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
        const double di = kBaseCov[i * (i + 1) / 2 + i], dj = kBaseCov[j * (j + 1) / 2 + j];
        rho[i * 6 + j] = kBaseCov[e] / std::sqrt(di * dj);
      }
    for (int step = 0; step <= 20; ++step) {
      c.lam = 1.0 - 0.05 * step;
      double A[36];
      for (int i = 0; i < 36; ++i) A[i] = c.lam * rho[i] + ((i % 7 == 0) ? 1.0 - c.lam : 0.0);
      if (cholesky6(A, c.L, 0.05 * 0.05)) break;
    }
    return c;
  }();
  return corr;
}

// Crossing of a track's exact helix with the cylinder of radius r, closed form.
// In the reference's parametrisation (tbb.cpp:738-742, :773) the track point
// after turning angle a is C + k*pt*(sin(a+phi), -cos(a+phi)) with centre
// C = (x0 - k*py, y0 + k*px), z = z0 + k*a*pt*cot(theta); motion along the
// momentum has a of the sign of ipt/k.  Returns false when the helix never
// reaches radius r.
bool helix_cylinder(const double p[6], double k, double r, double out[3]) {
  const double pt = 1.0 / p[3], px = std::cos(p[4]) * pt, py = std::sin(p[4]) * pt;
  const double cx = p[0] - k * py, cy = p[1] + k * px, R = std::fabs(k * pt), d = std::hypot(cx, cy);
  if (!(d > 0) || r > d + R || r < std::fabs(d - R)) return false;
  const double aa = (r * r - R * R + d * d) / (2 * d), h = std::sqrt(std::max(0.0, r * r - aa * aa));
  const double ux = cx / d, uy = cy / d, dir = (p[3] / k) > 0 ? 1.0 : -1.0;
  double best_t = 1e300, best_a = 0;
  for (int sgn = -1; sgn <= 1; sgn += 2) {
    const double qx = aa * ux - sgn * h * uy, qy = aa * uy + sgn * h * ux;
    const double s = (qx - cx) / (k * pt), c = -(qy - cy) / (k * pt);
    double t = dir * (std::atan2(s, c) - p[4]);
    t = std::fmod(t, 6.283185307179586);
    if (t < 0) t += 6.283185307179586;
    if (t < best_t) {
      best_t = t;
      best_a = dir * t;
    }
  }
  out[0] = cx + k * pt * std::sin(best_a + p[4]);
  out[1] = cy - k * pt * std::cos(best_a + p[4]);
  out[2] = p[2] + k * best_a * pt * std::cos(p[5]) / std::sin(p[5]);
  return true;
}

// Track of one lane; returns the true state used to place the hits.
void fill_track(float* blk, int bs, int it, uint64_t item, const p2r_gen_config& c, double truth[6], int& q) {
  q = (counter(c.seed, kCharge, item, 0) >> 63) ? 1 : -1;
  std::memcpy(blk + 27 * bs + it, &q, sizeof(int));
  if (c.mode == P2R_GEN_LITERAL) {
    for (int f = 0; f < 6; ++f) truth[f] = kBasePar[f] * (1.0 + c.smear * normal(c.seed, kPar, item, f));
    for (int f = 0; f < 6; ++f) blk[f * bs + it] = (float)truth[f];
    for (int f = 0; f < 21; ++f)
      blk[(6 + f) * bs + it] = (float)(kBaseCov[f] * (1.0 + c.smear * normal(c.seed, kCov, item, f)));
    return;
  }
  // Truth smears truncated at 3 sigma: with the base position (r = 15 cm) every track then starts inside
  // r = 19.5 cm, so the first cylinder (r0 = 20 cm by default) lies ahead of it.  Untruncated, about 8 in
  // 1e6 tracks start beyond r0, their first propagation runs backwards and the fit diverges.
  for (int f = 0; f < 6; ++f) truth[f] = kBasePar[f] * (1.0 + c.smear * normal_trunc3(c.seed, kPar, item, f));
  const Corr& corr = shrunk_correlation();
  double sig[6], z[6];
  for (int i = 0; i < 6; ++i) {
    sig[i] = std::sqrt(kBaseCov[i * (i + 1) / 2 + i]) * std::fabs(1.0 + c.smear * normal(c.seed, kCov, item, i));
    z[i] = normal(c.seed, kSeed, item, i);
  }
  for (int i = 0; i < 6; ++i) {  // input state = truth + diag(sig) L z
    double v = 0;
    for (int k = 0; k <= i; ++k) v += corr.L[i * 6 + k] * z[k];
    blk[i * bs + it] = (float)(truth[i] + sig[i] * v);
  }
  for (int i = 0; i < 6; ++i)  // cov = diag(sig) (L L^T) diag(sig)
    for (int j = 0; j <= i; ++j) {
      double v = 0;
      for (int k = 0; k <= j; ++k) v += corr.L[i * 6 + k] * corr.L[j * 6 + k];
      blk[(6 + i * (i + 1) / 2 + j) * bs + it] = (float)(sig[i] * sig[j] * v);
    }
}

void fill_hit(float* hb, int bs, int it, uint64_t item, int layer, const LayerHits& lh, const p2r_gen_config& c,
              const double truth[6], int q) {
  double p[3];
  if (c.mode == P2R_GEN_LITERAL) {
    for (int f = 0; f < 3; ++f)
      p[f] = lh.base[3 * layer + f] * (1.0 + c.smear * normal(c.seed, kHitPos, item, (uint64_t)layer * 3 + f));
  } else {
    const double k = q * (double)p2r_default_params().kfact / (double)c.b_scale_truth;
    const double r = (double)c.r0 + layer * (double)c.dr;
    if (!helix_cylinder(truth, k, r, p)) {  // never reaches r: put the hit radially outward of the start
      const double rr = std::hypot(truth[0], truth[1]);
      p[0] = truth[0] * r / rr;
      p[1] = truth[1] * r / rr;
      p[2] = truth[2];
    }
    const double sh = std::sqrt((double)c.hit_var);
    for (int f = 0; f < 3; ++f) p[f] += sh * normal(c.seed, kHitPos, item, (uint64_t)layer * 3 + f);
  }
  if (c.layer_offsets) {
    const float* o = c.layer_offsets + 4 * layer;  // dx, dy, dz, dphi
    const double cs = std::cos((double)o[3]), sn = std::sin((double)o[3]);
    const double x = cs * p[0] - sn * p[1], y = sn * p[0] + cs * p[1];
    p[0] = x + o[0];
    p[1] = y + o[1];
    p[2] = p[2] + o[2];
  }
  for (int f = 0; f < 3; ++f) hb[f * bs + it] = (float)p[f];
  const float cv[6] = {c.hit_var, 0.f, c.hit_var, 0.f, 0.f, c.hit_var};
  for (int f = 0; f < 6; ++f) hb[(3 + f) * bs + it] = cv[f];
}

bool shape_ok(const p2r_shape* s) {
  return s && s->nevts >= 0 && s->ntrks >= 0 && s->nlayer >= 1 && s->nlayer <= 1024 && s->bsize >= 1 &&
         s->bsize <= 1024 && s->ntrks % s->bsize == 0;
}

}  // namespace

extern "C" {

int p2r_gen_reference(float* trk, float* hits, const p2r_shape* s) {
  if (!trk || !hits || !s) return P2R_ERR_ARG;
  if (!shape_ok(s) || s->nlayer > 22) return P2R_ERR_SHAPE;
  const int bs = s->bsize, nb = s->ntrks / bs, nlayer = s->nlayer;
  const double smear = 0.00001;  // tbb.cpp:27
  srand(1);
  PolarRandn randn;
  float tpar[6], tcov[21];
  for (int i = 0; i < 6; ++i) tpar[i] = (float)kC0TrackPar[i];
  for (int i = 0; i < 21; ++i) tcov[i] = (float)kC0TrackCov[i];
  for (int64_t ie = 0; ie < s->nevts; ++ie)  // tbb.cpp:246-261
    for (int64_t ib = 0; ib < nb; ++ib) {
      float* blk = trk + (ib + (int64_t)nb * ie) * 28 * bs;
      for (int it = 0; it < bs; ++it) {
        for (int ip = 0; ip < 6; ++ip) blk[ip * bs + it] = (1 + smear * randn(0, 1)) * tpar[ip];
        for (int ip = 0; ip < 21; ++ip) blk[(6 + ip) * bs + it] = (1 + smear * randn(0, 1)) * tcov[ip] * 100;
        const int q = kC0TrackQ;
        std::memcpy(blk + 27 * bs + it, &q, sizeof(int));
      }
    }
  for (int lay = 0; lay < nlayer; ++lay) {  // tbb.cpp:268-291
    float hp[3], hc[6];
    for (int i = 0; i < 3; ++i) hp[i] = (float)kC0Hit[21 - lay][i];
    for (int i = 0; i < 6; ++i) hc[i] = (float)kC0Hit[21 - lay][3 + i];
    for (int64_t ie = 0; ie < s->nevts; ++ie)
      for (int64_t ib = 0; ib < nb; ++ib) {
        float* blk = hits + (lay + nlayer * (ib + (int64_t)nb * ie)) * 9 * bs;
        for (int it = 0; it < bs; ++it) {
          for (int ip = 0; ip < 3; ++ip) blk[ip * bs + it] = (1 + smear * randn(0, 1)) * hp[ip];
          for (int ip = 0; ip < 6; ++ip) blk[(3 + ip) * bs + it] = (1 + smear * randn(0, 1)) * hc[ip];
        }
      }
  }
  return P2R_OK;
}

int p2r_reference_hit(int layer, float pos[3]) {
  if (!pos || layer < 0 || layer > 21) return P2R_ERR_ARG;
  for (int i = 0; i < 3; ++i) pos[i] = (float)kC0Hit[21 - layer][i];
  return P2R_OK;
}

p2r_gen_config p2r_gen_default_config(void) {
  p2r_gen_config c;
  c.mode = P2R_GEN_CONSISTENT;
  c.seed = 0x50325221ull;
  c.smear = 0.1f;
  c.r0 = 20.f;
  c.dr = 5.f;
  c.hit_var = 1e-4f;
  c.layer_offsets = nullptr;
  c.b_scale_truth = 1.f;
  return c;
}

int p2r_gen_uniform(float* trk, float* hits, const p2r_shape* s, const p2r_gen_config* cfg) {
  if (!trk || !hits || !s) return P2R_ERR_ARG;
  if (!shape_ok(s)) return P2R_ERR_SHAPE;
  const p2r_gen_config c = cfg ? *cfg : p2r_gen_default_config();
  if (!(c.b_scale_truth > 0.f)) return P2R_ERR_ARG;
  const int bs = s->bsize, nlayer = s->nlayer;
  const int64_t nb = s->ntrks / bs, nblk = (int64_t)s->nevts * nb;
  const LayerHits lh = layer_bases(nlayer, c);
  run_parallel(nblk, [&](int64_t b0, int64_t b1) {
    for (int64_t b = b0; b < b1; ++b) {
      const int64_t ev = b / nb, ib = b % nb;
      for (int it = 0; it < bs; ++it) {
        const uint64_t item = track_item(ev, ib * bs + it);
        double truth[6];
        int q;
        fill_track(trk + b * 28 * bs, bs, it, item, c, truth, q);
        for (int l = 0; l < nlayer; ++l)
          fill_hit(hits + (b * nlayer + l) * 9 * bs, bs, it, item, l, lh, c, truth, q);
      }
    }
  });
  return P2R_OK;
}

int p2r_gen_ragged_layout(const int32_t* ntrk_per_evt, int32_t nevts, int32_t bsize, int64_t* blk_offsets,
                          int64_t* nblk) {
  if (!ntrk_per_evt || !blk_offsets || !nblk || nevts < 0 || bsize < 1 || bsize > 1024) return P2R_ERR_ARG;
  int64_t acc = 0;
  for (int32_t e = 0; e < nevts; ++e) {
    if (ntrk_per_evt[e] < 0) return P2R_ERR_SHAPE;
    blk_offsets[e] = acc;
    acc += (ntrk_per_evt[e] + bsize - 1) / bsize;
  }
  blk_offsets[nevts] = acc;
  *nblk = acc;
  return P2R_OK;
}

int p2r_gen_ragged(float* trk, float* hits, uint32_t* layer_mask, const int32_t* ntrk_per_evt, int32_t nevts,
                   const int64_t* blk_offsets, int32_t nlayer, int32_t bsize, int32_t kmin, int32_t kmax,
                   const p2r_gen_config* cfg) {
  if (!trk || !hits || !layer_mask || !ntrk_per_evt || !blk_offsets) return P2R_ERR_ARG;
  if (nevts < 0 || nlayer < 1 || nlayer > 32 || bsize < 1 || bsize > 1024 || kmin < 0 || kmax < kmin ||
      kmax > nlayer)
    return P2R_ERR_SHAPE;
  const p2r_gen_config c = cfg ? *cfg : p2r_gen_default_config();
  const int bs = bsize;
  const LayerHits lh = layer_bases(nlayer, c);
  run_parallel(nevts, [&](int64_t e0, int64_t e1) {
    for (int64_t ev = e0; ev < e1; ++ev) {
      const int64_t b0 = blk_offsets[ev], b1 = blk_offsets[ev + 1];
      for (int64_t b = b0; b < b1; ++b)
        for (int it = 0; it < bs; ++it) {
          const int64_t tk = (b - b0) * bs + it;
          const uint64_t item = track_item(ev, tk);
          const int64_t slot = b * bs + it;
          double truth[6];
          int q;
          fill_track(trk + b * 28 * bs, bs, it, item, c, truth, q);
          for (int l = 0; l < nlayer; ++l)
            fill_hit(hits + (b * nlayer + l) * 9 * bs, bs, it, item, l, lh, c, truth, q);
          if (tk >= ntrk_per_evt[ev]) {
            layer_mask[slot] = 0u;  // padding lane
            continue;
          }
          const int nk = kmin + (int)(counter(c.seed, kMask, item, 0) % (uint64_t)(kmax - kmin + 1));
          int order[32];
          for (int l = 0; l < nlayer; ++l) order[l] = l;
          for (int l = 0; l < nk; ++l) {  // partial Fisher-Yates
            const int j = l + (int)(counter(c.seed, kMaskPerm, item, l) % (uint64_t)(nlayer - l));
            std::swap(order[l], order[j]);
          }
          uint32_t m = 0;
          for (int l = 0; l < nk; ++l) m |= 1u << order[l];
          layer_mask[slot] = m;
        }
    }
  });
  return P2R_OK;
}

int p2r_gen_lognormal_counts(int32_t* out, int32_t nevts, double median, double sigma, int32_t lo, int32_t hi,
                             uint64_t seed) {
  if (!out || nevts < 0 || !(median > 0) || sigma < 0 || lo > hi) return P2R_ERR_ARG;
  for (int32_t e = 0; e < nevts; ++e) {
    const double v = std::exp(std::log(median) + sigma * normal(seed, kCount, (uint64_t)e, 0));
    out[e] = (int32_t)std::min<double>(hi, std::max<double>(lo, std::round(v)));
  }
  return P2R_OK;
}

}  // extern "C"
