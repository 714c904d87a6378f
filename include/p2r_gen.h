/*
 * p2r on MI355X -- host-side synthetic input generators (C-ABI).
 *
 * Buffers are host memory in the Matriplex layout of p2r.h.  These replace the
 * reference's prepareTracks / prepareHits (propagate-tor-test_tbb.cpp:243-293)
 * and add the seeded generator of the benchmark workloads.
 */
#ifndef P2R_GEN_H
#define P2R_GEN_H

#include <stdint.h>

#include "p2r.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reference-compatible inputs ("C0"): the input_track.h track and hits
 * (layer l uses inputhit[21-l], tbb.cpp:818-820), smeared by 1e-5 with the
 * reference's polar randn on glibc rand() in its unseeded state, covariance x100,
 * q = 1 -- the same values, in the same order, as tbb.cpp:243-293.
 * Requires nlayer <= 22 and ntrks % bsize == 0.  Single-threaded (rand()). */
int p2r_gen_reference(float* trk, float* hits, const p2r_shape* shape);

/* Unsmeared position of the reference input hit used at `layer` (the point the
 * reference summary compares against, tbb.cpp:948-951). */
int p2r_reference_hit(int layer, float pos[3]);

/* Seeded counter-based generator for the benchmark workloads: the base
 * state of propagate-tor-test_acc.v3.cpp:808-812 smeared by (1 + smear*N(0,1)),
 * q = +-1 with equal probability, one hit per layer l on the cylinder
 * r_l = r0 + l*dr, hit cov diag(hit_var) (entries xx, yy, zz).  Every value is
 * a pure function of (seed, stream, event, track, field), so any subset
 * regenerates identically and generation runs on all host threads.
 *
 * Generator modes.
 * P2R_GEN_CONSISTENT (default): each track's true state is base*(1+smear*N),
 *   N truncated to |N| <= 3 (every track starts inside r = 19.5 cm);
 *   its hits lie on that track's own helix (charge q) at r_l, plus Gaussian
 *   measurement noise of variance hit_var per coordinate (the hit covariance
 *   given to the fit); the fit's input state is the true state plus noise
 *   drawn from the input covariance, which has the base variances smeared as
 *   sigma*(1+smear*N) and the base correlations shrunk to a positive-definite
 *   matrix.
 * P2R_GEN_LITERAL: the configuration text taken word for word -- par, every
 *   covariance entry and every hit coordinate (hits on the base track's
 *   crossing) multiplied by (1+smear*N) independently.  Covariances need not
 *   be positive-definite and hits are ~5 cm off any helix, so chained fits
 *   mostly diverge; kept for parity testing. */
enum { P2R_GEN_CONSISTENT = 0, P2R_GEN_LITERAL = 1 };

typedef struct p2r_gen_config {
  int32_t mode;
  uint64_t seed;
  float smear;    /* 0.1 */
  float r0, dr;   /* cylinder radii r_l = r0 + l*dr (cm) */
  float hit_var;  /* 1e-4 */
  /* alignment truth (nullable): per-layer rigid offsets applied to the hits
   * after smearing: (dx, dy, dz, dphi) x nlayer, rotation about z then shift */
  const float* layer_offsets;
  float b_scale_truth; /* field scale used to place the hits (1 = nominal) */
} p2r_gen_config;

p2r_gen_config p2r_gen_default_config(void);

/* nlayer 1..1024, ntrks % bsize == 0. */
int p2r_gen_uniform(float* trk, float* hits, const p2r_shape* shape, const p2r_gen_config* cfg);

/* Uneven batch ("ragged"): nevts events whose track counts are given in
 * ntrk_per_evt.  Each event is padded to whole Matriplex blocks: event e owns
 * blocks [blk_offsets[e], blk_offsets[e+1]) and lanes >= its count are padding
 * (layer mask 0).  hits are dense (nlayer per block, as in the uniform layout).
 * layer_mask[slot] (slot = block*bsize + lane): bit l set when layer l has a
 * hit; each real track gets a hit count drawn uniformly in [kmin, kmax] and a
 * uniformly random subset of that size.  p2r_gen_ragged_layout fills
 * blk_offsets (nevts+1 entries) and returns the total block count in *nblk. */
int p2r_gen_ragged_layout(const int32_t* ntrk_per_evt, int32_t nevts, int32_t bsize, int64_t* blk_offsets,
                          int64_t* nblk);
int p2r_gen_ragged(float* trk, float* hits, uint32_t* layer_mask, const int32_t* ntrk_per_evt, int32_t nevts,
                   const int64_t* blk_offsets, int32_t nlayer, int32_t bsize, int32_t kmin, int32_t kmax,
                   const p2r_gen_config* cfg);

/* Track counts per event: round(exp(N(ln median, sigma^2))) clipped to
 * [lo, hi], seeded. */
int p2r_gen_lognormal_counts(int32_t* out, int32_t nevts, double median, double sigma, int32_t lo, int32_t hi,
                             uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* P2R_GEN_H */
