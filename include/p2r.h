/*
 * p2r on MI355X -- C-ABI of the propagate-to-radius + Kalman-update hot path.
 *
 * Data layout (byte-identical to the reference's Matriplex,
 * cerati/p2r-tests src/propagate-tor-test_tbb.cpp:68-125), bs = bsize:
 *   MPTRK block: float par[6*bs] | float cov[21*bs] | int32 q[bs]   (112*bs bytes)
 *   MPHIT block: float pos[3*bs] | float cov[6*bs]                  ( 36*bs bytes)
 *   field f of lane it at data[f*bs + it]; par = (x, y, z, 1/pt, phi, theta);
 *   cov packed lower-triangular symmetric, (i,j) i>=j at i*(i+1)/2 + j.
 *   Uniform batch: track block (ev, ib) at index ib + nb*ev, nb = ntrks/bs
 *   (tbb.cpp:147-153); hit block (ev, ib, layer) at layer + nlayer*(ib + nb*ev)
 *   (tbb.cpp:215-217).  A flat track index is block*bs + lane (= ev*ntrks + tk).
 *
 * Conventions: every pointer passed to an entry point that takes a `stream` is
 * device memory; the caller owns all buffers (as in the reference driver);
 * calls enqueue on `stream` (a hipStream_t, NULL = default stream) and do not
 * synchronize.  Every entry point returns a p2r_status and never aborts.
 */
#ifndef P2R_H
#define P2R_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum p2r_status {
  P2R_OK = 0,
  P2R_ERR_ARG = 1,         /* null pointer / invalid enum value */
  P2R_ERR_SHAPE = 2,       /* sizes the path does not accept (e.g. ntrks % bsize != 0) */
  P2R_ERR_HIP = 3,         /* a HIP runtime call failed; see p2r_last_error */
  P2R_ERR_UNSUPPORTED = 4, /* valid request the library does not implement */
  P2R_ERR_NOMEM = 5        /* workspace allocation failed */
} p2r_status;

/* Layer semantics.
 * P2R_REFERENCE: every layer restarts from the input track and the output is
 *   the last layer's update -- the reference CPU path (tbb.cpp:880-911).
 * P2R_CHAINED: layer L starts from layer L-1's updated state -- the in-place
 *   GPU variant (propagate-tor-test_hip.cu:839-845), a real Kalman filter. */
typedef enum p2r_semantics { P2R_REFERENCE = 0, P2R_CHAINED = 1 } p2r_semantics;

typedef struct p2r_ctx p2r_ctx;

/* Batch shape: the reference's compile-time nevts, ntrks, nlayer, bsize
 * (tbb.cpp:15-34), here runtime values.  bsize in [1, 1024]; nlayer in
 * [1, P2R_MAX_NLAYER] (at most 32 for p2r_forward_ragged, whose layer mask has
 * 32 bits); other values return P2R_ERR_SHAPE. */
enum { P2R_MAX_NLAYER = 64 };
typedef struct p2r_shape {
  int32_t nevts;
  int32_t ntrks;
  int32_t nlayer;
  int32_t bsize;
} p2r_shape;

/* Arithmetic of the per-lane math.
 * P2R_EXACT: every FP32 operation rounds as the reference's source does (no
 *   fused multiply-add, IEEE division and sqrt, FP64 determinant, sin/cos
 *   rounded from FP64); bit-identical to the IEEE-strict CPU build of the
 *   reference algorithm wherever the CPU's sinf/cosf are correctly rounded.
 * P2R_FAST: fused multiply-adds, hardware reciprocal/sqrt and an FP32 sin/cos
 *   (within 2 units of 2^-24 of the true values for |angle| < 6000; larger
 *   angles lose accuracy, NaN and infinities give NaN); same algorithm,
 *   different FP32 rounding. */
typedef enum p2r_precision { P2R_EXACT = 0, P2R_FAST = 1 } p2r_precision;

/* kfact: curvature constant, 100/(-0.299792458*Bz) in the reference
 * (tbb.cpp:654, Bz = 3.8112 T -> -87.52206); b_scale multiplies the field:
 * the kernel uses kfact / b_scale.  p2r_default_params() gives the reference's
 * constant, b_scale 1 and P2R_EXACT. */
typedef struct p2r_params {
  float kfact;
  float b_scale;
  int32_t precision;
  int32_t options; /* bitwise OR of p2r_option; 0 (the reference's behaviour) by default */
} p2r_params;

/* Physics the reference leaves commented out after the parameter update
 * (tbb.cpp:527 note on the charge, :635 squashPhiMPlex), applied to every
 * layer's updated state when set; forward entry points only (the adjoint
 * returns P2R_ERR_UNSUPPORTED for options != 0).
 * P2R_OPT_WRAP_PHI: phi <- phi - 2pi floor((phi + pi) / 2pi), i.e. into
 *   [-pi, pi) up to FP32 rounding (the multiply by 1/2pi is an FP32 constant).
 * P2R_OPT_CHARGE_FROM_IPT: an update that makes 1/pt negative flips the charge
 *   and 1/pt's sign; the next layer propagates with the new charge, and the
 *   output block carries it.
 * P2R_OPT_CHI2_PER_LAYER: `chi2` (when given) receives each update's chi2
 *   (the reference's Chi2Similarity, commented out at tbb.cpp:631) instead of
 *   the per-track sum: chi2[L * n + i] for layer L and track (or slot) i, with
 *   n = nevts*ntrks in p2r_forward and the slot count in p2r_forward_ragged;
 *   absent layers of an uneven batch get 0.  The buffer holds nlayer * n
 *   floats. */
typedef enum p2r_option {
  P2R_OPT_WRAP_PHI = 1,
  P2R_OPT_CHARGE_FROM_IPT = 2,
  P2R_OPT_CHI2_PER_LAYER = 4
} p2r_option;

/* Context: device binding and workspace.  Replaces the reference's
 * p2r_get_compute_device_id / p2r_get_streams set-up
 * (propagate-tor-test_hip_local.cu:112-136).  With P2R_ROCTX=1 in the
 * environment at creation, every entry point called on the context is
 * wrapped in a roctx range named after it (rocprofv3 --marker-trace); the
 * marker library is opened at run time. */
int p2r_create(int device, p2r_ctx** ctx);
int p2r_destroy(p2r_ctx* ctx);
p2r_params p2r_default_params(void);
const char* p2r_error_string(int status);
/* Detail text of the last failure on this context (never NULL). */
const char* p2r_last_error(const p2r_ctx* ctx);

/* Whole batch, all layers in one kernel: the reference's layer loop
 * (tbb.cpp:873-925 per repetition; GPU twins launch_p2r_kernel,
 * propagate-tor-test_hip_local.cu:911-939, and GPUsequence,
 * propagate-tor-test_hip.cu:820-849).  trk_in / trk_out: nevts*nb MPTRK
 * blocks; hits: nlayer*nevts*nb MPHIT blocks.  trk_out may equal trk_in
 * (also when the context's tuning splits the layers over several launches).
 * chi2 (nullable): per-track sum over layers of r^T S^-1 r (the local chi2
 * the reference leaves commented out, tbb.cpp:631), float[nevts*ntrks] by flat
 * track index.  Output q equals input q (the reference copies it, tbb.cpp:880). */
int p2r_forward(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* shape,
                int semantics, const p2r_params* params, float* chi2, void* stream);

/* Uneven batch in one persistent launch.  Events of different sizes, each
 * padded to whole Matriplex blocks: event e owns blocks
 * [blk_offsets[e], blk_offsets[e+1]) (nevts+1 offsets, blk_offsets[0] = 0,
 * non-decreasing); a flat slot index is block*bsize + lane.  hits: nlayer
 * MPHIT blocks per Matriplex block, dense (hit block index block*nlayer + L).
 * layer_mask[slot]: bit L set when layer L has a hit; absent layers are skipped
 * (no propagation, no update); padding lanes carry mask 0 and are copied
 * through unchanged.  semantics / params / chi2 as in p2r_forward (chi2 indexed
 * by slot).  `tile_rounds` picks the schedule:
 *   P2R_RAGGED_LAYER_ORDER (-1): lanes in slot order; each wavefront walks the
 *     union of its lanes' present layers and a lane updates at its own ones
 *     (hit loads coalesced as in p2r_forward, lanes idle at layers they lack);
 *   1, 2 or 4: persistent workgroups over tiles of 256*tile_rounds slots whose
 *     tracks are ordered by hit count before they are dealt to wavefronts
 *     (no idle lanes; each lane reads its own layers, so hit loads gather
 *     from many cache lines);
 *   P2R_RAGGED_PAIR_STREAM (-2): each wavefront keeps the state of a tile of
 *     112 slots in LDS and deals the tile's present (layer, slot) pairs, in
 *     layer-major order, 64 per step to its lanes (no idle lanes; the hit rows
 *     of a step belong to at most two layers, so loads stay near-coalesced);
 *   P2R_RAGGED_AUTO (0): the library's choice (the context's ragged_schedule).
 * The reference has no uneven batches
 * (every event has ntrks tracks and every track nlayer hits, tbb.cpp:15-34);
 * with all masks full and equal event sizes the result equals p2r_forward's. */
enum { P2R_RAGGED_AUTO = 0, P2R_RAGGED_LAYER_ORDER = -1, P2R_RAGGED_PAIR_STREAM = -2 };
int p2r_forward_ragged(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const int64_t* blk_offsets,
                       int32_t nevts, const uint32_t* layer_mask, int32_t nlayer, int32_t bsize, int semantics,
                       const p2r_params* params, float* chi2, int32_t tile_rounds, void* stream);

/* Gradient precision of p2r_forward_adjoint.  P2R_GRAD_MIXED is the one whose
 * per-track gradients stay within 1e-3 of FP64 central differences (the
 * Python mirror's default); all-FP32 gradients of ill-conditioned tracks can
 * be off by ~1e-2, because the FP32 covariance update cancels.
 * The forward pass (trk_out, chi2) uses params->precision as p2r_forward does;
 * the reverse sweep uses fused multiply-adds whatever that precision.
 * P2R_GRAD_F32:   every layer in FP32.
 * P2R_GRAD_MIXED: the state carried from layer to layer (and the checkpoints)
 *                 in FP64, each layer's reverse sweep in FP32 but for the
 *                 first two layers' (FP64), where the fit is least constrained
 *                 and the covariance update cancels most.
 * P2R_GRAD_F64:   everything in FP64.
 * The FP64 arithmetic divides, takes square roots and sin/cos within a few
 * ulps (refined hardware estimates, its own reduction), not IEEE-rounded. */
typedef enum p2r_grad_precision { P2R_GRAD_F32 = 0, P2R_GRAD_MIXED = 1, P2R_GRAD_F64 = 2 } p2r_grad_precision;

/* Forward pass plus hand-written reverse-mode adjoint of the whole chained fit
 * (P2R_CHAINED semantics, uniform batch as in p2r_forward).  Objective per
 * track: the sum over layers of the local chi2 r^T S^-1 r of each update.
 * Memory: the track state is checkpointed every `checkpoint_stride` layers
 * (1, 2, 4 or 8; 0 = 4) in context workspace; the reverse sweep recomputes
 * each segment from its checkpoint instead of storing every intermediate
 * (stride 1 stores every layer's input state and recomputes nothing but the
 * layer's own intermediates).  Workspace: (ceil(nlayer/stride) - 1) * 27 *
 * ntrk states of 4 bytes (P2R_GRAD_F32) or 8 bytes, plus (stride - 2) * 27 *
 * ntrk values for the recomputed layer inputs of a segment where they do not
 * fit in LDS (stride 8, and stride 4 with F64), of 4 bytes (F32, MIXED) or 8
 * (F64).  The context holds one
 * workspace: adjoint calls on one context must not overlap in time (issue
 * them on one stream, or synchronize between streams or threads).
 * trk_out must not overlap trk_in (the reverse sweep rereads the input
 * tracks after the forward pass has written trk_out): P2R_ERR_ARG.
 * Outputs (device memory; any may be NULL except trk_out):
 *   trk_out     final fitted tracks: with P2R_GRAD_F32 and P2R_EXACT word for
 *               word those of p2r_forward (P2R_CHAINED); with P2R_GRAD_F32 and
 *               P2R_FAST the same fused formulas, which the compiler may fuse
 *               differently in this kernel (fast-mode accuracy, not p2r_forward's
 *               bits); with MIXED and F64 the fit carried in FP64, rounded to
 *               float.  With P2R_EXACT and P2R_GRAD_F32 the gradients are the same
 *               words at every checkpoint stride; otherwise they agree across
 *               strides to the arithmetic's precision.
 *   chi2        float[ntrk]: total chi2 per track (flat track index)
 *   grad_par    float, MP6F blocks: d chi2 / d (initial x, y, z, 1/pt, phi, theta)
 *   grad_hits   float, one MP3F block per hit block (index block*nlayer + L):
 *               d chi2 / d (hit x, y, z)
 *   grad_kfact  float[ntrk]: d chi2 / d kfact (at the kfact/b_scale used)
 *   grad_global double[4*nlayer + 1], summed over all tracks in FP64 in a fixed
 *               order (bit-reproducible): for each layer L the derivative with
 *               respect to a rigid offset of that layer's hits --
 *               [4L+0..2] translation dx, dy, dz, [4L+3] rotation dphi about z
 *               (x' = x cos dphi - y sin dphi, ...) -- then [4*nlayer] the
 *               derivative with respect to b_scale (field = nominal * b_scale).
 *               Tracks whose total chi2 is not finite (diverged fits) are left
 *               out of these sums; their per-track outputs are still written.
 * The reference has no derivatives (this path is new). */
int p2r_forward_adjoint(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* shape,
                        const p2r_params* params, int grad_precision, int32_t checkpoint_stride, float* chi2,
                        float* grad_par, float* grad_hits, float* grad_kfact, double* grad_global, void* stream);

/* p2r_forward_adjoint without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, checks and results as p2r_forward_adjoint, but the checkpoints,
 * chi2 totals and global-sum partials live in a static device array of
 * P2R_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on
 * until the process ends (its address is resolved once per device at
 * p2r_create); the call itself allocates nothing.  A call that needs more
 * scratch returns P2R_ERR_SHAPE and launches nothing.  Scratch of a call:
 * (ceil(nlayer/stride) - 1) * 27 * ntrk checkpoint states plus ntrk chi2
 * totals, of 4 bytes (P2R_GRAD_F32) or 8, plus 8 * (4*nlayer + 1) *
 * ceil(ntrk/256) bytes when grad_global is requested, each part rounded up to
 * 256 bytes, plus the segment states of p2r_forward_adjoint's workspace where
 * they are not in LDS, plus 256: at most about 150 000 tracks x 20 layers at
 * stride 4 with P2R_GRAD_F32, about 68 000 with P2R_GRAD_MIXED.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (any stream, thread or context) race and must be serialized by
 * the caller.  Nothing in the array needs clearing between calls. */
enum { P2R_NOALLOC_SCRATCH_BYTES = 64 << 20 };
int p2r_forward_adjoint_noalloc(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out,
                                const p2r_shape* shape, const p2r_params* params, int grad_precision,
                                int32_t checkpoint_stride, float* chi2, float* grad_par, float* grad_hits,
                                float* grad_kfact, double* grad_global, void* stream);

/* Block-level mirrors of the reference entry points, over nblk Matriplex
 * blocks.  Pointers address the field arrays of the first block; consecutive
 * blocks are `trk_stride` floats apart for track-side arrays and `hit_stride`
 * floats apart for hit-side arrays (e.g. an MPTRK array: trk_stride = 28*bs;
 * layer L of a uniform hit array: msP = hits + L*9*bs, hit_stride = nlayer*9*bs).
 *
 * propagateToR(inErr, inPar, inChg, msP, outErr, outPar), tbb.cpp:656-657.
 * Output arrays use the same stride as the inputs.  Both block-level entry
 * points use P2R_EXACT arithmetic. */
int p2r_propagateToR(p2r_ctx* ctx, const float* inErr, const float* inPar, const int32_t* inChg, const float* msP,
                     float* outErr, float* outPar, int32_t bsize, int64_t nblk, int64_t trk_stride,
                     int64_t hit_stride, float kfact, void* stream);

/* KalmanUpdate(trkErr, inPar, hitErr, msP), in place, tbb.cpp:433. */
int p2r_KalmanUpdate(p2r_ctx* ctx, float* trkErr, float* inPar, const float* hitErr, const float* msP,
                     int32_t bsize, int64_t nblk, int64_t trk_stride, int64_t hit_stride, void* stream);

/* Host-to-host forward in the reference's streamed form (its default timing
 * mode, propagate-tor-test_hip.cu:905-1013 with MEASURE_H2D/D2H_TRANSFER): the
 * batch's Matriplex blocks are cut into `nstreams` contiguous chunks; each
 * chunk's tracks and hits are copied host -> device, run through all layers
 * and copied back on a stream of its own (created by the context on first use,
 * default priority), so the copies of one chunk overlap the kernel of another.
 * host_*: the batch in host memory, pinned (hipHostMalloc) for the copies to be
 * asynchronous; dev_*: device buffers of the same sizes, the caller's.
 * Returns when every chunk is back on the host.  nstreams in 1..16.
 * The reference times this with a wall clock around the same three steps. */
int p2r_forward_streamed(p2r_ctx* ctx, const void* host_trk_in, const void* host_hits, void* host_trk_out,
                         void* dev_trk_in, void* dev_hits, void* dev_trk_out, const p2r_shape* shape, int semantics,
                         const p2r_params* params, int32_t nstreams);

/* ---- HIP graphs ----------------------------------------------------------------
 * One forward / uneven-batch / adjoint call, captured once with its arguments
 * (pointers, shape, params, the context's tuning at capture time) into an
 * instantiated HIP graph, then replayed with p2r_graph_launch on any stream.
 * The enqueue paths issue kernel launches only (no allocation, no host
 * synchronisation, no memset nodes).  An adjoint graph owns a workspace of its
 * own, allocated before the capture and freed by p2r_graph_destroy, so later
 * calls on the context leave it alone; replays of one adjoint graph must not
 * overlap each other in time.  The ragged entry reads blk_offsets on the host
 * at capture time.
 * Not in the reference (its drivers launch directly, hip.cu:963-975). */
typedef struct p2r_graph p2r_graph;
int p2r_graph_forward(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* shape,
                      int semantics, const p2r_params* params, float* chi2, p2r_graph** out);
int p2r_graph_forward_ragged(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out,
                             const int64_t* blk_offsets, int32_t nevts, const uint32_t* layer_mask, int32_t nlayer,
                             int32_t bsize, int semantics, const p2r_params* params, float* chi2,
                             int32_t tile_rounds, p2r_graph** out);
int p2r_graph_forward_adjoint(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out,
                              const p2r_shape* shape, const p2r_params* params, int grad_precision,
                              int32_t checkpoint_stride, float* chi2, float* grad_par, float* grad_hits,
                              float* grad_kfact, double* grad_global, p2r_graph** out);
int p2r_graph_launch(p2r_graph* graph, void* stream);
int p2r_graph_destroy(p2r_graph* graph);

/* ---- Launch choices and the autotuner ---------------------------------------
 * The reference fixes Matriplex width, threads per block and layers per kernel
 * at compile time (build.py:68-80 sweeps them by rebuilding; hip_local.cu:911
 * templates the kernel on <bsize, layers>).  Here they are runtime choices held
 * by the context, set by hand or measured by the autotuners below. */
typedef struct p2r_tuning {
  int32_t fwd_threads;           /* workgroup size of p2r_forward: 64, 128 or 256 (default 256) */
  int32_t fwd_layers_per_launch; /* 0 = all layers in one launch (default); n = launches of n layers,
                                    the state passing through trk_out in between */
  int32_t ragged_schedule;       /* what P2R_RAGGED_AUTO runs: P2R_RAGGED_PAIR_STREAM (default),
                                    P2R_RAGGED_LAYER_ORDER, 1, 2, 4 */
  int32_t ragged_threads;        /* workgroup size of the layer-ordered schedule: 64, 128 or 256 (256) */
  int32_t checkpoint_stride;     /* what checkpoint_stride = 0 means in p2r_forward_adjoint: 1, 2, 4, 8 (4) */
  int32_t fwd_waves;             /* register budget of the p2r_forward kernels, in waves per SIMD:
                                    0 = the library's choice (default: 3 for P2R_EXACT, 4 for
                                    P2R_FAST), 1 = the compiler's allocation, 3 or 4 (fewer
                                    registers, possibly spills) */
} p2r_tuning;
int p2r_get_tuning(const p2r_ctx* ctx, p2r_tuning* out);
int p2r_set_tuning(p2r_ctx* ctx, const p2r_tuning* in); /* P2R_ERR_ARG on a value outside the lists above */

/* One timed candidate: the launch choices and the mean time of one call (ms). */
typedef struct p2r_tune_entry {
  int32_t bsize, threads, layers_per_launch, schedule, checkpoint_stride, waves;
  float ms;
} p2r_tune_entry;
typedef struct p2r_tune_report {
  int32_t count; /* entries used */
  int32_t best;  /* index of the fastest */
  p2r_tune_entry entry[32];
} p2r_tune_report;

/* Autotuners.  Each times its candidates on the caller's buffers (2 untimed
 * calls, then the mean of `reps` calls, HIP events on `stream`), stores the
 * fastest in the context's tuning and fills `report` (nullable).  They
 * synchronize the stream and create events: not for use under graph capture.
 * Outputs are overwritten with valid results of the last candidate run.
 * trk_out must not overlap trk_in (every candidate starts from the same
 * input): P2R_ERR_ARG.
 *   forward: fwd_threads {64, 128, 256} x fwd_layers_per_launch {all,
 *            ceil(nlayer/2), ceil(nlayer/4)}, then fwd_waves {1, 3, 4} at the
 *            best of those;
 *   ragged:  ragged_schedule {layer order x ragged_threads {64, 128, 256},
 *            pair stream, sorted tiles of 1, 2, 4 rounds};
 *   adjoint: checkpoint_stride {1, 2, 4, 8} (grad_precision as given). */
int p2r_autotune_forward(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* shape,
                         int semantics, const p2r_params* params, int32_t reps, p2r_tune_report* report,
                         void* stream);
int p2r_autotune_ragged(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const int64_t* blk_offsets,
                        int32_t nevts, const uint32_t* layer_mask, int32_t nlayer, int32_t bsize, int semantics,
                        const p2r_params* params, int32_t reps, p2r_tune_report* report, void* stream);
int p2r_autotune_adjoint(p2r_ctx* ctx, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* shape,
                         const p2r_params* params, int grad_precision, float* chi2, float* grad_par,
                         float* grad_hits, float* grad_kfact, double* grad_global, int32_t reps,
                         p2r_tune_report* report, void* stream);

/* Matriplex width.  The width is the caller's data layout, so it is not a
 * context setting: p2r_repack converts a batch between widths on the device,
 * and p2r_autotune_width times p2r_forward (current tuning) on copies of the
 * caller's batch repacked to each width in `widths` (each must divide ntrks;
 * at most 32), reporting the fastest width.  It allocates its copies with
 * hipMalloc and frees them before returning.
 *   p2r_repack: `nsub` sub-blocks of `words` words per lane (MPTRK: nsub 1,
 *   words 28; MPHIT: nsub nlayer, words 9) for nevts events of ntrks tracks;
 *   track t of event e moves from (block e*ntrks/src_bsize + t/src_bsize,
 *   lane t%src_bsize) to the same place in the dst_bsize layout.  src and dst
 *   must not overlap. */
int p2r_repack(p2r_ctx* ctx, const void* src, int32_t src_bsize, void* dst, int32_t dst_bsize, int32_t nevts,
               int32_t ntrks, int32_t nsub, int32_t words, void* stream);
int p2r_autotune_width(p2r_ctx* ctx, const void* trk_in, const void* hits, const p2r_shape* shape, int semantics,
                       const p2r_params* params, const int32_t* widths, int32_t nwidths, int32_t reps,
                       p2r_tune_report* report, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* P2R_H */
