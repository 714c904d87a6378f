// C-ABI of the p2r hot path on MI355X (declarations and layout: include/p2r.h).
//
// Kernels: one lane per track over a flat track index g = block*bs + lane.
// Each lane loads its track's 27 floats + charge from the Matriplex SoA block
// (coalesced: consecutive lanes read consecutive words of each field), keeps
// the state in VGPRs through every layer, loads the 9 hit floats of each
// layer as it starts (latency hidden by the other wavefronts of the SIMD) and
// writes the result once.  HBM traffic per track: 112 B in + 112 B out + 36 B per layer.
#include "../../include/p2r.h"
#include "p2r_ctx.hpp"
#include "p2r_lane.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <new>

using namespace p2r;

// Scratch of p2r_forward_adjoint_noalloc: device memory every GPU that loads the library holds until the
// process ends.  Host code never uses this name as an address (on the host it is a stand-in, not device
// memory): it uses the per-device address that resolve_scratch() obtained from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[P2R_NOALLOC_SCRATCH_BYTES];

// An instantiated graph of one captured call (p2r_graph_* in include/p2r.h).
struct p2r_graph {
  int device = 0;
  hipGraphExec_t exec = nullptr;  // null when the captured call launched nothing (empty batch)
  void* ws = nullptr;             // workspace owned by the graph (adjoint), freed with it
};

namespace {

constexpr int kThreads = 256;
constexpr int kAllOptions = P2R_OPT_WRAP_PHI | P2R_OPT_CHARGE_FROM_IPT | P2R_OPT_CHI2_PER_LAYER;

__device__ __forceinline__ void load_trk(const float* tb, uint32_t bs, Trk& t) {
#pragma unroll
  for (int f = 0; f < 6; ++f) t.par[f] = tb[(size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 21; ++f) t.cov[f] = tb[(size_t)(6 + f) * bs];
}
__device__ __forceinline__ void store_trk(float* ob, uint32_t bs, const Trk& t) {
#pragma unroll
  for (int f = 0; f < 6; ++f) ob[(size_t)f * bs] = t.par[f];
#pragma unroll
  for (int f = 0; f < 21; ++f) ob[(size_t)(6 + f) * bs] = t.cov[f];
}
__device__ __forceinline__ void load_hit(const float* hb, uint32_t bs, Hit& h) {
#pragma unroll
  for (int f = 0; f < 3; ++f) h.pos[f] = __builtin_nontemporal_load(hb + (size_t)f * bs);
#pragma unroll
  for (int f = 0; f < 6; ++f) h.cov[f] = __builtin_nontemporal_load(hb + (size_t)(3 + f) * bs);
}
// Hit loads through the caches with the default policy: for schedules whose hit lines are shared by
// consecutive steps of one wavefront.
__device__ __forceinline__ void load_hit_cached(const float* hb, uint32_t bs, Hit& h) {
#pragma unroll
  for (int f = 0; f < 3; ++f) h.pos[f] = hb[(size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 6; ++f) h.cov[f] = hb[(size_t)(3 + f) * bs];
}

// Hit rows of the blocks a wavefront's lanes belong to, read through a buffer resource: each lane's
// byte offset is fixed for the whole pass and the layer and field go in the scalar offset, so a
// layer's 9 loads need no address arithmetic.  The active lanes of a wavefront are a prefix of its
// 64 consecutive tracks (only the batch's end cuts it), so they span at most 63 / bs + 2 blocks:
// under 4.8 MB of hit rows for bs <= 1024 and nlayer <= P2R_MAX_NLAYER (64), which the entry points
// enforce, so the 32-bit record count and offsets cannot wrap.  The record count is that span.
struct WaveHits {
  __amdgpu_buffer_rsrc_t rs;
  uint32_t voff, field_bytes;
  __device__ __forceinline__ WaveHits(const float* hits, uint32_t g, uint32_t ntrk, uint32_t bs, int nlayer,
                                      uint32_t blk, uint32_t lane) {
    const uint32_t g0 = __builtin_amdgcn_readfirstlane(g);  // the wavefront's first track
    const uint32_t blk0 = __builtin_amdgcn_readfirstlane(g0 / bs);
    const uint32_t blk1 = (g0 + 63u < ntrk ? g0 + 63u : ntrk - 1u) / bs;
    const uint32_t row_bytes = 36u * bs;
    field_bytes = 4u * bs;
    const uint64_t a = reinterpret_cast<uint64_t>(hits + (size_t)blk0 * nlayer * 9 * bs);
    // readfirstlane returns int: each half goes back through uint32_t so the low word is not sign-extended
    const uint64_t au = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
                        (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a);
    rs = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(au), (short)0,
                                           (int)__builtin_amdgcn_readfirstlane((blk1 - blk0 + 1u) * (uint32_t)nlayer *
                                                                               row_bytes),
                                           0x00020000);
    voff = ((blk - blk0) * (uint32_t)nlayer * 9u * bs + lane) * 4u;
  }
  // Layer l's hit, streamed (non-temporal: each hit is read once).
  __device__ __forceinline__ void load(int l, Hit& h) const {
    auto ld = [&](int f) {
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                           rs, voff, __builtin_amdgcn_readfirstlane((uint32_t)(l * 9 + f) * field_bytes),
                                           2 /* nt */));
    };
#pragma unroll
    for (int f = 0; f < 3; ++f) h.pos[f] = ld(f);
#pragma unroll
    for (int f = 0; f < 6; ++f) h.cov[f] = ld(3 + f);
  }
};

// Whole batch, all layers per lane.  SEM: P2R_REFERENCE / P2R_CHAINED;
// EXACT selects the per-lane math instance (p2r_lane.hpp).
// Layers [l0, l1) of nlayer (all of them for one launch; a layer range when the
// tuning splits the loop over several launches: then chi2 accumulates if acc).
// Workgroups of 64..256 lanes (blockDim.x).
// MINW: minimum waves per SIMD the register allocation must allow (1 = the compiler's choice; 3 or 4
// trade spills for occupancy, a tuning choice).
// In reference semantics every layer restarts from the input track, whose covariance the propagation
// reads only in the transport at its end: it waits in LDS (21 consecutive words per lane, [blockDim.x][21])
// instead of holding 21 registers through the helix iterations.  The parameters stay in registers.  With
// the lane's words consecutive, one base address serves all 21 reads (the fields are immediate offsets, read
// in pairs), and the odd stride of 21 words puts consecutive lanes in distinct banks.
struct LdsCov {
  const float* row;  // the lane's 21 words
  __device__ __forceinline__ float operator[](int i) const { return row[i]; }
};

template <int SEM, bool CHI2, bool EXACT, int MINW>
__global__ __launch_bounds__(kThreads, MINW) void fwd_uniform_kernel(const float* trk, const float* __restrict__ hits,
                                                               float* out, float* __restrict__ chi2, uint32_t ntrk,
                                                               uint32_t bs, int nlayer, float kfact, int l0, int l1,
                                                               bool acc, bool store, int opts) {
  extern __shared__ float cov_lds[];  // [blockDim.x][21], reference semantics only
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntrk) return;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  Trk cur;  // chained: the state between layers
  float in_par[6];
  if (SEM == P2R_REFERENCE) {
#pragma unroll
    for (int f = 0; f < 6; ++f) in_par[f] = tb[(size_t)f * bs];
#pragma unroll
    for (int f = 0; f < 21; ++f) cov_lds[threadIdx.x * 21u + f] = tb[(size_t)(6 + f) * bs];
  } else {
    load_trk(tb, bs, cur);
  }
  const int q_in = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  int q = q_in;
  float k = (float)q * kfact;  // tbb.cpp:667

  const WaveHits wh(hits, g, ntrk, bs, nlayer, blk, lane);
  const bool per_layer = opts & P2R_OPT_CHI2_PER_LAYER;
  // a later launch of a split layer loop continues the running sum (same rounding as one launch)
  float c2 = (CHI2 && acc && !per_layer) ? chi2[g] : 0.f;
  // Each layer loads its hit at the top of the iteration.  Prefetching the next layer's hit into
  // registers (9 values live across the whole update) took the kernel to 152-168 VGPRs, 3 waves per
  // SIMD; loading it as the layer starts leaves the extra waves to hide the load latency.
  for (int l = l0; l < l1; ++l) {
    Hit h;
    wh.load(l, h);
    Trk prop;
    if (SEM == P2R_REFERENCE) {
      // tbb.cpp:894: every layer starts from the input track.  The barriers, on the input parameters in
      // place and on the covariance column's offset, make every layer recompute and reread it; nothing is
      // hoisted across layers.
#pragma unroll
      for (int f = 0; f < 6; ++f) opaque(in_par[f]);
      uint32_t col = threadIdx.x;
      asm volatile("" : "+v"(col));
      q = q_in;
      k = (float)q * kfact;
      Lane<EXACT>::propagate(in_par, LdsCov{cov_lds + col * 21u}, k, h.pos[0], h.pos[1], prop);
    } else {
      Lane<EXACT>::propagate(cur, k, h.pos[0], h.pos[1], prop);  // propagate-tor-test_hip.cu:842-845
    }
    float c = Lane<EXACT>::update(prop, h);
    if (opts) apply_options<EXACT>(opts, prop, q, k, kfact);
    if (SEM == P2R_REFERENCE) {
      // Only the last layer's result is returned in reference semantics; the
      // barrier makes every layer's result observable so the compiler cannot
      // sink the update out of the loop and skip layers 0..n-2.
      opaque(prop);
      opaque(c);
    }
    if (CHI2) {
      if (per_layer) chi2[(size_t)l * ntrk + g] = c;
      else c2 += c;
    }
    cur = prop;
  }
  if (store) {
    float* ob = out + (size_t)blk * 28 * bs + lane;
    store_trk(ob, bs, cur);
    reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
  }
  if (CHI2 && !per_layer) chi2[g] = c2;
}


// ---------------------------------------------------------------------------
// Uneven batch: persistent workgroups walk tiles of kRagThreads*ROUNDS
// consecutive slots.  Within a tile, slots are ordered by hit count (stable,
// descending) and dealt to the workgroup's lanes round by round, so the 64
// lanes of a wavefront carry tracks with (nearly) the same number of present
// layers; each lane then iterates only its own present layers.  Tiles are
// assigned statically (tile = blockIdx.x + i*gridDim.x): deterministic, no
// device-side counters to reset.
constexpr int kRagThreads = 256;
constexpr int kRagWaves = kRagThreads / 64;

template <int SEM, bool CHI2, bool EXACT>
__device__ __forceinline__ void ragged_slot(const float* __restrict__ trk, const float* __restrict__ hits,
                                            float* __restrict__ out, float* __restrict__ chi2, uint32_t slot,
                                            uint32_t m, uint32_t bs, int nlayer, float kfact, int opts,
                                            uint32_t nslot) {
  const bool per_layer = opts & P2R_OPT_CHI2_PER_LAYER;
  const uint32_t m0 = m;
  const uint32_t blk = slot / bs, lane = slot - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  Trk in;
  load_trk(tb, bs, in);
  const int q_in = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  int q = q_in;
  float k = (float)q * kfact;
  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  Trk cur = in;
  float c2 = 0.f;
  Hit hn;
  if (m) load_hit(hb + (size_t)__builtin_ctz(m) * hstride, bs, hn);
  while (m) {
    const int l = __builtin_ctz(m);
    m &= m - 1u;
    const Hit h = hn;
    if (m) load_hit(hb + (size_t)__builtin_ctz(m) * hstride, bs, hn);
    Trk src;
    if (SEM == P2R_REFERENCE) {
      opaque(in);  // in place, no copy: recompute per layer
      src = in;
      q = q_in;
      k = (float)q * kfact;
    } else {
      src = cur;
    }
    Trk prop;
    Lane<EXACT>::propagate(src, k, h.pos[0], h.pos[1], prop);
    float c = Lane<EXACT>::update(prop, h);
    if (opts) apply_options<EXACT>(opts, prop, q, k, kfact);
    if (SEM == P2R_REFERENCE) {
      opaque(prop);
      opaque(c);
    }
    if (CHI2) {
      if (per_layer) chi2[(size_t)l * nslot + slot] = c;
      else c2 += c;
    }
    cur = prop;
  }
  float* ob = out + (size_t)blk * 28 * bs + lane;
  store_trk(ob, bs, cur);
  reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
  if (CHI2) {
    if (!per_layer) chi2[slot] = c2;
    else
      for (int l = 0; l < nlayer; ++l)
        if (!((m0 >> l) & 1u)) chi2[(size_t)l * nslot + slot] = 0.f;
  }
}

template <int SEM, bool CHI2, bool EXACT, int ROUNDS>
__global__ __launch_bounds__(kRagThreads) void fwd_ragged_kernel(const float* __restrict__ trk,
                                                                 const float* __restrict__ hits,
                                                                 float* __restrict__ out, float* __restrict__ chi2,
                                                                 const uint32_t* __restrict__ mask, uint32_t nslot,
                                                                 uint32_t bs, int nlayer, float kfact,
                                                                 uint32_t ntiles, int opts) {
  constexpr int kTile = kRagThreads * ROUNDS;
  __shared__ uint32_t hist[kRagWaves * ROUNDS][33];  // per (wave, round): tracks with each hit count
  __shared__ uint32_t start[33];                       // first rank of each hit count in the tile
  __shared__ uint16_t order[kTile];                    // rank -> tile-local slot
  __shared__ uint32_t smask[kTile];
  const int tid = threadIdx.x, wave = tid >> 6, wl = tid & 63;
  const uint64_t lt_mask = (wl == 0) ? 0ull : (~0ull >> (64 - wl));
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t base = tile * (uint32_t)kTile;
    uint32_t key[ROUNDS], lrank[ROUNDS];
    // bits at or above nlayer name no hit block: ignored
    const uint32_t layer_bits = nlayer >= 32 ? ~0u : ((1u << nlayer) - 1u);
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
      const uint32_t s = base + r * kRagThreads + tid;
      const uint32_t m = s < nslot ? (mask[s] & layer_bits) : 0u;
      smask[r * kRagThreads + tid] = m;
      key[r] = s < nslot ? (uint32_t)__builtin_popcount(m) : 33u;  // out-of-range slots sort last
      for (uint32_t v = 0; v < 33; ++v) {
        const uint64_t b = __ballot(key[r] == v);
        if (key[r] == v) lrank[r] = (uint32_t)__popcll(b & lt_mask);
        if (wl == 0) hist[r * kRagWaves + wave][v] = (uint32_t)__popcll(b);
      }
    }
    __syncthreads();
    if (tid < 33) {  // total per count, then exclusive prefix in descending count order
      uint32_t t = 0;
      for (int g = 0; g < kRagWaves * ROUNDS; ++g) t += hist[g][tid];
      start[tid] = t;
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t acc = 0;
      for (int v = 32; v >= 0; --v) {
        const uint32_t t = start[v];
        start[v] = acc;
        acc += t;
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
      if (key[r] < 33) {
        uint32_t rank = start[key[r]] + lrank[r];
        const int g0 = r * kRagWaves + wave;
        for (int g = 0; g < g0; ++g) rank += hist[g][key[r]];
        order[rank] = (uint16_t)(r * kRagThreads + tid);
      }
    }
    __syncthreads();
    const uint32_t nvalid = nslot > base ? min((uint32_t)kTile, nslot - base) : 0u;
#pragma unroll 1
    for (int r = 0; r < ROUNDS; ++r) {
      const uint32_t rank = r * kRagThreads + tid;
      if (rank < nvalid) {
        const uint32_t local = order[rank];
        ragged_slot<SEM, CHI2, EXACT>(trk, hits, out, chi2, base + local, smask[local], bs, nlayer, kfact, opts,
                                      nslot);
      }
    }
    __syncthreads();  // order/smask/hist are rewritten by the next tile
  }
}


// Uneven batch, layer-ordered schedule: lanes keep slot order (consecutive
// lanes = consecutive tracks of a Matriplex block), and each wavefront walks
// the union of its lanes' present layers in increasing order; a lane updates
// only at its own present layers.  Every hit load is then as coalesced as in
// the uniform kernel; the price is idle lanes at layers other lanes of the
// wavefront have and they lack.
template <int SEM, bool CHI2, bool EXACT>
__global__ __launch_bounds__(kThreads) void fwd_ragged_layer_kernel(const float* __restrict__ trk,
                                                                    const float* __restrict__ hits,
                                                                    float* __restrict__ out, float* __restrict__ chi2,
                                                                    const uint32_t* __restrict__ mask, uint32_t nslot,
                                                                    uint32_t bs, int nlayer, float kfact, int opts) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;  // workgroups of 64..256 lanes
  const bool active = g < nslot;
  // bits at or above nlayer name no hit block: ignored
  const uint32_t layer_bits = nlayer >= 32 ? ~0u : ((1u << nlayer) - 1u);
  const uint32_t m = active ? (mask[g] & layer_bits) : 0u;
  uint32_t wm = m;  // union over the wavefront (all 64 lanes take part in the shuffles)
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) wm |= (uint32_t)__shfl_xor((int)wm, o);
  wm = __builtin_amdgcn_readfirstlane(wm);
  if (!active) return;
  const bool per_layer = opts & P2R_OPT_CHI2_PER_LAYER;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  Trk in;
  load_trk(tb, bs, in);
  const int q_in = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  int q = q_in;
  float k = (float)q * kfact;
  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  Trk cur = in;
  float c2 = 0.f;
  uint32_t w = wm;
  while (w) {
    const int l = __builtin_ctz(w);
    w &= w - 1u;
    if ((m >> l) & 1u) {
      Hit h;  // loaded as the layer starts (no prefetch: fewer registers, more waves; measured faster)
      load_hit(hb + (size_t)l * hstride, bs, h);
      Trk src;
      if (SEM == P2R_REFERENCE) {
        opaque(in);  // in place, no copy: recompute per layer
        src = in;
        q = q_in;
        k = (float)q * kfact;
      } else {
        src = cur;
      }
      Trk prop;
      Lane<EXACT>::propagate(src, k, h.pos[0], h.pos[1], prop);
      float c = Lane<EXACT>::update(prop, h);
      if (opts) apply_options<EXACT>(opts, prop, q, k, kfact);
      if (SEM == P2R_REFERENCE) {
        opaque(prop);
        opaque(c);
      }
      if (CHI2) {
        if (per_layer) chi2[(size_t)l * nslot + g] = c;
        else c2 += c;
      }
      cur = prop;
    }
  }
  float* ob = out + (size_t)blk * 28 * bs + lane;
  store_trk(ob, bs, cur);
  reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
  if (CHI2) {
    if (!per_layer) chi2[g] = c2;
    else
      for (int l = 0; l < nlayer; ++l)
        if (!((m >> l) & 1u)) chi2[(size_t)l * nslot + g] = 0.f;
  }
}

// ---------------------------------------------------------------------------
// Uneven batch, pair-stream schedule.  Each wavefront owns a tile of kPsTile
// slots whose state (27 floats, charge, chi2 running sum) lives in LDS for the
// whole pass.  The tile's present (layer, slot) pairs are taken in layer-major
// order, 64 at a time, and dealt to the lanes: no lane idles on an absent
// layer, and the lanes of one step read the hit rows of at most two layers of
// the tile's blocks (near-coalesced).  A step takes its pairs from a window that
// holds each of the tile's slots once (slots s0.. of one layer, then slots ..s0-1
// of the next), so no slot occurs twice in one step; steps run in order within the wavefront, so a track's layers are
// updated in increasing order exactly as in the one-lane-per-track kernels
// (same per-lane arithmetic, same chi2 summation order: bit-identical).
// Wavefronts are independent (no workgroup barrier); LDS is private per
// wavefront.
// 112 slots: 12.5 KB of state per wavefront without chi2, so 12 wavefronts fit a CU's LDS (128 slots fit
// 11; measured 1-2 % slower), while a window still holds 70 pairs at the uneven workload's density.
constexpr int kPsTile = 112;   // slots per wavefront tile (lane l holds slots l and l + 64, those below the tile)
constexpr int kPsPos = 128;    // position stride: position = layer * kPsPos + slot (slot in 7 bits)
constexpr int kPsWords = 29;   // per slot: 27 state floats, charge, chi2 sum (when returned)
constexpr int kPsWaves = 1;    // wavefronts per workgroup
static_assert(kPsTile > 64 && kPsTile <= kPsPos, "a tile holds 65..128 slots");
constexpr uint32_t kPsNone = 0xffffffffu;

// Orders this wavefront's LDS accesses across lanes (LDS operations of one
// wavefront complete in issue order; this keeps the compiler from reordering).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One step of the schedule from window start `c` (position = layer*kPsPos + slot, slot < kPsTile).
// Returns this lane's pair code (slot | layer << 7 | last-layer flag << 12 |
// lane-in-block << 13 | block - the tile's first block << 23) or kPsNone, and
// the next window start in `next` (wave-uniform).  m: mask of this lane's two
// slots (lane, lane + 64); sc: their constant code bits (slot, lane-in-block,
// relative block).
__device__ __forceinline__ uint32_t ps_step(uint32_t c, const uint32_t m[2], const uint32_t sc[2], int lane,
                                           uint32_t* tbl, uint32_t& next) {
  c = __builtin_amdgcn_readfirstlane(c);  // wave-uniform: keeps the window arithmetic scalar
  const uint32_t L0 = c >> 7, s0 = c & (kPsPos - 1);  // L0 < nlayer <= 32, s0 < kPsTile
  // Window order: slots s0..kPsTile-1 at layer L0, then slots 0..s0-1 at layer L0 + 1.  A slot is present
  // in the window if its mask has the layer the window gives it (lane-slots at or above the tile have
  // mask 0).
  bool first[2], pres[2];
  uint64_t P[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    first[r] = (uint32_t)lane + 64u * r >= s0;
    pres[r] = ((m[r] >> L0) >> (first[r] ? 0u : 1u)) & 1u;
    P[r] = __ballot(pres[r]);
  }
  const uint32_t c0 = __popcll(P[0]), total = c0 + __popcll(P[1]);
  const uint32_t before_s0 =  // present slots in [0, s0): the window's layer-(L0 + 1) part
      s0 < 64u ? __popcll(P[0] & ((1ull << s0) - 1ull)) : c0 + __popcll(P[1] & ((1ull << (s0 - 64u)) - 1ull));
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    // present slots in [0, slot), then the position in window order (mod 2^32: the second part wraps)
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(P[r] >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)P[r], r ? c0 : 0u));
    const uint32_t rank = below - before_s0 + (first[r] ? 0u : total);
    if (pres[r] && rank <= 64) {
      const uint32_t L = L0 + (first[r] ? 0u : 1u);  // a present layer, so L < nlayer <= 32
      const uint32_t last = (m[r] >> L) <= 1u;       // no present layer above L
      tbl[rank] = sc[r] | (L << 7) | (last << 12);
    }
  }
  wave_lds_sync();
  const uint32_t take = total < 64 ? total : 64u;
  const uint32_t code = (uint32_t)lane < take ? tbl[lane] : kPsNone;
  const uint32_t t64 = tbl[64];
  next = __builtin_amdgcn_readfirstlane(total > 64 ? (t64 & 0xfffu) : c + kPsPos);
  wave_lds_sync();
  return code;
}

template <int SEM, bool CHI2, bool EXACT>
__global__ __launch_bounds__(64 * kPsWaves) void fwd_ragged_stream_kernel(const float* __restrict__ trk,
                                                                           const float* __restrict__ hits,
                                                                           float* __restrict__ out,
                                                                           float* __restrict__ chi2,
                                                                           const uint32_t* __restrict__ mask,
                                                                           uint32_t nslot, uint32_t bs, int nlayer,
                                                                           float kfact, int opts) {
  constexpr int kW = CHI2 ? kPsWords : kPsWords - 1;  // the chi2 word only when it is returned
  const bool per_layer = opts & P2R_OPT_CHI2_PER_LAYER;
  __shared__ float st_all[kPsWaves][kW][kPsTile];
  __shared__ uint32_t tbl_all[kPsWaves][65];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;  // wv: wave-uniform
  const uint32_t base = (blockIdx.x * kPsWaves + wv) * (uint32_t)kPsTile;
  if (base >= nslot) return;  // whole wavefront; no workgroup barriers below
  float(*st)[kPsTile] = st_all[wv];
  uint32_t* tbl = tbl_all[wv];
  const uint32_t layer_bits = nlayer >= 32 ? ~0u : ((1u << nlayer) - 1u);
  const uint32_t blk0 = __builtin_amdgcn_readfirstlane(base / bs);  // the tile's blocks: blk0 .. blk0 + 128 at most
  uint32_t m[2], b[2], ln[2], sc[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const uint32_t slot = (uint32_t)lane + 64u * r, g = base + slot;
    m[r] = 0u;
    b[r] = g / bs;
    ln[r] = g - b[r] * bs;
    sc[r] = slot | (ln[r] << 13) | ((b[r] - blk0) << 23);
    if (slot < kPsTile && g < nslot) {
      const float* tb = trk + (size_t)b[r] * 28 * bs + ln[r];
#pragma unroll
      for (int f = 0; f < 28; ++f) st[f][slot] = tb[(size_t)f * bs];
      if (CHI2) st[kW - 1][slot] = 0.f;
      m[r] = mask[g] & layer_bits;
    }
  }
  wave_lds_sync();
  // Hits of the tile's blocks through a buffer resource over their rows: a 32-bit byte offset per
  // pair, the field stride as the scalar offset.  The tile spans at most 127 / bs + 2 blocks, under
  // 2.4 MB of hit rows for bs <= 1024 and nlayer <= 32; the record count is that span, so an offset
  // past it would read 0 rather than another tile's memory.
  const uint32_t row_bytes = 36u * bs, field_bytes = 4u * bs;
  const uint32_t blk_last = (base + kPsTile - 1 < nslot ? base + kPsTile - 1 : nslot - 1) / bs;
  const uint64_t hrow = reinterpret_cast<uint64_t>(hits + (size_t)blk0 * nlayer * 9 * bs);
  // readfirstlane returns int: each half goes back through uint32_t so the low word is not sign-extended
  const uint64_t hrow_u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(hrow >> 32)) << 32) |
                          (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)hrow);  // wave-uniform
  const __amdgpu_buffer_rsrc_t hrs = __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(hrow_u), (short)0,
      (int)__builtin_amdgcn_readfirstlane((blk_last - blk0 + 1) * (uint32_t)nlayer * row_bytes), 0x00020000);
  auto load_pair_hit = [&](uint32_t code, Hit& h) {
    const uint32_t off = ((code >> 23) * (uint32_t)nlayer + ((code >> 7) & 31u)) * row_bytes + ((code >> 11) & 0xffcu);
    auto ld = [&](int f) {
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                           hrs, off, __builtin_amdgcn_readfirstlane(f * field_bytes), 0));
    };
#pragma unroll
    for (int f = 0; f < 3; ++f) h.pos[f] = ld(f);
#pragma unroll
    for (int f = 0; f < 6; ++f) h.cov[f] = ld(3 + f);
  };
  auto process = [&](uint32_t code, const Hit& h) {
    if (code != kPsNone) {
      const uint32_t slot = code & (kPsPos - 1);
      Trk src;
#pragma unroll
      for (int f = 0; f < 6; ++f) src.par[f] = st[f][slot];
#pragma unroll
      for (int f = 0; f < 21; ++f) src.cov[f] = st[6 + f][slot];
      int q = __float_as_int(st[27][slot]);
      float k = (float)q * kfact;
      if (SEM == P2R_REFERENCE) opaque(src);
      Trk prop;
      Lane<EXACT>::propagate(src, k, h.pos[0], h.pos[1], prop);
      float cc = Lane<EXACT>::update(prop, h);
      if (opts) apply_options<EXACT>(opts, prop, q, k, kfact);
      if (SEM == P2R_REFERENCE) {
        opaque(prop);
        opaque(cc);
      }
      if (CHI2) {
        if (per_layer) chi2[(size_t)((code >> 7) & 31u) * nslot + base + slot] = cc;
        else st[kW - 1][slot] += cc;
      }
      if (SEM == P2R_CHAINED) {
#pragma unroll
        for (int f = 0; f < 6; ++f) st[f][slot] = prop.par[f];
#pragma unroll
        for (int f = 0; f < 21; ++f) st[6 + f][slot] = prop.cov[f];
        st[27][slot] = __int_as_float(q);
      } else if ((code >> 12) & 1u) {  // reference semantics: the last present layer's result is the output
        float* ob = out + (size_t)(blk0 + (code >> 23)) * 28 * bs + ((code >> 13) & 1023u);
        store_trk(ob, bs, prop);
        reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
      }
    }
    wave_lds_sync();
  };
  // Steps alternate between two hit registers sets: the next step's pairs and hit loads are issued
  // before the current step's arithmetic, with no copies between them.
  const uint32_t cend = (uint32_t)nlayer * kPsPos;
  uint32_t cn = 0;
  uint32_t code_a = ps_step(0u, m, sc, lane, tbl, cn), code_b = kPsNone;
  Hit ha, hb;
  if (code_a != kPsNone) load_pair_hit(code_a, ha);
  for (;;) {
    const bool more_b = cn < cend;
    if (more_b) {
      code_b = ps_step(cn, m, sc, lane, tbl, cn);
      if (code_b != kPsNone) load_pair_hit(code_b, hb);
    }
    process(code_a, ha);
    if (!more_b) break;
    const bool more_a = cn < cend;
    if (more_a) {
      code_a = ps_step(cn, m, sc, lane, tbl, cn);
      if (code_a != kPsNone) load_pair_hit(code_a, ha);
    }
    process(code_b, hb);
    if (!more_a) break;
  }
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const uint32_t slot = (uint32_t)lane + 64u * r, g = base + slot;
    if (slot >= kPsTile || g >= nslot) continue;  // another tile's slot, or past the batch
    if (SEM == P2R_CHAINED || m[r] == 0u) {  // chained result, or a track with no present layer copied through
      float* ob = out + (size_t)b[r] * 28 * bs + ln[r];
#pragma unroll
      for (int f = 0; f < 28; ++f) ob[(size_t)f * bs] = st[f][slot];
    }
    if (CHI2) {
      if (!per_layer) chi2[g] = st[kW - 1][slot];
      else
        for (int l = 0; l < nlayer; ++l)
          if (!((m[r] >> l) & 1u)) chi2[(size_t)l * nslot + g] = 0.f;
    }
  }
}

// ---------------------------------------------------------------------------
// Forward + adjoint of the chained fit.  One lane per track.  Forward pass in
// the state precision RS, checkpoints every CS layers (checkpoint 0 is the
// input track, not stored); reverse sweep per segment: the segment's layer
// inputs are recomputed from its checkpoint (into LDS or a global buffer, see
// SegStore), then each layer's values are recomputed and back-propagated in
// the layer precision.
constexpr int kAdjThreads = 256;

template <class R, bool EXACT> struct MathOf;
// fwd: one layer of the forward pass (the updated state and its chi2): p2r_forward's own per-lane code
// (propagate_to_r, kalman_update) in the instance's arithmetic, shared by the adjoint's forward pass, the
// reverse sweep's segment recompute and the FP64 tail.
template <> struct MathOf<float, true> {
  using T = exact::Trk; using H = exact::Hit;
  static __device__ __forceinline__ float fwd(const T& in, float k, const H& h, T& out) {
    exact::propagate_to_r(in, k, h.pos[0], h.pos[1], out);
    return exact::kalman_update(out, h);
  }
  static __device__ __forceinline__ void adj(const T& in, float k, const H& h, T& bar, float& kb, float hb[3]) {
    exact::layer_adjoint(in, k, h, bar, kb, hb);
  }
};
template <> struct MathOf<float, false> {
  using T = fast::Trk; using H = fast::Hit;
  static __device__ __forceinline__ float fwd(const T& in, float k, const H& h, T& out) {
    fast::propagate_to_r(in, k, h.pos[0], h.pos[1], out);
    return fast::kalman_update(out, h);
  }
  static __device__ __forceinline__ void adj(const T& in, float k, const H& h, T& bar, float& kb, float hb[3]) {
    fast::layer_adjoint(in, k, h, bar, kb, hb);
  }
};
template <bool EXACT> struct MathOf<double, EXACT> {
  using T = dbl::Trk; using H = dbl::Hit;
  static __device__ __forceinline__ double fwd(const T& in, double k, const H& h, T& out) {
    dbl::propagate_to_r(in, k, h.pos[0], h.pos[1], out);
    return dbl::kalman_update(out, h);
  }
  static __device__ __forceinline__ void adj(const T& in, double k, const H& h, T& bar, double& kb, double hb[3]) {
    dbl::layer_adjoint(in, k, h, bar, kb, hb);
  }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One wavefront's contributions of layer l to the global sums -- rigid translation dx, dy, dz and
// rotation dphi of the layer's hits, from the per-hit gradient hbar -- written by lane 0 into the
// wavefront's row of `lds` ([wave][4*nlayer + 1]).  Tracks out of the sums contribute 0.
template <class R>
__device__ __forceinline__ void layer_global_sums(double* lds_row, int l, bool in_global, const R hbar[3], double hx,
                                                  double hy, int wl) {
  const double a = in_global ? (double)hbar[0] : 0.0, b = in_global ? (double)hbar[1] : 0.0,
               cz = in_global ? (double)hbar[2] : 0.0;
  const double v0 = wave_sum(a), v1 = wave_sum(b), v2 = wave_sum(cz), v3 = wave_sum(-hy * a + hx * b);
  if (wl == 0) {
    lds_row[4 * l + 0] = v0;
    lds_row[4 * l + 1] = v1;
    lds_row[4 * l + 2] = v2;
    lds_row[4 * l + 3] = v3;
  }
}
// The workgroup's sums of global entries [j0, j1) over its wavefronts' rows, into partials[j][workgroup].
__device__ __forceinline__ void write_partials(const double* lds, int nglob, int j0, int j1, double* partials) {
  __syncthreads();
  for (int j = j0 + (int)threadIdx.x; j < j1; j += (int)blockDim.x) {
    double acc = 0.0;
    for (int w = 0; w < (int)blockDim.x / 64; ++w) acc += lds[w * nglob + j];
    partials[(size_t)j * gridDim.x + blockIdx.x] = acc;  // [global entry][workgroup]
  }
}

template <class R> __device__ __forceinline__ void load_hit_as(const float* hb, uint32_t bs, HitT<R>& h) {
  Hit hf;
  load_hit(hb, bs, hf);
  convert(hf, h);
}

// Forward pass of the adjoint: the chained fit in the state precision RS, storing the state every CS
// layers (checkpoint 0 is the input track, not stored) and each track's total chi2 (for the reverse
// kernel), and writing the fitted track and chi2 as p2r_forward does.  Its own kernel so it runs at the
// forward's occupancy (MINW waves per SIMD), not at the reverse sweep's register budget.
template <class RS, bool EXACT, int CS, int MINW>
__global__ __launch_bounds__(kAdjThreads, MINW) void adjoint_fwd_kernel(
    const float* __restrict__ trk, const float* __restrict__ hits, float* __restrict__ out, float* __restrict__ chi2,
    RS* __restrict__ ckpt, RS* __restrict__ c2_total, uint32_t ntrk, uint32_t bs, int nlayer, float kf) {
  using MS = MathOf<RS, EXACT>;
  const uint32_t g = blockIdx.x * kAdjThreads + threadIdx.x;
  if (g >= ntrk) return;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  const int q = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  const float k = (float)q * kf;
  Trk in_f;
  load_trk(tb, bs, in_f);
  TrkT<RS> cur;
  convert(in_f, cur);
  RS c2 = 0;
  for (int l = 0; l < nlayer; ++l) {
    if (l % CS == 0 && l > 0) {
      RS* cp = ckpt + ((size_t)(l / CS - 1) * 27) * ntrk + g;
#pragma unroll
      for (int f = 0; f < 6; ++f) cp[(size_t)f * ntrk] = cur.par[f];
#pragma unroll
      for (int f = 0; f < 21; ++f) cp[(size_t)(6 + f) * ntrk] = cur.cov[f];
    }
    HitT<RS> h;  // loaded as the layer starts (no prefetch: 127 instead of 153 VGPRs in FP32, 4 waves)
    load_hit_as(hb + l * hstride, bs, h);
    TrkT<RS> nxt;
    c2 += MS::fwd(cur, (RS)k, h, nxt);
    cur = nxt;
  }
  float* ob = out + (size_t)blk * 28 * bs + lane;
  Trk of;
  convert(cur, of);
  store_trk(ob, bs, of);
  reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
  if (chi2) chi2[g] = (float)c2;
  c2_total[g] = c2;
}

// Where the reverse sweep keeps the recomputed layer inputs of a segment between its forward recompute and
// their reverse: the inputs of layers l0+1 .. l0+len-2 (CS-2 states at most; the last one stays in
// registers, the first is the checkpoint, read again), in the precision RL of the per-layer reverse, which
// is all the reverse reads of them.  In LDS when RW workgroups of them fit beside the global-sum partials,
// else in a global buffer [slot][field][track] of the workspace.
// Register budget of the reverse sweep: 2 waves per SIMD with the FP32 reverse, 1 with the FP64 one.
template <class RL> constexpr int kRevWaves = sizeof(RL) == sizeof(float) ? 2 : 1;
constexpr size_t kAdjPartialsLdsMax = (size_t)(kAdjThreads / 64) * (4 * P2R_MAX_NLAYER + 1) * sizeof(double);
template <class RL, int CS, int RW> struct SegStore {
  static constexpr int kSlots = CS > 2 ? CS - 2 : 0;
  static constexpr size_t kLdsBytes = (size_t)kSlots * 27 * sizeof(RL) * kAdjThreads;
  // within 64 KB per workgroup, and RW workgroups (RW waves per SIMD) within the CU's 160 KB
  static constexpr bool kInLds = kSlots == 0 || (kLdsBytes + kAdjPartialsLdsMax <= 64 * 1024 &&
                                                 (size_t)RW * (kLdsBytes + kAdjPartialsLdsMax) <= 160 * 1024);
  static size_t global_bytes(uint64_t ntrk) { return kInLds ? 0 : (size_t)kSlots * 27 * sizeof(RL) * ntrk; }
};

// Reverse sweep of the adjoint, after adjoint_fwd_kernel: per segment, from the last, reload the checkpoint,
// recompute the segment's layer inputs, then per layer from the last recompute its values and back-propagate.
// The layer loops are not unrolled (one copy of each layer body in the code).  RW is the register budget in
// waves per SIMD.
template <class RS, class RL, bool EXACT, int CS, int RW = kRevWaves<RL>>
__global__ __launch_bounds__(kAdjThreads, RW) void adjoint_kernel(
    const float* __restrict__ trk, const float* __restrict__ hits, const RS* __restrict__ c2_total,
    float* __restrict__ gpar, float* __restrict__ ghit, float* __restrict__ gk, double* __restrict__ partials,
    const RS* __restrict__ ckpt, RL* __restrict__ seg_global, uint32_t ntrk, uint32_t bs, int nlayer, float kf,
    int l_stop, float* __restrict__ bar_out) {
  using MS = MathOf<RS, EXACT>;
  // The segment recomputation runs in the requested arithmetic (the layer inputs of the forward pass);
  // the reverse sweep always in the fused instance: gradients have no reference rounding to reproduce,
  // and they are checked against FP64 to a tolerance.
  using ML = MathOf<RL, false>;
  using SS = SegStore<RL, CS, RW>;
  constexpr int kWaves = kAdjThreads / 64;
  // dynamic LDS: [kWaves][4*nlayer + 1] doubles when partials != nullptr, then the segment states when in LDS
  extern __shared__ double lds_global[];
  const uint32_t g = blockIdx.x * kAdjThreads + threadIdx.x;
  const bool active = g < ntrk;
  const uint32_t gg = active ? g : ntrk - 1;  // inactive lanes recompute a valid track, write nothing
  const uint32_t blk = gg / bs, lane = gg - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  const int q = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  const float k = (float)q * kf;
  const int nseg = (nlayer + CS - 1) / CS;
  const RS c2 = c2_total[gg];
  const int nglob = 4 * nlayer + 1;
  RL* seg_lds = reinterpret_cast<RL*>(lds_global + (partials ? kWaves * nglob : 0));
  // slot e of the current lane: LDS [slot][field][thread] or global [slot][field][track]
  auto seg_at = [&](int slot, int f) -> RL& {
    if constexpr (SS::kInLds) return seg_lds[((size_t)slot * 27 + f) * kAdjThreads + threadIdx.x];
    else return seg_global[((size_t)slot * 27 + f) * ntrk + gg];
  };
  // Lanes past the batch share the last track's global slots: only the track's own lane writes them (the
  // others read what it wrote, or stale words; their results are not stored).
  const bool seg_writer = SS::kInLds || active;
  auto seg_put = [&](int slot, const TrkT<RS>& t) {
    if (!seg_writer) return;
#pragma unroll
    for (int f = 0; f < 6; ++f) seg_at(slot, f) = (RL)t.par[f];
#pragma unroll
    for (int f = 0; f < 21; ++f) seg_at(slot, 6 + f) = (RL)t.cov[f];
  };
  auto seg_get = [&](int slot, TrkT<RL>& t) {
#pragma unroll
    for (int f = 0; f < 6; ++f) t.par[f] = seg_at(slot, f);
#pragma unroll
    for (int f = 0; f < 21; ++f) t.cov[f] = seg_at(slot, 6 + f);
  };
  // input of the segment's first layer: the track itself or a checkpoint.  Read again for the segment's first
  // layer: the addresses pass through an opaque copy so that the compiler reloads instead of keeping the
  // values of the first read live through the whole segment.
  auto seg_start = [&](int s, auto& t) {  // t: TrkT<RS> or TrkT<RL>
    if (s == 0) {
      const float* tp = tb;
      asm volatile("" : "+v"(tp));
      Trk in_f;
      load_trk(tp, bs, in_f);
      convert(in_f, t);
    } else {
      const RS* cp = ckpt + ((size_t)(s - 1) * 27) * ntrk + gg;
      asm volatile("" : "+v"(cp));
      TrkT<RS> c;
#pragma unroll
      for (int f = 0; f < 6; ++f) c.par[f] = cp[(size_t)f * ntrk];
#pragma unroll
      for (int f = 0; f < 21; ++f) c.cov[f] = cp[(size_t)(6 + f) * ntrk];
      convert(c, t);
    }
  };
  // ---- reverse sweep ----
  // a track whose fit diverged (non-finite total chi2) adds nothing to the global sums
  const bool in_global = active && isfinite((double)c2);
  const int wave = threadIdx.x >> 6, wl = threadIdx.x & 63;
  TrkT<RL> bar;
#pragma unroll
  for (int f = 0; f < 6; ++f) bar.par[f] = 0;
#pragma unroll
  for (int f = 0; f < 21; ++f) bar.cov[f] = 0;
  RL kbar = 0;
  for (int s = nseg - 1; s >= 0; --s) {
    const int l0 = s * CS, len = min(CS, nlayer - l0);
    if (l0 + len <= l_stop) break;  // the layers below l_stop are left to adjoint_tail_kernel
    TrkT<RS> cur;
    seg_start(s, cur);
    // inputs of layers l0+1 .. l0+len-1: the last stays in `cur`, the others go to the segment store
#pragma nounroll
    for (int j = 1; j < len; ++j) {
      if (j >= 2) seg_put(j - 2, cur);
      HitT<RS> h;
      load_hit_as(hb + (l0 + j - 1) * hstride, bs, h);
      TrkT<RS> nxt;
      MS::fwd(cur, (RS)k, h, nxt);
      cur = nxt;
    }
    // reverse of the segment's layers from the last; each iteration ends by reading the next one's input
    // (the segment store, then the checkpoint), so that no input is held through another layer's reverse
    TrkT<RL> in_l;
    convert(cur, in_l);
#pragma nounroll
    for (int j = len - 1; j >= 0 && l0 + j >= l_stop; --j) {
      const int l = l0 + j;
      HitT<RL> h;
      load_hit_as(hb + l * hstride, bs, h);
      RL hbar[3];
      ML::adj(in_l, (RL)k, h, bar, kbar, hbar);
      if (active && ghit) {
        float* gh = ghit + ((size_t)blk * nlayer + l) * 3 * bs + lane;
#pragma unroll
        for (int c = 0; c < 3; ++c) gh[(size_t)c * bs] = (float)hbar[c];
      }
      if (partials) layer_global_sums(lds_global + wave * nglob, l, in_global, hbar, h.pos[0], h.pos[1], wl);
      if (j == 1) seg_start(s, in_l);
      else if (j > 1) seg_get(j - 2, in_l);
    }
  }
  if (active) {
    if (l_stop > 0) {
      // handed to adjoint_tail_kernel: the adjoint of layer l_stop's input state, and d/dk so far
      float* bo = bar_out + g;
#pragma unroll
      for (int f = 0; f < 6; ++f) bo[(size_t)f * ntrk] = (float)bar.par[f];
#pragma unroll
      for (int f = 0; f < 21; ++f) bo[(size_t)(6 + f) * ntrk] = (float)bar.cov[f];
      bo[(size_t)27 * ntrk] = (float)kbar;
    } else {
      if (gpar) {
        float* gp = gpar + (size_t)blk * 6 * bs + lane;
#pragma unroll
        for (int f = 0; f < 6; ++f) gp[(size_t)f * bs] = (float)bar.par[f];
      }
      if (gk) gk[g] = (float)((RL)q * kbar);  // d/dkfact = q d/dk
    }
  }
  if (partials) {
    if (l_stop == 0) {
      // b_scale: kfact_eff = kfact / b  =>  d/db = -kfact_eff * d/dkfact_eff at b = 1 scaled by 1/b; the
      // host passes kf = kfact/b, so d/db = -(kf / b) q kbar; written here for b folded into kf (b = 1 form)
      const double vb = wave_sum(in_global ? -(double)kf * (double)q * (double)kbar : 0.0);
      if (wl == 0) lds_global[wave * nglob + 4 * nlayer] = vb;
    }
    // this kernel's entries: its layers, and the field scale when no tail follows
    write_partials(lds_global, nglob, 4 * l_stop, l_stop > 0 ? 4 * nlayer : nglob, partials);
  }
}

// The first layers of a MIXED adjoint in FP64.  With the state in FP64 the per-layer FP32 reverse loses
// accuracy mostly at the first layers, where the fit is least constrained and the covariance update
// cancels most: adjoint_kernel stops at layer l_stop (<= kTailLayers) and hands over the adjoint of that
// layer's input; this kernel recomputes the inputs of layers 0 .. l_stop-1 in the requested arithmetic
// (FP64 states, as in the forward pass), back-propagates them in FP64 and writes the per-track outputs,
// those layers' global-sum partials and the field scale's.  Its own kernel so that the FP64 reverse's
// registers (1 wavefront per SIMD) do not set the occupancy of the FP32 sweep.
constexpr int kTailLayers = 2;
template <bool EXACT>
__global__ __launch_bounds__(kAdjThreads) void adjoint_tail_kernel(
    const float* __restrict__ trk, const float* __restrict__ hits, const double* __restrict__ c2_total,
    const float* __restrict__ bar_in, float* __restrict__ gpar, float* __restrict__ ghit, float* __restrict__ gk,
    double* __restrict__ partials, uint32_t ntrk, uint32_t bs, int nlayer, float kf, int l_stop) {
  using MS = MathOf<double, EXACT>;
  using ML = MathOf<double, false>;
  extern __shared__ double lds_global[];  // [kWaves][4*nlayer + 1] when partials != nullptr
  const uint32_t g = blockIdx.x * kAdjThreads + threadIdx.x;
  const bool active = g < ntrk;
  const uint32_t gg = active ? g : ntrk - 1;  // lanes past the batch recompute the last track, write nothing
  const uint32_t blk = gg / bs, lane = gg - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  const int q = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  const float k = (float)q * kf;
  const bool in_global = active && isfinite(c2_total[gg]);
  const int nglob = 4 * nlayer + 1;
  const int wave = threadIdx.x >> 6, wl = threadIdx.x & 63;
  // inputs of layers 0 .. l_stop-1: the track, then the fit in FP64
  TrkT<double> st[kTailLayers];
  {
    Trk in_f;
    load_trk(tb, bs, in_f);
    convert(in_f, st[0]);
  }
#pragma unroll
  for (int j = 1; j < kTailLayers; ++j) {
    if (j < l_stop) {
      HitT<double> h;
      load_hit_as(hb + (j - 1) * hstride, bs, h);
      MS::fwd(st[j - 1], (double)k, h, st[j]);
    }
  }
  TrkT<double> bar;
  double kbar;
  {
    const float* bi = bar_in + gg;
#pragma unroll
    for (int f = 0; f < 6; ++f) bar.par[f] = bi[(size_t)f * ntrk];
#pragma unroll
    for (int f = 0; f < 21; ++f) bar.cov[f] = bi[(size_t)(6 + f) * ntrk];
    kbar = bi[(size_t)27 * ntrk];
  }
#pragma unroll
  for (int l = kTailLayers - 1; l >= 0; --l) {
    if (l >= l_stop) continue;
    HitT<double> h;
    load_hit_as(hb + l * hstride, bs, h);
    double hbar[3];
    ML::adj(st[l], (double)k, h, bar, kbar, hbar);
    if (active && ghit) {
      float* gh = ghit + ((size_t)blk * nlayer + l) * 3 * bs + lane;
#pragma unroll
      for (int c = 0; c < 3; ++c) gh[(size_t)c * bs] = (float)hbar[c];
    }
    if (partials) layer_global_sums(lds_global + wave * nglob, l, in_global, hbar, h.pos[0], h.pos[1], wl);
  }
  if (active) {
    if (gpar) {
      float* gp = gpar + (size_t)blk * 6 * bs + lane;
#pragma unroll
      for (int f = 0; f < 6; ++f) gp[(size_t)f * bs] = (float)bar.par[f];
    }
    if (gk) gk[g] = (float)((double)q * kbar);  // d/dkfact = q d/dk
  }
  if (partials) {
    // the field scale, as in adjoint_kernel
    const double vb = wave_sum(in_global ? -(double)kf * (double)q * kbar : 0.0);
    if (wl == 0) lds_global[wave * nglob + 4 * nlayer] = vb;
    write_partials(lds_global, nglob, 0, 4 * l_stop, partials);
    if (threadIdx.x == 0) {  // the field-scale entry, summed over the wavefronts like the others
      double acc = 0.0;
      for (int w = 0; w < kAdjThreads / 64; ++w) acc += lds_global[w * nglob + 4 * nlayer];
      partials[(size_t)(4 * nlayer) * gridDim.x + blockIdx.x] = acc;
    }
  }
}

// grad_global[j] = sum over workgroups of partials[j][b]: one workgroup per entry j; thread t sums
// b = t, t + 256, ... in order (coalesced rows), then a fixed pairwise tree in LDS.  The order depends
// only on the number of workgroups, so the result is bit-reproducible.
constexpr int kFinThreads = 256;
__global__ __launch_bounds__(kFinThreads) void finalize_global_kernel(const double* __restrict__ partials,
                                                                      double* __restrict__ out, int nblocks,
                                                                      int nglob, double bscale_factor) {
  __shared__ double red[kFinThreads];
  const int j = blockIdx.x, t = threadIdx.x;
  const double* row = partials + (size_t)j * nblocks;
  double acc = 0.0;
  for (int b = t; b < nblocks; b += kFinThreads) acc += row[b];
  red[t] = acc;
  __syncthreads();
  for (int h = kFinThreads / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) out[j] = (j == nglob - 1) ? red[0] * bscale_factor : red[0];
}

// Block-level propagateToR over nblk blocks with caller strides.
__global__ __launch_bounds__(kThreads) void propagate_blocks_kernel(const float* __restrict__ inErr,
                                                                    const float* __restrict__ inPar,
                                                                    const int* __restrict__ inChg,
                                                                    const float* __restrict__ msP,
                                                                    float* __restrict__ outErr,
                                                                    float* __restrict__ outPar, uint32_t ntrk,
                                                                    uint32_t bs, int64_t tstride, int64_t hstride,
                                                                    float kfact) {
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= ntrk) return;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const size_t to = (size_t)blk * tstride + lane, ho = (size_t)blk * hstride + lane;
  Trk in;
#pragma unroll
  for (int f = 0; f < 6; ++f) in.par[f] = inPar[to + (size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 21; ++f) in.cov[f] = inErr[to + (size_t)f * bs];
  const float k = (float)inChg[to] * kfact;
  Trk o;
  Lane<true>::propagate(in, k, msP[ho], msP[ho + bs], o);
#pragma unroll
  for (int f = 0; f < 6; ++f) outPar[to + (size_t)f * bs] = o.par[f];
#pragma unroll
  for (int f = 0; f < 21; ++f) outErr[to + (size_t)f * bs] = o.cov[f];
}

// Block-level KalmanUpdate (in place) over nblk blocks with caller strides.
__global__ __launch_bounds__(kThreads) void kalman_blocks_kernel(float* __restrict__ trkErr,
                                                                 float* __restrict__ inPar,
                                                                 const float* __restrict__ hitErr,
                                                                 const float* __restrict__ msP, uint32_t ntrk,
                                                                 uint32_t bs, int64_t tstride, int64_t hstride) {
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= ntrk) return;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const size_t to = (size_t)blk * tstride + lane, ho = (size_t)blk * hstride + lane;
  Trk t;
  Hit h;
#pragma unroll
  for (int f = 0; f < 6; ++f) t.par[f] = inPar[to + (size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 21; ++f) t.cov[f] = trkErr[to + (size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 3; ++f) h.pos[f] = msP[ho + (size_t)f * bs];
#pragma unroll
  for (int f = 0; f < 6; ++f) h.cov[f] = hitErr[ho + (size_t)f * bs];
  Lane<true>::update(t, h);
#pragma unroll
  for (int f = 0; f < 6; ++f) inPar[to + (size_t)f * bs] = t.par[f];
#pragma unroll
  for (int f = 0; f < 21; ++f) trkErr[to + (size_t)f * bs] = t.cov[f];
}

inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline unsigned grid_for_n(uint64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

struct UniformLaunch {
  const float* T;
  const float* H;
  float* O;
  float* chi2;
  uint32_t n, bs;
  int nlayer;
  float kf;
  int opts;
};
template <int SEM, bool CHI2, bool EXACT, int MINW>
void launch_uniform_w(dim3 grid, dim3 block, hipStream_t st, const UniformLaunch& a, const float* in, int l0, int l1,
                      bool acc) {
  // Reference semantics: only the launch that runs the last layer writes the track (an earlier launch's
  // result would be overwritten anyway), so trk_in stays intact for every launch even when trk_out
  // aliases it.
  const bool store = SEM == P2R_CHAINED || l1 == a.nlayer;
  fwd_uniform_kernel<SEM, CHI2, EXACT, MINW><<<grid, block, SEM == P2R_REFERENCE ? 21 * block.x * sizeof(float) : 0,
                                               st>>>(
      in, a.H, a.O, a.chi2, a.n, a.bs, a.nlayer, a.kf, l0, l1, acc, store, a.opts);
}
template <int SEM, bool CHI2, bool EXACT>
void launch_uniform_e(int waves, dim3 grid, dim3 block, hipStream_t st, const UniformLaunch& a, const float* in,
                      int l0, int l1, bool acc) {
  // 0 = the library's choice: the 3-wave budget for the exact instance, 4 waves for the fast one.  The
  // current kernels fit 5 or more waves per SIMD at the compiler's own allocation, and the three budgets
  // compile to the same instructions; the choice stays for the autotuner.
  if (waves == 0) waves = EXACT ? 3 : 4;
  if (waves == 3) launch_uniform_w<SEM, CHI2, EXACT, 3>(grid, block, st, a, in, l0, l1, acc);
  else if (waves == 4) launch_uniform_w<SEM, CHI2, EXACT, 4>(grid, block, st, a, in, l0, l1, acc);
  else launch_uniform_w<SEM, CHI2, EXACT, 1>(grid, block, st, a, in, l0, l1, acc);
}
template <int SEM, bool CHI2>
void launch_uniform_p(bool exact, int waves, dim3 grid, dim3 block, hipStream_t st, const UniformLaunch& a,
                      const float* in, int l0, int l1, bool acc) {
  if (exact) launch_uniform_e<SEM, CHI2, true>(waves, grid, block, st, a, in, l0, l1, acc);
  else launch_uniform_e<SEM, CHI2, false>(waves, grid, block, st, a, in, l0, l1, acc);
}
// The whole forward: one launch over all layers, or launches of `per_launch`
// layers (the state passes through trk_out between launches; in reference
// semantics every launch restarts from trk_in, as every layer does, and only
// the last launch writes trk_out).
void launch_uniform(int sem, bool exact, int threads, int per_launch, int waves, hipStream_t st,
                    const UniformLaunch& a) {
  const dim3 grid(grid_for_n(a.n, threads)), block(threads);
  const int step = (per_launch > 0 && per_launch < a.nlayer) ? per_launch : a.nlayer;
  for (int l0 = 0; l0 < a.nlayer; l0 += step) {
    const int l1 = std::min(a.nlayer, l0 + step);
    const float* in = (sem == P2R_CHAINED && l0 > 0) ? a.O : a.T;
    const bool acc = l0 > 0;
    if (sem == P2R_REFERENCE) {
      if (a.chi2) launch_uniform_p<P2R_REFERENCE, true>(exact, waves, grid, block, st, a, in, l0, l1, acc);
      else launch_uniform_p<P2R_REFERENCE, false>(exact, waves, grid, block, st, a, in, l0, l1, acc);
    } else {
      if (a.chi2) launch_uniform_p<P2R_CHAINED, true>(exact, waves, grid, block, st, a, in, l0, l1, acc);
      else launch_uniform_p<P2R_CHAINED, false>(exact, waves, grid, block, st, a, in, l0, l1, acc);
    }
  }
}


template <int SEM, bool CHI2, bool EXACT, int ROUNDS>
int launch_ragged_t(p2r_ctx* c, hipStream_t st, const float* T, const float* H, float* O, float* chi2,
                    const uint32_t* M, uint32_t nslot, uint32_t bs, int nlayer, float kf, int opts) {
  constexpr int kTile = kRagThreads * ROUNDS;
  const uint32_t ntiles = (nslot + kTile - 1) / kTile;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fwd_ragged_kernel<SEM, CHI2, EXACT, ROUNDS>,
                                                   kRagThreads, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)(c->num_cus * per_cu));
  fwd_ragged_kernel<SEM, CHI2, EXACT, ROUNDS><<<grid, kRagThreads, 0, st>>>(T, H, O, chi2, M, nslot, bs, nlayer, kf,
                                                                           ntiles, opts);
  return P2R_OK;
}
template <int SEM, bool CHI2, bool EXACT>
int launch_ragged_r(int rounds, p2r_ctx* c, hipStream_t st, const float* T, const float* H, float* O, float* chi2,
                    const uint32_t* M, uint32_t nslot, uint32_t bs, int nlayer, float kf, int opts) {
  if (rounds == P2R_RAGGED_PAIR_STREAM) {
    const uint32_t ntiles = (nslot + kPsTile - 1) / kPsTile;
    fwd_ragged_stream_kernel<SEM, CHI2, EXACT><<<(ntiles + kPsWaves - 1) / kPsWaves, 64 * kPsWaves, 0, st>>>(
        T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
    return P2R_OK;
  }
  if (rounds == P2R_RAGGED_LAYER_ORDER) {
    const int threads = c->tune.ragged_threads;
    fwd_ragged_layer_kernel<SEM, CHI2, EXACT><<<grid_for_n(nslot, threads), threads, 0, st>>>(T, H, O, chi2, M, nslot,
                                                                                            bs, nlayer, kf, opts);
    return P2R_OK;
  }
  switch (rounds) {
    case 1: return launch_ragged_t<SEM, CHI2, EXACT, 1>(c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
    case 2: return launch_ragged_t<SEM, CHI2, EXACT, 2>(c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
    default: return launch_ragged_t<SEM, CHI2, EXACT, 4>(c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
  }
}
template <int SEM>
int launch_ragged_s(bool chi2_on, bool exact, int rounds, p2r_ctx* c, hipStream_t st, const float* T, const float* H,
                    float* O, float* chi2, const uint32_t* M, uint32_t nslot, uint32_t bs, int nlayer, float kf, int opts) {
  if (chi2_on) {
    if (exact) return launch_ragged_r<SEM, true, true>(rounds, c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
    return launch_ragged_r<SEM, true, false>(rounds, c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
  }
  if (exact) return launch_ragged_r<SEM, false, true>(rounds, c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
  return launch_ragged_r<SEM, false, false>(rounds, c, st, T, H, O, chi2, M, nslot, bs, nlayer, kf, opts);
}

template <class RS, class RL, bool EXACT, int CS>
void launch_adjoint_t(unsigned grid, size_t lds, hipStream_t st, const float* T, const float* H, float* O, float* chi2,
                      float* gpar, float* ghit, float* gk, double* partials, void* ckpt, void* c2_total, void* seg,
                      float* bar, uint32_t ntrk, uint32_t bs, int nlayer, float kf) {
  // the forward pass at the forward kernels' register budget (3 waves per SIMD in FP32)
  constexpr int kFwdWaves = sizeof(RS) == sizeof(float) ? 3 : 1;
  adjoint_fwd_kernel<RS, EXACT, CS, kFwdWaves><<<grid, kAdjThreads, 0, st>>>(
      T, H, O, chi2, static_cast<RS*>(ckpt), static_cast<RS*>(c2_total), ntrk, bs, nlayer, kf);
  using SS = SegStore<RL, CS, kRevWaves<RL>>;
  const size_t lds_all = lds + (SS::kInLds ? SS::kLdsBytes : 0);
  // MIXED (FP64 state, FP32 reverse): the first layers' reverse in FP64, in adjoint_tail_kernel
  constexpr bool kTail = sizeof(RS) == sizeof(double) && sizeof(RL) == sizeof(float);
  const int l_stop = kTail ? std::min(kTailLayers, nlayer) : 0;
  adjoint_kernel<RS, RL, EXACT, CS><<<grid, kAdjThreads, lds_all, st>>>(
      T, H, static_cast<const RS*>(c2_total), gpar, ghit, gk, partials, static_cast<const RS*>(ckpt),
      static_cast<RL*>(seg), ntrk, bs, nlayer, kf, l_stop, bar);
  if (kTail)
    adjoint_tail_kernel<EXACT><<<grid, kAdjThreads, lds, st>>>(T, H, static_cast<const double*>(c2_total), bar, gpar,
                                                               ghit, gk, partials, ntrk, bs, nlayer, kf, l_stop);
}
template <bool EXACT, int CS>
void launch_adjoint_p(int gp, unsigned grid, size_t lds, hipStream_t st, const float* T, const float* H, float* O,
                      float* chi2, float* gpar, float* ghit, float* gk, double* partials, void* ckpt, void* c2_total,
                      void* seg, float* bar, uint32_t ntrk, uint32_t bs, int nlayer, float kf) {
  if (gp == P2R_GRAD_F32)
    launch_adjoint_t<float, float, EXACT, CS>(grid, lds, st, T, H, O, chi2, gpar, ghit, gk, partials, ckpt, c2_total,
                                              seg, bar, ntrk, bs, nlayer, kf);
  else if (gp == P2R_GRAD_MIXED)
    launch_adjoint_t<double, float, EXACT, CS>(grid, lds, st, T, H, O, chi2, gpar, ghit, gk, partials, ckpt,
                                               c2_total, seg, bar, ntrk, bs, nlayer, kf);
  else
    launch_adjoint_t<double, double, EXACT, CS>(grid, lds, st, T, H, O, chi2, gpar, ghit, gk, partials, ckpt,
                                                c2_total, seg, bar, ntrk, bs, nlayer, kf);
}

// Bytes of the global segment store of one adjoint call (0 when it is in LDS).
template <class RS, class RL> size_t seg_bytes_rs(int cs, uint64_t ntrk) {
  if (cs == 4) return SegStore<RL, 4, kRevWaves<RL>>::global_bytes(ntrk);
  if (cs == 8) return SegStore<RL, 8, kRevWaves<RL>>::global_bytes(ntrk);
  return 0;
}
size_t seg_global_bytes(int gp, int cs, uint64_t ntrk) {
  if (gp == P2R_GRAD_F32) return seg_bytes_rs<float, float>(cs, ntrk);
  if (gp == P2R_GRAD_MIXED) return seg_bytes_rs<double, float>(cs, ntrk);
  return seg_bytes_rs<double, double>(cs, ntrk);
}

inline size_t ck_bytes_aligned(size_t b) { return (b + 255) & ~(size_t)255; }

// Address of `scratch` on each device, resolved once per device with that device current.
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxScratchDevices] = {};

int resolve_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return P2R_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    DeviceGuard g(device);
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return P2R_ERR_HIP;
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return P2R_OK;
}

// Validated arguments and workspace layout of one adjoint call.
struct AdjointPlan {
  uint32_t n = 0, bs = 0;
  int nlayer = 0, nglob = 0, cs = 0, grad_precision = 0;
  unsigned grid = 0;
  // workspace: checkpoints | chi2 totals | partials | segment store (when not in LDS) | the MIXED tail's
  // hand-over (the adjoint state at layer kTailLayers, 28 floats per track)
  size_t ck_bytes = 0, c2_bytes = 0, part_bytes = 0, seg_bytes = 0, bar_bytes = 0, ws_bytes = 0, lds = 0;
  p2r_params p{};
};

int adjoint_plan(p2r_ctx* c, const void* trk_in, const void* hits, const void* trk_out, const p2r_shape* s,
                 const p2r_params* params, int grad_precision, int32_t checkpoint_stride, bool global,
                 AdjointPlan& pl) {
  if (!trk_in || !hits || !trk_out || !s) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (s->nevts < 0 || s->ntrks < 0 || s->nlayer < 1 || s->nlayer > P2R_MAX_NLAYER || s->bsize < 1 ||
      s->bsize > 1024)
    return c->fail(P2R_ERR_SHAPE, "bad shape nevts=%d ntrks=%d nlayer=%d bsize=%d", s->nevts, s->ntrks, s->nlayer,
                   s->bsize);
  if (s->ntrks % s->bsize != 0) return c->fail(P2R_ERR_SHAPE, "ntrks must be a multiple of bsize");
  if (grad_precision < P2R_GRAD_F32 || grad_precision > P2R_GRAD_F64)
    return c->fail(P2R_ERR_ARG, "bad grad_precision %d", grad_precision);
  const int cs = checkpoint_stride ? checkpoint_stride : c->tune.checkpoint_stride;
  if (cs != 1 && cs != 2 && cs != 4 && cs != 8)
    return c->fail(P2R_ERR_ARG, "checkpoint_stride must be 0, 1, 2, 4 or 8");
  const uint64_t ntrk = (uint64_t)s->nevts * (uint64_t)s->ntrks;
  if (ntrk >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks");
  // the reverse sweep rereads trk_in after the forward pass has written trk_out
  if (overlaps(trk_in, ntrk * 112, trk_out, ntrk * 112))
    return c->fail(P2R_ERR_ARG, "trk_out overlaps trk_in (the adjoint rereads the input tracks)");
  const p2r_params p = params ? *params : p2r_default_params();
  if (!(p.b_scale != 0.f)) return c->fail(P2R_ERR_ARG, "b_scale must be non-zero");
  if (p.precision != P2R_EXACT && p.precision != P2R_FAST) return c->fail(P2R_ERR_ARG, "bad precision %d", p.precision);
  if (p.options != 0)
    return c->fail(P2R_ERR_UNSUPPORTED, "p2r_forward_adjoint differentiates the reference's update only (options 0)");
  pl.n = (uint32_t)ntrk;
  pl.bs = (uint32_t)s->bsize;
  pl.nlayer = s->nlayer;
  pl.nglob = 4 * s->nlayer + 1;
  pl.cs = cs;
  pl.grad_precision = grad_precision;
  pl.p = p;
  if (ntrk == 0) return P2R_OK;
  pl.grid = grid_for_n(ntrk, kAdjThreads);
  const size_t sz = grad_precision == P2R_GRAD_F32 ? sizeof(float) : sizeof(double);
  const size_t nck = (size_t)((pl.nlayer + cs - 1) / cs - 1);
  pl.ck_bytes = std::max<size_t>(nck, 1) * 27 * ntrk * sz;
  pl.c2_bytes = ntrk * sz;
  pl.part_bytes = global ? (size_t)pl.grid * pl.nglob * sizeof(double) : 0;
  pl.seg_bytes = seg_global_bytes(grad_precision, cs, ntrk);
  pl.bar_bytes = grad_precision == P2R_GRAD_MIXED ? (size_t)28 * sizeof(float) * ntrk : 0;
  pl.ws_bytes = ck_bytes_aligned(pl.ck_bytes) + ck_bytes_aligned(pl.c2_bytes) + ck_bytes_aligned(pl.part_bytes) +
                ck_bytes_aligned(pl.seg_bytes) + pl.bar_bytes + 256;
  pl.lds = global ? (size_t)(kAdjThreads / 64) * pl.nglob * sizeof(double) : 0;
  return P2R_OK;
}

// Launches one adjoint call on `st` with the workspace `ws` (at least pl.ws_bytes).  Kernel launches only, so
// the call can be captured into a graph.
int adjoint_enqueue(p2r_ctx* c, const AdjointPlan& pl, const void* trk_in, const void* hits, void* trk_out,
                    float* chi2, float* grad_par, float* grad_hits, float* grad_kfact, double* grad_global, char* ws,
                    hipStream_t st) {
  char* c2_total = ws + ck_bytes_aligned(pl.ck_bytes);
  char* part = c2_total + ck_bytes_aligned(pl.c2_bytes);
  double* partials = grad_global ? reinterpret_cast<double*>(part) : nullptr;
  char* seg = pl.seg_bytes ? part + ck_bytes_aligned(pl.part_bytes) : nullptr;
  float* bar = pl.bar_bytes
                   ? reinterpret_cast<float*>(part + ck_bytes_aligned(pl.part_bytes) + ck_bytes_aligned(pl.seg_bytes))
                   : nullptr;
  const float kf = pl.p.kfact / pl.p.b_scale;
  auto T = static_cast<const float*>(trk_in);
  auto H = static_cast<const float*>(hits);
  auto O = static_cast<float*>(trk_out);
  const int cs = pl.cs, gp = pl.grad_precision;
  const bool exact = pl.p.precision == P2R_EXACT;
#define P2R_ADJ(E, CS)                                                                                         \
  launch_adjoint_p<E, CS>(gp, pl.grid, pl.lds, st, T, H, O, chi2, grad_par, grad_hits, grad_kfact, partials, ws, \
                          c2_total, seg, bar, pl.n, pl.bs, pl.nlayer, kf)
  if (exact) {
    if (cs == 1) P2R_ADJ(true, 1); else if (cs == 2) P2R_ADJ(true, 2); else if (cs == 4) P2R_ADJ(true, 4);
    else P2R_ADJ(true, 8);
  } else {
    if (cs == 1) P2R_ADJ(false, 1); else if (cs == 2) P2R_ADJ(false, 2); else if (cs == 4) P2R_ADJ(false, 4);
    else P2R_ADJ(false, 8);
  }
#undef P2R_ADJ
  const int rc = c->check_launch("p2r_forward_adjoint");
  if (rc != P2R_OK || !grad_global) return rc;
  // d/d b_scale = (d/d kfact_eff) * (-kfact / b^2); the kernel summed -kf q dchi2/dk = -kf dchi2/dkfact_eff,
  // which equals b * d/db: divide by b.
  finalize_global_kernel<<<pl.nglob, kFinThreads, 0, st>>>(partials, grad_global, (int)pl.grid, pl.nglob,
                                                          1.0 / (double)pl.p.b_scale);
  return c->check_launch("p2r_forward_adjoint (reduction)");
}

// Captures what `enqueue(stream)` launches into an instantiated graph owned by a new p2r_graph (which takes
// `ws` on success).  Relaxed capture mode: the enqueue paths query occupancy, which the stricter modes may
// refuse while a capture is running.
template <class F>
int capture_graph(p2r_ctx* c, void* ws, p2r_graph** out, F&& enqueue) {
  DeviceGuard g(c->device);
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    return c->fail(P2R_ERR_HIP, "hipStreamCreateWithFlags failed");
  if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipStreamDestroy(st);
    return c->fail(P2R_ERR_HIP, "hipStreamBeginCapture failed");
  }
  const int rc = enqueue(st);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  (void)hipStreamDestroy(st);
  if (rc != P2R_OK || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc != P2R_OK ? rc : c->fail(P2R_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  }
  size_t nodes = 0;
  if (hipGraphGetNodes(graph, nullptr, &nodes) != hipSuccess) nodes = 1;
  hipGraphExec_t exec = nullptr;
  if (nodes > 0 && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return c->fail(P2R_ERR_HIP, "hipGraphInstantiate failed");
  }
  (void)hipGraphDestroy(graph);
  p2r_graph* gr = new (std::nothrow) p2r_graph();
  if (!gr) {
    if (exec) (void)hipGraphExecDestroy(exec);
    return c->fail(P2R_ERR_NOMEM, "graph handle");
  }
  gr->device = c->device;
  gr->exec = exec;
  gr->ws = ws;
  *out = gr;
  return P2R_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

p2r_params p2r_default_params(void) {
  p2r_params p;
  p.kfact = (float)(100 / (-0.299792458 * 3.8112));  // tbb.cpp:654
  p.b_scale = 1.f;
  p.precision = P2R_EXACT;
  p.options = 0;
  return p;
}

const char* p2r_error_string(int status) {
  switch (status) {
    case P2R_OK: return "success";
    case P2R_ERR_ARG: return "invalid argument";
    case P2R_ERR_SHAPE: return "unsupported shape";
    case P2R_ERR_HIP: return "HIP runtime error";
    case P2R_ERR_UNSUPPORTED: return "not supported";
    case P2R_ERR_NOMEM: return "out of device memory";
    default: return "unknown status";
  }
}

int p2r_create(int device, p2r_ctx** out) {
  if (!out) return P2R_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return P2R_ERR_ARG;
  p2r_ctx* c = new (std::nothrow) p2r_ctx();
  if (!c) return P2R_ERR_NOMEM;
  c->device = device;
  DeviceGuard g(device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete c;
    return P2R_ERR_HIP;
  }
  c->num_cus = prop.multiProcessorCount;
  // the no-allocation entry point's static scratch on this device (its status is kept: only that entry
  // point needs it)
  c->scratch_status = resolve_scratch(device, &c->scratch);
  c->roctx.load();
  *out = c;
  return P2R_OK;
}

int p2r_destroy(p2r_ctx* c) {
  if (!c) return P2R_ERR_ARG;
  {
    DeviceGuard g(c->device);
    c->release_workspace();
    c->release_streams();
  }
  delete c;
  return P2R_OK;
}

const char* p2r_last_error(const p2r_ctx* c) { return c ? c->msg : "null context"; }

int p2r_forward(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s, int semantics,
                const p2r_params* params, float* chi2, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_forward");
  if (!trk_in || !hits || !trk_out || !s) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (semantics != P2R_REFERENCE && semantics != P2R_CHAINED) return c->fail(P2R_ERR_ARG, "bad semantics %d", semantics);
  if (s->nevts < 0 || s->ntrks < 0 || s->nlayer < 1 || s->nlayer > P2R_MAX_NLAYER || s->bsize < 1 ||
      s->bsize > 1024)
    return c->fail(P2R_ERR_SHAPE, "bad shape nevts=%d ntrks=%d nlayer=%d bsize=%d", s->nevts, s->ntrks, s->nlayer,
                   s->bsize);
  if (s->ntrks % s->bsize != 0)
    return c->fail(P2R_ERR_SHAPE, "ntrks=%d is not a multiple of bsize=%d (use the ragged entry point)", s->ntrks,
                   s->bsize);
  const uint64_t ntrk = (uint64_t)s->nevts * (uint64_t)s->ntrks;
  if (ntrk >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks (%llu)", (unsigned long long)ntrk);
  const p2r_params p = params ? *params : p2r_default_params();
  if (!(p.b_scale != 0.f)) return c->fail(P2R_ERR_ARG, "b_scale must be non-zero");
  if (p.precision != P2R_EXACT && p.precision != P2R_FAST) return c->fail(P2R_ERR_ARG, "bad precision %d", p.precision);
  if (p.options & ~kAllOptions) return c->fail(P2R_ERR_ARG, "bad options %d", p.options);
  if (ntrk == 0) return P2R_OK;
  DeviceGuard g(c->device);
  const UniformLaunch a{static_cast<const float*>(trk_in), static_cast<const float*>(hits), static_cast<float*>(trk_out),
                        chi2, (uint32_t)ntrk, (uint32_t)s->bsize, s->nlayer, p.kfact / p.b_scale, p.options};
  launch_uniform(semantics, p.precision == P2R_EXACT, c->tune.fwd_threads, c->tune.fwd_layers_per_launch,
                 c->tune.fwd_waves, static_cast<hipStream_t>(stream), a);
  return c->check_launch("p2r_forward");
}


int p2r_forward_streamed(p2r_ctx* c, const void* host_trk_in, const void* host_hits, void* host_trk_out,
                         void* dev_trk_in, void* dev_hits, void* dev_trk_out, const p2r_shape* s, int semantics,
                         const p2r_params* params, int32_t nstreams) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_forward_streamed");
  if (!host_trk_in || !host_hits || !host_trk_out || !dev_trk_in || !dev_hits || !dev_trk_out || !s)
    return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (semantics != P2R_REFERENCE && semantics != P2R_CHAINED) return c->fail(P2R_ERR_ARG, "bad semantics %d", semantics);
  if (nstreams < 1 || nstreams > 16) return c->fail(P2R_ERR_ARG, "nstreams must be in 1..16");
  if (s->nevts < 0 || s->ntrks < 0 || s->nlayer < 1 || s->nlayer > P2R_MAX_NLAYER || s->bsize < 1 ||
      s->bsize > 1024 || s->ntrks % s->bsize != 0)
    return c->fail(P2R_ERR_SHAPE, "bad shape nevts=%d ntrks=%d nlayer=%d bsize=%d", s->nevts, s->ntrks, s->nlayer,
                   s->bsize);
  const uint64_t nblk = (uint64_t)s->nevts * (uint64_t)(s->ntrks / s->bsize);
  if (nblk * (uint64_t)s->bsize >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks");
  const p2r_params p = params ? *params : p2r_default_params();
  if (!(p.b_scale != 0.f)) return c->fail(P2R_ERR_ARG, "b_scale must be non-zero");
  if (p.precision != P2R_EXACT && p.precision != P2R_FAST) return c->fail(P2R_ERR_ARG, "bad precision %d", p.precision);
  if (p.options & ~kAllOptions) return c->fail(P2R_ERR_ARG, "bad options %d", p.options);
  if (nblk == 0) return P2R_OK;
  DeviceGuard g(c->device);
  const size_t bs = (size_t)s->bsize, tw = 28 * bs * 4, hw = (size_t)s->nlayer * 9 * bs * 4;  // bytes per block
  const int ns = (int)std::min<uint64_t>((uint64_t)nstreams, nblk);
  // On an error, the chunks already queued are drained before returning, so no copy or launch of this call
  // is still reading or writing the caller's buffers when it returns.
  auto drain = [&](int upto) {
    for (int k = 0; k < upto; ++k)
      if (hipStream_t st = c->stream(k)) (void)hipStreamSynchronize(st);
  };
  for (int k = 0; k < ns; ++k) {
    hipStream_t st = c->stream(k);
    if (!st) {
      drain(k);
      return c->fail(P2R_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    const uint64_t b0 = nblk * k / ns, b1 = nblk * (k + 1) / ns;  // contiguous block range of this chunk
    if (b1 == b0) continue;
    const size_t nb = b1 - b0;
    auto off = [](const void* base, size_t bytes) { return static_cast<const char*>(base) + bytes; };
    char* dt = static_cast<char*>(dev_trk_in) + b0 * tw;
    char* dh = static_cast<char*>(dev_hits) + b0 * hw;
    char* dout = static_cast<char*>(dev_trk_out) + b0 * tw;
    if (hipMemcpyAsync(dt, off(host_trk_in, b0 * tw), nb * tw, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dh, off(host_hits, b0 * hw), nb * hw, hipMemcpyHostToDevice, st) != hipSuccess) {
      drain(k + 1);
      return c->fail(P2R_ERR_HIP, "host-to-device copy of chunk %d failed", k);
    }
    const UniformLaunch a{reinterpret_cast<const float*>(dt), reinterpret_cast<const float*>(dh),
                          reinterpret_cast<float*>(dout), nullptr, (uint32_t)(nb * bs), (uint32_t)bs, s->nlayer,
                          p.kfact / p.b_scale, p.options};
    launch_uniform(semantics, p.precision == P2R_EXACT, c->tune.fwd_threads, c->tune.fwd_layers_per_launch,
                   c->tune.fwd_waves, st, a);
    const int rc = c->check_launch("p2r_forward_streamed");
    if (rc != P2R_OK) {
      drain(k + 1);
      return rc;
    }
    if (hipMemcpyAsync(static_cast<char*>(host_trk_out) + b0 * tw, dout, nb * tw, hipMemcpyDeviceToHost, st) !=
        hipSuccess) {
      drain(k + 1);
      return c->fail(P2R_ERR_HIP, "device-to-host copy of chunk %d failed", k);
    }
  }
  for (int k = 0; k < ns; ++k)
    if (hipStreamSynchronize(c->stream(k)) != hipSuccess) return c->fail(P2R_ERR_HIP, "stream %d failed", k);
  return P2R_OK;
}


int p2r_forward_ragged(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const int64_t* blk_offsets,
                       int32_t nevts, const uint32_t* layer_mask, int32_t nlayer, int32_t bsize, int semantics,
                       const p2r_params* params, float* chi2, int32_t tile_rounds, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_forward_ragged");
  if (!trk_in || !hits || !trk_out || !blk_offsets || !layer_mask) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (semantics != P2R_REFERENCE && semantics != P2R_CHAINED) return c->fail(P2R_ERR_ARG, "bad semantics %d", semantics);
  if (nevts < 0 || nlayer < 1 || nlayer > 32 || bsize < 1 || bsize > 1024)
    return c->fail(P2R_ERR_SHAPE, "bad shape nevts=%d nlayer=%d bsize=%d (nlayer <= 32 for the 32-bit mask)", nevts,
                   nlayer, bsize);
  if (tile_rounds != P2R_RAGGED_AUTO && tile_rounds != P2R_RAGGED_LAYER_ORDER &&
      tile_rounds != P2R_RAGGED_PAIR_STREAM && tile_rounds != 1 && tile_rounds != 2 && tile_rounds != 4)
    return c->fail(P2R_ERR_ARG,
                   "tile_rounds must be P2R_RAGGED_AUTO, P2R_RAGGED_LAYER_ORDER, P2R_RAGGED_PAIR_STREAM, 1, 2 or 4");
  if (blk_offsets[0] != 0) return c->fail(P2R_ERR_SHAPE, "blk_offsets[0] must be 0");
  for (int32_t e = 0; e < nevts; ++e)
    if (blk_offsets[e + 1] < blk_offsets[e]) return c->fail(P2R_ERR_SHAPE, "blk_offsets decrease at event %d", e);
  const uint64_t nslot = (uint64_t)blk_offsets[nevts] * (uint64_t)bsize;
  if (nslot >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many slots (%llu)", (unsigned long long)nslot);
  const p2r_params p = params ? *params : p2r_default_params();
  if (!(p.b_scale != 0.f)) return c->fail(P2R_ERR_ARG, "b_scale must be non-zero");
  if (p.precision != P2R_EXACT && p.precision != P2R_FAST) return c->fail(P2R_ERR_ARG, "bad precision %d", p.precision);
  if (p.options & ~kAllOptions) return c->fail(P2R_ERR_ARG, "bad options %d", p.options);
  if (nslot == 0) return P2R_OK;
  const int rounds = tile_rounds != P2R_RAGGED_AUTO ? tile_rounds : c->tune.ragged_schedule;
  DeviceGuard g(c->device);
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto T = static_cast<const float*>(trk_in);
  auto H = static_cast<const float*>(hits);
  auto O = static_cast<float*>(trk_out);
  const float kf = p.kfact / p.b_scale;
  const bool exact = p.precision == P2R_EXACT;
  if (semantics == P2R_REFERENCE)
    launch_ragged_s<P2R_REFERENCE>(chi2 != nullptr, exact, rounds, c, st, T, H, O, chi2, layer_mask, (uint32_t)nslot,
                                   (uint32_t)bsize, nlayer, kf, p.options);
  else
    launch_ragged_s<P2R_CHAINED>(chi2 != nullptr, exact, rounds, c, st, T, H, O, chi2, layer_mask, (uint32_t)nslot,
                                 (uint32_t)bsize, nlayer, kf, p.options);
  return c->check_launch("p2r_forward_ragged");
}


int p2r_forward_adjoint(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s,
                        const p2r_params* params, int grad_precision, int32_t checkpoint_stride, float* chi2,
                        float* grad_par, float* grad_hits, float* grad_kfact, double* grad_global, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_forward_adjoint");
  AdjointPlan pl;
  const int rc = adjoint_plan(c, trk_in, hits, trk_out, s, params, grad_precision, checkpoint_stride,
                              grad_global != nullptr, pl);
  if (rc != P2R_OK || pl.n == 0) return rc;
  DeviceGuard g(c->device);
  char* ws = static_cast<char*>(c->workspace(pl.ws_bytes));
  if (!ws) return c->fail(P2R_ERR_NOMEM, "workspace of %zu bytes", pl.ws_bytes);
  return adjoint_enqueue(c, pl, trk_in, hits, trk_out, chi2, grad_par, grad_hits, grad_kfact, grad_global, ws,
                         static_cast<hipStream_t>(stream));
}

int p2r_forward_adjoint_noalloc(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s,
                                const p2r_params* params, int grad_precision, int32_t checkpoint_stride, float* chi2,
                                float* grad_par, float* grad_hits, float* grad_kfact, double* grad_global,
                                void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_forward_adjoint_noalloc");
  AdjointPlan pl;
  const int rc = adjoint_plan(c, trk_in, hits, trk_out, s, params, grad_precision, checkpoint_stride,
                              grad_global != nullptr, pl);
  if (rc != P2R_OK || pl.n == 0) return rc;
  if (pl.ws_bytes > (size_t)P2R_NOALLOC_SCRATCH_BYTES)
    return c->fail(P2R_ERR_SHAPE,
                   "p2r_forward_adjoint_noalloc: %u tracks x %d layers need %zu bytes of scratch, the static "
                   "scratch holds %d (fewer tracks per call, a larger checkpoint stride or P2R_GRAD_F32)",
                   pl.n, pl.nlayer, pl.ws_bytes, (int)P2R_NOALLOC_SCRATCH_BYTES);
  if (c->scratch_status != P2R_OK || !c->scratch)
    return c->fail(P2R_ERR_HIP, "p2r_forward_adjoint_noalloc: no static scratch on device %d", c->device);
  DeviceGuard g(c->device);
  // every word of the scratch this call reads it writes first (checkpoints, chi2 totals, one partial per
  // workgroup and global entry): nothing needs clearing between calls
  return adjoint_enqueue(c, pl, trk_in, hits, trk_out, chi2, grad_par, grad_hits, grad_kfact, grad_global,
                         static_cast<char*>(c->scratch), static_cast<hipStream_t>(stream));
}

// ---- HIP graphs ------------------------------------------------------------------------------------------------
int p2r_graph_forward(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s,
                      int semantics, const p2r_params* params, float* chi2, p2r_graph** out) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_graph_forward");
  if (!out) return c->fail(P2R_ERR_ARG, "null graph pointer");
  *out = nullptr;
  return capture_graph(c, nullptr, out, [&](hipStream_t st) {
    return p2r_forward(c, trk_in, hits, trk_out, s, semantics, params, chi2, st);
  });
}

int p2r_graph_forward_ragged(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out,
                             const int64_t* blk_offsets, int32_t nevts, const uint32_t* layer_mask, int32_t nlayer,
                             int32_t bsize, int semantics, const p2r_params* params, float* chi2,
                             int32_t tile_rounds, p2r_graph** out) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_graph_forward_ragged");
  if (!out) return c->fail(P2R_ERR_ARG, "null graph pointer");
  *out = nullptr;
  return capture_graph(c, nullptr, out, [&](hipStream_t st) {
    return p2r_forward_ragged(c, trk_in, hits, trk_out, blk_offsets, nevts, layer_mask, nlayer, bsize, semantics,
                              params, chi2, tile_rounds, st);
  });
}

int p2r_graph_forward_adjoint(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out,
                              const p2r_shape* s, const p2r_params* params, int grad_precision,
                              int32_t checkpoint_stride, float* chi2, float* grad_par, float* grad_hits,
                              float* grad_kfact, double* grad_global, p2r_graph** out) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_graph_forward_adjoint");
  if (!out) return c->fail(P2R_ERR_ARG, "null graph pointer");
  *out = nullptr;
  AdjointPlan pl;
  int rc = adjoint_plan(c, trk_in, hits, trk_out, s, params, grad_precision, checkpoint_stride,
                        grad_global != nullptr, pl);
  if (rc != P2R_OK) return rc;
  // The graph owns its workspace (allocated here, before the capture), so later calls on the context, of any
  // size, leave it alone.
  DeviceGuard g(c->device);
  void* ws = nullptr;
  if (pl.n > 0 && hipMalloc(&ws, pl.ws_bytes) != hipSuccess)
    return c->fail(P2R_ERR_NOMEM, "graph workspace of %zu bytes", pl.ws_bytes);
  rc = capture_graph(c, ws, out, [&](hipStream_t st) {
    if (pl.n == 0) return (int)P2R_OK;
    return adjoint_enqueue(c, pl, trk_in, hits, trk_out, chi2, grad_par, grad_hits, grad_kfact, grad_global,
                           static_cast<char*>(ws), st);
  });
  if (rc != P2R_OK && ws) (void)hipFree(ws);
  return rc;
}

int p2r_graph_launch(p2r_graph* graph, void* stream) {
  if (!graph) return P2R_ERR_ARG;
  if (!graph->exec) return P2R_OK;  // captured an empty batch
  DeviceGuard g(graph->device);
  return hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? P2R_OK : P2R_ERR_HIP;
}

int p2r_graph_destroy(p2r_graph* graph) {
  if (!graph) return P2R_ERR_ARG;
  {
    DeviceGuard g(graph->device);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->ws) {
      (void)hipDeviceSynchronize();  // replays still in flight may use the workspace
      (void)hipFree(graph->ws);
    }
  }
  delete graph;
  return P2R_OK;
}

int p2r_propagateToR(p2r_ctx* c, const float* inErr, const float* inPar, const int32_t* inChg, const float* msP,
                     float* outErr, float* outPar, int32_t bsize, int64_t nblk, int64_t trk_stride,
                     int64_t hit_stride, float kfact, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_propagateToR");
  if (!inErr || !inPar || !inChg || !msP || !outErr || !outPar) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (bsize < 1 || bsize > 1024 || nblk < 0 || trk_stride < 0 || hit_stride < 0)
    return c->fail(P2R_ERR_SHAPE, "bad bsize/nblk/stride");
  const uint64_t ntrk = (uint64_t)nblk * (uint64_t)bsize;
  if (ntrk >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks");
  if (ntrk == 0) return P2R_OK;
  DeviceGuard g(c->device);
  propagate_blocks_kernel<<<grid_for(ntrk), kThreads, 0, static_cast<hipStream_t>(stream)>>>(
      inErr, inPar, inChg, msP, outErr, outPar, (uint32_t)ntrk, (uint32_t)bsize, trk_stride, hit_stride, kfact);
  return c->check_launch("p2r_propagateToR");
}

int p2r_KalmanUpdate(p2r_ctx* c, float* trkErr, float* inPar, const float* hitErr, const float* msP, int32_t bsize,
                     int64_t nblk, int64_t trk_stride, int64_t hit_stride, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_KalmanUpdate");
  if (!trkErr || !inPar || !hitErr || !msP) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (bsize < 1 || bsize > 1024 || nblk < 0 || trk_stride < 0 || hit_stride < 0)
    return c->fail(P2R_ERR_SHAPE, "bad bsize/nblk/stride");
  const uint64_t ntrk = (uint64_t)nblk * (uint64_t)bsize;
  if (ntrk >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks");
  if (ntrk == 0) return P2R_OK;
  DeviceGuard g(c->device);
  kalman_blocks_kernel<<<grid_for(ntrk), kThreads, 0, static_cast<hipStream_t>(stream)>>>(
      trkErr, inPar, hitErr, msP, (uint32_t)ntrk, (uint32_t)bsize, trk_stride, hit_stride);
  return c->check_launch("p2r_KalmanUpdate");
}

}  // extern "C"
