// C-ABI of the p2r hot path on MI355X (declarations and layout: include/p2r.h).
//
// Kernels: one lane per track over a flat track index g = block*bs + lane.
// Each lane loads its track's 27 floats + charge from the Matriplex SoA block
// (coalesced: consecutive lanes read consecutive words of each field), keeps
// the state in VGPRs through every layer, streams the 9 hit floats per layer
// (next layer prefetched while the current one computes) and writes the
// result once.  HBM traffic per track: 112 B in + 112 B out + 36 B per layer.
#include "../../include/p2r.h"
#include "p2r_ctx.hpp"
#include "p2r_lane.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>

using namespace p2r;

namespace {

constexpr int kThreads = 256;

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

// Whole batch, all layers per lane.  SEM: P2R_REFERENCE / P2R_CHAINED;
// EXACT selects the per-lane math instance (p2r_lane.hpp).
template <int SEM, bool CHI2, bool EXACT>
__global__ __launch_bounds__(kThreads) void fwd_uniform_kernel(const float* __restrict__ trk,
                                                               const float* __restrict__ hits,
                                                               float* __restrict__ out, float* __restrict__ chi2,
                                                               uint32_t ntrk, uint32_t bs, int nlayer, float kfact) {
  const uint32_t g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= ntrk) return;
  const uint32_t blk = g / bs, lane = g - blk * bs;
  const float* tb = trk + (size_t)blk * 28 * bs + lane;
  Trk in;
  load_trk(tb, bs, in);
  const int q = reinterpret_cast<const int*>(tb)[(size_t)27 * bs];
  const float k = (float)q * kfact;  // tbb.cpp:667

  const float* hb = hits + (size_t)blk * nlayer * 9 * bs + lane;
  const size_t hstride = (size_t)9 * bs;
  Trk cur = in;
  float c2 = 0.f;
  Hit hn;
  load_hit(hb, bs, hn);
  for (int l = 0; l < nlayer; ++l) {
    const Hit h = hn;
    if (l + 1 < nlayer) load_hit(hb + (l + 1) * hstride, bs, hn);
    Trk src;
    if (SEM == P2R_REFERENCE) {
      src = in;      // tbb.cpp:894: every layer starts from the input track
      opaque(src);   // recompute per layer; nothing hoisted across layers
    } else {
      src = cur;     // propagate-tor-test_hip.cu:842-845
    }
    Trk prop;
    Lane<EXACT>::propagate(src, k, h.pos[0], h.pos[1], prop);
    float c = Lane<EXACT>::update(prop, h);
    if (SEM == P2R_REFERENCE) {
      // Only the last layer's result is returned in reference semantics; the
      // barrier makes every layer's result observable so the compiler cannot
      // sink the update out of the loop and skip layers 0..n-2.
      opaque(prop);
      opaque(c);
    }
    if (CHI2) c2 += c;
    cur = prop;
  }
  float* ob = out + (size_t)blk * 28 * bs + lane;
  store_trk(ob, bs, cur);
  reinterpret_cast<int*>(ob)[(size_t)27 * bs] = q;
  if (CHI2) chi2[g] = c2;
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

template <int SEM, bool CHI2>
void launch_uniform_p(bool exact, dim3 grid, dim3 block, hipStream_t st, const float* T, const float* H, float* O,
                      float* chi2, uint32_t n, uint32_t bs, int nlayer, float kf) {
  if (exact) fwd_uniform_kernel<SEM, CHI2, true><<<grid, block, 0, st>>>(T, H, O, chi2, n, bs, nlayer, kf);
  else fwd_uniform_kernel<SEM, CHI2, false><<<grid, block, 0, st>>>(T, H, O, chi2, n, bs, nlayer, kf);
}
void launch_uniform(int sem, bool chi2_on, bool exact, dim3 grid, dim3 block, hipStream_t st, const float* T,
                    const float* H, float* O, float* chi2, uint32_t n, uint32_t bs, int nlayer, float kf) {
  if (sem == P2R_REFERENCE) {
    if (chi2_on) launch_uniform_p<P2R_REFERENCE, true>(exact, grid, block, st, T, H, O, chi2, n, bs, nlayer, kf);
    else launch_uniform_p<P2R_REFERENCE, false>(exact, grid, block, st, T, H, O, chi2, n, bs, nlayer, kf);
  } else {
    if (chi2_on) launch_uniform_p<P2R_CHAINED, true>(exact, grid, block, st, T, H, O, chi2, n, bs, nlayer, kf);
    else launch_uniform_p<P2R_CHAINED, false>(exact, grid, block, st, T, H, O, chi2, n, bs, nlayer, kf);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

p2r_params p2r_default_params(void) {
  p2r_params p;
  p.kfact = (float)(100 / (-0.299792458 * 3.8112));  // tbb.cpp:654
  p.b_scale = 1.f;
  p.precision = P2R_EXACT;
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
  *out = c;
  return P2R_OK;
}

int p2r_destroy(p2r_ctx* c) {
  if (!c) return P2R_ERR_ARG;
  {
    DeviceGuard g(c->device);
    c->release_workspace();
  }
  delete c;
  return P2R_OK;
}

const char* p2r_last_error(const p2r_ctx* c) { return c ? c->msg : "null context"; }

int p2r_forward(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s, int semantics,
                const p2r_params* params, float* chi2, void* stream) {
  if (!c) return P2R_ERR_ARG;
  if (!trk_in || !hits || !trk_out || !s) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (semantics != P2R_REFERENCE && semantics != P2R_CHAINED) return c->fail(P2R_ERR_ARG, "bad semantics %d", semantics);
  if (s->nevts < 0 || s->ntrks < 0 || s->nlayer < 1 || s->bsize < 1 || s->bsize > 1024)
    return c->fail(P2R_ERR_SHAPE, "bad shape nevts=%d ntrks=%d nlayer=%d bsize=%d", s->nevts, s->ntrks, s->nlayer,
                   s->bsize);
  if (s->ntrks % s->bsize != 0)
    return c->fail(P2R_ERR_SHAPE, "ntrks=%d is not a multiple of bsize=%d (use the ragged entry point)", s->ntrks,
                   s->bsize);
  const uint64_t ntrk = (uint64_t)s->nevts * (uint64_t)s->ntrks;
  if (ntrk >= (1ull << 31)) return c->fail(P2R_ERR_SHAPE, "too many tracks (%llu)", (unsigned long long)ntrk);
  const p2r_params p = params ? *params : p2r_default_params();
  if (!(p.b_scale != 0.f)) return c->fail(P2R_ERR_ARG, "b_scale must be non-zero");
  if (ntrk == 0) return P2R_OK;
  const float kf = p.kfact / p.b_scale;
  DeviceGuard g(c->device);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid(grid_for(ntrk)), block(kThreads);
  auto T = static_cast<const float*>(trk_in);
  auto H = static_cast<const float*>(hits);
  auto O = static_cast<float*>(trk_out);
  const uint32_t n = (uint32_t)ntrk, bs = (uint32_t)s->bsize;
  if (p.precision != P2R_EXACT && p.precision != P2R_FAST) return c->fail(P2R_ERR_ARG, "bad precision %d", p.precision);
  launch_uniform(semantics, chi2 != nullptr, p.precision == P2R_EXACT, grid, block, st, T, H, O, chi2, n, bs,
                 s->nlayer, kf);
  return c->check_launch("p2r_forward");
}

int p2r_propagateToR(p2r_ctx* c, const float* inErr, const float* inPar, const int32_t* inChg, const float* msP,
                     float* outErr, float* outPar, int32_t bsize, int64_t nblk, int64_t trk_stride,
                     int64_t hit_stride, float kfact, void* stream) {
  if (!c) return P2R_ERR_ARG;
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
