// Launch choices of a context and the autotuners over them (include/p2r.h).
//
// The autotuners are host code over the public entry points: they set a
// candidate tuning, time the entry point with HIP events on the caller's
// stream, and keep the fastest.  The only kernel here is the Matriplex width
// repack used to time other widths.
#include "../../include/p2r.h"
#include "p2r_ctx.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <vector>

namespace {

bool threads_ok(int t) { return t == 64 || t == 128 || t == 256; }
bool schedule_ok(int s) {
  return s == P2R_RAGGED_LAYER_ORDER || s == P2R_RAGGED_PAIR_STREAM || s == 1 || s == 2 || s == 4;
}
bool stride_ok(int s) { return s == 1 || s == 2 || s == 4 || s == 8; }

// Mean time (ms) of `reps` calls after 2 untimed ones; P2R status of the first failing call otherwise.
int time_calls(p2r_ctx* c, hipStream_t st, int reps, const std::function<int()>& call, float* ms) {
  for (int i = 0; i < 2; ++i) {
    const int rc = call();
    if (rc != P2R_OK) return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return c->fail(P2R_ERR_HIP, "hipEventCreate failed");
  }
  int rc = P2R_OK;
  if (hipEventRecord(e0, st) != hipSuccess) rc = c->fail(P2R_ERR_HIP, "hipEventRecord failed");
  for (int i = 0; rc == P2R_OK && i < reps; ++i) rc = call();
  if (rc == P2R_OK && hipEventRecord(e1, st) != hipSuccess) rc = c->fail(P2R_ERR_HIP, "hipEventRecord failed");
  if (rc == P2R_OK && hipEventSynchronize(e1) != hipSuccess) rc = c->fail(P2R_ERR_HIP, "hipEventSynchronize failed");
  float t = 0.f;
  if (rc == P2R_OK && hipEventElapsedTime(&t, e0, e1) != hipSuccess)
    rc = c->fail(P2R_ERR_HIP, "hipEventElapsedTime failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == P2R_OK) *ms = t / (float)reps;
  return rc;
}

// Runs `call` until >= 100 ms of GPU time have passed (at most 200 rounds of 12 calls): the clocks ramp up in the
// first tens of milliseconds of work, which would otherwise favour the candidates timed last.
int warm_up(p2r_ctx* c, hipStream_t st, const std::function<int()>& call) {
  float total = 0.f;
  for (int round = 0; round < 200 && total < 100.f; ++round) {
    float ms = 0.f;
    const int rc = time_calls(c, st, 10, call, &ms);
    if (rc != P2R_OK) return rc;
    total += 12.f * ms;
  }
  return P2R_OK;
}

// Times candidate i (selected by set(i)) for i in [0, n) in two interleaved passes and keeps the smaller
// mean of each; set(i) must leave the context ready for call().
int time_candidates(p2r_ctx* c, hipStream_t st, int reps, int n, const std::function<void(int)>& set,
                    const std::function<int()>& call, float* best_ms) {
  set(0);
  int rc = warm_up(c, st, call);
  for (int pass = 0; rc == P2R_OK && pass < 2; ++pass)
    for (int i = 0; rc == P2R_OK && i < n; ++i) {
      set(i);
      float ms = 0.f;
      rc = time_calls(c, st, reps, call, &ms);
      if (rc == P2R_OK && (pass == 0 || ms < best_ms[i])) best_ms[i] = ms;
    }
  return rc;
}

struct Collect {
  p2r_tune_report rep{};
  void add(const p2r_tune_entry& e) {
    if (rep.count < 32) rep.entry[rep.count++] = e;
  }
  void finish(p2r_tune_report* out) {
    rep.best = 0;
    for (int i = 1; i < rep.count; ++i)
      if (rep.entry[i].ms < rep.entry[rep.best].ms) rep.best = i;
    if (out) *out = rep;
  }
};

p2r_tune_entry entry_of(const p2r_tuning& t, int bsize, float ms, int schedule) {
  return p2r_tune_entry{bsize, schedule == P2R_RAGGED_LAYER_ORDER ? t.ragged_threads : t.fwd_threads,
                        t.fwd_layers_per_launch, schedule, t.checkpoint_stride, t.fwd_waves, ms};
}

// dst (dbs layout) <- src (sbs layout), one lane-word per thread; consecutive
// threads take consecutive tracks of one (event, sub-block, word), so reads and
// writes run along Matriplex rows.
__global__ void repack_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint32_t sbs,
                              uint32_t dbs, uint32_t ntrks, uint32_t nsub, uint32_t words, uint64_t total) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const uint64_t t = i % ntrks;
    uint64_t r = i / ntrks;
    const uint64_t w = r % words;
    r /= words;
    const uint64_t sub = r % nsub;
    const uint64_t e = r / nsub;
    const uint64_t sblk = e * (ntrks / sbs) + t / sbs, dblk = e * (ntrks / dbs) + t / dbs;
    dst[((dblk * nsub + sub) * words + w) * dbs + t % dbs] = src[((sblk * nsub + sub) * words + w) * sbs + t % sbs];
  }
}

}  // namespace

extern "C" {

int p2r_get_tuning(const p2r_ctx* c, p2r_tuning* out) {
  if (!c || !out) return P2R_ERR_ARG;
  *out = c->tune;
  return P2R_OK;
}

int p2r_set_tuning(p2r_ctx* c, const p2r_tuning* in) {
  if (!c) return P2R_ERR_ARG;
  if (!in) return c->fail(P2R_ERR_ARG, "null tuning");
  if (!threads_ok(in->fwd_threads) || !threads_ok(in->ragged_threads))
    return c->fail(P2R_ERR_ARG, "threads must be 64, 128 or 256");
  if (in->fwd_layers_per_launch < 0) return c->fail(P2R_ERR_ARG, "fwd_layers_per_launch must be >= 0");
  if (!schedule_ok(in->ragged_schedule))
    return c->fail(P2R_ERR_ARG, "ragged_schedule must be P2R_RAGGED_LAYER_ORDER, P2R_RAGGED_PAIR_STREAM, 1, 2 or 4");
  if (!stride_ok(in->checkpoint_stride)) return c->fail(P2R_ERR_ARG, "checkpoint_stride must be 1, 2, 4 or 8");
  if (in->fwd_waves != 0 && in->fwd_waves != 1 && in->fwd_waves != 3 && in->fwd_waves != 4)
    return c->fail(P2R_ERR_ARG, "fwd_waves must be 0, 1, 3 or 4");
  c->tune = *in;
  return P2R_OK;
}

int p2r_autotune_forward(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s,
                         int semantics, const p2r_params* params, int32_t reps, p2r_tune_report* report,
                         void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_autotune_forward");
  if (!s || reps < 1) return c->fail(P2R_ERR_ARG, "null shape or reps < 1");
  if (s->nevts > 0 && s->ntrks > 0 &&
      overlaps(trk_in, (size_t)s->nevts * s->ntrks * 112, trk_out, (size_t)s->nevts * s->ntrks * 112))
    return c->fail(P2R_ERR_ARG, "trk_out overlaps trk_in (every candidate must start from the same input)");
  const p2r_tuning keep = c->tune;
  hipStream_t st = static_cast<hipStream_t>(stream);
  std::vector<p2r_tuning> cand;
  const int per[3] = {0, (s->nlayer + 1) / 2, (s->nlayer + 3) / 4};
  for (int pi = 0; pi < 3; ++pi) {
    if (pi > 0 && (per[pi] >= s->nlayer || per[pi] == per[pi - 1])) continue;
    for (int threads : {64, 128, 256}) {
      p2r_tuning t = keep;
      t.fwd_threads = threads;
      t.fwd_layers_per_launch = per[pi];
      t.fwd_waves = 0;
      cand.push_back(t);
    }
  }
  auto run = [&] { return p2r_forward(c, trk_in, hits, trk_out, s, semantics, params, nullptr, stream); };
  std::vector<float> ms(cand.size(), 0.f);
  int rc = time_candidates(c, st, reps, (int)cand.size(), [&](int i) { c->tune = cand[i]; }, run, ms.data());
  if (rc != P2R_OK) {
    c->tune = keep;
    return rc;
  }
  // register budget at the best workgroup size / layer split
  const size_t b0 = std::min_element(ms.begin(), ms.end()) - ms.begin();
  for (int waves : {1, 3, 4}) {
    p2r_tuning t = cand[b0];
    t.fwd_waves = waves;
    cand.push_back(t);
  }
  float wms[3] = {};
  rc = time_candidates(c, st, reps, 3, [&](int i) { c->tune = cand[cand.size() - 3 + i]; }, run, wms);
  c->tune = keep;
  if (rc != P2R_OK) return rc;
  for (float w : wms) ms.push_back(w);
  Collect col;
  for (size_t i = 0; i < cand.size(); ++i) {
    p2r_tune_entry e = entry_of(cand[i], s->bsize, ms[i], 0);
    e.waves = cand[i].fwd_waves;
    col.add(e);
  }
  col.finish(report);
  const p2r_tuning& b = cand[col.rep.best];
  c->tune.fwd_threads = b.fwd_threads;
  c->tune.fwd_layers_per_launch = b.fwd_layers_per_launch;
  c->tune.fwd_waves = b.fwd_waves;
  return P2R_OK;
}

int p2r_autotune_ragged(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const int64_t* blk_offsets,
                        int32_t nevts, const uint32_t* layer_mask, int32_t nlayer, int32_t bsize, int semantics,
                        const p2r_params* params, int32_t reps, p2r_tune_report* report, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_autotune_ragged");
  if (reps < 1) return c->fail(P2R_ERR_ARG, "reps < 1");
  if (blk_offsets && nevts > 0 && bsize > 0 && blk_offsets[nevts] > 0) {
    const size_t bytes = (size_t)blk_offsets[nevts] * (size_t)bsize * 112;
    if (overlaps(trk_in, bytes, trk_out, bytes))
      return c->fail(P2R_ERR_ARG, "trk_out overlaps trk_in (every candidate must start from the same input)");
  }
  const p2r_tuning keep = c->tune;
  hipStream_t st = static_cast<hipStream_t>(stream);
  struct Cand {
    int schedule, threads;
  };
  const Cand cand[] = {{P2R_RAGGED_LAYER_ORDER, 64}, {P2R_RAGGED_LAYER_ORDER, 128}, {P2R_RAGGED_LAYER_ORDER, 256},
                       {P2R_RAGGED_PAIR_STREAM, 128}, {1, 256}, {2, 256}, {4, 256}};
  constexpr int n = sizeof(cand) / sizeof(cand[0]);
  float ms[n] = {};
  const int rc = time_candidates(
      c, st, reps, n,
      [&](int i) {
        c->tune = keep;
        c->tune.ragged_schedule = cand[i].schedule;
        c->tune.ragged_threads = cand[i].threads;
      },
      [&] {
        return p2r_forward_ragged(c, trk_in, hits, trk_out, blk_offsets, nevts, layer_mask, nlayer, bsize, semantics,
                                  params, nullptr, P2R_RAGGED_AUTO, stream);
      },
      ms);
  c->tune = keep;
  if (rc != P2R_OK) return rc;
  Collect col;
  for (int i = 0; i < n; ++i) {
    p2r_tune_entry e = entry_of(keep, bsize, ms[i], cand[i].schedule);
    e.threads = cand[i].threads;
    col.add(e);
  }
  col.finish(report);
  const p2r_tune_entry& b = col.rep.entry[col.rep.best];
  c->tune.ragged_schedule = b.schedule;
  if (b.schedule == P2R_RAGGED_LAYER_ORDER) c->tune.ragged_threads = b.threads;
  return P2R_OK;
}

int p2r_autotune_adjoint(p2r_ctx* c, const void* trk_in, const void* hits, void* trk_out, const p2r_shape* s,
                         const p2r_params* params, int grad_precision, float* chi2, float* grad_par,
                         float* grad_hits, float* grad_kfact, double* grad_global, int32_t reps,
                         p2r_tune_report* report, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_autotune_adjoint");
  if (!s || reps < 1) return c->fail(P2R_ERR_ARG, "null shape or reps < 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int strides[4] = {1, 2, 4, 8};
  int cs = strides[0];
  float ms[4] = {};
  const int rc = time_candidates(
      c, st, reps, 4, [&](int i) { cs = strides[i]; },
      [&] {
        return p2r_forward_adjoint(c, trk_in, hits, trk_out, s, params, grad_precision, cs, chi2, grad_par,
                                   grad_hits, grad_kfact, grad_global, stream);
      },
      ms);
  if (rc != P2R_OK) return rc;
  Collect col;
  for (int i = 0; i < 4; ++i) {
    p2r_tuning t = c->tune;
    t.checkpoint_stride = strides[i];
    col.add(entry_of(t, s->bsize, ms[i], 0));
  }
  col.finish(report);
  c->tune.checkpoint_stride = col.rep.entry[col.rep.best].checkpoint_stride;
  return P2R_OK;
}

int p2r_repack(p2r_ctx* c, const void* src, int32_t sbs, void* dst, int32_t dbs, int32_t nevts, int32_t ntrks,
               int32_t nsub, int32_t words, void* stream) {
  if (!c) return P2R_ERR_ARG;
  if (!src || !dst) return c->fail(P2R_ERR_ARG, "null pointer argument");
  if (sbs < 1 || sbs > 1024 || dbs < 1 || dbs > 1024 || nevts < 0 || ntrks < 0 || nsub < 1 || words < 1 ||
      ntrks % sbs != 0 || ntrks % dbs != 0)
    return c->fail(P2R_ERR_SHAPE, "bad repack shape (widths must divide ntrks)");
  const uint64_t total = (uint64_t)nevts * (uint64_t)ntrks * (uint64_t)nsub * (uint64_t)words;
  if (total == 0) return P2R_OK;
  const uint64_t bytes = total * 4;
  const char *a = static_cast<const char*>(src), *b = static_cast<const char*>(dst);
  if (a < b + bytes && b < a + bytes) return c->fail(P2R_ERR_ARG, "src and dst overlap");
  DeviceGuard g(c->device);
  const unsigned grid = (unsigned)std::min<uint64_t>((total + 255) / 256, (uint64_t)c->num_cus * 32);
  repack_kernel<<<grid, 256, 0, static_cast<hipStream_t>(stream)>>>(
      static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), (uint32_t)sbs, (uint32_t)dbs, (uint32_t)ntrks,
      (uint32_t)nsub, (uint32_t)words, total);
  return c->check_launch("p2r_repack");
}

int p2r_autotune_width(p2r_ctx* c, const void* trk_in, const void* hits, const p2r_shape* s, int semantics,
                       const p2r_params* params, const int32_t* widths, int32_t nwidths, int32_t reps,
                       p2r_tune_report* report, void* stream) {
  if (!c) return P2R_ERR_ARG;
  RangeScope range(c, "p2r_autotune_width");
  if (!s || !widths || nwidths < 1 || nwidths > 32 || reps < 1)
    return c->fail(P2R_ERR_ARG, "null shape/widths, nwidths outside 1..32 or reps < 1");
  for (int i = 0; i < nwidths; ++i)
    if (widths[i] < 1 || widths[i] > 1024 || s->ntrks % widths[i] != 0)
      return c->fail(P2R_ERR_SHAPE, "width %d does not divide ntrks=%d", widths[i], s->ntrks);
  DeviceGuard g(c->device);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t ntrk = (size_t)s->nevts * (size_t)s->ntrks;
  const size_t tb = ntrk * 28 * 4, hb = ntrk * (size_t)s->nlayer * 9 * 4;
  void *t = nullptr, *h = nullptr, *o = nullptr;
  if (hipMalloc(&t, tb) != hipSuccess || hipMalloc(&h, hb) != hipSuccess || hipMalloc(&o, tb) != hipSuccess) {
    if (t) (void)hipFree(t);
    if (h) (void)hipFree(h);
    return c->fail(P2R_ERR_NOMEM, "autotune_width copies (%zu bytes)", 2 * tb + hb);
  }
  Collect col;
  int rc = warm_up(c, st, [&] { return p2r_forward(c, trk_in, hits, o, s, semantics, params, nullptr, stream); });
  for (int i = 0; rc == P2R_OK && i < nwidths; ++i) {
    const int w = widths[i];
    rc = p2r_repack(c, trk_in, s->bsize, t, w, s->nevts, s->ntrks, 1, 28, stream);
    if (rc == P2R_OK) rc = p2r_repack(c, hits, s->bsize, h, w, s->nevts, s->ntrks, s->nlayer, 9, stream);
    const p2r_shape sw{s->nevts, s->ntrks, s->nlayer, w};
    float ms = 0.f;
    if (rc == P2R_OK)
      rc = time_calls(c, st, reps, [&] { return p2r_forward(c, t, h, o, &sw, semantics, params, nullptr, stream); },
                      &ms);
    if (rc == P2R_OK) col.add(entry_of(c->tune, w, ms, 0));
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(t);
  (void)hipFree(h);
  (void)hipFree(o);
  if (rc != P2R_OK) return rc;
  col.finish(report);
  return P2R_OK;
}

}  // extern "C"
