// Thin PyTorch-ROCm extension over the C-ABI of include/p2r.h.
//
// Tensors in, tensors out, on torch's current HIP stream; one p2r_ctx per
// device, created on first use.  All arithmetic happens in libp2r_amd.so: this
// file only checks tensor shapes / devices / dtypes, allocates outputs through
// torch's caching allocator and forwards device pointers.  Used by
// p2r_amd.autograd (the torch.autograd.Function of the chained fit).
#include <torch/extension.h>
// ROCm builds of torch give device tensors the device type "cuda": the guard and
// stream accessors are the HIP implementations registered under that type.
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>

#include <mutex>
#include <vector>

#include "../../include/p2r.h"

namespace {

p2r_ctx* ctx_for(int device) {
  static std::mutex mu;
  static std::vector<p2r_ctx*> ctxs;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)ctxs.size() <= device) ctxs.resize(device + 1, nullptr);
  if (!ctxs[device]) {
    p2r_ctx* c = nullptr;
    const int st = p2r_create(device, &c);
    TORCH_CHECK(st == P2R_OK, "p2r_create(", device, ") failed: ", p2r_error_string(st));
    ctxs[device] = c;  // lives until the process ends
  }
  return ctxs[device];
}

void check(int st, p2r_ctx* c, const char* what) {
  TORCH_CHECK(st == P2R_OK, what, ": ", p2r_error_string(st), ": ", p2r_last_error(c));
}

p2r_shape make_shape(int64_t nevts, int64_t ntrks, int64_t nlayer, int64_t bsize) {
  TORCH_CHECK(nevts >= 0 && ntrks >= 0 && nlayer >= 1 && bsize >= 1 && nevts < (1 << 30) && ntrks < (1 << 30) &&
                  nlayer <= P2R_MAX_NLAYER && bsize <= 1024,
              "bad shape");
  return p2r_shape{(int32_t)nevts, (int32_t)ntrks, (int32_t)nlayer, (int32_t)bsize};
}

void check_buffers(const at::Tensor& trk, const at::Tensor& hits, const p2r_shape& s) {
  TORCH_CHECK(trk.is_cuda() && hits.is_cuda(), "trk and hits must be device tensors");
  TORCH_CHECK(trk.device() == hits.device(), "trk and hits must be on the same device");
  TORCH_CHECK(trk.scalar_type() == at::kFloat && hits.scalar_type() == at::kFloat, "trk and hits must be float32");
  TORCH_CHECK(trk.is_contiguous() && hits.is_contiguous(), "trk and hits must be contiguous Matriplex buffers");
  TORCH_CHECK(s.ntrks % s.bsize == 0, "ntrks must be a multiple of bsize");
  const int64_t nblk = (int64_t)s.nevts * (s.ntrks / s.bsize);
  TORCH_CHECK(trk.numel() == nblk * 28 * s.bsize, "trk must hold nevts*nb MPTRK blocks (", nblk * 28 * s.bsize,
              " words), got ", trk.numel());
  TORCH_CHECK(hits.numel() == nblk * s.nlayer * 9 * s.bsize, "hits must hold nlayer*nevts*nb MPHIT blocks (",
              nblk * s.nlayer * 9 * s.bsize, " words), got ", hits.numel());
}

p2r_params make_params(double kfact, int64_t precision) {
  p2r_params p = p2r_default_params();
  p.kfact = (float)kfact;
  p.b_scale = 1.f;
  p.precision = (int32_t)precision;
  return p;
}

// Whole batch through all layers (p2r_forward).  Returns {trk_out, chi2 (per track; empty if !with_chi2)}.
std::vector<at::Tensor> forward(const at::Tensor& trk, const at::Tensor& hits, int64_t nevts, int64_t ntrks,
                                int64_t nlayer, int64_t bsize, int64_t semantics, double kfact, int64_t precision,
                                bool with_chi2) {
  const p2r_shape s = make_shape(nevts, ntrks, nlayer, bsize);
  check_buffers(trk, hits, s);
  const c10::hip::HIPGuardMasqueradingAsCUDA guard(trk.device());
  p2r_ctx* c = ctx_for(trk.device().index());
  at::Tensor out = at::empty_like(trk);
  at::Tensor chi2 = with_chi2 ? at::empty({(int64_t)s.nevts * s.ntrks}, trk.options()) : at::Tensor();
  const p2r_params p = make_params(kfact, precision);
  hipStream_t st = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(trk.device().index()).stream();
  check(p2r_forward(c, trk.data_ptr(), hits.data_ptr(), out.data_ptr(), &s, (int)semantics, &p,
                    with_chi2 ? chi2.data_ptr<float>() : nullptr, st),
        c, "p2r_forward");
  return {out, chi2};
}

// Chained fit + reverse-mode gradient of sum chi2 (p2r_forward_adjoint).
// Returns {trk_out, chi2, grad_par (MP6F blocks), grad_hits (MP3F per hit block), grad_kfact (per track),
//          grad_global (float64 [4*nlayer+1]; empty if !with_global)}.
std::vector<at::Tensor> forward_adjoint(const at::Tensor& trk, const at::Tensor& hits, int64_t nevts, int64_t ntrks,
                                        int64_t nlayer, int64_t bsize, double kfact, int64_t precision,
                                        int64_t grad_precision, int64_t checkpoint_stride, bool with_global) {
  const p2r_shape s = make_shape(nevts, ntrks, nlayer, bsize);
  check_buffers(trk, hits, s);
  const c10::hip::HIPGuardMasqueradingAsCUDA guard(trk.device());
  p2r_ctx* c = ctx_for(trk.device().index());
  const int64_t nblk = (int64_t)s.nevts * (s.ntrks / s.bsize);
  at::Tensor out = at::empty_like(trk);
  at::Tensor chi2 = at::empty({nblk * s.bsize}, trk.options());
  at::Tensor gpar = at::empty({nblk * 6 * s.bsize}, trk.options());
  at::Tensor ghit = at::empty({nblk * s.nlayer * 3 * s.bsize}, trk.options());
  at::Tensor gk = at::empty({nblk * s.bsize}, trk.options());
  at::Tensor gg = with_global ? at::empty({4 * (int64_t)s.nlayer + 1}, trk.options().dtype(at::kDouble))
                              : at::Tensor();
  const p2r_params p = make_params(kfact, precision);
  hipStream_t st = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(trk.device().index()).stream();
  check(p2r_forward_adjoint(c, trk.data_ptr(), hits.data_ptr(), out.data_ptr(), &s, &p, (int)grad_precision,
                            (int32_t)checkpoint_stride, chi2.data_ptr<float>(), gpar.data_ptr<float>(),
                            ghit.data_ptr<float>(), gk.data_ptr<float>(),
                            with_global ? gg.data_ptr<double>() : nullptr, st),
        c, "p2r_forward_adjoint");
  return {out, chi2, gpar, ghit, gk, gg};
}

double default_kfact() { return (double)p2r_default_params().kfact; }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "p2r on MI355X: torch tensors through the C-ABI of libp2r_amd.so";
  m.def("forward", &forward, "p2r_forward on torch tensors (current HIP stream)");
  m.def("forward_adjoint", &forward_adjoint, "p2r_forward_adjoint on torch tensors (current HIP stream)");
  m.def("default_kfact", &default_kfact, "kfact of the reference (tbb.cpp:654) as float");
}
