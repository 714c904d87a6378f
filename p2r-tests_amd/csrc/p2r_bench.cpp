// p2r_bench: command-line driver for the MI355X p2r path.
//
// Mirrors the reference program's driver (cerati/p2r-tests
// src/propagate-tor-test_tbb.cpp:814-1055 and the GPU variants'
// propagate-tor-test_hip.cu:851-1147): generate inputs, copy them to the GPU,
// run NITER timed repetitions of the whole batch, copy back, print the
// reference's stdout protocol -- the `done ntracks=... tot time=... time/trk=...`
// line parsed by the reference's build.py:83, the `formatted ...` line and the
// summary of tbb.cpp:1036-1048.  The compile-time macros of the reference
// (ntrks, nevts, nlayer, bsize, NITER, NWARMUP, num_streams) are flags here.
//
//   p2r_bench [--ntrks N] [--nevts N] [--nlayer N] [--bsize N] [--NITER N]
//             [--NWARMUP N] [--semantics reference|chained]
//             [--input reference|baseline] [--transfers none|h2d|d2h|both]
//             [--include-transfers (= both)] [--num-streams N] [--threads 64|128|256]
//             [--graph] [--device D]
//
// --graph captures the kernel-only repetition once (p2r_graph_forward) and
// replays the graph NITER times; with transfers it is ignored.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/p2r.h"
#include "../../include/p2r_gen.h"

#define CHECK_HIP(x)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) {                                                                       \
      std::fprintf(stderr, "\n%s, line %d:\nhipCheck(%s);\n%s: %s\n", __FILE__, __LINE__, #x,     \
                   hipGetErrorName(e_), hipGetErrorString(e_));                                   \
      std::exit(2);                                                                               \
    }                                                                                             \
  } while (0)
#define CHECK_P2R(ctx, x)                                                                         \
  do {                                                                                            \
    int s_ = (x);                                                                                 \
    if (s_ != P2R_OK) {                                                                           \
      std::fprintf(stderr, "%s failed: %s (%s)\n", #x, p2r_error_string(s_), p2r_last_error(ctx)); \
      std::exit(3);                                                                               \
    }                                                                                             \
  } while (0)

namespace {

struct Opts {
  int ntrks = 9600, nevts = 100, nlayer = 20, bsize = 32, niter = 5, nwarmup = 2, device = 0, nstreams = 1;
  int semantics = P2R_REFERENCE;
  int threads = 0;  // workgroup size of the forward kernel (0 = library default)
  bool graph = false;  // replay one captured launch (kernel-only timing)
  bool reference_input = true, h2d = false, d2h = false;
};

Opts parse(int argc, char** argv) {
  Opts o;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "--ntrks") o.ntrks = std::atoi(next());
    else if (a == "--nevts") o.nevts = std::atoi(next());
    else if (a == "--nlayer") o.nlayer = std::atoi(next());
    else if (a == "--bsize") o.bsize = std::atoi(next());
    else if (a == "--NITER") o.niter = std::atoi(next());
    else if (a == "--NWARMUP") o.nwarmup = std::atoi(next());
    else if (a == "--device") o.device = std::atoi(next());
    else if (a == "--semantics") o.semantics = std::string(next()) == "chained" ? P2R_CHAINED : P2R_REFERENCE;
    else if (a == "--input") o.reference_input = std::string(next()) != "baseline";
    else if (a == "--include-transfers") o.h2d = o.d2h = true;
    else if (a == "--transfers") {
      const std::string v = next();
      o.h2d = v == "both" || v == "h2d";
      o.d2h = v == "both" || v == "d2h";
      if (v != "both" && v != "h2d" && v != "d2h" && v != "none") {
        std::fprintf(stderr, "--transfers takes none, h2d, d2h or both\n");
        std::exit(2);
      }
    } else if (a == "--num-streams") o.nstreams = std::atoi(next());
    else if (a == "--threads") o.threads = std::atoi(next());
    else if (a == "--graph") o.graph = true;
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      std::exit(1);
    }
  }
  return o;
}

// The reference's validation summary (tbb.cpp:935-1048), same formats and
// float accumulation, against the unsmeared hit of the last layer.
void print_summary(const std::vector<float>& out, int nevts, int ntrks, int bs, const float h[3]) {
  const int nb = ntrks / bs;
  auto fld = [&](long ie, long tk, int f) { return out[(size_t)((tk / bs) + (long)nb * ie) * 28 * bs + f * bs + tk % bs]; };
  const float hx_ = h[0], hy_ = h[1], hz_ = h[2], hr_ = sqrtf(hx_ * hx_ + hy_ * hy_);
  int nnans = 0, nfail = 0;
  float avgx = 0, avgy = 0, avgz = 0, avgr = 0, avgpt = 0, avgphi = 0, avgtheta = 0;
  float avgdx = 0, avgdy = 0, avgdz = 0, avgdr = 0;
  auto fails = [&](float x_, float y_, float z_) {
    return std::fabs((x_ - hx_) / hx_) > 1. || std::fabs((y_ - hy_) / hy_) > 1. || std::fabs((z_ - hz_) / hz_) > 1.;
  };
  for (long ie = 0; ie < nevts; ++ie)
    for (long it = 0; it < ntrks; ++it) {
      const float x_ = fld(ie, it, 0), y_ = fld(ie, it, 1), z_ = fld(ie, it, 2), r_ = sqrtf(x_ * x_ + y_ * y_);
      const float pt_ = std::abs(1. / fld(ie, it, 3)), phi_ = fld(ie, it, 4), theta_ = fld(ie, it, 5);
      if (!std::isfinite(x_) || !std::isfinite(y_) || !std::isfinite(z_) || !std::isfinite(pt_) ||
          !std::isfinite(phi_) || !std::isfinite(theta_)) {
        nnans++;
        continue;
      }
      if (fails(x_, y_, z_)) {
        nfail++;
        continue;
      }
      avgpt += pt_; avgphi += phi_; avgtheta += theta_;
      avgx += x_; avgy += y_; avgz += z_; avgr += r_;
      avgdx += (x_ - hx_) / x_; avgdy += (y_ - hy_) / y_; avgdz += (z_ - hz_) / z_; avgdr += (r_ - hr_) / r_;
    }
  const float n = float(nevts * ntrks);
  avgpt /= n; avgphi /= n; avgtheta /= n; avgx /= n; avgy /= n; avgz /= n; avgr /= n;
  avgdx /= n; avgdy /= n; avgdz /= n; avgdr /= n;
  float stdx = 0, stdy = 0, stdz = 0, stdr = 0, stddx = 0, stddy = 0, stddz = 0, stddr = 0;
  for (long ie = 0; ie < nevts; ++ie)
    for (long it = 0; it < ntrks; ++it) {
      const float x_ = fld(ie, it, 0), y_ = fld(ie, it, 1), z_ = fld(ie, it, 2), r_ = sqrtf(x_ * x_ + y_ * y_);
      if (!std::isfinite(x_) || !std::isfinite(y_) || !std::isfinite(z_) || fails(x_, y_, z_)) continue;
      stdx += (x_ - avgx) * (x_ - avgx); stdy += (y_ - avgy) * (y_ - avgy);
      stdz += (z_ - avgz) * (z_ - avgz); stdr += (r_ - avgr) * (r_ - avgr);
      stddx += ((x_ - hx_) / x_ - avgdx) * ((x_ - hx_) / x_ - avgdx);
      stddy += ((y_ - hy_) / y_ - avgdy) * ((y_ - hy_) / y_ - avgdy);
      stddz += ((z_ - hz_) / z_ - avgdz) * ((z_ - hz_) / z_ - avgdz);
      stddr += ((r_ - hr_) / r_ - avgdr) * ((r_ - hr_) / r_ - avgdr);
    }
  stdx = sqrtf(stdx / n); stdy = sqrtf(stdy / n); stdz = sqrtf(stdz / n); stdr = sqrtf(stdr / n);
  stddx = sqrtf(stddx / n); stddy = sqrtf(stddy / n); stddz = sqrtf(stddz / n); stddr = sqrtf(stddr / n);
  std::printf("track x avg=%f std/avg=%f\n", avgx, std::fabs(stdx / avgx));
  std::printf("track y avg=%f std/avg=%f\n", avgy, std::fabs(stdy / avgy));
  std::printf("track z avg=%f std/avg=%f\n", avgz, std::fabs(stdz / avgz));
  std::printf("track r avg=%f std/avg=%f\n", avgr, std::fabs(stdr / avgz));  // as tbb.cpp:1039
  std::printf("track dx/x avg=%f std=%f\n", avgdx, stddx);
  std::printf("track dy/y avg=%f std=%f\n", avgdy, stddy);
  std::printf("track dz/z avg=%f std=%f\n", avgdz, stddz);
  std::printf("track dr/r avg=%f std=%f\n", avgdr, stddr);
  std::printf("track pt avg=%f\n", avgpt);
  std::printf("track phi avg=%f\n", avgphi);
  std::printf("track theta avg=%f\n", avgtheta);
  std::printf("number of tracks with nans=%i\n", nnans);
  std::printf("number of tracks failed=%i\n", nfail);
}

}  // namespace

int main(int argc, char** argv) {
  const Opts o = parse(argc, argv);
  if (o.bsize < 1 || o.ntrks % o.bsize != 0) {
    std::fprintf(stderr, "ntrks must be a multiple of bsize\n");
    return 1;
  }
  const int nb = o.ntrks / o.bsize;
  const p2r_shape shape{o.nevts, o.ntrks, o.nlayer, o.bsize};
  const size_t trk_floats = (size_t)o.nevts * nb * 28 * o.bsize;
  const size_t hit_floats = (size_t)o.nlayer * o.nevts * nb * 9 * o.bsize;
  std::printf("produce nevts=%i ntrks=%i nlayer=%i bsize=%i input=%s semantics=%s\n", o.nevts, o.ntrks, o.nlayer,
              o.bsize, o.reference_input ? "reference" : "baseline",
              o.semantics == P2R_CHAINED ? "chained" : "reference");
  std::printf("NITER=%d\n", o.niter);

  const auto setup_start = std::chrono::steady_clock::now();
  float *h_trk = nullptr, *h_hit = nullptr, *h_out = nullptr;
  CHECK_HIP(hipHostMalloc((void**)&h_trk, trk_floats * 4));
  CHECK_HIP(hipHostMalloc((void**)&h_hit, hit_floats * 4));
  CHECK_HIP(hipHostMalloc((void**)&h_out, trk_floats * 4));
  float hit_ref[3] = {0, 0, 0};
  if (o.reference_input) {
    if (p2r_gen_reference(h_trk, h_hit, &shape) != P2R_OK) {
      std::fprintf(stderr, "reference inputs need nlayer <= 22\n");
      return 1;
    }
    p2r_reference_hit(o.nlayer - 1, hit_ref);
  } else {
    const p2r_gen_config cfg = p2r_gen_default_config();
    if (p2r_gen_uniform(h_trk, h_hit, &shape, &cfg) != P2R_OK) return 1;
    // summary reference point: the base track's crossing of the last layer
    // (hit of any track with zero smear); take track 0's hit as the reference
    for (int f = 0; f < 3; ++f) hit_ref[f] = h_hit[(size_t)(o.nlayer - 1) * 9 * o.bsize + f * o.bsize];
  }
  CHECK_HIP(hipSetDevice(o.device));
  p2r_ctx* ctx = nullptr;
  if (p2r_create(o.device, &ctx) != P2R_OK) {
    std::fprintf(stderr, "p2r_create failed\n");
    return 1;
  }
  hipDeviceProp_t prop;
  CHECK_HIP(hipGetDeviceProperties(&prop, o.device));
  std::printf(" Device Number: %d\n Device name: %s\n", o.device, prop.name);
  hipStream_t st;
  CHECK_HIP(hipStreamCreate(&st));
  float *d_trk, *d_hit, *d_out;
  CHECK_HIP(hipMalloc((void**)&d_trk, trk_floats * 4));
  CHECK_HIP(hipMalloc((void**)&d_hit, hit_floats * 4));
  CHECK_HIP(hipMalloc((void**)&d_out, trk_floats * 4));
  CHECK_HIP(hipMemcpyAsync(d_trk, h_trk, trk_floats * 4, hipMemcpyHostToDevice, st));
  CHECK_HIP(hipMemcpyAsync(d_hit, h_hit, hit_floats * 4, hipMemcpyHostToDevice, st));
  CHECK_HIP(hipStreamSynchronize(st));
  const double setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - setup_start).count();
  std::printf("done preparing!\n");
  std::printf("Size of struct MPTRK trk[] = %ld\n", (long)(trk_floats * 4));
  std::printf("Size of struct struct MPHIT hit[] = %ld\n", (long)(hit_floats * 4));

  const p2r_params params = p2r_default_params();
  if (o.threads) {
    p2r_tuning t;
    CHECK_P2R(ctx, p2r_get_tuning(ctx, &t));
    t.fwd_threads = o.threads;
    CHECK_P2R(ctx, p2r_set_tuning(ctx, &t));
  }
  // Timed region as the reference's MEASURE_H2D_TRANSFER / MEASURE_D2H_TRANSFER (hip.cu:17-22,958-1006):
  // both: the streamed form, chunks of blocks over num_streams streams; one of them: that copy on the
  // stream around the kernel (the other happens outside the timed region); none: the kernel only.
  p2r_graph* graph = nullptr;
  if (o.graph && !o.h2d && !o.d2h)
    CHECK_P2R(ctx, p2r_graph_forward(ctx, d_trk, d_hit, d_out, &shape, o.semantics, &params, nullptr, &graph));
  auto one = [&]() {
    if (graph) {
      if (p2r_graph_launch(graph, st) != P2R_OK) {
        std::fprintf(stderr, "p2r_graph_launch failed\n");
        std::exit(1);
      }
      return;
    }
    if (o.h2d && o.d2h) {
      CHECK_P2R(ctx, p2r_forward_streamed(ctx, h_trk, h_hit, h_out, d_trk, d_hit, d_out, &shape, o.semantics, &params,
                                          o.nstreams));
      return;
    }
    if (o.h2d) {
      CHECK_HIP(hipMemcpyAsync(d_trk, h_trk, trk_floats * 4, hipMemcpyHostToDevice, st));
      CHECK_HIP(hipMemcpyAsync(d_hit, h_hit, hit_floats * 4, hipMemcpyHostToDevice, st));
    }
    CHECK_P2R(ctx, p2r_forward(ctx, d_trk, d_hit, d_out, &shape, o.semantics, &params, nullptr, st));
    if (o.d2h) CHECK_HIP(hipMemcpyAsync(h_out, d_out, trk_floats * 4, hipMemcpyDeviceToHost, st));
  };
  for (int i = 0; i < o.nwarmup; ++i) one();
  CHECK_HIP(hipStreamSynchronize(st));
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < o.niter; ++i) one();
  CHECK_HIP(hipStreamSynchronize(st));
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  CHECK_HIP(hipMemcpy(h_out, d_out, trk_floats * 4, hipMemcpyDeviceToHost));

  const long ntot = (long)o.nevts * o.ntrks * o.niter;
  std::printf("setup time time=%f (s)\n", setup_s);
  std::printf("done ntracks=%i tot time=%f (s) time/trk=%e (s)\n", (int)ntot, wall, wall / ntot);
  std::printf("formatted %i %i %i %i %i %f 0 %f %i\n", o.niter, o.nevts, o.ntrks, o.bsize, nb, wall, setup_s, -1);
  std::printf("track-layer updates/s=%e\n", (double)ntot * o.nlayer / wall);
  print_summary(std::vector<float>(h_out, h_out + trk_floats), o.nevts, o.ntrks, o.bsize, hit_ref);

  if (graph) p2r_graph_destroy(graph);
  p2r_destroy(ctx);
  CHECK_HIP(hipFree(d_trk));
  CHECK_HIP(hipFree(d_hit));
  CHECK_HIP(hipFree(d_out));
  CHECK_HIP(hipHostFree(h_trk));
  CHECK_HIP(hipHostFree(h_hit));
  CHECK_HIP(hipHostFree(h_out));
  CHECK_HIP(hipStreamDestroy(st));
  return 0;
}
