// Reference CPU timing harness -- TEST INFRASTRUCTURE ONLY (output: oracle/_ref/).
//
// Compiles the reference's own src/propagate-tor-test_acc.v3.cpp, unmodified and where it lies, with
// the reference's CPU flags (-O3 -ffast-math, build.py:91; the recipe in oracle/Makefile adds
// -I<reference>/src, -fopenmp and renames its main), and runs its propagateToR / KalmanUpdate block
// routines in the driver loop of tbb.cpp:870-926: for every (event, block), copy the input block to the
// output, then for every layer propagate from the input block and update the output block (reference
// semantics, tbb.cpp:894).  OpenMP over blocks stands in for the TBB parallel_for of tbb.cpp:874-877,
// whose headers this image lacks.  KalmanUpdate's covariance is taken from its newErr scratch as
// tbb.cpp:641 does (acc.v3.cpp:634 assigns the pointer instead).  Used as the CPU baseline of the
// benchmark (kind "reference": the timed arithmetic is the reference's compiled code).
#include "propagate-tor-test_acc.v3.cpp"

#include <omp.h>

extern "C" {
int ref_acc_bench_bsize(void) { return bsize; }

// One repetition over nblk Matriplex blocks (MPTRK / MPHIT layout, hits [block][layer]).
int ref_acc_bench_run(const float* trk, const float* hits, float* out, long nblk, int nl, int nthreads) {
  if (nblk < 0 || nl < 1) return 1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
  for (long b = 0; b < nblk; ++b) {
    const MPTRK* bt = reinterpret_cast<const MPTRK*>(trk) + b;
    MPTRK* ob = reinterpret_cast<MPTRK*>(out) + b;
    const MPHIT* bh0 = reinterpret_cast<const MPHIT*>(hits) + b * nl;
    MP6x6F errorProp, temp;
    MP1F rotT00, rotT01;
    MP2x2SF resErr_loc;
    MP3x6 kGain;
    MP2F res_loc;
    MP6x6SF newErr;
    *ob = *bt;
    for (int l = 0; l < nl; ++l) {
      const MPHIT* bh = bh0 + l;
      propagateToR(&bt->cov, &bt->par, &bt->q, &bh->pos, &ob->cov, &ob->par, &errorProp, &temp);
      KalmanUpdate(&ob->cov, &ob->par, &bh->cov, &bh->pos, &rotT00, &rotT01, &resErr_loc, &kGain, &res_loc,
                   &newErr);
      ob->cov = newErr;
    }
  }
  return 0;
}
}
