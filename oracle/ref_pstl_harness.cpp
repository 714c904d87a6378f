// Reference build harness -- TEST INFRASTRUCTURE ONLY (output: oracle/_ref/).
//
// Compiles the reference's own src/propagate-tor-test_pstl.cpp, unmodified and
// where it lies (oracle/Makefile adds -I<reference>/src, -Dmain=..., and the
// -Dbsize/-Dntrks/-Dnevts/-Dnlayer sizes), and exposes its input generator:
// randn (pstl.cpp's copy of tbb.cpp:127-145) and prepareTracks/prepareHits
// (same draw order and smear as tbb.cpp:243-293) fed with the reference's
// src/input_track.h in the reversed layer order of tbb.cpp:818-820.
#include "propagate-tor-test_pstl.cpp"

#include <cstring>

extern "C" {
int ref_pstl_bsize(void) { return bsize; }
int ref_pstl_ntrks(void) { return ntrks; }
int ref_pstl_nevts(void) { return nevts; }
int ref_pstl_nlayer(void) { return nlayer; }

// Fills trk_out (nevts*nb MPTRK blocks) and hits_out (nlayer*nevts*nb MPHIT
// blocks) from glibc rand() after srand(1) -- the reference's unseeded state.
int ref_pstl_generate(float* trk_out, float* hits_out) {
#include "input_track.h"
  std::vector<AHIT> inputhits{inputhit21, inputhit20, inputhit19, inputhit18, inputhit17, inputhit16,
                              inputhit15, inputhit14, inputhit13, inputhit12, inputhit11, inputhit10,
                              inputhit09, inputhit08, inputhit07, inputhit06, inputhit05, inputhit04,
                              inputhit03, inputhit02, inputhit01, inputhit00};
  srand(1);
  std::vector<MPTRK> trcks(nevts * nb);
  std::vector<MPHIT> hits(nlayer * nevts * nb);
  prepareTracks(trcks, inputtrk);
  prepareHits(hits, inputhits);
  std::memcpy(trk_out, trcks.data(), trcks.size() * sizeof(MPTRK));
  std::memcpy(hits_out, hits.data(), hits.size() * sizeof(MPHIT));
  return (int)(sizeof(MPTRK) / sizeof(float));
}
}
