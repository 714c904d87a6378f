// Reference build harness -- TEST INFRASTRUCTURE ONLY (output: oracle/_ref/).
//
// Compiles the reference's own src/propagate-tor-test_acc.v3.cpp, unmodified
// and where it lies (the recipe in oracle/Makefile adds -I<reference>/src and
// renames its main with -Dmain=...; g++ ignores its `#pragma acc` lines), and
// exposes its block-level routines through extern "C" wrappers so the tests can
// compare the oracle against the reference's compiled code.  The routines of
// that file (propagateToR, KalmanUpdate, MultHelixProp[Transp], sincos4, hipo)
// are arithmetically identical to those of propagate-tor-test_tbb.cpp; only its
// kfact constant differs (100/3.8, acc.v3.cpp:647, vs tbb.cpp:654), which the
// wrappers report so the oracle can be run with the same value.
#include "propagate-tor-test_acc.v3.cpp"

extern "C" {
int ref_acc_bsize(void) { return bsize; }
float ref_acc_kfact(void) { return kfact; }

// One Matriplex block: pointers to the MP6x6SF / MP6F / MP1I / MP3F arrays.
void ref_acc_propagateToR(const float* inErr, const float* inPar, const int* inChg, const float* msP, float* outErr,
                          float* outPar) {
  MP6x6F errorProp, temp;
  propagateToR(reinterpret_cast<const MP6x6SF*>(inErr), reinterpret_cast<const MP6F*>(inPar),
               reinterpret_cast<const MP1I*>(inChg), reinterpret_cast<const MP3F*>(msP),
               reinterpret_cast<MP6x6SF*>(outErr), reinterpret_cast<MP6F*>(outPar), &errorProp, &temp);
}

void ref_acc_KalmanUpdate(float* trkErr, float* inPar, const float* hitErr, const float* msP) {
  MP1F rotT00, rotT01;
  MP2x2SF resErr_loc;
  MP3x6 kGain;
  MP2F res_loc;
  MP6x6SF newErr;
  KalmanUpdate(reinterpret_cast<MP6x6SF*>(trkErr), reinterpret_cast<MP6F*>(inPar),
               reinterpret_cast<const MP3x3SF*>(hitErr), reinterpret_cast<const MP3F*>(msP), &rotT00, &rotT01,
               &resErr_loc, &kGain, &res_loc, &newErr);
  // acc.v3.cpp:634 ends KalmanUpdate with `trkErr = newErr;`, which assigns
  // the local pointer and so never hands the updated covariance back; tbb.cpp
  // copies it (`(*trkErr) = newErr;`, tbb.cpp:641).  Copy the reference's own
  // result out of its newErr scratch to get tbb.cpp's behaviour.
  *reinterpret_cast<MP6x6SF*>(trkErr) = newErr;
}

// MultHelixProp / MultHelixPropTransp on one block (A, B full 6x6; packed sym).
void ref_acc_MultHelixProp(const float* A, const float* B, float* C) {
  MultHelixProp(reinterpret_cast<const MP6x6F*>(A), reinterpret_cast<const MP6x6SF*>(B), reinterpret_cast<MP6x6F*>(C));
}
void ref_acc_MultHelixPropTransp(const float* A, const float* B, float* C) {
  MultHelixPropTransp(reinterpret_cast<const MP6x6F*>(A), reinterpret_cast<const MP6x6F*>(B),
                      reinterpret_cast<MP6x6SF*>(C));
}
}
