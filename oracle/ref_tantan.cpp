// Test infrastructure, not product code: exposes the reference's own tantan
// (tantan-wrapper.hh and tantan/tantan.cc, compiled from where they lie in the
// reference tree by oracle/Makefile's `ref` target) through one C entry point,
// so that tests can pin the oracle's restatement against it.
#include "tantan-wrapper.hh"

extern "C" void ref_tantan_probs(const unsigned char *seq, int n, int is_protein, float *out) {
    calcTantanProbabilities(seq, n, is_protein != 0, out);
}
