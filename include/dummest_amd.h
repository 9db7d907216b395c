/*
 * dummest_amd.h -- C ABI of the MI355X (gfx950) implementation of DUMMEST's
 * frameshift-aware DNA-vs-profile-HMM forward/backward DP.
 *
 * Conventions: every buffer is owned by the caller; every call returns 0 on
 * success and a negative DM_E* code on failure (dm_engine_error() gives the
 * message); no exceptions cross the ABI; one engine = one HIP stream on one
 * device; calls on one engine are not thread-safe, separate engines are.
 *
 * Arithmetic: FP32 DP with power-of-two rescaling (the dummestl_core build of
 * the reference), ExpScore results as (mantissa in [0.5,1) or 0, exponent).
 */
#ifndef DUMMEST_AMD_H
#define DUMMEST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DM_OK 0
#define DM_EINVAL -1   /* bad argument */
#define DM_EHIP -2     /* HIP runtime error (allocation, launch, copy) */
#define DM_EFORMAT -3  /* unreadable profile or sequence data */
#define DM_EUNSUP -4   /* input outside what the kernels cover */

typedef struct dm_engine dm_engine;

/* Create an engine on HIP device `device` (its own stream and buffers). */
int dm_engine_create(int device, dm_engine **out);
void dm_engine_destroy(dm_engine *e);
/* Message of the last failing call on this engine ("" if none). */
const char *dm_engine_error(const dm_engine *e);

/* Frameshift / stop-codon rates: insert1 insert2 delete1 delete2 stop bg_stop.
 * Replaces the globals set by the engine CLI options (dummest-core.hh:115-131,
 * dummest.cc:227-261). Defaults are the reference's OPT_* values. */
int dm_set_rates(dm_engine *e, const double rates[6]);

/* Read HMMER3/f profiles and finalize them (readProfiles + finalizeProfile,
 * dummest-profile.hh:105-329; tantan masking off, i.e. -m 0). bg_type is 'G'
 * (geometric mean, default), 'A' (arithmetic) or 'M' (median). Profiles get
 * consecutive ids starting at *first_id. */
int dm_read_hmm(dm_engine *e, const char *path, int bg_type, int32_t *first_id, int32_t *count);

/* Parse HMMER3/f text and finalize it without a device (as dm_read_hmm);
 * writes profile number `index` (M, values (M+1)x29, params (M+2)x11, bg 29;
 * pass NULL arrays to query M) and the number of profiles in the text. */
int dm_parse_hmm_text(const char *text, int bg_type, const double rates[6], int32_t index, int32_t *M,
                      float *values, float *params, float *bg, int32_t *count);

/* Add one finalized profile: values (M+1)x29, params (M+2)x11 (alpha'[3]
 * beta'[3] delta'[3] eps' enter, Params of dummest-core.hh:233-239), bg 29
 * (bg_probs). Replaces Profile construction in readProfiles. */
int dm_add_profile(dm_engine *e, int32_t M, const float *values, const float *params, const float *bg,
                   int32_t *id);

/* Copy out a profile's finalized arrays (sizes as in dm_add_profile). */
int dm_get_profile(const dm_engine *e, int32_t id, int32_t *M, float *values, float *params, float *bg);
int dm_profile_count(const dm_engine *e);

/* Decode DNA letters into codon codes (decodeSequence, dummest-sequence.hh:181-204):
 * codes[j] is the amino-acid code of the codon starting at j; the last two are 0. */
int dm_decode(const char *dna, int64_t n, uint8_t *codes);

/* Append decoded sequences (codes concatenated; offsets/lengths per sequence).
 * Replaces the per-request decoded vectors (SequenceData, dummest-core.hh:277-283). */
int dm_add_sequences(dm_engine *e, int32_t nseq, const uint8_t *codes, const int64_t *offsets,
                     const int32_t *lengths, int32_t *first_id);
int dm_clear_sequences(dm_engine *e);

/* Score-only filter DP for (profile, sequence) pairs: best over cells of
 * W1 * prefixSeed in FP32 with no rescaling (+inf when it overflows).
 * Replaces findSimilaritiesBackwardOnly (dummest-dp-backward.hh:300-438). */
int dm_filter_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, float *best);

/* Forward+backward DP scores: best anchored score max wMid = w0*W1 and, if
 * want_bwd, the backward start score max W1*prefixSeed, both exact ExpScores.
 * Replaces the score part of findSimilarities with needAlign=false
 * (dummest-search.hh:9-17, dummest-dp-backward.hh:444-615,
 * dummest-dp-forward.hh:9-250) as used by estimateK (dummest-calibration.hh:496-497). */
int dm_fwdbwd_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, int want_bwd,
                     float *mid_m, int32_t *mid_e, float *bwd_m, int32_t *bwd_e);

/* DP matrices of one pair for post-gate processing: w0 (forward) and W1
 * (backward), (M+1) x n each, as ExpScores (row-major, row i = profile row).
 * The full W1 matrix is what backwardPass keeps in DPScratch::W1 (dummest-dp.hh:41). */
int dm_pair_matrices(dm_engine *e, int32_t prof, int32_t seq, float *w0_m, int32_t *w0_e, float *W1_m,
                     int32_t *W1_e);

/* Null model of one pair (buildBatchNullModels + buildNullSeedModels,
 * dummest-dp-backward.hh:6-188): prefixSeed[n], suffixSeed[n], scales[3] = s1 s2 sC. */
int dm_pair_null(dm_engine *e, int32_t prof, int32_t seq, float *prefix_seed, float *suffix_seed, float *scales);

/* Local hits of one pair after the gate (forwardPass epilogue + traceback,
 * dummest-dp-forward.hh:252-545, dummest-traceback.hh:6-118). min_prob_ratio as
 * in SequenceRequest (dummest-core.hh:285-288). Writes up to hit_cap hits as
 * (lnratio, mantissa, exponent, anchor1, anchor2, nseg) and their segments as
 * (start1, start2, length); *nhits / *nsegs receive the counts. */
int dm_pair_hits(dm_engine *e, int32_t prof, int32_t seq, double min_prob_ratio, double *hits, int32_t hit_cap,
                 int32_t *segs, int32_t seg_cap, int32_t *nhits, int32_t *nsegs);

/* Options: "plain" = 1 launches the DP once per pair (the unbatched form);
 * "grid" = persistent grid size (0 = automatic). */
int dm_set_option(dm_engine *e, const char *name, int64_t value);

/* Device time of the DP kernels of the last scoring call (HIP events on the
 * engine's stream), in milliseconds, and the number of DP cells it covered
 * (sum over pairs of (M+1)*n). */
double dm_last_dp_ms(const dm_engine *e);
int64_t dm_last_cells(const dm_engine *e);

#ifdef __cplusplus
}
#endif
#endif /* DUMMEST_AMD_H */
