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
#define DM_EIO -5      /* a file could not be written */
#define DM_ERANGE -6   /* output buffers too small; the call reports the sizes it needs */

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
 * dummest-profile.hh:105-329; profile masking per dm_set_masking). bg_type is 'G'
 * (geometric mean, default), 'A' (arithmetic) or 'M' (median). Profiles get
 * consecutive ids starting at *first_id. */
int dm_read_hmm(dm_engine *e, const char *path, int bg_type, int32_t *first_id, int32_t *count);

/* Parse HMMER3/f text and finalize it without a device (as dm_read_hmm);
 * writes profile number `index` (M, values (M+1)x29, params (M+2)x11, bg 29;
 * pass NULL arrays to query M), the number of profiles in the text and the
 * positions masked. mask_threshold >= 0 turns on profile masking (-m 1) at that
 * tantan threshold; negative leaves it off. filter_std_dev > 0 applies the
 * Gaussian letter filter (as dm_set_profile_filter) instead of masking. */
int dm_parse_hmm_text(const char *text, int bg_type, const double rates[6], double mask_threshold,
                      double filter_std_dev, int keep_nonvarying, int32_t index, int32_t *M, float *values,
                      float *params, float *bg, int32_t *count, int32_t *masked);

/* Add profiles from values as the reader stores them (probabilities, (M+1)x29
 * per profile, concatenated: transitions m->i i->i m->d d->d then node i+1's
 * 20 emissions, dummest-profile.hh:264-314) and finalize them. */
int dm_add_raw_profiles(dm_engine *e, int32_t nprof, const int32_t *M, const float *raw, int bg_type,
                        int32_t *first_id);

/* Add one finalized profile: values (M+1)x29, params (M+2)x11 (alpha'[3]
 * beta'[3] delta'[3] eps' enter, Params of dummest-core.hh:233-239), bg 29
 * (bg_probs). Replaces Profile construction in readProfiles. */
int dm_add_profile(dm_engine *e, int32_t M, const float *values, const float *params, const float *bg,
                   int32_t *id);

/* Replace profile `id` with new finalized arrays (as dm_add_profile; the
 * length may change; with the same length and the device tables in sync only
 * this profile's tables are rewritten). For callers that re-evaluate a profile
 * many times, such as a training loop over dm_grad; no counterpart in the
 * reference. */
int dm_set_profile(dm_engine *e, int32_t id, int32_t M, const float *values, const float *params, const float *bg);

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
/* Empty the sequence store (ids restart at 0). Plans built on the old
 * sequences and registered gradient sets (dm_grad_set_create) become invalid:
 * executing such a plan or evaluating such a set returns DM_EINVAL. */
int dm_clear_sequences(dm_engine *e);

/* Masking of simple regions with tantan (dummest.cc -m and --tantan-threshold,
 * :181-184, :257-261). mask bit 1: profiles read or added afterwards are masked
 * (finalizeProfile's isMask, dummest-profile.hh:105-233); bit 2: DNA added with
 * dm_add_dna is masked (makeMaskedSequence, dummest-sequence.hh:206-236).
 * threshold in [0, 1] (default 0.5). */
int dm_set_masking(dm_engine *e, int mask, double threshold);
/* Gaussian letter filter of the profiles read or added afterwards (-d and
 * -D, dummest.cc:185-195; filterLetterProbabilities, dummest-profile.hh:64-103,
 * applied by finalizeProfile at :119-121): each letter's log probability less
 * the position's mean log probability, minus its Gaussian smoothing along the
 * profile (std_dev positions, circular), renormalized per position;
 * keep_nonvarying (-D) keeps the mean level. std_dev in [2, 1000], or 0 to turn
 * it off. While it is on, profile masking (mask bit 1) is not applied, as the
 * reference turns it off (dummest.cc:284-285). */
int dm_set_profile_filter(dm_engine *e, double std_dev, int keep_nonvarying);
/* Profile positions masked by tantan ("# Positions masked by tantan", dummest.cc:389-395);
 * -1 for an invalid id. */
int dm_profile_masked(const dm_engine *e, int32_t id);
/* tantan repeat probabilities of DNA letters on the device (calcTantanProbabilities
 * with isProtein = false, tantan-wrapper.hh:13-64; A,C,G,T in either case, every
 * other letter counts as A): probs[offsets[k] + j] for letter j of sequence k. */
int dm_tantan_dna(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
                  float *probs);
/* Append DNA sequences given as letters: masked on the device when mask bit 2
 * is set, then decoded (replaces makeMaskedSequence + decodeSequence per
 * contig, dummest.cc:557-573). Ids as dm_add_sequences. */
int dm_add_dna(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
               int32_t *first_id);

/* Score-only filter DP for (profile, sequence) pairs: best over cells of
 * W1 * prefixSeed in FP32 with no rescaling (+inf when it overflows).
 * Replaces findSimilaritiesBackwardOnly (dummest-dp-backward.hh:300-438). */
int dm_filter_scores(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, float *best);

/* dm_filter_scores for one profile against nseq sequences, allocating no
 * device memory. No counterpart in the reference: for callers that cannot
 * call hipMalloc (serving loops, harnesses that forbid allocation).
 * codes: device pointer to the sequences' decoded codes (as dm_decode writes
 * them), sequence s at codes + offsets[s] with lengths[s] >= 3 letters
 * (offsets and lengths are host arrays); best: device pointer to nseq floats,
 * filled in the caller's order with dm_filter_scores' values. The caller's
 * buffers must be ready when the call starts; it returns after the results
 * are written. Everything else the kernels use lives in one static device
 * array of 64 MiB per GPU (about 20 bytes per letter plus 84 per sequence,
 * so roughly 3 million letters per call; larger inputs return DM_EUNSUP).
 * The profile tables must already be on the device (any earlier scoring
 * call, plan or in-place dm_set_profile puts them there; DM_EINVAL
 * otherwise). Calls on one GPU share the array: calls that overlap on one
 * GPU (several engines, several threads) race and must be serialized by the
 * caller. */
int dm_filter_scores_static(dm_engine *e, int32_t prof, int32_t nseq, const uint8_t *codes, const int64_t *offsets,
                            const int32_t *lengths, float *best);

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
 * in SequenceRequest (dummest-core.hh:285-288). Writes the hits as
 * (lnratio, mantissa, exponent, anchor1, anchor2, nseg) and their segments as
 * (start1, start2, length); *nhits / *nsegs receive the counts. When more than
 * hit_cap hits or seg_cap segments are found, nothing is written, *nhits /
 * *nsegs receive the sizes needed and the call returns DM_ERANGE; the same
 * call with larger buffers then returns the kept result without running the
 * pair again (the reference has no per-pair hit limit). */
int dm_pair_hits(dm_engine *e, int32_t prof, int32_t seq, double min_prob_ratio, double *hits, int32_t hit_cap,
                 int32_t *segs, int32_t seg_cap, int32_t *nhits, int32_t *nsegs);

/* The hits of many pairs in one call, as the reference forms them for every
 * lane of a lane batch after its forward pass (findFinalSimilaritiesBatched,
 * dummest-search.hh:76-285, with dummest-dp-forward.hh:252-545): every pair's
 * DP matrices, posterior match matrix, flank totals, MEA tables and
 * tracebacks on the device, batched, with no per-pair synchronization; the
 * same hits as dm_pair_hits pair by pair. Pair k's hits are rows
 * first_hit[k] .. first_hit[k+1]-1 of hits (first_hit: npairs + 1 entries;
 * rows as in dm_pair_hits), their segments consecutive in segs in hit order.
 * *nhits / *nsegs receive the totals; DM_ERANGE when they pass hit_cap /
 * seg_cap (nothing written; the same call with larger buffers returns the
 * kept result). dm_last_hits_ms: device time of the last run. */
int dm_pairs_hits(dm_engine *e, int64_t npairs, const int32_t *prof, const int32_t *seq, const double *min_prob_ratio,
                  double *hits, int64_t hit_cap, int32_t *segs, int64_t seg_cap, int64_t *first_hit, int64_t *nhits,
                  int64_t *nsegs);
double dm_last_hits_ms(const dm_engine *e);

/* Plans: the work queue of a pair list (configuration choice, length sort,
 * units, device-resident pair descriptors) built once, then executed any
 * number of times with only kernel launches on the device. mode 0 = filter
 * scores, 1 = forward+backward scores. Replaces the per-call job list of
 * findFinalSimilaritiesBatched (dummest-search.hh:76-285). */
typedef struct dm_plan dm_plan;
int dm_plan_create(dm_engine *e, int mode, int64_t npairs, const int32_t *prof, const int32_t *seq, int want_bwd,
                   dm_plan **out);
/* mode 0 fills best[npairs]; mode 1 fills mid_m/mid_e (and bwd_m/bwd_e if want_bwd).
 * Mode 1 takes sequences of at most 2^23 / R - G + 1 letters, G lanes and R
 * rows per lane being the configuration chosen for the pair (about a million
 * letters at R = 8, the largest); longer ones return DM_EUNSUP. */
int dm_plan_execute(dm_engine *e, dm_plan *plan, float *best, float *mid_m, int32_t *mid_e, float *bwd_m,
                    int32_t *bwd_e);
void dm_plan_destroy(dm_engine *e, dm_plan *plan);
/* Capture a plan's launches into a hipGraph; later dm_plan_execute calls
 * replay the graph.
 * A plan stays valid while its profiles keep their lengths and the sequence
 * store is not cleared: profiles and sequences may be added, and profiles
 * rewritten at the same length (dm_set_profile), between executions; a plan
 * then reads the current tables (a captured graph recorded against tables
 * that have since moved, or against other rates, is captured again). After
 * dm_clear_sequences, or when a profile of the plan changed length,
 * dm_plan_execute and dm_plan_capture return DM_EINVAL. */
int dm_plan_capture(dm_engine *e, dm_plan *plan);

/* Per kernel configuration of the last execution: G (lanes per pair), R (rows
 * per lane), row-strip kernel or not, device milliseconds (HIP events around
 * each launch), DP cells, launches. Returns the number of configurations used
 * (fills up to cap). */
int dm_kernel_stats(const dm_engine *e, int32_t cap, int32_t *G, int32_t *R, int32_t *strips, double *ms,
                    int64_t *cells, int64_t *launches);

/* Measurement aid with no reference counterpart: one timed launch of a plain
 * device-to-device copy of `bytes` (a multiple of 16), the reference point
 * that scales the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) of the DP
 * kernels. */
int dm_copy_probe(dm_engine *e, int64_t bytes, double *ms);
/* Hash of the kernel sources this library was built from (measurement files
 * record it, so that they are matched to the build they describe). */
const char *dm_build_id(void);

/* Post-gate processing on caller-supplied DP matrices (no device): the same
 * host code dm_pair_hits runs after its dump, fed with w0/W1 ExpScore matrices
 * ((M+1) x n), the pair's prefixSeed[n] and scales s1 s2 sC, and the rates. */
int dm_post_process(int32_t M, const float *values, const float *params, const float *bg, int32_t n,
                    const uint8_t *codes, const float *w0_m, const int32_t *w0_e, const float *W1_m,
                    const int32_t *W1_e, const float *prefix_seed, const float *scales, const double rates[6],
                    double min_prob_ratio, double *hits, int32_t hit_cap, int32_t *segs, int32_t seg_cap,
                    int32_t *nhits, int32_t *nsegs);

/* Calibration of the E-value parameters of nprof profiles (estimateK,
 * dummest-calibration.hh:338-600, called per profile from dummest.cc:384-408):
 * per profile, `trials` random frameshifting codon sequences of seq_len +
 * border letters drawn from the profile's background (trial t seeded with
 * 5489 + t), all scored in one batched forward+backward pass captured as one
 * hipGraph (dm_last_dp_ms gives its device time). Writes the
 * natural-log scores (mid-anchored and filter, nprof x trials each; either
 * may be NULL) and per profile 16 fits: estimateGumbel of the mid scores
 * (MM lambda, MM K, MM K lambda=1, ML lambda, ML K, ML K lambda=1, LM lambda,
 * LM K) then the same 8 of the filter scores. The reference keeps lambda =
 * fits[0], gumbelKmidAnchored = fits[1], lambda_filter = fits[8],
 * gumbel_k_filter = fits[9]. Its end- and start-anchored samples are all
 * -inf in this build (dummest-dp-forward.hh:536-543) and are not produced. */
int dm_calibrate(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border, int32_t trials,
                 double *mid_scores, double *flt_scores, double *fits);

/* The per-trial table that estimateK prints when the engine runs without
 * sequences (dummest-calibration.hh:392-396, 511-519, the verbose branch of
 * its trial loop): for the same trials as dm_calibrate, the anchor of each
 * trial's best mid-anchored score as profile position and sequence position
 * (anchors[2 (k trials + t)], [.. + 1]) and that score in nats (mid_scores,
 * may be NULL), taken from the same dumped matrices as dm_pairs_hits with a
 * negative minimum ratio (dummest-dp-forward.hh:536-543), so the anchor is the
 * reference's: the first column holding the highest cell, and within it the
 * first row. */
int dm_calibrate_anchors(dm_engine *e, int32_t nprof, const int32_t *prof, int32_t seq_len, int32_t border,
                         int32_t trials, int32_t *anchors, double *mid_scores);

/* Host-only pieces of dm_calibrate: the DNA letters (L + border) of trial
 * `trial` for a profile with background bg[29] (bg_probs, letters at 4..23)
 * under the given rates (dummest-calibration.hh:436-482), and estimateGumbel
 * of n scores (dummest-calibration.hh:7-104; out as in dm_calibrate's 8 fits). */
int dm_calib_sequence(const float *bg, const double rates[6], int32_t L, int32_t border, int32_t trial, char *out);
int dm_gumbel_fits(const double *scores, int32_t n, double seq_len, double out[8]);

/* Calibration cache shared with the reference engine (ProfileCache,
 * dummest-calibration.hh:191-336).
 * dm_xxh3_128: XXH3 128-bit hash, seed 0 (out[0] low half, out[1] high half;
 * the reference's Hash128, :106-146).
 * dm_calibration_key: computeCacheKey (:201-226) for a finalized profile
 * (values (M+1) x 29 as dm_get_profile returns them) and the calibration
 * options: 32 hex digits and a NUL. binary_hash is the hex hash of the
 * executable (getBinaryHash, :148-159), or NULL / "" when it is left out
 * (DUMMEST_CACHE_IGNORE_BINARY_HASH).
 * dm_cache_file_path: the cache file that getCacheFilePath (:233-265) picks
 * (creates its directory).
 * dm_cache_read / dm_cache_write: the cache file's layout (cereal binary
 * archive of unordered_map<string, CacheEntry>, 27 doubles per entry in
 * CacheEntry order, :148-181). dm_cache_read returns the number of entries
 * (a missing or empty file has none) and copies up to cap of them (keys
 * NUL-terminated at key_stride bytes each); DM_EFORMAT for a malformed file. */
int dm_xxh3_128(const void *data, int64_t len, uint64_t out[2]);
int dm_calibration_key(const char *name, int32_t M, const float *values, const char *binary_hash, int32_t seq_len,
                       int32_t border, int32_t trials, const double rates[6], double tantan_threshold, char out[33]);
int dm_cache_file_path(char *out, int32_t cap);
int64_t dm_cache_read(const char *path, int64_t cap, char *keys, int32_t key_stride, double *values);
int dm_cache_write(const char *path, int64_t n, const char *keys, int32_t key_stride, const double *values);

/* Profile builder's Baum-Welch E-step on the GPU (dummest-build.cc:561-736).
 * dm_bw_prepare: the ungapped sequences of one alignment (letter codes 0..A,
 * A = unknown letter; A = 20 protein or 4 DNA), their weights and the profile
 * length L; uploads them once for the iterations that follow.
 * dm_bw_estep: one E-step with the profile probs ((L+1) x (7+A) in the
 * builder's order: match, insStart, delStart, insEnd, insExt, delEnd, delExt,
 * letters): forward, backward and calculateTransitionCounts over every
 * sequence, weighted by weight / (total 2^-480), summed over sequences in
 * order into counts ((L+1) x (7+A); the delete-start column is left 0 for the
 * caller to derive), and each sequence's forward total (totals, nseq; may be
 * NULL). dm_last_bw_ms: device time of the last E-step. */
int dm_bw_prepare(dm_engine *e, int32_t A, int32_t L, int32_t nseq, const uint8_t *letters, const int64_t *offsets,
                  const int32_t *lengths, const double *weights);
int dm_bw_estep(dm_engine *e, const double *probs, double *counts, double *totals);
double dm_last_bw_ms(const dm_engine *e);

/* Gradient of the summed forward log-likelihood of one profile over DNA
 * sequences, on the GPU (FP64 forward, backward and expected counts over the
 * reference's recurrences, dummest-dp-forward.hh:9-250 and
 * dummest-dp-backward.hh:444-615, scaled by a power of two per 64-row strip
 * and 64-column block, so any ln Z fits; sequences of 3..2^28 letters). Per
 * sequence s the total of
 * all local alignments Z_s = sum_{i,j} prefixSeed(j) W1(i,j) (ln_z[s]; may be
 * NULL), and counts = sum_s weights[s] (NULL = 1) times the expected uses of
 * each parameter: (M+1) x 76 doubles, per profile row the 11 Params of
 * dm_add_profile (d ln Z / d ln param; entries 1, 2, 7, 8 are the global
 * frameshift rates) then the match emissions by codon (64 codons in ACGT
 * order, then codons holding another letter; d ln Z / d ln e of that codon).
 * The background and the null model stay fixed. dm_last_grad_ms: device time
 * of the DP, total and count kernels of the last call; dm_last_grad_cells: its cells
 * (sum of (M+1) n). "grad_budget" (dm_set_option) bounds the FP64 matrices of
 * one launch batch (default 8 GiB; 48 B per cell). A sequence whose matrices
 * alone pass it is checkpointed: both passes run once keeping only the edge
 * row of every 64-row strip (48 B per strip and letter), then each segment of
 * strips that fits the budget (at least one strip, 3 KiB per letter) is
 * computed again from its edge rows and counted; the results are the same bit
 * for bit. dm_last_grad_segments: the segments so recomputed by the last call
 * (0 when every sequence was held in full). Replaces the reference's full
 * O(M n) matrix per lane batch (dummest-dp-backward.hh:252). */
int dm_grad(dm_engine *e, int32_t prof, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
            const double *weights, double *ln_z, double *counts);
double dm_last_grad_ms(const dm_engine *e);
int64_t dm_last_grad_cells(const dm_engine *e);
int64_t dm_last_grad_segments(const dm_engine *e);
/* The same gradient over a registered sequence set, for callers that evaluate
 * it many times over the same sequences (gradient descent on a profile):
 * dm_grad_set_create decodes the sequences once, keeps them in the engine's
 * store and their codon indices on the device, and writes the set's id;
 * dm_grad_set computes what dm_grad computes for them (bit-identical results);
 * dm_grad_set_destroy frees the set's device memory. No counterpart in the
 * reference (its search never forms the total likelihood). */
int dm_grad_set_create(dm_engine *e, int32_t nseq, const char *dna, const int64_t *offsets, const int32_t *lengths,
                       int32_t *set);
int dm_grad_set(dm_engine *e, int32_t prof, int32_t set, const double *weights, double *ln_z, double *counts);
int dm_grad_set_destroy(dm_engine *e, int32_t set);

/* Autotuner: times every kernel configuration (G lanes per pair, R rows per
 * lane, and the row-strip kernel) on this device with synthetic pairs of
 * seq_len letters whose profile fills the configuration's rows, and keeps the
 * measured device time per wave-step; later plans choose, per pair, the
 * configuration with the least measured cost (n + G - 1) * strips * G/64 * t.
 * Writes up to cap times (ns) in configuration order; returns the number of
 * configurations. Replaces the reference's fixed lane packing
 * (simdWidth, DUMMEST_MAX_LANES; dummest-search.hh:85-87). */
int dm_autotune(dm_engine *e, int32_t seq_len, double *ns_per_step, int32_t cap);
/* Load a cost table saved from an earlier dm_autotune on the same device (n =
 * the number of configurations; n = 0 returns to the untuned cost model). */
int dm_set_config_costs(dm_engine *e, int32_t n, const double *ns_per_step);
/* The configuration table (G, R, strip kernel or not, tuned ns per wave-step or 0). */
int dm_config_table(const dm_engine *e, int32_t cap, int32_t *G, int32_t *R, int32_t *strips, double *ns_per_step);

/* Options: "plain" = 1 launches the DP once per pair (the unbatched form);
 * "grid" = persistent grid size (0 = automatic); "config" = index into the
 * configuration table (dm_config_table) that every pair whose profile fits it
 * runs on (-1 = automatic choice); "bw_budget" = bytes of Baum-Welch matrices
 * per launch batch (default 8 GiB; dm_bw_prepare splits the sequences by it);
 * "chunk_pairs" = most pairs in one work-queue chunk of a plan (default 2^22; a
 * plan's chunks run one after the other and share its device buffers). */
int dm_set_option(dm_engine *e, const char *name, int64_t value);

/* Device time of the DP kernels of the last scoring call (HIP events on the
 * engine's stream), in milliseconds, and the number of DP cells it covered
 * (sum over pairs of (M+1)*n). */
double dm_last_dp_ms(const dm_engine *e);
/* Device time of the tantan kernel of the last dm_tantan_dna / masked dm_add_dna call, ms. */
double dm_last_tantan_ms(const dm_engine *e);
int64_t dm_last_cells(const dm_engine *e);

#ifdef __cplusplus
}
#endif
#endif /* DUMMEST_AMD_H */
