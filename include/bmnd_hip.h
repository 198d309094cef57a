/*
 * C-ABI of libbmnd_hip.so -- the MI355X (gfx950) kernels behind bmnd_amd.
 *
 * All pointers are device pointers unless the name ends in _h (host).
 * Volumes are C-contiguous float32 and always described as 3-D: a 2-D image
 * of shape (H, W) is passed as (1, H, W) with block (1, bh, bw) and radius
 * (0, rh, rw).  Patch origins ("lin") are C-order linear indices into the
 * patch grid counts[d] = shape[d] - block[d] + 1, exactly as the reference's
 * blockmatching.multi_index_to_linear (src/bmnd/blockmatching.py:458-465).
 * `stream` is a hipStream_t (0 = the null stream).  Every entry point returns
 * 0 on success and a negative status otherwise; bmnd_hip_last_error() then
 * holds a message.  No exceptions cross the ABI; the caller owns every
 * buffer.
 */
#ifndef BMND_HIP_H
#define BMND_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMND_HIP_OK 0
#define BMND_HIP_EINVAL -1      /* argument outside what the kernels support */
#define BMND_HIP_ELAUNCH -2     /* HIP runtime error */
#define BMND_HIP_ECAPACITY -3   /* a per-tile capacity was exceeded */

/* Library version (major*10000 + minor*100 + patch). */
int32_t bmnd_hip_version(void);

/* Thread-local message of the last failing call. */
const char *bmnd_hip_last_error(void);

/* Names of the block/radius configurations compiled into this library,
 * one per line; for diagnostics. */
const char *bmnd_hip_configs(void);

/* Poisson variance volume: out = max(scale * max(v, 0), floor) in float32.
 * Replaces variance.build_poisson_variance_model (variance.py:556-598). */
int32_t bmnd_hip_variance_volume(const float *vol, float *out, int64_t n,
                                 double scale, double floor_, void *stream);

/* Count check used to pick the block-matching mode: flags[0] = 1 when every
 * voxel is a non-negative integer <= 2^24, flags[1] = the largest such value
 * (flags is a device int32[2]).  No reference counterpart; it guards the
 * integer-count table of bmnd_hip_blockmatch mode 2. */
int32_t bmnd_hip_count_range(const float *vol, int64_t n, int32_t *flags,
                             void *stream);

/* Fused block matching for every reference of a stage.
 * Replaces blockmatching.blockmatch_groups -> _fused_blockmatch_groups_batch_numba
 * (blockmatching.py:1393-1511, 1045-1205) for the Poisson deviance distance
 * (blockmatching.py:322-371) with tie canonicalisation (957-1042) and
 * finalisation (913-954).
 *   ref_lin[nrefs]            patch-grid linear index of each reference
 *   sel[nrefs*max_matches]    out: selected lins, reference first, -1 padded
 *   counts[nrefs]             out: final group size
 *   accepted[nrefs]           out: number of candidates under the threshold
 * mode 0 evaluates each deviance term as the reference writes it (two
 * divisions, two logarithms in float64); mode 1 uses the one-log form
 * r log 2r + c log 2c - s log s; mode 2 requires integer counts in
 * [0, table_max] (table_max <= 100, count_scale == 1) and reads each term
 * from a float64 table built on the host with the reference's expression
 * (the caller checks the precondition, e.g. with bmnd_hip_count_range; the
 * kernel clamps counts to [0, table_max] so that no other input reads past
 * the table, but the lists of such an input are meaningless);
 * mode 3 (certified) evaluates the one-log form with a degree-4 logarithm
 * together with an error bound per reference and re-evaluates the reference
 * expression wherever the bound leaves the tie-grid key open, so its output
 * equals mode 0's; mode 4 (fast certified) first bounds every candidate's
 * key with a float32 pass (hardware log2) and evaluates mode 3's keys only
 * for the candidates that can change the selection or the accepted count,
 * so its output equals mode 0's too.  Modes 1, 3 and 4 fall back to mode 0
 * when count_scale != 1; mode 4 falls back to mode 3 on a device whose
 * hardware log2 is not within the bound bmnd_hip_log2_check measures.
 * Modes 5 and 6 match with the Pearson and the Anscombe SSD distances
 * (blockmatching.py:356-369), evaluated as the reference writes them in
 * float64.  Modes 7-9 (standardized) need bmnd_hip_blockmatch_standardized. */
int32_t bmnd_hip_blockmatch(const float *vol, const int64_t *shape_h,
                            const int32_t *block_h, const int32_t *radius_h,
                            const int32_t *ref_lin, int64_t nrefs,
                            float threshold, int32_t max_matches,
                            int32_t min_matches, int32_t pow2,
                            int32_t stabilize, double count_scale,
                            int32_t mode, int32_t table_max, int32_t *sel,
                            int32_t *counts, int32_t *accepted, void *stream);

/* Accuracy check behind block-matching mode 4 (no reference counterpart):
 * *max_ratio = max |v_log_f32(x) - log2 x| / (|log2 x| + 1) over every normal
 * float x >= 2^-126 on the current device (measured once per device and
 * cached).  Mode 4 runs when it is <= 2^-20. */
int32_t bmnd_hip_log2_check(double *max_ratio);

/* Matrix table consumed by bmnd_hip_filter_groups, float32, packed as
 *   G_n (n x n, row-major, out index first) for n = 1..max_matches,
 *   then Ginv_n for n = 1..max_matches,
 *   then the spatial analysis matrices of axes 0,1,2 (b_d x b_d),
 *   then the spatial synthesis matrices of axes 0,1,2,
 * then, from the next even float offset, float64 values
 * (aggregationweights.py:79-120, poisson.py:120-160):
 *   the spatial metric diagonal diag(S^T diag(w^2) S) [P] (S the Kronecker
 *   product of the synthesis matrices, w the float32 Kaiser window),
 *   the group metric diagonals diag(Ginv_n^T Ginv_n) for n = 1..max (n each),
 *   the constant-synthesis directions solve(Ginv_n, 1) for n = 1..max,
 *   then solve(S_d, 1) of each spatial axis d,
 * then float32 pair tables for bmnd_hip_exact_std_pairs: for each axis
 *   d = 0,1,2 and offset e = 0..2 b_d - 2, the b_d x b_d matrix
 *   H[t][p] = f32(A_d[t][p] * A_d[t][p - e + b_d - 1]), 0 where that column
 *   leaves 0..b_d-1 (A_d the analysis matrix of axis d),
 * then the float32 spatial metric of the exact group risk,
 *   SM = S^T diag(w^2) S [P][P] (S the Kronecker product of the synthesis
 *   matrices, w the float32 Kaiser window; variance.py:1543-1556).
 * bmnd_hip_matrix_table_size() gives its length in floats. */
int64_t bmnd_hip_matrix_table_size(int32_t max_matches, const int32_t *block_h);

/* bmnd_hip_blockmatch with one threshold per reference (ref_thresholds,
 * device, nrefs float32): the "reference_finite_count" stage-1 policy, whose
 * thresholds compute_poisson_reference_thresholds forms on the host
 * (blockmatching.py:95-148, bmndalgo.py:269-284); each reference's tie-grid
 * key threshold is its threshold on the 32-ulp grid. */
int32_t bmnd_hip_blockmatch_thresholds(const float *vol, const int64_t *shape_h,
                                       const int32_t *block_h, const int32_t *radius_h,
                                       const int32_t *ref_lin, const float *ref_thresholds,
                                       int64_t nrefs, int32_t max_matches, int32_t min_matches,
                                       int32_t pow2, int32_t stabilize, double count_scale,
                                       int32_t mode, int32_t table_max, int32_t *sel,
                                       int32_t *counts, int32_t *accepted, void *stream);

/* bmnd_hip_blockmatch with the candidate-standardized distance of the
 * "candidate_standardized" stage-1 policy (blockmatching.py:374-426,
 * bmndalgo.py:311-316): per candidate (sum of the distance's terms - sum of
 * the null means) / sqrt(sum of the null variances) over the patch voxels of
 * positive pooled count, null sums in float32, signed and not divided by the
 * patch size.  distance: 1 deviance, 2 Pearson, 3 Anscombe SSD.
 * null_moments (device float32, 2 * (null_max + 1)): the null means for pooled
 * counts 0..null_max, then the null variances (poisson_null_moment_table cast
 * to float32, blockmatching.py:288-319); pooled counts above null_max take 1
 * and 2. */
int32_t bmnd_hip_blockmatch_standardized(const float *vol, const int64_t *shape_h,
                                         const int32_t *block_h, const int32_t *radius_h,
                                         const int32_t *ref_lin, int64_t nrefs, float threshold,
                                         int32_t max_matches, int32_t min_matches, int32_t pow2,
                                         int32_t stabilize, double count_scale, int32_t distance,
                                         const float *null_moments, int32_t null_max,
                                         int32_t *sel, int32_t *counts, int32_t *accepted,
                                         void *stream);

/* SSD block matching of the stationary Gaussian model (blockmatching.py:
 * 320-334, 1046-1205 with distance "ssd"): per candidate the float32 SSD
 * summed in patch C order, less gamma * noise_ssd[lag] when noise_ssd
 * (device, the (2R+1) search box in C order, zero lag 0) is given
 * (blockmatching.py:244-286, 1133-1139); accepted below threshold, sorted
 * by (distance, index), no tie grid; then the same group assembly as
 * bmnd_hip_blockmatch.  The threshold is the profile's times the squared
 * value range of vol (blockmatching.py:168-184; bmnd_hip_value_range).
 * Refuses a negative or non-finite gamma or threshold and a nonzero gamma
 * without noise_ssd.  Replaces blockmatching.blockmatch_groups
 * (blockmatching.py:1393-1505) on the Gaussian path (bmndalgo.py:291-326,
 * 722-743). */
int32_t bmnd_hip_blockmatch_ssd(const float *vol, const int64_t *shape_h,
                                const int32_t *block_h, const int32_t *radius_h,
                                const int32_t *ref_lin, int64_t nrefs, float threshold,
                                int32_t max_matches, int32_t min_matches, int32_t pow2,
                                const float *noise_ssd, float gamma, int32_t *sel,
                                int32_t *counts, int32_t *accepted, void *stream);

/* Stage-1 SSD block matching of the Poisson model (blockmatch_distance
 * "ssd" with noise_model "poisson": bmndalgo.py:292-294, 321-322) with the
 * Poisson gamma correction (blockmatching.py:432-443, 1140-1145): every
 * candidate other than the reference has gamma times its Poisson noise bias
 * (the float64 serial sum over the patch of count_scale * (r + c) where the
 * float32 r + c is positive) taken off its float32 SSD.  The corrected
 * distance is a float64, accepted when below threshold and ordered as the
 * reference's float32 candidate buffers order it (a later candidate goes
 * first when its float64 distance is below the earlier one's float32 value);
 * then the same group assembly as bmnd_hip_blockmatch.  gamma == 0 is plain
 * SSD.  The threshold is the profile's times the squared value range of vol
 * (blockmatching.py:168-184).  Refuses a negative or non-finite gamma or
 * threshold and a count_scale that is not finite and positive.  Replaces
 * blockmatching.blockmatch_groups (blockmatching.py:1393-1505) in stage 1 of
 * the Poisson path under the SSD distance; stage 2 is bmnd_hip_blockmatch_ssd
 * without a noise SSD (bmndalgo.py:728-741). */
int32_t bmnd_hip_blockmatch_ssd_poisson(const float *vol, const int64_t *shape_h,
                                        const int32_t *block_h, const int32_t *radius_h,
                                        const int32_t *ref_lin, int64_t nrefs, float threshold,
                                        int32_t max_matches, int32_t min_matches, int32_t pow2,
                                        float gamma, double count_scale, int32_t *sel,
                                        int32_t *counts, int32_t *accepted, void *stream);

/* Median and median absolute deviation behind the Gaussian sigma estimate
 * (sigmaestimation.py:7-45): of the finite nonzero values of vol (the finite
 * ones when every finite value is zero), out (device, 2 floats) <- numpy's
 * median m (the float32 mean of the two central values for an even count)
 * and the numpy median of |x - m|; counts (device, 2 x uint64) <- the number
 * of values used and 1 when they are the finite ones.  The estimate is
 * out[1] / 0.6745 when positive; the reference's IQR and trimmed-std
 * fallbacks take over otherwise. */
int32_t bmnd_hip_median_mad(const float *vol, int64_t n, float *out,
                            unsigned long long *counts, void *stream);

/* range (device, 2 floats) <- (min, max) of the n values of vol: the data
 * scale of compute_blockmatch_threshold (blockmatching.py:168-184). */
int32_t bmnd_hip_value_range(const float *vol, int64_t n, float *range, void *stream);

/* Exact-plane noise variance of the stationary Gaussian model for the first
 * min(k, N) planes of every group (variance.py:1889-2033 as called by
 * get_stationary_gaussian_group_noise_variance, variance.py:2036-2106): for
 * plane q and spatial coefficient s, over the nonzero entries f of row q of
 * the member matrix in increasing order, the float32 serial sum of
 * (G[q,f] C_s(0)) G[q,f] and, for each nonzero t < f, of
 * (G[q,f] G[q,t]) (C_s(l) + C_s(-l)) with l the members' origin difference
 * modulo the domain, floored at 0.
 *   cov        device, [D][P] float32: the cross-covariance maps
 *              (variance.py:1791-1886) lag-major, D = domain[0]*[1]*[2]
 *              lags in C order (2-D inputs as (1, d0, d1)); each map even in
 *              the lag
 *   exact_var  out: (nrefs, min(k, max_matches), P) */
int32_t bmnd_hip_gaussian_variance(const float *cov, const int32_t *domain_h,
                                   const int64_t *shape_h, const int32_t *block_h,
                                   const int32_t *radius_h, const int32_t *sel,
                                   const int32_t *counts, int64_t nrefs, int32_t max_matches,
                                   const float *matrices, int32_t k, float *exact_var,
                                   void *stream);

/* Exact overlap-aware noise std of the first min(k, N) group planes of every
 * group: var[q,s] = sum over union voxels v (C order) of
 * var_v * (sum over entries (i,p) with x_i + p = v of G[q,i]*S[s,p])^2,
 * every product and sum a float32 operation in that serial order; out is
 * sqrt(max(var, 0)).  exact_std holds (nrefs, min(k, max_matches), P).
 * Replaces variance._build_poisson_overlap_structure_numba and
 * _accumulate_exact_poisson_variance_from_entry_transform_numba
 * (variance.py:652-851, 920-975) as called by get_direct_poisson_noise_std
 * (variance.py:1634-1697). */
int32_t bmnd_hip_exact_std(const float *var, const int64_t *shape_h,
                           const int32_t *block_h, const int32_t *radius_h,
                           const int32_t *sel, const int32_t *counts,
                           int64_t nrefs, int32_t max_matches,
                           const float *matrices, int32_t k, float *exact_std,
                           void *stream);

/* The same exact-plane std summed in another order, for speed: expanding the
 * square over the members covering each voxel,
 *   var[q,s] = sum over member pairs i <= j whose patches overlap of
 *              G[q,i] G[q,j] (x2 for i < j) *
 *              sum over the overlap p of var(x_i + p) S[s,p] S[s,p - (x_j - x_i)],
 * each pair's inner sum a separable transform of its overlap region (three
 * float32 passes, one per axis) and the plane products added on the matrix
 * cores.  The same quantity as bmnd_hip_exact_std (variance.py:652-975) to
 * float32 rounding, not bit for bit: its sums run in a different order.
 * Same arguments and output. */
int32_t bmnd_hip_exact_std_pairs(const float *var, const int64_t *shape_h,
                                 const int32_t *block_h, const int32_t *radius_h,
                                 const int32_t *sel, const int32_t *counts, int64_t nrefs,
                                 int32_t max_matches, const float *matrices, int32_t k,
                                 float *exact_std, void *stream);

/* The exact covariance-aware group risk of the windowed-synthesis,
 * group-scope weights (bmndalgo.py:359-377, 493-505, 752-758, 913-926;
 * wiener.py:638-844): bmnd_hip_filter_groups_sigma with the weights of the
 * variance model (HT) or the variance / risk model (Wiener, code bits 3-4)
 * in the windowed-synthesis domain (bit 5) with group scope (bit 6) and k >
 * 0 replaced by w = f32(1 / (risk + 1e-10)) for every member, where risk is
 * the trace of the windowed synthesis of the filtered group's exact noise
 * covariance over its first min(k, n) planes (the overlap-aware covariance
 * from var), plus the diagonal variance of the later planes and, for the
 * risk model, the bias.  The reference builds that covariance per union
 * voxel; here it is regrouped by overlapping member pairs and transformed
 * separably, the same quantity to float32 rounding.  The Poisson model only.
 * Same arguments as bmnd_hip_filter_groups_sigma (code need not carry the
 * bit this call adds), plus a workspace of bmnd_hip_exact_risk_workspace()
 * bytes. */
int64_t bmnd_hip_exact_risk_workspace(int64_t nrefs, int32_t max_matches, int32_t k,
                                      const int32_t *block_h);
int32_t bmnd_hip_filter_groups_exact(int32_t code, const float *noisy, const float *pilot,
                                     const float *var, const int64_t *shape_h,
                                     const int32_t *block_h, const int32_t *radius_h,
                                     const int32_t *sel, const int32_t *counts, int64_t nrefs,
                                     int32_t max_matches, const float *matrices,
                                     const float *exact_std, float shrink, int32_t k,
                                     double count_scale, float *filt, float *weights,
                                     float *sigma_out, void *workspace, int64_t workspace_bytes,
                                     void *stream);

/* Flags OR'd into the k argument of bmnd_hip_stage, bmnd_hip_stage_noalloc,
 * bmnd_hip_stage_workspace and into bmnd_hip_pipeline_desc.k (the exact
 * planes are k & BMND_HIP_K_PLANES): the HT or the Wiener stage computes its
 * exact planes with bmnd_hip_exact_std_pairs instead of bmnd_hip_exact_std;
 * BMND_HIP_MFMA_WIENER sets filter code bit 10 (the matrix-core member
 * transforms) for the Wiener stage. */
#define BMND_HIP_K_PLANES 0xffff
#define BMND_HIP_EXACT_PAIRS_HT (1 << 16)
#define BMND_HIP_EXACT_PAIRS_WIENER (1 << 17)
#define BMND_HIP_MFMA_WIENER (1 << 18)

/* Collaborative filtering of every group of a stage: gather, separable
 * forward transform, coefficient noise std (the exact planes q < k from
 * bmnd_hip_exact_std plus the rescaled approximate planes), shrinkage,
 * inverse transform and per-patch aggregation weights.
 * Replaces bmndalgo.filter_ht_group (bmndalgo.py:335-559) when wiener == 0
 * and bmndalgo.filter_wiener_group (bmndalgo.py:760-981) when wiener != 0,
 * with variance.get_direct_poisson_noise_std (variance.py:1634-1711),
 * poisson._poisson_hard_threshold_mask (poisson.py:30-46),
 * wiener.compute_poisson_wiener_gain (wiener.py:78-95) and
 * aggregationweights.compute_aggregation_weights (aggregationweights.py:690-800).
 *   noisy      observation volume (HT: the coefficients to shrink)
 *   pilot      Wiener only: pilot volume whose coefficients drive the gain
 *   var        variance volume of the stage
 *   filt       out: (nrefs*max_matches, P) filtered patches, row g*max+i
 *   weights    out: (nrefs*max_matches) float32 weights
 * shrink is ht_lambda_threshold (HT) or wiener_variance_scale (Wiener).
 * wiener is the filter code: bits 0-1 select the stage and the Wiener gain's
 * signal power (wiener.py:24-53): 0 hard threshold, 1 variance-scaled
 * max(r^2 - shrink sigma^2, 0), 2 classic r^2, 3 noise-floor
 * max(r^2 - sigma^2, 0); bit 2 (HT only) soft thresholding
 * (bmndalgo.py:475-483); bits 3-4 the aggregation weight model
 * (aggregationweights.py:214-259): 0 variance, 1 classic, 2 risk (Wiener
 * only; the gain's signal power as S); bit 5 the windowed-synthesis weight
 * domain; bit 6 the group weight scope (one risk per group, numpy's float64
 * sum; aggregationweights.py:175-211); bit 7 writes the shrunk coefficients
 * to filt instead of the inverse-transformed patches (the input of
 * bmnd_hip_mass_correct); bit 8 the stationary Gaussian model: exact_std
 * then holds the exact planes' variances from bmnd_hip_gaussian_variance,
 * var the P coefficient variances (entry 0 of each cross-covariance map),
 * the remaining planes share (n pv - sum of the exact planes) / (n - k)
 * (variance.py:2096-2106), the hard threshold keeps |C| >= lambda sigma and
 * the group's first coefficient (bmndalgo.py:470-474), and the Wiener gain
 * is |r|^2 / (|r|^2 + shrink max(sigma^2, 1e-8)) (wiener.py:56-75), its
 * signal power |r|^2 for the risk model.  Bit 10 (Wiener only) runs the
 * member-axis products of the four group transforms on the matrix cores
 * (v_mfma_f32_16x16x4f32): the same products summed in the hardware's
 * order, so the output equals the serial fmaf chains of the reference's
 * matrix products up to float32 rounding (the bit is off for the serial
 * order).  The same values in every entry that takes a filter code. */
int32_t bmnd_hip_filter_groups(int32_t wiener, const float *noisy,
                               const float *pilot, const float *var,
                               const int64_t *shape_h, const int32_t *block_h,
                               const int32_t *radius_h, const int32_t *sel,
                               const int32_t *counts, int64_t nrefs,
                               int32_t max_matches, const float *matrices,
                               const float *exact_std, float shrink, int32_t k,
                               double count_scale, float *filt, float *weights,
                               void *stream);

/* bmnd_hip_filter_groups plus the noise std that bmnd(return_sigma_map=True)
 * scatters into its sigma maps (bmndalgo.py:557, 576-595 for HT;
 * bmndalgo.py:874-875, 979, 1001-1011 for Wiener):
 *   sigma_out  HT: (nrefs*max_matches, P) limited std of every coefficient,
 *              row g*max+i; Wiener: (nrefs) float32 mean of each group's
 *              stds, in numpy's summation order (np.mean).  NULL = none. */
int32_t bmnd_hip_filter_groups_sigma(int32_t wiener, const float *noisy,
                                     const float *pilot, const float *var,
                                     const int64_t *shape_h, const int32_t *block_h,
                                     const int32_t *radius_h, const int32_t *sel,
                                     const int32_t *counts, int64_t nrefs,
                                     int32_t max_matches, const float *matrices,
                                     const float *exact_std, float shrink, int32_t k,
                                     double count_scale, float *filt, float *weights,
                                     float *sigma_out, void *stream);

/* Reference-position map used by the aggregation: ref_index[lin] = g for
 * every scheduled reference g, -1 elsewhere (num_patches entries). */
int32_t bmnd_hip_reference_map(const int32_t *ref_lin, int64_t nrefs,
                               int32_t *ref_index, int64_t num_patches,
                               void *stream);

/* Ordered aggregation and normalisation: every output voxel sums its
 * contributions num += f*(w*k), den += w*k sequentially in (schedule index,
 * member index) order -- the order of the reference's serial scatter -- and
 * writes est = den > 0 ? num / (den + 1e-10f) : fallback.
 * Replaces patchaggregation._scatter_add_patches_numba
 * (patchaggregation.py:170-199, 373-397) plus the normalisation of
 * bmndalgo.py:626-631 / 1039-1044.  Deterministic; no atomics on floats. */
int32_t bmnd_hip_aggregate(const float *filt, const float *weights,
                           const int32_t *sel, const int32_t *counts,
                           const int32_t *ref_index, int64_t nrefs,
                           int32_t max_matches, const int64_t *shape_h,
                           const int32_t *block_h, const int32_t *radius_h,
                           const float *kaiser, const float *fallback,
                           float *est, void *stream);

/* bmnd_hip_aggregate plus one stage's sigma map: the same ordered walk sums
 * snum += s*(w*k) with s = group_sigma[row*P + p] (per_group == 0, the HT
 * layout of bmnd_hip_filter_groups_sigma) or group_sigma[g] (per_group != 0,
 * the Wiener layout), and writes sigma_map = den > 0 ? snum / den : fill.
 * Replaces patchaggregation._scatter_add_patches_with_sigma_numba and
 * _scatter_add_patches_with_scalar_sigma_numba (patchaggregation.py:229-294,
 * 400-470) and the sigma-map normalisation of bmndalgo.py:1048-1062; fill is
 * the global sigma (bmnd_hip_poisson_global_sigma) as float32. */
int32_t bmnd_hip_aggregate_sigma(const float *filt, const float *weights,
                                 const int32_t *sel, const int32_t *counts,
                                 const int32_t *ref_index, int64_t nrefs,
                                 int32_t max_matches, const int64_t *shape_h,
                                 const int32_t *block_h, const int32_t *radius_h,
                                 const float *kaiser, const float *fallback,
                                 float *est, const float *group_sigma,
                                 int32_t per_group, float fill, float *sigma_map,
                                 void *stream);

/* bmnd_hip_aggregate_sigma with a choice of normalisation: norm_mode 0 is
 * num / (den + 1e-10) where den > 0 (the others' rule, bmndalgo.py:626-631,
 * 1039-1044); 1 is num / den where den > 0, the fallback elsewhere (the
 * np.divide of a group-mass-conserving stage, bmndalgo.py:618-626,
 * 1031-1038); 2 writes den itself into est (the denominator the mass
 * correction reads).  group_sigma / sigma_map may be NULL. */
int32_t bmnd_hip_aggregate_norm(int32_t norm_mode, const float *filt, const float *weights,
                                const int32_t *sel, const int32_t *counts,
                                const int32_t *ref_index, int64_t nrefs, int32_t max_matches,
                                const int64_t *shape_h, const int32_t *block_h,
                                const int32_t *radius_h, const float *kaiser,
                                const float *fallback, float *est, const float *group_sigma,
                                int32_t per_group, float fill, float *sigma_map, void *stream);

/* Group-mass correction of every group (poisson.py:166-265,
 * patchaggregation.py:71-98), in place on filt, which holds each group's
 * shrunk coefficients (bmnd_hip_filter_groups with filter code bit 7): the
 * output functional O = f32(w_i k_p) / den at every member voxel, the target
 * sum O * source, and the correction of the coefficients along the
 * constant-synthesis direction (matrix table) that makes the synthesised
 * group carry it; filt then holds the inverse-transformed patches for
 * bmnd_hip_aggregate_norm(1, ...).  den is the stage's aggregation
 * denominator (bmnd_hip_aggregate_norm(2, ...)), source the observation.
 * Float64 sums in a fixed order (a tolerance restatement of the reference's
 * numpy sums). */
int32_t bmnd_hip_mass_correct(float *filt, const float *weights, const int32_t *sel,
                              const int32_t *counts, int64_t nrefs, int32_t max_matches,
                              const int64_t *shape_h, const int32_t *block_h,
                              const int32_t *radius_h, const float *den, const float *source,
                              const float *kaiser, const float *matrices, void *stream);

/* Global noise std of the Poisson variance model, the "global" entry of the
 * sigma maps: sigmaestimation.estimate_poisson_global_sigma
 * (sigmaestimation.py:48-93) on the n-voxel variance volume var (device),
 * with eps = max(variance_floor, 1e-10) as variance.build_poisson_variance_model
 * passes it (variance.py:581-593): sqrt(max(mean, eps)) of the float32 mean,
 * in numpy's summation order, of the finite positive values (a 10%-trimmed
 * mean of the sorted values when they are fewer than a tenth of the finite
 * ones and at least 3).  Synchronises the stream; *global_sigma is a host
 * pointer. */
int32_t bmnd_hip_poisson_global_sigma(const float *var, int64_t n,
                                      double variance_floor, double *global_sigma,
                                      void *stream);

/* ---- one whole stage ---------------------------------------------------------
 * Device workspace bytes bmnd_hip_stage needs (the exact-plane std, filtered
 * patches, weights and reference map of one stage), or -1 for an invalid
 * geometry.  Host-only. */
int64_t bmnd_hip_stage_workspace(const int64_t *shape_h, const int32_t *block_h,
                                 int64_t nrefs, int32_t max_matches, int32_t k);

/* One denoising stage of the Poisson bmnd(), the HT pilot (wiener == 0) or the
 * Wiener estimate: block matching, exact-plane variance, group filtering,
 * reference map and ordered aggregation, in one call on the caller's stream
 * (bmndalgo.py:241-633 for HT, 676-1046 for Wiener).  The caller owns every
 * buffer: match_vol (the observation for HT, the pilot for Wiener), noisy
 * (the observation), var (the stage's variance volume, bmnd_hip_variance_volume),
 * fallback (the observation), the schedule ref_lin, the matrix table
 * (bmnd_hip_matrix_table_size), the Kaiser window (P floats), the workspace of
 * bmnd_hip_stage_workspace() bytes, and the outputs est, sel, counts,
 * accepted.  mode as in bmnd_hip_blockmatch (0-6; the standardized modes
 * need bmnd_hip_blockmatch_standardized's null table and are refused here),
 * or -1 for the automatic choice (the count table for integer counts <= 100
 * at count_scale 1, which reads two integers back from the device, else mode
 * 4).  An explicit mode 2 reads the count range back too, takes the table
 * size from it and returns BMND_HIP_EINVAL unless match_vol holds
 * non-negative integers <= 100.  wiener as in bmnd_hip_filter_groups (0 HT,
 * 1-3 Wiener gain).  The Wiener gains come from match_vol's coefficients, so
 * the stage driver covers the reference's default Wiener match source (the
 * pilot, bmndalgo.py:714-720); a caller matching on the observation
 * (poisson_wiener_match_source = observation) calls the per-kernel entry
 * points and passes the pilot to bmnd_hip_filter_groups. */
int32_t bmnd_hip_stage(int32_t wiener, const float *match_vol, const float *noisy,
                       const float *var, const float *fallback, const int64_t *shape_h,
                       const int32_t *block_h, const int32_t *radius_h, const int32_t *ref_lin,
                       int64_t nrefs, float threshold, int32_t max_matches, int32_t min_matches,
                       double count_scale, int32_t mode, const float *matrices,
                       const float *kaiser, float shrink, int32_t k, void *workspace,
                       int64_t workspace_bytes, float *est, int32_t *sel, int32_t *counts,
                       int32_t *accepted, void *stream);

/* Library set-up for the current device (no counterpart in the reference):
 * builds every table the kernels read (float64 log tables, the count-table
 * terms for every table size 0..100, the hardware-log2 accuracy check) and
 * looks up the device address of the no-allocation scratch.  Allocates; call
 * it once per device before bmnd_hip_stage_noalloc, outside any region that
 * must not allocate.  Repeated calls return at once. */
int32_t bmnd_hip_init(void);

/* bmnd_hip_stage without the workspace argument and without any allocation
 * (no counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The stage's workspace, the fast matcher's patch sums
 * and the aggregation's capacity flag live in one static device array of
 * 64 MiB per device, held until the process ends; a stage needing more
 * (bmnd_hip_stage_workspace() + 256 + its patch sums) returns
 * BMND_HIP_ECAPACITY, and a device on which bmnd_hip_init() has not run
 * returns BMND_HIP_EINVAL.  Flags and maps are cleared by kernels (no memset).
 * Every call on one device shares that array and leaves its contents behind:
 * calls that overlap on one device (other streams, other threads) race.
 * Otherwise identical to bmnd_hip_stage, bit for bit; like it, it reads the
 * count range and the capacity flag back with a stream synchronisation. */
int32_t bmnd_hip_stage_noalloc(int32_t wiener, const float *match_vol, const float *noisy,
                               const float *var, const float *fallback, const int64_t *shape_h,
                               const int32_t *block_h, const int32_t *radius_h,
                               const int32_t *ref_lin, int64_t nrefs, float threshold,
                               int32_t max_matches, int32_t min_matches, double count_scale,
                               int32_t mode, const float *matrices, const float *kaiser,
                               float shrink, int32_t k, float *est, int32_t *sel,
                               int32_t *counts, int32_t *accepted, void *stream);

/* ---- batches of inputs of different shapes --------------------------------
 * No counterpart in the reference, which denoises one array per call: a batch
 * runs every input of one profile (one block / search geometry) through each
 * kernel in a single launch over the flat list of all inputs' references
 * (or aggregation tiles).  The inputs' voxels, patch grids and tiles are
 * packed one after the other in flat buffers, in input order; input i is
 * described by */
typedef struct {
  int64_t vox_off;   /* first voxel in the packed volumes: sum of the earlier inputs' voxels */
  int64_t grid_off;  /* first patch origin in packed patch-grid buffers: sum of earlier grids */
  int64_t tile_off;  /* first aggregation tile: sum of earlier inputs' tile counts */
  int32_t shape[3];  /* voxels per axis; images as (1, H, W) */
  int32_t grid[3];   /* patch-grid counts shape - block + 1 of the stage */
} bmnd_hip_input;

/* ---- both stages captured as one HIP graph ---------------------------------
 * No counterpart in the reference (which has no device): the whole two-stage
 * Poisson pipeline of bmnd() (bmndalgo.py:213-1046) recorded once as a HIP
 * graph and relaunched with one call, for inputs small enough that kernel
 * launches and the per-stage read-backs dominate.  The matching modes are
 * fixed at creation (0..6 as bmnd_hip_blockmatch; mode 2 with table_max);
 * the library allocates the pipeline's workspace and builds its tables at
 * creation; a launch reads desc.input and writes desc.pilot, desc.output
 * and each stage's sel / counts / accepted, with no allocation, memset or
 * host synchronisation.  After a launch, bmnd_hip_pipeline_status (which
 * synchronises the stream) returns BMND_HIP_EINVAL if stage 1 ran the count
 * table on an input it does not hold (not integer counts <= table_max: the
 * results are not valid and the caller reruns or recreates), or
 * BMND_HIP_ECAPACITY as bmnd_hip_aggregate. */
typedef struct {
  int64_t shape[3];             /* voxels per axis; images as (1, H, W) */
  int32_t block[2][3];          /* [stage][axis], stage 0 HT, 1 Wiener */
  int32_t radius[2][3];
  const int32_t *ref_lin[2];    /* device: each stage's references */
  int64_t nrefs[2];
  float threshold[2];
  int32_t max_matches[2], min_matches[2];
  int32_t mode[2];              /* matching modes 0..6 */
  int32_t table_max;            /* stage 1 in mode 2: the count table's size */
  const float *matrices[2];     /* device, bmnd_hip_matrix_table_size layout */
  const float *kaiser[2];       /* device, P floats */
  float shrink[2];              /* lambda (HT), mu (Wiener) */
  int32_t k;
  int32_t wiener_gain;          /* 1 variance-scaled, 2 classic, 3 noise-floor */
  double count_scale, variance_floor;
  int32_t match_observation;    /* Wiener matching on the observation */
  int32_t variance_observation; /* Wiener variance model from the observation */
  const float *input;           /* device volumes */
  float *pilot, *output;
  int32_t *sel[2], *counts[2], *accepted[2];
  /* Optional batch.  ninputs = 0: one input of `shape`.  ninputs > 0: input,
   * pilot and output hold every input packed one after another; batch[st]
   * (host) and batch_d[st] (device) describe them for stage st as the
   * bmnd_hip_*_batch calls take them (that stage's grids, offsets and
   * aggregation tiles); ref_lin, ref_input, sel, counts and accepted are the
   * stage's flat lists over all inputs (ref_input: each reference's input);
   * `shape` is not read.  Stage 1's mode applies to every input. */
  int32_t ninputs;
  const bmnd_hip_input *batch[2];
  const bmnd_hip_input *batch_d[2];
  const int32_t *ref_input[2];
} bmnd_hip_pipeline_desc;
typedef struct bmnd_hip_pipeline bmnd_hip_pipeline;
int32_t bmnd_hip_pipeline_create(const bmnd_hip_pipeline_desc *desc, bmnd_hip_pipeline **out);
/* One graph over `count` descriptors, run one after another on the launch's
 * stream (for instance a batch of images and a batch of volumes, whose
 * geometries differ): one bmnd_hip_pipeline_launch runs them all, and
 * bmnd_hip_pipeline_status checks every descriptor's flags. */
int32_t bmnd_hip_pipeline_create_many(const bmnd_hip_pipeline_desc *descs, int32_t count,
                                      bmnd_hip_pipeline **out);
int32_t bmnd_hip_pipeline_launch(bmnd_hip_pipeline *pipeline, void *stream);
int32_t bmnd_hip_pipeline_status(bmnd_hip_pipeline *pipeline, void *stream);
void bmnd_hip_pipeline_destroy(bmnd_hip_pipeline *pipeline);

/* ---- the reference schedule on the device ----------------------------------
 * Replaces shifts.get_reference_schedule in GENERATED mode
 * (shifts.py:405-423, 519-576, with the shift tables of 229-266): the
 * stage's reference patch origins as patch-grid linear indices, in the
 * reference's order, built on the device (ndim <= 4; counts = shape - block
 * + 1 per axis).  ref_lin (and, unless null, ref_slot: each reference's
 * candidate index slot * schedule_density + phase, from which the host can
 * rebuild its shifted position) need bmnd_hip_schedule_capacity() entries;
 * *nrefs receives the count (the call synchronises the stream to read it). */
int64_t bmnd_hip_schedule_capacity(int32_t ndim, const int64_t *counts_h, const int32_t *block_h,
                                   const int32_t *step_h, double shift_density,
                                   int32_t schedule_density);
int32_t bmnd_hip_schedule_generated(int32_t ndim, const int64_t *counts_h, const int32_t *block_h,
                                    const int32_t *step_h, double shift_density,
                                    int32_t schedule_density, int32_t *ref_lin, int32_t *ref_slot,
                                    int64_t capacity, int64_t *nrefs, void *stream);
/* The shift table of one axis (shifts.py:229-247) into out[64]; returns its
 * length.  Host-only. */
int32_t bmnd_hip_schedule_axis_shifts(int32_t block, int32_t step, int32_t allow_multiple,
                                      double density, int32_t *out);

/* Each batched entry point takes the descriptors twice: inputs_h (host, used
 * for checks and launch sizes) and inputs_d (the same ninputs records in
 * device memory, read by the kernels).  ref_lin holds every reference's
 * patch-grid index within its own input, ref_input its input; sel lists,
 * counts, weights and filtered rows are indexed by the flat reference index
 * as in the single-input calls, and their patch indices are per input.  Each
 * input's result equals that of the single-input call on it alone. */

/* Aggregation tile of a block configuration (tile[3], voxels per axis), for
 * bmnd_hip_input.tile_off: an input has prod(ceil(shape[d] / tile[d])) tiles.
 * (A single-input aggregation picks among three tiles of 512, 256 and 128
 * voxels by timing each once per device, block configuration and geometry;
 * every tile sums each voxel's contributions in the same order, so the
 * estimate is the same bit for bit.  Batches use the first.) */
int32_t bmnd_hip_aggregate_tile(const int32_t *block_h, int32_t *tile);
/* Diagnostic control of that choice, process-wide: index 0..2 forces a tile
 * of the candidates (512, 256, 128 voxels) for later single-input calls, -1
 * (the default) lets the tuner choose. */
int32_t bmnd_hip_aggregate_force_tile(int32_t index);

/* bmnd_hip_blockmatch over a batch (vol: packed volumes); ref_thresholds as
 * in bmnd_hip_blockmatch_thresholds, or NULL for the scalar threshold; modes
 * 7-9 are the standardized distances 1-3 of bmnd_hip_blockmatch_standardized
 * with its null_moments / null_max (else NULL / 0). */
int32_t bmnd_hip_blockmatch_batch(const float *vol, const bmnd_hip_input *inputs_h,
                                  const bmnd_hip_input *inputs_d, int32_t ninputs,
                                  const int32_t *block_h, const int32_t *radius_h,
                                  const int32_t *ref_lin, const int32_t *ref_input,
                                  const float *ref_thresholds, int64_t nrefs, float threshold,
                                  int32_t max_matches, int32_t min_matches, int32_t pow2,
                                  int32_t stabilize, double count_scale, int32_t mode,
                                  int32_t table_max, const float *null_moments, int32_t null_max,
                                  int32_t *sel, int32_t *counts, int32_t *accepted,
                                  void *stream);

/* bmnd_hip_exact_std over a batch (var: packed variance volumes). */
int32_t bmnd_hip_exact_std_batch(const float *var, const bmnd_hip_input *inputs_h,
                                 const bmnd_hip_input *inputs_d, int32_t ninputs,
                                 const int32_t *block_h, const int32_t *radius_h,
                                 const int32_t *sel, const int32_t *counts,
                                 const int32_t *ref_input, int64_t nrefs,
                                 int32_t max_matches, const float *matrices, int32_t k,
                                 float *exact_std, void *stream);

/* bmnd_hip_exact_std_pairs over a batch. */
int32_t bmnd_hip_exact_std_pairs_batch(const float *var, const bmnd_hip_input *inputs_h,
                                       const bmnd_hip_input *inputs_d, int32_t ninputs,
                                       const int32_t *block_h, const int32_t *radius_h,
                                       const int32_t *sel, const int32_t *counts,
                                       const int32_t *ref_input, int64_t nrefs,
                                       int32_t max_matches, const float *matrices, int32_t k,
                                       float *exact_std, void *stream);

/* bmnd_hip_filter_groups over a batch (noisy, pilot, var: packed volumes). */
int32_t bmnd_hip_filter_groups_batch(int32_t wiener, const float *noisy, const float *pilot,
                                     const float *var, const bmnd_hip_input *inputs_h,
                                     const bmnd_hip_input *inputs_d, int32_t ninputs,
                                     const int32_t *block_h, const int32_t *radius_h,
                                     const int32_t *sel, const int32_t *counts,
                                     const int32_t *ref_input, int64_t nrefs,
                                     int32_t max_matches, const float *matrices,
                                     const float *exact_std, float shrink, int32_t k,
                                     double count_scale, float *filt, float *weights,
                                     void *stream);

/* bmnd_hip_reference_map over a batch: ref_index (packed patch grids,
 * num_patches in all) gets the flat reference index at every scheduled
 * position of every input, -1 elsewhere. */
int32_t bmnd_hip_reference_map_batch(const int32_t *ref_lin, const int32_t *ref_input,
                                     int64_t nrefs, const bmnd_hip_input *inputs_d,
                                     int32_t *ref_index, int64_t num_patches, void *stream);

/* bmnd_hip_aggregate over a batch (fallback, est: packed volumes; ref_index
 * from bmnd_hip_reference_map_batch). */
int32_t bmnd_hip_aggregate_batch(const float *filt, const float *weights, const int32_t *sel,
                                 const int32_t *counts, const int32_t *ref_index, int64_t nrefs,
                                 int32_t max_matches, const bmnd_hip_input *inputs_h,
                                 const bmnd_hip_input *inputs_d, int32_t ninputs,
                                 const int32_t *block_h, const int32_t *radius_h,
                                 const float *kaiser, const float *fallback, float *est,
                                 void *stream);

/* Host-only check of the summation plans: numpy's float32 order for x[0..n)
 * (host memory) evaluated on the CPU with the plans and leaf sums the
 * kernels use.  For tests without a GPU. */
float bmnd_hip_np_sum_host(const float *x, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
