/*
 * CPU oracle for the Poisson bmnd() path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the arithmetic of the reference
 * denoiser (cduhme/bmnd v0.3.0, src/bmnd/) for the default Poisson branch.
 * It is the checker the GPU path is compared against and the CPU baseline
 * timed beside it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (bmnd_amd) never does.
 *
 * Conventions: arrays are C-contiguous; positions and linear indices are
 * int64; the patch grid has counts[d] = shape[d] - block[d] + 1 origins.
 *
 * Pinning (tests/test_oracle_*.py and tests/test_*_cpu.py): every function
 * is checked against the reference's own known answers or properties, a
 * closed form, numpy/scipy, or brute force.  Not fully pinned:
 *   - the last bit of log() in the deviance terms (the reference calls the
 *     platform log through numba; nothing in its tests fixes its ulps):
 *     parity unpinned at that bit;
 *   - oracle_patch_distance with ORACLE_DIST_SSD: its float64 accumulation is
 *     checked on integer-valued known answers only (parity unpinned for
 *     non-integer data; the Poisson path never uses SSD);
 *   - the per-voxel values of the sigma maps (oracle_bmnd_poisson_sigma):
 *     invariants only, parity partially pinned;
 *   - end-to-end outputs: the reference cannot run here, so the pipeline
 *     rests on its components' pins.
 */
#ifndef BMND_ORACLE_H
#define BMND_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_NDIM 4

/* distance codes, blockmatching.py:13-16 */
#define ORACLE_DIST_SSD 0
#define ORACLE_DIST_DEVIANCE 1
#define ORACLE_DIST_PEARSON 2
#define ORACLE_DIST_ANSCOMBE 3

/* ---- host geometry ---------------------------------------------------- */
void oracle_kaiser(int m, double beta, double *out);
void oracle_kaiser_nd(int ndim, const int *block, double beta, float *out);
int oracle_haar_matrix(int n, float *fwd);
void oracle_dct_matrix(int n, float *fwd, float *inv);
int oracle_bior15_matrix(int n, float *fwd, float *inv);
int64_t oracle_schedule_generated(int ndim, const int64_t *counts,
                                  const int *block, const int *step,
                                  double shift_density, int schedule_density,
                                  int64_t *out_abs, int64_t cap);

/* ---- block matching ---------------------------------------------------- */
float oracle_canonical_key(float d);
float oracle_patch_distance(const float *ref, const float *cand, int p,
                            int code, double count_scale);
void oracle_blockmatch(const float *patches_flat, const float *volume,
                       int ndim, const int64_t *vol_shape,
                       const int64_t *counts, const int *block,
                       const int64_t *refs, int64_t nrefs, const int *radius,
                       const float *thresholds, int max_matches,
                       int min_matches, int pow2, int stabilize, int code,
                       double count_scale, int nthreads, int64_t *sel_out,
                       int32_t *count_out, int64_t *accepted_out);

float oracle_standardized_distance(const float *ref, const float *cand, int p, int code,
                                   double count_scale, const float *null_means,
                                   const float *null_vars, int null_max);
void oracle_blockmatch_standardized(const float *patches_flat, const float *volume, int ndim,
                                    const int64_t *vol_shape, const int64_t *counts,
                                    const int *block, const int64_t *refs, int64_t nrefs,
                                    const int *radius, const float *thresholds, int max_matches,
                                    int min_matches, int pow2, int stabilize, int code,
                                    double count_scale, const float *null_means,
                                    const float *null_vars, int null_max, int nthreads,
                                    int64_t *sel_out, int32_t *count_out,
                                    int64_t *accepted_out);

void oracle_blockmatch_ssd(const float *volume, int ndim, const int64_t *vol_shape,
                           const int64_t *counts, const int *block, const int64_t *refs,
                           int64_t nrefs, const int *radius, const float *thresholds,
                           int max_matches, int min_matches, int pow2, const float *noise_ssd,
                           float gamma, int nthreads, int64_t *sel_out, int32_t *count_out,
                           int64_t *accepted_out);

/* stage-1 SSD matching of the Poisson model with the Poisson gamma
 * correction (blockmatching.py:432-443, 1140-1145): float64 corrected
 * distances, ordered against the float32 values the candidate buffers keep */
void oracle_blockmatch_ssd_poisson(const float *volume, int ndim, const int64_t *vol_shape,
                                   const int64_t *counts, const int *block, const int64_t *refs,
                                   int64_t nrefs, const int *radius, const float *thresholds,
                                   int max_matches, int min_matches, int pow2, float gamma,
                                   double count_scale, int nthreads, int64_t *sel_out,
                                   int32_t *count_out, int64_t *accepted_out);

/* ---- group statistics --------------------------------------------------- */
void oracle_gaussian_group_variance(const float *maps, int P, int64_t D, const int64_t *pair_lag,
                                    const float *gmat, int n, int k, float *var_out);
void oracle_ht_gaussian(float *coef, const float *sigma, int64_t n, float lambda_,
                        float *atten);
void oracle_wiener_gaussian(float *coef, const float *ref_coef, const float *sigma, int64_t n,
                            float mu, float *atten, float *spow);
void oracle_apply_transform(float *data, int ndim_total, const int *dims,
                            const float *const *mats);
int oracle_exact_variance(const float *var_volume, int ndim,
                          const int64_t *vol_shape, const int *block, int n,
                          const int64_t *abs_pos, const float *group_mat,
                          const float *const *spatial_mats, int planes,
                          float *var_out);
int oracle_noise_std(const float *var_volume, int ndim,
                     const int64_t *vol_shape, const int *block, int n,
                     const int64_t *abs_pos, const float *group_mat,
                     const float *const *spatial_mats, int k,
                     float *sigma_out);
void oracle_limit_planes(const float *exact_std, const float *approx_std,
                         int n, int p, int k, float *out);
void oracle_patch_weights(const float *atten, const float *sigma, int n,
                          int p, const float *group_inv, double eps,
                          float *w_out);

void oracle_ht_shrink(float *coef, const float *sigma, int64_t n, float lambda_,
                      double count_scale, float *atten);
void oracle_ht_soft(float *coef, const float *sigma, int64_t n, float lambda_, float *atten);
double oracle_np_sum_f64(const double *a, int64_t n);
int oracle_weight_risks(const float *atten, const float *sigma, const float *spow, int n, int p,
                        const float *group_inv, const double *smd, const double *gmd,
                        int model, int windowed, int group_scope, double *risks);
void oracle_mass_correct(int nax, const int *dims, const float *const *mats,
                         const double *const *dirs, const double *O, const float *src_vox,
                         float *coef, double eps);
double oracle_exact_group_risk(const float *var_volume, int ndim, const int64_t *vol_shape,
                               const int *block, int n, const int64_t *abs_pos,
                               const float *group_mat, const float *group_inv,
                               const float *const *spatial_mats,
                               const float *const *spatial_inv, const float *window,
                               int planes, const float *gain, const float *sigma,
                               const float *spow);
void oracle_weights_ex(const float *atten, const float *sigma, const float *spow, int n, int p,
                       const float *group_inv, const double *smd, const double *gmd, int model,
                       int windowed, int group_scope, double eps, float *w_out);
void oracle_wiener_shrink(float *coef, const float *ref_coef, const float *sigma,
                          int64_t n, float mu, float *atten);
void oracle_wiener_shrink_mode(float *coef, const float *ref_coef, const float *sigma,
                               int64_t n, float mu, int mode, float *atten);
void oracle_scatter_add(float *num, float *den, int ndim, const int64_t *vol_shape,
                        const int *block, const float *kaiser, const float *filt,
                        const float *weights, const int64_t *lins, int n);

/* numpy's float32 pairwise sum / mean of a contiguous array (np.sum, np.mean) */
float oracle_np_sum_f32(const float *a, int64_t n);
float oracle_np_mean_f32(const float *a, int64_t n);
/* scatter with the sigma-weighted sum; group_sigma (n x P) or, when NULL,
 * the scalar sigma for every element */
void oracle_scatter_add_sigma(float *num, float *den, float *sigma_num, int ndim,
                              const int64_t *vol_shape, const int *block,
                              const float *kaiser, const float *filt,
                              const float *weights, const float *group_sigma,
                              float sigma, const int64_t *lins, int n);
/* Poisson global sigma of a variance volume with a positive floor */
double oracle_poisson_global_sigma(const float *var_vol, int64_t n, double floor_);

/* ---- full pipeline ----------------------------------------------------- */
/* Entry order inside one union voxel of the exact variance: 0 (default)
 * increasing entry index; 1 numba's quicksort argsort order, as the
 * reference's np.argsort of the union labels leaves it (variance.py:804).
 * Process-wide; set it before a run. */
void oracle_set_exact_order(int quicksort);
/* R[0..n) = argsort(A) in numba's default (quicksort) order. */
void oracle_numba_argsort(const int64_t *A, int64_t n, int32_t *R);

typedef struct {
  int ndim;
  int64_t shape[ORACLE_MAX_NDIM];
  int ht_block[ORACLE_MAX_NDIM], ht_step[ORACLE_MAX_NDIM],
      ht_radius[ORACLE_MAX_NDIM];
  int wi_block[ORACLE_MAX_NDIM], wi_step[ORACLE_MAX_NDIM],
      wi_radius[ORACLE_MAX_NDIM];
  int ht_max, ht_min, wi_max, wi_min;
  float ht_threshold, wi_threshold;
  float ht_lambda, wi_mu;
  double ht_beta, wi_beta;
  double shift_density;
  int schedule_density;
  int k;
  double count_scale;
  double variance_floor;
  int nthreads;
  /* optional spatial matrices per stage: for each axis d the analysis
   * matrix (b_d x b_d), then for each axis the synthesis matrix; NULL builds
   * them here (bior1.5 for HT, DCT for Wiener) */
  const float *ht_spatial;
  const float *wi_spatial;
  /* non-default Poisson options (profiles.py:42-72): Wiener matching on the
   * observation instead of the pilot (poisson_wiener_match_source), the
   * Wiener variance model from the observation (poisson_variance_source_wiener),
   * and per-reference HT thresholds (poisson_ht_match_policy
   * "reference_finite_count", one per HT reference in schedule order; NULL
   * for the scalar ht_threshold) */
  int wi_match_observation;
  int wi_var_observation;
  const float *ht_thresholds;
  /* distance of both stages (ORACLE_DIST_*; 0 means the deviance) and the
   * candidate-standardized stage-1 distance with its float32 null tables */
  int distance;
  int ht_standardize;
  const float *null_means;
  const float *null_vars;
  int null_max;
  /* Wiener gain signal power: 0 variance-scaled, 1 classic, 2 noise-floor */
  int wi_gain_mode;
  /* HT soft thresholding (ht_use_soft_thresholding) and each stage's
   * aggregation weights: model 0 variance, 1 classic, 2 risk (Wiener);
   * the windowed-synthesis domain with its spatial metric diagonal smd [P]
   * and group metric diagonals gmd (n entries at offset n(n-1)/2); the group
   * scope */
  int ht_soft;
  int ht_wmodel, wi_wmodel;
  int ht_windowed, wi_windowed;
  int ht_group_scope, wi_group_scope;
  const double *ht_smd, *wi_smd;
  const double *ht_gmd, *wi_gmd;
  /* group-mass conservation (poisson_group_mass_conservation): bit 0 HT,
   * bit 1 Wiener, with each stage's constant-synthesis directions: mdir n
   * entries for each n at n(n-1)/2, sdir the block's axes one after another */
  int mass_mode;
  const double *ht_mdir, *wi_mdir;
  const double *ht_sdir, *wi_sdir;
  /* the stationary Gaussian noise model (noise_model "gaussian",
   * bmndalgo.py:203-228, 291-316, 404-474, 826-870, variance.py:452-553,
   * 1791-2106): SSD matching with thresholds ht_match_d / wi_match_d times
   * the squared value range of the matched volume; the HT search-box noise
   * SSD scaled by ht_gamma (NULL: no correction); each stage's
   * cross-covariance maps (P x D over the nf domain cov_domain) for the
   * group noise variance with k exact planes; global_sigma fills the sigma
   * maps */
  int gaussian;
  double ht_match_d, wi_match_d;
  const float *ht_cov, *wi_cov;
  int cov_domain[ORACLE_MAX_NDIM];
  const float *ht_noise_ssd;
  float ht_gamma;
  double global_sigma;
  /* the SSD distance under the Poisson model (blockmatch_distance "ssd",
   * bmndalgo.py:285-322, 721-741): thresholds ht_match_d / wi_match_d times
   * the squared value range of the matched volume as above, stage 1 with the
   * Poisson gamma correction (ht_gamma, count_scale; 0: none), stage 2 plain
   * SSD; everything else is the Poisson path */
  int poisson_ssd;
} oracle_params;

/* Runs both stages.  pilot may be NULL.  The match lists of each stage can
 * be captured: *_sel has room for nrefs*max entries, *_cnt for nrefs; pass
 * NULL to skip.  Returns 0, or a negative code on invalid arguments. */
int oracle_bmnd_poisson(const float *vol, const oracle_params *prm,
                        float *out, float *pilot, int64_t *ht_sel,
                        int32_t *ht_cnt, int64_t ht_cap, int64_t *wi_sel,
                        int32_t *wi_cnt, int64_t wi_cap);
/* As oracle_bmnd_poisson, with the reference's return_sigma_map outputs:
 * the per-stage sigma maps (NULL to skip) and the global sigma. */
int oracle_bmnd_poisson_sigma(const float *vol, const oracle_params *prm,
                              float *out, float *pilot, int64_t *ht_sel,
                              int32_t *ht_cnt, int64_t ht_cap, int64_t *wi_sel,
                              int32_t *wi_cnt, int64_t wi_cap, float *ht_sigma_map,
                              float *wi_sigma_map, double *global_sigma);

#ifdef __cplusplus
}
#endif
#endif
