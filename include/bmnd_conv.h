/* C-ABI of the learned stage-1 pilot kernels (libbmnd_hip.so).
 *
 * No counterpart in the reference: the reference's stage-1 pilot is the
 * hard-threshold pass (bmndalgo.py:230-633), which the pilot network stands
 * in for at the seam where stage 2 takes its inputs (bmndalgo.py:636-644
 * variance model, 660-667 pilot patches, 714-721 matching).
 *
 * Tensors are device pointers.  Activations are channels-last bf16
 * [B][D][H][W][C] (raw 16-bit words); a 3x3x3 layer's weights are bf16
 * [COUT][KP] with column k = tap * CIN + ci, tap = 9*(dz+1) + 3*(dy+1) +
 * (dx+1), KP = bmnd_conv_kpad(CIN) (padding columns zero).  The 1-channel
 * first and last layers take float32 weights [16][27].  Every call is
 * asynchronous on `stream` and returns 0 or a negative BMND_HIP_E* code
 * (message in bmnd_hip_last_error()).  No reduction uses float atomics:
 * results are independent of tiling.
 */
#ifndef BMND_CONV_H
#define BMND_CONV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 27*cin rounded up to a multiple of 32 */
int64_t bmnd_conv_kpad(int32_t cin);

/* 3x3x3 "same" convolution on bf16 MFMA (FP32 accumulation).  Input: in_a
 * [B][D][H][W][ca] or, with upsample != 0, [B][D/2][H/2][W/2][ca] read as
 * its 2x nearest upsampling; cb > 0 appends in_b [B][D][H][W][cb] as
 * further channels (the decoder's skip concatenation, never materialised).
 * out_float == 0: out = bf16(max(conv + bias, 0)); out_float != 0: out =
 * float conv (no bias, the backward-data pass with flipped weights).  W must
 * be a multiple of 16.  mt: voxel tiles of 16 per wave (1, 2 or 4; W must be
 * a multiple of 16*mt) or 0 for the autotuner, which times every variant
 * once per (layer, geometry, device) on the call's own data, keeps the
 * fastest and reports it in *mt_used (may be null).  Compiled channel
 * configurations (ca+cb -> co): 16->16, 16->32, 32->32, 32->64, 64->64,
 * up(64)+32->32, up(32)+16->16 (bias + ReLU); 16->16, 32->16, 32->32,
 * 64->32, 64->64, 32->96, 16->48 (float out). */
int32_t bmnd_conv3d_forward(const void *in_a, const void *in_b, int32_t ca, int32_t cb,
                            int32_t upsample, const void *weights, const float *bias, int32_t co,
                            void *out, int32_t out_float, int32_t B, int32_t D, int32_t H,
                            int32_t W, int32_t mt, int32_t *mt_used, void *stream);

/* Weight gradient of a layer: grad[co][k] (float [co][KP]) = sum over voxels
 * v of gy[v][co] * X[v + tap(k)][ci(k)], X the layer's input as in
 * bmnd_conv3d_forward, on MFMA per chunk of `chunk` voxels (a multiple of
 * 32) into parts (bmnd_conv3d_wgrad_parts() * co * KP floats), then summed
 * over the chunks in order. */
int64_t bmnd_conv3d_wgrad_parts(int32_t B, int32_t D, int32_t H, int32_t W, int32_t chunk);
int32_t bmnd_conv3d_wgrad(const void *x_a, const void *x_b, int32_t ca, int32_t cb,
                          int32_t upsample, const void *gy, int32_t co, float *parts,
                          int32_t chunk, float *grad, int32_t B, int32_t D, int32_t H, int32_t W,
                          void *stream);

/* First layer (1 -> 16): out = bf16(max(bias + conv(x), 0)), x float. */
int32_t bmnd_conv_in_forward(const float *x, const float *w, const float *bias, void *out,
                             int32_t B, int32_t D, int32_t H, int32_t W, void *stream);
/* Last layer (16 -> 1) with the residual: y = bias + conv(d) + res. */
int32_t bmnd_conv_out_forward(const void *d, const float *w, const float *bias, const float *res,
                              float *y, int32_t B, int32_t D, int32_t H, int32_t W, void *stream);
/* Its input gradient, masked by the ReLU that made d: gd = bf16((d > 0) *
 * convT(gy)). */
int32_t bmnd_conv_out_backward(const float *gy, const float *w, const void *d, void *gd,
                               int32_t B, int32_t D, int32_t H, int32_t W, void *stream);
/* Weight gradients of the 1-channel layers ([16][27] float): last == 0, the
 * first layer (x float input, gb bf16 [..][16] gradient); last != 0, the
 * last layer (xb bf16 [..][16] input, gf float gradient).  parts holds
 * ceil(voxels/chunk) * 432 floats. */
int32_t bmnd_conv_small_wgrad(int32_t last, const float *xf, const void *xb, const float *gf,
                              const void *gb, float *parts, int32_t chunk, float *grad, int32_t B,
                              int32_t D, int32_t H, int32_t W, void *stream);
/* Bias gradient grad[c] = sum_v g[v][c] (gb bf16 [n][C]) or sum_v gf[v]
 * (gb null, C = 1); parts holds ceil(n/chunk) * C floats. */
int32_t bmnd_conv_bias_grad(const void *gb, const float *gf, int32_t C, float *parts,
                            int32_t chunk, float *grad, int64_t n, void *stream);

/* 2x2x2 mean pooling of bf16 [B][D][H][W][C] (even axes). */
int32_t bmnd_conv_pool2(const void *in, void *out, int32_t C, int32_t B, int32_t D, int32_t H,
                        int32_t W, void *stream);
/* Gradient of a skip activation act (C channels at [B][D][H][W]): out =
 * bf16((act > 0) * (gin[v][off + c] + gpool[parent(v)][c] / 8)), gin rows of
 * ctot floats, either term may be null. */
int32_t bmnd_conv_skip_grad(const float *gin, int32_t ctot, int32_t off, const float *gpool,
                            const void *act, void *out, int32_t C, int32_t B, int32_t D, int32_t H,
                            int32_t W, void *stream);
/* Gradient of an upsampled input: out[V][c] = bf16((act[V][c] > 0) * sum of
 * gin[v][c] over the 8 children v of coarse voxel V), c < C, gin rows of
 * ctot floats at [B][D][H][W], out and act at [B][D/2][H/2][W/2][C]. */
int32_t bmnd_conv_up_grad(const float *gin, int32_t ctot, const void *act, void *out, int32_t C,
                          int32_t B, int32_t D, int32_t H, int32_t W, void *stream);
/* out = bf16((act > 0) * g), n elements. */
int32_t bmnd_conv_mask_grad(const float *g, const void *act, void *out, int64_t n, void *stream);
/* Backward-data weights: wt[ci][tap' * co_n + co] = w[co][(26 - tap') *
 * ci_n + ci], shape [ci_n][KP(co_n)]. */
int32_t bmnd_conv_flip_weights(const void *w, int32_t ci, int32_t co, void *wt, void *stream);

/* Squared error over the voxels at least `margin` from every face of their
 * sample: sample_sums[b] (device, B floats) = sample b's sum of squared
 * differences, summed in a fixed order; gy = dL/dy for L = the mean squared
 * error over the batch (log_loss == 0) or L = sum over samples of
 * log(mse_b + eps) (log_loss != 0, equal weight per decibel of each
 * sample).  parts holds B * ceil(D*H*W / 256) floats. */
int32_t bmnd_conv_mse(const float *y, const float *t, float *gy, float *parts, int32_t margin,
                      int32_t B, int32_t D, int32_t H, int32_t W, int32_t log_loss, float eps,
                      float *sample_sums, void *stream);
/* Adam step `step` (from 1) on n float parameters with bias correction;
 * wb (may be null) receives the bf16 copy. */
int32_t bmnd_adam_step(float *p, const float *g, float *m, float *v, void *wb, int64_t n, float lr,
                       float beta1, float beta2, float eps, int32_t step, void *stream);
int32_t bmnd_to_bf16(const float *x, void *out, int64_t n, void *stream);

/* Standardised Anscombe transform (A(x) - center) / spread with A(x) =
 * 2 sqrt(max(x, 0) + 3/8) (inverse == 0), or the algebraic inverse
 * max(((center + spread * x) / 2)^2 - 3/8, 0). */
int32_t bmnd_anscombe(const float *x, float *out, int64_t n, int32_t inverse, float center,
                      float spread, void *stream);
/* sums[0..2] (device doubles) = sum of A(x), sum of A(x)^2 (in a fixed
 * order) and max of x over n counts; part holds 3072 doubles. */
int32_t bmnd_anscombe_moments(const float *x, int64_t n, double *part, double *sums,
                              void *stream);
/* Edge-replicating pad of a float volume (n0, n1, n2) into (m0, m1, m2)
 * (crop == 0), or the leading (n0, n1, n2) box of an (m0, m1, m2) volume
 * (crop != 0). */
int32_t bmnd_pad_edge(const float *in, float *out, int32_t n0, int32_t n1, int32_t n2, int32_t m0,
                      int32_t m1, int32_t m2, int32_t crop, void *stream);
/* Agreement of a pilot with plain pilots of sub-volumes: for crop r (packed
 * crops of c0*c1*c2 voxels with origins[r] in the full volume of n0*n1*n2),
 * block j writes parts[2 * (r * blocks_per_crop + j) + {0, 1}] = (its
 * partial sum of (A(full) - A(crop))^2, A the Anscombe transform, its
 * partial max of crop) over the crop's interior (margin voxels in from every
 * face). */
int32_t bmnd_compare_crops(const float *full, const float *crops, const int32_t *origins,
                           int32_t ncrops, int32_t n0, int32_t n1, int32_t n2, int32_t c0,
                           int32_t c1, int32_t c2, int32_t margin, float *parts,
                           int32_t blocks_per_crop, void *stream);

#ifdef __cplusplus
}
#endif

#endif
