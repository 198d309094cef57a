// MI355X (gfx950) kernels of the learned stage-1 pilot: a 3-level 3-D U-Net
// (channels 16/32/64, 3x3x3 "same" convolutions, ReLU) in BF16 with FP32
// accumulation, trained and run on the device.
//
// Layout: activations are channels-last bf16, [B][D][H][W][C]; a layer's
// weights are bf16 [COUT][KP] with k = tap * CIN + ci (tap = 9*dz + 3*dy + dx
// over the 3x3x3 neighbourhood, offsets -1..1) and KP = 27*CIN rounded up to
// 32 (zero columns).  Master weights, gradients and optimiser moments are
// float32 in the same [COUT][KP] layout.
//
// Kernels
//   k_conv_mfma      implicit-GEMM convolution on v_mfma_f32_16x16x32_bf16:
//                    M = 16 output voxels along x per tile (MT tiles per
//                    wave), N = output channels, K = 27*CIN.  The decoder's
//                    2x nearest upsampling and skip concatenation are read
//                    on the fly ([up(A) | B] as the input, never stored).
//                    Also the backward-data pass, with flipped weights.
//   k_conv_wgrad     weight gradients on the same MFMA: M = COUT, N = k,
//                    K = voxels, per voxel chunk (fixed-order reduction after)
//   k_conv_in / k_conv_out and their gradients: the 1-channel first and last
//                    layers on VALU (27*16 MACs per voxel)
//   pooling, gradient routing, MSE loss, Adam, Anscombe transforms
// No float atomics: every reduction has a fixed order, so results do not
// depend on the launch shape the autotuner picks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bmnd_conv.h"
#include "bmnd_internal.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

int cfail(int code, const std::string &msg) { return bmnd_internal_fail(code, msg.c_str()); }

#define CONV_TRY(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) return cfail(-2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (inputs are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

struct Geom {
  int B, D, H, W;  // full-resolution voxels per axis (the output's)
};

struct Vox {
  int b, z, y, x;
};

// voxel index -> (sample, z, y, x); every conv buffer holds < 2^31 voxels,
// so the divisions are 32-bit
__device__ __forceinline__ Vox decode(uint32_t v, const Geom &g) {
  Vox r;
  r.x = (int)(v % (uint32_t)g.W);
  v /= (uint32_t)g.W;
  r.y = (int)(v % (uint32_t)g.H);
  v /= (uint32_t)g.H;
  r.z = (int)(v % (uint32_t)g.D);
  r.b = (int)(v / (uint32_t)g.D);
  return r;
}

constexpr int kTaps = 27;
__host__ __device__ constexpr int kpad(int cin) { return (kTaps * cin + 31) / 32 * 32; }

// ---------------------------------------------------------------------------
// implicit-GEMM convolution
// ---------------------------------------------------------------------------
struct ConvArgs {
  const uint16_t *in_a;  // [B][Da][Ha][Wa][CA]; Da = D/2 etc. when UP
  const uint16_t *in_b;  // [B][D][H][W][CB] (CB > 0)
  const uint16_t *w;     // [CO][KP]
  const float *bias;     // [CO] (OUTM 0)
  void *out;             // bf16 (OUTM 0) or float (OUTM 1) [B][D][H][W][CO]
  Geom g;
  int64_t segments;      // B*D*H*(W / (16*MT))
};

// Each wave computes MT tiles of 16 voxels along x (16*MT consecutive
// voxels of one row) for all CO channels.  Lane l holds row l%16 of each
// A tile and column l%16 of each B tile, with k = 32*c + 8*(l/16) + j,
// j = 0..7; the accumulators hold rows 4*(l/16) + r, r = 0..3.
// OUTM 0: out = bf16(max(acc + bias, 0)); OUTM 1: out = float(acc).
template <int CA, int CB, int CO, int MT, bool UP, int OUTM>
__global__ __launch_bounds__(256) void k_conv_mfma(ConvArgs a) {
  constexpr int CIN = CA + CB;
  constexpr int KP = kpad(CIN);
  constexpr int NT = CO / 16;
  static_assert(CIN % 8 == 0 && CO % 16 == 0, "channel counts");
  const int lane = threadIdx.x & 63;
  const int64_t seg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (seg >= a.segments) return;
  const int W = a.g.W, H = a.g.H, D = a.g.D;
  const uint32_t spr = (uint32_t)(W / (16 * MT));  // segments per row
  const int xs = (int)((uint32_t)seg % spr) * 16 * MT;
  uint32_t row = (uint32_t)seg / spr;
  const int y = (int)(row % (uint32_t)H);
  row /= (uint32_t)H;
  const int z = (int)(row % (uint32_t)D);
  const int b = (int)(row / (uint32_t)D);
  const int i = lane & 15, grp = lane >> 4;
  const int Da = UP ? D / 2 : D, Ha = UP ? H / 2 : H, Wa = UP ? W / 2 : W;

  floatx4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int c = 0; c < KP / 32; ++c) {
    const int k0 = 32 * c + 8 * grp;
    const int tap = k0 / CIN, ci = k0 % CIN;
    const bool tap_ok = tap < kTaps;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int zz = z + dz, yy = y + dy;
    const bool zy_ok = tap_ok && zz >= 0 && zz < D && yy >= 0 && yy < H;
    bf16x8 bfr[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n)
      bfr[n] = as_bf16x8(*reinterpret_cast<const uint4 *>(a.w + (size_t)(16 * n + i) * KP + k0));
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int xx = xs + 16 * m + i + dx;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (zy_ok && xx >= 0 && xx < W) {
        if (CB == 0 || ci < CA) {
          const int64_t vox = UP ? (((int64_t)b * Da + (zz >> 1)) * Ha + (yy >> 1)) * Wa + (xx >> 1)
                                 : (((int64_t)b * D + zz) * H + yy) * W + xx;
          v = *reinterpret_cast<const uint4 *>(a.in_a + vox * CA + ci);
        } else {
          const int64_t vox = (((int64_t)b * D + zz) * H + yy) * W + xx;
          v = *reinterpret_cast<const uint4 *>(a.in_b + vox * CB + (ci - CA));
        }
      }
      const bf16x8 afr = as_bf16x8(v);
#pragma unroll
      for (int n = 0; n < NT; ++n)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr, bfr[n], acc[m][n], 0, 0, 0);
    }
  }
  const int64_t rowbase = (((int64_t)b * D + z) * H + y) * W;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int co = 16 * n + i;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t vox = rowbase + xs + 16 * m + 4 * grp + r;
        if constexpr (OUTM == 0) {
          const float v = fmaxf(acc[m][n][r] + a.bias[co], 0.0f);
          static_cast<uint16_t *>(a.out)[vox * CO + co] = f2bf(v);
        } else {
          static_cast<float *>(a.out)[vox * CO + co] = acc[m][n][r];
        }
      }
    }
}

// ---------------------------------------------------------------------------
// weight gradients: part[chunk][co][k] = sum over the chunk's voxels v of
// gy[v][co] * X[v + tap(k)][ci(k)], X the layer's (virtual) input
// ---------------------------------------------------------------------------
struct WgradArgs {
  const uint16_t *x_a, *x_b;  // as ConvArgs in_a / in_b
  const uint16_t *gy;         // [B][D][H][W][CO]
  float *part;                // [nchunks][CO][KP]
  Geom g;
  int64_t voxels;             // B*D*H*W
  int chunk;                  // voxels per chunk (multiple of 32)
};

template <int CA, int CB, bool UP>
__device__ __forceinline__ float wg_load_x(const WgradArgs &a, int b, int z, int y, int x,
                                           int ci) {
  const int D = a.g.D, H = a.g.H, W = a.g.W;
  if (z < 0 || z >= D || y < 0 || y >= H || x < 0 || x >= W) return 0.0f;
  if (CB == 0 || ci < CA) {
    const int64_t vox = UP ? (((int64_t)b * (D / 2) + (z >> 1)) * (H / 2) + (y >> 1)) * (W / 2) +
                                 (x >> 1)
                           : (((int64_t)b * D + z) * H + y) * W + x;
    return bf2f(a.x_a[vox * CA + ci]);
  }
  const int64_t vox = (((int64_t)b * D + z) * H + y) * W + x;
  return bf2f(a.x_b[vox * CB + (ci - CA)]);
}

// One wave per (chunk, group of 4 k-tiles); M tiles cover CO.  A = gy^T
// (rows co, k = voxels), B = the shifted input (k = voxels, columns k).
template <int CA, int CB, int CO, bool UP>
__global__ __launch_bounds__(256) void k_conv_wgrad(WgradArgs a) {
  constexpr int CIN = CA + CB;
  constexpr int KP = kpad(CIN);
  constexpr int MTL = CO / 16;
  constexpr int NTL = 4;
  constexpr int KT = KP / 16;                 // k tiles
  constexpr int KG = (KT + NTL - 1) / NTL;    // groups of NTL k tiles
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int kg = wave % KG;
  const int chunk = wave / KG;
  const int64_t v0 = (int64_t)chunk * a.chunk;
  if (v0 >= a.voxels) return;
  const int64_t v1 = min(v0 + (int64_t)a.chunk, a.voxels);
  const int i = lane & 15, grp = lane >> 4;

  // this lane's B columns: k = 16*(NTL*kg + n) + i
  int tapz[NTL], tapy[NTL], tapx[NTL], cis[NTL];
  bool kok[NTL];
#pragma unroll
  for (int n = 0; n < NTL; ++n) {
    const int k = 16 * (NTL * kg + n) + i;
    const int tap = k / CIN;
    kok[n] = (16 * (NTL * kg + n) < KP) && tap < kTaps;
    cis[n] = k % CIN;
    tapz[n] = tap / 9 - 1;
    tapy[n] = (tap / 3) % 3 - 1;
    tapx[n] = tap % 3 - 1;
  }
  floatx4 acc[MTL][NTL];
#pragma unroll
  for (int m = 0; m < MTL; ++m)
#pragma unroll
    for (int n = 0; n < NTL; ++n) acc[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int64_t vs = v0; vs < v1; vs += 32) {
    // this lane's 8 voxels: vs + 8*grp + j
    uint16_t av[MTL][8];
    float bv[NTL][8];
    // W % 16 == 0 and vs % 32 == 0: the 8 voxels share one row (x = x0 + j),
    // and all or none of them lie before v1 (a multiple of 16)
    const int64_t vb = vs + 8 * grp;
    const bool vok = vb < v1;
    const Vox o = decode((uint32_t)vb, a.g);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t v = vb + j;
#pragma unroll
      for (int m = 0; m < MTL; ++m) av[m][j] = vok ? a.gy[v * CO + 16 * m + i] : (uint16_t)0;
#pragma unroll
      for (int n = 0; n < NTL; ++n)
        bv[n][j] = (vok && kok[n]) ? wg_load_x<CA, CB, UP>(a, o.b, o.z + tapz[n], o.y + tapy[n],
                                                            o.x + j + tapx[n], cis[n])
                                   : 0.0f;
    }
    bf16x8 afr[MTL], bfr[NTL];
#pragma unroll
    for (int m = 0; m < MTL; ++m) {
      uint4 u;
      u.x = av[m][0] | ((uint32_t)av[m][1] << 16);
      u.y = av[m][2] | ((uint32_t)av[m][3] << 16);
      u.z = av[m][4] | ((uint32_t)av[m][5] << 16);
      u.w = av[m][6] | ((uint32_t)av[m][7] << 16);
      afr[m] = as_bf16x8(u);
    }
#pragma unroll
    for (int n = 0; n < NTL; ++n) {
      // exact: the values are bf16 already
      uint4 u;
      u.x = (__float_as_uint(bv[n][0]) >> 16) | (__float_as_uint(bv[n][1]) & 0xffff0000u);
      u.y = (__float_as_uint(bv[n][2]) >> 16) | (__float_as_uint(bv[n][3]) & 0xffff0000u);
      u.z = (__float_as_uint(bv[n][4]) >> 16) | (__float_as_uint(bv[n][5]) & 0xffff0000u);
      u.w = (__float_as_uint(bv[n][6]) >> 16) | (__float_as_uint(bv[n][7]) & 0xffff0000u);
      bfr[n] = as_bf16x8(u);
    }
#pragma unroll
    for (int m = 0; m < MTL; ++m)
#pragma unroll
      for (int n = 0; n < NTL; ++n)
        acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afr[m], bfr[n], acc[m][n], 0, 0, 0);
  }
  float *part = a.part + (size_t)chunk * CO * KP;
#pragma unroll
  for (int m = 0; m < MTL; ++m)
#pragma unroll
    for (int n = 0; n < NTL; ++n) {
      const int kt = NTL * kg + n;
      if (16 * kt >= KP) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = 16 * m + 4 * grp + r;
        part[(size_t)co * KP + 16 * kt + i] = acc[m][n][r];
      }
    }
}

// ---------------------------------------------------------------------------
// 1-channel layers (VALU)
// ---------------------------------------------------------------------------
// a1[v][co] = bf16(max(b[co] + sum_tap w[co][tap] x[v + tap], 0)), CO = 16
__global__ void k_conv_in(const float *__restrict__ x, const float *__restrict__ w,
                          const float *__restrict__ bias, uint16_t *__restrict__ out, Geom g,
                          int64_t n) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= n) return;
  const Vox o = decode((uint32_t)v, g);
  const int x0 = o.x, y0 = o.y, z0 = o.z, b = o.b;
  float nb[kTaps];
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int zz = z0 + tap / 9 - 1, yy = y0 + (tap / 3) % 3 - 1, xx = x0 + tap % 3 - 1;
    nb[tap] = (zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
                  ? x[(((int64_t)b * g.D + zz) * g.H + yy) * g.W + xx]
                  : 0.0f;
  }
  uint32_t packed[8];
#pragma unroll
  for (int co = 0; co < 16; ++co) {
    float s = bias[co];
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) s = fmaf(w[co * kTaps + tap], nb[tap], s);
    const uint16_t h = f2bf(fmaxf(s, 0.0f));
    if (co & 1) packed[co >> 1] |= (uint32_t)h << 16;
    else packed[co >> 1] = h;
  }
  uint4 *dst = reinterpret_cast<uint4 *>(out + v * 16);
  dst[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
  dst[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
}

// y[v] = b + sum_tap sum_ci w[ci][tap] d[v + tap][ci] + res[v], CI = 16
__global__ void k_conv_out(const uint16_t *__restrict__ d, const float *__restrict__ w,
                           const float *__restrict__ bias, const float *__restrict__ res,
                           float *__restrict__ y, Geom g, int64_t n) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= n) return;
  const Vox o = decode((uint32_t)v, g);
  const int x0 = o.x, y0 = o.y, z0 = o.z, b = o.b;
  float s = bias[0];
  for (int tap = 0; tap < kTaps; ++tap) {
    const int zz = z0 + tap / 9 - 1, yy = y0 + (tap / 3) % 3 - 1, xx = x0 + tap % 3 - 1;
    if (zz < 0 || zz >= g.D || yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) continue;
    const uint4 *p = reinterpret_cast<const uint4 *>(
        d + ((((int64_t)b * g.D + zz) * g.H + yy) * g.W + xx) * 16);
    const uint4 lo = p[0], hi = p[1];
    const uint32_t words[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
    for (int ci = 0; ci < 16; ++ci) {
      const uint16_t h = (uint16_t)(words[ci >> 1] >> (16 * (ci & 1)));
      s = fmaf(w[ci * kTaps + tap], bf2f(h), s);
    }
  }
  y[v] = s + res[v];
}

// gd[v][ci] = bf16((d[v][ci] > 0) * sum_tap w[ci][tap] gy[v - tap])  (CI = 16)
__global__ void k_conv_out_bwd(const float *__restrict__ gy, const float *__restrict__ w,
                               const uint16_t *__restrict__ d, uint16_t *__restrict__ gd, Geom g,
                               int64_t n) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= n) return;
  const Vox o = decode((uint32_t)v, g);
  const int x0 = o.x, y0 = o.y, z0 = o.z, b = o.b;
  float nb[kTaps];
#pragma unroll
  for (int tap = 0; tap < kTaps; ++tap) {
    const int zz = z0 - (tap / 9 - 1), yy = y0 - ((tap / 3) % 3 - 1), xx = x0 - (tap % 3 - 1);
    nb[tap] = (zz >= 0 && zz < g.D && yy >= 0 && yy < g.H && xx >= 0 && xx < g.W)
                  ? gy[(((int64_t)b * g.D + zz) * g.H + yy) * g.W + xx]
                  : 0.0f;
  }
  for (int ci = 0; ci < 16; ++ci) {
    float s = 0.0f;
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap) s = fmaf(w[ci * kTaps + tap], nb[tap], s);
    gd[v * 16 + ci] = bf2f(d[v * 16 + ci]) > 0.0f ? f2bf(s) : (uint16_t)0;
  }
}

// weight gradients of the 1-channel layers, per chunk of voxels: the first
// layer (MODE 0): part[chunk][co][tap] = sum_v g[v][co] x[v + tap] (CO = 16);
// the last (MODE 1): part[chunk][ci][tap] = sum_v gy[v] d[v + tap][ci].
// One thread per (channel, tap) = 432 outputs; the chunk's voxels in order.
template <int MODE>
__global__ void k_small_wgrad(const float *__restrict__ xf, const uint16_t *__restrict__ xb,
                              const float *__restrict__ gf, const uint16_t *__restrict__ gb,
                              float *__restrict__ part, Geom g, int64_t n, int chunk) {
  const int o = threadIdx.x + (blockIdx.y & 1) * 256;
  const int c = blockIdx.x;
  if (o >= 16 * kTaps) return;
  const int ch = o / kTaps, tap = o % kTaps;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int64_t v0 = (int64_t)c * chunk, v1 = min(v0 + (int64_t)chunk, n);
  float s = 0.0f;
  for (int64_t v = v0; v < v1; ++v) {
    const Vox o = decode((uint32_t)v, g);
    const int x0 = o.x, y0 = o.y, z0 = o.z, b = o.b;
    const int zz = z0 + dz, yy = y0 + dy, xx = x0 + dx;
    if (zz < 0 || zz >= g.D || yy < 0 || yy >= g.H || xx < 0 || xx >= g.W) continue;
    const int64_t u = (((int64_t)b * g.D + zz) * g.H + yy) * g.W + xx;
    if (MODE == 0) s = fmaf(bf2f(gb[v * 16 + ch]), xf[u], s);
    else s = fmaf(gf[v], bf2f(xb[u * 16 + ch]), s);
  }
  part[(size_t)c * 16 * kTaps + o] = s;
}

// bias gradients per chunk: part[chunk][c] = sum_v g[v][c] (bf16 g) or
// sum_v gy[v] (float, C = 1)
__global__ void k_bias_grad(const uint16_t *__restrict__ gb, const float *__restrict__ gf, int C,
                            float *__restrict__ part, int64_t n, int chunk) {
  const int c = threadIdx.x;
  if (c >= C) return;
  const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = min(v0 + (int64_t)chunk, n);
  float s = 0.0f;
  for (int64_t v = v0; v < v1; ++v) s += gb ? bf2f(gb[v * C + c]) : gf[v];
  part[(size_t)blockIdx.x * C + c] = s;
}

// out[i] = sum_c part[c][i], chunks in order
__global__ void k_reduce_parts(const float *__restrict__ part, int nparts, int64_t len,
                               float *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.0f;
  for (int c = 0; c < nparts; ++c) s += part[(size_t)c * len + i];
  out[i] = s;
}

// ---------------------------------------------------------------------------
// pooling and gradient routing
// ---------------------------------------------------------------------------
// 2x2x2 mean: out [B][D/2][H/2][W/2][C] from in [B][D][H][W][C]
__global__ void k_pool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, Geom g,
                        int C, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c = (int)(e % C);
  const Vox o = decode((uint32_t)(e / C), Geom{g.B, g.D / 2, g.H / 2, g.W / 2});
  const int x = o.x, y = o.y, z = o.z, b = o.b;
  float s = 0.0f;
  for (int q = 0; q < 8; ++q) {
    const int64_t v = (((int64_t)b * g.D + 2 * z + (q >> 2)) * g.H + 2 * y + ((q >> 1) & 1)) * g.W +
                      2 * x + (q & 1);
    s += bf2f(in[v * C + c]);
  }
  out[e] = f2bf(0.125f * s);
}

// gradient of a level's skip activation s (C channels, full res):
// g = (skip part of the decoder's input gradient, channel offset off in a
// row of ctot) + (pool gradient / 8 at the parent voxel), masked by s > 0.
__global__ void k_skip_grad(const float *__restrict__ gin, int ctot, int off,
                            const float *__restrict__ gpool, const uint16_t *__restrict__ act,
                            uint16_t *__restrict__ out, Geom g, int C, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c = (int)(e % C);
  const int64_t v = e / C;
  float s = gin ? gin[v * ctot + off + c] : 0.0f;
  if (gpool) {
    const Vox o = decode((uint32_t)v, g);
    const int64_t p =
        (((int64_t)o.b * (g.D / 2) + o.z / 2) * (g.H / 2) + o.y / 2) * (g.W / 2) + o.x / 2;
    s += 0.125f * gpool[p * C + c];
  }
  out[e] = bf2f(act[e]) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// gradient of the upsampled part of a decoder input: channels [0, C) of
// gin (full res, rows of ctot) summed over each coarse voxel's 8 children,
// masked by the coarse activation > 0; out coarse [B][D/2][H/2][W/2][C]
__global__ void k_up_grad(const float *__restrict__ gin, int ctot, const uint16_t *__restrict__ act,
                          uint16_t *__restrict__ out, Geom g, int C, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c = (int)(e % C);
  const Vox o = decode((uint32_t)(e / C), Geom{g.B, g.D / 2, g.H / 2, g.W / 2});
  const int x = o.x, y = o.y, z = o.z, b = o.b;
  float s = 0.0f;
  for (int q = 0; q < 8; ++q) {
    const int64_t v = (((int64_t)b * g.D + 2 * z + (q >> 2)) * g.H + 2 * y + ((q >> 1) & 1)) * g.W +
                      2 * x + (q & 1);
    s += gin[v * ctot + c];
  }
  out[e] = bf2f(act[e]) > 0.0f ? f2bf(s) : (uint16_t)0;
}

// out = bf16(g * (act > 0)), or (pooled == true) the gradient of a pooled
// input that feeds nothing else: g is then float on the coarse grid... (not
// needed); plain masked conversion
__global__ void k_mask_grad(const float *__restrict__ g, const uint16_t *__restrict__ act,
                            uint16_t *__restrict__ out, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  out[e] = bf2f(act[e]) > 0.0f ? f2bf(g[e]) : (uint16_t)0;
}

// W'[ci][tap' * CO + co] = W[co][(26 - tap') * CI + ci]: the backward-data
// weights (flipped taps, swapped channels), zero padding columns
__global__ void k_flip_weights(const uint16_t *__restrict__ w, int ci_n, int co_n,
                               uint16_t *__restrict__ wt) {
  const int kpo = kpad(co_n);
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= (int64_t)ci_n * kpo) return;
  const int ci = (int)(e / kpo), k = (int)(e % kpo);
  uint16_t v = 0;
  if (k < kTaps * co_n) {
    const int tp = k / co_n, co = k % co_n;
    v = w[(size_t)co * kpad(ci_n) + (kTaps - 1 - tp) * ci_n + ci];
  }
  wt[e] = v;
}

// ---------------------------------------------------------------------------
// loss, optimiser, transforms
// ---------------------------------------------------------------------------
// squared error over voxels at least `margin` from every face of their
// sample: one grid row (blockIdx.y) per sample, part[b][block] = the block's
// sum of squares, gy = 2 (y - t) * gscale inside, 0 outside
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ tgt, float *gy,
                      float *__restrict__ part, Geom g, int margin, float gscale) {
  __shared__ float red[256];
  const int64_t per = (int64_t)g.D * g.H * g.W;
  const int64_t local = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  float sq = 0.0f;
  if (local < per) {
    const int64_t v = (int64_t)b * per + local;
    const Vox o = decode((uint32_t)local, Geom{1, g.D, g.H, g.W});
    const bool in = o.x >= margin && o.x < g.W - margin && o.y >= margin && o.y < g.H - margin &&
                    o.z >= margin && o.z < g.D - margin;
    const float d = in ? y[v] - tgt[v] : 0.0f;
    sq = d * d;
    gy[v] = 2.0f * d * gscale;
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(size_t)b * gridDim.x + blockIdx.x] = red[0];
}

// per-sample sums of the block partials, in order
__global__ void k_sample_sums(const float *__restrict__ part, int nblk, int B,
                              float *__restrict__ sums) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.0f;
  for (int j = 0; j < nblk; ++j) s += part[(size_t)b * nblk + j];
  sums[b] = s;
}

// log loss: gy of sample b divided by (its mean squared error + eps)
__global__ void k_log_scale(float *__restrict__ gy, const float *__restrict__ sums, int64_t per,
                            float inv_inside, float eps, int64_t n) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int b = (int)(v / per);
  gy[v] *= 1.0f / (sums[b] * inv_inside + eps);
}

// Adam (decoupled step count t, bias-corrected); p float master weights, wb
// the bf16 copy (may be null)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ wb, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  const float upd = lr * (mm * c1) / (sqrtf(vv * c2) + eps);
  const float np_ = p[e] - upd;
  p[e] = np_;
  if (wb) wb[e] = f2bf(np_);
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ out, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e < n) out[e] = f2bf(x[e]);
}

// Anscombe transform of counts, standardised: (A(x) - center) / spread, and
// its algebraic inverse A^-1(center + spread * x)
__global__ void k_anscombe(const float *__restrict__ x, float *__restrict__ out, int64_t n,
                           int inverse, float center, float spread) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n) return;
  if (!inverse) {
    out[e] = (2.0f * sqrtf(fmaxf(x[e], 0.0f) + 0.375f) - center) / spread;
  } else {
    const float h = 0.5f * fmaf(spread, x[e], center);
    out[e] = fmaxf(h * h - 0.375f, 0.0f);
  }
}

// block partial sums of A(x) and A(x)^2 in float64 and the block maximum of
// x (grid-stride, fixed order), reduced in order by k_moments_sum
__global__ void k_moments(const float *__restrict__ x, int64_t n, double *__restrict__ part) {
  __shared__ double r1[256], r2[256], r3[256];
  double s1 = 0.0, s2 = 0.0, mx = -INFINITY;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double a = 2.0 * sqrt(fmax((double)x[e], 0.0) + 0.375);
    s1 += a;
    s2 += a * a;
    mx = fmax(mx, (double)x[e]);
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  r3[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      r1[threadIdx.x] += r1[threadIdx.x + k];
      r2[threadIdx.x] += r2[threadIdx.x + k];
      r3[threadIdx.x] = fmax(r3[threadIdx.x], r3[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = r1[0];
    part[3 * blockIdx.x + 1] = r2[0];
    part[3 * blockIdx.x + 2] = r3[0];
  }
}

__global__ void k_moments_sum(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double s1 = 0.0, s2 = 0.0, mx = -INFINITY;
  for (int j = 0; j < nblk; ++j) {
    s1 += part[3 * j];
    s2 += part[3 * j + 1];
    mx = fmax(mx, part[3 * j + 2]);
  }
  out[0] = s1;
  out[1] = s2;
  out[2] = mx;
}

// agreement of a full-volume pilot with plain pilots of sub-volumes: for
// crop r (dims c, origin org[r] in the full volume of dims n), block j
// writes part[(r * blocks_per_crop + j) * 2 + {0, 1}] = (sum of (A(full) -
// A(crop))^2 with A the Anscombe transform, max of crop) over its share of
// the crop's interior (margin voxels in from every face)
struct CompareArgs {
  const float *full;
  const float *crops;  // packed crops, each c0*c1*c2
  const int32_t *org;  // [ncrops][3]
  int n0, n1, n2;      // full volume
  int c0, c1, c2;      // crop
  int margin;
  float *part;         // [ncrops][blocks_per_crop][2]
  int blocks_per_crop;
};

__global__ void k_compare_crops(CompareArgs a) {
  __shared__ float red[256], top[256];
  const int r = blockIdx.y;
  const int i0 = a.c0 - 2 * a.margin, i1 = a.c1 - 2 * a.margin, i2 = a.c2 - 2 * a.margin;
  const int64_t nin = (int64_t)i0 * i1 * i2;
  float s = 0.0f, mx = 0.0f;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nin;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(e % i2) + a.margin;
    const int y = (int)((e / i2) % i1) + a.margin;
    const int z = (int)(e / ((int64_t)i2 * i1)) + a.margin;
    const float p = a.crops[(size_t)r * a.c0 * a.c1 * a.c2 + ((int64_t)z * a.c1 + y) * a.c2 + x];
    const int fz = a.org[3 * r] + z, fy = a.org[3 * r + 1] + y, fx = a.org[3 * r + 2] + x;
    const float f = a.full[((int64_t)fz * a.n1 + fy) * a.n2 + fx];
    const float d = 2.0f * sqrtf(fmaxf(f, 0.0f) + 0.375f) - 2.0f * sqrtf(fmaxf(p, 0.0f) + 0.375f);
    s += d * d;
    mx = fmaxf(mx, p);
  }
  red[threadIdx.x] = s;
  top[threadIdx.x] = mx;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      red[threadIdx.x] += red[threadIdx.x + k];
      top[threadIdx.x] = fmaxf(top[threadIdx.x], top[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.part[2 * (r * a.blocks_per_crop + blockIdx.x)] = red[0];
    a.part[2 * (r * a.blocks_per_crop + blockIdx.x) + 1] = top[0];
  }
}

// edge-replicating copy of vol (n0, n1, n2) into out (m0, m1, m2), or the
// inverse crop (crop == 1: out of shape n takes the leading box of vol m)
__global__ void k_pad_edge(const float *__restrict__ in, float *__restrict__ out, int n0, int n1,
                           int n2, int m0, int m1, int m2, int crop) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (!crop) {
    if (e >= (int64_t)m0 * m1 * m2) return;
    const int x = (int)(e % m2), y = (int)((e / m2) % m1), z = (int)(e / ((int64_t)m2 * m1));
    out[e] = in[((int64_t)min(z, n0 - 1) * n1 + min(y, n1 - 1)) * n2 + min(x, n2 - 1)];
  } else {
    if (e >= (int64_t)n0 * n1 * n2) return;
    const int x = (int)(e % n2), y = (int)((e / n2) % n1), z = (int)(e / ((int64_t)n2 * n1));
    out[e] = in[((int64_t)z * m1 + y) * m2 + x];
  }
}

// ---------------------------------------------------------------------------
// host side: dispatch and the tile autotuner
// ---------------------------------------------------------------------------
inline unsigned nblocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

using FwdKey = std::tuple<int, int, int, int, int, int>;  // ca, cb, co, up, outm, mt
using LaunchFn = void (*)(const ConvArgs &, hipStream_t);

template <int CA, int CB, int CO, int MT, bool UP, int OUTM>
void launch_fwd(const ConvArgs &a, hipStream_t s) {
  hipLaunchKernelGGL((k_conv_mfma<CA, CB, CO, MT, UP, OUTM>), dim3(nblocks(a.segments, 4)),
                     dim3(256), 0, s, a);
}

#define FWD_ALL_MT(ca, cb, co, up, om)                                          \
  {FwdKey{ca, cb, co, up, om, 1}, &launch_fwd<ca, cb, co, 1, up, om>},          \
      {FwdKey{ca, cb, co, up, om, 2}, &launch_fwd<ca, cb, co, 2, up, om>},      \
      {FwdKey{ca, cb, co, up, om, 4}, &launch_fwd<ca, cb, co, 4, up, om>}

const std::map<FwdKey, LaunchFn> &fwd_table() {
  static const std::map<FwdKey, LaunchFn> t = {
      // forward layers (bias + ReLU, bf16 out)
      FWD_ALL_MT(16, 0, 16, false, 0), FWD_ALL_MT(16, 0, 32, false, 0),
      FWD_ALL_MT(32, 0, 32, false, 0), FWD_ALL_MT(32, 0, 64, false, 0),
      FWD_ALL_MT(64, 0, 64, false, 0), FWD_ALL_MT(64, 32, 32, true, 0),
      FWD_ALL_MT(32, 16, 16, true, 0),
      // backward-data passes (flipped weights, float out)
      FWD_ALL_MT(16, 0, 16, false, 1), FWD_ALL_MT(32, 0, 16, false, 1),
      FWD_ALL_MT(32, 0, 32, false, 1), FWD_ALL_MT(64, 0, 32, false, 1),
      FWD_ALL_MT(64, 0, 64, false, 1), FWD_ALL_MT(32, 0, 96, false, 1),
      FWD_ALL_MT(16, 0, 48, false, 1),
  };
  return t;
}
#undef FWD_ALL_MT

using WgKey = std::tuple<int, int, int, int>;  // ca, cb, co, up
using WgFn = void (*)(const WgradArgs &, int64_t waves, hipStream_t);

template <int CA, int CB, int CO, bool UP>
void launch_wgrad(const WgradArgs &a, int64_t waves, hipStream_t s) {
  hipLaunchKernelGGL((k_conv_wgrad<CA, CB, CO, UP>), dim3(nblocks(waves, 4)), dim3(256), 0, s, a);
}

const std::map<WgKey, WgFn> &wgrad_table() {
  static const std::map<WgKey, WgFn> t = {
      {WgKey{16, 0, 16, 0}, &launch_wgrad<16, 0, 16, false>},
      {WgKey{16, 0, 32, 0}, &launch_wgrad<16, 0, 32, false>},
      {WgKey{32, 0, 32, 0}, &launch_wgrad<32, 0, 32, false>},
      {WgKey{32, 0, 64, 0}, &launch_wgrad<32, 0, 64, false>},
      {WgKey{64, 0, 64, 0}, &launch_wgrad<64, 0, 64, false>},
      {WgKey{64, 32, 32, 1}, &launch_wgrad<64, 32, 32, true>},
      {WgKey{32, 16, 16, 1}, &launch_wgrad<32, 16, 16, true>},
  };
  return t;
}

// autotuned tile per (layer, geometry, device): MT (tiles of 16 voxels per
// wave) among the variants the row length allows
std::mutex g_tune_mutex;
std::map<std::tuple<int, FwdKey, int, int, int, int>, int> g_tuned;

int check_geom(const Geom &g, bool up) {
  if (g.B < 1 || g.D < 1 || g.H < 1 || g.W < 16)
    return cfail(-1, "conv geometry: every axis positive and W >= 16");
  if (g.W % 16) return cfail(-1, "conv geometry: W must be a multiple of 16");
  if (up && (g.D % 2 || g.H % 2 || g.W % 2))
    return cfail(-1, "conv geometry: an upsampled input needs even D, H and W");
  if ((int64_t)g.B * g.D * g.H * g.W >= ((int64_t)1 << 31))
    return cfail(-1, "conv geometry: too many voxels");
  return 0;
}

}  // namespace

extern "C" {

int64_t bmnd_conv_kpad(int32_t cin) { return kpad(cin); }

int32_t bmnd_conv3d_forward(const void *in_a, const void *in_b, int32_t ca, int32_t cb,
                            int32_t upsample, const void *weights, const float *bias, int32_t co,
                            void *out, int32_t out_float, int32_t B, int32_t D, int32_t H,
                            int32_t W, int32_t mt, int32_t *mt_used, void *stream) {
  const Geom g{B, D, H, W};
  if (int rc = check_geom(g, upsample != 0)) return rc;
  if (!in_a || !weights || !out || (cb > 0 && !in_b) || (!out_float && !bias))
    return cfail(-1, "conv: null pointer");
  const auto &tab = fwd_table();
  hipStream_t s = (hipStream_t)stream;
  auto make = [&](int m) {
    ConvArgs a;
    a.in_a = static_cast<const uint16_t *>(in_a);
    a.in_b = static_cast<const uint16_t *>(in_b);
    a.w = static_cast<const uint16_t *>(weights);
    a.bias = bias;
    a.out = out;
    a.g = g;
    a.segments = (int64_t)B * D * H * (W / (16 * m));
    return a;
  };
  std::vector<int> cands;
  for (int m : {1, 2, 4})
    if (W % (16 * m) == 0 &&
        tab.count(FwdKey{ca, cb, co, upsample != 0, out_float != 0, m}))
      cands.push_back(m);
  if (cands.empty())
    return cfail(-1, "conv: channel configuration (" + std::to_string(ca) + "+" +
                         std::to_string(cb) + " -> " + std::to_string(co) +
                         ") not compiled into this library");
  int pick = mt;
  if (mt == 0) {
    int dev = 0;
    CONV_TRY(hipGetDevice(&dev));
    const auto key = std::make_tuple(dev, FwdKey{ca, cb, co, upsample != 0, out_float != 0, 0},
                                     B, D, H, W);
    {
      std::lock_guard<std::mutex> lock(g_tune_mutex);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) pick = it->second;
    }
    if (pick == 0) {
      // time every variant on this call's own data (each writes the same
      // output, bit for bit), keep the fastest
      hipEvent_t e0, e1;
      CONV_TRY(hipEventCreate(&e0));
      CONV_TRY(hipEventCreate(&e1));
      float best = 1e30f;
      for (int rep = 0; rep < 2; ++rep)
        for (int m : cands) {
          const auto fn = tab.at(FwdKey{ca, cb, co, upsample != 0, out_float != 0, m});
          CONV_TRY(hipEventRecord(e0, s));
          fn(make(m), s);
          CONV_TRY(hipGetLastError());
          CONV_TRY(hipEventRecord(e1, s));
          CONV_TRY(hipEventSynchronize(e1));
          float ms = 0.f;
          CONV_TRY(hipEventElapsedTime(&ms, e0, e1));
          if (rep == 1 && ms < best) {
            best = ms;
            pick = m;
          }
        }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      std::lock_guard<std::mutex> lock(g_tune_mutex);
      g_tuned[key] = pick;
    }
  } else if (std::find(cands.begin(), cands.end(), mt) == cands.end()) {
    return cfail(-1, "conv: tile of " + std::to_string(mt) + " x 16 voxels does not divide W");
  }
  tab.at(FwdKey{ca, cb, co, upsample != 0, out_float != 0, pick})(make(pick), s);
  CONV_TRY(hipGetLastError());
  if (mt_used) *mt_used = pick;
  return 0;
}

int64_t bmnd_conv3d_wgrad_parts(int32_t B, int32_t D, int32_t H, int32_t W, int32_t chunk) {
  if (chunk <= 0 || chunk % 32) return -1;
  const int64_t v = (int64_t)B * D * H * W;
  return (v + chunk - 1) / chunk;
}

int32_t bmnd_conv3d_wgrad(const void *x_a, const void *x_b, int32_t ca, int32_t cb,
                          int32_t upsample, const void *gy, int32_t co, float *parts,
                          int32_t chunk, float *grad, int32_t B, int32_t D, int32_t H, int32_t W,
                          void *stream) {
  const Geom g{B, D, H, W};
  if (int rc = check_geom(g, upsample != 0)) return rc;
  if (chunk <= 0 || chunk % 32) return cfail(-1, "wgrad: chunk must be a positive multiple of 32");
  const auto &tab = wgrad_table();
  auto it = tab.find(WgKey{ca, cb, co, upsample != 0});
  if (it == tab.end()) return cfail(-1, "wgrad: channel configuration not compiled");
  WgradArgs a;
  a.x_a = static_cast<const uint16_t *>(x_a);
  a.x_b = static_cast<const uint16_t *>(x_b);
  a.gy = static_cast<const uint16_t *>(gy);
  a.part = parts;
  a.g = g;
  a.voxels = (int64_t)B * D * H * W;
  a.chunk = chunk;
  const int64_t nch = (a.voxels + chunk - 1) / chunk;
  const int kp = kpad(ca + cb);
  const int64_t kg = (kp / 16 + 3) / 4;
  hipStream_t s = (hipStream_t)stream;
  it->second(a, nch * kg, s);
  CONV_TRY(hipGetLastError());
  const int64_t len = (int64_t)co * kp;
  hipLaunchKernelGGL(k_reduce_parts, dim3(nblocks(len)), dim3(256), 0, s, parts, (int)nch, len,
                     grad);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_in_forward(const float *x, const float *w, const float *bias, void *out,
                             int32_t B, int32_t D, int32_t H, int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * D * H * W;
  hipLaunchKernelGGL(k_conv_in, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, x, w, bias,
                     static_cast<uint16_t *>(out), g, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_out_forward(const void *d, const float *w, const float *bias, const float *res,
                              float *y, int32_t B, int32_t D, int32_t H, int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * D * H * W;
  hipLaunchKernelGGL(k_conv_out, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const uint16_t *>(d), w, bias, res, y, g, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_out_backward(const float *gy, const float *w, const void *d, void *gd,
                               int32_t B, int32_t D, int32_t H, int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * D * H * W;
  hipLaunchKernelGGL(k_conv_out_bwd, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, gy, w,
                     static_cast<const uint16_t *>(d), static_cast<uint16_t *>(gd), g, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_small_wgrad(int32_t last, const float *xf, const void *xb, const float *gf,
                              const void *gb, float *parts, int32_t chunk, float *grad, int32_t B,
                              int32_t D, int32_t H, int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * D * H * W;
  if (chunk <= 0) return cfail(-1, "small wgrad: chunk must be positive");
  const int nch = (int)((n + chunk - 1) / chunk);
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(nch, 2);
  if (last)
    hipLaunchKernelGGL(k_small_wgrad<1>, grid, dim3(256), 0, s, xf,
                       static_cast<const uint16_t *>(xb), gf, static_cast<const uint16_t *>(gb),
                       parts, g, n, chunk);
  else
    hipLaunchKernelGGL(k_small_wgrad<0>, grid, dim3(256), 0, s, xf,
                       static_cast<const uint16_t *>(xb), gf, static_cast<const uint16_t *>(gb),
                       parts, g, n, chunk);
  CONV_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_parts, dim3(nblocks(16 * kTaps)), dim3(256), 0, s, parts, nch,
                     (int64_t)16 * kTaps, grad);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_bias_grad(const void *gb, const float *gf, int32_t C, float *parts,
                            int32_t chunk, float *grad, int64_t n, void *stream) {
  if (C < 1 || C > 256 || chunk <= 0) return cfail(-1, "bias grad: 1 <= C <= 256, chunk > 0");
  const int nch = (int)((n + chunk - 1) / chunk);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bias_grad, dim3(nch), dim3(256), 0, s, static_cast<const uint16_t *>(gb),
                     gf, C, parts, n, chunk);
  CONV_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, s, parts, nch, (int64_t)C, grad);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_pool2(const void *in, void *out, int32_t C, int32_t B, int32_t D, int32_t H,
                        int32_t W, void *stream) {
  if (D % 2 || H % 2 || W % 2) return cfail(-1, "pool: even axes");
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * (D / 2) * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(k_pool2, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const uint16_t *>(in), static_cast<uint16_t *>(out), g, C, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_skip_grad(const float *gin, int32_t ctot, int32_t off, const float *gpool,
                            const void *act, void *out, int32_t C, int32_t B, int32_t D, int32_t H,
                            int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * D * H * W * C;
  hipLaunchKernelGGL(k_skip_grad, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, gin, ctot,
                     off, gpool, static_cast<const uint16_t *>(act), static_cast<uint16_t *>(out),
                     g, C, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_up_grad(const float *gin, int32_t ctot, const void *act, void *out, int32_t C,
                          int32_t B, int32_t D, int32_t H, int32_t W, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t n = (int64_t)B * (D / 2) * (H / 2) * (W / 2) * C;
  hipLaunchKernelGGL(k_up_grad, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, gin, ctot,
                     static_cast<const uint16_t *>(act), static_cast<uint16_t *>(out), g, C, n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_mask_grad(const float *g, const void *act, void *out, int64_t n, void *stream) {
  hipLaunchKernelGGL(k_mask_grad, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, g,
                     static_cast<const uint16_t *>(act), static_cast<uint16_t *>(out), n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_flip_weights(const void *w, int32_t ci, int32_t co, void *wt, void *stream) {
  const int64_t n = (int64_t)ci * kpad(co);
  hipLaunchKernelGGL(k_flip_weights, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const uint16_t *>(w), ci, co, static_cast<uint16_t *>(wt));
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_conv_mse(const float *y, const float *t, float *gy, float *parts, int32_t margin,
                      int32_t B, int32_t D, int32_t H, int32_t W, int32_t log_loss, float eps,
                      float *sample_sums, void *stream) {
  const Geom g{B, D, H, W};
  const int64_t per = (int64_t)D * H * W;
  const int64_t inside = (int64_t)std::max(D - 2 * margin, 0) * std::max(H - 2 * margin, 0) *
                         std::max(W - 2 * margin, 0);
  if (inside <= 0) return cfail(-1, "mse: the margin leaves no voxel");
  if (B < 1 || B > 65535) return cfail(-1, "mse: 1 <= B <= 65535");
  const unsigned nb = nblocks(per);
  hipStream_t s = (hipStream_t)stream;
  // mean over the batch (log_loss == 0) or per sample (then rescaled)
  const float gscale = log_loss ? 1.0f / (float)inside : 1.0f / (float)(inside * B);
  hipLaunchKernelGGL(k_mse, dim3(nb, B), dim3(256), 0, s, y, t, gy, parts, g, margin, gscale);
  CONV_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_sample_sums, dim3(nblocks(B)), dim3(256), 0, s, parts, (int)nb, B,
                     sample_sums);
  CONV_TRY(hipGetLastError());
  if (log_loss) {
    const int64_t n = per * B;
    hipLaunchKernelGGL(k_log_scale, dim3(nblocks(n)), dim3(256), 0, s, gy, sample_sums, per,
                       1.0f / (float)inside, eps, n);
    CONV_TRY(hipGetLastError());
  }
  return 0;
}

int32_t bmnd_adam_step(float *p, const float *g, float *m, float *v, void *wb, int64_t n, float lr,
                       float beta1, float beta2, float eps, int32_t step, void *stream) {
  if (step < 1) return cfail(-1, "adam: step counts from 1");
  const float c1 = 1.0f / (1.0f - powf(beta1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(beta2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                     static_cast<uint16_t *>(wb), n, lr, beta1, beta2, eps, c1, c2);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_to_bf16(const float *x, void *out, int64_t n, void *stream) {
  hipLaunchKernelGGL(k_to_bf16, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, x,
                     static_cast<uint16_t *>(out), n);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_anscombe(const float *x, float *out, int64_t n, int32_t inverse, float center,
                      float spread, void *stream) {
  if (!(spread > 0.0f)) return cfail(-1, "anscombe: spread must be positive");
  hipLaunchKernelGGL(k_anscombe, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, x, out, n,
                     inverse, center, spread);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_anscombe_moments(const float *x, int64_t n, double *part, double *sums,
                              void *stream) {
  const int nb = 1024;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_moments, dim3(nb), dim3(256), 0, s, x, n, part);
  CONV_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_moments_sum, dim3(1), dim3(64), 0, s, part, nb, sums);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_pad_edge(const float *in, float *out, int32_t n0, int32_t n1, int32_t n2, int32_t m0,
                      int32_t m1, int32_t m2, int32_t crop, void *stream) {
  if (m0 < n0 || m1 < n1 || m2 < n2) return cfail(-1, "pad: the padded box must contain the volume");
  const int64_t n = crop ? (int64_t)n0 * n1 * n2 : (int64_t)m0 * m1 * m2;
  hipLaunchKernelGGL(k_pad_edge, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, in, out, n0,
                     n1, n2, m0, m1, m2, crop);
  CONV_TRY(hipGetLastError());
  return 0;
}

int32_t bmnd_compare_crops(const float *full, const float *crops, const int32_t *origins,
                           int32_t ncrops, int32_t n0, int32_t n1, int32_t n2, int32_t c0,
                           int32_t c1, int32_t c2, int32_t margin, float *parts,
                           int32_t blocks_per_crop, void *stream) {
  if (2 * margin >= std::min(c0, std::min(c1, c2))) return cfail(-1, "compare: margin too wide");
  CompareArgs a{full, crops, origins, n0, n1, n2, c0, c1, c2, margin, parts, blocks_per_crop};
  hipLaunchKernelGGL(k_compare_crops, dim3(blocks_per_crop, ncrops), dim3(256), 0,
                     (hipStream_t)stream, a);
  CONV_TRY(hipGetLastError());
  return 0;
}

}  // extern "C"
