// Median and median absolute deviation of a volume on the device: the first
// estimate of estimate_gaussian_global_sigma (sigmaestimation.py:7-45), which
// takes the finite nonzero values (the finite ones when every finite value
// is zero), their numpy median m, then the numpy median of |x - m|.  numpy's
// median of an even count is the float32 mean of the two central values,
// fl(fl(a + b) / 2); of an odd count, the central value.
//
// Order statistics are exact, so the k-th smallest value is found by a
// radix select over an order-preserving 32-bit key of each float, 8 bits per
// pass, for the two central ranks at once; no value is copied or sorted.
//
// Kernels
//   k_mad_count   finite-nonzero and finite counts                 HBM stream
//   k_mad_hist    per pass: 256-bin histograms of the keys that
//                 match each rank's resolved prefix                HBM stream
//   k_mad_pick    per pass: each rank's bin and remaining rank     one wave
//   k_mad_centre  the median from the two ranks; re-arms the ranks
//                 for the deviations                                one lane
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/bmnd_hip.h"
#include "bmnd_internal.h"

namespace {

constexpr int kMadThreads = 256;
constexpr int kMadBlocks = 2048;

struct MadState {
  unsigned long long count_nz, count_fin;
  unsigned long long rank[2];  // remaining rank of each target within its prefix
  unsigned prefix[2];
  unsigned mask;
  int use_finite;  // every finite value is zero: take the finite ones
  float median, mad;
  unsigned long long hist[2][256];
};

__device__ __forceinline__ unsigned key_of(float x) {
  const unsigned b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float value_of(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__device__ __forceinline__ bool used(float x, int use_finite) {
  return isfinite(x) && (use_finite || x != 0.0f);
}

__global__ void k_mad_init(MadState *st) {
  if (threadIdx.x == 0) {
    st->count_nz = 0ull;
    st->count_fin = 0ull;
  }
  for (int i = threadIdx.x; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0ull;
}

__global__ __launch_bounds__(kMadThreads) void k_mad_count(const float *__restrict__ v, int64_t n,
                                                          MadState *st) {
  unsigned long long nz = 0ull, fin = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * kMadThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kMadThreads) {
    const float x = v[i];
    fin += isfinite(x) ? 1ull : 0ull;
    nz += (isfinite(x) && x != 0.0f) ? 1ull : 0ull;
  }
  for (int o = 32; o > 0; o >>= 1) {
    nz += __shfl_xor(nz, o, 64);
    fin += __shfl_xor(fin, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&st->count_nz, nz);
    atomicAdd(&st->count_fin, fin);
  }
}

// the two central ranks of the used values, every prefix empty
__device__ void arm_ranks(MadState *st) {
  const unsigned long long m = st->use_finite ? st->count_fin : st->count_nz;
  st->rank[0] = m > 0 ? (m - 1) / 2 : 0;
  st->rank[1] = m / 2;
  st->prefix[0] = st->prefix[1] = 0u;
  st->mask = 0u;
}

__global__ void k_mad_arm(MadState *st) {
  if (threadIdx.x == 0) {
    st->use_finite = st->count_nz == 0ull ? 1 : 0;
    arm_ranks(st);
  }
}

// DEV: keys of the values (false) or of |x - median| (true)
template <bool DEV>
__global__ __launch_bounds__(kMadThreads) void k_mad_hist(const float *__restrict__ v, int64_t n,
                                                         MadState *st, int shift) {
  __shared__ unsigned h[2][256];
  for (int i = threadIdx.x; i < 512; i += kMadThreads) (&h[0][0])[i] = 0u;
  __syncthreads();
  const unsigned mask = st->mask, p0 = st->prefix[0], p1 = st->prefix[1];
  const int use_finite = st->use_finite;
  const float med = st->median;
  for (int64_t i = (int64_t)blockIdx.x * kMadThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kMadThreads) {
    const float x = v[i];
    if (!used(x, use_finite)) continue;
    const unsigned k = key_of(DEV ? fabsf(x - med) : x);
    const unsigned bin = (k >> shift) & 255u;
    if ((k & mask) == p0) atomicAdd(&h[0][bin], 1u);
    if ((k & mask) == p1) atomicAdd(&h[1][bin], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 512; i += kMadThreads) {
    const unsigned c = (&h[0][0])[i];
    if (c) atomicAdd(&(&st->hist[0][0])[i], (unsigned long long)c);
  }
}

// one wave: lane t scans the 256 bins of rank t for the bin holding it and
// moves the rank into that bin; then the histograms are cleared
__global__ void k_mad_pick(MadState *st, int shift) {
  const int t = threadIdx.x;
  if (t < 2) {
    unsigned long long below = 0ull;
    const unsigned long long r = st->rank[t];
    int bin = 255;
    for (int b = 0; b < 256; ++b) {
      const unsigned long long c = st->hist[t][b];
      if (below + c > r) {
        bin = b;
        break;
      }
      below += c;
    }
    st->rank[t] = r - below;
    st->prefix[t] |= (unsigned)bin << shift;
  }
  __syncthreads();
  if (t == 0) st->mask |= 255u << shift;
  for (int i = t; i < 512; i += blockDim.x) (&st->hist[0][0])[i] = 0ull;
}

__device__ float central(const MadState *st) {
  const unsigned long long m = st->use_finite ? st->count_fin : st->count_nz;
  if (m == 0ull) return 0.0f;  // the reference's [0.0] when nothing is finite
  const float a = value_of(st->prefix[0]), b = value_of(st->prefix[1]);
  return (m & 1ull) ? a : (a + b) / 2.0f;
}

__global__ void k_mad_centre(MadState *st, int deviations) {
  if (threadIdx.x == 0) {
    if (!deviations) {
      st->median = central(st);
      arm_ranks(st);
    } else {
      st->mad = central(st);
    }
  }
}

__global__ void k_mad_out(const MadState *st, float *out, unsigned long long *counts) {
  if (threadIdx.x == 0) {
    out[0] = st->median;
    out[1] = st->mad;
    counts[0] = st->use_finite ? st->count_fin : st->count_nz;
    counts[1] = (unsigned long long)st->use_finite;
  }
}

}  // namespace

extern "C" {

int32_t bmnd_hip_median_mad(const float *vol, int64_t n, float *out,
                            unsigned long long *counts, void *stream) {
  if (!vol || !out || !counts) return bmnd_internal_fail(BMND_HIP_EINVAL, "null argument");
  if (n <= 0) return bmnd_internal_fail(BMND_HIP_EINVAL, "an empty volume has no median");
  hipStream_t s = (hipStream_t)stream;
  MadState *st = nullptr;
  hipError_t e = hipMallocAsync((void **)&st, sizeof(MadState), s);
  if (e != hipSuccess) return bmnd_internal_fail(BMND_HIP_ELAUNCH, hipGetErrorString(e));
  const int blocks = (int)std::min<int64_t>((n + kMadThreads - 1) / kMadThreads, kMadBlocks);
  hipLaunchKernelGGL(k_mad_init, dim3(1), dim3(256), 0, s, st);
  hipLaunchKernelGGL(k_mad_count, dim3(blocks), dim3(kMadThreads), 0, s, vol, n, st);
  hipLaunchKernelGGL(k_mad_arm, dim3(1), dim3(64), 0, s, st);
  for (int dev = 0; dev < 2; ++dev) {
    for (int shift = 24; shift >= 0; shift -= 8) {
      if (dev)
        hipLaunchKernelGGL(k_mad_hist<true>, dim3(blocks), dim3(kMadThreads), 0, s, vol, n, st,
                           shift);
      else
        hipLaunchKernelGGL(k_mad_hist<false>, dim3(blocks), dim3(kMadThreads), 0, s, vol, n, st,
                           shift);
      hipLaunchKernelGGL(k_mad_pick, dim3(1), dim3(64), 0, s, st, shift);
    }
    hipLaunchKernelGGL(k_mad_centre, dim3(1), dim3(64), 0, s, st, dev);
  }
  hipLaunchKernelGGL(k_mad_out, dim3(1), dim3(64), 0, s, st, out, counts);
  e = hipGetLastError();
  (void)hipFreeAsync(st, s);
  if (e != hipSuccess) return bmnd_internal_fail(BMND_HIP_ELAUNCH, hipGetErrorString(e));
  return 0;
}

}  // extern "C"
