// Global noise std of the Poisson variance model for bmnd(return_sigma_map=True):
// estimate_poisson_global_sigma (sigmaestimation.py:48-93) on the variance
// volume, with every mean taken in numpy's float32 summation order.
//
// Kernels
//   k_chunk_counts   finite / finite-positive counts per 8192-voxel chunk   HBM stream
//   k_compact        finite positive values, in order (only when some
//                    voxel is not positive: a zero variance floor)        HBM stream
//   k_chunk_np_sum   numpy's pairwise sum of every 8192-element chunk     HBM stream
//   k_finish_global  chunk sums added in order, mean, sqrt(max(mean, eps))  one lane
// The trimmed mean of the reference's sparse case sorts the positive values
// with rocPRIM's radix sort.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/bmnd_hip.h"
#include "bmnd_internal.h"
#include "np_pairwise.h"

namespace {

#define SIG_TRY(expr)                                              \
  do {                                                             \
    hipError_t e_ = (expr);                                        \
    if (e_ != hipSuccess) {                                        \
      release();                                                   \
      return bmnd_internal_fail(BMND_HIP_ELAUNCH, hipGetErrorString(e_)); \
    }                                                              \
  } while (0)

constexpr int kSigThreads = 256;
constexpr int kPerThread = kNpChunk / kSigThreads;  // 32

__device__ __forceinline__ bool finite_positive(float v) { return isfinite(v) && v > 0.0f; }

// counts[2c] = finite values of chunk c, counts[2c+1] = finite positive ones
__global__ __launch_bounds__(kSigThreads) void k_chunk_counts(const float *__restrict__ v,
                                                              int64_t n,
                                                              int32_t *__restrict__ counts) {
  __shared__ int s_f[kSigThreads / 64], s_p[kSigThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kNpChunk;
  int nf = 0, npos = 0;
#pragma unroll 4
  for (int it = 0; it < kPerThread; ++it) {
    const int64_t i = base + it * kSigThreads + threadIdx.x;
    if (i < n) {
      const float x = v[i];
      nf += isfinite(x) ? 1 : 0;
      npos += finite_positive(x) ? 1 : 0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    nf += __shfl_xor(nf, o, 64);
    npos += __shfl_xor(npos, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    s_f[threadIdx.x >> 6] = nf;
    s_p[threadIdx.x >> 6] = npos;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tf = 0, tp = 0;
    for (int w = 0; w < kSigThreads / 64; ++w) {
      tf += s_f[w];
      tp += s_p[w];
    }
    counts[2 * blockIdx.x] = tf;
    counts[2 * blockIdx.x + 1] = tp;
  }
}

// the finite positive values of chunk c, in order, from out[offsets[c]]
__global__ __launch_bounds__(kSigThreads) void k_compact(const float *__restrict__ v, int64_t n,
                                                         const int64_t *__restrict__ offsets,
                                                         float *__restrict__ out) {
  __shared__ float s_x[kNpChunk];
  __shared__ int s_scan[kSigThreads];
  const int64_t base = (int64_t)blockIdx.x * kNpChunk;
  const int len = (int)(n - base < kNpChunk ? n - base : kNpChunk);
  for (int i = threadIdx.x; i < len; i += kSigThreads) s_x[i] = v[base + i];
  __syncthreads();
  // thread t owns the contiguous run [t*32, t*32+32) of the chunk
  const int lo = threadIdx.x * kPerThread;
  int mine = 0;
  for (int j = 0; j < kPerThread; ++j)
    if (lo + j < len && finite_positive(s_x[lo + j])) ++mine;
  s_scan[threadIdx.x] = mine;
  __syncthreads();
  for (int d = 1; d < kSigThreads; d <<= 1) {  // inclusive scan
    const int add = threadIdx.x >= d ? s_scan[threadIdx.x - d] : 0;
    __syncthreads();
    s_scan[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t w = offsets[blockIdx.x] + (s_scan[threadIdx.x] - mine);
  for (int j = 0; j < kPerThread; ++j)
    if (lo + j < len && finite_positive(s_x[lo + j])) out[w++] = s_x[lo + j];
}

// numpy's pairwise sum of every 8192-element chunk of x[0..m)
__global__ __launch_bounds__(kSigThreads) void k_chunk_np_sum(const float *__restrict__ x,
                                                              int64_t m,
                                                              const int32_t *__restrict__ full,
                                                              const int32_t *__restrict__ tail,
                                                              float *__restrict__ sums) {
  __shared__ float s_x[kNpChunk];
  __shared__ float s_slot[kNpMaxSlots];
  const int64_t base = (int64_t)blockIdx.x * kNpChunk;
  const int len = (int)(m - base < kNpChunk ? m - base : kNpChunk);
  for (int i = threadIdx.x; i < len; i += kSigThreads) s_x[i] = x[base + i];
  __syncthreads();
  const float r = np_block_sum<kSigThreads>(s_x, len == kNpChunk ? full : tail, s_slot);
  if (threadIdx.x == 0) sums[blockIdx.x] = r;
}

// the chunk sums added in turn (numpy's reduction loop), the mean as
// np.mean forms it (float64 quotient rounded to float32), and the estimate
__global__ void k_finish_global(const float *__restrict__ sums, int64_t nchunks, int64_t m,
                                double eps, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  float r = 0.0f;
  for (int64_t c = 0; c < nchunks; ++c) r = r + sums[c];
  const double mean = (double)(float)((double)r / (double)m);
  const double g = sqrt(eps > mean ? eps : mean);
  out[0] = g > 0.0 ? g : 0.0;
}

}  // namespace

extern "C" {

int32_t bmnd_hip_poisson_global_sigma(const float *var, int64_t n, double variance_floor,
                                      double *global_sigma, void *stream) {
  if (!global_sigma || !var) return bmnd_internal_fail(BMND_HIP_EINVAL, "null pointer");
  if (n <= 0) return bmnd_internal_fail(BMND_HIP_EINVAL, "the volume must not be empty");
  if (!(variance_floor >= 0.0) || !std::isfinite(variance_floor))
    return bmnd_internal_fail(BMND_HIP_EINVAL, "variance_floor must be finite and non-negative");
  const double eps = variance_floor > 1e-10 ? variance_floor : 1e-10;
  hipStream_t s = (hipStream_t)stream;
  const int64_t nchunks = (n + kNpChunk - 1) / kNpChunk;
  if (nchunks > INT32_MAX) return bmnd_internal_fail(BMND_HIP_EINVAL, "volume too large");

  int32_t *d_counts = nullptr, *d_plans = nullptr;
  int64_t *d_off = nullptr;
  float *d_packed = nullptr, *d_sorted = nullptr, *d_sums = nullptr;
  double *d_out = nullptr;
  void *d_tmp = nullptr;
  auto release = [&]() {
    void *ptrs[] = {d_counts, d_plans, d_off, d_packed, d_sorted, d_sums, d_out, d_tmp};
    for (void *p : ptrs)
      if (p) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
  };

  SIG_TRY(hipMallocAsync((void **)&d_counts, sizeof(int32_t) * 2 * nchunks, s));
  hipLaunchKernelGGL(k_chunk_counts, dim3((unsigned)nchunks), dim3(kSigThreads), 0, s, var, n,
                     d_counts);
  SIG_TRY(hipGetLastError());
  std::vector<int32_t> counts(2 * nchunks);
  SIG_TRY(hipMemcpyAsync(counts.data(), d_counts, sizeof(int32_t) * 2 * nchunks,
                         hipMemcpyDeviceToHost, s));
  SIG_TRY(hipStreamSynchronize(s));
  int64_t nfin = 0, npos = 0;
  std::vector<int64_t> off(nchunks);
  for (int64_t c = 0; c < nchunks; ++c) {
    off[c] = npos;
    nfin += counts[2 * c];
    npos += counts[2 * c + 1];
  }
  if (npos == 0) {  // the mean of the (zero) non-negative values
    release();
    *global_sigma = std::sqrt(eps);
    return 0;
  }
  const float *src = var;
  int64_t m = n;
  if (npos != n) {
    SIG_TRY(hipMallocAsync((void **)&d_off, sizeof(int64_t) * nchunks, s));
    SIG_TRY(hipMallocAsync((void **)&d_packed, sizeof(float) * npos, s));
    SIG_TRY(hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * nchunks, hipMemcpyHostToDevice,
                           s));
    hipLaunchKernelGGL(k_compact, dim3((unsigned)nchunks), dim3(kSigThreads), 0, s, var, n, d_off,
                       d_packed);
    SIG_TRY(hipGetLastError());
    src = d_packed;
    m = npos;
    const double frac = (double)npos / (double)(nfin > 1 ? nfin : 1);
    if (frac < 0.1 && npos >= 3) {
      SIG_TRY(hipMallocAsync((void **)&d_sorted, sizeof(float) * npos, s));
      size_t bytes = 0;
      SIG_TRY(rocprim::radix_sort_keys(nullptr, bytes, d_packed, d_sorted, (size_t)npos, 0, 32, s));
      SIG_TRY(hipMallocAsync(&d_tmp, bytes > 0 ? bytes : 1, s));
      SIG_TRY(rocprim::radix_sort_keys(d_tmp, bytes, d_packed, d_sorted, (size_t)npos, 0, 32, s));
      int64_t trim = (int64_t)(0.1 * (double)npos);
      if (trim < 1) trim = 1;
      if (npos - 2 * trim > 0) {
        src = d_sorted + trim;
        m = npos - 2 * trim;
      } else {
        src = d_sorted;
        m = npos;
      }
    }
  }
  const int64_t mchunks = (m + kNpChunk - 1) / kNpChunk;
  const std::vector<int32_t> full = np_pairwise_plan(kNpChunk);
  const int tail_len = (int)(m % kNpChunk);
  const std::vector<int32_t> tail = tail_len ? np_pairwise_plan(tail_len) : full;
  std::vector<int32_t> plans(full);
  plans.insert(plans.end(), tail.begin(), tail.end());
  SIG_TRY(hipMallocAsync((void **)&d_plans, sizeof(int32_t) * plans.size(), s));
  SIG_TRY(hipMallocAsync((void **)&d_sums, sizeof(float) * mchunks, s));
  SIG_TRY(hipMallocAsync((void **)&d_out, sizeof(double), s));
  SIG_TRY(hipMemcpyAsync(d_plans, plans.data(), sizeof(int32_t) * plans.size(),
                         hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_chunk_np_sum, dim3((unsigned)mchunks), dim3(kSigThreads), 0, s, src, m,
                     d_plans, d_plans + full.size(), d_sums);
  SIG_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_finish_global, dim3(1), dim3(64), 0, s, d_sums, mchunks, m, eps, d_out);
  SIG_TRY(hipGetLastError());
  double g = 0.0;
  SIG_TRY(hipMemcpyAsync(&g, d_out, sizeof(double), hipMemcpyDeviceToHost, s));
  SIG_TRY(hipStreamSynchronize(s));
  release();
  *global_sigma = g;
  return 0;
}

float bmnd_hip_np_sum_host(const float *x, int64_t n) {
  float r = 0.0f;
  for (int64_t c = 0; c < n; c += kNpChunk) {
    const int len = (int)(n - c < kNpChunk ? n - c : kNpChunk);
    r = r + np_plan_sum_host(x + c, np_pairwise_plan(len));
  }
  return r;
}

}  // extern "C"
