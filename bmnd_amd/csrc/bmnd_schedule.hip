// The GENERATED reference schedule (the work queue of a stage) built on the
// device: shifts.py:229-266 (per-axis shift tables) and 519-576 (slot tuples
// in C order with the phase innermost, the shift of axis d chosen by the slot
// of axis d+1 plus the phase, the last slot of an axis pinned to the far
// boundary, duplicates dropped keeping the first occurrence).
//
// Every candidate (slot tuple, phase) computes its patch origin; an integer
// atomicMin per origin finds each origin's first candidate (a minimum, so
// the result is order-free), and an order-preserving compaction (block
// counts, one scan, block-local scans) writes the kept candidates in
// candidate order.  Only the reference count is read back.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "../../include/bmnd_hip.h"
#include "bmnd_internal.h"

namespace {

constexpr int kMaxDim = 4;
constexpr int kMaxShifts = 16;

int sfail(int code, const std::string &msg) { return bmnd_internal_fail(code, msg.c_str()); }

#define SCHED_TRY(expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return sfail(BMND_HIP_ELAUNCH, std::string(#expr) + ": " +        \
                                                             hipGetErrorString(e_));       \
  } while (0)

struct SchedArgs {
  int ndim, phases;
  int slots[kMaxDim], step[kMaxDim], last[kMaxDim];
  int64_t cstride[kMaxDim];  // C-order strides of the patch grid
  int nshift[kMaxDim];
  int shifts[kMaxDim][kMaxShifts];
  int64_t ncand;
};

__device__ int64_t candidate_origin(const SchedArgs &a, int64_t c) {
  int64_t s = c / a.phases;
  const int ph = (int)(c % a.phases);
  int slot[kMaxDim];
  for (int d = a.ndim - 1; d >= 0; --d) {
    slot[d] = (int)(s % a.slots[d]);
    s /= a.slots[d];
  }
  int64_t lin = 0;
  for (int d = 0; d < a.ndim; ++d) {
    const int base = min(slot[d] * a.step[d], a.last[d]);
    const int offset = d < a.ndim - 1 ? a.shifts[d][(slot[d + 1] + ph) % a.nshift[d]] : 0;
    int pos = min(max(base - offset, 0), a.last[d]);
    if (slot[d] == a.slots[d] - 1) pos = a.last[d];
    lin += (int64_t)pos * a.cstride[d];
  }
  return lin;
}

__global__ void k_sched_fill(int32_t *__restrict__ first, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    first[i] = INT_MAX;
}

__global__ void k_sched_candidates(SchedArgs a, int32_t *__restrict__ cand_lin,
                                   int32_t *__restrict__ first) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= a.ncand) return;
  const int64_t lin = candidate_origin(a, c);
  cand_lin[c] = (int32_t)lin;
  atomicMin(&first[lin], (int32_t)c);
}

// kept candidates per block of 256
__global__ void k_sched_count(const int32_t *__restrict__ cand_lin,
                              const int32_t *__restrict__ first, int64_t ncand,
                              int32_t *__restrict__ bcount) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool keep = c < ncand && first[cand_lin[c]] == (int32_t)c;
  const unsigned long long m = __ballot(keep);
  if ((threadIdx.x & 63) == 0) atomicAdd(&cnt, __popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) bcount[blockIdx.x] = cnt;
}

// exclusive scan of the block counts by one workgroup; total at boff[nblk]
__global__ void k_sched_scan(const int32_t *__restrict__ bcount, int nblk,
                             int32_t *__restrict__ boff) {
  __shared__ int sums[1024];
  const int per = (nblk + 1023) / 1024;
  const int t = threadIdx.x;
  int s = 0;
  for (int j = 0; j < per; ++j) {
    const int i = t * per + j;
    if (i < nblk) s += bcount[i];
  }
  sums[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? sums[t - off] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? sums[t - 1] : 0;
  for (int j = 0; j < per; ++j) {
    const int i = t * per + j;
    if (i < nblk) {
      boff[i] = run;
      run += bcount[i];
    }
  }
  if (t == 1023) boff[nblk] = sums[1023];
}

__global__ void k_sched_write(const int32_t *__restrict__ cand_lin,
                              const int32_t *__restrict__ first, int64_t ncand,
                              const int32_t *__restrict__ boff, int32_t *__restrict__ ref_lin,
                              int32_t *__restrict__ ref_slot) {
  __shared__ int wave_base[4];
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool keep = c < ncand && first[cand_lin[c]] == (int32_t)c;
  const unsigned long long m = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wave_base[wave] = __popcll(m);
  __syncthreads();
  int base = boff[blockIdx.x];
  for (int w = 0; w < wave; ++w) base += wave_base[w];
  if (keep) {
    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
    ref_lin[pos] = cand_lin[c];
    if (ref_slot) ref_slot[pos] = (int32_t)c;
  }
}

// offsets of one axis (shifts.py:229-247): round(linspace(0, block - 1, n))
// with n = clip(ceil(density * (block - 1) / step), 1, block), half to even,
// as a sorted set with 0 added
int axis_shifts(int block, int step, bool allow_multiple, double density, int *out) {
  if (!allow_multiple || block < 3 || step < 1 || step > block) {
    out[0] = 0;
    return 1;
  }
  const double safe = density > 0.0 ? density : 0.0;
  int count = (int)std::ceil(safe * (double)(block - 1) / (double)step);
  if (count < 1) count = 1;
  if (count > block) count = block;
  if (count == 1) {
    out[0] = 0;
    return 1;
  }
  // numpy.linspace: start + j * ((stop - start) / (num - 1)), the last exactly stop
  const double delta = (double)(block - 1) / (double)(count - 1);
  std::set<int> vals{0};
  for (int j = 0; j < count; ++j) {
    const double p = (j == count - 1) ? (double)(block - 1) : (double)j * delta;
    vals.insert((int)std::nearbyint(p));
  }
  int n = 0;
  for (int v : vals) out[n++] = v;
  return n;
}

int build_args(int ndim, const int64_t *counts, const int32_t *block, const int32_t *step,
               double shift_density, int schedule_density, SchedArgs *a) {
  if (ndim < 1 || ndim > kMaxDim) return sfail(BMND_HIP_EINVAL, "schedule: 1 <= ndim <= 4");
  a->ndim = ndim;
  a->phases = schedule_density > 1 ? schedule_density : 1;
  int64_t ncand = a->phases, patches = 1;
  for (int d = 0; d < ndim; ++d) {
    if (counts[d] < 1 || block[d] < 1 || step[d] < 1)
      return sfail(BMND_HIP_EINVAL, "schedule: counts, block and step must be positive");
    a->last[d] = (int)(counts[d] - 1);
    a->step[d] = step[d];
    a->slots[d] = (a->last[d] + step[d] - 1) / step[d] + 1;
    ncand *= a->slots[d];
    patches *= counts[d];
    if (d < ndim - 1) {
      int tmp[64];
      const int n = axis_shifts(block[d], step[d], true, shift_density, tmp);
      if (n > kMaxShifts) return sfail(BMND_HIP_EINVAL, "schedule: more than 16 shifts per axis");
      a->nshift[d] = n;
      for (int j = 0; j < n; ++j) a->shifts[d][j] = tmp[j];
    } else {
      a->nshift[d] = 1;
      a->shifts[d][0] = 0;
    }
  }
  if (patches >= ((int64_t)1 << 31) || ncand >= ((int64_t)1 << 31))
    return sfail(BMND_HIP_EINVAL, "schedule: more than 2^31 patch origins or candidates");
  int64_t st = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    a->cstride[d] = st;
    st *= counts[d];
  }
  a->ncand = ncand;
  return 0;
}

}  // namespace

extern "C" {

int32_t bmnd_hip_schedule_axis_shifts(int32_t block, int32_t step, int32_t allow_multiple,
                                      double density, int32_t *out) {
  if (!out) return sfail(BMND_HIP_EINVAL, "out must hold 64 values");
  int tmp[64];
  const int n = axis_shifts(block, step, allow_multiple != 0, density, tmp);
  for (int j = 0; j < n; ++j) out[j] = tmp[j];
  return n;
}

int64_t bmnd_hip_schedule_capacity(int32_t ndim, const int64_t *counts_h, const int32_t *block_h,
                                   const int32_t *step_h, double shift_density,
                                   int32_t schedule_density) {
  SchedArgs a;
  if (build_args(ndim, counts_h, block_h, step_h, shift_density, schedule_density, &a)) return -1;
  return a.ncand;
}

int32_t bmnd_hip_schedule_generated(int32_t ndim, const int64_t *counts_h, const int32_t *block_h,
                                    const int32_t *step_h, double shift_density,
                                    int32_t schedule_density, int32_t *ref_lin, int32_t *ref_slot,
                                    int64_t capacity, int64_t *nrefs, void *stream) {
  SchedArgs a;
  if (int rc = build_args(ndim, counts_h, block_h, step_h, shift_density, schedule_density, &a))
    return rc;
  if (!ref_lin || !nrefs) return sfail(BMND_HIP_EINVAL, "schedule: null output");
  if (capacity < a.ncand)
    return sfail(BMND_HIP_EINVAL, "schedule: capacity below bmnd_hip_schedule_capacity()");
  int64_t patches = 1;
  for (int d = 0; d < ndim; ++d) patches *= counts_h[d];
  hipStream_t s = (hipStream_t)stream;
  const int nblk = (int)((a.ncand + 255) / 256);
  int32_t *first = nullptr, *cand = nullptr, *bcount = nullptr;
  SCHED_TRY(hipMallocAsync((void **)&first, sizeof(int32_t) * patches, s));
  SCHED_TRY(hipMallocAsync((void **)&cand, sizeof(int32_t) * a.ncand, s));
  SCHED_TRY(hipMallocAsync((void **)&bcount, sizeof(int32_t) * (2 * (size_t)nblk + 1), s));
  int32_t *boff = bcount + nblk;
  hipLaunchKernelGGL(k_sched_fill, dim3((unsigned)std::min<int64_t>((patches + 255) / 256, 4096)),
                     dim3(256), 0, s, first, patches);
  hipLaunchKernelGGL(k_sched_candidates, dim3(nblk), dim3(256), 0, s, a, cand, first);
  hipLaunchKernelGGL(k_sched_count, dim3(nblk), dim3(256), 0, s, cand, first, a.ncand, bcount);
  hipLaunchKernelGGL(k_sched_scan, dim3(1), dim3(1024), 0, s, bcount, nblk, boff);
  hipLaunchKernelGGL(k_sched_write, dim3(nblk), dim3(256), 0, s, cand, first, a.ncand, boff,
                     ref_lin, ref_slot);
  const hipError_t launched = hipGetLastError();
  int32_t total = 0;
  const hipError_t copied = hipMemcpyAsync(&total, boff + nblk, sizeof(int32_t),
                                           hipMemcpyDeviceToHost, s);
  (void)hipFreeAsync(first, s);
  (void)hipFreeAsync(cand, s);
  (void)hipFreeAsync(bcount, s);
  if (launched != hipSuccess)
    return sfail(BMND_HIP_ELAUNCH, std::string("schedule launch: ") + hipGetErrorString(launched));
  if (copied != hipSuccess)
    return sfail(BMND_HIP_ELAUNCH, std::string("schedule copy: ") + hipGetErrorString(copied));
  SCHED_TRY(hipStreamSynchronize(s));
  *nrefs = total;
  return 0;
}

}  // extern "C"
