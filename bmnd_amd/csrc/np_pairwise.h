// numpy's float32 summation order (np.sum / np.mean of a contiguous array),
// evaluated by one workgroup.
//
// numpy reduces a contiguous array in buffer-sized chunks of 8192 elements,
// each summed by its pairwise routine and added in turn to the running total.
// The pairwise routine sums a range of m elements as follows:
//   m < 8:    sequentially, from zero;
//   m <= 128: eight interleaved accumulators r[j] = x[j] + x[j+8] + ...,
//             combined as ((r0+r1)+(r2+r3)) + ((r4+r5)+(r6+r7)), then the
//             elements past the last multiple of 8 added in order;
//   m > 128:  the sums of the two halves split at h = m/2 - (m/2) % 8.
// For a given m the ranges of <= 128 elements (leaves) and the additions that
// join them form a fixed binary tree.  A plan lists the leaves and then the
// internal additions grouped by height, so that a workgroup sums the leaves in
// parallel and joins one height per step.
//
// Plan layout (int32): [L, H, count[0..H), leaf_start[L], leaf_len[L],
// (a, b) per addition]; addition j writes slot L + j, slots 0..L-1 hold the
// leaves, and the root is the last addition (or slot 0 when L == 1).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>
#include <vector>

constexpr int kNpChunk = 8192;  // numpy's reduction buffer, in elements

// a range of more than 128 elements splits into halves of at least 64, so
// every leaf of a range longer than 128 has at least 64 elements
constexpr int kNpMaxLeaves = kNpChunk / 64;
constexpr int kNpMaxSlots = 2 * kNpMaxLeaves;

__host__ __device__ __forceinline__ float np_leaf_sum(const float *__restrict__ x, int m) {
  if (m < 8) {
    float r = 0.0f;
    for (int i = 0; i < m; ++i) r = r + x[i];
    return r;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = x[j];
  int i = 8;
  const int m8 = m - (m % 8);
  for (; i < m8; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + x[i + j];
  }
  float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < m; ++i) res = res + x[i];
  return res;
}

// Sum of x[0..m) in numpy's order for m <= 8192, by all NT threads of the
// block (every thread must call it); s_slot holds >= kNpMaxSlots floats.  The
// result is returned to every thread.
template <int NT>
__device__ __forceinline__ float np_block_sum(const float *__restrict__ x,
                                              const int32_t *__restrict__ plan,
                                              float *__restrict__ s_slot) {
  const int L = plan[0], H = plan[1];
  const int32_t *cnt = plan + 2;
  const int32_t *lstart = cnt + H;
  const int32_t *llen = lstart + L;
  const int32_t *ops = llen + L;
  for (int l = threadIdx.x; l < L; l += NT) s_slot[l] = np_leaf_sum(x + lstart[l], llen[l]);
  __syncthreads();
  int base = 0;
  for (int h = 0; h < H; ++h) {
    const int c = cnt[h];
    for (int j = threadIdx.x; j < c; j += NT) {
      const int o = base + j;
      s_slot[L + o] = s_slot[ops[2 * o]] + s_slot[ops[2 * o + 1]];
    }
    __syncthreads();
    base += c;
  }
  const float r = s_slot[base == 0 ? 0 : L + base - 1];
  __syncthreads();  // s_slot may be reused after the return
  return r;
}

// The same two routines over values produced on the fly: val(i) gives
// element i of the range; T is the element and accumulator type (numpy's
// float64 sum follows the same tree as its float32 one).
template <class T, class F>
__host__ __device__ __forceinline__ T np_leaf_sum_fn(F val, int lo, int m) {
  if (m < 8) {
    T r = T(0);
    for (int i = 0; i < m; ++i) r = r + val(lo + i);
    return r;
  }
  T r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = val(lo + j);
  int i = 8;
  const int m8 = m - (m % 8);
  for (; i < m8; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + val(lo + i + j);
  }
  T res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < m; ++i) res = res + val(lo + i);
  return res;
}

template <int NT, class T, class F>
__device__ __forceinline__ T np_block_sum_fn(F val, const int32_t *__restrict__ plan,
                                             T *__restrict__ s_slot) {
  const int L = plan[0], H = plan[1];
  const int32_t *cnt = plan + 2;
  const int32_t *lstart = cnt + H;
  const int32_t *llen = lstart + L;
  const int32_t *ops = llen + L;
  for (int l = threadIdx.x; l < L; l += NT) s_slot[l] = np_leaf_sum_fn<T>(val, lstart[l], llen[l]);
  __syncthreads();
  int base = 0;
  for (int h = 0; h < H; ++h) {
    const int c = cnt[h];
    for (int j = threadIdx.x; j < c; j += NT) {
      const int o = base + j;
      s_slot[L + o] = s_slot[ops[2 * o]] + s_slot[ops[2 * o + 1]];
    }
    __syncthreads();
    base += c;
  }
  const T r = s_slot[base == 0 ? 0 : L + base - 1];
  __syncthreads();
  return r;
}

// Host: the plan for a range of m elements (1 <= m <= 8192).
inline std::vector<int32_t> np_pairwise_plan(int m) {
  struct Node {
    int lo, len, height, left, right;  // children are node indices, -1 for a leaf
  };
  std::vector<Node> nodes;
  // depth-first, left before right, so leaves come out in address order
  struct Build {
    std::vector<Node> &nodes;
    int operator()(int lo, int len) {
      const int id = (int)nodes.size();
      nodes.push_back({lo, len, 0, -1, -1});
      if (len > 128) {
        int h = len / 2;
        h -= h % 8;
        const int l = (*this)(lo, h);
        const int r = (*this)(lo + h, len - h);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].height = 1 + std::max(nodes[l].height, nodes[r].height);
      }
      return id;
    }
  } build{nodes};
  build(0, m);
  std::vector<int> slot(nodes.size(), -1);
  std::vector<int32_t> starts, lens;
  int H = 0;
  for (size_t i = 0; i < nodes.size(); ++i) {
    if (nodes[i].left < 0) {
      slot[i] = (int)starts.size();
      starts.push_back(nodes[i].lo);
      lens.push_back(nodes[i].len);
    }
    H = std::max(H, nodes[i].height);
  }
  const int L = (int)starts.size();
  std::vector<int32_t> counts(H, 0), ops;
  int next = L;
  for (int h = 1; h <= H; ++h) {
    for (size_t i = 0; i < nodes.size(); ++i) {
      if (nodes[i].left >= 0 && nodes[i].height == h) {
        slot[i] = next++;
        ops.push_back(slot[nodes[i].left]);
        ops.push_back(slot[nodes[i].right]);
        ++counts[h - 1];
      }
    }
  }
  std::vector<int32_t> plan;
  plan.push_back(L);
  plan.push_back(H);
  plan.insert(plan.end(), counts.begin(), counts.end());
  plan.insert(plan.end(), starts.begin(), starts.end());
  plan.insert(plan.end(), lens.begin(), lens.end());
  plan.insert(plan.end(), ops.begin(), ops.end());
  return plan;
}

// Host: a plan evaluated serially, in the order np_block_sum follows.
inline float np_plan_sum_host(const float *x, const std::vector<int32_t> &plan) {
  const int L = plan[0], H = plan[1];
  const int32_t *cnt = plan.data() + 2;
  const int32_t *lstart = cnt + H;
  const int32_t *llen = lstart + L;
  const int32_t *ops = llen + L;
  std::vector<float> slot(2 * (size_t)L, 0.0f);
  for (int l = 0; l < L; ++l) slot[l] = np_leaf_sum(x + lstart[l], llen[l]);
  int base = 0;
  for (int h = 0; h < H; ++h) {
    for (int j = 0; j < cnt[h]; ++j) {
      const int o = base + j;
      slot[L + o] = slot[ops[2 * o]] + slot[ops[2 * o + 1]];
    }
    base += cnt[h];
  }
  return slot[base == 0 ? 0 : L + base - 1];
}
