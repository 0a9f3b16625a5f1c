// pff_block.hpp -- block-level helpers (256-thread blocks, 64-wide waves):
// exclusive scan and list append with one global atomic per block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pff {

constexpr int kBT = 256;  // threads per block for block helpers

// Exclusive prefix of v over a block of BT threads (1-D or flattened);
// *total receives the block sum.  Must be reached by every thread.
template <int BT = kBT>
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[BT / 64];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int woff = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BT / 64; ++w) {
    const int s = wsum[w];
    woff += (w < wid) ? s : 0;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return woff + x - v;
}

// Append the taken values of every thread (up to ITEMS each) to list[*count..]
// with a single atomicAdd per block.  Must be reached by every thread.
template <int ITEMS, int BT = kBT>
__device__ __forceinline__ void block_append(const int32_t (&val)[ITEMS], const bool (&take)[ITEMS], int32_t* list,
                                             int64_t* count) {
  __shared__ int64_t base_s;
  int mine = 0;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) mine += take[j] ? 1 : 0;
  int total;
  const int off = block_exclusive_scan<BT>(mine, &total);
  if (threadIdx.x == 0 && threadIdx.y == 0) base_s = total ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)total) : 0;
  __syncthreads();
  int64_t p = base_s + off;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
    if (take[j]) list[p++] = val[j];
  __syncthreads();
}

// Append in (round, wave, lane) order with one atomicAdd per block: item r of
// every thread is one round, so items a block visits row by row (one wave
// per row) land in the list in row-major order, and items taken from a list
// keep that list's order.  Must be reached by every thread.
template <int R, int BT = kBT>
__device__ __forceinline__ void block_append_ordered(const int32_t (&val)[R], const bool (&take)[R], int32_t* list,
                                                     int64_t* count) {
  constexpr int W = BT / 64;
  __shared__ int off_s[R * W];
  __shared__ int64_t base_s;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  uint64_t mask[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    mask[r] = __ballot(take[r]);
    if (lane == 0) off_s[r * W + wid] = __popcll(mask[r]);
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int k = 0; k < R * W; ++k) {
      const int c = off_s[k];
      off_s[k] = acc;
      acc += c;
    }
    base_s = acc ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)acc) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (take[r]) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask[r] >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask[r], 0u));
      list[base_s + off_s[r * W + wid] + below] = val[r];
    }
  }
  __syncthreads();
}

// The same in (wave, round, lane) order: a wave that visits R consecutive
// rows, one per round, with the waves stacked, also appends row by row.
template <int R, int BT = kBT>
__device__ __forceinline__ void block_append_wave_major(const int32_t (&val)[R], const bool (&take)[R],
                                                        int32_t* list, int64_t* count) {
  constexpr int W = BT / 64;
  __shared__ int off_s[R * W];
  __shared__ int64_t base_s;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  uint64_t mask[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    mask[r] = __ballot(take[r]);
    if (lane == 0) off_s[wid * R + r] = __popcll(mask[r]);
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int k = 0; k < R * W; ++k) {
      const int c = off_s[k];
      off_s[k] = acc;
      acc += c;
    }
    base_s = acc ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)acc) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (take[r]) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask[r] >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask[r], 0u));
      list[base_s + off_s[wid * R + r] + below] = val[r];
    }
  }
  __syncthreads();
}

// block_append_wave_major of the values taken by `take`, followed in the same
// list (one atomicAdd for both) by the values taken by `take2`, each with
// `flag2` or'ed in (any order).  Must be reached by every thread.
template <int R, int BT = kBT>
__device__ __forceinline__ void block_append_wave_major2(const int32_t (&val)[R], const bool (&take)[R],
                                                         const bool (&take2)[R], uint32_t flag2, int32_t* list,
                                                         int64_t* count) {
  constexpr int W = BT / 64;
  __shared__ int off_s[R * W];
  __shared__ int64_t base_s;
  __shared__ int second_s;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  uint64_t mask[R];
  int mine2 = 0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    mask[r] = __ballot(take[r]);
    if (lane == 0) off_s[wid * R + r] = __popcll(mask[r]);
    mine2 += take2[r] ? 1 : 0;
  }
  int total2;
  const int off2 = block_exclusive_scan<BT>(mine2, &total2);  // synchronises
  if (tid == 0) {
    int acc = 0;
    for (int k = 0; k < R * W; ++k) {
      const int c = off_s[k];
      off_s[k] = acc;
      acc += c;
    }
    second_s = acc;
    base_s = acc + total2 ? (int64_t)atomicAdd((unsigned long long*)count, (unsigned long long)(acc + total2)) : 0;
  }
  __syncthreads();
  int64_t p2 = base_s + second_s + off2;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (take[r]) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask[r] >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)mask[r], 0u));
      list[base_s + off_s[wid * R + r] + below] = val[r];
    }
    if (take2[r]) list[p2++] = (int32_t)((uint32_t)val[r] | flag2);
  }
  __syncthreads();
}

// Compact the taken values of every thread into this block's own segment
// (seg, room for every item of the block) and store the count: no global
// atomic.  Appends of thousands of blocks to one shared counter serialise at
// the counter's cache line (about 12 ns each); per-block segments are
// joined afterwards by k_seg_scan and a copy.  Must be reached by every
// thread.
template <int ITEMS, int BT = kBT>
__device__ __forceinline__ void seg_append(const int32_t (&val)[ITEMS], const bool (&take)[ITEMS], int32_t* seg,
                                           int32_t* count) {
  int mine = 0;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) mine += take[j] ? 1 : 0;
  int total;
  int p = block_exclusive_scan<BT>(mine, &total);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
    if (take[j]) seg[p++] = val[j];
  if (threadIdx.x == 0 && threadIdx.y == 0) *count = total;
}

// seg_append with the taken items as bits of a mask and their values from
// val(j)
template <int ITEMS, int BT, class F>
__device__ __forceinline__ void seg_append_mask(uint32_t take, F val, int32_t* seg, int32_t* count) {
  int total;
  int p = block_exclusive_scan<BT>(__popc(take), &total);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j)
    if ((take >> j) & 1u) seg[p++] = val(j);
  if (threadIdx.x == 0 && threadIdx.y == 0) *count = total;
}

// Block-wide sum with one atomicAdd on a 64-bit counter.
template <int BT = kBT>
__device__ __forceinline__ void block_count(int mine, int64_t* counter) {
  int total;
  (void)block_exclusive_scan<BT>(mine, &total);
  if (threadIdx.x == 0 && threadIdx.y == 0 && total) atomicAdd((unsigned long long*)counter, (unsigned long long)total);
}

}  // namespace pff
