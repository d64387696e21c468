// vl_dev.h — device helpers shared by the gfx950 kernels (wave64 primitives).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vl {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive wave64 scan.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = __lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    T o = __shfl_up(v, d, WAVE);
    if (l >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, WAVE);
  return v;
}

// Block-wide exclusive scan of one value per thread.  `ws` must hold
// (blockDim.x/64) elements of shared memory.  Returns the exclusive prefix;
// *total receives the block sum.  Contains two __syncthreads().
template <typename T, int NT>
__device__ __forceinline__ T block_excl_scan(T v, T* ws, T* total) {
  constexpr int NW = NT / WAVE;
  const int w = threadIdx.x / WAVE;
  const int l = __lane_id();
  T inc = wave_incl_scan(v);
  if (l == WAVE - 1) ws[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) {
    T x = ws[i];
    if (i < w) base += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// Block-wide exclusive scan of 0/1 flags through ballots.
template <int NT>
__device__ __forceinline__ int block_excl_flag_scan(bool f, int* ws, int* total) {
  constexpr int NW = NT / WAVE;
  const int w = threadIdx.x / WAVE;
  const uint64_t m = __ballot(f);
  const int pre = __popcll(m & lanemask_lt());
  if (__lane_id() == 0) ws[w] = __popcll(m);
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; i++) {
    int x = ws[i];
    if (i < w) base += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return base + pre;
}

__device__ __forceinline__ uint64_t bits_range(int lo, int hi) {  // bits lo..hi inclusive, 0<=lo<=hi<=63
  return (~0ull >> (63 - hi)) & (~0ull << lo);
}

}  // namespace vl
