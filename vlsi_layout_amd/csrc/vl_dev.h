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

// Device memory reached through a pointer the compiler cannot trace to a
// kernel argument (one read back from LDS, or a member of a by-value struct
// in some kernels) is accessed with flat instructions.  Those count against
// the LDS wait counter as well, so every later LDS wait also waits for the
// store to reach memory.  These stores name the global address space.
#define VL_GLOBAL __attribute__((address_space(1)))
typedef int32_t vl_v4i32 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ void gstore(T* p, T v) {
  *(VL_GLOBAL T*)p = v;
}

template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const VL_GLOBAL T*)p;
}

// A 16-byte box record (x1, y1, x2, y2) at p.
__device__ __forceinline__ void gstore_box(void* p, int32_t x1, int32_t y1, int32_t x2, int32_t y2) {
  *(VL_GLOBAL vl_v4i32*)p = vl_v4i32{x1, y1, x2, y2};
}

// Lane exchange v(lane ^ M) without the LDS crossbar where the hardware has a
// cheaper path: DPP quad permutes for 1 and 2, the DPP 16-lane row rotate by
// 8 for 8, ds_swizzle in bit mode (no LDS access) for 4 and 16, and
// ds_bpermute only for 32.
template <int M>
__device__ __forceinline__ int xor_lanes(int v) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "xor distance");
  if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);        // quad_perm 1,0,3,2
  else if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm 2,3,0,1
  else if constexpr (M == 8) return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  else if constexpr (M == 4 || M == 16) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (M << 10));
  else return __shfl_xor(v, 32, WAVE);
}

// The same with the distance known only after unrolling (the switch folds).
__device__ __forceinline__ int xor_lanes(int v, int m) {
  switch (m) {
    case 1: return xor_lanes<1>(v);
    case 2: return xor_lanes<2>(v);
    case 4: return xor_lanes<4>(v);
    case 8: return xor_lanes<8>(v);
    case 16: return xor_lanes<16>(v);
    case 32: return xor_lanes<32>(v);
    default: return __shfl_xor(v, m, WAVE);
  }
}

__device__ __forceinline__ uint64_t xor_lanes64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)xor_lanes((int)(uint32_t)v, m);
  const uint32_t hi = (uint32_t)xor_lanes((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// 32- or 64-bit value of lane ^ M / of the lane N below in the same 16-lane
// row (0 where that lane is outside the row), through DPP or swizzle.
template <int M, typename T>
__device__ __forceinline__ T xor_lanes_t(T v) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)xor_lanes<M>((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)xor_lanes<M>((int)(uint32_t)(u >> 32));
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)xor_lanes<M>((int)v);
  }
}

template <int N, typename T>
__device__ __forceinline__ T row_shr(T v) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, 0x110 + N, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), 0x110 + N, 0xF, 0xF, false);
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + N, 0xF, 0xF, false);
  }
}

template <typename T>
__device__ __forceinline__ T read_lane(T v, int l) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_readlane((int)v, l);
  }
}

// v, the same in every active lane, as a scalar (wave-uniform) value: values
// read from LDS land in vector registers, and the compiler cannot tell that
// they are uniform, so arithmetic on them would otherwise stay in the VALU.
template <typename T>
__device__ __forceinline__ T uniform(T v) {
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (T)(((uint64_t)hi << 32) | lo);
  } else {
    return (T)__builtin_amdgcn_readfirstlane((int)v);
  }
}

// Inclusive wave64 scan: DPP shifts within each 16-lane row, then the row
// totals (lanes 15, 31, 47) are added to the rows above them.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = __lane_id();
  const int r = l & 15;
  T t = row_shr<1>(v);
  if (r >= 1) v += t;
  t = row_shr<2>(v);
  if (r >= 2) v += t;
  t = row_shr<4>(v);
  if (r >= 4) v += t;
  t = row_shr<8>(v);
  if (r >= 8) v += t;
  const T s0 = read_lane(v, 15), s1 = read_lane(v, 31), s2 = read_lane(v, 47);
  const int row = l >> 4;
  if (row >= 1) v += s0;
  if (row >= 2) v += s1;
  if (row >= 3) v += s2;
  return v;
}

// Sum over the wave, returned in every lane.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += xor_lanes_t<1>(v);
  v += xor_lanes_t<2>(v);
  v += xor_lanes_t<4>(v);
  v += xor_lanes_t<8>(v);
  v += xor_lanes_t<16>(v);
  return read_lane(v, 0) + read_lane(v, 32);
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
