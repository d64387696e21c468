// vl_stitch.hip — seam resolution: side-stream partition and the x-seam stitch.
//
// Strips whose run touches an internal row border of their tile (SEAM_B/T)
// are pieces of runs that continue vertically; they are re-tiled at the next
// level with coarser rows.  Strips that only touch internal column borders
// (SEAM_L/R) carry an exact global run and only need joining along x: a strip
// ending at x=s and a strip starting at x=s with the same [y1,y2) are the same
// maximal strip (the pieces are disjoint, so the corner (s, y1) names at most
// one start).  Starts go into an open-addressing hash table keyed by their
// lower-left corner (64-bit CAS, linear probing); every end looks up the
// start at its lower-right corner; chain heads walk their successors and emit
// the joined strip.  The result does not depend on the table's layout.
#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (unsigned)g;
}

__device__ __forceinline__ uint64_t corner_key(int32_t x, int32_t y) {
  return ((uint64_t)((uint32_t)x ^ 0x80000000u) << 32) | ((uint32_t)y ^ 0x80000000u);
}

__global__ void k_flag_border(const uint8_t* __restrict__ tag, int64_t n, uint32_t* flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (tag[i] & (SEAM_B | SEAM_T)) ? 1u : 0u;
}

__global__ void k_partition(const vl_box* __restrict__ side, const uint8_t* __restrict__ tag, int64_t n,
                            const uint32_t* __restrict__ pre, vl_box* next, vl_box* stitch, uint8_t* stitch_tag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t t = tag[i];
    if (t & (SEAM_B | SEAM_T)) {
      next[pre[i]] = side[i];
    } else {
      const int64_t k = i - pre[i];
      stitch[k] = side[i];
      stitch_tag[k] = t;
    }
  }
}

__global__ void k_part_counts(const uint32_t* pre, const uint8_t* tag, int64_t n, uint64_t* counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint64_t nb = n ? (uint64_t)pre[n - 1] + ((tag[n - 1] & (SEAM_B | SEAM_T)) ? 1u : 0u) : 0u;
    counts[0] = nb;
    counts[1] = (uint64_t)n - nb;
  }
}

// Key of a corner relative to the domain origin: x - x0 above the low yb bits,
// y - y0 in them (both below 2^32 for int32 coordinates).
__device__ __forceinline__ uint64_t rel_key(int32_t x, int32_t y, int64_t x0, int64_t y0, int yb) {
  return ((uint64_t)((int64_t)x - x0) << yb) | (uint64_t)((int64_t)y - y0);
}

// Hash table of the starts: 2^k slots (at least twice the strips), key
// HASH_EMPTY = free (no corner key reaches it: x - x0 < 2^32 above y - y0).
constexpr uint64_t HASH_EMPTY = ~0ull;

__device__ __forceinline__ uint64_t hash_slot(uint64_t k, uint64_t mask) {
  k ^= k >> 31;
  k *= 0x9E3779B97F4A7C15ull;
  k ^= k >> 29;
  return k & mask;
}

__global__ void k_hash_starts(const vl_box* __restrict__ s, const uint8_t* __restrict__ tag, int64_t n, int64_t x0,
                              int64_t y0, int yb, unsigned long long* __restrict__ tk, uint32_t* __restrict__ tv,
                              uint64_t mask) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (!(tag[i] & SEAM_L)) continue;
    const unsigned long long key = rel_key(s[i].x1, s[i].y1, x0, y0, yb);
    uint64_t slot = hash_slot(key, mask);
    for (uint64_t probe = 0; probe <= mask; probe++) {   // fewer starts than slots: ends at a free slot
      const unsigned long long prev = atomicCAS(&tk[slot], (unsigned long long)HASH_EMPTY, key);
      if (prev == HASH_EMPTY || prev == key) {
        tv[slot] = (uint32_t)i;
        break;
      }
      slot = (slot + 1) & mask;
    }
  }
}

__global__ void k_hash_link(const vl_box* __restrict__ s, const uint8_t* __restrict__ tag, int64_t n, int64_t x0,
                            int64_t y0, int yb, const unsigned long long* __restrict__ tk,
                            const uint32_t* __restrict__ tv, uint64_t mask, int32_t* next, uint8_t* has_prev) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t nx = -1;
    if (tag[i] & SEAM_R) {
      const vl_box b = s[i];
      const unsigned long long key = rel_key(b.x2, b.y1, x0, y0, yb);
      uint64_t slot = hash_slot(key, mask);
      for (uint64_t probe = 0; probe <= mask; probe++) {
        const unsigned long long k = tk[slot];
        if (k == HASH_EMPTY) break;
        if (k == key) {
          const uint32_t j = tv[slot];
          if (s[j].y2 == b.y2) {
            nx = (int32_t)j;
            has_prev[j] = 1;
          }
          break;
        }
        slot = (slot + 1) & mask;
      }
    }
    next[i] = nx;
  }
}

__global__ void k_zero_u8(uint8_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_head_flags(const uint8_t* __restrict__ has_prev, int64_t n, uint32_t* flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = has_prev[i] ? 0u : 1u;
}

__global__ void k_stitch_emit(const vl_box* __restrict__ s, int64_t n, const int32_t* __restrict__ next,
                              const uint8_t* __restrict__ has_prev, const uint32_t* __restrict__ pre,
                              vl_box* out, uint64_t out_base) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (has_prev[i]) continue;
    vl_box b = s[i];
    int64_t cur = i;
    int64_t guard = 0;
    while (next[cur] >= 0 && guard++ < n) cur = next[cur];
    b.x2 = s[cur].x2;
    out[out_base + pre[i]] = b;
  }
}

__global__ void k_head_count(const uint32_t* pre, const uint8_t* has_prev, int64_t n, uint64_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *count = n ? (uint64_t)pre[n - 1] + (has_prev[n - 1] ? 0u : 1u) : 0u;
}

}  // namespace

size_t partition_tmp_bytes(int64_t n) {
  return (size_t)(n > 0 ? n : 1) * sizeof(uint32_t) * 2 + scan_tmp_bytes((size_t)(n > 0 ? n : 1)) + 512;
}

void launch_partition_side(const vl_box* side, const uint8_t* tag, int64_t n, vl_box* next, vl_box* stitch,
                           uint8_t* stitch_tag, uint64_t* counts, void* tmp, hipStream_t st) {
  if (n <= 0) {
    launch_fill_u64(counts, 0ull, 2, st);
    return;
  }
  uint32_t* flag = (uint32_t*)tmp;
  uint32_t* pre = flag + n;
  char* stmp = (char*)(((uintptr_t)(pre + n) + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL(k_flag_border, dim3(grid_for(n)), dim3(256), 0, st, tag, n, flag);
  scan_exclusive_u32(flag, pre, (size_t)n, nullptr, stmp, st);
  hipLaunchKernelGGL(k_partition, dim3(grid_for(n)), dim3(256), 0, st, side, tag, n, pre, next, stitch, stitch_tag);
  hipLaunchKernelGGL(k_part_counts, dim3(1), dim3(64), 0, st, pre, tag, n, counts);
}

uint64_t stitch_table_slots(int64_t n) {
  uint64_t h = 1024;
  while (h < 2 * (uint64_t)(n > 0 ? n : 1)) h <<= 1;
  return h;
}

void launch_stitch_link(const vl_box* s, const uint8_t* tag, int64_t n, int64_t x0, int64_t y0, int yb,
                        unsigned long long* tk, uint32_t* tv, int32_t* next, uint8_t* has_prev, hipStream_t st) {
  if (n <= 0) return;
  const uint64_t slots = stitch_table_slots(n);
  launch_fill_u64((uint64_t*)tk, HASH_EMPTY, slots, st);
  hipLaunchKernelGGL(k_zero_u8, dim3(grid_for(n)), dim3(256), 0, st, has_prev, n);
  hipLaunchKernelGGL(k_hash_starts, dim3(grid_for(n)), dim3(256), 0, st, s, tag, n, x0, y0, yb, tk, tv, slots - 1);
  hipLaunchKernelGGL(k_hash_link, dim3(grid_for(n)), dim3(256), 0, st, s, tag, n, x0, y0, yb, tk, tv, slots - 1,
                     next, has_prev);
}

void launch_stitch_emit(const vl_box* s, int64_t n, const int32_t* next, const uint8_t* has_prev, vl_box* out,
                        uint64_t out_base, void* tmp, uint64_t* count, hipStream_t st) {
  if (n <= 0) {
    launch_fill_u64(count, 0ull, 1, st);
    return;
  }
  uint32_t* flag = (uint32_t*)tmp;
  uint32_t* pre = flag + n;
  char* stmp = (char*)(((uintptr_t)(pre + n) + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL(k_head_flags, dim3(grid_for(n)), dim3(256), 0, st, has_prev, n, flag);
  scan_exclusive_u32(flag, pre, (size_t)n, nullptr, stmp, st);
  hipLaunchKernelGGL(k_stitch_emit, dim3(grid_for(n)), dim3(256), 0, st, s, n, next, has_prev, pre, out, out_base);
  hipLaunchKernelGGL(k_head_count, dim3(1), dim3(64), 0, st, pre, has_prev, n, count);
}

}  // namespace vl
