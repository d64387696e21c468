// vl_stitch.hip — seam resolution: side-stream partition and the x-seam stitch.
//
// Strips whose run touches an internal row border of their tile (SEAM_B/T)
// are pieces of runs that continue vertically; they are re-tiled at the next
// level with coarser rows.  Strips that only touch internal column borders
// (SEAM_L/R) carry an exact global run and only need joining along x: a strip
// ending at x=s and a strip starting at x=s with the same [y1,y2) are the same
// maximal strip (runs at one x are disjoint, so (s, y1) names at most one
// start).  Starts are radix-sorted by (x1,y1); every end looks up its
// successor by binary search; chain heads walk their successors and emit the
// joined strip.
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

// Starts are the strips on a left seam: flag, then (after a scan) their keys
// and indices at the scanned positions, so only starts are sorted.
__global__ void k_start_flags(const uint8_t* __restrict__ tag, int64_t n, uint32_t* flag) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    flag[i] = (tag[i] & SEAM_L) ? 1u : 0u;
}

__global__ void k_start_keys(const vl_box* __restrict__ s, const uint8_t* __restrict__ tag, int64_t n,
                             const uint32_t* __restrict__ pre, int64_t x0, int64_t y0, int yb, uint64_t* keys,
                             uint32_t* vals) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (tag[i] & SEAM_L) {
      keys[pre[i]] = rel_key(s[i].x1, s[i].y1, x0, y0, yb);
      vals[pre[i]] = (uint32_t)i;
    }
}

__global__ void k_start_count(const uint32_t* pre, const uint8_t* tag, int64_t n, uint64_t* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = n ? (uint64_t)pre[n - 1] + ((tag[n - 1] & SEAM_L) ? 1u : 0u) : 0u;
}

__global__ void k_stitch_link(const vl_box* __restrict__ s, const uint8_t* __restrict__ tag, int64_t n,
                              const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ svals, int64_t ns,
                              int64_t x0, int64_t y0, int yb, int32_t* next, uint8_t* has_prev) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t nx = -1;
    if (tag[i] & SEAM_R) {
      const vl_box b = s[i];
      const uint64_t k = rel_key(b.x2, b.y1, x0, y0, yb);
      int64_t lo = 0, hi = ns;
      while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (skeys[mid] < k) lo = mid + 1; else hi = mid;
      }
      if (lo < ns && skeys[lo] == k) {
        const uint32_t j = svals[lo];
        if (s[j].y2 == b.y2) {
          nx = (int32_t)j;
          has_prev[j] = 1;
        }
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

void launch_stitch_starts(const vl_box* s, const uint8_t* tag, int64_t n, int64_t x0, int64_t y0, int yb,
                          uint64_t* keys, uint32_t* vals, uint64_t* count, void* tmp, hipStream_t st) {
  if (n <= 0) {
    launch_fill_u64(count, 0ull, 1, st);
    return;
  }
  uint32_t* flag = (uint32_t*)tmp;
  uint32_t* pre = flag + n;
  char* stmp = (char*)(((uintptr_t)(pre + n) + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL(k_start_flags, dim3(grid_for(n)), dim3(256), 0, st, tag, n, flag);
  scan_exclusive_u32(flag, pre, (size_t)n, nullptr, stmp, st);
  hipLaunchKernelGGL(k_start_keys, dim3(grid_for(n)), dim3(256), 0, st, s, tag, n, pre, x0, y0, yb, keys, vals);
  hipLaunchKernelGGL(k_start_count, dim3(1), dim3(64), 0, st, pre, tag, n, count);
}

void launch_stitch_link(const vl_box* s, const uint8_t* tag, int64_t n, const uint64_t* skeys,
                        const uint32_t* svals, int64_t ns, int64_t x0, int64_t y0, int yb, int32_t* next,
                        uint8_t* has_prev, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u8, dim3(grid_for(n)), dim3(256), 0, st, has_prev, n);
  hipLaunchKernelGGL(k_stitch_link, dim3(grid_for(n)), dim3(256), 0, st, s, tag, n, skeys, svals, ns, x0, y0, yb,
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
