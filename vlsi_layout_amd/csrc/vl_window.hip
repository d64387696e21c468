// vl_window.hip — window extraction from a cell-indexed layer, and an
// order-independent hash of a box set (used by the query server).
//
// A layer index is the layer's boxes binned into a uniform grid of cells
// (the tile binning of vl_tile.hip, cells in Morton order): entries in cell
// order plus the start of every cell.  A window query gathers the entry
// ranges of the runs of cells the window covers, clips every candidate to the
// window and compacts the ones that keep a positive area.  A box spanning
// several cells has a copy in each; only the copy in the cell holding the
// lower-left corner of the box's part inside the window is kept, so the
// window's boxes come out once each.
#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

// Candidate i of the concatenated ranges: range r with off[r] <= i < off[r+1].
__device__ __forceinline__ int range_of(const uint64_t* off, int nr, uint64_t i) {
  int lo = 0, hi = nr;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_window_clip(const vl_box* __restrict__ entries,
                                                     const uint64_t* __restrict__ rbeg,
                                                     const uint64_t* __restrict__ roff, int nr, uint64_t n,
                                                     vl_box w, const uint64_t* __restrict__ start, Grid g,
                                                     vl_box* __restrict__ clip, uint32_t* __restrict__ keep) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const int r = range_of(roff, nr, i);
    const uint64_t e = rbeg[r] + (i - roff[r]);
    const vl_box b = entries[e];
    vl_box c;
    c.x1 = max(b.x1, w.x1);
    c.y1 = max(b.y1, w.y1);
    c.x2 = min(b.x2, w.x2);
    c.y2 = min(b.y2, w.y2);
    bool ok = c.x1 < c.x2 && c.y1 < c.y2;
    if (ok) {
      // the cell of (c.x1, c.y1): inside the domain, since the box is
      const int cc = (int)(((int64_t)c.x1 - g.x0) / g.tw), rr = (int)(((int64_t)c.y1 - g.y0) / g.th);
      const uint64_t t = tile_number(rr, cc, g);
      ok = start[2 * t] <= e && e < start[2 * t + 1];
    }
    clip[i] = c;
    keep[i] = ok ? 1u : 0u;
  }
}

__global__ __launch_bounds__(256) void k_window_compact(const vl_box* __restrict__ clip,
                                                        const uint32_t* __restrict__ keep,
                                                        const uint32_t* __restrict__ pos, uint64_t n,
                                                        vl_box* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (keep[i]) out[pos[i]] = clip[i];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {   // SplitMix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_box_hash(const vl_box* __restrict__ b, int64_t n,
                                                  unsigned long long* out) {
  __shared__ unsigned long long ws[256 / WAVE];
  unsigned long long h = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const vl_box v = b[i];
    const uint64_t lo = ((uint64_t)(uint32_t)v.x1 << 32) | (uint32_t)v.y1;
    const uint64_t hi = ((uint64_t)(uint32_t)v.x2 << 32) | (uint32_t)v.y2;
    h += mix64(lo ^ mix64(hi));
  }
  h = wave_sum(h);
  if (__lane_id() == 0) ws[threadIdx.x / WAVE] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int k = 0; k < 256 / WAVE; k++) s += ws[k];
    if (s) atomicAdd(out, s);
  }
}

unsigned grid_of(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  return (unsigned)g;
}

}  // namespace

void launch_window_clip(const vl_box* entries, const uint64_t* rbeg, const uint64_t* roff, int nr, uint64_t n,
                        vl_box w, const uint64_t* start, Grid g, vl_box* clip, uint32_t* keep, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_window_clip, dim3(grid_of(n)), dim3(256), 0, st, entries, rbeg, roff, nr, n, w, start, g,
                       clip, keep);
}

void launch_window_compact(const vl_box* clip, const uint32_t* keep, const uint32_t* pos, uint64_t n, vl_box* out,
                           hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_window_compact, dim3(grid_of(n)), dim3(256), 0, st, clip, keep, pos, n, out);
}

void launch_box_hash(const vl_box* b, int64_t n, unsigned long long* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_box_hash, dim3(grid_of((uint64_t)n)), dim3(256), 0, st, b, n, out);
}

}  // namespace vl
