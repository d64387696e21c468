// vl_pnum.hip — polygon numbering and overlap certification of a layer
// (the PNUM state of the reference's PreprocessLayer).
//
// Box i and box k "touch" when their closed extents meet:
// max(x1) <= min(x2) and max(y1) <= min(y2), so shared edges and corners
// count (PerformMerge, src/pnummerge.f90:279-385).  The connected components
// of that relation are numbered the reference's way: every pair is united in
// a union-find that hooks the larger root under the smaller one, so a
// component's root is its smallest box index, and the components with two or
// more boxes get the labels 1, 2, ... in the order of their roots
// (uf_fullreduce, src/datastructures.f90:143-171); a box that touches no
// other box keeps 0.
//
// Pairs are found per tile of a uniform grid over the layer, the boxes binned
// by their closed extent (x2 and y2 included), a wave per tile.  A pair is
// taken in exactly one tile: the one holding the lower-left corner
// (max x1, max y1) of the closed intersection, which both boxes reach.  Per
// pair: the union (lock-free: compare-and-swap on the larger root's parent),
// the touch flags of both boxes, and the exact int64 sums of the positive
// intersection areas and of the perimeters 2 (w + h) of the zero-area ones
// (the reference accumulates both in real64).  A second mode only counts the
// pairs that overlap with positive area (ComputeInteractionsGPU,
// src/interaction.f90:204-310).
#include "vl_tile_dev.h"

namespace vl {

namespace {

__device__ __forceinline__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (p != x) {
    // path halving: point x at its grandparent
    const uint32_t g = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g != p) atomicCAS(&parent[x], p, g);
    x = p;
    p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return x;
}

// Unites the sets of a and b: the larger root is hooked under the smaller.
__device__ __forceinline__ void uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    // a > b: a's parent becomes b, unless a stopped being a root meanwhile
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

struct PairArgs {
  const TileDesc* desc;
  int ntiles;
  const uint32_t* idx;       // entries: box indices, tile by tile
  const vl_box* boxes;
  Grid g;
  int mode;                  // 0: PNUM (unions, flags, sums); 1: count positive-area pairs
  uint32_t* parent;
  uint32_t* touched;
  unsigned long long* sums;  // [0] positive areas, [1] touch perimeters, [2] positive-area pairs
};

__global__ __launch_bounds__(256) void k_pairs(PairArgs a) {
  const int lane = __lane_id();
  const int nwaves = gridDim.x * (blockDim.x / WAVE);
  long long area = 0, perim = 0, npos = 0;
  for (int t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; t < a.ntiles; t += nwaves) {
    const int64_t off = uniform(a.desc[t].off);
    const int n = uniform(a.desc[t].nA);
    const int32_t tx0 = uniform(a.desc[t].x0), ty0 = uniform(a.desc[t].y0);
    const int32_t tx1 = uniform(a.desc[t].x1), ty1 = uniform(a.desc[t].y1);
    if (n < 2) continue;
    for (int jb = 0; jb < n; jb += WAVE) {
      const int j = jb + lane;
      uint32_t bj = 0;
      vl_box B{0, 0, 0, 0};
      if (j < n) {
        bj = a.idx[off + j];
        B = a.boxes[bj];
      }
      for (int i = 0; i < n; i++) {
        const uint32_t bi = uniform(a.idx[off + i]);
        const vl_box A = a.boxes[bi];   // the same address in every lane
        if (j >= n || bj <= bi) continue;   // each unordered pair once: bi < bj
        const int32_t ix1 = max(A.x1, B.x1), iy1 = max(A.y1, B.y1);
        const int32_t ix2 = min(A.x2, B.x2), iy2 = min(A.y2, B.y2);
        if (ix1 > ix2 || iy1 > iy2) continue;                 // closed extents apart
        // the pair belongs to the tile holding (ix1, iy1) (the last column /
        // row also holds the far border when the domain ends at INT32_MAX)
        const bool xin = ix1 >= tx0 && (ix1 < tx1 || (ix1 == tx1 && (int64_t)tx1 == a.g.x1));
        const bool yin = iy1 >= ty0 && (iy1 < ty1 || (iy1 == ty1 && (int64_t)ty1 == a.g.y1));
        if (!xin || !yin) continue;
        const bool positive = ix1 < ix2 && iy1 < iy2;
        if (a.mode == 1) {
          npos += positive;
          continue;
        }
        uf_unite(a.parent, bi, bj);
        a.touched[bi] = 1u;
        a.touched[bj] = 1u;
        if (positive) area += (long long)(ix2 - ix1) * (long long)(iy2 - iy1);
        else perim += 2 * ((long long)(ix2 - ix1) + (long long)(iy2 - iy1));
      }
    }
  }
  area = wave_sum(area);
  perim = wave_sum(perim);
  npos = wave_sum(npos);
  if (lane == 0) {
    if (area) atomicAdd(&a.sums[0], (unsigned long long)area);
    if (perim) atomicAdd(&a.sums[1], (unsigned long long)perim);
    if (npos) atomicAdd(&a.sums[2], (unsigned long long)npos);
  }
}

__global__ void k_uf_init(uint32_t* parent, uint32_t* touched, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    parent[i] = (uint32_t)i;
    touched[i] = 0u;
  }
}

// root[i] = the final root of i; is_root[i] = 1 for the roots of components
// with two or more boxes.
__global__ void k_uf_roots(uint32_t* parent, const uint32_t* touched, uint32_t* root, uint32_t* is_root, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t r = uf_find(parent, (uint32_t)i);
    root[i] = r;
    is_root[i] = (touched[i] && r == (uint32_t)i) ? 1u : 0u;
  }
}

// labels[i] = rank[root[i]] + 1 for touched boxes, else 0.
__global__ void k_uf_labels(const uint32_t* root, const uint32_t* touched, const uint32_t* rank, int64_t* labels,
                            int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    labels[i] = touched[i] ? (int64_t)rank[root[i]] + 1 : 0;
}

unsigned grid_of(int64_t n, int per) {
  int64_t g = (n + per - 1) / per;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

}  // namespace

void launch_uf_init(uint32_t* parent, uint32_t* touched, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_uf_init, dim3(grid_of(n, 256)), dim3(256), 0, st, parent, touched, n);
}

void launch_pairs(const TileDesc* desc, int ntiles, const uint32_t* idx, const vl_box* boxes, Grid g, int mode,
                  uint32_t* parent, uint32_t* touched, unsigned long long* sums, hipStream_t st) {
  if (ntiles <= 0) return;
  PairArgs a{desc, ntiles, idx, boxes, g, mode, parent, touched, sums};
  hipLaunchKernelGGL(k_pairs, dim3(grid_of(ntiles, 4)), dim3(256), 0, st, a);
}

void launch_uf_roots(uint32_t* parent, const uint32_t* touched, uint32_t* root, uint32_t* is_root, int64_t n,
                     hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_uf_roots, dim3(grid_of(n, 256)), dim3(256), 0, st, parent, touched, root, is_root, n);
}

void launch_uf_labels(const uint32_t* root, const uint32_t* touched, const uint32_t* rank, int64_t* labels, int64_t n,
                      hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_uf_labels, dim3(grid_of(n, 256)), dim3(256), 0, st, root, touched, rank, labels, n);
}

}  // namespace vl
