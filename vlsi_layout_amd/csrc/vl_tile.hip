// vl_tile.hip — tile binning and the per-tile Boolean kernels for gfx950.
//
// Binning: every box whose interior meets a tile of a uniform grid (or, for
// erosion, the tile grown by the erosion distance) is copied, after the op's
// per-box transform, into that tile's entry range.  The ranges come from a
// stable radix sort of (tile, layer) keys, so binning needs no atomics.
//
// Per-tile kernels.  A tile's result inside the tile is computed on the
// compressed grid of its own coordinates:
//  1. clip the tile's boxes to the tile; sort their x and y coordinates (plus
//     the tile borders) and drop duplicates: nx x-breakpoints give ncol = nx-1
//     slabs, ny y-breakpoints give nrows = ny-1 elementary y-intervals;
//  2. rasterise each layer into per-slab coverage bitmasks (one 64-bit word
//     per 64 elementary intervals) with LDS atomics, and combine the layers
//     word by word with the operator (AND, A&~B, OR, XOR);
//  3. at every slab boundary k, the change mask D = dilate1(R_k ^ R_{k-1})
//     marks where runs differ: a run of R_k that meets D starts a strip at k
//     (mark S_k at the clear row just above the run), a run of R_{k-1} that
//     meets D ends one (mark E_k).  These are exactly the maximal strips of
//     heal_boxes (src/polygon_fracture.f90:881-960) inside the tile;
//  4. the strips touching an internal tile border are seam pieces (SEAM_* tag,
//     side stream); the rest are final.  In area mode step 3 is replaced by
//     the int64 sum of run lengths times slab widths
//     (calculate_union_area_fast, src/geometry.f90:1840-1927).
//
// The main kernel runs one wave64 per tile: the coordinate sort is a bitonic
// network in registers (shuffles across lanes), duplicates are dropped with a
// wave scan, each lane owns slab boundaries {lane, lane+64, ...} and marks
// S_k/E_k with one multi-word addition per column (the run's D bits carry up
// to the row above it); the E marks are transposed so that a lane holds one
// row's end boundaries, and every S mark is paired with the first E mark to
// its right in its row.  Tiles above its capacity run on the 256-thread
// kernel (capacity 255); larger tiles are split into x-sub-columns by the
// host.  Each workgroup writes its strips into chunks of a scratch stream; a
// reorder pass restores tile order.
#include <utility>

#include "vl_tile_dev.h"

namespace vl {

namespace {

// Quotients of grid offsets (below 2^32) by a tile size d >= 1 without a
// division per box: m = floor(2^32 / d) gives floor(x * m / 2^32) = q or q - 1
// for q = floor(x / d), and one comparison of the remainder settles it.
struct Div32 {
  uint32_t d, m;
  int mode;   // 0: 2 <= d <= 2^31; 1: d == 1; 2: 2^31 < d < 2^32 (q is 0 or 1); 3: d >= 2^32 (q = 0)
};

__device__ __forceinline__ Div32 make_div(int64_t d) {
  Div32 v{1u, 0u, 1};
  if (d >= ((int64_t)1 << 32)) v = Div32{0u, 0u, 3};
  else if (d > ((int64_t)1 << 31)) v = Div32{(uint32_t)d, 0u, 2};
  else if (d >= 2) v = Div32{(uint32_t)d, (uint32_t)((1ull << 32) / (uint64_t)d), 0};
  return v;
}

__device__ __forceinline__ uint32_t div_q(uint32_t x, const Div32& v) {
  if (v.mode == 0) {
    uint32_t q = __umulhi(x, v.m);
    if (x - q * v.d >= v.d) q++;
    return q;
  }
  if (v.mode == 1) return x;
  if (v.mode == 2) return x >= v.d ? 1u : 0u;
  return 0u;
}

// The tile sizes of grid g as reciprocals (computed once per thread).
struct GridDiv {
  Div32 x, y;
};
__device__ __forceinline__ GridDiv grid_div(const Grid& g) { return GridDiv{make_div(g.tw), make_div(g.th)}; }

// Tile index range of box b grown by xr (64-bit arithmetic, clipped to the grid).
__device__ __forceinline__ bool tile_range(const vl_box& b, const Xform& xr, const Grid& g, const GridDiv& gd,
                                           int& c1, int& c2, int& r1, int& r2) {
  const int64_t bx1 = max((int64_t)b.x1 - xr.w, g.x0), bx2 = min((int64_t)b.x2 + xr.e, g.x1);
  const int64_t by1 = max((int64_t)b.y1 - xr.s, g.y0), by2 = min((int64_t)b.y2 + xr.n, g.y1);
  if (bx1 >= bx2 || by1 >= by2) return false;
  const uint64_t ox2 = (uint64_t)(bx2 - 1 - g.x0), oy2 = (uint64_t)(by2 - 1 - g.y0);
  if (((ox2 | oy2) >> 32) == 0) {
    // offsets below 2^32 (every grid whose domain lies in the int32 range)
    c1 = (int)div_q((uint32_t)(bx1 - g.x0), gd.x);
    c2 = (int)div_q((uint32_t)ox2, gd.x);
    r1 = (int)div_q((uint32_t)(by1 - g.y0), gd.y);
    r2 = (int)div_q((uint32_t)oy2, gd.y);
  } else {
    c1 = (int)((bx1 - g.x0) / g.tw);
    c2 = (int)((int64_t)ox2 / g.tw);
    r1 = (int)((by1 - g.y0) / g.th);
    r2 = (int)((int64_t)oy2 / g.th);
  }
  return true;
}

__device__ __forceinline__ vl_box xform(vl_box b, const Xform& x) {
  b.x1 -= x.w;
  b.y1 -= x.s;
  b.x2 += x.e;
  b.y2 += x.n;
  return b;
}

// Binning is a stable sort of (tile, layer) keys: every (box, tile) pair gets
// the key 2*tile + layer, an LSD radix sort orders the pairs, and each tile's
// entries (layer A, then layer B, each in input order) end up contiguous.  No
// atomics, and the entry order is deterministic.
//   span:   tiles per box (layer A boxes 0..na-1, layer B boxes na..na+nb-1)
//   emit:   keys/values at the scanned positions
//   bounds: start[k] = first position with key >= k
//   gather: entries[i] = stored form of box vals[i]
__global__ __launch_bounds__(256) void k_bin_span(const vl_box* __restrict__ a, int64_t na, Xform xra,
                                                  const vl_box* __restrict__ b, int64_t nb, Grid g,
                                                  uint32_t* __restrict__ cnt) {
  const int64_t n = na + nb;
  const GridDiv gd = grid_div(g);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int c1, c2, r1, r2;
    const bool isA = i < na;
    const vl_box v = isA ? a[i] : b[i - na];
    uint32_t k = 0;
    if (tile_range(v, isA ? xra : Xform{0, 0, 0, 0}, g, gd, c1, c2, r1, r2))
      k = (uint32_t)(c2 - c1 + 1) * (uint32_t)(r2 - r1 + 1);
    cnt[i] = k;
  }
}

__global__ __launch_bounds__(256) void k_bin_emit(const vl_box* __restrict__ a, int64_t na, Xform xra,
                                                  const vl_box* __restrict__ b, int64_t nb, Grid g,
                                                  const uint64_t* __restrict__ pos, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
  const int64_t n = na + nb;
  const GridDiv gd = grid_div(g);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int c1, c2, r1, r2;
    const bool isA = i < na;
    const vl_box v = isA ? a[i] : b[i - na];
    if (!tile_range(v, isA ? xra : Xform{0, 0, 0, 0}, g, gd, c1, c2, r1, r2)) continue;
    uint64_t p = pos[i];
    const uint32_t layer = isA ? 0u : 1u;
    for (int r = r1; r <= r2; r++)
      for (int c = c1; c <= c2; c++) {
        keys[p] = (tile_number(r, c, g) << 1) | layer;
        vals[p] = (uint32_t)i;
        p++;
      }
  }
}

__global__ __launch_bounds__(256) void k_bin_bounds(const uint32_t* __restrict__ keys, uint64_t total,
                                                    int64_t nkeys, uint64_t* __restrict__ start) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= nkeys;
       k += (int64_t)gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = total;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((int64_t)keys[mid] < k) lo = mid + 1; else hi = mid;
    }
    start[k] = lo;
  }
}

__global__ __launch_bounds__(256) void k_bin_gather(const vl_box* __restrict__ a, int64_t na, Xform xsa,
                                                    const vl_box* __restrict__ b, const uint32_t* __restrict__ vals,
                                                    uint64_t total, vl_box* __restrict__ entries) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t v = vals[i];
    entries[i] = v < na ? xform(a[v], xsa) : b[v - na];
  }
}

// Atomic binning (the default): a count pass and a scatter pass over the
// boxes in input order.  The (tile, layer) key of every (box, tile) pair is
// 2*tile + layer; lanes of a wave whose keys are equal and adjacent form a
// segment, and only the segment's first lane touches the key's counter (one
// atomic per segment: inputs produced by earlier operators are in tile order,
// so neighbouring boxes usually share a tile).  The scatter pass takes slots
// from the back of each key's range by counting the counters down to zero.
// Entries of one tile end up in an order that depends on scheduling; every
// tile kernel's result depends only on the set of its entries, so results
// and their order do not.
struct Seg {
  int head;   // lane of this lane's segment head
  int rank;   // position in the segment
  int len;    // segment length (valid in the head lane)
  bool is_head;
};

__device__ __forceinline__ Seg segment_of(bool act, uint32_t key, int lane) {
  const uint32_t pk = (uint32_t)__shfl_up((int)key, 1, WAVE);
  const int pa = __shfl_up((int)act, 1, WAVE);
  Seg s;
  s.is_head = act && (lane == 0 || !pa || pk != key);
  const uint64_t hm = __ballot(s.is_head);
  const uint64_t am = __ballot(act);
  const uint64_t upto = (2ull << lane) - 1ull;   // lanes 0..lane (lane 63: all)
  const uint64_t below = hm & upto;
  s.head = below ? 63 - __clzll((long long)below) : lane;
  s.rank = lane - s.head;
  const uint64_t stop = (hm | ~am) & ~upto;      // first head or inactive lane above
  s.len = (stop ? __ffsll((unsigned long long)stop) - 1 : WAVE) - lane;
  return s;
}

// Grouped form: the active lanes of the wave are grouped by equal key, not
// only adjacent equal keys, one group per distinct key (up to GROUP_ITERS
// keys; lanes left after that are groups of their own).  Each group's leader
// issues its atomic, all leaders in one instruction.  Boxes produced in
// another grid's tile order interleave the tiles of this grid, so runs are
// short while the wave still holds few distinct keys.
constexpr int GROUP_ITERS = 16;

struct Grp {
  int leader;   // lane of this lane's group leader
  int rank;     // position in the group
  int len;      // group size (valid in every member)
  bool is_leader;
};

__device__ __forceinline__ Grp group_of(bool act, uint32_t key, int lane) {
  uint64_t mine = 0;   // this lane's group (its lanes)
  uint64_t pend = __ballot(act);
  for (int it = 0; pend && it < GROUP_ITERS; it++) {
    const int ld = __ffsll((unsigned long long)pend) - 1;
    const uint32_t k = (uint32_t)__builtin_amdgcn_readlane((int)key, ld);
    const bool in = act && key == k;
    const uint64_t grp = __ballot(in);   // every lane of key k is still pending
    if (in) mine = grp;
    pend &= ~grp;
  }
  Grp r{lane, 0, 1, act};
  if (mine) {
    r.leader = __ffsll((unsigned long long)mine) - 1;
    r.rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mine >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mine, 0u));
    r.len = __popcll(mine);
    r.is_leader = r.leader == lane;
  }
  return r;
}

struct Lead {
  int leader;   // lane of the group's leader
  int rank;     // position in the group
  int len;      // group size
  bool lead;    // this lane issues the group's atomic
};

template <bool GROUPED>
__device__ __forceinline__ Lead lead_of(bool act, uint32_t key, int lane) {
  if constexpr (GROUPED) {
    const Grp gr = group_of(act, key, lane);
    return Lead{gr.leader, gr.rank, gr.len, act && gr.is_leader};
  } else {
    const Seg sg = segment_of(act, key, lane);
    return Lead{sg.head, sg.rank, sg.len, sg.is_head};
  }
}

// A wave bins BIN_U consecutive 64-box chunks per iteration: all their loads
// are issued first, and (in the scatter) the atomics of every chunk's first
// tile are in flight together, so the wave waits on memory once per
// iteration rather than once per chunk.
constexpr int BIN_U = 4;

__device__ __forceinline__ uint32_t bin_key(int r, int c, const Grid& g, uint32_t layer) {
  return (tile_number(r, c, g) << 1) | layer;
}

template <bool GROUPED>
__global__ __launch_bounds__(256) void k_bin_count(const vl_box* __restrict__ a, int64_t na, Xform xra,
                                                   const vl_box* __restrict__ b, int64_t nb, Grid g,
                                                   uint32_t* __restrict__ cnt) {
  const int64_t n = na + nb;
  const int lane = __lane_id();
  const GridDiv gd = grid_div(g);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / WAVE);
  for (int64_t base = wave * (WAVE * BIN_U); base < n; base += nwaves * (WAVE * BIN_U)) {
    vl_box v[BIN_U];
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t i = base + u * WAVE + lane;
      v[u] = vl_box{0, 0, 0, 0};
      if (i < n) v[u] = i < na ? a[i] : b[i - na];
    }
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t i = base + u * WAVE + lane;
      const bool isA = i < na;
      int c1 = 0, c2 = -1, r1 = 0, r2 = -1;
      if (i < n && !tile_range(v[u], isA ? xra : Xform{0, 0, 0, 0}, g, gd, c1, c2, r1, r2)) { c2 = c1 - 1; r2 = r1 - 1; }
      const int m = (c2 >= c1 && r2 >= r1) ? (c2 - c1 + 1) * (r2 - r1 + 1) : 0;
      const uint32_t layer = isA ? 0u : 1u;
      int c = c1, r = r1;   // the box's tiles in row-major order
      for (int jt = 0; __ballot(jt < m); jt++) {
        const bool act = jt < m;
        const uint32_t key = act ? bin_key(r, c, g, layer) : 0u;
        const Lead ld = lead_of<GROUPED>(act, key, lane);
        if (ld.lead) atomicAdd(&cnt[key], (uint32_t)ld.len);
        if (++c > c2) { c = c1; ++r; }
      }
    }
  }
}

template <bool GROUPED>
__global__ __launch_bounds__(256) void k_bin_scatter(const vl_box* __restrict__ a, int64_t na, Xform xra, Xform xsa,
                                                     const vl_box* __restrict__ b, int64_t nb, Grid g,
                                                     const uint64_t* __restrict__ start, uint32_t* __restrict__ cnt,
                                                     vl_box* __restrict__ entries, uint32_t* __restrict__ idx) {
  const int64_t n = na + nb;
  const int lane = __lane_id();
  const GridDiv gd = grid_div(g);
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / WAVE);
  for (int64_t base = wave * (WAVE * BIN_U); base < n; base += nwaves * (WAVE * BIN_U)) {
    vl_box v[BIN_U];
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t i = base + u * WAVE + lane;
      v[u] = vl_box{0, 0, 0, 0};
      if (i < n) v[u] = i < na ? a[i] : b[i - na];
    }
    int c1[BIN_U], c2[BIN_U], r1[BIN_U], m[BIN_U];
    uint32_t key0[BIN_U], old0[BIN_U];
    uint64_t st0[BIN_U];
    Lead l0[BIN_U];
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t i = base + u * WAVE + lane;
      const bool isA = i < na;
      int x1 = 0, x2 = -1, y1 = 0, y2 = -1;
      if (i < n && !tile_range(v[u], isA ? xra : Xform{0, 0, 0, 0}, g, gd, x1, x2, y1, y2)) { x2 = x1 - 1; y2 = y1 - 1; }
      c1[u] = x1;
      c2[u] = x2;
      r1[u] = y1;
      m[u] = (x2 >= x1 && y2 >= y1) ? (x2 - x1 + 1) * (y2 - y1 + 1) : 0;
      key0[u] = m[u] > 0 ? bin_key(y1, x1, g, isA ? 0u : 1u) : 0u;
      l0[u] = lead_of<GROUPED>(m[u] > 0, key0[u], lane);
    }
    // the first tile of every box: atomics and range starts of all chunks in flight
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      old0[u] = l0[u].lead ? atomicSub(&cnt[key0[u]], (uint32_t)l0[u].len) : 0u;
      st0[u] = m[u] > 0 ? start[key0[u]] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < BIN_U; u++) {
      const int64_t i = base + u * WAVE + lane;
      const bool isA = i < na;
      const vl_box stored = isA ? xform(v[u], xsa) : v[u];
      const uint32_t old = (uint32_t)__shfl((int)old0[u], l0[u].leader, WAVE);
      if (m[u] > 0) {
        const uint64_t pos = st0[u] + old - 1u - (uint32_t)l0[u].rank;
        if (idx) idx[pos] = (uint32_t)i;   // index binning: the box's position
        else entries[pos] = stored;
      }
      // the further tiles of boxes spanning several, in row-major order
      const uint32_t layer = isA ? 0u : 1u;
      int c = c1[u] + 1, r = r1[u];
      if (c > c2[u]) { c = c1[u]; ++r; }
      for (int jt = 1; __ballot(jt < m[u]); jt++) {
        const bool act = jt < m[u];
        const uint32_t key = act ? bin_key(r, c, g, layer) : 0u;
        const Lead ld = lead_of<GROUPED>(act, key, lane);
        uint32_t o = 0;
        if (ld.lead) o = atomicSub(&cnt[key], (uint32_t)ld.len);
        o = (uint32_t)__shfl((int)o, ld.leader, WAVE);
        if (act) {
          const uint64_t pos = start[key] + o - 1u - (uint32_t)ld.rank;
          if (idx) idx[pos] = (uint32_t)i;
          else entries[pos] = stored;
        }
        if (++c > c2[u]) { c = c1[u]; ++r; }
      }
    }
  }
}

// AND pair filter.  A tile of A AND B has an empty result unless some box of
// A, clipped to the tile, overlaps some box of B with positive area.  Tiles
// with such a pair, or with more than a wave of boxes of either layer, are
// appended to `list`; the others get empty result counts here and are never
// loaded by the tile kernels.
//
// Most tiles hold a few boxes of each layer, so a wave tests four tiles at
// once, one per 16-lane row: lane s of a row holds A box s (clipped to the
// tile) and B box s (empty boxes past the counts), and sixteen rotations of
// the B boxes within the row (DPP row_ror) test every A box against every B
// box, straight-line code with no memory traffic.  Tiles with more than 16
// boxes of a layer are then tested one at a time by the whole wave (lane j
// holds B box j, the A boxes broadcast with readlane).
constexpr int AF_G = 16;   // lanes per tile in the grouped pass (one DPP row)

template <int K>
__device__ __forceinline__ int32_t row_ror(int32_t x) {
  if constexpr (K == 0) return x;
  else return __builtin_amdgcn_update_dpp(0, x, 0x120 + K, 0xF, 0xF, false);   // row_ror:K
}

// A box of this lane against the B box K lanes along its row (open overlap).
template <int K>
__device__ __forceinline__ bool hit_rot(int32_t ax1, int32_t ay1, int32_t ax2, int32_t ay2, const vl_box& b) {
  const int32_t bx1 = row_ror<K>(b.x1), by1 = row_ror<K>(b.y1), bx2 = row_ror<K>(b.x2), by2 = row_ror<K>(b.y2);
  return max(ax1, bx1) < min(ax2, bx2) && max(ay1, by1) < min(ay2, by2);
}

template <int... K>
__device__ __forceinline__ bool hit_any(std::integer_sequence<int, K...>, int32_t ax1, int32_t ay1, int32_t ax2,
                                        int32_t ay2, const vl_box& b) {
  return (hit_rot<K>(ax1, ay1, ax2, ay2, b) | ...);
}

__device__ __forceinline__ bool overlap_open(int32_t ax1, int32_t ay1, int32_t ax2, int32_t ay2, const vl_box& b) {
  return max(ax1, b.x1) < min(ax2, b.x2) && max(ay1, b.y1) < min(ay2, b.y2);
}

// Whole-wave test of one tile (nA, nB <= WAVE; arguments uniform).
__device__ __forceinline__ bool and_tile_wave(const vl_box* __restrict__ entries, int64_t off, int nA, int nB,
                                              int32_t x0, int32_t y0, int32_t x1, int32_t y1, int lane) {
  int32_t ax1 = 0, ax2 = 0, ay1 = 0, ay2 = 0;
  if (lane < nA) {
    const vl_box a = entries[off + lane];
    ax1 = max(a.x1, x0); ax2 = min(a.x2, x1); ay1 = max(a.y1, y0); ay2 = min(a.y2, y1);
  }
  vl_box b{0, 0, 0, 0};
  if (lane < nB) b = entries[off + nA + lane];
  for (int i = 0; i < nA; i++) {
    const int32_t qx1 = __builtin_amdgcn_readlane(ax1, i), qx2 = __builtin_amdgcn_readlane(ax2, i);
    const int32_t qy1 = __builtin_amdgcn_readlane(ay1, i), qy2 = __builtin_amdgcn_readlane(ay2, i);
    if (__ballot(lane < nB && overlap_open(qx1, qy1, qx2, qy2, b))) return true;
  }
  return false;
}

__global__ __launch_bounds__(256) void k_and_filter(const TileDesc* __restrict__ desc, int ntiles,
                                                    const vl_box* __restrict__ entries, uint32_t* __restrict__ list,
                                                    uint32_t* __restrict__ list_n, uint32_t* __restrict__ tab_n,
                                                    int ntab) {
  constexpr int TPW = WAVE / AF_G;   // tiles per wave
  const int lane = __lane_id();
  const int sub = lane & (AF_G - 1), grp = lane / AF_G;
  const int base = lane & ~(AF_G - 1);
  const uint64_t gmask = ((1ull << AF_G) - 1ull) << base;
  const int nwaves = gridDim.x * (blockDim.x / WAVE);
  for (int t0 = (blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * TPW; t0 < ntiles; t0 += nwaves * TPW) {
    const int t = t0 + grp;
    const bool tv = t < ntiles;
    int32_t x0 = 0, y0 = 0, x1 = 0, y1 = 0, nA = 0, nB = 0;
    int64_t off = 0;
    if (tv) {
      const TileDesc d = desc[t];
      x0 = d.x0; y0 = d.y0; x1 = d.x1; y1 = d.y1;
      off = d.off;
      nA = d.nA;
      nB = d.nB;
    }
    const bool both = tv && nA > 0 && nB > 0;
    const bool small = nA <= AF_G && nB <= AF_G;
    // grouped pass over the small tiles
    const bool need = both && small;
    int32_t ax1 = 0, ax2 = 0, ay1 = 0, ay2 = 0;
    vl_box b{0, 0, 0, 0};
    if (need && sub < nA) {
      const vl_box a = entries[off + sub];
      ax1 = max(a.x1, x0); ax2 = min(a.x2, x1); ay1 = max(a.y1, y0); ay2 = min(a.y2, y1);
    }
    if (need && sub < nB) b = entries[off + nA + sub];
    // lanes past a tile's counts (and tiles not tested here) hold empty
    // boxes, which overlap nothing
    const bool hit = hit_any(std::make_integer_sequence<int, AF_G>{}, ax1, ay1, ax2, ay2, b);
    bool keep = (__ballot(hit) & gmask) != 0ull;
    // the larger tiles, one at a time with the whole wave
    uint64_t big = __ballot(both && !small && sub == 0);
    while (big) {
      const int l = __ffsll((unsigned long long)big) - 1;
      big &= big - 1;
      const int bA = __builtin_amdgcn_readlane(nA, l), bB = __builtin_amdgcn_readlane(nB, l);
      bool k = true;
      if (bA <= WAVE && bB <= WAVE) {
        const int64_t boff = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(off >> 32), l) << 32) |
                                       (uint32_t)__builtin_amdgcn_readlane((int)off, l));
        k = and_tile_wave(entries, boff, bA, bB, __builtin_amdgcn_readlane(x0, l), __builtin_amdgcn_readlane(y0, l),
                          __builtin_amdgcn_readlane(x1, l), __builtin_amdgcn_readlane(y1, l), lane);
      }
      if (base == l) keep = k;
    }
    if (tv && sub == 0) {
      if (keep) {
        list[atomicAdd(list_n, 1u)] = (uint32_t)t;
      } else {
        tab_n[t] = 0u;
        tab_n[ntab + t] = 0u;
      }
    }
  }
}

// Semi-join test of an AND: the boxes of one input were binned alone into g
// (tile t's boxes at entries[start[2t] .. start[2t+1])); a lane takes a box
// of the other input and tests it against the binned boxes of every tile its
// interior meets.  A positive-area overlap of any pair lies inside a tile
// both boxes meet, so *hit stays 0 exactly when the AND is empty.  With mark,
// mark[i] says whether box i overlaps some binned box (the semi-join
// reduction); without, a lane stops testing after its first overlap.  Inputs
// in spatial order keep neighbouring lanes on the same tiles (cache hits);
// the binned boxes of a tile are loaded PROBE_U at a time (one memory round
// trip for most tiles) and the lane's next box is loaded before the walk.
constexpr int PROBE_U = 8;

__device__ __forceinline__ bool overlaps(const vl_box& q, const vl_box& b) {
  return max(q.x1, b.x1) < min(q.x2, b.x2) && max(q.y1, b.y1) < min(q.y2, b.y2);
}

__global__ __launch_bounds__(256) void k_and_probe(const vl_box* __restrict__ p, int64_t n, Xform xp, Grid g,
                                                   const vl_box* __restrict__ entries,
                                                   const uint64_t* __restrict__ start, uint32_t* __restrict__ hit,
                                                   uint32_t* __restrict__ mark) {
  const GridDiv gd = grid_div(g);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool found = false;   // some box of this lane overlaps
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  vl_box nq = i < n ? p[i] : vl_box{0, 0, 0, 0};
  for (; i < n; i += stride) {
    const vl_box q = xform(nq, xp);
    if (i + stride < n) nq = p[i + stride];
    bool h = false;
    int c1 = 0, c2 = -1, r1 = 0, r2 = -1;
    if ((mark || !found) && tile_range(q, Xform{0, 0, 0, 0}, g, gd, c1, c2, r1, r2)) {
      for (int r = r1; r <= r2 && !h; r++) {
        for (int c = c1; c <= c2 && !h; c++) {
          const uint64_t t = (uint64_t)r * (uint64_t)g.cols + (uint64_t)c;
          const uint64_t s0 = start[2 * t], s1 = start[2 * t + 1];
          for (uint64_t j = s0; j < s1; j += PROBE_U) {
            vl_box b[PROBE_U];
#pragma unroll
            for (int u = 0; u < PROBE_U; u++) b[u] = j + u < s1 ? entries[j + u] : vl_box{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < PROBE_U; u++) h |= overlaps(q, b[u]);
          }
        }
      }
    }
    if (mark) mark[i] = h ? 1u : 0u;
    found |= h;
  }
  if (__ballot(found) != 0ull && __lane_id() == 0) gstore(hit, 1u);
}

// out[pos[i]] = p[i] transformed by xp, for the marked boxes.
__global__ __launch_bounds__(256) void k_compact_marked(const vl_box* __restrict__ p, int64_t n, Xform xp,
                                                        const uint32_t* __restrict__ mark,
                                                        const uint32_t* __restrict__ pos, vl_box* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (mark[i]) out[pos[i]] = xform(p[i], xp);
}

__global__ void k_build_desc(Grid g, const uint64_t* __restrict__ start, int32_t op, TileDesc* desc) {
  const int64_t nt = (int64_t)g.cols * g.rows;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
    int c = (int)(t % g.cols), r = (int)(t / g.cols);
    TileDesc d;
    d.x0 = (int32_t)(g.x0 + (int64_t)c * g.tw);
    d.x1 = (int32_t)min(g.x0 + (int64_t)(c + 1) * g.tw, g.x1);
    d.y0 = (int32_t)(g.y0 + (int64_t)r * g.th);
    d.y1 = (int32_t)min(g.y0 + (int64_t)(r + 1) * g.th, g.y1);
    const uint64_t s0 = start[2 * t], s1 = start[2 * t + 1], s2 = start[2 * t + 2];
    d.off = (int64_t)s0;
    d.nA = (int32_t)(s1 - s0);
    d.nB = (int32_t)(s2 - s1);
    d.flags = (c > 0 ? SEAM_L : 0u) | (c < g.cols - 1 ? SEAM_R : 0u) | (r > 0 ? SEAM_B : 0u) |
              (r < g.rows - 1 ? SEAM_T : 0u);
    d.op = op;
    desc[t] = d;
  }
}

// ---- wave-per-tile kernel --------------------------------------------------------
template <int NMAX, bool TWO, bool AREA, bool ERO>
__global__ __launch_bounds__(64) void k_tile_w(const TileDesc* __restrict__ desc, int ntiles,
                                               const vl_box* __restrict__ entries, TileOut o) {
  constexpr int CPB = ERO ? 4 : 2;    // coordinates per box and axis (erosion adds the shifted edges)
  constexpr int P = CPB * (NMAX + 1); // coordinate slots per axis
  constexpr int K = P / WAVE;         // slots per lane
  constexpr int NCOL = P - 1;
  constexpr int NW = P / WAVE;        // words per column
  __shared__ int32_t ux[P], uy[P];
  __shared__ uint16_t eup[ERO ? P : 1], edn[ERO ? P : 1];   // row maps of the y erosion
  __shared__ uint64_t MA[NCOL * NW];
  __shared__ uint64_t MX[(TWO || ERO) ? NCOL * NW : 1];   // layer B mask / x-eroded mask
  constexpr int NL = (NMAX + WAVE) / WAVE;                // box loads per lane and tile
  __shared__ vl_box sbuf[TILE_BATCH * (NMAX + 1)];        // the batch's boxes
  __shared__ TileDesc sdesc[TILE_BATCH];                  // the batch's descriptors
  uint32_t* const sdw = (uint32_t*)sdesc;                 // ... as 10 dwords each
  __shared__ vl_box sk[TWO ? NMAX + 1 : 1];               // boxes kept after pruning
  __shared__ uint32_t sovf[OVF_BUF];                      // overflow tiles not yet flushed
  static_assert(TILE_BATCH * 10 <= WAVE, "one descriptor dword per lane");

  const int lane = threadIdx.x;
  const int nt = o.list_n ? (int)*o.list_n : ntiles;
  uint64_t cbase = 0, cused = 0, ccap = 0;   // this workgroup's current chunk (uniform)
  uint64_t used = 0, side_total = 0;
  bool short_cap = false;   // a range fell past the stream's capacity (uniform)
  uint32_t novf = 0;        // entries in sovf (uniform)
  long long area_acc = 0;   // area mode: this workgroup's share (uniform)
  // Workgroup g takes batches of TILE_BATCH consecutive tiles (list entries)
  // g, g+G, ...; the descriptors and boxes of a batch are loaded together, so
  // the memory latency is paid once per batch rather than once per tile.
  for (int b0 = blockIdx.x * TILE_BATCH; b0 < nt; b0 += gridDim.x * TILE_BATCH) {
    const int nbt = min(TILE_BATCH, nt - b0);
    uint32_t my_ti = 0;
    if (lane < nbt) my_ti = o.list ? o.list[b0 + lane] : (uint32_t)(b0 + lane);
    {
      const int dj = lane / 10, dw = lane - 10 * (lane / 10);
      const uint32_t tj = (uint32_t)__shfl((int)my_ti, dj < TILE_BATCH ? dj : 0, WAVE);
      if (lane < nbt * 10) sdw[lane] = ((const uint32_t*)desc)[(size_t)tj * 10 + dw];
    }
    __syncthreads();
    {
      int4 v[TILE_BATCH * NL];
      int nj[TILE_BATCH];
#pragma unroll
      for (int j = 0; j < TILE_BATCH; j++) {
        const int64_t off = sdesc[j].off;
        const int c = (j < nbt) ? sdesc[j].nA + sdesc[j].nB : 0;
        nj[j] = c <= NMAX ? c : 0;
#pragma unroll
        for (int q = 0; q < NL; q++) {
          const int i = lane + q * WAVE;
          v[j * NL + q] = make_int4(0, 0, 0, 0);
          if (i < nj[j]) v[j * NL + q] = ((const int4*)entries)[off + i];
        }
      }
#pragma unroll
      for (int j = 0; j < TILE_BATCH; j++) {
#pragma unroll
        for (int q = 0; q < NL; q++) {
          const int i = lane + q * WAVE;
          if (i < nj[j]) ((int4*)sbuf)[j * (NMAX + 1) + i] = v[j * NL + q];
        }
      }
    }
    __syncthreads();
  for (int jb = 0; jb < nbt; jb++) {
    const uint32_t ti = read_lane(my_ti, jb);
    const TileDesc d = sdesc[jb];
    vl_box* sb = sbuf + jb * (NMAX + 1);
    int nA = d.nA;
    int n = d.nA + d.nB;
    const int op = d.op;
    const bool twol = TWO && (op == TOP_AND || op == TOP_NOT || op == TOP_XOR || op == TOP_OVERLAP);
    if (n > NMAX || n == 0) {
      if (lane == 0) {
        if (n > NMAX) sovf[novf] = ti;
        if (!AREA) {
          o.tab_n[ti] = 0;
          o.tab_n[o.ntab + ti] = 0;
        }
      }
      if (n > NMAX && ++novf == OVF_BUF) {
        __syncthreads();
        flush_overflow(sovf, novf, o.ovf, o.ovf_n, lane, WAVE);
        novf = 0;
        __syncthreads();
      }
      continue;
    }
    // 0. the boxes are in LDS.  AND prunes first: a box of either layer that
    //    meets no box of the other inside the tile cannot change A AND B, and
    //    an AND tile without any such pair has an empty result.  (The same
    //    test would let NOT drop B boxes, but measured on the OP1 ring the
    //    test costs more than it saves there.)
    const vl_box* bx = sb;
    if (TWO && op == TOP_AND) {
      constexpr bool TWOREG = NMAX >= WAVE;   // boxes lane and lane+64
      const int nB = n - nA;
      const uint64_t lt = lanemask_lt();
      // lane l holds A boxes l (and l+64) and B boxes l (and l+64); the A
      // boxes are clipped to the tile for the test and broadcast by readlane
      vl_box a1{0, 0, 0, 0}, a2{0, 0, 0, 0}, b1{0, 0, 0, 0}, b2{0, 0, 0, 0};
      if (lane < nA) a1 = sb[lane];
      if (lane < nB) b1 = sb[nA + lane];
      if (TWOREG) {
        if (lane + WAVE < nA) a2 = sb[lane + WAVE];
        if (lane + WAVE < nB) b2 = sb[nA + lane + WAVE];
      }
      const int32_t c1x1 = max(a1.x1, d.x0), c1x2 = min(a1.x2, d.x1), c1y1 = max(a1.y1, d.y0), c1y2 = min(a1.y2, d.y1);
      const int32_t c2x1 = max(a2.x1, d.x0), c2x2 = min(a2.x2, d.x1), c2y1 = max(a2.y1, d.y0), c2y2 = min(a2.y2, d.y1);
      uint64_t hitA1 = 0ull, hitA2 = 0ull;
      bool keepB1 = false, keepB2 = false;
      for (int i = 0; i < nA; i++) {
        const bool lo = i < WAVE;
        const int src = i & (WAVE - 1);
        const int32_t ax1 = __builtin_amdgcn_readlane(lo ? c1x1 : c2x1, src);
        const int32_t ax2 = __builtin_amdgcn_readlane(lo ? c1x2 : c2x2, src);
        const int32_t ay1 = __builtin_amdgcn_readlane(lo ? c1y1 : c2y1, src);
        const int32_t ay2 = __builtin_amdgcn_readlane(lo ? c1y2 : c2y2, src);
        bool h = lane < nB && max(ax1, b1.x1) < min(ax2, b1.x2) && max(ay1, b1.y1) < min(ay2, b1.y2);
        keepB1 |= h;
        if (TWOREG) {
          const bool h2 = lane + WAVE < nB && max(ax1, b2.x1) < min(ax2, b2.x2) && max(ay1, b2.y1) < min(ay2, b2.y2);
          keepB2 |= h2;
          h |= h2;
        }
        if (__ballot(h)) {
          if (lo) hitA1 |= 1ull << src;
          else hitA2 |= 1ull << src;
        }
      }
      // compact the kept boxes (A first, then B), keeping their order
      const bool kA1 = lane < nA && ((hitA1 >> lane) & 1ull);
      const bool kA2 = TWOREG && lane + WAVE < nA && ((hitA2 >> lane) & 1ull);
      const uint64_t mA1 = __ballot(kA1), mA2 = __ballot(kA2);
      const uint64_t mB1 = __ballot(keepB1), mB2 = __ballot(keepB2);
      const int cA1 = __popcll(mA1), cA = cA1 + __popcll(mA2);
      const int cB1 = __popcll(mB1), cB = cB1 + __popcll(mB2);
      if (kA1) sk[__popcll(mA1 & lt)] = a1;
      if (kA2) sk[cA1 + __popcll(mA2 & lt)] = a2;
      if (keepB1) sk[cA + __popcll(mB1 & lt)] = b1;
      if (keepB2) sk[cA + cB1 + __popcll(mB2 & lt)] = b2;
      __syncthreads();
      nA = cA;
      n = cA + cB;
      bx = sk;
      if (cA == 0 || cB == 0) {
        if (lane == 0) {
          o.tab_n[ti] = 0;
          o.tab_n[o.ntab + ti] = 0;
        }
        __syncthreads();
        continue;
      }
    }
    // 1. coordinates: slot e = lane*K + k holds component e % CPB of box
    //    e / CPB (x1, x2, and when eroding x1+R, x2-R clamped to the tile),
    //    clipped to the window: the tile, grown by R when eroding.  Slot
    //    group n holds the window and tile borders; the rest are padding.
    const int32_t R = ERO ? o.erode : 0;
    const int32_t wx0 = clamp64((int64_t)d.x0 - R, INT32_MIN, INT32_MAX);
    const int32_t wx1 = clamp64((int64_t)d.x1 + R, INT32_MIN, INT32_MAX);
    const int32_t wy0 = clamp64((int64_t)d.y0 - R, INT32_MIN, INT32_MAX);
    const int32_t wy1 = clamp64((int64_t)d.y1 + R, INT32_MIN, INT32_MAX);
    int32_t vx[K], vy[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = lane * K + k;
      const int i = e / CPB, comp = e % CPB;
      int32_t cx = INT32_MAX, cy = INT32_MAX;
      if (i < n) {
        const vl_box b = bx[i];
        const int32_t x1 = max(b.x1, wx0), x2 = min(b.x2, wx1);
        const int32_t y1 = max(b.y1, wy0), y2 = min(b.y2, wy1);
        if (comp == 0) { cx = x1; cy = y1; }
        else if (comp == 1) { cx = x2; cy = y2; }
        else if (comp == 2) { cx = clamp64((int64_t)x1 + R, d.x0, d.x1); cy = clamp64((int64_t)y1 + R, d.y0, d.y1); }
        else { cx = clamp64((int64_t)x2 - R, d.x0, d.x1); cy = clamp64((int64_t)y2 - R, d.y0, d.y1); }
      } else if (i == n) {
        if (comp == 0) { cx = wx0; cy = wy0; }
        else if (comp == 1) { cx = wx1; cy = wy1; }
        else if (comp == 2) { cx = d.x0; cy = d.y0; }
        else { cx = d.x1; cy = d.y1; }
      }
      vx[k] = cx;
      vy[k] = cy;
    }
    const int cnt = CPB * n + CPB;
    // bitonic sort of P elements, element e = lane*K + k
#pragma unroll
    for (int size = 2; size <= P; size <<= 1) {
#pragma unroll
      for (int j = size >> 1; j > 0; j >>= 1) {
        if (j >= K) {
          const int lj = j / K;
#pragma unroll
          for (int k = 0; k < K; k++) {
            const int e = lane * K + k;
            const int px = xor_lanes(vx[k], lj);
            const int py = xor_lanes(vy[k], lj);
            const bool take_min = ((e & j) == 0) == ((e & size) == 0);
            vx[k] = take_min ? min(vx[k], px) : max(vx[k], px);
            vy[k] = take_min ? min(vy[k], py) : max(vy[k], py);
          }
        } else {
#pragma unroll
          for (int k = 0; k < K; k++) {
            if ((k & j) == 0) {
              const int k2 = k | j;
              const bool asc = ((lane * K + k) & size) == 0;
              int32_t a = vx[k], b = vx[k2];
              if ((a > b) == asc) { vx[k] = b; vx[k2] = a; }
              a = vy[k]; b = vy[k2];
              if ((a > b) == asc) { vy[k] = b; vy[k2] = a; }
            }
          }
        }
      }
    }
    // drop duplicates
    int nxu, nyu;
    {
      const int32_t px = __shfl_up(vx[K - 1], 1, WAVE), py = __shfl_up(vy[K - 1], 1, WAVE);
      bool fx[K], fy[K];
      uint32_t cx = 0, cy = 0;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int e = lane * K + k;
        const int32_t qx = k ? vx[k - 1] : px, qy = k ? vy[k - 1] : py;
        fx[k] = e < cnt && (e == 0 || vx[k] != qx);
        fy[k] = e < cnt && (e == 0 || vy[k] != qy);
        cx += fx[k];
        cy += fy[k];
      }
      uint32_t tx, ty;
      uint32_t ox = (uint32_t)wave_excl_u32(cx, &tx), oy = (uint32_t)wave_excl_u32(cy, &ty);
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (fx[k]) ux[ox++] = vx[k];
        if (fy[k]) uy[oy++] = vy[k];
      }
      nxu = (int)tx;
      nyu = (int)ty;
    }
    // one row above the last is kept clear for the run end marks
    // (a 64-slot grid has at most 63 rows: one word, W folds to 1)
    const int ncol = nxu - 1, nrows = nyu - 1, W = NW == 1 ? 1 : (nrows + 64) >> 6;
    const int words = ncol * W;
    for (int k = lane; k < words; k += WAVE) {
      MA[k] = 0ull;
      if (TWO) MX[k] = 0ull;
    }
    __syncthreads();
    // 2. compress + rasterise (box-parallel: lane takes boxes lane, lane+64, ...)
    for (int i = lane; i < n; i += WAVE) {
      const vl_box b = bx[i];
      const int cx1 = lower_bound_fixed<P>(ux, nxu, max(b.x1, wx0)), cx2 = lower_bound_fixed<P>(ux, nxu, min(b.x2, wx1));
      const int cy1 = lower_bound_fixed<P>(uy, nyu, max(b.y1, wy0)), cy2 = lower_bound_fixed<P>(uy, nyu, min(b.y2, wy1));
      if (cx1 >= cx2 || cy1 >= cy2) continue;
      uint64_t* M = (twol && i >= nA) ? MX : MA;
      const int w1 = cy1 >> 6, w2 = (cy2 - 1) >> 6;
      for (int c = cx1; c < cx2; c++)
        for (int w = w1; w <= w2; w++) {
          const int lo = (w == w1) ? (cy1 & 63) : 0;
          const int hi = (w == w2) ? ((cy2 - 1) & 63) : 63;
          const uint64_t bits = bits_range(lo, hi);
          if (TWO && op == TOP_OVERLAP) {
            // cells this box finds covered already are covered twice
            const uint64_t old = atomicOr((unsigned long long*)&MA[c * W + w], (unsigned long long)bits);
            if (old & bits) atomicOr((unsigned long long*)&MX[c * W + w], (unsigned long long)(old & bits));
          } else {
            atomicOr((unsigned long long*)&M[c * W + w], (unsigned long long)bits);
          }
        }
    }
    __syncthreads();
    if (twol) {
      for (int k = lane; k < words; k += WAVE) MA[k] = combine(op, MA[k], MX[k]);
      __syncthreads();
    }
    // the tile's own columns [colT0, colT1) and rows [rowT0, rowT1)
    int colT0 = 0, colT1 = ncol, rowT0 = 0, rowT1 = nrows;
    if (ERO) {
      colT0 = count_below_wave<K>(ux, nxu, d.x0, lane);
      colT1 = count_below_wave<K>(ux, nxu, d.x1, lane);
      rowT0 = count_below_wave<K>(uy, nyu, d.y0, lane);
      rowT1 = count_below_wave<K>(uy, nyu, d.y1, lane);
      erosion_row_maps(uy, nyu, d.y0, d.y1, R, eup, edn, lane, WAVE);
      erode_x<NW>(MA, MX, ux, nxu, ncol, W, colT0, colT1, R, lane, WAVE);
      __syncthreads();
      erode_y<NW>(MX, MA, eup, edn, ncol, W, lane, WAVE);
      __syncthreads();
    }
    if (AREA) {
      long long acc = 0;
      for (int j = lane; j < ncol; j += WAVE) {
        const uint64_t* col = MA + j * W;
        long long h = 0;
        int a = -1;
        for (int w = 0; w < W; w++) {
          const uint64_t v = col[w];
          int pos = 0;
          while (pos < 64) {
            if (a < 0) {
              const uint64_t rem = v >> pos;
              if (rem == 0) break;
              pos += __ffsll((unsigned long long)rem) - 1;
              a = w * 64 + pos;
            } else {
              const uint64_t rem = (~v) >> pos;
              if (rem == 0) break;
              pos += __ffsll((unsigned long long)rem) - 1;
              h += (long long)uy[w * 64 + pos] - uy[a];
              a = -1;
            }
          }
        }
        if (a >= 0) h += (long long)uy[W * 64] - uy[a];
        acc += h * ((long long)ux[j + 1] - ux[j]);
      }
      area_acc += wave_sum(acc);
      __syncthreads();
      continue;
    }
    // 3. start / end masks at every slab boundary k = 0..ncol, kept in
    //    registers: lane owns boundaries k = 64*kb + lane (Sr/Er[kb][w], bit =
    //    row).  Every strip ends exactly once, so the popcount of the end
    //    masks is the strip count.
    uint64_t Sr[NW][NW], Er[NW][NW];
    uint32_t myT = 0;
#pragma unroll
    for (int kb = 0; kb < NW; kb++) {
      const int k = kb * WAVE + lane;
#pragma unroll
      for (int w = 0; w < NW; w++) { Sr[kb][w] = 0ull; Er[kb][w] = 0ull; }
      if (k <= ncol) {
        const uint64_t* cur = (k < ncol) ? MA + k * W : nullptr;
        const uint64_t* prv = (k > 0) ? MA + (k - 1) * W : nullptr;
        uint64_t D[NW];
        change_mask<NW>(cur, prv, W, D);
        ends_meeting<NW>(cur, W, D, Sr[kb]);
        ends_meeting<NW>(prv, W, D, Er[kb]);
#pragma unroll
        for (int w = 0; w < NW; w++) myT += (uint32_t)__popcll(Er[kb][w]);
      }
    }
    const uint32_t T = wave_sum(myT);
    if (T == 0) {
      if (lane == 0) {
        o.tab_n[ti] = 0;
        o.tab_n[o.ntab + ti] = 0;
      }
      continue;
    }
    if (cused + T > ccap) {   // a new chunk
      const uint64_t want = T > o.chunk ? (uint64_t)T : o.chunk;
      unsigned long long nb = 0;
      if (lane == 0) nb = atomicAdd(o.next, (unsigned long long)want);
      cbase = read_lane((uint64_t)nb, 0);
      cused = 0;
      ccap = want;
    }
    const uint64_t base = cbase + cused;
    const bool ok = base + T <= o.stream_cap;
    short_cap = short_cap || !ok;
    cused += T;
    used += T;
    if (lane == 0) o.tab_off[ti] = base;
    // 4. strips.  Lane l owns boundary j = 64*kb + l and walks its S marks:
    //    the rows b just above the runs [a, b) of strips starting there.  The
    //    E marks are transposed in 64x64 bit blocks so that lane l holds, for
    //    row 64*wb + l, the boundaries at which a strip marked in that row
    //    ends; a strip starting at (j, b) ends at the first E mark after j in
    //    row b, read from lane b & 63, and its run starts one above the
    //    highest clear row of column j below b.  Every lane stays in the loop
    //    until the whole wave is done, so the cross-lane reads see a full
    //    wave.  Output order: row block, boundary block, pass, boundary.
    const uint32_t flags = d.flags;
    const int KW = (ncol + 1 + 63) >> 6;   // boundary words per row
    uint32_t nside = 0;   // seam strips of this tile (uniform)
    uint64_t pos = base;
#pragma unroll
    for (int wb = 0; wb < NW; wb++) {
      if (wb >= W) break;
      uint64_t re[NW];
#pragma unroll
      for (int kb = 0; kb < NW; kb++) re[kb] = kb < KW ? transpose64(Er[kb][wb], lane) : 0ull;
#pragma unroll
      for (int kb = 0; kb < NW; kb++) {
        if (kb >= KW) break;
        uint64_t m = Sr[kb][wb];
        const int j = kb * WAVE + lane;
        for (;;) {
          const uint64_t act = __ballot(m != 0ull);
          if (!act) break;
          const bool has = m != 0ull;
          const int t = has ? __ffsll((unsigned long long)m) - 1 : 0;
          m &= m - 1;
          uint64_t rr[NW];
#pragma unroll
          for (int kq = 0; kq < NW; kq++) rr[kq] = lane_read64(re[kq], t);
          uint32_t tg = 0u;
          if (has) {
            const int b = wb * WAVE + t;
            const int k = next_bit_above<NW>(rr, j, ncol);
            const int a = run_start_w<NW>(MA + j * W, W, b);
            tg = ((a == rowT0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                 ((b == rowT1 && (flags & SEAM_T)) ? SEAM_T : 0u) |
                 ((j == colT0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                 ((k == colT1 && (flags & SEAM_R)) ? SEAM_R : 0u);
            const uint64_t p = pos + lanes_below(act);
            if (ok) {
              vl_box st;
              st.x1 = ux[j]; st.y1 = uy[a]; st.x2 = ux[k]; st.y2 = uy[b];
              o.strips[p] = st;
              o.tag[p] = (uint8_t)tg;
            }
          }
          nside += (uint32_t)__popcll(__ballot(tg != 0u));
          pos += (uint64_t)__popcll(act);
        }
      }
    }

    side_total += nside;
    if (lane == 0) {
      o.tab_n[ti] = T - nside;
      o.tab_n[o.ntab + ti] = nside;
    }
    __syncthreads();
  }
    __syncthreads();
  }
  if (novf) {
    __syncthreads();
    flush_overflow(sovf, novf, o.ovf, o.ovf_n, lane, WAVE);
  }
  if (lane == 0 && AREA && area_acc) atomicAdd((unsigned long long*)o.area, (unsigned long long)area_acc);
  if (lane == 0 && !AREA && used) {
    atomicAdd(&o.cnt[0], (unsigned long long)used);
    if (side_total) atomicAdd(&o.cnt[1], (unsigned long long)side_total);
    if (short_cap) atomicOr(o.err, ERR_OVERFLOW_CAP);
  }
}

// ---- workgroup-per-tile kernel (large tiles) ------------------------------------

template <int NMAX, int NT, bool TWO, bool AREA, bool ERO>
__global__ __launch_bounds__(NT) void k_tile_b(const TileDesc* __restrict__ desc, int ntiles,
                                               const vl_box* __restrict__ entries, TileOut o) {
  constexpr int CPB = ERO ? 4 : 2;   // coordinates per box and axis
  constexpr int P = CPB * (NMAX + 1);
  constexpr int NCOL = P - 1;
  constexpr int NW = (NCOL + 63) / 64;
  __shared__ int32_t sx[P], sy[P];
  __shared__ int32_t ux[P], uy[P];
  __shared__ uint16_t eup[ERO ? P : 1], edn[ERO ? P : 1];   // row maps of the y erosion
  __shared__ int16_t cb[NMAX][4];
  __shared__ uint64_t MA[NCOL * NW];
  __shared__ uint64_t MB[(TWO || ERO) ? NCOL * NW : 1];
  __shared__ uint64_t colc[NCOL];
  __shared__ int wsI[NT / WAVE];
  __shared__ uint64_t wsU[NT / WAVE];
  __shared__ long long wsA[NT / WAVE];

  __shared__ uint32_t sovf[OVF_BUF];
  const int tid = threadIdx.x;
  const int nt = o.list_n ? (int)*o.list_n : ntiles;
  __shared__ unsigned long long s_chunk;
  uint64_t cbase = 0, cused = 0, ccap = 0;   // this workgroup's current chunk (uniform)
  uint64_t used = 0, side_total = 0;   // uniform across the workgroup
  bool short_cap = false;
  uint32_t novf = 0;
  long long area_acc = 0;
  for (int tk = blockIdx.x; tk < nt; tk += gridDim.x) {
    const uint32_t ti = o.list ? gload(o.list + tk) : (uint32_t)tk;
    const TileDesc d = desc[ti];
    const int nA = d.nA;
    const int n = d.nA + d.nB;
    const int op = d.op;
    const bool twol = TWO && (op == TOP_AND || op == TOP_NOT || op == TOP_XOR || op == TOP_OVERLAP);
    if (n > NMAX || n == 0) {
      if (tid == 0) {
        if (n > NMAX) sovf[novf] = ti;
        if (!AREA) {
          gstore(o.tab_n + ti, 0u);
          gstore(o.tab_n + o.ntab + ti, 0u);
        }
      }
      __syncthreads();
      if (n > NMAX && ++novf == OVF_BUF) {
        flush_overflow(sovf, novf, o.ovf, o.ovf_n, tid, NT);
        novf = 0;
      }
      continue;
    }
    uint64_t F = 0, S = 0;
    int ncol = 0, nrows = 0, W = 0;
    long long my_area = 0;
    // clip window: the tile, grown by R when eroding
    const int32_t R = ERO ? o.erode : 0;
    const int32_t wx0 = clamp64((int64_t)d.x0 - R, INT32_MIN, INT32_MAX);
    const int32_t wx1 = clamp64((int64_t)d.x1 + R, INT32_MIN, INT32_MAX);
    const int32_t wy0 = clamp64((int64_t)d.y0 - R, INT32_MIN, INT32_MAX);
    const int32_t wy1 = clamp64((int64_t)d.y1 + R, INT32_MIN, INT32_MAX);
    int colT0 = 0, colT1 = 0, rowT0 = 0, rowT1 = 0;
    {
      const int cnt = CPB * n + CPB;
      int Pn = 2;
      while (Pn < cnt) Pn <<= 1;
      for (int i = tid; i < Pn; i += NT) {
        int32_t vx = INT32_MAX, vy = INT32_MAX;
        const int q = i / CPB, part = i % CPB;
        if (q < n) {
          const vl_box b = entries[d.off + q];
          const int32_t x1 = max(b.x1, wx0), x2 = min(b.x2, wx1), y1 = max(b.y1, wy0), y2 = min(b.y2, wy1);
          switch (part) {
            case 0: vx = x1; vy = y1; break;
            case 1: vx = x2; vy = y2; break;
            case 2: vx = clamp64((int64_t)x1 + R, d.x0, d.x1); vy = clamp64((int64_t)y1 + R, d.y0, d.y1); break;
            default: vx = clamp64((int64_t)x2 - R, d.x0, d.x1); vy = clamp64((int64_t)y2 - R, d.y0, d.y1); break;
          }
        } else if (q == n) {
          switch (part) {
            case 0: vx = wx0; vy = wy0; break;
            case 1: vx = wx1; vy = wy1; break;
            case 2: vx = d.x0; vy = d.y0; break;
            default: vx = d.x1; vy = d.y1; break;
          }
        }
        sx[i] = vx;
        sy[i] = vy;
      }
      __syncthreads();
      for (int k = 2; k <= Pn; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < Pn; i += NT) {
            const int ixj = i ^ j;
            if (ixj > i) {
              const bool asc = (i & k) == 0;
              int32_t a = sx[i], b = sx[ixj];
              if ((a > b) == asc) { sx[i] = b; sx[ixj] = a; }
              a = sy[i]; b = sy[ixj];
              if ((a > b) == asc) { sy[i] = b; sy[ixj] = a; }
            }
          }
          __syncthreads();
        }
      }
      int nxu = 0, nyu = 0;
      for (int c0 = 0; c0 < Pn; c0 += NT) {
        const int i = c0 + tid;
        bool fx = false, fy = false;
        int32_t vx = 0, vy = 0;
        if (i < cnt) {
          vx = sx[i]; vy = sy[i];
          fx = (i == 0) || (vx != sx[i - 1]);
          fy = (i == 0) || (vy != sy[i - 1]);
        }
        int tx, ty;
        const int px = block_excl_flag_scan<NT>(fx, wsI, &tx);
        if (fx) ux[nxu + px] = vx;
        const int py = block_excl_flag_scan<NT>(fy, wsI, &ty);
        if (fy) uy[nyu + py] = vy;
        nxu += tx;
        nyu += ty;
      }
      ncol = nxu - 1;
      nrows = nyu - 1;
      W = (nrows + 63) >> 6;
      __syncthreads();
      for (int i = tid; i < n; i += NT) {
        const vl_box b = entries[d.off + i];
        cb[i][0] = (int16_t)lower_bound_i32(ux, nxu, max(b.x1, wx0));
        cb[i][1] = (int16_t)lower_bound_i32(ux, nxu, min(b.x2, wx1));
        cb[i][2] = (int16_t)lower_bound_i32(uy, nyu, max(b.y1, wy0));
        cb[i][3] = (int16_t)lower_bound_i32(uy, nyu, min(b.y2, wy1));
      }
      const int words = ncol * W;
      for (int k = tid; k < words; k += NT) {
        MA[k] = 0ull;
        if (TWO) MB[k] = 0ull;
      }
      __syncthreads();
      for (int i = tid; i < n; i += NT) {
        const int cx1 = cb[i][0], cx2 = cb[i][1], cy1 = cb[i][2], cy2 = cb[i][3];
        if (cx1 >= cx2 || cy1 >= cy2) continue;
        uint64_t* M = (twol && i >= nA) ? MB : MA;
        const int w1 = cy1 >> 6, w2 = (cy2 - 1) >> 6;
        for (int c = cx1; c < cx2; c++) {
          uint64_t* col = M + c * W;
          for (int w = w1; w <= w2; w++) {
            const int lo = (w == w1) ? (cy1 & 63) : 0;
            const int hi = (w == w2) ? ((cy2 - 1) & 63) : 63;
            const uint64_t bits = bits_range(lo, hi);
            if (TWO && op == TOP_OVERLAP) {
              const uint64_t old = atomicOr((unsigned long long*)&MA[c * W + w], (unsigned long long)bits);
              if (old & bits) atomicOr((unsigned long long*)&MB[c * W + w], (unsigned long long)(old & bits));
            } else {
              atomicOr((unsigned long long*)&col[w], (unsigned long long)bits);
            }
          }
        }
      }
      __syncthreads();
      if (twol) {
        for (int k = tid; k < words; k += NT) MA[k] = combine(op, MA[k], MB[k]);
        __syncthreads();
      }
      colT0 = 0; colT1 = ncol; rowT0 = 0; rowT1 = nrows;
      if (ERO) {
        colT0 = count_below(ux, nxu, d.x0);
        colT1 = count_below(ux, nxu, d.x1);
        rowT0 = count_below(uy, nyu, d.y0);
        rowT1 = count_below(uy, nyu, d.y1);
        erosion_row_maps(uy, nyu, d.y0, d.y1, R, eup, edn, tid, NT);
        erode_x<NW>(MA, MB, ux, nxu, ncol, W, colT0, colT1, R, tid, NT);
        __syncthreads();
        erode_y<NW>(MB, MA, eup, edn, ncol, W, tid, NT);
        __syncthreads();
      }
      const uint32_t flags = d.flags;
      for (int j = tid; j < ncol; j += NT) {
        const uint64_t* cur = MA + j * W;
        const uint64_t* prv = j ? (MA + (j - 1) * W) : nullptr;
        if (AREA) {
          long long h = 0;
          for_each_run(cur, W, [&](int a, int b) { h += (long long)uy[b] - uy[a]; });
          my_area += h * ((long long)ux[j + 1] - ux[j]);
        } else {
          uint32_t nf = 0, ns = 0;
          for_each_run(cur, W, [&](int a, int b) {
            const int lo = a > 0 ? a - 1 : 0;
            const int hi = b < nrows ? b : nrows - 1;
            if (!any_diff(cur, prv, lo, hi)) return;
            int k = j + 1;
            while (k < ncol && !any_diff(MA + k * W, MA + (k - 1) * W, lo, hi)) k++;
            const bool side = (a == rowT0 && (flags & SEAM_B)) || (b == rowT1 && (flags & SEAM_T)) ||
                              (j == colT0 && (flags & SEAM_L)) || (k == colT1 && (flags & SEAM_R));
            if (side) ns++; else nf++;
          });
          colc[j] = ((uint64_t)nf << 32) | ns;
        }
      }
      __syncthreads();
      if (!AREA) {
        uint64_t run = 0;
        for (int c0 = 0; c0 < ncol; c0 += NT) {
          const int j = c0 + tid;
          const uint64_t v = j < ncol ? colc[j] : 0ull;
          uint64_t tot;
          const uint64_t ex = block_excl_scan<uint64_t, NT>(v, wsU, &tot);
          if (j < ncol) colc[j] = run + ex;
          run += tot;
        }
        F = run >> 32;
        S = run & 0xffffffffull;
      }
    }

    if (AREA) {
      long long a = wave_sum(my_area);
      if (__lane_id() == 0) wsA[tid / WAVE] = a;
      __syncthreads();
      long long s = 0;
      for (int i = 0; i < NT / WAVE; i++) s += wsA[i];
      area_acc += s;
      __syncthreads();
      continue;
    }

    // one range per tile, from the workgroup's chunk: final strips first, then seam strips
    const uint64_t T = F + S;
    if (T > 0 && cused + T > ccap) {   // a new chunk
      const uint64_t want = T > o.chunk ? T : o.chunk;
      if (tid == 0) s_chunk = atomicAdd(o.next, (unsigned long long)want);
      __syncthreads();
      cbase = s_chunk;
      cused = 0;
      ccap = want;
      __syncthreads();
    }
    const uint64_t b0 = cbase + cused;
    const bool ok = b0 + T <= o.stream_cap;
    short_cap = short_cap || !ok;
    cused += T;
    used += T;
    side_total += S;
    if (tid == 0) {
      gstore(o.tab_off + ti, b0);
      gstore(o.tab_n + ti, (uint32_t)F);
      gstore(o.tab_n + o.ntab + ti, (uint32_t)S);
    }

    if ((F + S) > 0) {
      const uint64_t offF = b0, offS = b0 + F;
      const bool okF = ok, okS = ok;
      const uint32_t flags = d.flags;
      for (int j = tid; j < ncol; j += NT) {
        const uint64_t* cur = MA + j * W;
        const uint64_t* prv = j ? (MA + (j - 1) * W) : nullptr;
        uint64_t fi = offF + (colc[j] >> 32);
        uint64_t si = offS + (colc[j] & 0xffffffffull);
        for_each_run(cur, W, [&](int a, int b) {
          const int lo = a > 0 ? a - 1 : 0;
          const int hi = b < nrows ? b : nrows - 1;
          if (!any_diff(cur, prv, lo, hi)) return;
          int k = j + 1;
          while (k < ncol && !any_diff(MA + k * W, MA + (k - 1) * W, lo, hi)) k++;
          const uint32_t tag = ((a == rowT0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                               ((b == rowT1 && (flags & SEAM_T)) ? SEAM_T : 0u) |
                               ((j == colT0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                               ((k == colT1 && (flags & SEAM_R)) ? SEAM_R : 0u);
          if (tag) {
            if (okS) {
              gstore_box(o.strips + si, ux[j], uy[a], ux[k], uy[b]);
              gstore(o.tag + si, (uint8_t)tag);
            }
            si++;
          } else {
            if (okF) {
              gstore_box(o.strips + fi, ux[j], uy[a], ux[k], uy[b]);
              gstore(o.tag + fi, (uint8_t)0);
            }
            fi++;
          }
        });
      }
    }
    __syncthreads();
  }
  if (novf) {
    __syncthreads();
    flush_overflow(sovf, novf, o.ovf, o.ovf_n, tid, NT);
  }
  if (tid == 0 && AREA && area_acc) atomicAdd((unsigned long long*)o.area, (unsigned long long)area_acc);
  if (tid == 0 && !AREA && used) {
    atomicAdd(&o.cnt[0], (unsigned long long)used);
    if (side_total) atomicAdd(&o.cnt[1], (unsigned long long)side_total);
    if (short_cap) atomicOr(o.err, ERR_OVERFLOW_CAP);
  }
}

// ---- reorder into tile order -------------------------------------------------------
constexpr int REORDER_TILES = 256;   // tiles per 256-thread workgroup (64 per wave)

// A wave takes 64 consecutive tiles and copies their strips, concatenated in
// tile order, in chunks of 64 (one strip per lane), so small tiles do not
// leave most lanes idle.  A strip's tile is found by binary search over the
// tiles' starts in the concatenation; its rank among its tile's final strips
// comes from a ballot over the chunk plus the tile's finals in earlier chunks
// (only the tile crossing a chunk border carries any).  Final strips go to
// the output and seam strips to the side stream, each in tile order.
__global__ __launch_bounds__(256) void k_reorder(int ntiles, const uint64_t* __restrict__ tab_off,
                                                 const uint32_t* __restrict__ tab_n, int ntab,
                                                 const uint64_t* __restrict__ dstF, const uint64_t* __restrict__ dstS,
                                                 const vl_box* __restrict__ scr, const uint8_t* __restrict__ scrTag,
                                                 vl_box* __restrict__ out, uint64_t out_base,
                                                 vl_box* __restrict__ side, uint8_t* __restrict__ side_tag,
                                                 uint64_t side_base) {
  __shared__ uint32_t s_seg[4][WAVE];
  __shared__ uint64_t s_src[4][WAVE], s_dF[4][WAVE], s_dS[4][WAVE];
  const int lane = __lane_id();
  const int wave = threadIdx.x / WAVE;
  const uint64_t lt = lanemask_lt();
  const int t = (blockIdx.x * 4 + wave) * WAVE + lane;
  uint32_t T = 0;
  uint64_t src = 0, dF = 0, dS = 0;
  if (t < ntiles) {
    T = tab_n[t] + tab_n[ntab + t];
    if (T) {
      src = tab_off[t];
      dF = out_base + dstF[t];
      dS = side_base + dstS[t];
    }
  }
  uint32_t total;
  s_seg[wave][lane] = (uint32_t)wave_excl_u32(T, &total);
  s_src[wave][lane] = src;
  s_dF[wave][lane] = dF;
  s_dS[wave][lane] = dS;
  __syncthreads();
  uint32_t carry_tile = ~0u, carry_f = 0;   // the tile crossing into this chunk and its earlier finals
  for (uint32_t c0 = 0; c0 < total; c0 += WAVE) {
    const uint32_t f = c0 + lane;
    const bool valid = f < total;
    int lo = 0, hi = WAVE;   // tile of strip f: the last tile starting at or before f
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_seg[wave][mid] <= f) lo = mid + 1; else hi = mid;
    }
    const int own = lo - 1;
    const uint32_t sg = s_seg[wave][own];
    uint8_t tg = 0;
    vl_box b;
    if (valid) {
      const uint64_t sp = s_src[wave][own] + (f - sg);
      tg = scrTag[sp];
      b = scr[sp];
    }
    const bool isF = valid && tg == 0;
    const uint64_t mF = __ballot(isF);
    const int first = sg > c0 ? (int)(sg - c0) : 0;   // first lane of this tile in the chunk
    if (valid) {
      const uint32_t rf = (uint32_t)__popcll(mF & lt & (~0ull << first)) + ((uint32_t)own == carry_tile ? carry_f : 0u);
      if (isF) {
        out[s_dF[wave][own] + rf] = b;
      } else {
        const uint64_t p = s_dS[wave][own] + ((f - sg) - rf);
        side[p] = b;
        side_tag[p] = tg;
      }
    }
    const int own63 = __builtin_amdgcn_readlane(own, WAVE - 1);
    const int first63 = __builtin_amdgcn_readlane(first, WAVE - 1);
    const uint32_t fin63 = (uint32_t)__popcll(mF & (~0ull << first63));
    carry_f = ((uint32_t)own63 == carry_tile ? carry_f : 0u) + fin63;
    carry_tile = (uint32_t)own63;
  }
}

__global__ void k_gather_desc(const TileDesc* __restrict__ desc, const uint32_t* __restrict__ idx, int n,
                              TileDesc* out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = desc[idx[i]];
}

// ---- overflow splitting ----------------------------------------------------------
// One workgroup per parent tile; its sub-tiles form a kx x ky grid (row-major;
// all but the last column / row have the first sub-tile's width / height).
// A box belongs to every sub-tile its extent grown by halo meets.
__device__ __forceinline__ void sub_rect(const vl_box& b, const TileDesc& p, const TileDesc* subs, int first,
                                         int nsub, int32_t halo, int& c1, int& c2, int& r1, int& r2, int& kx) {
  const int64_t sw = (int64_t)subs[first].x1 - subs[first].x0;
  const int64_t sh = (int64_t)subs[first].y1 - subs[first].y0;
  kx = (int)(((int64_t)p.x1 - p.x0 + sw - 1) / sw);
  const int ky = nsub / kx;
  const int64_t x1 = max((int64_t)b.x1 - halo, (int64_t)p.x0), x2 = min((int64_t)b.x2 + halo, (int64_t)p.x1);
  const int64_t y1 = max((int64_t)b.y1 - halo, (int64_t)p.y0), y2 = min((int64_t)b.y2 + halo, (int64_t)p.y1);
  c1 = min((int)((x1 - p.x0) / sw), kx - 1);
  c2 = min((int)((x2 - 1 - p.x0) / sw), kx - 1);
  r1 = min((int)((y1 - p.y0) / sh), ky - 1);
  r2 = min((int)((y2 - 1 - p.y0) / sh), ky - 1);
}

__global__ void k_split_count(const TileDesc* __restrict__ parents, const int32_t* __restrict__ sub_first,
                              const TileDesc* __restrict__ subs, uint32_t* cntA, uint32_t* cntB,
                              const vl_box* __restrict__ entries, int32_t halo) {
  const TileDesc p = parents[blockIdx.x];
  const int first = sub_first[blockIdx.x], nsub = sub_first[blockIdx.x + 1] - first;
  if (nsub == 0) return;   // a parent left whole (it goes to the dense kernel)
  const int n = p.nA + p.nB;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const vl_box b = entries[p.off + i];
    int c1, c2, r1, r2, kx;
    sub_rect(b, p, subs, first, nsub, halo, c1, c2, r1, r2, kx);
    for (int r = r1; r <= r2; r++)
      for (int c = c1; c <= c2; c++) atomicAdd(i < p.nA ? &cntA[first + r * kx + c] : &cntB[first + r * kx + c], 1u);
  }
}

__global__ void k_split_scatter(const TileDesc* __restrict__ parents, const int32_t* __restrict__ sub_first,
                                const TileDesc* __restrict__ subs, const uint64_t* __restrict__ off,
                                const uint32_t* __restrict__ cntA, uint32_t* fillA, uint32_t* fillB,
                                const vl_box* __restrict__ entries, vl_box* __restrict__ sub_entries,
                                int32_t halo) {
  const TileDesc p = parents[blockIdx.x];
  const int first = sub_first[blockIdx.x], nsub = sub_first[blockIdx.x + 1] - first;
  if (nsub == 0) return;   // a parent left whole (it goes to the dense kernel)
  const int n = p.nA + p.nB;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const vl_box b = entries[p.off + i];
    int c1, c2, r1, r2, kx;
    sub_rect(b, p, subs, first, nsub, halo, c1, c2, r1, r2, kx);
    for (int r = r1; r <= r2; r++)
      for (int c = c1; c <= c2; c++) {
        const int q = first + r * kx + c;
        uint64_t pos = (i < p.nA) ? off[q] + atomicAdd(&fillA[q], 1u)
                                  : off[q] + cntA[q] + atomicAdd(&fillB[q], 1u);
        sub_entries[pos] = b;
      }
  }
}

__global__ void k_fill_subdesc(TileDesc* subs, int ns, const uint64_t* __restrict__ off,
                               const uint32_t* __restrict__ cntA, const uint32_t* __restrict__ cntB) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
    subs[i].off = (int64_t)off[i];
    subs[i].nA = (int32_t)cntA[i];
    subs[i].nB = (int32_t)cntB[i];
  }
}

unsigned grid_for(int64_t n, int per = 256) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 16384) g = 16384;
  return (unsigned)g;
}


template <int NMAX, bool TWO, bool AREA, bool ERO>
void launch_w(const TileDesc* desc, int ntiles, const vl_box* entries, TileOut o, int wg, hipStream_t st) {
  const int g = wg < 1 ? 1 : wg;
  hipLaunchKernelGGL((k_tile_w<NMAX, TWO, AREA, ERO>), dim3(g), dim3(WAVE), 0, st, desc, ntiles, entries, o);
}

constexpr int TILE_BLOCK_NMAX_ERODE = 127;

template <bool TWO, bool AREA, bool ERO>
void launch_b(const TileDesc* desc, int ntiles, const vl_box* entries, TileOut o, int wg, hipStream_t st) {
  const int g = wg < 1 ? 1 : wg;
  constexpr int NM = ERO ? TILE_BLOCK_NMAX_ERODE : TILE_BLOCK_NMAX;
  hipLaunchKernelGGL((k_tile_b<NM, 256, TWO, AREA, ERO>), dim3(g), dim3(256), 0, st, desc, ntiles, entries, o);
}

}  // namespace

void launch_bin_span(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g, uint32_t* cnt,
                     hipStream_t st) {
  if (na + nb > 0)
    hipLaunchKernelGGL(k_bin_span, dim3(grid_for(na + nb)), dim3(256), 0, st, a, na, xra, b, nb, g, cnt);
}
void launch_bin_emit(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g,
                     const uint64_t* pos, uint32_t* keys, uint32_t* vals, hipStream_t st) {
  if (na + nb > 0)
    hipLaunchKernelGGL(k_bin_emit, dim3(grid_for(na + nb)), dim3(256), 0, st, a, na, xra, b, nb, g, pos, keys,
                       vals);
}
void launch_bin_bounds(const uint32_t* keys, uint64_t total, int64_t nkeys, uint64_t* start, hipStream_t st) {
  hipLaunchKernelGGL(k_bin_bounds, dim3(grid_for(nkeys + 1)), dim3(256), 0, st, keys, total, nkeys, start);
}
void launch_bin_gather(const vl_box* a, int64_t na, Xform xsa, const vl_box* b, const uint32_t* vals,
                       uint64_t total, vl_box* entries, hipStream_t st) {
  if (total > 0)
    hipLaunchKernelGGL(k_bin_gather, dim3(grid_for((int64_t)total)), dim3(256), 0, st, a, na, xsa, b, vals, total,
                       entries);
}
void launch_bin_count(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g, uint32_t* cnt,
                      bool grouped, hipStream_t st) {
  const int64_t n = na + nb;
  if (n <= 0) return;
  const dim3 grid(grid_for((n + BIN_U - 1) / BIN_U));
  if (grouped) hipLaunchKernelGGL(k_bin_count<true>, grid, dim3(256), 0, st, a, na, xra, b, nb, g, cnt);
  else hipLaunchKernelGGL(k_bin_count<false>, grid, dim3(256), 0, st, a, na, xra, b, nb, g, cnt);
}
void launch_bin_scatter(const vl_box* a, int64_t na, Xform xra, Xform xsa, const vl_box* b, int64_t nb, Grid g,
                        const uint64_t* start, uint32_t* cnt, vl_box* entries, bool grouped, hipStream_t st) {
  const int64_t n = na + nb;
  if (n <= 0) return;
  const dim3 grid(grid_for((n + BIN_U - 1) / BIN_U));
  if (grouped)
    hipLaunchKernelGGL(k_bin_scatter<true>, grid, dim3(256), 0, st, a, na, xra, xsa, b, nb, g, start, cnt, entries,
                       (uint32_t*)nullptr);
  else
    hipLaunchKernelGGL(k_bin_scatter<false>, grid, dim3(256), 0, st, a, na, xra, xsa, b, nb, g, start, cnt, entries,
                       (uint32_t*)nullptr);
}
void launch_bin_scatter_idx(const vl_box* a, int64_t na, Xform xra, Grid g, const uint64_t* start, uint32_t* cnt,
                            uint32_t* idx, bool grouped, hipStream_t st) {
  if (na <= 0) return;
  const dim3 grid(grid_for((na + BIN_U - 1) / BIN_U));
  if (grouped)
    hipLaunchKernelGGL(k_bin_scatter<true>, grid, dim3(256), 0, st, a, na, xra, Xform{0, 0, 0, 0},
                       (const vl_box*)nullptr, (int64_t)0, g, start, cnt, (vl_box*)nullptr, idx);
  else
    hipLaunchKernelGGL(k_bin_scatter<false>, grid, dim3(256), 0, st, a, na, xra, Xform{0, 0, 0, 0},
                       (const vl_box*)nullptr, (int64_t)0, g, start, cnt, (vl_box*)nullptr, idx);
}
void launch_and_filter(const TileDesc* desc, int ntiles, const vl_box* entries, uint32_t* list, uint32_t* list_n,
                       uint32_t* tab_n, int ntab, hipStream_t st) {
  if (ntiles > 0)
    hipLaunchKernelGGL(k_and_filter, dim3(grid_for(ntiles, 16)), dim3(256), 0, st, desc, ntiles, entries, list,
                       list_n, tab_n, ntab);
}
void launch_and_probe(const vl_box* p, int64_t n, Xform xp, Grid g, const vl_box* entries, const uint64_t* start,
                      uint32_t* hit, uint32_t* mark, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_and_probe, dim3(grid_for(n)), dim3(256), 0, st, p, n, xp, g, entries, start, hit, mark);
}
void launch_compact_marked(const vl_box* p, int64_t n, Xform xp, const uint32_t* mark, const uint32_t* pos,
                           vl_box* out, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_compact_marked, dim3(grid_for(n)), dim3(256), 0, st, p, n, xp, mark, pos, out);
}
void launch_build_desc(Grid g, const uint64_t* start, int32_t op, TileDesc* desc, hipStream_t st) {
  int64_t nt = (int64_t)g.cols * g.rows;
  hipLaunchKernelGGL(k_build_desc, dim3(grid_for(nt)), dim3(256), 0, st, g, start, op, desc);
}

#define VL_W_DISPATCH(NM)                                                                  \
  if (nmax == NM) {                                                                        \
    if (area_mode) {                                                                       \
      if (two_layers) launch_w<NM, true, true, false>(desc, ntiles, entries, o, wg, st);   \
      else launch_w<NM, false, true, false>(desc, ntiles, entries, o, wg, st);             \
    } else {                                                                               \
      if (two_layers) launch_w<NM, true, false, false>(desc, ntiles, entries, o, wg, st);  \
      else launch_w<NM, false, false, false>(desc, ntiles, entries, o, wg, st);            \
    }                                                                                      \
    return;                                                                                \
  }

void launch_tile_wave(int nmax, bool two_layers, bool area_mode, bool erode, const TileDesc* desc, int ntiles,
                      const vl_box* entries, TileOut o, int wg, hipStream_t st) {
  if (erode) {
    if (nmax == 15) launch_w<15, false, false, true>(desc, ntiles, entries, o, wg, st);
    else launch_w<31, false, false, true>(desc, ntiles, entries, o, wg, st);
    return;
  }
  VL_W_DISPATCH(31)
  VL_W_DISPATCH(63)
  VL_W_DISPATCH(127)
}

#define VL_W_MAXWG(NM)                                                                                  \
  if (nmax == NM) {                                                                                     \
    if (area_mode)                                                                                      \
      return two_layers ? occupancy_wg(k_tile_w<NM, true, true, false>, WAVE)                           \
                        : occupancy_wg(k_tile_w<NM, false, true, false>, WAVE);                         \
    return two_layers ? occupancy_wg(k_tile_w<NM, true, false, false>, WAVE)                            \
                      : occupancy_wg(k_tile_w<NM, false, false, false>, WAVE);                          \
  }

int tile_wave_max_wg(int nmax, bool two_layers, bool area_mode, bool erode) {
  static int cache[4][2][2][2] = {};
  const int ni = nmax == 15 ? 0 : nmax == 31 ? 1 : nmax == 63 ? 2 : 3;
  int& c = cache[ni][two_layers][area_mode][erode];
  if (c) return c;
  auto f = [&]() -> int {
    if (erode)
      return nmax == 15 ? occupancy_wg(k_tile_w<15, false, false, true>, WAVE)
                        : occupancy_wg(k_tile_w<31, false, false, true>, WAVE);
    VL_W_MAXWG(31)
    VL_W_MAXWG(63)
    VL_W_MAXWG(127)
    return 1024;
  };
  c = f();
  return c;
}

void launch_tile_block(bool two_layers, bool area_mode, bool erode, const TileDesc* desc, int ntiles,
                       const vl_box* entries, TileOut o, int wg, hipStream_t st) {
  if (erode) {
    launch_b<false, false, true>(desc, ntiles, entries, o, wg, st);
  } else if (area_mode) {
    if (two_layers) launch_b<true, true, false>(desc, ntiles, entries, o, wg, st);
    else launch_b<false, true, false>(desc, ntiles, entries, o, wg, st);
  } else {
    if (two_layers) launch_b<true, false, false>(desc, ntiles, entries, o, wg, st);
    else launch_b<false, false, false>(desc, ntiles, entries, o, wg, st);
  }
}

int tile_block_max_wg(bool two_layers, bool area_mode, bool erode) {
  static int cache[2][2][2] = {};
  int& c = cache[two_layers][area_mode][erode];
  if (c) return c;
  if (erode)
    c = occupancy_wg(k_tile_b<TILE_BLOCK_NMAX_ERODE, 256, false, false, true>, 256);
  else if (area_mode)
    c = two_layers ? occupancy_wg(k_tile_b<TILE_BLOCK_NMAX, 256, true, true, false>, 256)
                   : occupancy_wg(k_tile_b<TILE_BLOCK_NMAX, 256, false, true, false>, 256);
  else
    c = two_layers ? occupancy_wg(k_tile_b<TILE_BLOCK_NMAX, 256, true, false, false>, 256)
                   : occupancy_wg(k_tile_b<TILE_BLOCK_NMAX, 256, false, false, false>, 256);
  return c;
}

void launch_tile_reorder(int ntiles, const uint64_t* tab_off, const uint32_t* tab_n, int ntab, const uint64_t* dstF,
                         const uint64_t* dstS, const vl_box* scr, const uint8_t* scrTag, vl_box* out,
                         uint64_t out_base, vl_box* side, uint8_t* side_tag, uint64_t side_base, hipStream_t st) {
  if (ntiles <= 0) return;
  const int g = (ntiles + REORDER_TILES - 1) / REORDER_TILES;
  hipLaunchKernelGGL(k_reorder, dim3(g), dim3(256), 0, st, ntiles, tab_off, tab_n, ntab, dstF, dstS, scr, scrTag,
                     out, out_base, side, side_tag, side_base);
}

void launch_gather_desc(const TileDesc* desc, const uint32_t* idx, int n, TileDesc* out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_gather_desc, dim3(grid_for(n)), dim3(256), 0, st, desc, idx, n, out);
}

void launch_split_count(const TileDesc* parents, int np, const int32_t* sub_first, const TileDesc* subs,
                        uint32_t* sub_cntA, uint32_t* sub_cntB, const vl_box* entries, int32_t halo,
                        hipStream_t st) {
  if (np > 0)
    hipLaunchKernelGGL(k_split_count, dim3(np), dim3(256), 0, st, parents, sub_first, subs, sub_cntA, sub_cntB,
                       entries, halo);
}

void launch_split_scatter(const TileDesc* parents, int np, const int32_t* sub_first, TileDesc* subs, int ns,
                          const uint64_t* off, const uint32_t* cntA, const uint32_t* cntB, uint32_t* fillA,
                          uint32_t* fillB, const vl_box* entries, vl_box* sub_entries, int32_t halo,
                          hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_split_scatter, dim3(np), dim3(256), 0, st, parents, sub_first, subs, off, cntA, fillA,
                     fillB, entries, sub_entries, halo);
  hipLaunchKernelGGL(k_fill_subdesc, dim3(grid_for(ns)), dim3(256), 0, st, subs, ns, off, cntA, cntB);
}

}  // namespace vl
