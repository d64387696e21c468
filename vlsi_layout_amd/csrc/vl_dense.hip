// vl_dense.hip — the per-tile Boolean for tiles no capacity-bounded kernel
// can hold: more boxes over one x-position than the workgroup kernel's 255
// (a bus of several hundred parallel wires, hundreds of identical boxes),
// which splitting the tile into x-sub-columns cannot reduce.
//
// One 256-thread workgroup per tile, every array in global memory, sized by
// the host from a first pass:
//   k_dense_coords  clips the tile's boxes to the tile, sorts their x and y
//                   coordinates (bitonic network over a power-of-two array)
//                   and drops duplicates: nx x-breakpoints, ny y-breakpoints;
//   k_dense_tile    rasterises every layer into per-slab coverage bitmasks
//                   (W = (ny - 1 + 64) / 64 words per slab, one clear row on
//                   top), combines the layers word by word, then extracts
//                   the tile's maximal strips column by column: a run [a, b)
//                   of slab j starts a strip when slab j - 1 does not hold
//                   exactly that run, and the strip ends at the first slab
//                   that does not (heal_boxes' slice matching,
//                   src/polygon_fracture.f90:939-960).  Strips are counted
//                   per slab, scanned, and written in slab order, so the
//                   output order is fixed; area mode sums run lengths times
//                   slab widths instead.
// Strips touching an internal tile border carry the same SEAM_* tags as the
// other tile kernels and take the same seam-level / x-stitch path.
#include "vl_tile_dev.h"

namespace vl {

namespace {

constexpr int DNT = 256;   // threads per dense-tile workgroup

__device__ __forceinline__ int64_t dense_pow2(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// In-place ascending bitonic sort of two arrays of P (a power of two)
// elements by the whole workgroup.
__device__ void block_bitonic2(int32_t* a, int32_t* b, int64_t P) {
  for (int64_t size = 2; size <= P; size <<= 1) {
    for (int64_t j = size >> 1; j > 0; j >>= 1) {
      for (int64_t e = threadIdx.x; e < P; e += DNT) {
        const int64_t q = e ^ j;
        if (q > e) {
          const bool asc = (e & size) == 0;
          int32_t x = a[e], y = a[q];
          if ((x > y) == asc) { a[e] = y; a[q] = x; }
          x = b[e]; y = b[q];
          if ((x > y) == asc) { b[e] = y; b[q] = x; }
        }
      }
      __syncthreads();
    }
  }
}

// Distinct values of the sorted s[0..cnt) into u; returns their number.
__device__ int block_unique(const int32_t* s, int64_t cnt, int32_t* u, int* ws) {
  int carry = 0;
  for (int64_t c0 = 0; c0 < cnt; c0 += DNT) {
    const int64_t e = c0 + threadIdx.x;
    const bool f = e < cnt && (e == 0 || s[e] != s[e - 1]);
    int tot = 0;
    const int pre = block_excl_flag_scan<DNT>(f, ws, &tot);
    if (f) u[carry + pre] = s[e];
    carry += tot;
  }
  __syncthreads();
  return carry;
}

__global__ __launch_bounds__(DNT) void k_dense_coords(const TileDesc* __restrict__ desc,
                                                      const uint32_t* __restrict__ list,
                                                      const vl_box* __restrict__ entries,
                                                      const int64_t* __restrict__ coff, int32_t* __restrict__ cs,
                                                      int32_t* __restrict__ nxy) {
  __shared__ int ws[DNT / WAVE];
  const int i = blockIdx.x;
  const TileDesc d = desc[list[i]];
  const int64_t n = (int64_t)d.nA + d.nB;
  const int64_t cnt = 2 * n + 2;
  const int64_t P = dense_pow2(cnt);
  int32_t* xs = cs + coff[i];
  int32_t* ys = xs + P;
  int32_t* ux = ys + P;
  int32_t* uy = ux + P;
  for (int64_t e = threadIdx.x; e < P; e += DNT) {
    int32_t x = INT32_MAX, y = INT32_MAX;
    if (e < 2 * n) {
      const vl_box b = entries[d.off + (e >> 1)];
      if (e & 1) { x = min(max(b.x2, d.x0), d.x1); y = min(max(b.y2, d.y0), d.y1); }
      else { x = min(max(b.x1, d.x0), d.x1); y = min(max(b.y1, d.y0), d.y1); }
    } else if (e == 2 * n) {
      x = d.x0; y = d.y0;
    } else if (e == 2 * n + 1) {
      x = d.x1; y = d.y1;
    }
    xs[e] = x;
    ys[e] = y;
  }
  __syncthreads();
  block_bitonic2(xs, ys, P);
  const int nx = block_unique(xs, cnt, ux, ws);
  const int ny = block_unique(ys, cnt, uy, ws);
  if (threadIdx.x == 0) {
    nxy[2 * i] = nx;
    nxy[2 * i + 1] = ny;
  }
}

// Does column col hold exactly the run [a, b) (rows a..b-1 set, a-1 and b clear)?
__device__ __forceinline__ bool has_run(const uint64_t* col, int a, int b) {
  if (a > 0 && ((col[(a - 1) >> 6] >> ((a - 1) & 63)) & 1ull)) return false;
  if ((col[b >> 6] >> (b & 63)) & 1ull) return false;
  for (int w = a >> 6; w <= (b - 1) >> 6; w++) {
    const int lo = max(a - 64 * w, 0), hi = min(b - 64 * w, 64) - 1;
    const uint64_t m = bits_range(lo, hi);
    if ((col[w] & m) != m) return false;
  }
  return true;
}

// First row >= r whose bit equals want (rows < limit; returns limit if none).
__device__ __forceinline__ int next_row(const uint64_t* col, int r, int limit, bool want) {
  while (r < limit) {
    const int w = r >> 6;
    uint64_t v = want ? col[w] : ~col[w];
    v &= ~0ull << (r & 63);
    if (v) return min(64 * w + __ffsll((unsigned long long)v) - 1, limit);
    r = 64 * (w + 1);
  }
  return limit;
}

struct DenseArgs {
  const TileDesc* desc;
  const uint32_t* list;
  const vl_box* entries;
  const int64_t* coff;
  const int32_t* cs;
  const int32_t* nxy;
  uint64_t* grid;
  const uint64_t* goff;    // per tile: first word of its grids
  uint32_t* colcnt;
  const uint64_t* cofs;    // per tile: first column counter
  vl_box* strips;
  uint8_t* tag;
  const uint64_t* soff;    // per tile: first strip slot
  uint64_t* tab_off;
  uint32_t* tab_n;         // [i] final count, [nlist + i] seam count
  int nlist;
  int two;                 // two-layer operators rasterise layer B separately
  unsigned long long* cnt; // [0] strips, [1] seam strips
  long long* area;         // area mode when non-null
};

__global__ __launch_bounds__(DNT) void k_dense_tile(DenseArgs g) {
  __shared__ int ws[DNT / WAVE];
  __shared__ uint32_t wsu[DNT / WAVE];
  __shared__ long long wsl[DNT / WAVE];
  const int i = blockIdx.x;
  const int tid = threadIdx.x;
  const TileDesc d = g.desc[g.list[i]];
  const int64_t n = (int64_t)d.nA + d.nB;
  const int64_t P = dense_pow2(2 * n + 2);
  const int32_t* ux = g.cs + g.coff[i] + 2 * P;
  const int32_t* uy = ux + P;
  const int nxu = g.nxy[2 * i], nyu = g.nxy[2 * i + 1];
  const int ncol = nxu - 1, nrows = nyu - 1;
  const int W = (nrows + 64) >> 6;
  const bool twol = g.two && (d.op == TOP_AND || d.op == TOP_NOT || d.op == TOP_XOR || d.op == TOP_OVERLAP);
  const int64_t gw = (int64_t)ncol * W;
  uint64_t* MA = g.grid + g.goff[i];
  uint64_t* MB = MA + gw;
  // 1. clear, rasterise (a thread per box), combine
  for (int64_t k = tid; k < (twol ? 2 : 1) * gw; k += DNT) MA[k] = 0ull;
  __syncthreads();
  for (int64_t bi = tid; bi < n; bi += DNT) {
    const vl_box b = g.entries[d.off + bi];
    const int32_t x1 = max(b.x1, d.x0), x2 = min(b.x2, d.x1), y1 = max(b.y1, d.y0), y2 = min(b.y2, d.y1);
    if (x1 >= x2 || y1 >= y2) continue;
    const int cx1 = lower_bound_i32(ux, nxu, x1), cx2 = lower_bound_i32(ux, nxu, x2);
    const int cy1 = lower_bound_i32(uy, nyu, y1), cy2 = lower_bound_i32(uy, nyu, y2);
    uint64_t* M = (twol && bi >= d.nA) ? MB : MA;
    const int w1 = cy1 >> 6, w2 = (cy2 - 1) >> 6;
    for (int c = cx1; c < cx2; c++)
      for (int w = w1; w <= w2; w++) {
        const int lo = (w == w1) ? (cy1 & 63) : 0;
        const int hi = (w == w2) ? ((cy2 - 1) & 63) : 63;
        const uint64_t bits = bits_range(lo, hi);
        if (d.op == TOP_OVERLAP && twol) {
          // cells this box finds covered already are covered twice
          const uint64_t old = atomicOr((unsigned long long*)&MA[(int64_t)c * W + w], (unsigned long long)bits);
          if (old & bits) atomicOr((unsigned long long*)&MB[(int64_t)c * W + w], (unsigned long long)(old & bits));
        } else {
          atomicOr((unsigned long long*)&M[(int64_t)c * W + w], (unsigned long long)bits);
        }
      }
  }
  // the atomics land in L2: the device-scope fence also drops any stale
  // line of the masks from this CU's L1 before they are read back
  __threadfence();
  __syncthreads();
  if (twol) {
    for (int64_t k = tid; k < gw; k += DNT) MA[k] = combine(d.op, MA[k], MB[k]);
    __syncthreads();
  }
  // 2a. area mode: run lengths times slab widths
  if (g.area) {
    long long acc = 0;
    for (int j = tid; j < ncol; j += DNT) {
      const uint64_t* col = MA + (int64_t)j * W;
      long long h = 0;
      for (int r = next_row(col, 0, nrows, true); r < nrows;) {
        const int e = next_row(col, r, nrows, false);
        h += (long long)uy[e] - uy[r];
        r = next_row(col, e, nrows, true);
      }
      acc += h * ((long long)ux[j + 1] - ux[j]);
    }
    long long tot = 0;
    const long long pre = block_excl_scan<long long, DNT>(acc, wsl, &tot);
    (void)pre;
    if (tid == 0 && tot) atomicAdd((unsigned long long*)g.area, (unsigned long long)tot);
    return;
  }
  // 2b. strips, counted per slab, then written at the scanned slab offsets
  uint32_t* cc = g.colcnt + g.cofs[i];
  const uint32_t flags = d.flags;
  auto walk = [&](int j, auto&& emit) {
    const uint64_t* col = MA + (int64_t)j * W;
    const uint64_t* prv = j > 0 ? col - W : nullptr;
    for (int a = next_row(col, 0, nrows, true); a < nrows;) {
      const int b = next_row(col, a, nrows, false);
      if (!prv || !has_run(prv, a, b)) {
        int k = j + 1;
        while (k < ncol && has_run(MA + (int64_t)k * W, a, b)) k++;
        const uint32_t tg = ((a == 0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                            ((b == nrows && (flags & SEAM_T)) ? SEAM_T : 0u) |
                            ((j == 0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                            ((k == ncol && (flags & SEAM_R)) ? SEAM_R : 0u);
        emit(a, b, k, tg);
      }
      a = next_row(col, b, nrows, true);
    }
  };
  uint32_t my_side = 0;
  for (int j = tid; j < ncol; j += DNT) {
    uint32_t c = 0;
    walk(j, [&](int, int, int, uint32_t tg) {
      c++;
      my_side += tg != 0u;
    });
    cc[j] = c;
  }
  __syncthreads();
  uint32_t carry = 0;
  for (int c0 = 0; c0 < ncol; c0 += DNT) {
    const int j = c0 + tid;
    const uint32_t v = j < ncol ? cc[j] : 0u;
    uint32_t tot = 0;
    const uint32_t pre = block_excl_scan<uint32_t, DNT>(v, wsu, &tot);
    if (j < ncol) cc[j] = carry + pre;
    carry += tot;
  }
  uint32_t side = 0;
  {
    const uint32_t pre = block_excl_scan<uint32_t, DNT>(my_side, wsu, &side);
    (void)pre;
  }
  __syncthreads();
  const uint64_t base = g.soff[i];
  for (int j = tid; j < ncol; j += DNT) {
    uint64_t p = base + cc[j];
    walk(j, [&](int a, int b, int k, uint32_t tg) {
      vl_box s;
      s.x1 = ux[j]; s.y1 = uy[a]; s.x2 = ux[k]; s.y2 = uy[b];
      g.strips[p] = s;
      g.tag[p] = (uint8_t)tg;
      p++;
    });
  }
  if (tid == 0) {
    g.tab_off[i] = base;
    g.tab_n[i] = carry - side;
    g.tab_n[g.nlist + i] = side;
    atomicAdd(&g.cnt[0], (unsigned long long)carry);
    atomicAdd(&g.cnt[1], (unsigned long long)side);
  }
}

}  // namespace

int64_t dense_coord_slots(int64_t nboxes) {
  int64_t p = 1;
  while (p < 2 * nboxes + 2) p <<= 1;
  return 4 * p;
}

void launch_dense_coords(const TileDesc* desc, const uint32_t* list, int nlist, const vl_box* entries,
                         const int64_t* coff, int32_t* cs, int32_t* nxy, hipStream_t st) {
  if (nlist > 0) hipLaunchKernelGGL(k_dense_coords, dim3(nlist), dim3(DNT), 0, st, desc, list, entries, coff, cs, nxy);
}

void launch_dense_tile(const TileDesc* desc, const uint32_t* list, int nlist, const vl_box* entries,
                       const int64_t* coff, const int32_t* cs, const int32_t* nxy, uint64_t* grid,
                       const uint64_t* goff, uint32_t* colcnt, const uint64_t* cofs, vl_box* strips, uint8_t* tag,
                       const uint64_t* soff, uint64_t* tab_off, uint32_t* tab_n, bool two,
                       unsigned long long* cnt, long long* area, hipStream_t st) {
  if (nlist <= 0) return;
  DenseArgs a;
  a.desc = desc; a.list = list; a.entries = entries; a.coff = coff; a.cs = cs; a.nxy = nxy;
  a.grid = grid; a.goff = goff; a.colcnt = colcnt; a.cofs = cofs; a.strips = strips; a.tag = tag; a.soff = soff;
  a.tab_off = tab_off; a.tab_n = tab_n; a.nlist = nlist; a.two = two ? 1 : 0; a.cnt = cnt; a.area = area;
  hipLaunchKernelGGL(k_dense_tile, dim3(nlist), dim3(DNT), 0, st, a);
}

}  // namespace vl
