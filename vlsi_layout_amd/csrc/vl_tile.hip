// vl_tile.hip — tile binning and the per-tile Boolean kernel for gfx950.
//
// Binning: every box whose interior meets a tile of a uniform grid is copied
// (after the op's per-box transform) into that tile's entry range: a counting
// pass (atomics per tile), an exclusive scan, and a scatter pass.
//
// Tile kernel (one workgroup per tile, persistent over an atomic ticket):
//  1. load the tile's boxes, clip them to the tile, gather their x and y
//     coordinates plus the tile borders, bitonic-sort both lists in LDS and
//     remove duplicates: nx x-breakpoints give ncol = nx-1 slabs, ny
//     y-breakpoints give nrows = ny-1 elementary y-intervals;
//  2. map every box to compressed indices (binary search) and rasterise the
//     layers into per-slab coverage bitmasks in LDS, one 64-bit word per 64
//     elementary y-intervals (wave64-sized words, ds_or_b64 atomics);
//  3. combine the layers word by word with the operator (AND, A&~B, OR, XOR);
//  4. per slab, list the maximal runs of set bits; a run starts a strip when
//     the change mask (this slab XOR previous slab, dilated by one row) meets
//     it, and the strip ends at the first later slab whose change mask meets
//     it.  This is exactly heal_boxes' maximal-strip form
//     (src/polygon_fracture.f90:881-960) restricted to the tile;
//  5. strips that touch an internal tile border are seam pieces and go to the
//     side stream with SEAM_* tags; the others are final;
//  6. output offsets come from a decoupled look-back over tile order, so the
//     output order is deterministic (tile order, then slab, then y).
// In area mode step 4 sums the covered run lengths times slab widths (int64)
// instead (calculate_union_area_fast, src/geometry.f90:1840-1927).
#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

__device__ __forceinline__ bool tile_range(const vl_box& b, const Grid& g, int& c1, int& c2, int& r1, int& r2) {
  int64_t bx1 = max((int64_t)b.x1, g.x0), bx2 = min((int64_t)b.x2, g.x1);
  int64_t by1 = max((int64_t)b.y1, g.y0), by2 = min((int64_t)b.y2, g.y1);
  if (bx1 >= bx2 || by1 >= by2) return false;
  c1 = (int)((bx1 - g.x0) / g.tw);
  c2 = (int)((bx2 - 1 - g.x0) / g.tw);
  r1 = (int)((by1 - g.y0) / g.th);
  r2 = (int)((by2 - 1 - g.y0) / g.th);
  return true;
}

__device__ __forceinline__ vl_box xform(vl_box b, const Xform& x) {
  b.x1 -= x.w;
  b.y1 -= x.s;
  b.x2 += x.e;
  b.y2 += x.n;
  return b;
}

__global__ void k_bin_count(const vl_box* __restrict__ b, int64_t n, Xform xf, Grid g, uint32_t* cnt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    vl_box v = xform(b[i], xf);
    int c1, c2, r1, r2;
    if (!tile_range(v, g, c1, c2, r1, r2)) continue;
    for (int r = r1; r <= r2; r++)
      for (int c = c1; c <= c2; c++) atomicAdd(&cnt[(int64_t)r * g.cols + c], 1u);
  }
}

__global__ void k_bin_scatter(const vl_box* __restrict__ b, int64_t n, Xform xf, Grid g,
                              const uint64_t* __restrict__ off, const uint32_t* __restrict__ base_add,
                              uint32_t* fill, vl_box* __restrict__ entries) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    vl_box v = xform(b[i], xf);
    int c1, c2, r1, r2;
    if (!tile_range(v, g, c1, c2, r1, r2)) continue;
    for (int r = r1; r <= r2; r++)
      for (int c = c1; c <= c2; c++) {
        int64_t t = (int64_t)r * g.cols + c;
        uint64_t pos = off[t] + (base_add ? base_add[t] : 0u) + atomicAdd(&fill[t], 1u);
        entries[pos] = v;
      }
  }
}

__global__ void k_build_desc(Grid g, const uint64_t* __restrict__ off, const uint32_t* __restrict__ cntA,
                             const uint32_t* __restrict__ cntB, int32_t op, TileDesc* desc) {
  const int64_t nt = (int64_t)g.cols * g.rows;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += (int64_t)gridDim.x * blockDim.x) {
    int c = (int)(t % g.cols), r = (int)(t / g.cols);
    TileDesc d;
    d.x0 = (int32_t)(g.x0 + (int64_t)c * g.tw);
    d.x1 = (int32_t)min(g.x0 + (int64_t)(c + 1) * g.tw, g.x1);
    d.y0 = (int32_t)(g.y0 + (int64_t)r * g.th);
    d.y1 = (int32_t)min(g.y0 + (int64_t)(r + 1) * g.th, g.y1);
    d.off = (int64_t)off[t];
    d.nA = (int32_t)cntA[t];
    d.nB = cntB ? (int32_t)cntB[t] : 0;
    d.flags = (c > 0 ? SEAM_L : 0u) | (c < g.cols - 1 ? SEAM_R : 0u) | (r > 0 ? SEAM_B : 0u) |
              (r < g.rows - 1 ? SEAM_T : 0u);
    d.op = op;
    desc[t] = d;
  }
}

// ---- tile kernel -------------------------------------------------------------

__device__ __forceinline__ int lower_bound_i32(const int32_t* a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Calls f(a, b) for every maximal run [a, b) of set bits in a column of W words.
template <class F>
__device__ __forceinline__ void for_each_run(const uint64_t* col, int W, F&& f) {
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
        f(a, w * 64 + pos);
        a = -1;
      }
    }
  }
  if (a >= 0) f(a, W * 64);
}

// True when columns c and p differ in any bit of rows lo..hi (p may be null = empty).
__device__ __forceinline__ bool any_diff(const uint64_t* c, const uint64_t* p, int lo, int hi) {
  const int w0 = lo >> 6, w1 = hi >> 6;
  for (int w = w0; w <= w1; w++) {
    const int l = (w == w0) ? (lo & 63) : 0;
    const int h = (w == w1) ? (hi & 63) : 63;
    const uint64_t pv = p ? p[w] : 0ull;
    if ((c[w] ^ pv) & bits_range(l, h)) return true;
  }
  return false;
}

__device__ __forceinline__ uint64_t combine(int op, uint64_t a, uint64_t b) {
  switch (op) {
    case TOP_AND: return a & b;
    case TOP_NOT: return a & ~b;
    case TOP_XOR: return a ^ b;
    default: return a | b;
  }
}

constexpr int LOOKBACK_SPIN_LIMIT = 1 << 26;

template <int NMAX, int NT, bool TWO, bool AREA>
__global__ __launch_bounds__(NT) void k_tile(const TileDesc* __restrict__ desc, int ntiles,
                                             const vl_box* __restrict__ entries, TileOut o) {
  constexpr int P = 2 * (NMAX + 1);
  constexpr int NCOL = 2 * NMAX + 1;
  constexpr int NW = (NCOL + 63) / 64;
  __shared__ int32_t sx[P], sy[P];
  __shared__ int32_t ux[P], uy[P];
  __shared__ int16_t cb[NMAX][4];
  __shared__ uint64_t MA[NCOL * NW];
  __shared__ uint64_t MB[TWO ? NCOL * NW : 1];
  __shared__ uint64_t colc[NCOL];
  __shared__ int wsI[NT / WAVE];
  __shared__ uint64_t wsU[NT / WAVE];
  __shared__ long long wsA[NT / WAVE];
  __shared__ int s_tile, s_nx, s_ny;
  __shared__ uint64_t s_offF, s_offS;
  __shared__ int s_okF, s_okS;

  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_tile = (int)atomicAdd(o.ticket, 1u);
    __syncthreads();
    const int t = s_tile;
    if (t >= ntiles) break;
    const TileDesc d = desc[t];
    const int nA = d.nA;
    const int n = d.nA + d.nB;
    const int op = d.op;
    const bool twol = TWO && (op == TOP_AND || op == TOP_NOT || op == TOP_XOR);
    const bool big = n > NMAX;

    uint64_t F = 0, S = 0;
    int ncol = 0, nrows = 0, W = 0;
    long long my_area = 0;
    if (!big && n > 0) {
      // 1. coordinates
      const int cnt = 2 * n + 2;
      int Pn = 2;
      while (Pn < cnt) Pn <<= 1;
      for (int i = tid; i < Pn; i += NT) {
        int32_t vx = INT32_MAX, vy = INT32_MAX;
        if (i < 2 * n) {
          const vl_box b = entries[d.off + (i >> 1)];
          if (i & 1) { vx = min(b.x2, d.x1); vy = min(b.y2, d.y1); }
          else { vx = max(b.x1, d.x0); vy = max(b.y1, d.y0); }
        } else if (i == 2 * n) { vx = d.x0; vy = d.y0; }
        else if (i == 2 * n + 1) { vx = d.x1; vy = d.y1; }
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
      // 2. compressed coordinates, clear masks, rasterise
      for (int i = tid; i < n; i += NT) {
        const vl_box b = entries[d.off + i];
        cb[i][0] = (int16_t)lower_bound_i32(ux, nxu, max(b.x1, d.x0));
        cb[i][1] = (int16_t)lower_bound_i32(ux, nxu, min(b.x2, d.x1));
        cb[i][2] = (int16_t)lower_bound_i32(uy, nyu, max(b.y1, d.y0));
        cb[i][3] = (int16_t)lower_bound_i32(uy, nyu, min(b.y2, d.y1));
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
            atomicOr((unsigned long long*)&col[w], (unsigned long long)bits_range(lo, hi));
          }
        }
      }
      __syncthreads();
      // 3. combine
      if (twol) {
        for (int k = tid; k < words; k += NT) MA[k] = combine(op, MA[k], MB[k]);
        __syncthreads();
      }
      // 4. count strips per slab (or area)
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
            const bool side = (a == 0 && (flags & SEAM_B)) || (b == nrows && (flags & SEAM_T)) ||
                              (j == 0 && (flags & SEAM_L)) || (k == ncol && (flags & SEAM_R));
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
      if (tid == 0) {
        long long s = 0;
        for (int i = 0; i < NT / WAVE; i++) s += wsA[i];
        if (s) atomicAdd((unsigned long long*)o.area, (unsigned long long)s);
        if (big) o.ovf[atomicAdd(o.ovf_n, 1u)] = (uint32_t)t;
      }
      __syncthreads();
      continue;
    }

    // 5. decoupled look-back over tile order for output offsets
    if (tid == 0) {
      if (big) o.ovf[atomicAdd(o.ovf_n, 1u)] = (uint32_t)t;
      const uint64_t agg = lb_pack(F, S);
      uint64_t excl = 0;
      if (t == 0) {
        __hip_atomic_store(&o.status[0], LB_PRE | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_store(&o.status[t], LB_AGG | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int j = t - 1;
        int spins = 0;
        for (;;) {
          const uint64_t v = __hip_atomic_load(&o.status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t fl = v & LB_FLAGMASK;
          if (fl == 0) {
            if (++spins > LOOKBACK_SPIN_LIMIT) { atomicOr(o.err, ERR_LOOKBACK); break; }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          excl += v & LB_VALMASK;
          if (fl == LB_PRE || j == 0) break;
          j--;
        }
        __hip_atomic_store(&o.status[t], LB_PRE | (excl + agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
      const uint64_t eF = excl >> 31, eS = excl & 0x7fffffffull;
      s_offF = o.out_base + eF;
      s_offS = o.side_base + eS;
      s_okF = (o.out_base + eF + F) <= o.out_cap;
      s_okS = (o.side_base + eS + S) <= o.side_cap;
      if (!s_okF || !s_okS) atomicOr(o.err, ERR_OVERFLOW_CAP);
      if (t == ntiles - 1) {
        o.totals[0] = eF + F;
        o.totals[1] = eS + S;
      }
    }
    __syncthreads();

    // 6. emit
    if (!big && n > 0 && (F + S) > 0) {
      const uint64_t offF = s_offF, offS = s_offS;
      const bool okF = s_okF, okS = s_okS;
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
          const uint32_t tag = ((a == 0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                               ((b == nrows && (flags & SEAM_T)) ? SEAM_T : 0u) |
                               ((j == 0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                               ((k == ncol && (flags & SEAM_R)) ? SEAM_R : 0u);
          vl_box s;
          s.x1 = ux[j]; s.y1 = uy[a]; s.x2 = ux[k]; s.y2 = uy[b];
          if (tag) {
            if (okS) { o.side[si] = s; o.side_tag[si] = (uint8_t)tag; }
            si++;
          } else {
            if (okF) o.out[fi] = s;
            fi++;
          }
        });
      }
    }
    __syncthreads();
  }
}

// ---- overflow splitting ------------------------------------------------------
// One workgroup per parent tile; sub-tiles are x-sub-columns of the parent.
__device__ __forceinline__ void sub_range(const vl_box& b, const TileDesc& p, const TileDesc* subs, int first,
                                          int nsub, int& s1, int& s2) {
  const int32_t x1 = max(b.x1, p.x0), x2 = min(b.x2, p.x1);
  const int64_t sw = (int64_t)subs[first].x1 - subs[first].x0;   // all but the last have width sw
  s1 = (int)(((int64_t)x1 - p.x0) / sw);
  s2 = (int)(((int64_t)x2 - 1 - p.x0) / sw);
  if (s1 >= nsub) s1 = nsub - 1;
  if (s2 >= nsub) s2 = nsub - 1;
}

__global__ void k_split_count(const TileDesc* __restrict__ parents, const int32_t* __restrict__ sub_first,
                              const TileDesc* __restrict__ subs, uint32_t* cntA, uint32_t* cntB,
                              const vl_box* __restrict__ entries) {
  const TileDesc p = parents[blockIdx.x];
  const int first = sub_first[blockIdx.x], nsub = sub_first[blockIdx.x + 1] - first;
  const int n = p.nA + p.nB;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const vl_box b = entries[p.off + i];
    int s1, s2;
    sub_range(b, p, subs, first, nsub, s1, s2);
    for (int s = s1; s <= s2; s++) atomicAdd(i < p.nA ? &cntA[first + s] : &cntB[first + s], 1u);
  }
}

__global__ void k_split_scatter(const TileDesc* __restrict__ parents, const int32_t* __restrict__ sub_first,
                                const TileDesc* __restrict__ subs, const uint64_t* __restrict__ off,
                                const uint32_t* __restrict__ cntA, uint32_t* fillA, uint32_t* fillB,
                                const vl_box* __restrict__ entries, vl_box* __restrict__ sub_entries) {
  const TileDesc p = parents[blockIdx.x];
  const int first = sub_first[blockIdx.x], nsub = sub_first[blockIdx.x + 1] - first;
  const int n = p.nA + p.nB;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const vl_box b = entries[p.off + i];
    int s1, s2;
    sub_range(b, p, subs, first, nsub, s1, s2);
    for (int s = s1; s <= s2; s++) {
      const int q = first + s;
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

template <int NMAX, bool TWO, bool AREA>
void launch_tile_t(const TileDesc* desc, int ntiles, const vl_box* entries, TileOut o, int max_wg, hipStream_t st) {
  int g = ntiles < max_wg ? ntiles : max_wg;
  if (g < 1) g = 1;
  hipLaunchKernelGGL((k_tile<NMAX, 256, TWO, AREA>), dim3(g), dim3(256), 0, st, desc, ntiles, entries, o);
}

template <int NMAX, bool TWO, bool AREA>
int max_wg_t() {
  int dev = 0, ncu = 0, per = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_tile<NMAX, 256, TWO, AREA>, 256, 0);
  if (per < 1) per = 1;
  return ncu * per;
}

}  // namespace

int tile_nmax_max() { return 255; }

void launch_bin_count(const vl_box* b, int64_t n, Xform xf, Grid g, uint32_t* cnt, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_bin_count, dim3(grid_for(n)), dim3(256), 0, st, b, n, xf, g, cnt);
}
void launch_bin_scatter(const vl_box* b, int64_t n, Xform xf, Grid g, const uint64_t* off,
                        const uint32_t* base_add, uint32_t* fill, vl_box* entries, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(n)), dim3(256), 0, st, b, n, xf, g, off, base_add, fill,
                       entries);
}
void launch_build_desc(Grid g, const uint64_t* off, const uint32_t* cntA, const uint32_t* cntB, int32_t op,
                       TileDesc* desc, hipStream_t st) {
  int64_t nt = (int64_t)g.cols * g.rows;
  hipLaunchKernelGGL(k_build_desc, dim3(grid_for(nt)), dim3(256), 0, st, g, off, cntA, cntB, op, desc);
}

#define VL_TILE_DISPATCH(NM)                                                              \
  if (nmax == NM) {                                                                       \
    if (area_mode) {                                                                      \
      if (two_layers) launch_tile_t<NM, true, true>(desc, ntiles, entries, o, max_wg, st); \
      else launch_tile_t<NM, false, true>(desc, ntiles, entries, o, max_wg, st);           \
    } else {                                                                              \
      if (two_layers) launch_tile_t<NM, true, false>(desc, ntiles, entries, o, max_wg, st);\
      else launch_tile_t<NM, false, false>(desc, ntiles, entries, o, max_wg, st);          \
    }                                                                                     \
    return;                                                                               \
  }

void launch_tile_kernel(int nmax, bool two_layers, bool area_mode, const TileDesc* desc, int ntiles,
                        const vl_box* entries, TileOut o, int max_wg, hipStream_t st) {
  VL_TILE_DISPATCH(63)
  VL_TILE_DISPATCH(127)
  VL_TILE_DISPATCH(255)
}

#define VL_TILE_MAXWG(NM)                                                   \
  if (nmax == NM) {                                                         \
    if (area_mode) return two_layers ? max_wg_t<NM, true, true>() : max_wg_t<NM, false, true>(); \
    return two_layers ? max_wg_t<NM, true, false>() : max_wg_t<NM, false, false>();              \
  }

int tile_kernel_max_wg(int nmax, bool two_layers, bool area_mode) {
  VL_TILE_MAXWG(63)
  VL_TILE_MAXWG(127)
  VL_TILE_MAXWG(255)
  return 256;
}

void launch_split_count(const TileDesc* parents, int np, const int32_t* sub_first, const TileDesc* subs,
                        uint32_t* sub_cntA, uint32_t* sub_cntB, const vl_box* entries, hipStream_t st) {
  if (np > 0)
    hipLaunchKernelGGL(k_split_count, dim3(np), dim3(256), 0, st, parents, sub_first, subs, sub_cntA, sub_cntB,
                       entries);
}

void launch_split_scatter(const TileDesc* parents, int np, const int32_t* sub_first, TileDesc* subs, int ns,
                           const uint64_t* off, const uint32_t* cntA, const uint32_t* cntB, uint32_t* fillA,
                           uint32_t* fillB, const vl_box* entries, vl_box* sub_entries, hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_split_scatter, dim3(np), dim3(256), 0, st, parents, sub_first, subs, off, cntA, fillA,
                     fillB, entries, sub_entries);
  hipLaunchKernelGGL(k_fill_subdesc, dim3(grid_for(ns)), dim3(256), 0, st, subs, ns, off, cntA, cntB);
}

}  // namespace vl
