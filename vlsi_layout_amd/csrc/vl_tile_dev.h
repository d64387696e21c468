// vl_tile_dev.h — device helpers of the per-tile kernels (vl_tile.hip,
// vl_fused.hip): compressed-grid lookups, run masks, 64x64 bit transposes,
// erosion passes, overflow lists, and the occupancy-sized persistent grids.
#pragma once
#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {
namespace {

// ---- shared device helpers -----------------------------------------------------

__device__ __forceinline__ int lower_bound_i32(const int32_t* a, int n, int32_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// lower_bound over a sorted array of n <= P entries (P a power of two) in a
// fixed log2(P) + 1 steps: no data-dependent loop, so the lanes of a wave
// stay converged (a step moves lo up by `step` when the entry just below
// lo + step is still below v).
template <int P>
__device__ __forceinline__ int lower_bound_fixed(const int32_t* a, int n, int32_t v) {
  int lo = 0;
#pragma unroll
  for (int step = P; step > 0; step >>= 1) {
    const int t = lo + step;
    const bool up = t <= n && a[max(min(t, n) - 1, 0)] < v;
    lo = up ? t : lo;
  }
  return lo;
}

// Bits of rows [r0, r1) in word w.
__device__ __forceinline__ uint64_t row_bits(int w, int r0, int r1) {
  const int lo = max(r0 - 64 * w, 0), hi = min(r1 - 64 * w, 64);
  return lo < hi ? bits_range(lo, hi - 1) : 0ull;
}

__device__ __forceinline__ uint64_t combine(int op, uint64_t a, uint64_t b) {
  switch (op) {
    case TOP_AND: return a & b;
    case TOP_NOT: return a & ~b;
    case TOP_XOR: return a ^ b;
    case TOP_OVERLAP: return b;   // b holds the cells a second box reached
    default: return a | b;
  }
}

// Bit w of the dilated change mask of columns cur/prv (null = empty column).
template <int NW>
__device__ __forceinline__ void change_mask(const uint64_t* cur, const uint64_t* prv, int W, uint64_t* D) {
  uint64_t dl[NW + 2];
  dl[0] = 0;
  dl[W + 1] = 0;
#pragma unroll
  for (int w = 0; w < NW; w++)
    if (w < W) dl[w + 1] = (cur ? cur[w] : 0ull) ^ (prv ? prv[w] : 0ull);
#pragma unroll
  for (int w = 0; w < NW; w++)
    if (w < W) D[w] = dl[w + 1] | (dl[w + 1] << 1) | (dl[w] >> 63) | (dl[w + 1] >> 1) | (dl[w + 2] << 63);
}

// First set bit above position j of the NW-word bitset re, or `none`.  The
// word holding j+1 is picked directly and the words above only while nothing
// is found (an unrolled scan of every word kept far more registers live in
// the strip loop).
template <int NW>
__device__ __forceinline__ int next_bit_above(const uint64_t (&re)[NW], int j, int none) {
  const int t = j + 1;
  const int qw = t >> 6;
  uint64_t mm = 0ull;
#pragma unroll
  for (int q = 0; q < NW; q++)
    if (q == qw) mm = re[q] & (~0ull << (t & 63));
  int bq = qw;
#pragma unroll
  for (int s = 1; s < NW; s++)
    if (mm == 0ull && qw + s < NW) {
#pragma unroll
      for (int q = 0; q < NW; q++)
        if (q == qw + s) mm = re[q];
      bq = qw + s;
    }
  return mm ? bq * WAVE + __ffsll((unsigned long long)mm) - 1 : none;
}

// End marks of the runs of col (col null = empty; W <= NW words, bit i of
// word w = row 64w+i) that meet D: bit b is set when col has a run [a, b)
// holding a bit of D, b being the clear row just above the run (the grid
// keeps one row above its last: W = (nrows + 64) / 64).  Adding the run's D
// bits to the column carries through the rest of the run into row b and
// stops there, so one multi-word addition marks every run that meets D.
template <int NW>
__device__ __forceinline__ void ends_meeting(const uint64_t* col, int W, const uint64_t* D, uint64_t* out) {
  uint64_t carry = 0, below = 0;   // carry into word w; bit 63 of word w-1
#pragma unroll
  for (int w = 0; w < NW; w++) {
    out[w] = 0;
    if (w >= W || !col) continue;
    const uint64_t r = col[w];
    const uint64_t s1 = r + (D[w] & r);
    const uint64_t sum = s1 + carry;
    carry = (uint64_t)(s1 < r) | (uint64_t)(sum < s1);
    out[w] = sum & ~r & ((r << 1) | below);
    below = r >> 63;
  }
}

// Lowest row of the run of a column whose clear row just above is b (b >= 1):
// one above the highest clear bit below b, or 0.
template <int NW>
__device__ __forceinline__ int run_start_w(const uint64_t* col, int W, int b) {
  int w = (b - 1) >> 6;
  uint64_t m = ~col[w] & (((b >> 6) == w) ? ((1ull << (b & 63)) - 1) : ~0ull);
  if constexpr (NW == 1) {
    return m ? 64 - __clzll((unsigned long long)m) : 0;
  } else {
    while (m == 0) {
      if (--w < 0) return 0;
      m = ~col[w];
    }
    return 64 * w + 64 - __clzll((unsigned long long)m);
  }
}

// Value of v in lane l (l uniform across the wave).
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Value of v in lane l, where l may differ per lane (every lane of the wave
// must be active).
__device__ __forceinline__ uint64_t lane_read64(uint64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(l << 2, (int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(l << 2, (int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Number of set bits of a wave mask below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// 64x64 bit-matrix transpose across a wave: lane l holds row l (bit c = entry
// (l, c)) and gets column l.  Six block-swap stages: at stage s, entry (l, c)
// with bit s of l clear and bit s of c set trades places with (l^s, c^s).
// The lane-dependent choices are bit selects with an all-ones/all-zeros
// lane mask, not branches: the lower lane of a pair keeps its M bits and takes
// the partner's bits shifted up, the upper lane keeps its ~M bits and takes
// the partner's bits shifted down.
__device__ __forceinline__ uint64_t transpose64(uint64_t x, int lane) {
  constexpr uint64_t M[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                             0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
  const auto bsel = [](uint64_t m, uint64_t a, uint64_t b) { return (a & m) | (b & ~m); };
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const int sft = 32 >> i;
    // all ones in the upper lane of the pair (bit 5-i of the lane); computed
    // here rather than hoisted, which would hold 12 lane masks in registers
    uint32_t s32;
    asm volatile("v_bfe_i32 %0, %1, %2, 1" : "=v"(s32) : "v"(lane), "i"(5 - i));
    const uint64_t up = ((uint64_t)s32 << 32) | s32;
    const uint64_t y = xor_lanes64(x, sft);
    const uint64_t ys = bsel(up, y >> sft, y << sft);
    x = bsel(M[i] ^ up, x, ys);
  }
  return x;
}

__device__ __forceinline__ uint64_t wave_excl_u32(uint32_t v, uint32_t* tot) {
  uint32_t inc = wave_incl_scan(v);
  *tot = read_lane(inc, WAVE - 1);
  return inc - v;
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


// ---- erosion ------------------------------------------------------------------------
// Erosion by the square [-R, R]^2 on unit cells: cell p survives when every
// cell within R of p (in x and in y) is covered.  It is separable (erode in x,
// then in y).  The compressed grid of an erosion tile holds, besides the
// clipped box edges, every edge shifted inwards by R (clamped to the tile), so
// both passes map runs onto existing grid lines: a run [u_a, u_b) becomes
// [u_a + R, u_b - R).  Columns and rows outside the tile end up empty.

// Number of v[0..n) below x / at most x (v sorted).
__device__ __forceinline__ int count_below(const int32_t* v, int n, int64_t x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)v[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int count_at_most(const int32_t* v, int n, int64_t x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((int64_t)v[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// count_below(v, n, x) for x the same in every lane of the wave and n <= 64K:
// one ballot per group of 64 entries instead of a binary search, whose
// dependent LDS reads would run as a divergent loop.
template <int K>
__device__ __forceinline__ int count_below_wave(const int32_t* v, int n, int32_t x, int lane) {
  int c = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int e = k * WAVE + lane;
    c += __popcll(__ballot(e < n && v[min(e, n - 1)] < x));
  }
  return c;
}

// count_below(v, n, x) (n >= 1) when the answer is known to lie near index
// h: two branch-free steps from h (each reads v[h-1] and v[h] and moves h by
// at most one), then a binary search for the lanes whose walk has not
// settled.  Erosion lookups land within a step or two of the grid line they
// start from.
__device__ __forceinline__ int count_below_near(const int32_t* v, int n, int64_t x, int h) {
  h = h < 0 ? 0 : (h > n ? n : h);
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const int64_t lo = v[max(h - 1, 0)], hi = v[min(h, n - 1)];
    const bool up = h < n && hi < x;     // v sorted: up and dn exclude each other
    const bool dn = h > 0 && lo >= x;
    h += (int)up - (int)dn;
  }
  const int64_t lo = v[max(h - 1, 0)], hi = v[min(h, n - 1)];
  const bool settled = (h == n || hi >= x) && (h == 0 || lo < x);
  return settled ? h : count_below(v, n, x);
}

__device__ __forceinline__ int32_t clamp64(int64_t v, int64_t lo, int64_t hi) {
  return (int32_t)(v < lo ? lo : (v > hi ? hi : v));
}

// x pass: column c of the tile (c0 <= c < c1) keeps the rows covered in every
// column meeting [ux[c] - R, ux[c+1] + R); windows leaving the grid keep nothing.
template <int NW>
__device__ __forceinline__ void erode_x(const uint64_t* M, uint64_t* out, const int32_t* ux, int nxu, int ncol,
                                        int W, int c0, int c1, int32_t R, int t, int nt) {
  for (int c = t; c < ncol; c += nt) {
    uint64_t acc[NW];
    bool keep = c >= c0 && c < c1;
    int lo = 0, hi = -1;
    if (keep) {
      lo = count_below_near(ux, nxu, (int64_t)ux[c] - R + 1, c) - 1;   // last line <= ux[c] - R
      hi = count_below_near(ux, nxu, (int64_t)ux[c + 1] + R, c + 1) - 1;
      keep = lo >= 0 && hi < ncol;
    }
#pragma unroll
    for (int w = 0; w < NW; w++) acc[w] = keep ? ~0ull : 0ull;
    // the wave runs its longest window; a lane past its own window re-reads
    // its last column (AND is idempotent), a lane keeping nothing reads
    // column 0 into a zero accumulator
    const int len = keep ? hi - lo + 1 : 0;
    for (int s = 0; __ballot(s < len); s++) {
      const int j = keep ? min(lo + s, hi) : 0;
#pragma unroll
      for (int w = 0; w < NW; w++)
        if (w < W) acc[w] &= M[j * W + w];
    }
#pragma unroll
    for (int w = 0; w < NW; w++)
      if (w < W) out[c * W + w] = acc[w];
  }
}

// Row maps of the y pass: for every grid line i, the index of the line
// uy[i] + R (up) and of the line uy[i] - R (dn), both clamped to the tile's
// [Y0, Y1] (one walk per line, lane-parallel, instead of two per run).
__device__ __forceinline__ void erosion_row_maps(const int32_t* uy, int nyu, int32_t Y0, int32_t Y1, int32_t R,
                                                 uint16_t* up, uint16_t* dn, int t, int nt) {
  for (int i = t; i < nyu; i += nt) {
    up[i] = (uint16_t)count_below_near(uy, nyu, clamp64((int64_t)uy[i] + R, Y0, Y1), i);
    dn[i] = (uint16_t)count_below_near(uy, nyu, clamp64((int64_t)uy[i] - R, Y0, Y1), i);
  }
}

// y pass: each run [a, b) of a column becomes the rows [up[a], dn[b]), i.e.
// [clamp(uy[a] + R), clamp(uy[b] - R)) clamped to the tile's [Y0, Y1].  Run
// borders of the x-eroded mask are box edges or window borders, whose shifted
// lines are on the grid.
template <int NW>
__device__ __forceinline__ void erode_y(const uint64_t* M, uint64_t* out, const uint16_t* up, const uint16_t* dn,
                                        int ncol, int W, int t, int nt) {
  if constexpr (NW == 1) {
    // one word per column (bit 63 clear): the k-th run start pairs with the
    // k-th run end; the wave runs its largest run count, lanes predicated
    for (int c = t; c < ncol; c += nt) {
      const uint64_t v = M[c];
      uint64_t st = v & ~(v << 1), en = ~v & (v << 1), res = 0ull;
      while (__ballot(st != 0ull)) {
        const bool has = st != 0ull;
        const int a = has ? __ffsll((unsigned long long)st) - 1 : 0;
        const int b = has ? __ffsll((unsigned long long)en) - 1 : 0;
        st &= st - 1;
        en &= en - 1;
        const int ia = up[a], ib = dn[b];
        if (has && ia < ib) res |= bits_range(ia, ib - 1);
      }
      out[c] = res;
    }
    return;
  }
  if constexpr (NW == 2) {
    // two words (the grid's top row is clear, so every run ends inside the
    // column): the same pairing on the 128-bit start and end masks
    for (int c = t; c < ncol; c += nt) {
      const uint64_t v0 = M[c * W], v1 = W > 1 ? M[c * W + 1] : 0ull;
      uint64_t s0 = v0 & ~(v0 << 1), s1 = v1 & ~((v1 << 1) | (v0 >> 63));
      uint64_t e0 = ~v0 & (v0 << 1), e1 = ~v1 & ((v1 << 1) | (v0 >> 63));
      uint64_t r0 = 0ull, r1 = 0ull;
      while (__ballot((s0 | s1) != 0ull)) {
        const bool has = (s0 | s1) != 0ull;
        const int a = s0 ? __ffsll((unsigned long long)s0) - 1 : (s1 ? 64 + __ffsll((unsigned long long)s1) - 1 : 0);
        const int b = e0 ? __ffsll((unsigned long long)e0) - 1 : (e1 ? 64 + __ffsll((unsigned long long)e1) - 1 : 0);
        if (s0) s0 &= s0 - 1; else s1 &= s1 - 1;
        if (e0) e0 &= e0 - 1; else e1 &= e1 - 1;
        const int ia = up[a], ib = dn[b];
        if (has && ia < ib) {
          if (ia < 64) r0 |= bits_range(ia, min(ib, 64) - 1);
          if (ib > 64) r1 |= bits_range(max(ia, 64) - 64, ib - 65);
        }
      }
      out[c * W] = r0;
      if (W > 1) out[c * W + 1] = r1;
    }
    return;
  }
  for (int c = t; c < ncol; c += nt) {
    uint64_t res[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) res[w] = 0ull;
    for_each_run(M + c * W, W, [&](int a, int b) {
      const int ia = up[a], ib = dn[b];
      if (ia >= ib) return;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const int l = max(ia, 64 * w), h = min(ib, 64 * w + 64);
        if (l < h) res[w] |= bits_range(l - 64 * w, h - 1 - 64 * w);
      }
    });
#pragma unroll
    for (int w = 0; w < NW; w++)
      if (w < W) out[c * W + w] = res[w];
  }
}

// ---- overflow lists -------------------------------------------------------------
constexpr int TILE_BATCH = 4;   // tiles per load batch of the wave kernel
constexpr int OVF_BUF = 256;    // overflow tiles buffered in LDS per workgroup

// Appends the n buffered overflow tiles to the launch's overflow list: one
// atomic per flush of the workgroup, not one per tile (a single counter
// address serialises its atomics across the whole device).
__device__ __forceinline__ void flush_overflow(const uint32_t* buf, uint32_t n, uint32_t* ovf, uint32_t* ovf_n,
                                               int t, int nt) {
  __shared__ uint32_t s_base;
  if (t == 0) s_base = atomicAdd(ovf_n, n);
  __syncthreads();
  const uint32_t base = s_base;
  for (uint32_t k = t; k < n; k += nt) ovf[base + k] = buf[k];
  __syncthreads();
}

// ---- persistent grids -------------------------------------------------------------
int num_cus() {
  static int ncu = 0;
  if (!ncu) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (ncu < 1) ncu = 1;
  }
  return ncu;
}

// Workgroups of a persistent launch: the occupancy limit, but for the
// one-wave tile kernels at most 24 waves per CU (6 per SIMD): measured on the
// 10M-box OP1 ring, 7 or 8 resident waves per SIMD run the tile stage 7 %
// slower than 6 (and 4 run it 14 % slower).
constexpr int WAVE_KERNEL_WAVES_PER_CU = 24;

template <class KF>
int occupancy_wg(KF kern, int threads, int wave_cap = WAVE_KERNEL_WAVES_PER_CU) {
  int per = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)kern, threads, 0);
  if (per < 1) per = 1;
  if (threads == WAVE && per > wave_cap) per = wave_cap;
  // one-wave workgroups with a fixed share of the tiles each: a count per CU
  // that is not a multiple of the CU's 4 SIMDs leaves some SIMDs one wave
  // more, and their workgroups finish last
  if (threads == WAVE && per >= 4) per &= ~3;
  return per * num_cus();
}

}  // namespace
}  // namespace vl
