// vl_fused.hip — fused per-tile evaluation of several operators over one
// source layer.
//
// A recorded program often derives a whole family of layers from one source
// S: the OP1 ring (src/op_template.m4:14-22) computes GROW(S), SIZE(S,-d),
// GROW(S) NOT S, S NOT SIZE(S,-d) and two ANDs of those.  Evaluated one
// operator at a time, every intermediate layer is written, re-read, re-binned
// and re-sorted.  Here S is binned once, with tile windows grown by the
// largest morphological reach, and each tile builds, on one compressed grid,
//   * the coverage of S,
//   * up to two morphological images of S: GROW by (e, n, w, s) (the grown
//     boxes rasterised directly) or erosion by the square of half-side d (the
//     erosion passes of the SIZE kernel),
// and then every requested output as a Boolean function of those three masks
// (a truth table), clipped to the tile, through the same start/end-mask strip
// extraction as the single-operator kernel.  Each output has its own scratch
// stream, tables and seam side stream, so it is finished (seam levels,
// x-stitch) exactly like a single operator's result: the output is the
// canonical maximal-strip decomposition of the same point set.
//
// The compressed grid holds, per box, x1 and x2 clipped to the window and the
// two edges of each morphological image (grown edges clipped to the window,
// eroded edges x1+d, x2-d clamped to the tile): 6 coordinates per box and
// axis, plus the window's and the tile's borders, so a 64-slot grid takes 10
// boxes, a 128-slot grid 20 and a 256-slot grid 42.
#include "vl_tile_dev.h"

namespace vl {

namespace {

constexpr int FUSE_OVF_BUF = 16;   // overflow tiles buffered in LDS per workgroup
// resident-workgroup cap of the fused grid: 24 one-wave workgroups per CU
// (6 per SIMD) measured best on the 64-slot tier, ahead of 20, 28 and 32
constexpr int FUSED_WAVES_PER_CU = 24;

// arr[q] for a wave-uniform q, by selects over the kernel argument's array
// (no copy in LDS)
template <typename T>
__device__ __forceinline__ T pick_out(const T (&arr)[FUSE_MAX_OUT], int q) {
  T v = arr[0];
#pragma unroll
  for (int k = 1; k < FUSE_MAX_OUT; k++)
    if (k == q) v = arr[k];
  return v;
}

// Value of truth table tt (bit s | m0 << 1 | m1 << 2) on base words s, m0, m1
// (bit 0, the value outside every mask, is taken as 0).  Shannon expansion on
// m1, m0 and s with the table's bits as all-ones / all-zero words: seven
// bit-selects, no branches.
__device__ __forceinline__ uint64_t tt_eval(uint32_t tt, uint64_t s, uint64_t m0, uint64_t m1) {
  const auto c = [tt](int k) { return ((tt >> k) & 1u) ? ~0ull : 0ull; };
  const auto sel = [](uint64_t m, uint64_t a, uint64_t b) { return (m & a) | (~m & b); };
  const uint64_t f00 = s & c(1), f01 = sel(s, c(3), c(2));
  const uint64_t f10 = sel(s, c(5), c(4)), f11 = sel(s, c(7), c(6));
  return sel(m1, sel(m0, f11, f10), sel(m0, f01, f00));
}

// The same, with the tables of the outputs most groups have (a morphological
// image itself, and the ring differences m0 & ~s and s & ~m1) as plain word
// operations ahead of the general evaluation (tt uniform).
__device__ __forceinline__ uint64_t tt_eval_common(uint32_t tt, uint64_t s, uint64_t m0, uint64_t m1) {
  switch (tt) {
    case 0xCCu: return m0;
    case 0xF0u: return m1;
    case 0x44u: return m0 & ~s;
    case 0x0Au: return s & ~m1;
    default: return tt_eval(tt, s, m0, m1);
  }
}

// The 64-slot tier asks for 6 waves per SIMD (at most 80 VGPRs, no spills):
// its shared strip pass holds every output's transposed end marks, and
// without the bound the compiler takes 86 VGPRs (5 waves per SIMD), which
// costs more than the pass saves.
template <int P>
__global__ __launch_bounds__(64, P == 64 ? 6 : 1) void k_tile_f(const TileDesc* __restrict__ desc, int ntiles,
                                               const vl_box* __restrict__ entries, FusedSpec fs, FusedOut o) {
  constexpr int CPB = 6;               // coordinates per box and axis
  // the slot group after the boxes holds the window and tile borders: four
  // slots (components 0..3; components 4 and 5 repeat the window's, and a
  // grid the boxes fill to within four slots has no room for them)
  constexpr int NMAX = (P - 4) / CPB;
  constexpr int K = P / WAVE;          // slots per lane
  constexpr int NCOL = P - 1;
  constexpr int NW = P / WAVE;         // words per column
  constexpr int NL = (NMAX + WAVE) / WAVE;
  static_assert(NMAX == (P == 64 ? FUSE_NMAX_TINY : P == 128 ? FUSE_NMAX_SMALL : FUSE_NMAX_LARGE), "tier capacities");
  __shared__ int32_t ux[P], uy[P];
  __shared__ uint8_t sidx_x[P], sidx_y[P]; // column / row index of every coordinate slot
  // row maps of the y erosion; during output extraction the strip-start table
  // (4 entries per lane on 64-slot grids, 8 on larger ones: one round takes
  // nearly every tile's starts)
  constexpr int EROW = 2 * P > (P == 64 ? 4 : 8) * WAVE ? 2 * P : (P == 64 ? 4 : 8) * WAVE;
  __shared__ uint16_t erow[EROW];
  uint16_t* const eup = erow;
  uint16_t* const edn = erow + P;
  __shared__ uint64_t MS[NCOL * NW];       // S
  __shared__ uint64_t MM[2][NCOL * NW];    // morphological images of S
  constexpr int NQ = NW == 1 ? 4 : 1;      // outputs evaluated in one pass (one-word grids)
  __shared__ uint64_t MTQ[NQ][NW == 1 ? NCOL : 1];
  // erosion temporary, then the current output; on one-word grids the first
  // slot of MTQ, which is written only after the erosion passes and unused
  // when the outputs are evaluated one at a time
  __shared__ uint64_t MTbuf[NW == 1 ? 1 : NCOL * NW];
  uint64_t* const MT = NW == 1 ? &MTQ[0][0] : MTbuf;
  __shared__ vl_box sbuf[TILE_BATCH * (NMAX + 1)];
  __shared__ TileDesc sdesc[TILE_BATCH];
  uint32_t* const sdw = (uint32_t*)sdesc;
  __shared__ uint32_t sovf[FUSE_OVF_BUF];
  // per-output parameters, copied once from the kernel arguments
  __shared__ vl_box* s_strips[FUSE_MAX_OUT];
  __shared__ uint8_t* s_tag[FUSE_MAX_OUT];
  __shared__ uint64_t* s_toff[FUSE_MAX_OUT];
  __shared__ uint32_t* s_tn[FUSE_MAX_OUT];
  __shared__ uint64_t s_used[FUSE_MAX_OUT], s_side[FUSE_MAX_OUT];
  // each output's current chunk of its stream: next free strip, chunk end
  // (the LDS footprint sets how many workgroups a CU holds: 12.5 KB on the
  // 128-slot tier is 13 per CU)
  __shared__ uint64_t s_cpos[FUSE_MAX_OUT], s_cend[FUSE_MAX_OUT];
  __shared__ uint32_t s_short;
  __shared__ uint32_t s_tt[FUSE_MAX_OUT];
  static_assert(TILE_BATCH * 10 <= WAVE, "one descriptor dword per lane");

  const int lane = threadIdx.x;
  const int nt = o.list_n ? (int)*o.list_n : ntiles;
  const int nout = fs.nout;
  const int nmorph = fs.nmorph;
  const int32_t R = fs.halo;
  // images rasterised from their slots: grown images and the second layer
  const bool g0 = nmorph > 0 && (fs.kind[0] == MORPH_GROW || fs.kind[0] == MORPH_LAYER);
  const bool g1 = nmorph > 1 && (fs.kind[1] == MORPH_GROW || fs.kind[1] == MORPH_LAYER);
  const bool e0 = nmorph > 0 && fs.kind[0] == MORPH_ERODE, e1 = nmorph > 1 && fs.kind[1] == MORPH_ERODE;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < FUSE_MAX_OUT; k++) {
      s_strips[k] = o.strips[k];
      s_tag[k] = o.tag[k];
      s_toff[k] = o.tab_off[k];
      s_tn[k] = o.tab_n[k];
      s_used[k] = 0;
      s_side[k] = 0;
      s_cpos[k] = s_cend[k] = 0;
      s_tt[k] = fs.tt[k];
    }
    s_short = 0u;
  }
  uint32_t novf = 0;   // entries in sovf (uniform)
  // Per-slot constants of the coordinate step (slot e = lane*K + k holds
  // component e % CPB of box e / CPB): bit 0 of cflag = upper edge, bit 1 =
  // edge of a grown image (raw edge plus offset, clamped to the window), bit
  // 2 = edge of an eroded image (clipped edge plus offset, clamped to the
  // tile), bits 3-4 = which border the slot holds in the border group (window
  // low / high, tile low / high), bit 6 = edge of the second layer's image
  // (the clipped edge of a second-layer box).  Other slots hold the clipped
  // edge.  A second-layer box (the tile's entries past nA) fills only its
  // second-layer image slots; a first-layer box every slot but those.  A slot
  // a box does not fill holds the window corner in both edges (an empty
  // image).
  uint32_t cflag[K];
  int32_t coffx[K], coffy[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int e = lane * K + k;
    const int comp = e % CPB;
    const int hi = comp & 1;
    const int m = (comp - 2) >> 1;   // morph of components 2..5
    const uint32_t border = (comp == 0 || comp == 4) ? 0u : (comp == 1 || comp == 5) ? 1u : (comp == 2) ? 2u : 3u;
    uint32_t f = (uint32_t)hi | (border << 3);
    int32_t ox = 0, oy = 0;
    if (comp >= 2 && m < nmorph) {
      const int32_t kind = m ? fs.kind[1] : fs.kind[0];
      const int32_t ae = m ? fs.arg[1][0] : fs.arg[0][0], an = m ? fs.arg[1][1] : fs.arg[0][1];
      const int32_t aw = m ? fs.arg[1][2] : fs.arg[0][2], as = m ? fs.arg[1][3] : fs.arg[0][3];
      if (kind == MORPH_GROW) {
        f |= 2u;
        ox = hi ? ae : -aw;
        oy = hi ? an : -as;
      } else if (kind == MORPH_LAYER) {
        f |= 64u;
      } else {   // erosion by ae: edges shifted inwards
        f |= 4u;
        ox = oy = hi ? -ae : ae;
      }
    }
    cflag[k] = f;
    coffx[k] = ox;
    coffy[k] = oy;
  }
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
      TileDesc d;
#pragma unroll
      for (int k = 0; k < 10; k++) ((uint32_t*)&d)[k] = uniform(sdw[jb * 10 + k]);
      const vl_box* bx = sbuf + jb * (NMAX + 1);
      const int n = d.nA + d.nB;
      if (n > NMAX || n == 0) {
        if (lane == 0) {
          if (n > NMAX) sovf[novf] = ti;
          for (int k = 0; k < nout; k++) {
            gstore(s_tn[k] + ti, 0u);
            gstore(s_tn[k] + o.ntab + ti, 0u);
          }
        }
        if (n > NMAX && ++novf == FUSE_OVF_BUF) {
          __syncthreads();
          flush_overflow(sovf, novf, o.ovf, o.ovf_n, lane, WAVE);
          novf = 0;
          __syncthreads();
        }
        continue;
      }
      // 1. coordinates: slot e = lane*K + k holds component e % 6 of box e / 6:
      //    x1, x2 clipped to the window, then two per morphological image.
      //    Slot group n holds the window and tile borders.
      const int32_t wx0 = clamp64((int64_t)d.x0 - R, INT32_MIN, INT32_MAX);
      const int32_t wx1 = clamp64((int64_t)d.x1 + R, INT32_MIN, INT32_MAX);
      const int32_t wy0 = clamp64((int64_t)d.y0 - R, INT32_MIN, INT32_MAX);
      const int32_t wy1 = clamp64((int64_t)d.y1 + R, INT32_MIN, INT32_MAX);
      int32_t vx[K], vy[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int i = (lane * K + k) / CPB;
        const uint32_t f = cflag[k];
        const vl_box b = bx[min(i, NMAX)];
        const int32_t rx = (f & 1u) ? b.x2 : b.x1, ry = (f & 1u) ? b.y2 : b.y1;
        const int32_t kx = (f & 1u) ? min(b.x2, wx1) : max(b.x1, wx0);   // clipped to the window
        const int32_t ky = (f & 1u) ? min(b.y2, wy1) : max(b.y1, wy0);
        const bool grown = f & 2u, eroded = f & 4u;
        const int32_t cx = clamp64((int64_t)(grown ? rx : kx) + coffx[k], eroded ? d.x0 : wx0, eroded ? d.x1 : wx1);
        const int32_t cy = clamp64((int64_t)(grown ? ry : ky) + coffy[k], eroded ? d.y0 : wy0, eroded ? d.y1 : wy1);
        const uint32_t bsel = (f >> 3) & 3u;
        const int32_t bxv = bsel == 0 ? wx0 : bsel == 1 ? wx1 : bsel == 2 ? d.x0 : d.x1;
        const int32_t byv = bsel == 0 ? wy0 : bsel == 1 ? wy1 : bsel == 2 ? d.y0 : d.y1;
        const bool nul = (i >= d.nA) != ((f & 64u) != 0u);   // a slot this box does not fill
        vx[k] = i < n ? (nul ? wx0 : cx) : (i == n ? bxv : INT32_MAX);
        vy[k] = i < n ? (nul ? wy0 : cy) : (i == n ? byv : INT32_MAX);
      }
      const int cnt = CPB * n + CPB;
      // Sort keys: the coordinate relative to the window corner (below 2^24:
      // the host only fuses grids whose windows are narrower) above the slot
      // number, so a slot travels with its value through the sort; padding
      // slots get the largest key.
      constexpr int SB = P == 64 ? 6 : P == 128 ? 7 : 8;   // slot bits
      uint32_t kx[K], ky[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int e = lane * K + k;
        const bool real = e / CPB <= n;
        kx[k] = real ? (((uint32_t)vx[k] - (uint32_t)wx0) << SB) | (uint32_t)e : ~0u;
        ky[k] = real ? (((uint32_t)vy[k] - (uint32_t)wy0) << SB) | (uint32_t)e : ~0u;
      }
      // bitonic sort of P keys, element e = lane*K + k
#pragma unroll
      for (int size = 2; size <= P; size <<= 1) {
#pragma unroll
        for (int j = size >> 1; j > 0; j >>= 1) {
          if (j >= K) {
            const int lj = j / K;
#pragma unroll
            for (int k = 0; k < K; k++) {
              const int e = lane * K + k;
              const uint32_t px = (uint32_t)xor_lanes((int)kx[k], lj);
              const uint32_t py = (uint32_t)xor_lanes((int)ky[k], lj);
              const bool take_min = ((e & j) == 0) == ((e & size) == 0);
              kx[k] = take_min ? min(kx[k], px) : max(kx[k], px);
              ky[k] = take_min ? min(ky[k], py) : max(ky[k], py);
            }
          } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
              if ((k & j) == 0) {
                const int k2 = k | j;
                const bool asc = ((lane * K + k) & size) == 0;
                uint32_t a = kx[k], b = kx[k2];
                if ((a > b) == asc) { kx[k] = b; kx[k2] = a; }
                a = ky[k]; b = ky[k2];
                if ((a > b) == asc) { ky[k] = b; ky[k2] = a; }
              }
            }
          }
        }
      }
      int ex[K], ey[K];   // slot of origin of the sorted values
#pragma unroll
      for (int k = 0; k < K; k++) {
        vx[k] = (int32_t)((kx[k] >> SB) + (uint32_t)wx0);
        vy[k] = (int32_t)((ky[k] >> SB) + (uint32_t)wy0);
        ex[k] = (int)(kx[k] & ((1u << SB) - 1));
        ey[k] = (int)(ky[k] & ((1u << SB) - 1));
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
        // distinct values in order; every slot learns the index of its value
#pragma unroll
        for (int k = 0; k < K; k++) {
          if (fx[k]) ux[ox++] = vx[k];
          if (fy[k]) uy[oy++] = vy[k];
          sidx_x[ex[k]] = (uint8_t)(ox - 1);
          sidx_y[ey[k]] = (uint8_t)(oy - 1);
        }
        nxu = (int)tx;
        nyu = (int)ty;
      }
      // one row above the last is kept clear for the run end marks
      // (a 64-slot grid has at most 63 rows: one word, W folds to 1)
      const int ncol = nxu - 1, nrows = nyu - 1, W = NW == 1 ? 1 : (nrows + 64) >> 6;
      if constexpr (NW <= 2) {
        // 2. rasterise S and the grown images, column-parallel: lane l ORs,
        //    into its columns 64*cb + l, the row words of every box image
        //    whose column range holds the column.  The edges of a box image
        //    are read from the lanes of its two slots (slot e is component
        //    e % K on lane e / K), so its row words are uniform values.
        __syncthreads();
        static_assert(K <= 2, "column-parallel raster on one- and two-word grids");
        // Box images, one per lane in parallel: the image whose edges are
        // slots e0, e0+1 (e0 = 6i, 6i+2, 6i+4) is prepared on the lane of
        // slot e0: its column range packed as cx1 | cx2 << 8 (empty when the
        // image is) and its row words.  Slot e is component e % K of lane
        // e / K, so with K == 1 the upper edge comes from the next lane.
        const uint32_t lo = (uint32_t)sidx_x[lane * K] | ((uint32_t)sidx_y[lane * K] << 8);
        uint32_t hi;
        if constexpr (K == 1) hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, 0x101, 0xF, 0xF, false);   // row_shl:1
        else hi = (uint32_t)sidx_x[lane * K + 1] | ((uint32_t)sidx_y[lane * K + 1] << 8);
        const int bx1 = (int)(lo & 0xffu), by1 = (int)(lo >> 8), bx2 = (int)(hi & 0xffu), by2 = (int)(hi >> 8);
        const bool live = bx1 < bx2 && by1 < by2;
        const uint32_t span = live ? (uint32_t)bx1 | ((uint32_t)(bx2 - bx1) << 8) : 0u;
        uint64_t rw[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) rw[w] = live ? (NW == 1 ? bits_range(by1, by2 - 1) : row_bits(w, by1, by2)) : 0ull;
        uint64_t acc[3][NW][NW];
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
          for (int cb = 0; cb < NW; cb++)
#pragma unroll
            for (int w = 0; w < NW; w++) acc[m][cb][w] = 0ull;
        for (int i = 0; i < n; i++) {
#pragma unroll
          for (int m = -1; m < 2; m++) {
            if (m == 0 && !g0) continue;
            if (m == 1 && !g1) continue;
            const int l0 = (CPB * i + (m < 0 ? 0 : 2 + 2 * m)) / K;
            const uint32_t sp = (uint32_t)__builtin_amdgcn_readlane((int)span, l0);
            const uint32_t x1 = sp & 0xffu, width = sp >> 8;
#pragma unroll
            for (int cb = 0; cb < NW; cb++) {
              const bool in = (uint32_t)(cb * WAVE + lane) - x1 < width;
#pragma unroll
              for (int w = 0; w < NW; w++) {
                const uint64_t rows = read_lane(rw[w], l0);
                if (in) acc[m + 1][cb][w] |= rows;
              }
            }
          }
        }
#pragma unroll
        for (int cb = 0; cb < NW; cb++) {
          const int c = cb * WAVE + lane;
          if (c < ncol) {
#pragma unroll
            for (int w = 0; w < NW; w++) {
              if (w >= W) break;
              MS[c * W + w] = acc[0][cb][w];
              if (g0) MM[0][c * W + w] = acc[1][cb][w];
              if (g1) MM[1][c * W + w] = acc[2][cb][w];
            }
          }
        }
        __syncthreads();
      } else {
        // (256-slot grids) box-parallel raster with LDS atomics
        const int words = ncol * W;
        for (int k = lane; k < words; k += WAVE) {
          MS[k] = 0ull;
          if (g0) MM[0][k] = 0ull;
          if (g1) MM[1][k] = 0ull;
        }
        __syncthreads();
        // 2. rasterise S and the grown images (box-parallel)
        for (int i = lane; i < n; i += WAVE) {
#pragma unroll
          for (int m = -1; m < 2; m++) {
            if (m == 0 && !g0) continue;
            if (m == 1 && !g1) continue;
            // the box's edges (S: slots 0, 1; grown image m: slots 2+2m, 3+2m)
            const int e0 = CPB * i + (m < 0 ? 0 : 2 + 2 * m);
            const int cx1 = sidx_x[e0], cx2 = sidx_x[e0 + 1];
            const int cy1 = sidx_y[e0], cy2 = sidx_y[e0 + 1];
            if (cx1 >= cx2 || cy1 >= cy2) continue;
            uint64_t* M = m < 0 ? MS : MM[m];
            const int w1 = cy1 >> 6, w2 = (cy2 - 1) >> 6;
            for (int c = cx1; c < cx2; c++)
              for (int w = w1; w <= w2; w++) {
                const int lo = (w == w1) ? (cy1 & 63) : 0;
                const int hi = (w == w2) ? ((cy2 - 1) & 63) : 63;
                atomicOr((unsigned long long*)&M[c * W + w], (unsigned long long)bits_range(lo, hi));
              }
          }
        }
        __syncthreads();
      }
      // the tile's own columns [colT0, colT1) and rows [rowT0, rowT1)
      const int colT0 = count_below_wave<K>(ux, nxu, d.x0, lane), colT1 = count_below_wave<K>(ux, nxu, d.x1, lane);
      const int rowT0 = count_below_wave<K>(uy, nyu, d.y0, lane), rowT1 = count_below_wave<K>(uy, nyu, d.y1, lane);
      // 3. erosion images
      if (e0) {
        erosion_row_maps(uy, nyu, d.y0, d.y1, fs.arg[0][0], eup, edn, lane, WAVE);
        erode_x<NW>(MS, MT, ux, nxu, ncol, W, colT0, colT1, fs.arg[0][0], lane, WAVE);
        __syncthreads();
        erode_y<NW>(MT, MM[0], eup, edn, ncol, W, lane, WAVE);
        __syncthreads();
      }
      if (e1) {
        erosion_row_maps(uy, nyu, d.y0, d.y1, fs.arg[1][0], eup, edn, lane, WAVE);
        erode_x<NW>(MS, MT, ux, nxu, ncol, W, colT0, colT1, fs.arg[1][0], lane, WAVE);
        __syncthreads();
        erode_y<NW>(MT, MM[1], eup, edn, ncol, W, lane, WAVE);
        __syncthreads();
      }
      const uint32_t flags = d.flags;
      const int KW = (ncol + 1 + 63) >> 6;   // boundary words per row
      // 4. outputs.  On one-word grids up to four outputs get their column
      //    words from one read of the base masks (bit q of anymask: output q
      //    is not empty); otherwise each output is evaluated into MT in turn.
      uint32_t anymask = 0;
      const bool onepass = NW == 1 && nout <= NQ;
      if (onepass) {
        const int c = lane;
        const bool in = c < ncol && c >= colT0 && c < colT1;
        uint64_t b0 = 0ull, b1 = 0ull, b2 = 0ull;
        if (in) {
          b0 = MS[c];
          if (nmorph > 0) b1 = MM[0][c];
          if (nmorph > 1) b2 = MM[1][c];
        }
        const uint64_t rb = row_bits(0, rowT0, rowT1);
#pragma unroll 1
        for (int q = 0; q < nout; q++) {
          const uint64_t v = in ? tt_eval_common(uniform(s_tt[q]), b0, b1, b2) & rb : 0ull;
          if (c < ncol) MTQ[min(q, NQ - 1)][c] = v;
          if (__ballot(v != 0ull)) anymask |= 1u << q;
        }
        __syncthreads();
      }
      bool shared_pass = false;
      if constexpr (NW == 1) {
        if (onepass) {
          // One-word grids, up to four outputs: each output's marks, stream
          // range and transposed end marks as below, but the strip starts of
          // all outputs go into one table (output q's entries contiguous, in
          // output order), and one pass over the table writes every output's
          // strips: entry g of output q goes to position off[q] + g, where
          // off[q] = its range's base - the table index of its first entry
          // (the same positions, hence the same order, as one pass per
          // output).  The table is written out whenever it cannot take a
          // full round (a start per lane).
          shared_pass = true;
          constexpr int Q = EROW / WAVE;
          uint16_t* const stab = erow;   // (j << 8) | (q << 6) | b
          uint64_t reQ[NQ];              // transposed end marks of each output
          uint32_t TQ[NQ], sideQ[NQ];
          uint64_t off[NQ];
          uint32_t okmask = 0;           // outputs whose range fits their stream
          uint32_t tab = 0;              // entries in the table (uniform)
          uint64_t gbase = 0;            // table index of stab[0] over the tile
#pragma unroll
          for (int q = 0; q < NQ; q++) {
            reQ[q] = 0ull;
            TQ[q] = 0u;
            sideQ[q] = 0u;
            off[q] = 0ull;
          }
          auto flush = [&]() {
            __syncthreads();
            for (uint32_t g0 = 0; g0 < tab; g0 += WAVE) {
              const uint32_t g = g0 + (uint32_t)lane;
              const bool has = g < tab;
              const uint32_t ent = has ? (uint32_t)stab[g] : 0u;
              const int j = (int)(ent >> 8), qq = (int)((ent >> 6) & 3u), b = (int)(ent & 63u);
              uint64_t rr[1] = {0ull};
              uint64_t po = 0ull;
#pragma unroll
              for (int q = 0; q < NQ; q++) {
                if (q >= nout) break;
                const uint64_t v = lane_read64(reQ[q], b);   // every lane active
                if (qq == q) {
                  rr[0] = v;
                  po = off[q];
                }
              }
              uint32_t tg = 0u;
              if (has) {
                const int k = next_bit_above<1>(rr, j, ncol);
                const int a = run_start_w<1>(MTQ[qq] + j, 1, b);
                tg = ((a == rowT0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                     ((b == rowT1 && (flags & SEAM_T)) ? SEAM_T : 0u) |
                     ((j == colT0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                     ((k == colT1 && (flags & SEAM_R)) ? SEAM_R : 0u);
                if ((okmask >> qq) & 1u) {
                  const uint64_t p = po + gbase + g;
                  gstore_box(s_strips[qq] + p, ux[j], uy[a], ux[k], uy[b]);
                  gstore(s_tag[qq] + p, (uint8_t)tg);
                }
              }
#pragma unroll
              for (int q = 0; q < NQ; q++)
                sideQ[q] += (uint32_t)__popcll(__ballot(tg != 0u && qq == q));
            }
            gbase += tab;
            tab = 0;
            __syncthreads();
          };
#pragma unroll
          for (int q = 0; q < NQ; q++) {
            if (q >= nout) break;
            if (!((anymask >> q) & 1u)) {
              if (lane == 0) {
                gstore(s_tn[q] + ti, 0u);
                gstore(s_tn[q] + o.ntab + ti, 0u);
              }
              continue;
            }
            const uint64_t* MO = MTQ[q];
            uint64_t Sr = 0ull, Er = 0ull;
            uint32_t myT = 0;
            if (lane <= ncol) {
              const uint64_t* cur = (lane < ncol) ? MO + lane : nullptr;
              const uint64_t* prv = (lane > 0) ? MO + (lane - 1) : nullptr;
              uint64_t D[1];
              change_mask<1>(cur, prv, 1, D);
              ends_meeting<1>(cur, 1, D, &Sr);
              ends_meeting<1>(prv, 1, D, &Er);
              myT = (uint32_t)__popcll(Er);
            }
            const uint32_t T = wave_sum(myT);
            uint64_t cpos = uniform(s_cpos[q]), cend = uniform(s_cend[q]);
            if (cpos + T > cend) {
              const uint64_t want = (uint64_t)T > o.chunk ? (uint64_t)T : o.chunk;
              unsigned long long nb = 0;
              if (lane == 0) nb = atomicAdd(pick_out(o.cnt, q) + 2, (unsigned long long)want);
              cpos = read_lane((uint64_t)nb, 0);
              cend = cpos + want;
            }
            const bool ok = cpos + T <= pick_out(o.stream_cap, q);
            if (ok) okmask |= 1u << q;
            const uint64_t used = uniform(s_used[q]);
            if (lane == 0) {
              gstore(s_toff[q] + ti, cpos);
              s_used[q] = used + T;
              s_cpos[q] = cpos + T;
              s_cend[q] = cend;
              if (!ok) s_short = 1u;
            }
            TQ[q] = T;
            reQ[q] = transpose64(Er, lane);
            off[q] = cpos - (gbase + tab);
            // the starts of output q: lane l owns boundary l, its S marks are
            // the rows b just above the runs of the strips starting there
            uint64_t m = Sr;
            while (__ballot(m != 0ull)) {
              if (tab + WAVE > (uint32_t)EROW) flush();
#pragma unroll
              for (int r = 0; r < Q; r++) {
                if (tab + WAVE > (uint32_t)EROW) break;
                const uint64_t act = __ballot(m != 0ull);
                if (!act) break;
                if (m) {
                  const uint32_t t = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                  m &= m - 1;
                  stab[tab + lanes_below(act)] = (uint16_t)(((uint32_t)lane << 8) | ((uint32_t)q << 6) | t);
                }
                tab += (uint32_t)__popcll(act);
              }
            }
          }
          if (tab) flush();
          if (lane == 0) {
#pragma unroll
            for (int q = 0; q < NQ; q++) {
              if (q >= nout) break;
              if (!((anymask >> q) & 1u)) continue;
              gstore(s_tn[q] + ti, TQ[q] - sideQ[q]);
              gstore(s_tn[q] + o.ntab + ti, sideQ[q]);
              s_side[q] += sideQ[q];
            }
          }
          __syncthreads();
        }
      }
#pragma unroll 1
      for (int q = 0; q < (shared_pass ? 0 : nout); q++) {
        uint32_t* const tn = uniform(s_tn[q]);
        const uint64_t* MO = MT;
        if (onepass) {
          MO = MTQ[min(q, NQ - 1)];
          if (!((anymask >> q) & 1u)) {
            if (lane == 0) {
              gstore(tn + ti, 0u);
              gstore(tn + o.ntab + ti, 0u);
            }
            continue;
          }
        } else {
          const uint32_t tt = uniform(s_tt[q]);
          uint64_t any = 0;
          for (int c = lane; c < ncol; c += WAVE) {
            const bool in = c >= colT0 && c < colT1;
#pragma unroll
            for (int w = 0; w < NW; w++) {
              if (w >= W) break;
              const int ix = c * W + w;
              const uint64_t v = in ? tt_eval_common(tt, MS[ix], (nmorph > 0) ? MM[0][ix] : 0ull,
                                                     (nmorph > 1) ? MM[1][ix] : 0ull) & row_bits(w, rowT0, rowT1)
                                    : 0ull;
              MT[ix] = v;
              any |= v;
            }
          }
          if (__ballot(any != 0ull) == 0ull) {
            if (lane == 0) {
              gstore(tn + ti, 0u);
              gstore(tn + o.ntab + ti, 0u);
            }
            continue;
          }
          __syncthreads();
        }
        // start / end marks at every slab boundary (lane owns boundaries
        // 64*kb + lane): the runs of the column right of the boundary (S) and
        // left of it (E) that meet the change mask, each marked at the clear
        // row just above the run
        uint64_t Sr[NW][NW], Er[NW][NW];
        uint32_t myT = 0;
#pragma unroll
        for (int kb = 0; kb < NW; kb++) {
          const int k = kb * WAVE + lane;
#pragma unroll
          for (int w = 0; w < NW; w++) { Sr[kb][w] = 0ull; Er[kb][w] = 0ull; }
          if (k <= ncol) {
            const uint64_t* cur = (k < ncol) ? MO + k * W : nullptr;
            const uint64_t* prv = (k > 0) ? MO + (k - 1) * W : nullptr;
            uint64_t D[NW];
            change_mask<NW>(cur, prv, W, D);
            ends_meeting<NW>(cur, W, D, Sr[kb]);
            ends_meeting<NW>(prv, W, D, Er[kb]);
#pragma unroll
            for (int w = 0; w < NW; w++) myT += (uint32_t)__popcll(Er[kb][w]);
          }
        }
        const uint32_t T = wave_sum(myT);
        // the tile's range: from this workgroup's chunk of the output's stream,
        // or a new chunk of max(chunk, T) strips
        uint64_t cpos = uniform(s_cpos[q]), cend = uniform(s_cend[q]);
        if (cpos + T > cend) {
          const uint64_t want = (uint64_t)T > o.chunk ? (uint64_t)T : o.chunk;
          unsigned long long nb = 0;
          if (lane == 0) nb = atomicAdd(pick_out(o.cnt, q) + 2, (unsigned long long)want);
          cpos = read_lane((uint64_t)nb, 0);
          cend = cpos + want;
        }
        const uint64_t used = uniform(s_used[q]);
        const uint64_t base = cpos;
        const bool ok = base + T <= pick_out(o.stream_cap, q);
        vl_box* const strips = uniform(s_strips[q]);
        uint8_t* const tags = uniform(s_tag[q]);
        if (lane == 0) gstore(s_toff[q] + ti, base);
        // strips.  Lane l owns boundary j = 64*kb + l; its S marks are the
        // rows b just above the runs [a, b) of strips starting there.  In
        // rounds, every lane appends up to Q of its starts (j, b) to a table
        // in LDS, then the whole wave takes the table's entries one per lane,
        // so a round costs ceil(entries / 64) passes however the starts are
        // spread over the boundaries.  A strip starting at (j, b) ends at the
        // first E mark after j in row b, read from the row-major (transposed)
        // end marks held by lane b & 63; its run starts one above the highest
        // clear row of column j below b.  Strips are written in table order.
        constexpr int Q = EROW / WAVE;    // table entries per lane and round
        uint16_t* const stab = erow;      // (j << 8) | b
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
            const uint32_t jl = (uint32_t)(kb * WAVE + lane);
            while (__ballot(m != 0ull)) {
              uint32_t nent = 0;
#pragma unroll
              for (int r = 0; r < Q; r++) {
                const uint64_t act = __ballot(m != 0ull);
                if (!act) break;
                if (m) {
                  const uint32_t t = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                  m &= m - 1;
                  stab[nent + lanes_below(act)] = (uint16_t)((jl << 8) | ((uint32_t)wb * WAVE + t));
                }
                nent += (uint32_t)__popcll(act);
              }
              __syncthreads();
              for (uint32_t g0 = 0; g0 < nent; g0 += WAVE) {
                const uint32_t g = g0 + (uint32_t)lane;
                const bool has = g < nent;
                const uint32_t ent = has ? (uint32_t)stab[g] : 0u;
                const int j = (int)(ent >> 8), b = (int)(ent & 0xffu);
                uint64_t rr[NW];
#pragma unroll
                for (int kq = 0; kq < NW; kq++) rr[kq] = lane_read64(re[kq], b & (WAVE - 1));
                uint32_t tg = 0u;
                if (has) {
                  const int k = next_bit_above<NW>(rr, j, ncol);
                  const int a = run_start_w<NW>(MO + j * W, W, b);
                  tg = ((a == rowT0 && (flags & SEAM_B)) ? SEAM_B : 0u) |
                       ((b == rowT1 && (flags & SEAM_T)) ? SEAM_T : 0u) |
                       ((j == colT0 && (flags & SEAM_L)) ? SEAM_L : 0u) |
                       ((k == colT1 && (flags & SEAM_R)) ? SEAM_R : 0u);
                  const uint64_t p = pos + g;
                  if (ok) {
                    gstore_box(strips + p, ux[j], uy[a], ux[k], uy[b]);
                    gstore(tags + p, (uint8_t)tg);
                  }
                }
                nside += (uint32_t)__popcll(__ballot(tg != 0u));
              }
              pos += nent;
              __syncthreads();
            }
          }
        }
        __syncthreads();
        if (lane == 0) {
          gstore(tn + ti, T - nside);
          gstore(tn + o.ntab + ti, nside);
          s_used[q] = used + T;
          s_side[q] += nside;
          s_cpos[q] = cpos + T;
          s_cend[q] = cend;
          if (!ok) s_short = 1u;
        }
        __syncthreads();
      }
      __syncthreads();
    }
    __syncthreads();
  }
  if (novf) {
    __syncthreads();
    flush_overflow(sovf, novf, o.ovf, o.ovf_n, lane, WAVE);
  }
  __syncthreads();
  if (lane < nout) {
    const uint64_t used = s_used[lane];
    if (used) {
      unsigned long long* cnt = nullptr;
#pragma unroll
      for (int k = 0; k < FUSE_MAX_OUT; k++)
        if (k == lane) cnt = o.cnt[k];
      atomicAdd(&cnt[0], (unsigned long long)used);
      if (s_side[lane]) atomicAdd(&cnt[1], (unsigned long long)s_side[lane]);
    }
  }
  if (lane == 0 && s_short) atomicOr(o.err, ERR_OVERFLOW_CAP);
}

template <int P>
void launch_f(const TileDesc* desc, int ntiles, const vl_box* entries, const FusedSpec& fs, const FusedOut& o,
              int wg, hipStream_t st) {
  const int g = wg < 1 ? 1 : wg;
  hipLaunchKernelGGL((k_tile_f<P>), dim3(g), dim3(WAVE), 0, st, desc, ntiles, entries, fs, o);
}

}  // namespace

void launch_tile_fused(int nmax, const TileDesc* desc, int ntiles, const vl_box* entries, const FusedSpec& fs,
                       const FusedOut& o, int wg, hipStream_t st) {
  if (nmax == FUSE_NMAX_TINY) launch_f<64>(desc, ntiles, entries, fs, o, wg, st);
  else if (nmax == FUSE_NMAX_SMALL) launch_f<128>(desc, ntiles, entries, fs, o, wg, st);
  else launch_f<256>(desc, ntiles, entries, fs, o, wg, st);
}

int tile_fused_max_wg(int nmax) {
  static int cache[3] = {0, 0, 0};
  const int i = nmax == FUSE_NMAX_TINY ? 0 : nmax == FUSE_NMAX_SMALL ? 1 : 2;
  int& c = cache[i];
  if (!c)
    c = i == 0 ? occupancy_wg(k_tile_f<64>, WAVE, FUSED_WAVES_PER_CU)
               : i == 1 ? occupancy_wg(k_tile_f<128>, WAVE, FUSED_WAVES_PER_CU)
                        : occupancy_wg(k_tile_f<256>, WAVE, FUSED_WAVES_PER_CU);
  return c;
}

}  // namespace vl
