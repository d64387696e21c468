// vl_snap.hip — decode of SNAP box-codec chunks on the GPU, straight into a
// layer's device buffer (the codec of src/box_codec.f90:447-536).
//
// A chunk is a byte stream of variable-length records, one per box: an info
// byte, then optional varints / int16s whose presence the info byte encodes.
// Box i depends on boxes before it through the x/y cursor (deltas), the
// width/height dictionaries (new values appended, others referenced by
// index) and the current dictionary entries.  None of that needs a
// sequential decode:
//   1. length: for every byte, the length a record starting there would have
//      (bounded look-ahead over at most four varints);
//   2. walk: the chunk is cut into 2 KB blocks; from each of the 41 possible
//      entry offsets of a block (no record is longer) one lane walks the
//      lengths to the block end and records the exit offset and the record
//      count — speculative, so all blocks walk at once;
//   3. resolve: per chunk, the true entry of each block follows from its
//      predecessor's exit (one short loop over the blocks), with the record
//      index each block starts at;
//   4. starts: each block walks once more from its true entry and writes the
//      byte offset of every record;
//   5. decode: a workgroup per chunk parses its records 256 at a time; block
//      scans give x1, y1 (wrapping int32 sums of the deltas), the dictionary
//      index of every new width/height (count of new values so far), and the
//      current entries (the last record that set one, carried forward); a
//      second pass reads the dictionaries and writes the boxes.
// Malformed streams (a record past the chunk end, a count that does not
// match, an index outside the dictionary) set an error flag.
#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

constexpr int SNAP_BW = 2048;   // bytes per walk block
constexpr int SNAP_E = 41;      // entry offsets per block (longest record + 1)
constexpr int SNAP_NT = 256;    // threads of the decode workgroup

unsigned grid_of(int64_t n, int per = 256) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (unsigned)g;
}

// index of the segment of a sorted offset table holding position p
__device__ __forceinline__ int seg_of(const int64_t* off, int n, int64_t p) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= p) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ int64_t skip_varint(const uint8_t* b, int64_t q, int64_t end) {
  for (int k = 0; k < 10; k++) {
    if (q >= end) return -1;
    if (!(b[q++] & 0x80u)) return q;
  }
  return -1;
}

// 1. record lengths (255: no record can start here)
__global__ void k_snap_len(const uint8_t* __restrict__ raw, int64_t total, const int64_t* __restrict__ coff,
                           int nchunks, uint8_t* __restrict__ len) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (int64_t)gridDim.x * blockDim.x) {
    const int c = seg_of(coff, nchunks, p);
    const int64_t end = coff[c + 1];
    const uint32_t info = raw[p];
    int64_t q = p + 1;
    if (q >= 0 && (info & 32u)) q = skip_varint(raw, q, end);
    if (q >= 0 && (info & 64u)) q = skip_varint(raw, q, end);
    if (q >= 0 && (info & 1u)) q = (info & 2u) ? skip_varint(raw, q, end) : q + 2;
    if (q >= 0 && !(info & 16u) && (info & 4u)) q = (info & 8u) ? skip_varint(raw, q, end) : q + 2;
    len[p] = (q >= 0 && q <= end && q - p < SNAP_E) ? (uint8_t)(q - p) : (uint8_t)255;
  }
}

// 2. speculative walks: block b (global index), entry e
__global__ void k_snap_walk(const uint8_t* __restrict__ len, const int64_t* __restrict__ coff,
                            const int64_t* __restrict__ boff, int nchunks, int64_t nblocks,
                            uint8_t* __restrict__ exit_e, uint16_t* __restrict__ cnt) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nblocks * SNAP_E;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / SNAP_E;
    const int e = (int)(t - b * SNAP_E);
    const int c = seg_of(boff, nchunks, b);
    const int64_t bs = coff[c] + (b - boff[c]) * SNAP_BW;
    const int64_t be = min(bs + SNAP_BW, coff[c + 1]);
    int64_t pos = bs + e;
    uint32_t n = 0;
    bool bad = false;
    while (pos < be) {
      const uint32_t l = len[pos];
      if (l == 255u) { bad = true; break; }
      pos += l;
      n++;
    }
    const int64_t ex = pos - be;
    exit_e[t] = (bad || ex >= SNAP_E) ? (uint8_t)255 : (uint8_t)ex;
    cnt[t] = (uint16_t)n;
  }
}

// 3. true block entries and record bases, one thread per chunk
__global__ void k_snap_resolve(const uint8_t* __restrict__ exit_e, const uint16_t* __restrict__ cnt,
                               const int64_t* __restrict__ boff, const int64_t* __restrict__ roff, int nchunks,
                               uint8_t* __restrict__ entry, int64_t* __restrict__ rbase, uint32_t* err) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
    uint32_t e = 0;
    int64_t r = roff[c];
    bool bad = false;
    for (int64_t b = boff[c]; b < boff[c + 1]; b++) {
      entry[b] = (uint8_t)e;
      rbase[b] = r;
      const uint32_t ex = exit_e[b * SNAP_E + e];
      r += cnt[b * SNAP_E + e];
      if (ex == 255u) { bad = true; break; }
      e = ex;
    }
    if (bad) atomicOr(err, 1u);
    else if (e != 0 || r != roff[c + 1]) atomicOr(err, 2u);   // the stream must end on the chunk's last record
  }
}

// 4. record starts, one thread per block
__global__ void k_snap_starts(const uint8_t* __restrict__ len, const int64_t* __restrict__ coff,
                              const int64_t* __restrict__ boff, int nchunks, int64_t nblocks,
                              const uint8_t* __restrict__ entry, const int64_t* __restrict__ rbase,
                              const int64_t* __restrict__ roff, int64_t* __restrict__ start, const uint32_t* err) {
  if (*err) return;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks; b += (int64_t)gridDim.x * blockDim.x) {
    const int c = seg_of(boff, nchunks, b);
    const int64_t bs = coff[c] + (b - boff[c]) * SNAP_BW;
    const int64_t be = min(bs + SNAP_BW, coff[c + 1]);
    int64_t pos = bs + entry[b], r = rbase[b];
    const int64_t rend = roff[c + 1];
    while (pos < be && r < rend) {
      start[r++] = pos;
      pos += len[pos];
    }
  }
}

__device__ __forceinline__ uint64_t read_varint(const uint8_t* b, int64_t& q) {
  uint64_t v = 0;
  for (int sh = 0; sh < 70; sh += 7) {
    const uint32_t x = b[q++];
    v |= (uint64_t)(x & 127u) << sh;
    if (!(x & 128u)) break;
  }
  return v;
}

__device__ __forceinline__ int32_t zigzag_decode(uint64_t v) { return (int32_t)(uint32_t)((v >> 1) ^ (0 - (v & 1))); }

// block scans (SNAP_NT threads): inclusive sum (wrapping uint32) and forward
// fill (the last non-zero value at or before each thread)
__device__ __forceinline__ uint32_t block_incl_sum(uint32_t v, uint32_t* ws, uint32_t* total) {
  const int lane = __lane_id(), w = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)v, o, WAVE);
    if (lane >= o) v += u;
  }
  if (lane == WAVE - 1) ws[w] = v;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int k = 0; k < SNAP_NT / WAVE; k++) {
    if (k < w) pre += ws[k];
    tot += ws[k];
  }
  __syncthreads();
  *total = tot;
  return v + pre;
}
__device__ __forceinline__ uint32_t block_fill(uint32_t v, uint32_t* ws, uint32_t carry, uint32_t* last) {
  const int lane = __lane_id(), w = threadIdx.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)v, o, WAVE);
    if (lane >= o && v == 0u) v = u;
  }
  if (lane == WAVE - 1) ws[w] = v;
  __syncthreads();
  uint32_t pre = carry;
  for (int k = 0; k < w; k++)
    if (ws[k]) pre = ws[k];
  uint32_t lst = carry;
  for (int k = 0; k < SNAP_NT / WAVE; k++)
    if (ws[k]) lst = ws[k];
  __syncthreads();
  *last = lst;
  return v ? v : pre;
}

// 5. decode: a workgroup per chunk
__global__ __launch_bounds__(SNAP_NT) void k_snap_decode(const uint8_t* __restrict__ raw,
                                                         const int64_t* __restrict__ start,
                                                         const int64_t* __restrict__ roff, int32_t* __restrict__ dictw,
                                                         int32_t* __restrict__ dicth, int32_t* __restrict__ rx,
                                                         int32_t* __restrict__ ry, uint32_t* __restrict__ widx,
                                                         uint32_t* __restrict__ hidx, vl_box* __restrict__ out,
                                                         uint32_t* err) {
  __shared__ uint32_t ws[SNAP_NT / WAVE];
  __shared__ uint32_t s_err, s_bad;
  const int c = blockIdx.x;
  const int64_t r0 = roff[c], r1 = roff[c + 1];
  // the record starts exist only when the earlier stages found no error
  // (read once and shared: the decision must be the same for the whole block)
  if (threadIdx.x == 0) {
    s_err = *err & 3u;
    s_bad = 0u;
  }
  __syncthreads();
  if (s_err) return;
  uint32_t cx = 0, cy = 0, nw = 0, nh = 0, cw = 0, ch = 0;   // carries across tiles
  for (int64_t t0 = r0; t0 < r1; t0 += SNAP_NT) {
    const int64_t r = t0 + threadIdx.x;
    const bool live = r < r1;
    uint32_t dx = 0, dy = 0, newW = 0, newH = 0, setW = 0, setH = 0;
    int32_t vw = 0, vh = 0;
    bool sq = false, bad = false;
    if (live) {
      int64_t q = start[r];
      const uint32_t info = raw[q++];
      sq = (info & 16u) != 0u;
      if (info & 32u) dx = (uint32_t)zigzag_decode(read_varint(raw, q));
      if (info & 64u) dy = (uint32_t)zigzag_decode(read_varint(raw, q));
      // an int16 index is 1-based: 0 (or a negative value) is malformed
      if (info & 1u) {
        if (info & 2u) { vw = (int32_t)(uint32_t)read_varint(raw, q); newW = 1; }
        else { setW = (uint32_t)(uint16_t)(raw[q] | (raw[q + 1] << 8)); q += 2; if (!setW || setW > 32767u) bad = true; }
      }
      if (!sq && (info & 4u)) {
        if (info & 8u) { vh = (int32_t)(uint32_t)read_varint(raw, q); newH = 1; }
        else { setH = (uint32_t)(uint16_t)(raw[q] | (raw[q + 1] << 8)); q += 2; if (!setH || setH > 32767u) bad = true; }
      }
    }
    uint32_t tot;
    const uint32_t x = block_incl_sum(dx, ws, &tot) + cx;
    cx += tot;
    const uint32_t y = block_incl_sum(dy, ws, &tot) + cy;
    cy += tot;
    const uint32_t kw = block_incl_sum(newW, ws, &tot) + nw;   // dictionary size after this record
    nw += tot;
    const uint32_t kh = block_incl_sum(newH, ws, &tot) + nh;
    nh += tot;
    if (newW) { dictw[r0 + kw - 1] = vw; setW = kw; }
    if (newH) { dicth[r0 + kh - 1] = vh; setH = kh; }
    // an index set by an int16 must name an entry that exists by then
    if (live && (bad || (setW && setW > kw) || (setH && setH > kh))) {
      atomicOr(err, 4u);
      s_bad = 1u;
    }
    uint32_t lastW, lastH;
    const uint32_t curW = block_fill(setW, ws, cw, &lastW);
    cw = lastW;
    const uint32_t curH = block_fill(setH, ws, ch, &lastH);
    ch = lastH;
    if (live) {
      rx[r] = (int32_t)x;
      ry[r] = (int32_t)y;
      widx[r] = curW;
      hidx[r] = sq ? 0xFFFFFFFFu : curH;
      if (curW == 0u || (!sq && curH == 0u)) atomicOr(err, 8u);
    }
  }
  __threadfence_block();
  __syncthreads();
  // an index past the dictionary would read outside this chunk's entries
  if (s_bad) return;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += SNAP_NT) {
    const uint32_t iw = widx[r], ih = hidx[r];
    if (iw == 0u) continue;
    const int32_t w = dictw[r0 + iw - 1];
    const int32_t h = ih == 0xFFFFFFFFu ? w : (ih ? dicth[r0 + ih - 1] : 0);
    vl_box bx;
    bx.x1 = rx[r];
    bx.y1 = ry[r];
    bx.x2 = (int32_t)((uint32_t)bx.x1 + (uint32_t)w);
    bx.y2 = (int32_t)((uint32_t)bx.y1 + (uint32_t)h);
    out[r] = bx;
  }
}

}  // namespace

int64_t snap_blocks(int64_t raw_bytes) { return (raw_bytes + SNAP_BW - 1) / SNAP_BW; }

void launch_snap_decode(const SnapDecode& d, hipStream_t st) {
  if (d.total_raw <= 0 || d.nchunks <= 0) return;
  hipLaunchKernelGGL(k_snap_len, dim3(grid_of(d.total_raw)), dim3(256), 0, st, d.raw, d.total_raw, d.coff,
                     d.nchunks, d.len);
  hipLaunchKernelGGL(k_snap_walk, dim3(grid_of(d.nblocks * SNAP_E)), dim3(256), 0, st, d.len, d.coff, d.boff,
                     d.nchunks, d.nblocks, d.exit_e, d.cnt);
  hipLaunchKernelGGL(k_snap_resolve, dim3(grid_of(d.nchunks, 64)), dim3(64), 0, st, d.exit_e, d.cnt, d.boff, d.roff,
                     d.nchunks, d.entry, d.rbase, d.err);
  hipLaunchKernelGGL(k_snap_starts, dim3(grid_of(d.nblocks, 64)), dim3(64), 0, st, d.len, d.coff, d.boff,
                     d.nchunks, d.nblocks, d.entry, d.rbase, d.roff, d.start, d.err);
  hipLaunchKernelGGL(k_snap_decode, dim3((unsigned)d.nchunks), dim3(SNAP_NT), 0, st, d.raw, d.start, d.roff,
                     d.dictw, d.dicth, d.rx, d.ry, d.widx, d.hidx, d.out, d.err);
}

}  // namespace vl
