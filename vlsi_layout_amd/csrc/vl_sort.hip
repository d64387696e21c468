// vl_sort.hip — stable LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// Digits of 8 or 11 bits (an engine parameter the autotuner sets: 11-bit
// digits take fewer passes over the keys, 8-bit ones less LDS per workgroup).
// One pre-pass reads the keys once and builds the histograms of all digit
// positions; passes whose digit is the same for every key are skipped
// (coordinates that use only the low bits of int32 leave the high digits
// constant).  Each executed pass is reduce-then-scan:
//   upsweep   per-workgroup digit histogram of a 4096-key tile (LDS atomics)
//   scan      exclusive scan of the digit-major [bins x tiles] histogram
//   downsweep stable local ranking with wave64 ballots (one ballot per digit
//             bit gives the lanes that share a digit; popcount of the lower
//             peers is the rank), per-wave digit counters in LDS, a block
//             scan over digits, a local reorder in LDS, then a coalesced
//             scatter: consecutive LDS slots of one digit go to consecutive
//             global addresses.
// Key generation and payload gathers for the sort entry points live here too.
#include <vector>

#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

constexpr int RS_NT = 256;
constexpr int RS_NW = RS_NT / WAVE;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_NT * RS_ITEMS;

template <int DB>
struct Digits {
  static constexpr int BINS = 1 << DB;
  static constexpr int BPT = BINS / RS_NT;          // digits per thread
  static constexpr int POS = (64 + DB - 1) / DB;    // digit positions of a 64-bit key
  static constexpr uint32_t MASK = (uint32_t)BINS - 1u;
  static_assert(BINS % RS_NT == 0, "whole digits per thread");
};

template <int DB>
__global__ void k_rs_hist_all(const uint64_t* __restrict__ keys, int64_t n, uint32_t* ghist) {
  using D = Digits<DB>;
  __shared__ uint32_t h[D::POS * D::BINS];
  for (int i = threadIdx.x; i < D::POS * D::BINS; i += RS_NT) h[i] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * RS_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RS_NT) {
    uint64_t k = keys[i];
#pragma unroll
    for (int p = 0; p < D::POS; p++) atomicAdd(&h[p * D::BINS + ((k >> (DB * p)) & D::MASK)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < D::POS * D::BINS; i += RS_NT)
    if (h[i]) atomicAdd(&ghist[i], h[i]);
}

template <typename K, int DB>
__global__ __launch_bounds__(RS_NT) void k_rs_upsweep(const K* __restrict__ keys, int64_t n,
                                                       int shift, uint32_t* bhist, int nb) {
  using D = Digits<DB>;
  __shared__ uint32_t h[D::BINS];
  for (int d = threadIdx.x; d < D::BINS; d += RS_NT) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int k = 0; k < RS_ITEMS; k++) {
    int64_t i = base + (int64_t)k * RS_NT + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & D::MASK], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < D::BINS; d += RS_NT) bhist[(size_t)d * nb + blockIdx.x] = h[d];
}

template <typename K, int DB>
__global__ __launch_bounds__(RS_NT) void k_rs_downsweep(const K* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin,
                                                         K* __restrict__ kout,
                                                         uint32_t* __restrict__ vout, int64_t n,
                                                         int shift, const uint32_t* __restrict__ goff,
                                                         int nb) {
  using D = Digits<DB>;
  __shared__ K sk[RS_TILE];
  __shared__ uint32_t sv[RS_TILE];
  __shared__ uint32_t wcnt[RS_NW][D::BINS];
  __shared__ uint32_t dstart[D::BINS];
  __shared__ uint32_t gbase[D::BINS];
  __shared__ uint32_t ws[RS_NW];
  const int w = threadIdx.x / WAVE;
  const int lane = __lane_id();
  for (int i = threadIdx.x; i < RS_NW * D::BINS; i += RS_NT) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  K k[RS_ITEMS];
  uint32_t v[RS_ITEMS];
  uint32_t r[RS_ITEMS];
#pragma unroll
  for (int it = 0; it < RS_ITEMS; it++) {
    int64_t idx = base + (int64_t)w * (WAVE * RS_ITEMS) + it * WAVE + lane;
    bool valid = idx < n;
    k[it] = valid ? kin[idx] : (K)~(K)0;
    v[it] = valid ? vin[idx] : 0u;
  }
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; it++) {
    const uint32_t d = (uint32_t)(k[it] >> shift) & D::MASK;
    uint64_t peers = ~0ull;
#pragma unroll
    for (int b = 0; b < DB; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int leader = __ffsll((unsigned long long)peers) - 1;
    uint32_t old = 0;
    if (lane == leader) {
      old = wcnt[w][d];
      wcnt[w][d] = old + (uint32_t)__popcll(peers);
    }
    old = __shfl(old, leader, WAVE);   // per-lane source (each digit group has its own leader)
    r[it] = old + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();
  {
    // thread t owns digits t*BPT .. t*BPT+BPT-1: per-wave offsets, then a
    // block scan of the threads' totals
    uint32_t run[D::BPT];
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < D::BPT; j++) {
      const int d = threadIdx.x * D::BPT + j;
      uint32_t acc = 0;
#pragma unroll
      for (int ww = 0; ww < RS_NW; ww++) {
        const uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = acc;
        acc += c;
      }
      run[j] = acc;
      tsum += acc;
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan<uint32_t, RS_NT>(tsum, ws, &tot);
#pragma unroll
    for (int j = 0; j < D::BPT; j++) {
      const int d = threadIdx.x * D::BPT + j;
      dstart[d] = ex;
      ex += run[j];
      gbase[d] = goff[(size_t)d * nb + blockIdx.x];
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < RS_ITEMS; it++) {
    const uint32_t d = (uint32_t)(k[it] >> shift) & D::MASK;
    const uint32_t pos = dstart[d] + wcnt[w][d] + r[it];
    sk[pos] = k[it];
    sv[pos] = v[it];
  }
  __syncthreads();
  const int64_t rem = n - base;
  const int nvalid = rem < RS_TILE ? (int)rem : RS_TILE;
  for (int i = threadIdx.x; i < nvalid; i += RS_NT) {
    const K key = sk[i];
    const uint32_t d = (uint32_t)(key >> shift) & D::MASK;
    const uint32_t g = gbase[d] + ((uint32_t)i - dstart[d]);
    kout[g] = key;
    vout[g] = sv[i];
  }
}

template <typename K>
__global__ void k_copy_pairs(const K* ks, const uint32_t* vs, K* kd, uint32_t* vd, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    kd[i] = ks[i];
    vd[i] = vs[i];
  }
}

// ---- keys and gathers ------------------------------------------------------
__device__ __forceinline__ uint64_t spread32(uint32_t c) {
  uint64_t m = c;
  m = (m | (m << 16)) & 0x0000FFFF0000FFFFull;
  m = (m | (m << 8)) & 0x00FF00FF00FF00FFull;
  m = (m | (m << 4)) & 0x0F0F0F0F0F0F0F0Full;
  m = (m | (m << 2)) & 0x3333333333333333ull;
  m = (m | (m << 1)) & 0x5555555555555555ull;
  return m;
}

// Tracker records are 12 bytes (int32 X, int32 Y, int8 + padding).
__global__ void k_tracker_keys(const int32_t* __restrict__ t, uint64_t* keys, uint32_t* vals, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t ux = (uint32_t)t[3 * i] ^ 0x80000000u;
    uint32_t uy = (uint32_t)t[3 * i + 1] ^ 0x80000000u;
    keys[i] = ((uint64_t)ux << 32) | uy;
    vals[i] = (uint32_t)i;
  }
}
__global__ void k_tracker_gather(const int32_t* __restrict__ in, const uint32_t* __restrict__ perm,
                                 int32_t* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t s = perm[i];
    out[3 * i] = in[3 * s];
    out[3 * i + 1] = in[3 * s + 1];
    out[3 * i + 2] = in[3 * s + 2];
  }
}
// Morton code of the box centre (centre computed in 64 bits, truncating division).
__global__ void k_morton_keys(const vl_box* __restrict__ b, uint64_t* keys, uint32_t* vals, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    vl_box v = b[i];
    uint32_t cx = (uint32_t)(((int64_t)v.x1 + v.x2) / 2) ^ 0x80000000u;
    uint32_t cy = (uint32_t)(((int64_t)v.y1 + v.y2) / 2) ^ 0x80000000u;
    keys[i] = (spread32(cy) << 1) | spread32(cx);
    vals[i] = (uint32_t)i;
  }
}
// (x1, y1) as an order-preserving u64: sorts strips by lower-left corner.
__global__ void k_corner_keys(const vl_box* __restrict__ b, uint64_t* keys, uint32_t* vals, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    vl_box v = b[i];
    keys[i] = ((uint64_t)((uint32_t)v.x1 ^ 0x80000000u) << 32) | ((uint32_t)v.y1 ^ 0x80000000u);
    vals[i] = (uint32_t)i;
  }
}
__global__ void k_box_gather(const vl_box* __restrict__ in, const uint32_t* __restrict__ perm,
                             vl_box* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[perm[i]];
}
__global__ void k_perm_index1(const uint32_t* __restrict__ perm, int64_t* idx, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    idx[i] = (int64_t)perm[i] + 1;
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

}  // namespace

RadixPlan radix_plan(int64_t n, int digit_bits) {
  RadixPlan p;
  p.n = n;
  p.nblocks = (int)((n + RS_TILE - 1) / RS_TILE);
  const size_t bins = digit_bits == 11 ? 2048 : 256;
  const size_t pos = digit_bits == 11 ? 6 : 8;
  size_t hb = bins * (size_t)(p.nblocks > 0 ? p.nblocks : 1);
  p.tmp_bytes = pos * bins * sizeof(uint32_t) + 2 * hb * sizeof(uint32_t) + scan_tmp_bytes(hb) + 1024;
  return p;
}

namespace {

template <typename K, int DB>
void rs_pass(const K* ks, const uint32_t* vs, K* kd, uint32_t* vd, int64_t n, int shift, uint32_t* bhist,
             uint32_t* goff, char* stmp, int nb, hipStream_t st) {
  const size_t hb = (size_t)Digits<DB>::BINS * nb;
  hipLaunchKernelGGL((k_rs_upsweep<K, DB>), dim3(nb), dim3(RS_NT), 0, st, ks, n, shift, bhist, nb);
  scan_exclusive_u32(bhist, goff, hb, nullptr, stmp, st);
  hipLaunchKernelGGL((k_rs_downsweep<K, DB>), dim3(nb), dim3(RS_NT), 0, st, ks, vs, kd, vd, n, shift, goff, nb);
}

template <int DB>
void sort_pairs_hist(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n, void* tmp,
                     hipStream_t st) {
  using D = Digits<DB>;
  RadixPlan p = radix_plan(n, DB);
  const int nb = p.nblocks;
  const size_t hb = (size_t)D::BINS * nb;
  uint32_t* ghist = (uint32_t*)tmp;
  uint32_t* bhist = ghist + D::POS * D::BINS;
  uint32_t* goff = bhist + hb;
  char* stmp = (char*)(goff + hb);
  stmp = (char*)(((uintptr_t)stmp + 255) & ~(uintptr_t)255);
  launch_fill_u32(ghist, 0u, (size_t)D::POS * D::BINS, st);
  unsigned hg = (unsigned)((n + RS_NT * 16 - 1) / (RS_NT * 16));
  if (hg > 2048) hg = 2048;
  hipLaunchKernelGGL(k_rs_hist_all<DB>, dim3(hg), dim3(RS_NT), 0, st, keys, n, ghist);
  std::vector<uint32_t> h((size_t)D::POS * D::BINS);
  (void)hipMemcpyAsync(h.data(), ghist, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
  (void)hipStreamSynchronize(st);
  uint64_t* ks = keys;
  uint32_t* vs = vals;
  uint64_t* kd = keys_alt;
  uint32_t* vd = vals_alt;
  int done = 0;
  for (int pass = 0; pass < D::POS; pass++) {
    bool trivial = false;
    for (int d = 0; d < D::BINS; d++)
      if ((int64_t)h[(size_t)pass * D::BINS + d] == n) trivial = true;
    if (trivial) continue;
    rs_pass<uint64_t, DB>(ks, vs, kd, vd, n, DB * pass, bhist, goff, stmp, nb, st);
    uint64_t* tk = ks; ks = kd; kd = tk;
    uint32_t* tv = vs; vs = vd; vd = tv;
    done++;
  }
  if (done & 1)
    hipLaunchKernelGGL(k_copy_pairs<uint64_t>, dim3(grid_for(n)), dim3(256), 0, st, ks, vs, keys, vals, n);
}

template <typename K, int DB>
bool sort_bits(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n, int key_bits, void* tmp,
               hipStream_t st) {
  if (n <= 1 || key_bits <= 0) return false;
  using D = Digits<DB>;
  RadixPlan p = radix_plan(n, DB);
  const int nb = p.nblocks;
  const size_t hb = (size_t)D::BINS * nb;
  uint32_t* bhist = (uint32_t*)tmp + D::POS * D::BINS;
  uint32_t* goff = bhist + hb;
  char* stmp = (char*)(goff + hb);
  stmp = (char*)(((uintptr_t)stmp + 255) & ~(uintptr_t)255);
  K* ks = keys;
  uint32_t* vs = vals;
  K* kd = keys_alt;
  uint32_t* vd = vals_alt;
  bool alt = false;
  for (int shift = 0; shift < key_bits && shift < (int)(8 * sizeof(K)); shift += DB) {
    rs_pass<K, DB>(ks, vs, kd, vd, n, shift, bhist, goff, stmp, nb, st);
    K* tk = ks; ks = kd; kd = tk;
    uint32_t* tv = vs; vs = vd; vd = tv;
    alt = !alt;
  }
  return alt;
}

}  // namespace

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      int64_t n, void* tmp, hipStream_t st, int digit_bits) {
  if (n <= 1) return;
  if (digit_bits == 11) sort_pairs_hist<11>(keys, vals, keys_alt, vals_alt, n, tmp, st);
  else sort_pairs_hist<8>(keys, vals, keys_alt, vals_alt, n, tmp, st);
}

bool radix_sort_pairs_bits(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n,
                           int key_bits, void* tmp, hipStream_t st, int digit_bits) {
  if (digit_bits == 11) return sort_bits<uint64_t, 11>(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, st);
  return sort_bits<uint64_t, 8>(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, st);
}

bool radix_sort_pairs_bits32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n,
                             int key_bits, void* tmp, hipStream_t st, int digit_bits) {
  if (digit_bits == 11) return sort_bits<uint32_t, 11>(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, st);
  return sort_bits<uint32_t, 8>(keys, vals, keys_alt, vals_alt, n, key_bits, tmp, st);
}

void launch_tracker_keys(const vl_tracker* t, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_tracker_keys, dim3(grid_for(n)), dim3(256), 0, st, (const int32_t*)t, keys, vals, n);
}
void launch_tracker_gather(const vl_tracker* in, const uint32_t* perm, vl_tracker* out, int64_t n,
                           hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_tracker_gather, dim3(grid_for(n)), dim3(256), 0, st, (const int32_t*)in, perm,
                       (int32_t*)out, n);
}
void launch_morton_keys(const vl_box* b, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_morton_keys, dim3(grid_for(n)), dim3(256), 0, st, b, keys, vals, n);
}
void launch_corner_keys(const vl_box* b, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_corner_keys, dim3(grid_for(n)), dim3(256), 0, st, b, keys, vals, n);
}
void launch_box_gather(const vl_box* in, const uint32_t* perm, vl_box* out, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_box_gather, dim3(grid_for(n)), dim3(256), 0, st, in, perm, out, n);
}
void launch_perm_to_index1(const uint32_t* perm, int64_t* idx, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_perm_index1, dim3(grid_for(n)), dim3(256), 0, st, perm, idx, n);
}

}  // namespace vl
