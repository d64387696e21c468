// vl_prims.hip — fills, bounding-box reduction and device-wide exclusive scans.
#include <algorithm>

#include "vl_internal.h"
#include "vl_dev.h"

namespace vl {

namespace {

constexpr int SCAN_NT = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_NT * SCAN_ITEMS;

template <typename T>
__global__ void k_fill(T* p, T v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Sets *flag when any of the n dwords at p differs from the expected value
// (graph replays: a size the captured launches were planned with).
__global__ void k_expect(const uint32_t* __restrict__ p, Expect8 ex, int n, uint32_t* flag) {
  const int i = threadIdx.x;
  if (i < n && p[i] != ex.w[i]) atomicOr(flag, 1u);
}

// Bounding box of the transformed boxes; flags invalid input or output boxes.
// sums (may be null): += the boxes' summed widths and heights (sums[0], [1]).
__global__ void k_mbr(const vl_box* __restrict__ b, int64_t n, Xform xf, int32_t* mbr, uint32_t* err,
                      unsigned long long* sums) {
  __shared__ int32_t s[4][SCAN_NT / WAVE];
  __shared__ unsigned long long sw[2][SCAN_NT / WAVE];
  int32_t x1 = INT32_MAX, y1 = INT32_MAX, x2 = INT32_MIN, y2 = INT32_MIN;
  unsigned long long tw = 0, th = 0;
  uint32_t bad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    vl_box v = b[i];
    bad |= (v.x1 >= v.x2 || v.y1 >= v.y2) ? 1u : 0u;
    int64_t a1 = (int64_t)v.x1 - xf.w, a2 = (int64_t)v.x2 + xf.e;
    int64_t c1 = (int64_t)v.y1 - xf.s, c2 = (int64_t)v.y2 + xf.n;
    bad |= (a1 >= a2 || c1 >= c2 || a1 < INT32_MIN || a2 > INT32_MAX || c1 < INT32_MIN ||
            c2 > INT32_MAX) ? 1u : 0u;
    x1 = min(x1, (int32_t)max(a1, (int64_t)INT32_MIN));
    y1 = min(y1, (int32_t)max(c1, (int64_t)INT32_MIN));
    x2 = max(x2, (int32_t)min(a2, (int64_t)INT32_MAX));
    y2 = max(y2, (int32_t)min(c2, (int64_t)INT32_MAX));
    tw += a2 > a1 ? (unsigned long long)(a2 - a1) : 0ull;
    th += c2 > c1 ? (unsigned long long)(c2 - c1) : 0ull;
  }
  tw = wave_sum(tw);
  th = wave_sum(th);
  for (int d = WAVE / 2; d > 0; d >>= 1) {
    x1 = min(x1, __shfl_xor(x1, d, WAVE));
    y1 = min(y1, __shfl_xor(y1, d, WAVE));
    x2 = max(x2, __shfl_xor(x2, d, WAVE));
    y2 = max(y2, __shfl_xor(y2, d, WAVE));
  }
  if (__any(bad)) {
    if (__lane_id() == 0) atomicOr(err, ERR_INVALID_BOX);
  }
  const int w = threadIdx.x / WAVE;
  if (__lane_id() == 0) {
    s[0][w] = x1; s[1][w] = y1; s[2][w] = x2; s[3][w] = y2;
    sw[0][w] = tw; sw[1][w] = th;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x / WAVE); i++) {
      x1 = min(x1, s[0][i]); y1 = min(y1, s[1][i]); x2 = max(x2, s[2][i]); y2 = max(y2, s[3][i]);
      tw += sw[0][i];
      th += sw[1][i];
    }
    atomicMin(&mbr[0], x1); atomicMin(&mbr[1], y1);
    atomicMax(&mbr[2], x2); atomicMax(&mbr[3], y2);
    if (sums) {
      atomicAdd(&sums[0], tw);
      atomicAdd(&sums[1], th);
    }
  }
}

// ---- reduce-then-scan ------------------------------------------------------
template <typename T, typename Load>
__global__ void k_scan_reduce(Load ld, size_t n, T* sums) {
  __shared__ T ws[SCAN_NT / WAVE];
  size_t base = (size_t)blockIdx.x * SCAN_TILE;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    size_t i = base + (size_t)k * SCAN_NT + threadIdx.x;
    if (i < n) acc += ld(i);
  }
  acc = wave_sum(acc);
  if (__lane_id() == 0) ws[threadIdx.x / WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int i = 0; i < SCAN_NT / WAVE; i++) t += ws[i];
    sums[blockIdx.x] = t;
  }
}

// Scans one tile of SCAN_TILE elements (blocked per thread) and adds prefix[blockIdx.x].
template <typename T, typename Load>
__global__ void k_scan_tile(Load ld, size_t n, T* out, const T* prefix, T* total) {
  __shared__ T ws[SCAN_NT / WAVE];
  size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T acc = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    size_t i = base + k;
    v[k] = (i < n) ? ld(i) : (T)0;
    acc += v[k];
  }
  T tot;
  T ex = block_excl_scan<T, SCAN_NT>(acc, ws, &tot);
  T p = prefix ? prefix[blockIdx.x] : (T)0;
  T run = p + ex;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    size_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_NT - 1) *total = run;
}

template <typename T>
struct LoadArr {
  const T* p;
  __device__ T operator()(size_t i) const { return p[i]; }
};
struct LoadPair {
  const uint32_t* a;
  const uint32_t* b;
  __device__ uint64_t operator()(size_t i) const { return (uint64_t)a[i] + (uint64_t)b[i]; }
};
struct LoadWiden {
  const uint32_t* a;
  __device__ uint64_t operator()(size_t i) const { return (uint64_t)a[i]; }
};

size_t nblk(size_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

template <typename T, typename Load>
void scan_impl(Load ld, size_t n, T* out, T* total, char* tmp, hipStream_t st) {
  if (n == 0) {
    if (total) hipLaunchKernelGGL((k_fill<T>), dim3(1), dim3(64), 0, st, total, (T)0, (size_t)1);
    return;
  }
  size_t nb = nblk(n);
  if (nb == 1) {
    hipLaunchKernelGGL((k_scan_tile<T, Load>), dim3(1), dim3(SCAN_NT), 0, st, ld, n, out,
                       (const T*)nullptr, total);
    return;
  }
  T* sums = (T*)tmp;
  T* sums_sc = sums + nb;
  char* rest = (char*)(sums_sc + nb);
  hipLaunchKernelGGL((k_scan_reduce<T, Load>), dim3(nb), dim3(SCAN_NT), 0, st, ld, n, sums);
  scan_impl<T, LoadArr<T>>(LoadArr<T>{sums}, nb, sums_sc, (T*)nullptr, rest, st);
  hipLaunchKernelGGL((k_scan_tile<T, Load>), dim3(nb), dim3(SCAN_NT), 0, st, ld, n, out,
                     (const T*)sums_sc, total);
}

}  // namespace

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st) {
  if (!n) return;
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL((k_fill<uint32_t>), dim3(g), dim3(256), 0, st, p, v, n);
}
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t st) {
  if (!n) return;
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL((k_fill<uint64_t>), dim3(g), dim3(256), 0, st, p, v, n);
}

// rows x cols sub-boxes of the domain (x_lo, y_lo, x_hi, y_hi), row-major
// from the bottom-left: base width / height plus one unit for the first
// (extent mod count) columns / rows, padded by overlap/2 to the low side and
// overlap - overlap/2 to the high side, clamped to the domain
// (CreateGrid, src/design_impl.f90:159-275).
__global__ void k_grid(int64_t x_lo, int64_t y_lo, int64_t x_hi, int64_t y_hi, int rows, int cols, int64_t overlap,
                       vl_box* out) {
  const int64_t n = (int64_t)rows * cols;
  const int64_t w = x_hi - x_lo, h = y_hi - y_lo;
  const int64_t cw = w / cols, ew = w % cols, rh = h / rows, eh = h % rows;
  const int64_t pl = overlap / 2, pr = overlap - pl;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / cols, j = k % cols;   // row i, column j (0-based)
    const int64_t y0 = y_lo + i * rh + min(i, eh), y1 = y_lo + (i + 1) * rh + min(i + 1, eh);
    const int64_t x0 = x_lo + j * cw + min(j, ew), x1 = x_lo + (j + 1) * cw + min(j + 1, ew);
    vl_box b;
    b.x1 = (int32_t)max(x_lo, x0 - pl);
    b.x2 = (int32_t)min(x_hi, x1 + pr);
    b.y1 = (int32_t)max(y_lo, y0 - pl);
    b.y2 = (int32_t)min(y_hi, y1 + pr);
    out[k] = b;
  }
}

void launch_grid(int64_t x_lo, int64_t y_lo, int64_t x_hi, int64_t y_hi, int rows, int cols, int64_t overlap,
                 vl_box* out, hipStream_t st) {
  const int64_t n = (int64_t)rows * cols;
  if (n <= 0) return;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_grid, dim3((unsigned)g), dim3(256), 0, st, x_lo, y_lo, x_hi, y_hi, rows, cols, overlap, out);
}

void launch_expect(const void* dev, const void* expected, size_t bytes, uint32_t* flag, hipStream_t st) {
  const uint32_t* p = (const uint32_t*)dev;
  const uint32_t* x = (const uint32_t*)expected;
  for (size_t o = 0; o < bytes / 4; o += 8) {
    Expect8 ex{};
    const int n = (int)std::min<size_t>(8, bytes / 4 - o);
    for (int i = 0; i < n; i++) ex.w[i] = x[o + i];
    hipLaunchKernelGGL(k_expect, dim3(1), dim3(64), 0, st, p + o, ex, n, flag);
  }
}

void launch_mbr(const vl_box* b, int64_t n, Xform xf, int32_t* mbr, uint32_t* err, hipStream_t st,
                unsigned long long* sums) {
  if (n <= 0) return;
  int64_t g = (n + SCAN_NT - 1) / SCAN_NT;
  if (g > 1024) g = 1024;   // one set of 4 atomics per workgroup on shared addresses
  hipLaunchKernelGGL(k_mbr, dim3((unsigned)g), dim3(SCAN_NT), 0, st, b, n, xf, mbr, err, sums);
}

size_t scan_tmp_bytes(size_t n) {
  size_t bytes = 0;
  size_t nb = nblk(n);
  while (nb > 1) {
    bytes += 2 * nb * sizeof(uint64_t);
    nb = nblk(nb);
  }
  return bytes + 256;
}

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, void* tmp,
                        hipStream_t st) {
  scan_impl<uint32_t, LoadArr<uint32_t>>(LoadArr<uint32_t>{in}, n, out, total, (char*)tmp, st);
}
void scan_exclusive_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* total, void* tmp,
                        hipStream_t st) {
  scan_impl<uint64_t, LoadArr<uint64_t>>(LoadArr<uint64_t>{in}, n, out, total, (char*)tmp, st);
}
void scan_pair_counts(const uint32_t* cntA, const uint32_t* cntB, uint64_t* off, size_t n,
                      uint64_t* total, void* tmp, hipStream_t st) {
  scan_impl<uint64_t, LoadPair>(LoadPair{cntA, cntB}, n, off, total, (char*)tmp, st);
}
void scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, uint64_t* total, void* tmp,
                     hipStream_t st) {
  scan_impl<uint64_t, LoadWiden>(LoadWiden{in}, n, out, total, (char*)tmp, st);
}

}  // namespace vl
