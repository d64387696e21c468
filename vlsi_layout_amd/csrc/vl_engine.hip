// vl_engine.hip — host orchestration of the gfx950 Boolean engine and the C ABI
// declared in include/vl_api.h.
//
// One operator (plain form) runs as:
//   bounding boxes (one reduction kernel per input, validating every box)
//   -> uniform tile grid over the op's domain (host arithmetic)
//   -> binning: count, scan, scatter (boxes transformed on the fly for GROW)
//   -> tile kernel (level 0) -> final strips + seam side stream
//   -> seam levels: side strips touching internal row borders are re-tiled on
//      a grid with `level_factor` times taller rows until none remain; strips
//      only touching column borders are collected for the x-stitch
//   -> x-stitch joins pieces along x into maximal strips.
// Tiles with more boxes than the kernel capacity are split into x-sub-columns
// and processed in a follow-up launch.  Every launch writes at offsets from
// the tile kernel's look-back, so results are deterministic.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "vl_internal.h"

using namespace vl;

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;
static bool g_debug = getenv("VL_DEBUG") != nullptr;

#include <chrono>
static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int vl::set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCK(expr)                                                                      \
  do {                                                                                   \
    hipError_t _r = (expr);                                                              \
    if (_r != hipSuccess) return fail(VL_EHIP, "%s: %s", #expr, hipGetErrorString(_r)); \
  } while (0)

#define VLCK(expr)          \
  do {                      \
    int _s = (expr);        \
    if (_s != VL_OK) return _s; \
  } while (0)

// ---------------------------------------------------------------------------
// device memory pool (stream-ordered reuse: every engine launch uses one
// stream).  When the device is full, the pool frees its cached buffers and,
// for a worker engine, those of the engine that created it (`fallback`,
// idle while its workers run a program): worker streams running next to
// each other must not fail where one stream would not.
struct Pool {
  std::map<size_t, std::vector<void*>> free_;
  size_t live = 0;
  size_t cached = 0;   // bytes held in free_
  bool no_trim = false;
  Pool* fallback = nullptr;
  std::mutex m;   // the fallback's cache is freed from a worker's thread
  // Size classes: powers of two up to 1 MiB, then four per power of two
  // (a large buffer wastes at most a quarter of its size, not a half).
  static size_t round(size_t b) {
    size_t r = 256;
    while (r < b) r <<= 1;
    if (r <= ((size_t)1 << 20)) return r;
    const size_t q = r >> 3;   // r/2 < b <= r: the smallest of 5/8, 6/8, 7/8 and 8/8 of r that holds b
    for (size_t k = 5; k < 8; k++)
      if (k * q >= b) return k * q;
    return r;
  }
  void* get(size_t b, size_t* rounded) {
    size_t r = round(b ? b : 1);
    *rounded = r;
    std::unique_lock<std::mutex> lk(m);
    auto it = free_.find(r);
    if (it != free_.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      cached -= r;
      return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, r) != hipSuccess) {
      (void)hipGetLastError();
      if (no_trim) return nullptr;   // buffers may be referenced by a graph being captured
      trim_locked();
      bool ok = hipMalloc(&p, r) == hipSuccess;
      if (!ok && fallback) {
        (void)hipGetLastError();
        fallback->trim();
        ok = hipMalloc(&p, r) == hipSuccess;
      }
      if (!ok) {
        (void)hipGetLastError();
        return nullptr;
      }
    }
    live += r;
    return p;
  }
  void put(void* p, size_t r) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(m);
    free_[r].push_back(p);
    cached += r;
  }
  void trim() {
    std::lock_guard<std::mutex> lk(m);
    trim_locked();
  }
  void trim_locked() {
    for (auto& kv : free_)
      for (void* p : kv.second) {
        (void)hipFree(p);
        live -= kv.first;
      }
    free_.clear();
    cached = 0;
  }
  // Device bytes an allocation can count on: free memory plus this pool's cache.
  size_t available() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      (void)hipGetLastError();
      return SIZE_MAX;
    }
    std::lock_guard<std::mutex> lk(m);
    return fr + cached;
  }
};

struct vl_engine {
  int device = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::recursive_mutex mu;
  // Lifetime: the engine handle and every layer / program created on it hold
  // a reference; vl_engine_destroy drops the handle's reference and the
  // engine is freed when the last reference goes, whatever the destroy order.
  std::atomic<int> refs{1};
  // Worker engines of a program run on the same device with their own stream
  // and pool; they act on the layers of the engine that created them (root).
  vl_engine* root = this;
  Pool pool;
  uint32_t* d_err = nullptr;
  uint64_t* h_stage = nullptr;   // pinned, 4 KiB
  // tuning
  int nmax = 63;                 // wave-kernel tile capacity (63 or 127 boxes)
  int target = 20;               // mean boxes per tile the grid aims for
  int erode_target = 12;         // same for SIZE(-d) erosion grids (windows carry a halo)
  int level_factor = 1024;       // row coarsening per seam level
  int erode_kernel = 1;          // SIZE(-d) by the per-tile erosion kernel (0: by complement, grow, difference)
  int max_wg_cap = 0;            // 0 = occupancy limit
  int fuse = 1;                  // programs: fuse operator groups over one source layer
  int fuse2 = 0;                 // programs: groups may take a second layer as an image (set by the autotuner)
  int fuse_target = 6;           // mean boxes per tile of a fused group's grid
  int bin_mode = 2;              // 2: atomic binning, atomics grouped per wave (1: per run); 0: radix sort
  int radix_bits = 8;            // digit width of the radix sorts (8 or 11)
  int workers = 4;               // programs: chains run concurrently on up to this many worker streams
  int and_check = 1;             // programs: ANDs read only by ASSERT_ZERO are tested for emptiness first
  int check_target = 6;          // mean binned boxes per tile of that test's grid
  int tile_shape = 32;           // tiles shaped like the boxes, up to this aspect (0: square tiles)
  int64_t fused_groups = 0;      // groups of operators run fused so far (programs)
  // Strips per binned entry that a tile stage's scratch stream is sized for
  // (plain tile stage / each output of a fused group).  Raised when a stream
  // is short, so later operators of this engine do not repeat the retry.
  double region_scale = 4.0;
  double fuse_region_scale = 4.0;
  double last_ms = 0.0;
  // kernel statistics (collected while prof is set)
  struct KStat {
    int64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
  };
  bool prof = false;
  KStat kst[5];                  // KS_* below
  // Event pairs of profiled stages whose durations are read later (at the
  // next statistics query or the end of a program run), so that collecting
  // statistics adds no host waits to the stream.
  struct ProfRec {
    hipEvent_t a, b;
    int kind;
    int64_t launches;
    double bytes;
  };
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> ev_free;
  // Graph capture of programs.  Record mode runs eagerly and logs every
  // value the host reads back from the device; capture mode replays that
  // log instead of reading (no host waits, so the work can be captured into
  // a HIP graph), enqueues a check of each logged size against the device
  // value (d_err[4] is set on a mismatch), and keeps the host sources of its
  // uploads alive for the graph.
  struct Capture {
    int mode = 0;   // CAP_EAGER, CAP_RECORD, CAP_CAPTURE
    std::vector<std::vector<uint8_t>> log;
    size_t pos = 0;
    std::vector<void*> pinned;
  } cap;
};

enum { CAP_EAGER = 0, CAP_RECORD = 1, CAP_CAPTURE = 2 };

enum { KS_TILE = 0, KS_BIN = 1, KS_SORT = 2, KS_FUSED = 3, KS_SNAP = KSTAT_SNAP, KS_N = 5 };

struct vl_layer {
  vl_engine* e = nullptr;
  // Lifetime: the handle and every program that records the layer hold a
  // reference; the layer (and its engine reference) goes with the last one.
  std::atomic<int> refs{1};
  vl_box* d = nullptr;
  int64_t n = 0;
  size_t bytes = 0;   // pool-rounded allocation size
  // Bumped by every write (captured programs replay only while the versions
  // of the layers they read and write are the ones they were captured with).
  uint64_t version = 0;
  // Cached bounding box of the boxes (valid while mbr_ok; every write to the
  // layer clears it).  Operators read it instead of reducing the boxes again.
  mutable bool mbr_ok = false;
  mutable int64_t mbr[4] = {0, 0, 0, 0};
  mutable double sum_w = 0.0, sum_h = 0.0;   // summed box widths and heights (valid with mbr)
  // A box known to contain every box of the layer (not necessarily the
  // smallest), valid while version == bound_version: set by producers that
  // know it without a reduction (a fused group's outputs lie in its domain).
  uint64_t bound_version = ~0ull;
  int64_t bound[4] = {0, 0, 0, 0};
  double shape_hint = 1.0;   // tile height / width suited to the boxes (with the bound)
};

namespace {

// RAII scratch allocations returned to the pool at scope exit.
struct Scratch {
  vl_engine* e;
  std::vector<std::pair<void*, size_t>> v;
  bool oom = false;
  explicit Scratch(vl_engine* en) : e(en) {}
  ~Scratch() { release(); }
  // Returns every buffer to the pool (work already queued on the engine's
  // stream still sees them: later users are ordered behind it on the stream).
  void release() {
    for (auto& pr : v) e->pool.put(pr.first, pr.second);
    v.clear();
  }
  template <class T>
  T* get(size_t n) {
    size_t r = 0;
    void* p = e->pool.get(n * sizeof(T) + 64, &r);
    if (!p) {
      oom = true;
      return nullptr;
    }
    v.push_back({p, r});
    return (T*)p;
  }
};

// A growable device array owned by a Scratch-like holder.
template <class T>
struct DVec {
  vl_engine* e = nullptr;
  T* p = nullptr;
  size_t cap = 0;
  size_t bytes = 0;
  explicit DVec(vl_engine* en) : e(en) {}
  ~DVec() { release(); }
  void release() {
    if (p) e->pool.put(p, bytes);
    p = nullptr;
    cap = 0;
    bytes = 0;
  }
  // Ensure capacity >= n, keeping the first `keep` elements.
  int ensure(size_t n, size_t keep) {
    if (n <= cap) return VL_OK;
    size_t want = std::max(n, cap + cap / 2);
    size_t r = 0;
    T* q = (T*)e->pool.get(want * sizeof(T), &r);
    if (!q) return fail(VL_ENOMEM, "device allocation of %zu bytes failed", want * sizeof(T));
    if (p && keep) HIPCK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, e->st));
    if (p) e->pool.put(p, bytes);
    p = q;
    bytes = r;
    cap = r / sizeof(T);
    return VL_OK;
  }
  T* take(size_t* out_bytes) {   // hand ownership over
    T* q = p;
    *out_bytes = bytes;
    p = nullptr;
    cap = 0;
    bytes = 0;
    return q;
  }
};

// Host reads of device values.  Each segment is copied to its host address;
// `check`: in capture mode, a replay compares the device value with the
// logged one (segments whose value may differ between correct runs, such as
// a strip stream's demand, are not checked).  One stream wait per call.
struct RB {
  void* host;
  const void* dev;
  size_t bytes;
  bool check;
};

uint32_t* cap_flag(vl_engine* e) { return e->d_err + 4; }

// Status of a capture whose host decisions left the recorded path (the run
// falls back to eager execution).
constexpr int CAPTURE_DIVERGED = 1002;

int sync_read(vl_engine* e, const RB* r, int n) {
  size_t total = 0;
  for (int i = 0; i < n; i++) total += r[i].bytes;
  if (e->cap.mode == CAP_CAPTURE) {
    if (e->cap.pos >= e->cap.log.size() || e->cap.log[e->cap.pos].size() != total)
      return fail(CAPTURE_DIVERGED, "graph capture diverged from the recorded run");
    const std::vector<uint8_t>& v = e->cap.log[e->cap.pos++];
    size_t o = 0;
    for (int i = 0; i < n; i++) {
      memcpy(r[i].host, v.data() + o, r[i].bytes);
      if (r[i].check) launch_expect(r[i].dev, v.data() + o, r[i].bytes, cap_flag(e), e->st);
      o += r[i].bytes;
    }
    HIPCK(hipGetLastError());
    return VL_OK;
  }
  if (total <= 4096) {
    size_t o = 0;
    for (int i = 0; i < n; i++) {
      if (r[i].bytes) HIPCK(hipMemcpyAsync((char*)e->h_stage + o, r[i].dev, r[i].bytes, hipMemcpyDeviceToHost, e->st));
      o += r[i].bytes;
    }
    HIPCK(hipStreamSynchronize(e->st));
    o = 0;
    for (int i = 0; i < n; i++) {
      memcpy(r[i].host, (char*)e->h_stage + o, r[i].bytes);
      o += r[i].bytes;
    }
  } else {
    for (int i = 0; i < n; i++)
      if (r[i].bytes) HIPCK(hipMemcpyAsync(r[i].host, r[i].dev, r[i].bytes, hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
  }
  if (e->cap.mode == CAP_RECORD) {
    std::vector<uint8_t> v(total);
    size_t o = 0;
    for (int i = 0; i < n; i++) {
      memcpy(v.data() + o, r[i].host, r[i].bytes);
      o += r[i].bytes;
    }
    e->cap.log.push_back(std::move(v));
  }
  return VL_OK;
}

int readback(vl_engine* e, void* host, const void* dev, size_t bytes) {
  RB r{host, dev, bytes, true};
  return sync_read(e, &r, 1);
}

// Host-to-device copy of a small host array.  While capturing, the source is
// kept in pinned memory owned by the capture (the graph copies from it on
// every replay).
int upload(vl_engine* e, void* dev, const void* host, size_t bytes) {
  if (!bytes) return VL_OK;
  if (e->cap.mode == CAP_CAPTURE) {
    void* p = nullptr;
    HIPCK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    e->cap.pinned.push_back(p);
    memcpy(p, host, bytes);
    HIPCK(hipMemcpyAsync(dev, p, bytes, hipMemcpyHostToDevice, e->st));
    return VL_OK;
  }
  HIPCK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, e->st));
  return VL_OK;
}

int check_err(vl_engine* e, uint32_t* out) {
  uint32_t v = 0;
  VLCK(readback(e, &v, e->d_err, sizeof(v)));
  *out = v;
  return VL_OK;
}

int clear_err(vl_engine* e) {
  launch_fill_u32(e->d_err, 0u, 1, e->st);
  return VL_OK;
}

// Profiling marks: an event recorded on the engine stream (null when the
// engine does not collect statistics or no event can be made).
hipEvent_t prof_mark(vl_engine* e) {
  // none inside a graph capture: events recorded by graph replays do not
  // give usable elapsed times
  if (!e->prof || e->cap.mode == CAP_CAPTURE) return nullptr;
  hipEvent_t ev = nullptr;
  if (!e->ev_free.empty()) {
    ev = e->ev_free.back();
    e->ev_free.pop_back();
  } else if (hipEventCreate(&ev) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipEventRecord(ev, e->st) != hipSuccess) {
    (void)hipGetLastError();
    e->ev_free.push_back(ev);
    return nullptr;
  }
  return ev;
}

// Queues the stage between marks a and b (its duration is read later).
void prof_add(vl_engine* e, hipEvent_t a, hipEvent_t b, int kind, int64_t launches, double bytes) {
  if (!a || !b) {
    if (a) e->ev_free.push_back(a);
    if (b) e->ev_free.push_back(b);
    return;
  }
  e->prof_pending.push_back({a, b, kind, launches, bytes});
}

// Adds the durations of every queued stage to the statistics.
int prof_resolve(vl_engine* e) {
  int rc = VL_OK;
  for (auto& r : e->prof_pending) {
    float ms = 0.f;
    if (rc == VL_OK && (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess))
      rc = fail(VL_EHIP, "profiling events: %s", hipGetErrorString(hipGetLastError()));
    auto& k = e->kst[r.kind];
    k.launches += r.launches;
    k.ms += ms;
    k.bytes += r.bytes;
    e->ev_free.push_back(r.a);
    e->ev_free.push_back(r.b);
  }
  e->prof_pending.clear();
  return rc;
}

struct Box64 {
  int64_t x1, y1, x2, y2;
  bool empty() const { return x1 >= x2 || y1 >= y2; }
};

// Bounding boxes of A (transformed) and B; validates boxes.
// sums (may be null): summed widths and heights of A's (transformed) boxes,
// then of B's.
int mbr2(vl_engine* e, const vl_box* a, int64_t na, Xform xa, const vl_box* b, int64_t nb, Box64* ma, Box64* mb,
         double* sums = nullptr) {
  Scratch s(e);
  int32_t* d = s.get<int32_t>(8);
  unsigned long long* ds = s.get<unsigned long long>(4);
  if (!d || !ds) return fail(VL_ENOMEM, "scratch");
  // (INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN) twice, by fill kernels
  for (int k = 0; k < 8; k += 4) {
    launch_fill_u32((uint32_t*)d + k, (uint32_t)INT32_MAX, 2, e->st);
    launch_fill_u32((uint32_t*)d + k + 2, (uint32_t)INT32_MIN, 2, e->st);
  }
  launch_fill_u64((uint64_t*)ds, 0ull, 4, e->st);
  VLCK(clear_err(e));
  launch_mbr(a, na, xa, d, e->d_err, e->st, ds);
  launch_mbr(b, nb, Xform{0, 0, 0, 0}, d + 4, e->d_err, e->st, ds + 2);
  int32_t h[8];
  unsigned long long hs[4];
  uint32_t err = 0;
  const RB r[3] = {{h, d, sizeof(h), true}, {&err, e->d_err, 4, true}, {hs, ds, sizeof(hs), true}};
  VLCK(sync_read(e, r, 3));
  if (err & ERR_INVALID_BOX)
    return fail(VL_EINVAL, "invalid box: input has x1>=x2 or y1>=y2, or a transformed box became invalid");
  *ma = Box64{h[0], h[1], h[2], h[3]};
  *mb = Box64{h[4], h[5], h[6], h[7]};
  if (sums)
    for (int k = 0; k < 4; k++) sums[k] = (double)hs[k];
  return VL_OK;
}

// Tile height / width for boxes of summed widths sw and heights sh: their
// mean aspect, within [1/cap, cap] (cap 0: square tiles).  A tile shaped
// like the boxes is crossed by the fewest of them (for tiles of a given area,
// the expected tiles a box meets, (1 + w/tw)(1 + h/th), is least at
// th/tw = h/w on average), so the grid holds the fewest duplicate entries
// and seam pieces: vertical wires get tall tiles, horizontal wires wide ones.
double tile_shape(double sw, double sh, int cap) {
  if (cap <= 0 || !(sw > 0) || !(sh > 0)) return 1.0;
  return std::min((double)cap, std::max(1.0 / cap, sh / sw));
}

// Uniform grid of about nboxes / target tiles over dom, each tile about
// `shape` times as tall as it is wide.
Grid make_grid(Box64 dom, int64_t nboxes, int target, double shape = 1.0, int64_t max_tiles = (int64_t)1 << 26) {
  Grid g;
  g.x0 = dom.x1; g.y0 = dom.y1; g.x1 = dom.x2; g.y1 = dom.y2;
  const int64_t W = dom.x2 - dom.x1, H = dom.y2 - dom.y1;
  double nt = std::max(1.0, (double)nboxes / std::max(1, target));
  nt = std::min(nt, (double)max_tiles);
  double aspect = (double)W * shape / (double)H;
  int64_t cols = std::max<int64_t>(1, (int64_t)llround(std::sqrt(nt * aspect)));
  cols = std::min(cols, W);
  int64_t rows = std::max<int64_t>(1, (int64_t)std::ceil(nt / (double)cols));
  rows = std::min(rows, H);
  g.tw = (W + cols - 1) / cols;
  g.th = (H + rows - 1) / rows;
  g.cols = (int32_t)((W + g.tw - 1) / g.tw);
  g.rows = (int32_t)((H + g.th - 1) / g.th);
  return g;
}

// Entries a binning may make before its grid is coarsened: boxes that span
// many tiles (long strips of a complement, say, over a fine grid) would
// otherwise multiply the entries far beyond the boxes.
// (operators on ordinary layouts make 1 to 2.5 entries per box)
uint64_t bin_entry_limit(int64_t n) { return 4 * (uint64_t)n + ((uint64_t)1 << 16); }

// Grid g with tiles twice as wide and tall (none smaller than the domain).
bool coarsen_grid(Grid& g) {
  if (g.cols <= 1 && g.rows <= 1) return false;
  const int64_t W = g.x1 - g.x0, H = g.y1 - g.y0;
  g.tw = std::min<int64_t>(2 * g.tw, W);
  g.th = std::min<int64_t>(2 * g.th, H);
  g.cols = (int32_t)((W + g.tw - 1) / g.tw);
  g.rows = (int32_t)((H + g.th - 1) / g.th);
  return true;
}

// Output streams of one operator.
struct Outs {
  DVec<vl_box> out, side, next, stitch;
  DVec<uint8_t> side_tag, stitch_tag;
  uint64_t nout = 0, nside = 0, nstitch = 0;
  long long area = 0;
  // the current level is an AND: tiles without an overlapping pair are
  // skipped by a pair filter before the tile kernels
  bool and_filter = false;
  explicit Outs(vl_engine* e) : out(e), side(e), next(e), stitch(e), side_tag(e), stitch_tag(e) {}
};

// Strips per chunk of a tile stage's scratch stream (a workgroup takes a new
// chunk when its current one cannot hold the next tile's strips).
constexpr uint64_t STRIP_CHUNK = 1024;

bool two_layer(int top) { return top == TOP_AND || top == TOP_NOT || top == TOP_XOR; }

int process_tiles(vl_engine* e, const TileDesc* d_desc, int ntiles, const vl_box* d_entries, uint64_t nentries,
                  bool two, bool area, int32_t erode, Outs& o, int depth);

// Overflow tiles re-binned into x-sub-columns of about tgt boxes each.  halo >
// 0: windows grown by halo (a box belongs to every sub-column its extent grown
// by halo meets; sub-columns stay at least 2*halo wide).  Only x splits: the
// seam levels rely on row borders shared by every tile of a grid row (a strip
// crossing a row border is then cut in every tile it meets, and one not
// crossing one in none), which a y split of one tile would break.  The arrays
// live in s.
struct SplitTiles {
  TileDesc* desc = nullptr;
  int n = 0;
  vl_box* entries = nullptr;
  uint64_t total = 0;
};

// Overflow tiles of a launch, sorted by descriptor index (a deterministic
// order), and their descriptors.
int overflow_tiles(vl_engine* e, const TileDesc* d_desc, const uint32_t* d_ovf, uint32_t novf,
                   std::vector<uint32_t>& ovf, std::vector<TileDesc>& parents) {
  ovf.resize(novf);
  parents.resize(novf);
  // the list's order depends on scheduling: not checked on replays (its
  // length is, where the tile stage reads it)
  RB r{ovf.data(), d_ovf, novf * sizeof(uint32_t), false};
  VLCK(sync_read(e, &r, 1));
  std::sort(ovf.begin(), ovf.end());
  Scratch gs(e);
  uint32_t* d_idx = gs.get<uint32_t>(novf);
  TileDesc* d_par = gs.get<TileDesc>(novf);
  if (gs.oom) return fail(VL_ENOMEM, "scratch");
  VLCK(upload(e, d_idx, ovf.data(), novf * sizeof(uint32_t)));
  launch_gather_desc(d_desc, d_idx, (int)novf, d_par, e->st);
  RB rp{parents.data(), d_par, novf * sizeof(TileDesc), false};
  VLCK(sync_read(e, &rp, 1));
  return VL_OK;
}

// dense != null: tiles that no x-split can divide (too narrow) are listed
// there (descriptor indices, with their descriptors in dense_par) instead of
// failing.
int split_tiles(vl_engine* e, const TileDesc* d_desc, const uint32_t* d_ovf, uint32_t novf, const vl_box* d_entries,
                int32_t halo, int tgt, int cap, int depth, Scratch& s, SplitTiles* out,
                std::vector<uint32_t>* dense = nullptr, std::vector<TileDesc>* dense_par = nullptr) {
  if (depth > 24) return fail(VL_EINTERNAL, "tile splitting did not converge");
  std::vector<uint32_t> ovf;
  std::vector<TileDesc> parents;
  VLCK(overflow_tiles(e, d_desc, d_ovf, novf, ovf, parents));
  std::vector<TileDesc> subs;
  std::vector<int32_t> first(novf + 1);
  for (uint32_t i = 0; i < novf; i++) {
    const TileDesc& p = parents[i];
    first[i] = (int32_t)subs.size();
    const int64_t n = (int64_t)p.nA + p.nB;
    const int64_t width = (int64_t)p.x1 - p.x0;
    int64_t kx = std::max<int64_t>(2, (n + tgt - 1) / tgt);
    kx = std::min(kx, width);
    // windows overlap by twice the halo: keep sub-columns at least that wide
    if (halo > 0) kx = std::min<int64_t>(kx, width / (2 * (int64_t)halo));
    if (kx < 2) {
      if (!dense)
        return fail(VL_EINTERNAL, "a %lld-DBU-wide tile holds %lld boxes (more than the tile capacity %d)",
                    (long long)width, (long long)n, cap);
      dense->push_back(ovf[i]);
      dense_par->push_back(p);
      continue;
    }
    const int64_t sw = (width + kx - 1) / kx;
    kx = (width + sw - 1) / sw;
    for (int64_t c = 0; c < kx; c++) {
      TileDesc q = p;
      q.x0 = (int32_t)(p.x0 + c * sw);
      q.x1 = (int32_t)std::min<int64_t>(p.x0 + (c + 1) * sw, p.x1);
      q.flags = (p.flags & (SEAM_B | SEAM_T)) | ((c > 0 || (p.flags & SEAM_L)) ? SEAM_L : 0u) |
                ((c < kx - 1 || (p.flags & SEAM_R)) ? SEAM_R : 0u);
      q.nA = q.nB = 0;
      q.off = 0;
      subs.push_back(q);
    }
  }
  first[novf] = (int32_t)subs.size();
  const int ns = (int)subs.size();
  out->n = 0;
  out->total = 0;
  if (ns == 0) return VL_OK;
  TileDesc* d_par = s.get<TileDesc>(novf);
  TileDesc* d_sub = s.get<TileDesc>(ns);
  int32_t* d_first = s.get<int32_t>(novf + 1);
  uint32_t* cnt = s.get<uint32_t>(4 * (size_t)ns);
  uint64_t* off = s.get<uint64_t>(ns + 1);
  void* stmp = s.get<char>(scan_tmp_bytes(ns));
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  VLCK(upload(e, d_par, parents.data(), novf * sizeof(TileDesc)));
  VLCK(upload(e, d_sub, subs.data(), ns * sizeof(TileDesc)));
  VLCK(upload(e, d_first, first.data(), (novf + 1) * sizeof(int32_t)));
  uint32_t *cA = cnt, *cB = cnt + ns, *fA = cnt + 2 * ns, *fB = cnt + 3 * ns;
  launch_fill_u32(cnt, 0u, 4 * (size_t)ns, e->st);
  launch_split_count(d_par, (int)novf, d_first, d_sub, cA, cB, d_entries, halo, e->st);
  scan_pair_counts(cA, cB, off, ns, off + ns, stmp, e->st);
  uint64_t total = 0;
  VLCK(readback(e, &total, off + ns, sizeof(total)));
  {
    // boxes spanning many sub-columns (long strips) would multiply the
    // entries: those tiles go to the dense path whole instead
    uint64_t nin = 0;
    for (const TileDesc& p : parents) nin += (uint64_t)p.nA + p.nB;
    if (total > bin_entry_limit((int64_t)nin)) {
      if (!dense)
        return fail(VL_EINTERNAL, "splitting %u tiles would make %llu entries of %llu", novf,
                    (unsigned long long)total, (unsigned long long)nin);
      for (uint32_t i = 0; i < novf; i++) {
        if (first[i + 1] == first[i]) continue;   // already listed as dense
        dense->push_back(ovf[i]);
        dense_par->push_back(parents[i]);
      }
      return VL_OK;
    }
  }
  vl_box* sub_entries = s.get<vl_box>(total + 1);
  if (!sub_entries) return fail(VL_ENOMEM, "scratch");
  launch_split_scatter(d_par, (int)novf, d_first, d_sub, ns, off, cA, cB, fA, fB, d_entries, sub_entries, halo,
                       e->st);
  out->desc = d_sub;
  out->n = ns;
  out->entries = sub_entries;
  out->total = total;
  return VL_OK;
}

// Status of an erosion (SIZE(-d)) when a tile exceeds every erosion kernel
// and no x-split divides it: nothing is written, and SIZE runs as the
// complement composition instead, whose operators have the dense path.
constexpr int ERODE_FALLBACK = 1001;

// Tiles above every capacity-bounded kernel (list: descriptor indices in
// tile order, par: their descriptors): vl_dense.hip, one workgroup per tile,
// in batches whose scratch stays below a few GiB.  Their strips go to the
// operator's streams in list order.
int process_dense(vl_engine* e, const TileDesc* d_desc, const std::vector<uint32_t>& list,
                  const std::vector<TileDesc>& par, const vl_box* d_entries, bool two, bool area, Outs& o) {
  const size_t nl = list.size();
  constexpr int64_t COORD_BUDGET = (int64_t)1 << 27;   // int32 coordinate slots per batch
  constexpr uint64_t GRID_BUDGET = (uint64_t)1 << 28;  // grid words per launch
  for (size_t i0 = 0; i0 < nl;) {
    // batch [i0, i1) by coordinate scratch
    std::vector<int64_t> coff;
    int64_t cs_total = 0;
    size_t i1 = i0;
    while (i1 < nl) {
      const int64_t c = dense_coord_slots((int64_t)par[i1].nA + par[i1].nB);
      if (i1 > i0 && cs_total + c > COORD_BUDGET) break;
      coff.push_back(cs_total);
      cs_total += c;
      i1++;
    }
    const int nb = (int)(i1 - i0);
    Scratch s(e);
    uint32_t* d_list = s.get<uint32_t>(nb);
    int64_t* d_coff = s.get<int64_t>(nb);
    int32_t* d_cs = s.get<int32_t>((size_t)cs_total);
    int32_t* d_nxy = s.get<int32_t>(2 * (size_t)nb);
    if (s.oom) return fail(VL_ENOMEM, "dense-tile coordinate scratch (%lld slots)", (long long)cs_total);
    VLCK(upload(e, d_list, list.data() + i0, nb * sizeof(uint32_t)));
    VLCK(upload(e, d_coff, coff.data(), nb * sizeof(int64_t)));
    launch_dense_coords(d_desc, d_list, nb, d_entries, d_coff, d_cs, d_nxy, e->st);
    HIPCK(hipGetLastError());
    std::vector<int32_t> nxy(2 * (size_t)nb);
    {
      RB r{nxy.data(), d_nxy, nxy.size() * sizeof(int32_t), nxy.size() <= 64};
      VLCK(sync_read(e, &r, 1));
    }
    // launches [j0, j1) of this batch by grid size
    for (int j0 = 0; j0 < nb;) {
      std::vector<uint64_t> goff, cofs, soff;
      uint64_t gw_total = 0, col_total = 0, s_total = 0;
      int j1 = j0;
      while (j1 < nb) {
        const int64_t ncol = nxy[2 * j1] - 1, nrows = nxy[2 * j1 + 1] - 1;
        const int64_t W = (nrows + 64) >> 6;
        const uint64_t gw = (uint64_t)ncol * W * (two ? 2 : 1);
        if (j1 > j0 && gw_total + gw > GRID_BUDGET) break;
        goff.push_back(gw_total);
        cofs.push_back(col_total);
        soff.push_back(s_total);
        gw_total += gw;
        col_total += (uint64_t)ncol;
        s_total += (uint64_t)ncol * (uint64_t)(nrows / 2 + 1);
        j1++;
      }
      const int m = j1 - j0;
      Scratch t(e);
      uint64_t* d_goff = t.get<uint64_t>(m);
      uint64_t* d_cofs = t.get<uint64_t>(m);
      uint64_t* d_soff = t.get<uint64_t>(m);
      uint64_t* d_grid = t.get<uint64_t>(gw_total + 1);
      uint32_t* d_cc = t.get<uint32_t>(col_total + 1);
      vl_box* d_str = area ? nullptr : t.get<vl_box>(s_total + 1);
      uint8_t* d_tag = area ? nullptr : t.get<uint8_t>(s_total + 1);
      uint64_t* d_toff = t.get<uint64_t>(m);
      uint32_t* d_tn = t.get<uint32_t>(2 * (size_t)m);
      unsigned long long* d_cnt = t.get<unsigned long long>(2);
      long long* d_area = t.get<long long>(1);
      if (t.oom) return fail(VL_ENOMEM, "dense-tile grids (%llu words)", (unsigned long long)gw_total);
      VLCK(upload(e, d_goff, goff.data(), m * sizeof(uint64_t)));
      VLCK(upload(e, d_cofs, cofs.data(), m * sizeof(uint64_t)));
      VLCK(upload(e, d_soff, soff.data(), m * sizeof(uint64_t)));
      launch_fill_u64((uint64_t*)d_cnt, 0ull, 2, e->st);
      launch_fill_u64((uint64_t*)d_area, 0ull, 1, e->st);
      launch_dense_tile(d_desc, d_list + j0, m, d_entries, d_coff + j0, d_cs, d_nxy + 2 * j0, d_grid, d_goff, d_cc,
                        d_cofs, d_str, d_tag, d_soff, d_toff, d_tn, two, d_cnt, area ? d_area : nullptr, e->st);
      HIPCK(hipGetLastError());
      struct {
        unsigned long long cnt[2];
        long long area;
      } rb;
      {
        const RB r[2] = {{rb.cnt, d_cnt, 16, true}, {&rb.area, d_area, 8, true}};
        VLCK(sync_read(e, r, 2));
      }
      if (area) {
        o.area += rb.area;
      } else if (rb.cnt[0] > 0) {
        const uint64_t nside = rb.cnt[1], nfinal = rb.cnt[0] - nside;
        VLCK(o.out.ensure(o.nout + nfinal + 1, o.nout));
        VLCK(o.side.ensure(o.nside + nside + 1, o.nside));
        VLCK(o.side_tag.ensure(o.nside + nside + 1, o.nside));
        uint64_t* dst = t.get<uint64_t>(2 * (size_t)m + 2);
        void* stmp = t.get<char>(scan_tmp_bytes(m));
        if (t.oom) return fail(VL_ENOMEM, "scratch");
        scan_u32_to_u64(d_tn, dst, m, nullptr, stmp, e->st);
        scan_u32_to_u64(d_tn + m, dst + m, m, nullptr, stmp, e->st);
        launch_tile_reorder(m, d_toff, d_tn, m, dst, dst + m, d_str, d_tag, o.out.p, o.nout, o.side.p, o.side_tag.p,
                            o.nside, e->st);
        HIPCK(hipGetLastError());
        o.nout += nfinal;
        o.nside += nside;
      }
      if (g_debug)
        fprintf(stderr, "[vl]  dense tiles=%d strips=%llu seam=%llu grid words=%llu\n", m, rb.cnt[0], rb.cnt[1],
                (unsigned long long)gw_total);
      j0 = j1;
    }
    i0 = i1;
  }
  return VL_OK;
}

// Tiles above the workgroup kernel's capacity.  The first time (depth 0)
// they are split into x-sub-columns of about the grid's target box count and
// processed again; a sub-column still above capacity, and a tile too narrow
// to split, holds more boxes over one x-position than any x-split can
// separate, and runs on the dense kernel (erosion: ERODE_FALLBACK).
int split_and_process(vl_engine* e, const TileDesc* d_desc, const uint32_t* d_ovf, uint32_t novf,
                      const vl_box* d_entries, bool two, bool area, int32_t erode, Outs& o, int depth) {
  Scratch s(e);
  if (depth >= 1) {
    if (erode > 0) return ERODE_FALLBACK;
    std::vector<uint32_t> ovf;
    std::vector<TileDesc> par;
    VLCK(overflow_tiles(e, d_desc, d_ovf, novf, ovf, par));
    return process_dense(e, d_desc, ovf, par, d_entries, two, area, o);
  }
  SplitTiles sp;
  std::vector<uint32_t> dense;
  std::vector<TileDesc> dpar;
  const int rc = split_tiles(e, d_desc, d_ovf, novf, d_entries, erode, std::max(8, e->target), TILE_BLOCK_NMAX, depth,
                             s, &sp, erode > 0 ? nullptr : &dense, &dpar);
  if (rc == VL_EINTERNAL && erode > 0) return ERODE_FALLBACK;
  VLCK(rc);
  if (!dense.empty()) VLCK(process_dense(e, d_desc, dense, dpar, d_entries, two, area, o));
  return process_tiles(e, sp.desc, sp.n, sp.entries, sp.total, two, area, erode, o, depth + 1);
}

int process_tiles(vl_engine* e, const TileDesc* d_desc, int ntiles, const vl_box* d_entries, uint64_t nentries,
                  bool two, bool area, int32_t erode, Outs& o, int depth) {
  if (ntiles <= 0) return VL_OK;
  Scratch s(e);
  // Tiers: wave kernels of growing capacity, each taking the previous tier's
  // overflow list, then the 256-thread kernel; its overflow is split.
  int tiers[3];
  int ntier = 0;
  if (erode > 0) {
    tiers[ntier++] = 15;
    tiers[ntier++] = 31;
  } else {
    tiers[ntier++] = 31;
    tiers[ntier++] = 63;
    if (e->nmax == 127) tiers[ntier++] = 127;
  }
  uint32_t* small = s.get<uint32_t>(4);   // overflow count of tier t at [t], of the block kernel at [ntier]
  unsigned long long* cnt = s.get<unsigned long long>(3);   // [0] strips [1] seam strips [2] stream bump counter
  uint32_t* ovfl[4];
  for (int t = 0; t <= ntier; t++) ovfl[t] = s.get<uint32_t>(ntiles);
  uint32_t* ovf2 = ovfl[ntier];
  long long* d_area = s.get<long long>(1);
  uint64_t* tab_off = area ? nullptr : s.get<uint64_t>((size_t)ntiles);
  uint32_t* tab_n = area ? nullptr : s.get<uint32_t>(2 * (size_t)ntiles);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  // AND: the first tier takes only the tiles the pair filter keeps
  uint32_t* flist = nullptr;
  uint32_t* flist_n = nullptr;
  // the stage's time includes the filter (its bytes are the entry reads the
  // stage formula counts once)
  hipEvent_t pa_first = prof_mark(e);
  if (o.and_filter && two && !area && erode == 0) {
    flist = s.get<uint32_t>((size_t)ntiles);
    flist_n = s.get<uint32_t>(1);
    if (s.oom) return fail(VL_ENOMEM, "scratch");
    launch_fill_u32(flist_n, 0u, 1, e->st);
    launch_and_filter(d_desc, ntiles, d_entries, flist, flist_n, tab_n, ntiles, e->st);
  }
  // Exact grids: workgroup g of a launch takes tiles g, g+G, ... and takes
  // chunks of the scratch stream as it needs them (nregions bounds the chunks
  // left partly used).
  int G[4];
  int nregions = 0;
  for (int t = 0; t < ntier; t++) {
    int mw = tile_wave_max_wg(tiers[t], two, area, erode > 0);
    if (e->max_wg_cap > 0) mw = std::min(mw, e->max_wg_cap);
    G[t] = std::max(1, std::min(ntiles, mw));
    nregions += G[t];
  }
  G[ntier] = std::max(1, std::min(ntiles, tile_block_max_wg(two, area, erode > 0)));
  nregions += G[ntier];
  DVec<vl_box> scr(e);
  DVec<uint8_t> scrT(e);
  uint64_t total = 0, nsideT = 0;
  uint32_t nov = 0;
  const double t_pt = g_debug ? now_ms() : 0.0;
  // The scratch stream: the expected strips (region_scale per entry, learnt
  // from earlier operators) plus one chunk per workgroup (the most a
  // workgroup leaves unused).  A short stream makes the launch run again with
  // the demand the stream's bump counter measured, plus a quarter (the
  // chunks a rerun leaves unused depend on scheduling).
  uint64_t cap = (uint64_t)(e->region_scale * (double)nentries) + (uint64_t)nregions * STRIP_CHUNK + 1024;
  for (int attempt = 0; attempt < 4; attempt++) {
    if (!area) {
      VLCK(scr.ensure(cap, 0));
      VLCK(scrT.ensure(cap, 0));
    }
    launch_fill_u32(small, 0u, 4, e->st);
    launch_fill_u64((uint64_t*)cnt, 0ull, 3, e->st);
    launch_fill_u64((uint64_t*)d_area, 0ull, 1, e->st);
    VLCK(clear_err(e));
    TileOut to;
    to.strips = scr.p;
    to.tag = scrT.p;
    to.stream_cap = area ? 0 : cap;
    to.chunk = STRIP_CHUNK;
    to.next = cnt + 2;
    to.erode = erode;
    to.cnt = cnt;
    to.tab_off = tab_off;
    to.tab_n = tab_n;
    to.ntab = ntiles;
    to.area = d_area;
    to.err = e->d_err;
    hipEvent_t pa = attempt == 0 ? pa_first : prof_mark(e);
    for (int t = 0; t <= ntier; t++) {
      to.list = t ? ovfl[t - 1] : flist;          // tiles above the previous tier's capacity
      to.list_n = t ? small + (t - 1) : flist_n;  // (count known only on the device)
      to.ovf = ovfl[t];
      to.ovf_n = small + t;
      if (t < ntier) launch_tile_wave(tiers[t], two, area, erode > 0, d_desc, ntiles, d_entries, to, G[t], e->st);
      else launch_tile_block(two, area, erode > 0, d_desc, ntiles, d_entries, to, G[t], e->st);
    }
    HIPCK(hipGetLastError());
    hipEvent_t pb = prof_mark(e);
    struct {
      unsigned long long cnt[3];
      uint32_t err, nov;
      long long area;
    } rb;
    {
      // the stream's demand (cnt[2]) depends on scheduling: not checked
      const RB r[5] = {{rb.cnt, cnt, 16, true},
                       {rb.cnt + 2, cnt + 2, 8, false},
                       {&rb.err, e->d_err, 4, true},
                       {&rb.nov, small + ntier, 4, true},
                       {&rb.area, d_area, 8, true}};
      VLCK(sync_read(e, r, 5));
    }
    prof_add(e, pa, pb, KS_TILE, ntier + 1, 16.0 * (double)nentries + 16.0 * (double)rb.cnt[0] + 48.0 * ntiles);
    total = rb.cnt[0];
    nsideT = rb.cnt[1];
    nov = rb.nov;
    if (area) {
      o.area += rb.area;
      break;
    }
    if (rb.err & ERR_OVERFLOW_CAP) {
      if (attempt == 3) return fail(VL_EINTERNAL, "output capacity still short after resize");
      cap = std::max<uint64_t>(2 * cap, rb.cnt[2] + rb.cnt[2] / 4 + (uint64_t)nregions * STRIP_CHUNK + 1024);
      e->region_scale = std::max(e->region_scale, 1.25 * (double)rb.cnt[2] / std::max<double>(1.0, nentries));
      if (g_debug) fprintf(stderr, "[vl]  strip stream short: retry with %llu strips\n", (unsigned long long)cap);
      continue;
    }
    break;
  }
  const uint64_t nfinal = total - nsideT;
  if (!area && total > 0) {
    // tile order: exclusive scans of the per-tile counts, then a coalesced copy
    VLCK(o.out.ensure(o.nout + nfinal + 1, o.nout));
    VLCK(o.side.ensure(o.nside + nsideT + 1, o.nside));
    VLCK(o.side_tag.ensure(o.nside + nsideT + 1, o.nside));
    uint64_t* dst = s.get<uint64_t>(2 * (size_t)ntiles + 2);
    void* stmp = s.get<char>(scan_tmp_bytes(ntiles));
    if (s.oom) return fail(VL_ENOMEM, "scratch");
    scan_u32_to_u64(tab_n, dst, ntiles, nullptr, stmp, e->st);
    scan_u32_to_u64(tab_n + ntiles, dst + ntiles, ntiles, nullptr, stmp, e->st);
    launch_tile_reorder(ntiles, tab_off, tab_n, ntiles, dst, dst + ntiles, scr.p, scrT.p, o.out.p, o.nout, o.side.p,
                        o.side_tag.p, o.nside, e->st);
    HIPCK(hipGetLastError());
    o.nout += nfinal;
    o.nside += nsideT;
  }
  if (g_debug)
    fprintf(stderr, "[vl]  tiles=%d entries=%llu depth=%d two=%d area=%d final=%llu side=%llu overflow=%u %.2f ms\n",
            ntiles, (unsigned long long)nentries, depth, (int)two, (int)area, (unsigned long long)nfinal,
            (unsigned long long)nsideT, nov, now_ms() - t_pt);
  if (nov > 0) {
    const double t_sp = g_debug ? now_ms() : 0.0;
    VLCK(split_and_process(e, d_desc, ovf2, nov, d_entries, two, area, erode, o, depth));
    if (g_debug) fprintf(stderr, "[vl]  split %u tiles (depth %d): %.2f ms\n", nov, depth, now_ms() - t_sp);
  }
  return VL_OK;
}

// Status of run_fused when the group cannot run fused here (nothing written;
// the caller runs the group's operators one by one).
constexpr int FUSE_DECLINED = 1000;
constexpr int BIN_TOO_MANY = 1003;   // bin_entries without coarsening: the grid would need it

// Bins up to two inputs into grid g: entries of tile t (layer A, then B, each
// in input order) at entries[start[2t] .. start[2t+2]).  Layer A boxes are
// binned by their extent grown by xr and stored grown by xs.  The arrays are
// allocated from `keep` and live as long as it does.  When the boxes would
// make more than bin_entry_limit entries, g is coarsened (coarsen true: the
// caller's grid changes) or BIN_TOO_MANY is returned.
int bin_entries(vl_engine* e, Grid& g, const vl_box* a, int64_t na, Xform xs, Xform xr, const vl_box* b, int64_t nb,
                Scratch& keep, vl_box** entries_out, uint64_t** start_out, uint64_t* total_out,
                bool coarsen = true) {
  const int64_t nt0 = grid_tiles(g);
  const int64_t n = na + nb;
  if (e->bin_mode >= 1) {
    // atomic binning: count per (tile, layer) key, scan, scatter (bin_mode 2:
    // the atomics grouped by distinct key over the wave)
    uint64_t* start = keep.get<uint64_t>(2 * (size_t)nt0 + 1);
    Scratch s(e);
    uint32_t* cnt = s.get<uint32_t>(2 * (size_t)nt0);
    void* stmp = s.get<char>(scan_tmp_bytes(2 * (size_t)nt0));
    if (s.oom || keep.oom) return fail(VL_ENOMEM, "scratch for %lld tiles", (long long)nt0);
    hipEvent_t pa = prof_mark(e);
    uint64_t total = 0;
    int64_t nt = nt0;
    for (;;) {   // coarser grids have fewer tiles: the arrays of the first fit
      nt = grid_tiles(g);
      launch_fill_u32(cnt, 0u, 2 * (size_t)nt, e->st);
      launch_bin_count(a, na, xr, b, nb, g, cnt, e->bin_mode == 2, e->st);
      scan_u32_to_u64(cnt, start, 2 * (size_t)nt, start + 2 * nt, stmp, e->st);
      VLCK(readback(e, &total, start + 2 * nt, sizeof(total)));
      if (total <= bin_entry_limit(n)) break;
      if (!coarsen || g.zb) return BIN_TOO_MANY;
      if (g_debug) fprintf(stderr, "[vl]  %llu entries for %lld boxes: coarser grid\n", (unsigned long long)total,
                           (long long)n);
      if (!coarsen_grid(g)) break;
    }
    vl_box* entries = keep.get<vl_box>(total + 1);
    if (!entries) return fail(VL_ENOMEM, "binning scratch (%llu entries)", (unsigned long long)total);
    launch_bin_scatter(a, na, xr, xs, b, nb, g, start, cnt, entries, e->bin_mode == 2, e->st);
    HIPCK(hipGetLastError());
    hipEvent_t pb = prof_mark(e);
    // algorithmic bytes: the boxes read twice (count, scatter), 16 B written
    // per entry, the counters zeroed, scanned and counted down
    prof_add(e, pa, pb, KS_BIN, 1, 32.0 * (double)n + 16.0 * (double)total + 24.0 * (double)nt);
    *entries_out = entries;
    *start_out = start;
    *total_out = total;
    return VL_OK;
  }
  Scratch s(e);
  uint32_t* cnt = s.get<uint32_t>((size_t)n + 1);
  uint64_t* pos = s.get<uint64_t>((size_t)n + 1);
  void* stmp = s.get<char>(scan_tmp_bytes((size_t)n + 1));
  uint64_t* start = keep.get<uint64_t>(2 * (size_t)nt0 + 1);
  if (s.oom || keep.oom) return fail(VL_ENOMEM, "scratch for %lld tiles", (long long)nt0);
  uint64_t total = 0;
  for (;;) {
    launch_bin_span(a, na, xr, b, nb, g, cnt, e->st);
    scan_u32_to_u64(cnt, pos, (size_t)n, pos + n, stmp, e->st);
    VLCK(readback(e, &total, pos + n, sizeof(total)));
    if (total <= bin_entry_limit(n)) break;
    if (!coarsen || g.zb) return BIN_TOO_MANY;
    if (!coarsen_grid(g)) break;
  }
  const int64_t nt = grid_tiles(g);
  Scratch s2(e);
  uint32_t* keys = s2.get<uint32_t>(total + 1);
  uint32_t* vals = s2.get<uint32_t>(total + 1);
  uint32_t* keys2 = s2.get<uint32_t>(total + 1);
  uint32_t* vals2 = s2.get<uint32_t>(total + 1);
  void* rtmp = s2.get<char>(radix_plan((int64_t)total, e->radix_bits).tmp_bytes);
  vl_box* entries = keep.get<vl_box>(total + 1);
  if (s2.oom || !entries) return fail(VL_ENOMEM, "binning scratch (%llu entries)", (unsigned long long)total);
  hipEvent_t pa = prof_mark(e);
  launch_bin_emit(a, na, xr, b, nb, g, pos, keys, vals, e->st);
  int key_bits = 1;
  while (key_bits < 32 && ((uint64_t)1 << key_bits) < (uint64_t)(2 * nt)) key_bits++;
  hipEvent_t ps = prof_mark(e);
  if (radix_sort_pairs_bits32(keys, vals, keys2, vals2, (int64_t)total, key_bits, rtmp, e->st, e->radix_bits)) {
    std::swap(keys, keys2);
    std::swap(vals, vals2);
  }
  hipEvent_t pe = prof_mark(e);
  launch_bin_bounds(keys, total, 2 * nt, start, e->st);
  launch_bin_gather(a, na, xs, b, vals, total, entries, e->st);
  HIPCK(hipGetLastError());
  hipEvent_t pb = prof_mark(e);
  {
    // algorithmic bytes: the radix passes read and write 8-byte pairs; the
    // binning as a whole reads the boxes twice (span, emit), writes the
    // pairs, sorts them, gathers 16 B per entry and writes 16 B per entry
    const int passes = (key_bits + 7) / 8;
    const double sort_bytes = 16.0 * (double)total * passes;
    prof_add(e, ps, pe, KS_SORT, passes, sort_bytes);
    prof_add(e, pa, pb, KS_BIN, 1, 32.0 * (double)n + 8.0 * (double)total + sort_bytes + 32.0 * (double)total);
  }
  *entries_out = entries;
  *start_out = start;
  *total_out = total;
  return VL_OK;
}

// Bins up to two inputs into grid g and processes the tiles.  Layer A boxes
// are binned by their extent grown by xr and stored grown by xs; erode > 0
// selects the erosion tile kernel (xr is then the erosion halo, xs zero).
// g may come back coarser (bin_entries): callers go on with it.
int bin_and_process(vl_engine* e, Grid& g, int top, const vl_box* a, int64_t na, Xform xs, Xform xr,
                    const vl_box* b, int64_t nb, bool area, int32_t erode, Outs& o) {
  Scratch keep(e);
  vl_box* entries = nullptr;
  uint64_t* start = nullptr;
  uint64_t total = 0;
  VLCK(bin_entries(e, g, a, na, xs, xr, b, nb, keep, &entries, &start, &total));
  const int64_t nt = (int64_t)g.cols * g.rows;
  TileDesc* desc = keep.get<TileDesc>(nt);
  if (!desc) return fail(VL_ENOMEM, "scratch for %lld tiles", (long long)nt);
  launch_build_desc(g, start, top, desc, e->st);
  HIPCK(hipGetLastError());
  o.and_filter = top == TOP_AND && nb > 0 && !area && erode == 0;
  // single-layer AND rasterises into a second mask as well
  const bool two = (two_layer(top) && nb > 0) || top == TOP_OVERLAP;
  return process_tiles(e, desc, (int)nt, entries, total, two, area, erode, o, 0);
}

// Moves result strips into layer c.
int adopt(vl_engine* e, Outs& o, vl_layer* c) {
  size_t bytes = 0;
  vl_box* p = o.out.take(&bytes);
  if (c->d) e->pool.put(c->d, c->bytes);
  c->d = p;
  c->bytes = bytes;
  c->n = (int64_t)o.nout;
  c->mbr_ok = false;
  c->version++;
  return VL_OK;
}

int set_empty(vl_engine* e, vl_layer* c) {
  (void)e;
  c->n = 0;
  c->mbr_ok = false;
  c->version++;
  return VL_OK;
}

// The whole operator pipeline.  area_out != null selects area mode.
// MBRs of the inputs: from the layers' caches when the transform cannot make a
// box invalid (all components >= 0 and the grown MBR stays in int32), else by
// a validating reduction.
int op_mbrs(vl_engine* e, const vl_box* a, int64_t na, Xform xa, const vl_box* b, int64_t nb, const Box64* ka,
            const Box64* kb, Box64* ma, Box64* mb) {
  const bool grow_ok = xa.e >= 0 && xa.n >= 0 && xa.w >= 0 && xa.s >= 0;
  if (ka && (kb || nb <= 0) && grow_ok) {
    Box64 g{ka->x1 - xa.w, ka->y1 - xa.s, ka->x2 + xa.e, ka->y2 + xa.n};
    if (na <= 0 || (g.x1 >= INT32_MIN && g.y1 >= INT32_MIN && g.x2 <= INT32_MAX && g.y2 <= INT32_MAX)) {
      *ma = na > 0 ? g : Box64{INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
      *mb = nb > 0 ? *kb : Box64{INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
      return VL_OK;
    }
  }
  return mbr2(e, a, na, xa, b, nb, ma, mb);
}

// Seam levels and x-stitch of an operator's level-0 streams (grid g), then the
// result into layer c.
int finish_op(vl_engine* e, Grid g, Outs& o, vl_layer* c, double t_start) {
  // seam levels
  int level = 0;
  while (o.nside > 0) {
    VLCK(o.next.ensure(o.nside, 0));
    VLCK(o.stitch.ensure(o.nstitch + o.nside, o.nstitch));
    VLCK(o.stitch_tag.ensure(o.nstitch + o.nside, o.nstitch));
    Scratch ps(e);
    void* ptmp = ps.get<char>(partition_tmp_bytes((int64_t)o.nside));
    uint64_t* cnts = ps.get<uint64_t>(2);
    if (ps.oom) return fail(VL_ENOMEM, "scratch");
    launch_partition_side(o.side.p, o.side_tag.p, (int64_t)o.nside, o.next.p, o.stitch.p + o.nstitch,
                          o.stitch_tag.p + o.nstitch, cnts, ptmp, e->st);
    uint64_t hc[2];
    VLCK(readback(e, hc, cnts, sizeof(hc)));
    if (g_debug && e->cap.mode != CAP_CAPTURE) {
      std::vector<uint8_t> tags(o.nside);
      HIPCK(hipMemcpy(tags.data(), o.side_tag.p, o.nside, hipMemcpyDeviceToHost));
      int hist[16] = {0};
      for (uint8_t t : tags) hist[t & 15]++;
      fprintf(stderr, "[vl] level %d grid %dx%d tw=%lld th=%lld side=%llu border=%llu stitch=%llu tags:", level,
              g.cols, g.rows, (long long)g.tw, (long long)g.th, (unsigned long long)o.nside,
              (unsigned long long)hc[0], (unsigned long long)hc[1]);
      for (int i = 0; i < 16; i++)
        if (hist[i]) fprintf(stderr, " %x:%d", i, hist[i]);
      fprintf(stderr, "\n");
    }
    o.nstitch += hc[1];
    o.nside = 0;
    if (hc[0] == 0) break;
    if (g.rows <= 1) return fail(VL_EINTERNAL, "row-seam strips remain on a single-row level");
    // next level: same columns, taller rows
    Grid gn = g;
    gn.th = std::min<int64_t>(g.th * e->level_factor, g.y1 - g.y0);
    gn.rows = (int32_t)((g.y1 - g.y0 + gn.th - 1) / gn.th);
    if (gn.rows == g.rows && g.rows > 1) { gn.rows = 1; gn.th = g.y1 - g.y0; }
    // Columns follow this level's own strip count (x-seams are harmless: the
    // x-stitch joins pieces at any x), so its tiles stay near the target size.
    {
      const int64_t W = g.x1 - g.x0;
      const double nt = std::max(1.0, (double)hc[0] / std::max(1, e->target));
      int64_t cols = std::max<int64_t>(1, (int64_t)llround(nt / gn.rows));
      cols = std::min(cols, W);
      gn.tw = (W + cols - 1) / cols;
      gn.cols = (int32_t)((W + gn.tw - 1) / gn.tw);
    }
    g = gn;
    level++;
    // The level's input (next) is consumed before o.next is reused.
    DVec<vl_box> in(e);
    std::swap(in.p, o.next.p);
    std::swap(in.cap, o.next.cap);
    std::swap(in.bytes, o.next.bytes);
    const double t_lv = g_debug ? now_ms() : 0.0;
    VLCK(bin_and_process(e, g, TOP_UNION, in.p, (int64_t)hc[0], Xform{0, 0, 0, 0}, Xform{0, 0, 0, 0}, nullptr, 0,
                         false, 0, o));
    if (g_debug) fprintf(stderr, "[vl] level %d: %llu strips, %.2f ms\n", level, (unsigned long long)hc[0],
                         now_ms() - t_lv);
  }
  // x-stitch
  const double t_st = g_debug ? now_ms() : 0.0;
  if (o.nstitch > 0) {
    const int64_t n = (int64_t)o.nstitch;
    Scratch ss(e);
    const uint64_t slots = stitch_table_slots(n);
    unsigned long long* tk = ss.get<unsigned long long>(slots);
    uint32_t* tv = ss.get<uint32_t>(slots);
    int32_t* next = ss.get<int32_t>(n);
    uint8_t* hasp = ss.get<uint8_t>(n);
    void* ptmp = ss.get<char>(partition_tmp_bytes(n));
    uint64_t* cnt = ss.get<uint64_t>(1);
    if (ss.oom) return fail(VL_ENOMEM, "stitch scratch");
    // corner keys (x1, y1) relative to the domain: y in the low yb bits
    auto bits_for = [](int64_t v) {
      int b = 1;
      while (b < 63 && (v >> b) != 0) b++;
      return b;
    };
    const int yb = bits_for(g.y1 - g.y0);
    launch_stitch_link(o.stitch.p, o.stitch_tag.p, n, g.x0, g.y0, yb, tk, tv, next, hasp, e->st);
    VLCK(o.out.ensure(o.nout + (uint64_t)n, o.nout));
    launch_stitch_emit(o.stitch.p, n, next, hasp, o.out.p, o.nout, ptmp, cnt, e->st);
    uint64_t heads = 0;
    VLCK(readback(e, &heads, cnt, sizeof(heads)));
    o.nout += heads;
  }
  if (g_debug)
    fprintf(stderr, "[vl] stitch %llu strips %.2f ms; total op %.2f ms, out %llu\n", (unsigned long long)o.nstitch,
            now_ms() - t_st, now_ms() - t_start, (unsigned long long)o.nout);
  HIPCK(hipGetLastError());
  return adopt(e, o, c);
}

int run_op(vl_engine* e, int top, const vl_box* a, int64_t na, Xform xa, const vl_box* b, int64_t nb,
           vl_layer* c, long long* area_out, int32_t erode = 0, const Box64* ka = nullptr,
           const Box64* kb = nullptr, double shape = 1.0) {
  if (!two_layer(top) && top != TOP_OR) { b = nullptr; nb = 0; kb = nullptr; }
  Box64 ma, mb;
  VLCK(op_mbrs(e, a, na, xa, b, nb, ka, kb, &ma, &mb));
  Box64 dom;
  const bool ea = na <= 0, eb = nb <= 0;
  switch (top) {
    case TOP_AND:
      if (ea || eb) { dom = Box64{0, 0, 0, 0}; break; }
      dom = Box64{std::max(ma.x1, mb.x1), std::max(ma.y1, mb.y1), std::min(ma.x2, mb.x2), std::min(ma.y2, mb.y2)};
      break;
    case TOP_NOT:
    case TOP_UNION:
    case TOP_OVERLAP:
      dom = ea ? Box64{0, 0, 0, 0} : ma;
      break;
    default:   // OR, XOR
      if (ea && eb) dom = Box64{0, 0, 0, 0};
      else if (ea) dom = mb;
      else if (eb) dom = ma;
      else dom = Box64{std::min(ma.x1, mb.x1), std::min(ma.y1, mb.y1), std::max(ma.x2, mb.x2), std::max(ma.y2, mb.y2)};
  }
  if (dom.empty()) {
    if (area_out) *area_out = 0;
    if (c) set_empty(e, c);
    return VL_OK;
  }
  const bool area = area_out != nullptr;
  Outs o(e);
  Grid g = make_grid(dom, na + nb, erode > 0 ? e->erode_target : e->target, shape);
  const double t_start = g_debug ? now_ms() : 0.0;
  const Xform xr = erode > 0 ? Xform{erode, erode, erode, erode} : xa;
  VLCK(bin_and_process(e, g, top, a, na, xa, xr, b, nb, area, erode, o));
  if (g_debug) fprintf(stderr, "[vl] op %d na=%lld nb=%lld level0 %.2f ms\n", top, (long long)na, (long long)nb,
                       now_ms() - t_start);
  if (area) {
    *area_out = o.area;
    return VL_OK;
  }
  return finish_op(e, g, o, c, t_start);
}

struct OpTimer {
  vl_engine* e;
  explicit OpTimer(vl_engine* en) : e(en) {
    if (e->cap.mode != CAP_CAPTURE) (void)hipEventRecord(e->ev0, e->st);
  }
  ~OpTimer() {
    if (e->cap.mode == CAP_CAPTURE) return;   // no host waits inside a capture
    (void)hipEventRecord(e->ev1, e->st);
    if (hipEventSynchronize(e->ev1) == hipSuccess) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e->ev0, e->ev1) == hipSuccess) e->last_ms = ms;
    }
  }
};

int check_layers(const vl_engine* e, const vl_layer* a, const vl_layer* b, const vl_layer* c) {
  if (!e) return fail(VL_EINVAL, "null engine");
  if (!a || !c) return fail(VL_EINVAL, "null layer");
  if (a->e != e->root || (b && b->e != e->root) || c->e != e->root)
    return fail(VL_EINVAL, "layer belongs to another engine");
  return VL_OK;
}

// Copies a layer (used when the output aliases an input).
int snapshot(vl_engine* e, const vl_layer* l, Scratch& s, const vl_box** out) {
  if (l->n == 0) { *out = nullptr; return VL_OK; }
  vl_box* p = s.get<vl_box>(l->n);
  if (!p) return fail(VL_ENOMEM, "scratch");
  HIPCK(hipMemcpyAsync(p, l->d, l->n * sizeof(vl_box), hipMemcpyDeviceToDevice, e->st));
  *out = p;
  return VL_OK;
}

// Bounding box of a layer (cached on the layer).
int layer_mbr(vl_engine* e, const vl_layer* l, Box64* m) {
  if (!l->mbr_ok) {
    Box64 ma{INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN}, mb;
    double sums[4] = {0, 0, 0, 0};
    if (l->n > 0) VLCK(mbr2(e, l->d, l->n, Xform{0, 0, 0, 0}, nullptr, 0, &ma, &mb, sums));
    l->mbr[0] = ma.x1; l->mbr[1] = ma.y1; l->mbr[2] = ma.x2; l->mbr[3] = ma.y2;
    l->sum_w = sums[0];
    l->sum_h = sums[1];
    l->mbr_ok = true;
  }
  *m = Box64{l->mbr[0], l->mbr[1], l->mbr[2], l->mbr[3]};
  return VL_OK;
}

// A box containing every box of the layer: the cached bounding box, else a
// producer's bound for the current version, else the bounding box reduced now.
int layer_extent(vl_engine* e, const vl_layer* l, Box64* m) {
  if (!l->mbr_ok && l->bound_version == l->version) {
    *m = Box64{l->bound[0], l->bound[1], l->bound[2], l->bound[3]};
    return VL_OK;
  }
  return layer_mbr(e, l, m);
}

// Tile shape suited to a layer's boxes: from its summed widths and heights
// when its bounding box is cached, else a producer's hint for its current
// version, else square.
double layer_shape(const vl_engine* e, const vl_layer* l) {
  if (l->mbr_ok) return tile_shape(l->sum_w, l->sum_h, e->tile_shape);
  if (l->bound_version == l->version && e->tile_shape > 0) return l->shape_hint;
  return 1.0;
}

// ---- fused operator groups ------------------------------------------------------
// Level 0 of a fused group: the fused tile kernel's tiers (10, 20 and 42 boxes
// per window), then x-sub-column splits of the tiles above the second tier's
// capacity (windows grown by the halo), recursively.
int process_tiles_fused(vl_engine* e, const TileDesc* d_desc, int ntiles, const vl_box* d_entries,
                        uint64_t nentries, const FusedSpec& fs, Outs* const* outs, int depth) {
  if (ntiles <= 0) return VL_OK;
  const int nout = fs.nout;
  Scratch s(e);
  constexpr int NT = 3;
  const int tiers[NT] = {FUSE_NMAX_TINY, FUSE_NMAX_SMALL, FUSE_NMAX_LARGE};
  uint32_t* small = s.get<uint32_t>(4);   // overflow counts of the tiers
  unsigned long long* cnt = s.get<unsigned long long>(3 * (size_t)FUSE_MAX_OUT);
  uint32_t* ovfl[NT];
  for (int t = 0; t < NT; t++) ovfl[t] = s.get<uint32_t>(ntiles);
  uint64_t* tab_off = s.get<uint64_t>((size_t)nout * ntiles);
  uint32_t* tab_n = s.get<uint32_t>((size_t)nout * 2 * ntiles);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  int G[NT];
  int nregions = 0;
  for (int t = 0; t < NT; t++) {
    int mw = tile_fused_max_wg(tiers[t]);
    if (e->max_wg_cap > 0) mw = std::min(mw, e->max_wg_cap);
    G[t] = std::max(1, std::min(ntiles, mw));
    nregions += G[t];
  }
  // stream size per output: the expected strips plus one chunk per
  // workgroup; regrown as in the single-operator tile stage.  The streams of
  // all outputs exist at once: when they would take more than half of the
  // device memory left (a full-chip layer), the group runs operator by
  // operator instead, which holds one output's at a time.
  uint64_t scap[FUSE_MAX_OUT];
  for (int k = 0; k < nout; k++)
    scap[k] = (uint64_t)(e->fuse_region_scale * (double)nentries) + (uint64_t)nregions * STRIP_CHUNK + 1024;
  const double budget = 0.5 * (double)e->pool.available();
  auto fits = [&]() {
    double need = 0;
    for (int k = 0; k < nout; k++) need += (double)scap[k] * (sizeof(vl_box) + 1);
    return need <= budget;
  };
  if (e->cap.mode == CAP_EAGER && !fits()) return FUSE_DECLINED;
  std::unique_ptr<Scratch> rs;
  vl_box* scr[FUSE_MAX_OUT] = {};
  uint8_t* scrT[FUSE_MAX_OUT] = {};
  unsigned long long hc[3 * FUSE_MAX_OUT];
  uint32_t nov = 0;
  const double t_pt = g_debug ? now_ms() : 0.0;
  for (int attempt = 0; attempt < 4; attempt++) {
    rs.reset(new Scratch(e));
    for (int k = 0; k < nout; k++) {
      scr[k] = rs->get<vl_box>(scap[k]);
      scrT[k] = rs->get<uint8_t>(scap[k]);
    }
    if (rs->oom) return fail(VL_ENOMEM, "fused strip scratch");
    launch_fill_u32(small, 0u, 4, e->st);
    launch_fill_u64((uint64_t*)cnt, 0ull, 3 * (size_t)FUSE_MAX_OUT, e->st);
    VLCK(clear_err(e));
    FusedOut fo;
    memset(&fo, 0, sizeof(fo));
    for (int k = 0; k < nout; k++) {
      fo.strips[k] = scr[k];
      fo.tag[k] = scrT[k];
      fo.tab_off[k] = tab_off + (size_t)k * ntiles;
      fo.tab_n[k] = tab_n + (size_t)k * 2 * ntiles;
      fo.cnt[k] = cnt + 3 * k;
      fo.stream_cap[k] = scap[k];
    }
    fo.chunk = STRIP_CHUNK;
    fo.ntab = ntiles;
    fo.err = e->d_err;
    hipEvent_t pa = prof_mark(e);
    for (int t = 0; t < NT; t++) {
      fo.list = t ? ovfl[t - 1] : nullptr;        // tiles above the previous tier's capacity
      fo.list_n = t ? small + (t - 1) : nullptr;
      fo.ovf = ovfl[t];
      fo.ovf_n = small + t;
      launch_tile_fused(tiers[t], d_desc, ntiles, d_entries, fs, fo, G[t], e->st);
    }
    HIPCK(hipGetLastError());
    hipEvent_t pb = prof_mark(e);
    uint32_t err = 0;
    {
      // per output: strips and seam strips checked, the stream's demand not
      RB r[2 * FUSE_MAX_OUT + 2];
      int nr = 0;
      for (int k = 0; k < FUSE_MAX_OUT; k++) {
        r[nr++] = RB{hc + 3 * k, cnt + 3 * k, 16, true};
        r[nr++] = RB{hc + 3 * k + 2, cnt + 3 * k + 2, 8, false};
      }
      r[nr++] = RB{&err, e->d_err, 4, true};
      r[nr++] = RB{&nov, small + NT - 1, 4, true};
      VLCK(sync_read(e, r, nr));
    }
    {
      double strips = 0.0;
      for (int q = 0; q < nout; q++) strips += (double)hc[3 * q];
      prof_add(e, pa, pb, KS_FUSED, NT, 16.0 * (double)nentries + 16.0 * strips + 48.0 * ntiles);
    }
    if (err & ERR_OVERFLOW_CAP) {
      if (attempt == 3) return fail(VL_EINTERNAL, "fused output capacity still short after resize");
      for (int k = 0; k < nout; k++) {
        const uint64_t demand = hc[3 * k + 2];
        if (demand <= scap[k]) continue;
        scap[k] = std::max<uint64_t>(2 * scap[k], demand + demand / 4 + (uint64_t)nregions * STRIP_CHUNK + 1024);
        e->fuse_region_scale =
            std::max(e->fuse_region_scale, 1.25 * (double)demand / std::max<double>(1.0, nentries));
      }
      if (e->cap.mode == CAP_EAGER && !fits()) return FUSE_DECLINED;
      if (g_debug) fprintf(stderr, "[vl]  fused strip stream short: retry\n");
      continue;
    }
    break;
  }
  // every output into tile order (final and seam streams)
  for (int k = 0; k < nout; k++) {
    const uint64_t total = hc[3 * k], nsideT = hc[3 * k + 1];
    if (total == 0) continue;
    Outs& o = *outs[k];
    const uint64_t nfinal = total - nsideT;
    VLCK(o.out.ensure(o.nout + nfinal + 1, o.nout));
    VLCK(o.side.ensure(o.nside + nsideT + 1, o.nside));
    VLCK(o.side_tag.ensure(o.nside + nsideT + 1, o.nside));
    Scratch ks(e);
    uint64_t* dst = ks.get<uint64_t>(2 * (size_t)ntiles + 2);
    void* stmp = ks.get<char>(scan_tmp_bytes(ntiles));
    if (ks.oom) return fail(VL_ENOMEM, "scratch");
    const uint32_t* tn = tab_n + (size_t)k * 2 * ntiles;
    scan_u32_to_u64(tn, dst, ntiles, nullptr, stmp, e->st);
    scan_u32_to_u64(tn + ntiles, dst + ntiles, ntiles, nullptr, stmp, e->st);
    launch_tile_reorder(ntiles, tab_off + (size_t)k * ntiles, tn, ntiles, dst, dst + ntiles, scr[k], scrT[k],
                        o.out.p, o.nout, o.side.p, o.side_tag.p, o.nside, e->st);
    HIPCK(hipGetLastError());
    o.nout += nfinal;
    o.nside += nsideT;
  }
  rs.reset();
  if (g_debug) {
    fprintf(stderr, "[vl]  fused tiles=%d entries=%llu depth=%d overflow=%u %.2f ms; strips", ntiles,
            (unsigned long long)nentries, depth, nov, now_ms() - t_pt);
    for (int k = 0; k < nout; k++) fprintf(stderr, " %llu/%llu", hc[3 * k], hc[3 * k + 1]);
    fprintf(stderr, "\n");
  }
  if (nov > 0) {
    // a sub-column still above the largest tier holds more boxes over one
    // x-position than further x-splits separate: the group runs operator by
    // operator instead (the single-operator path has the dense kernel).
    // Rows are not split: every tile of a grid row must share its row
    // borders, or a strip crossing a column border would reach the seam
    // levels on one side and the x-stitch on the other.
    if (depth >= 1) return fail(VL_EINTERNAL, "%u fused sub-columns above the largest tier", nov);
    Scratch ss(e);
    SplitTiles sp;
    VLCK(split_tiles(e, d_desc, ovfl[NT - 1], nov, d_entries, fs.halo, FUSE_NMAX_SMALL / 2, FUSE_NMAX_LARGE, depth, ss,
                     &sp));
    VLCK(process_tiles_fused(e, sp.desc, sp.n, sp.entries, sp.total, fs, outs, depth + 1));
  }
  return VL_OK;
}


// Outputs c[k] = fs.tt[k](S, morphological images of S) of a fused group.
int run_fused(vl_engine* e, const vl_layer* S, const vl_layer* T, const FusedSpec& fs_in, vl_layer* const* c_in) {
  if (S->n == 0) {
    // with a second layer, outputs such as S OR T are not empty: the
    // operators run one by one
    if (T) return FUSE_DECLINED;
    for (int k = 0; k < fs_in.nout; k++) set_empty(e, c_in[k]);
    return VL_OK;
  }
  // Every output is evaluated per tile.  Programs never group an operator
  // whose truth table is 0 (an AND of a morphological difference with its
  // own source, say): such operators are checks, and they run on the
  // materialised layers instead.
  const FusedSpec& fs = fs_in;
  vl_layer* const* c = c_in;
  const int nout = fs.nout;
  if (nout <= 0 || nout > FUSE_MAX_OUT) return fail(VL_EINVAL, "fused group with %d outputs", nout);
  Box64 m;
  VLCK(layer_mbr(e, S, &m));
  // domain: MBR(S) grown by the largest GROW reach on each side (erosion and
  // the Boolean outputs stay inside S or its grown images)
  int64_t ge = 0, gn = 0, gw = 0, gs = 0, de = 0;
  for (int k = 0; k < fs.nmorph; k++) {
    if (fs.kind[k] == MORPH_GROW) {
      ge = std::max<int64_t>(ge, fs.arg[k][0]);
      gn = std::max<int64_t>(gn, fs.arg[k][1]);
      gw = std::max<int64_t>(gw, fs.arg[k][2]);
      gs = std::max<int64_t>(gs, fs.arg[k][3]);
    } else {
      de = std::max<int64_t>(de, fs.arg[k][0]);
    }
  }
  Box64 dom{m.x1 - gw, m.y1 - gs, m.x2 + ge, m.y2 + gn};
  if (T && T->n > 0) {   // outputs may cover parts of the second layer outside S
    Box64 mt;
    VLCK(layer_mbr(e, T, &mt));
    dom = Box64{std::min(dom.x1, mt.x1), std::min(dom.y1, mt.y1), std::max(dom.x2, mt.x2), std::max(dom.y2, mt.y2)};
  }
  const int64_t h = fs.halo;
  if (dom.x1 - h < INT32_MIN || dom.y1 - h < INT32_MIN || dom.x2 + h > INT32_MAX || dom.y2 + h > INT32_MAX)
    return FUSE_DECLINED;
  double shape = 1.0;
  if (e->tile_shape) {
    double sw = S->sum_w, sh = S->sum_h;
    if (T && T->mbr_ok) {
      sw += T->sum_w;
      sh += T->sum_h;
    }
    shape = tile_shape(sw, sh, e->tile_shape);
  }
  const Grid g = make_grid(dom, S->n + (T ? T->n : 0), e->fuse_target, shape);
  // windows must stay small against the tiles (each box meets a few windows),
  // and below 2^24 - 1 DBU wide and tall (the kernel sorts coordinates relative
  // to the window corner in 24 bits)
  if (g.tw < 4 * h || g.th < 4 * h || g.tw < 4 * de || g.th < 4 * de) return FUSE_DECLINED;
  if (std::max(g.tw, g.th) + 2 * h >= (int64_t)(1 << 24) - 1) return FUSE_DECLINED;
  const double t_start = g_debug ? now_ms() : 0.0;
  std::vector<std::unique_ptr<Outs>> outs;
  Outs* po[FUSE_MAX_OUT] = {};
  for (int k = 0; k < nout; k++) {
    outs.emplace_back(new Outs(e));
    po[k] = outs.back().get();
  }
  {
    const int64_t nt = (int64_t)g.cols * g.rows;
    Scratch keep(e);
    vl_box* entries = nullptr;
    uint64_t* start = nullptr;
    uint64_t total = 0;
    const int32_t R = fs.halo;
    // a fused grid keeps its tile windows (the kernel's coordinate range
    // depends on them): a layer whose boxes span too many tiles is declined
    Grid gb = g;
    // S's boxes by their extent grown by the halo (its images reach that far),
    // the second layer's by their extent (its image is the layer itself)
    const int brc = bin_entries(e, gb, S->d, S->n, Xform{0, 0, 0, 0}, Xform{R, R, R, R}, T ? T->d : nullptr,
                                T ? T->n : 0, keep, &entries, &start, &total, false);
    if (brc == BIN_TOO_MANY) return FUSE_DECLINED;
    VLCK(brc);
    TileDesc* desc = keep.get<TileDesc>(nt);
    if (!desc) return fail(VL_ENOMEM, "scratch for %lld tiles", (long long)nt);
    launch_build_desc(g, start, TOP_UNION, desc, e->st);
    HIPCK(hipGetLastError());
    const int rc = process_tiles_fused(e, desc, (int)nt, entries, total, fs, po, 0);
    if (rc == VL_EINTERNAL) {
      if (g_debug) fprintf(stderr, "[vl] fused group declined: %s\n", g_err.c_str());
      return FUSE_DECLINED;
    }
    VLCK(rc);
  }
  if (g_debug) fprintf(stderr, "[vl] fused level0 n=%lld outputs=%d %.2f ms\n", (long long)S->n, nout,
                       now_ms() - t_start);
  for (int k = 0; k < nout; k++) {
    VLCK(finish_op(e, g, *outs[k], c[k], t_start));
    // every output lies in the group's domain
    c[k]->bound[0] = dom.x1; c[k]->bound[1] = dom.y1; c[k]->bound[2] = dom.x2; c[k]->bound[3] = dom.y2;
    c[k]->bound_version = c[k]->version;
    c[k]->shape_hint = shape;   // outputs follow the outlines of the source's boxes
  }
  return VL_OK;
}

int binary(vl_engine* e, int top, const vl_layer* a, const vl_layer* b, vl_layer* c, Xform xa = Xform{0, 0, 0, 0}) {
  VLCK(check_layers(e, a, b, c));
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  Scratch s(e);
  Box64 ka, kb;
  VLCK(layer_mbr(e, a, &ka));
  if (b) VLCK(layer_mbr(e, b, &kb));
  const vl_box* pa = a->d;
  const vl_box* pb = b ? b->d : nullptr;
  if (c == a) VLCK(snapshot(e, a, s, &pa));
  if (b && c == b) VLCK(snapshot(e, b, s, &pb));
  // tiles shaped like the inputs' boxes (B's count only where it is binned)
  double shape = 1.0;
  if (e->tile_shape) {
    const bool with_b = b && (two_layer(top) || top == TOP_OR);
    shape = tile_shape(a->sum_w + (with_b ? b->sum_w : 0.0), a->sum_h + (with_b ? b->sum_h : 0.0), e->tile_shape);
  }
  return run_op(e, top, pa, a->n, xa, pb, b ? b->n : 0, c, nullptr, 0, &ka, b ? &kb : nullptr, shape);
}

// One input of a semi-join: boxes and the transform applied to them.
struct Side {
  const vl_box* d;
  int64_t n;
  Xform x;
};

// The boxes of p (transformed) that overlap a box of the binned input
// (entries/start over grid g) with positive area, compacted, transformed,
// into *out (allocated from keep); *m = their number.
int probe_reduce(vl_engine* e, const Side& p, const Grid& g, const vl_box* entries, const uint64_t* start,
                 Scratch& keep, vl_box** out, uint64_t* m) {
  Scratch s(e);
  uint32_t* hit = s.get<uint32_t>(1);
  uint32_t* mark = s.get<uint32_t>((size_t)p.n + 1);
  uint32_t* pos = s.get<uint32_t>((size_t)p.n + 1);
  void* stmp = s.get<char>(scan_tmp_bytes((size_t)p.n + 1));
  if (s.oom) return fail(VL_ENOMEM, "semi-join scratch");
  launch_fill_u32(hit, 0u, 1, e->st);
  launch_and_probe(p.d, p.n, p.x, g, entries, start, hit, mark, e->st);
  scan_exclusive_u32(mark, pos, (size_t)p.n, pos + p.n, stmp, e->st);
  uint32_t k = 0;
  VLCK(readback(e, &k, pos + p.n, sizeof(k)));
  *m = k;
  *out = keep.get<vl_box>((size_t)k + 1);
  if (!*out) return fail(VL_ENOMEM, "semi-join output");
  launch_compact_marked(p.d, p.n, p.x, mark, pos, *out, e->st);
  HIPCK(hipGetLastError());
  return VL_OK;
}

// C = A AND B (A's boxes transformed by A.x: a GROW of A's union) planned for
// a result that is expected to be empty (an AND read only by ASSERT_ZERO, a
// DRC check); ma, mb: bounding boxes of the transformed inputs.  The smaller
// input is binned alone and every box of the other is tested against the
// binned boxes of the tiles it meets (k_and_probe).  No positive-area
// overlap: C is empty and nothing else runs.  Otherwise the inputs are
// reduced to the boxes that take part in an overlap (the probed boxes that
// overlap a binned one, then the binned input's boxes that overlap one of
// those) and the AND of the reduced inputs, which is the same point set, is
// computed in full: a failing check costs a few passes over the inputs plus
// an AND of the violating boxes.  The caller holds the engine.
int and_check_core(vl_engine* e, const Side& A, const Side& B, const Box64& ma, const Box64& mb, vl_layer* c,
                   double shape_a = 1.0, double shape_b = 1.0) {
  const Box64 dom{std::max(ma.x1, mb.x1), std::max(ma.y1, mb.y1), std::min(ma.x2, mb.x2), std::min(ma.y2, mb.y2)};
  if (A.n == 0 || B.n == 0 || dom.empty()) return set_empty(e, c);
  const bool build_a = A.n <= B.n;
  const Side& build = build_a ? A : B;
  const Side& probe = build_a ? B : A;
  Grid g = make_grid(dom, build.n, e->check_target, build_a ? shape_a : shape_b);
  Scratch keep(e);
  vl_box* entries = nullptr;
  uint64_t* start = nullptr;
  uint64_t total = 0;
  VLCK(bin_entries(e, g, build.d, build.n, build.x, build.x, nullptr, 0, keep, &entries, &start, &total));
  uint32_t* hit = keep.get<uint32_t>(1);
  if (!hit) return fail(VL_ENOMEM, "scratch");
  hipEvent_t pa = prof_mark(e);
  launch_fill_u32(hit, 0u, 1, e->st);
  launch_and_probe(probe.d, probe.n, probe.x, g, entries, start, hit, nullptr, e->st);
  HIPCK(hipGetLastError());
  hipEvent_t pb = prof_mark(e);
  // algorithmic bytes: every probe box and every binned entry read once
  prof_add(e, pa, pb, KS_TILE, 1, 16.0 * (double)probe.n + 16.0 * (double)total);
  uint32_t h = 0;
  VLCK(readback(e, &h, hit, sizeof(h)));
  if (g_debug) fprintf(stderr, "[vl] and_check build=%lld probe=%lld entries=%llu hit=%u\n", (long long)build.n,
                       (long long)probe.n, (unsigned long long)total, h);
  if (h == 0) return set_empty(e, c);
  // semi-join reduction in both directions (the reduced boxes carry their
  // transform)
  vl_box* p2 = nullptr;
  uint64_t np2 = 0;
  VLCK(probe_reduce(e, probe, g, entries, start, keep, &p2, &np2));
  Grid g2 = make_grid(dom, (int64_t)np2, e->check_target, build_a ? shape_b : shape_a);
  vl_box* entries2 = nullptr;
  uint64_t* start2 = nullptr;
  uint64_t total2 = 0;
  VLCK(bin_entries(e, g2, p2, (int64_t)np2, Xform{0, 0, 0, 0}, Xform{0, 0, 0, 0}, nullptr, 0, keep, &entries2,
                   &start2, &total2));
  vl_box* b2 = nullptr;
  uint64_t nb2 = 0;
  VLCK(probe_reduce(e, build, g2, entries2, start2, keep, &b2, &nb2));
  if (g_debug) fprintf(stderr, "[vl] and_check reduced to %llu x %llu boxes\n", (unsigned long long)np2,
                       (unsigned long long)nb2);
  const vl_box* ra = build_a ? b2 : p2;
  const vl_box* rb = build_a ? p2 : b2;
  const int64_t na = (int64_t)(build_a ? nb2 : np2), nb = (int64_t)(build_a ? np2 : nb2);
  return run_op(e, TOP_AND, ra, na, Xform{0, 0, 0, 0}, rb, nb, c, nullptr);
}

// and_check_core on two layers (no transform; the layers' cached bounding boxes).
int and_check(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c) {
  VLCK(check_layers(e, a, b, c));
  if (!b) return fail(VL_EINVAL, "null layer");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  // a box containing each input is enough here (the test's grid only has to
  // cover every overlap)
  Box64 ka, kb;
  VLCK(layer_extent(e, a, &ka));
  VLCK(layer_extent(e, b, &kb));
  return and_check_core(e, Side{a->d, a->n, Xform{0, 0, 0, 0}}, Side{b->d, b->n, Xform{0, 0, 0, 0}}, ka, kb, c,
                        layer_shape(e, a), layer_shape(e, b));
}

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t at;
  memset(&at, 0, sizeof(at));
  hipError_t r = hipPointerGetAttributes(&at, p);
  if (r != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return at.type == hipMemoryTypeDevice;
}

int upload_tmp(vl_engine* e, Scratch& s, const void* host, size_t bytes, void** dev) {
  *dev = s.get<char>(bytes);
  if (!*dev) return fail(VL_ENOMEM, "scratch");
  return upload(e, *dev, host, bytes);
}

vl_engine* g_default = nullptr;
std::mutex g_default_mu;

vl_engine* default_engine() {
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default) {
    if (vl_engine_create(0, &g_default) != VL_OK) {
      fprintf(stderr, "vlsi_layout_amd: %s\n", g_err.c_str());
      abort();
    }
  }
  return g_default;
}

}  // namespace

// ===========================================================================
// C ABI
hipEvent_t vl::engine_prof_mark(vl_engine* e) { return prof_mark(e); }
void vl::engine_prof_add(vl_engine* e, hipEvent_t a, hipEvent_t b, int kind, int64_t launches, double bytes) {
  prof_add(e, a, b, kind, launches, bytes);
}

extern "C" {

const char* vl_last_error(void) { return g_err.c_str(); }

// Creates an engine on a device (the caller has checked the device).
static int engine_new(int device, vl_engine** out) {
  HIPCK(hipSetDevice(device));
  vl_engine* e = new vl_engine();
  e->device = device;
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
      hipMalloc((void**)&e->d_err, 256) != hipSuccess ||
      hipHostMalloc((void**)&e->h_stage, 4096, hipHostMallocDefault) != hipSuccess) {
    delete e;
    return fail(VL_EHIP, "engine initialisation failed");
  }
  *out = e;
  return VL_OK;
}

static void engine_release(vl_engine* e);

int vl_engine_create(int device, vl_engine** out) {
  if (!out) return fail(VL_EINVAL, "null out");
  *out = nullptr;
  int ndev = 0;
  HIPCK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(VL_EINVAL, "device %d not present (%d devices)", device, ndev);
  vl_engine* e = nullptr;
  VLCK(engine_new(device, &e));
  // set-up of the static scratch of vl_event_sort_noalloc on this device
  // (the device is current after engine_new)
  const int rs = noalloc_setup(device);
  if (rs != VL_OK) {
    engine_release(e);
    return rs;
  }
  if (const char* v = getenv("VL_TILE_NMAX")) e->nmax = atoi(v);
  if (e->nmax != 63 && e->nmax != 127) e->nmax = 63;
  *out = e;
  return VL_OK;
}

static void engine_release(vl_engine* e) {
  if (e->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->st);
  e->pool.trim();
  (void)hipFree(e->d_err);
  (void)hipHostFree(e->h_stage);
  (void)hipEventDestroy(e->ev0);
  (void)hipEventDestroy(e->ev1);
  (void)prof_resolve(e);
  for (hipEvent_t ev : e->ev_free) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(e->st);
  delete e;
}

vl::EngineLock::~EngineLock() {
  if (locked_) e_->mu.unlock();
}

int vl::EngineLock::begin() {
  if (!locked_) {
    e_->mu.lock();
    locked_ = true;
  }
  HIPCK(hipSetDevice(e_->device));
  return VL_OK;
}

hipStream_t vl::EngineLock::stream() const { return e_->st; }
int vl::EngineLock::device() const { return e_->device; }

int vl_engine_destroy(vl_engine* e) {
  if (!e) return VL_OK;
  engine_release(e);
  return VL_OK;
}

int vl_engine_set_param(vl_engine* e, const char* key, int64_t v) {
  if (!e || !key) return fail(VL_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  std::string k(key);
  if (k == "tile_nmax") {
    if (v != 63 && v != 127) return fail(VL_EINVAL, "tile_nmax must be 63 or 127");
    e->nmax = (int)v;
  } else if (k == "tile_target") {
    if (v < 4 || v > 4096) return fail(VL_EINVAL, "tile_target out of range");
    e->target = (int)v;
  } else if (k == "erode_target") {
    if (v < 4 || v > 4096) return fail(VL_EINVAL, "erode_target out of range");
    e->erode_target = (int)v;
  } else if (k == "level_factor") {
    if (v < 2 || v > 4096) return fail(VL_EINVAL, "level_factor out of range");
    e->level_factor = (int)v;
  } else if (k == "profile") {
    VLCK(prof_resolve(e));
    e->prof = v != 0;
    for (auto& s : e->kst) s = vl_engine::KStat();
  } else if (k == "max_wg") {
    e->max_wg_cap = (int)std::max<int64_t>(0, v);
  } else if (k == "erode_kernel") {
    e->erode_kernel = v != 0;
  } else if (k == "fuse") {
    if (v < 0 || v > 2) return fail(VL_EINVAL, "fuse must be 0, 1 or 2");
    e->fuse = (int)v;
  } else if (k == "fuse2") {
    e->fuse2 = v != 0;
  } else if (k == "fuse_target") {
    if (v < 2 || v > 4096) return fail(VL_EINVAL, "fuse_target out of range");
    e->fuse_target = (int)v;
  } else if (k == "bin_mode") {
    if (v < 0 || v > 2) return fail(VL_EINVAL, "bin_mode must be 0, 1 or 2");
    e->bin_mode = (int)v;
  } else if (k == "workers") {
    if (v < 1 || v > 16) return fail(VL_EINVAL, "workers must be 1 to 16");
    e->workers = (int)v;
  } else if (k == "radix_bits") {
    if (v != 8 && v != 11) return fail(VL_EINVAL, "radix_bits must be 8 or 11");
    e->radix_bits = (int)v;
  } else if (k == "and_check") {
    e->and_check = v != 0;
  } else if (k == "tile_shape") {
    if (v < 0 || v > 1024) return fail(VL_EINVAL, "tile_shape out of range (0..1024)");
    e->tile_shape = (int)v;
  } else if (k == "check_target") {
    if (v < 1 || v > 4096) return fail(VL_EINVAL, "check_target out of range");
    e->check_target = (int)v;
  } else {
    return fail(VL_EINVAL, "unknown parameter '%s'", key);
  }
  return VL_OK;
}

int vl_engine_get_param(vl_engine* e, const char* key, int64_t* v) {
  if (!e || !key || !v) return fail(VL_EINVAL, "null argument");
  std::string k(key);
  if (k == "tile_nmax") *v = e->nmax;
  else if (k == "tile_target") *v = e->target;
  else if (k == "erode_target") *v = e->erode_target;
  else if (k == "level_factor") *v = e->level_factor;
  else if (k == "max_wg") *v = e->max_wg_cap;
  else if (k == "erode_kernel") *v = e->erode_kernel;
  else if (k == "fuse") *v = e->fuse;
  else if (k == "fuse_target") *v = e->fuse_target;
  else if (k == "fuse2") *v = e->fuse2;
  else if (k == "bin_mode") *v = e->bin_mode;
  else if (k == "radix_bits") *v = e->radix_bits;
  else if (k == "workers") *v = e->workers;
  else if (k == "and_check") *v = e->and_check;
  else if (k == "check_target") *v = e->check_target;
  else if (k == "tile_shape") *v = e->tile_shape;
  else if (k == "fused_groups") *v = e->fused_groups;
  else return fail(VL_EINVAL, "unknown parameter '%s'", key);
  return VL_OK;
}

int vl_engine_sync(vl_engine* e) {
  if (!e) return fail(VL_EINVAL, "null engine");
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

int vl_engine_kernel_stats(vl_engine* e, const char* kernel, int64_t* launches, double* total_ms, double* bytes) {
  if (!e || !kernel || !launches || !total_ms || !bytes) return fail(VL_EINVAL, "null argument");
  std::string k(kernel);
  int i = k == "tile" ? KS_TILE : k == "bin" ? KS_BIN : k == "sort" ? KS_SORT : k == "fused" ? KS_FUSED
        : k == "snap" ? KS_SNAP : -1;
  if (i < 0) return fail(VL_EINVAL, "unknown kernel '%s'", kernel);
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  VLCK(prof_resolve(e));
  *launches = e->kst[i].launches;
  *total_ms = e->kst[i].ms;
  *bytes = e->kst[i].bytes;
  return VL_OK;
}

int vl_engine_last_time_ms(vl_engine* e, double* ms) {
  if (!e || !ms) return fail(VL_EINVAL, "null argument");
  *ms = e->last_ms;
  return VL_OK;
}

int vl_layer_create(vl_engine* e, vl_layer** out) {
  if (!e || !out) return fail(VL_EINVAL, "null argument");
  vl_layer* l = new vl_layer();
  l->e = e;
  e->refs.fetch_add(1);
  *out = l;
  return VL_OK;
}

static void layer_release(vl_layer* l) {
  if (l->refs.fetch_sub(1) != 1) return;
  vl_engine* e = l->e;
  if (l->d) {
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    e->pool.put(l->d, l->bytes);
  }
  delete l;
  engine_release(e);
}

int vl_layer_destroy(vl_layer* l) {
  if (!l) return VL_OK;
  layer_release(l);
  return VL_OK;
}

static int layer_reserve(vl_layer* l, int64_t n) {
  size_t need = (size_t)std::max<int64_t>(n, 1) * sizeof(vl_box);
  if (l->d && l->bytes >= need) return VL_OK;
  if (l->d) l->e->pool.put(l->d, l->bytes);
  l->d = nullptr;
  size_t r = 0;
  void* p = l->e->pool.get(need, &r);
  if (!p) return fail(VL_ENOMEM, "layer allocation of %zu bytes failed", need);
  l->d = (vl_box*)p;
  l->bytes = r;
  return VL_OK;
}

static int validate_host(const vl_box* h, int64_t n) {
  for (int64_t i = 0; i < n; i++)
    if (!(h[i].x1 < h[i].x2 && h[i].y1 < h[i].y2))
      return fail(VL_EINVAL, "invalid box %lld: (%d,%d,%d,%d)", (long long)i, h[i].x1, h[i].y1, h[i].x2, h[i].y2);
  return VL_OK;
}

int vl_layer_upload(vl_layer* l, const vl_box* host, int64_t n) {
  if (!l || (n > 0 && !host) || n < 0) return fail(VL_EINVAL, "bad argument");
  VLCK(validate_host(host, n));
  vl_engine* e = l->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  VLCK(layer_reserve(l, n));
  if (n > 0) HIPCK(hipMemcpyAsync(l->d, host, n * sizeof(vl_box), hipMemcpyHostToDevice, e->st));
  HIPCK(hipStreamSynchronize(e->st));
  l->n = n;
  l->mbr_ok = false;
  l->version++;
  return VL_OK;
}

int vl_layer_upload_device(vl_layer* l, const vl_box* dev, int64_t n) {
  if (!l || (n > 0 && !dev) || n < 0) return fail(VL_EINVAL, "bad argument");
  vl_engine* e = l->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  VLCK(layer_reserve(l, n));
  if (n > 0) HIPCK(hipMemcpyAsync(l->d, dev, n * sizeof(vl_box), hipMemcpyDeviceToDevice, e->st));
  l->n = n;
  l->mbr_ok = false;
  l->version++;
  if (n > 0) {
    Box64 m;
    VLCK(layer_mbr(e, l, &m));   // validates the boxes
  }
  return VL_OK;
}

int vl_layer_download(const vl_layer* l, vl_box* host, int64_t cap, int64_t* n) {
  if (!l || !n) return fail(VL_EINVAL, "null argument");
  *n = l->n;
  if (cap < l->n) return fail(VL_ENOSPC, "capacity %lld < %lld boxes", (long long)cap, (long long)l->n);
  if (l->n == 0) return VL_OK;
  if (!host) return fail(VL_EINVAL, "null host buffer");
  vl_engine* e = l->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  HIPCK(hipMemcpyAsync(host, l->d, l->n * sizeof(vl_box), hipMemcpyDeviceToHost, e->st));
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

int vl_layer_size(const vl_layer* l, int64_t* n) {
  if (!l || !n) return fail(VL_EINVAL, "null argument");
  *n = l->n;
  return VL_OK;
}

int vl_layer_device_ptr(const vl_layer* l, const vl_box** dev) {
  if (!l || !dev) return fail(VL_EINVAL, "null argument");
  *dev = l->d;
  return VL_OK;
}

int vl_and(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c) {
  if (!b) return fail(VL_EINVAL, "null layer");
  return binary(e, TOP_AND, a, b, c);
}
int vl_not(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c) {
  if (!b) return fail(VL_EINVAL, "null layer");
  return binary(e, TOP_NOT, a, b, c);
}
int vl_or(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c) {
  if (!b) return fail(VL_EINVAL, "null layer");
  return binary(e, TOP_OR, a, b, c);
}
int vl_xor(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c) {
  if (!b) return fail(VL_EINVAL, "null layer");
  return binary(e, TOP_XOR, a, b, c);
}
int vl_merge(vl_engine* e, const vl_layer* a, vl_layer* c) { return binary(e, TOP_UNION, a, nullptr, c); }

int vl_self_and(vl_engine* e, const vl_layer* a, vl_layer* c) { return binary(e, TOP_OVERLAP, a, nullptr, c); }

// Pairs of A's boxes whose closed extents meet, found per tile (vl_pnum.hip):
// the box indices are binned by their closed extent (the tile range grown by
// one unit east and north) over MBR(A) grown the same way.
static int layer_pairs(vl_engine* e, const vl_layer* a, int mode, uint32_t* parent, uint32_t* touched,
                       unsigned long long sums_out[3]) {
  Box64 m;
  VLCK(layer_mbr(e, a, &m));
  const Box64 dom{m.x1, m.y1, std::min<int64_t>(m.x2 + 1, INT32_MAX), std::min<int64_t>(m.y2 + 1, INT32_MAX)};
  const Grid g = make_grid(dom, a->n, e->target);
  const int64_t nt = (int64_t)g.cols * g.rows;
  const Xform closed{1, 1, 0, 0};
  Scratch s(e);
  uint64_t* start = s.get<uint64_t>(2 * (size_t)nt + 1);
  uint32_t* cnt = s.get<uint32_t>(2 * (size_t)nt);
  void* stmp = s.get<char>(scan_tmp_bytes(2 * (size_t)nt));
  TileDesc* desc = s.get<TileDesc>((size_t)nt);
  unsigned long long* sums = s.get<unsigned long long>(3);
  if (s.oom) return fail(VL_ENOMEM, "scratch for %lld tiles", (long long)nt);
  launch_fill_u32(cnt, 0u, 2 * (size_t)nt, e->st);
  launch_bin_count(a->d, a->n, closed, nullptr, 0, g, cnt, e->bin_mode == 2, e->st);
  scan_u32_to_u64(cnt, start, 2 * (size_t)nt, start + 2 * nt, stmp, e->st);
  uint64_t total = 0;
  VLCK(readback(e, &total, start + 2 * nt, sizeof(total)));
  uint32_t* idx = s.get<uint32_t>(total + 1);
  if (!idx) return fail(VL_ENOMEM, "pair binning (%llu entries)", (unsigned long long)total);
  launch_bin_scatter_idx(a->d, a->n, closed, g, start, cnt, idx, e->bin_mode == 2, e->st);
  launch_build_desc(g, start, TOP_UNION, desc, e->st);
  launch_fill_u64((uint64_t*)sums, 0ull, 3, e->st);
  launch_pairs(desc, (int)nt, idx, a->d, g, mode, parent, touched, sums, e->st);
  HIPCK(hipGetLastError());
  return readback(e, sums_out, sums, 3 * sizeof(unsigned long long));
}

int vl_pnum(vl_engine* e, const vl_layer* a, int64_t* labels, int64_t* ncomp, int64_t* overlap_area,
            int64_t* touch_perimeter) {
  if (!e || !a || !ncomp || !overlap_area || !touch_perimeter || (a->n > 0 && !labels))
    return fail(VL_EINVAL, "null argument");
  if (a->e != e->root) return fail(VL_EINVAL, "layer belongs to another engine");
  if (a->n > (int64_t)UINT32_MAX - 1) return fail(VL_EINVAL, "more than 2^32 - 1 boxes");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  *ncomp = *overlap_area = *touch_perimeter = 0;
  const int64_t n = a->n;
  if (n == 0) return VL_OK;
  Scratch s(e);
  uint32_t* parent = s.get<uint32_t>(n);
  uint32_t* touched = s.get<uint32_t>(n);
  uint32_t* root = s.get<uint32_t>(n);
  uint32_t* is_root = s.get<uint32_t>(n + 1);
  uint32_t* rank = s.get<uint32_t>(n + 1);
  void* stmp = s.get<char>(scan_tmp_bytes((size_t)n));
  const bool dev = is_device_ptr(labels);
  int64_t* d_labels = dev ? labels : s.get<int64_t>(n);
  if (s.oom) return fail(VL_ENOMEM, "polygon numbering scratch");
  launch_uf_init(parent, touched, n, e->st);
  unsigned long long sums[3] = {0, 0, 0};
  VLCK(layer_pairs(e, a, 0, parent, touched, sums));
  launch_uf_roots(parent, touched, root, is_root, n, e->st);
  scan_exclusive_u32(is_root, rank, (size_t)n, rank + n, stmp, e->st);
  launch_uf_labels(root, touched, rank, d_labels, n, e->st);
  HIPCK(hipGetLastError());
  uint32_t nroots = 0;
  VLCK(readback(e, &nroots, rank + n, sizeof(nroots)));
  if (!dev) {
    HIPCK(hipMemcpyAsync(labels, d_labels, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, e->st));
    HIPCK(hipStreamSynchronize(e->st));
  }
  *ncomp = nroots;
  *overlap_area = (int64_t)sums[0];
  // the reference reports touch perimeters only when nothing overlaps
  *touch_perimeter = sums[0] ? 0 : (int64_t)sums[1];
  return VL_OK;
}

int vl_interaction_count(vl_engine* e, const vl_layer* a, int64_t* count) {
  if (!e || !a || !count) return fail(VL_EINVAL, "null argument");
  if (a->e != e->root) return fail(VL_EINVAL, "layer belongs to another engine");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  *count = 0;
  if (a->n < 2) return VL_OK;
  unsigned long long sums[3] = {0, 0, 0};
  VLCK(layer_pairs(e, a, 1, nullptr, nullptr, sums));
  *count = (int64_t)sums[2];
  return VL_OK;
}

int vl_grid(vl_engine* e, const vl_layer* a, int32_t rows, int32_t cols, int32_t overlap, vl_layer* c) {
  VLCK(check_layers(e, a, nullptr, c));
  if (rows < 1 || cols < 1) return fail(VL_EINVAL, "GRID: rows = %d, cols = %d must be positive", rows, cols);
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  Box64 m;
  VLCK(layer_mbr(e, a, &m));
  if (a->n == 0 || m.empty()) return fail(VL_EINVAL, "GRID: the input layer has no extent");
  const int64_t n = (int64_t)rows * cols;
  size_t r = 0;
  vl_box* d = (vl_box*)e->pool.get((size_t)n * sizeof(vl_box), &r);
  if (!d) return fail(VL_ENOMEM, "GRID output of %lld boxes", (long long)n);
  launch_grid(m.x1, m.y1, m.x2, m.y2, rows, cols, overlap, d, e->st);
  HIPCK(hipGetLastError());
  if (c->d) e->pool.put(c->d, c->bytes);
  c->d = d;
  c->bytes = r;
  c->n = n;
  c->mbr_ok = false;
  c->version++;
  return VL_OK;
}

int vl_extent(vl_engine* e, const vl_layer* a, vl_layer* c) {
  VLCK(check_layers(e, a, nullptr, c));
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  Box64 m;
  VLCK(layer_mbr(e, a, &m));
  if (a->n == 0 || m.empty()) return fail(VL_EINVAL, "EXTENT: the input layer has no extent");
  const vl_box b{(int32_t)m.x1, (int32_t)m.y1, (int32_t)m.x2, (int32_t)m.y2};
  VLCK(layer_reserve(c, 1));   // (a's boxes are not read again: c may be a)
  VLCK(upload(e, c->d, &b, sizeof(b)));
  HIPCK(hipStreamSynchronize(e->st));
  c->n = 1;
  c->mbr_ok = false;
  c->version++;
  return VL_OK;
}

int vl_grow(vl_engine* e, const vl_layer* a, int32_t east, int32_t north, int32_t west, int32_t south, vl_layer* c) {
  return binary(e, TOP_UNION, a, nullptr, c, Xform{east, north, west, south});
}

int vl_size(vl_engine* e, const vl_layer* a, int32_t d, vl_layer* c) {
  if (d >= 0) return vl_grow(e, a, d, d, d, d, c);
  VLCK(check_layers(e, a, nullptr, c));
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  if (a->n == 0) return set_empty(e, c);
  const int32_t r = (d == INT32_MIN) ? INT32_MAX : -d;
  OpTimer tm(e);
  Box64 ma;
  VLCK(layer_mbr(e, a, &ma));
  if (ma.x1 - 1 < INT32_MIN || ma.y1 - 1 < INT32_MIN || ma.x2 + 1 > INT32_MAX || ma.y2 + 1 > INT32_MAX)
    return fail(VL_EINVAL, "SIZE frame leaves the int32 range");
  Scratch s(e);
  const vl_box* pa = a->d;
  if (c == a) VLCK(snapshot(e, a, s, &pa));
  // Per-tile erosion when the tiles of the operator's grid are at least 4d
  // wide and tall (each box then meets only a few tile windows).
  const double shape = tile_shape(a->sum_w, a->sum_h, e->tile_shape);
  const Grid g = make_grid(ma, a->n, e->erode_target, shape);
  if (e->erode_kernel && g.tw >= 4 * (int64_t)r && g.th >= 4 * (int64_t)r) {
    const int rc = run_op(e, TOP_UNION, pa, a->n, Xform{0, 0, 0, 0}, nullptr, 0, c, nullptr, r, &ma, nullptr, shape);
    if (rc != ERODE_FALLBACK) return rc;
    g_err.clear();   // a tile no erosion kernel holds: the composition below
  }
  vl_box frame{(int32_t)(ma.x1 - 1), (int32_t)(ma.y1 - 1), (int32_t)(ma.x2 + 1), (int32_t)(ma.y2 + 1)};
  void* dframe = nullptr;
  VLCK(upload_tmp(e, s, &frame, sizeof(frame), &dframe));
  vl_layer t1, t2;
  t1.e = t2.e = e;
  int rc = run_op(e, TOP_NOT, (const vl_box*)dframe, 1, Xform{0, 0, 0, 0}, pa, a->n, &t1, nullptr);
  if (rc == VL_OK) rc = run_op(e, TOP_UNION, t1.d, t1.n, Xform{r, r, r, r}, nullptr, 0, &t2, nullptr);
  if (rc == VL_OK) rc = run_op(e, TOP_NOT, pa, a->n, Xform{0, 0, 0, 0}, t2.d, t2.n, c, nullptr);
  if (t1.d) e->pool.put(t1.d, t1.bytes);
  if (t2.d) e->pool.put(t2.d, t2.bytes);
  return rc;
}

int vl_assert_zero(vl_engine* e, const vl_layer* a, int abort_on_nonzero, int64_t* count) {
  if (!e || !a || !count) return fail(VL_EINVAL, "null argument");
  *count = a->n;
  if (a->n > 0 && abort_on_nonzero) return fail(VL_EASSERT, "ASSERT_ZERO: layer has %lld shapes", (long long)a->n);
  return VL_OK;
}

int vl_klee_area(vl_engine* e, const vl_layer* a, int64_t* area) {
  if (!e || !a || !area) return fail(VL_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  long long ar = 0;
  VLCK(run_op(e, TOP_UNION, a->d, a->n, Xform{0, 0, 0, 0}, nullptr, 0, nullptr, &ar));
  *area = ar;
  return VL_OK;
}

int vl_xor_empty(vl_engine* e, const vl_layer* a, const vl_layer* b, int* empty) {
  if (!e || !a || !b || !empty) return fail(VL_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  long long ar = 0;
  VLCK(run_op(e, TOP_XOR, a->d, a->n, Xform{0, 0, 0, 0}, b->d, b->n, nullptr, &ar));
  *empty = ar == 0;
  return VL_OK;
}

int vl_canonicalize(vl_engine* e, const vl_layer* a, vl_layer* c) {
  VLCK(check_layers(e, a, nullptr, c));
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  VLCK(vl_merge(e, a, c));
  const int64_t n = c->n;
  if (n <= 1) return VL_OK;
  Scratch s(e);
  uint64_t* k = s.get<uint64_t>(n);
  uint32_t* v = s.get<uint32_t>(n);
  uint64_t* k2 = s.get<uint64_t>(n);
  uint32_t* v2 = s.get<uint32_t>(n);
  vl_box* tmpb = s.get<vl_box>(n);
  void* rt = s.get<char>(radix_plan(n, e->radix_bits).tmp_bytes);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  launch_corner_keys(c->d, k, v, n, e->st);
  radix_sort_pairs(k, v, k2, v2, n, rt, e->st, e->radix_bits);
  HIPCK(hipMemcpyAsync(tmpb, c->d, n * sizeof(vl_box), hipMemcpyDeviceToDevice, e->st));
  launch_box_gather(tmpb, v, c->d, n, e->st);
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

int vl_boolean_op(vl_engine* e, const vl_box* a, uint64_t an, const vl_box* b, uint64_t bn, vl_box* out,
                  uint64_t cap, uint64_t* out_n, uint64_t ctrl, int64_t val) {
  if (!e || !out_n || (an && !a) || (bn && !b)) return fail(VL_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  vl_layer A, B, C;
  A.e = B.e = C.e = e;
  auto cleanup = [&]() {
    for (vl_layer* l : {&A, &B, &C})
      if (l->d) e->pool.put(l->d, l->bytes);
  };
  int rc = vl_layer_upload(&A, a, (int64_t)an);
  if (rc == VL_OK) rc = vl_layer_upload(&B, b, (int64_t)bn);
  if (rc == VL_OK) {
    switch (ctrl) {
      case VL_OP_XOR: rc = vl_xor(e, &A, &B, &C); break;
      case VL_OP_OR: rc = vl_or(e, &A, &B, &C); break;
      case VL_OP_AND: rc = vl_and(e, &A, &B, &C); break;
      case VL_OP_NOT: rc = vl_not(e, &A, &B, &C); break;
      case VL_OP_MERGE: rc = vl_merge(e, &A, &C); break;
      case VL_OP_SIZE:
        if (val < INT32_MIN || val > INT32_MAX) rc = fail(VL_EINVAL, "SIZE amount out of range");
        else rc = vl_size(e, &A, (int32_t)val, &C);
        break;
      case VL_OP_GROW:
        if (val < INT32_MIN || val > INT32_MAX) rc = fail(VL_EINVAL, "GROW amount out of range");
        else rc = vl_grow(e, &A, (int32_t)val, (int32_t)val, (int32_t)val, (int32_t)val, &C);
        break;
      default:
        *out_n = 0;   // unknown control: empty result (src/boost_polygon_api.cpp:136-139)
        cleanup();
        return VL_OK;
    }
  }
  if (rc == VL_OK) {
    *out_n = (uint64_t)C.n;
    if ((uint64_t)C.n > cap) {
      rc = fail(VL_ENOSPC, "output needs %lld boxes, capacity %llu", (long long)C.n, (unsigned long long)cap);
    } else if (C.n > 0) {
      int64_t got = 0;
      rc = vl_layer_download(&C, out, (int64_t)cap, &got);
    }
  }
  cleanup();
  return rc;
}

// ---- sorts -------------------------------------------------------------------
int vl_event_sort(vl_engine* e, vl_tracker* arr, int64_t n) {
  if (!e || (n > 0 && !arr) || n < 0) return fail(VL_EINVAL, "bad argument");
  if (n <= 1) return VL_OK;
  if (n > (int64_t)UINT32_MAX) return fail(VL_EINVAL, "more than 2^32 trackers");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  Scratch s(e);
  const bool dev = is_device_ptr(arr);
  const size_t bytes = (size_t)n * sizeof(vl_tracker);
  vl_tracker* d = arr;
  if (!dev) {
    void* p = nullptr;
    VLCK(upload_tmp(e, s, arr, bytes, &p));
    d = (vl_tracker*)p;
  }
  uint64_t* k = s.get<uint64_t>(n);
  uint32_t* v = s.get<uint32_t>(n);
  uint64_t* k2 = s.get<uint64_t>(n);
  uint32_t* v2 = s.get<uint32_t>(n);
  vl_tracker* src = s.get<vl_tracker>(n);
  void* rt = s.get<char>(radix_plan(n, e->radix_bits).tmp_bytes);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  launch_tracker_keys(d, k, v, n, e->st);
  radix_sort_pairs(k, v, k2, v2, n, rt, e->st, e->radix_bits);
  HIPCK(hipMemcpyAsync(src, d, bytes, hipMemcpyDeviceToDevice, e->st));
  launch_tracker_gather(src, v, d, n, e->st);
  if (!dev) HIPCK(hipMemcpyAsync(arr, d, bytes, hipMemcpyDeviceToHost, e->st));
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

int vl_morton_sort_direct(vl_engine* e, vl_box* boxes, int64_t n) {
  if (!e || (n > 0 && !boxes) || n < 0) return fail(VL_EINVAL, "bad argument");
  if (n <= 1) return VL_OK;
  if (n > (int64_t)UINT32_MAX) return fail(VL_EINVAL, "more than 2^32 boxes");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  Scratch s(e);
  const bool dev = is_device_ptr(boxes);
  const size_t bytes = (size_t)n * sizeof(vl_box);
  vl_box* d = boxes;
  if (!dev) {
    void* p = nullptr;
    VLCK(upload_tmp(e, s, boxes, bytes, &p));
    d = (vl_box*)p;
  }
  uint64_t* k = s.get<uint64_t>(n);
  uint32_t* v = s.get<uint32_t>(n);
  uint64_t* k2 = s.get<uint64_t>(n);
  uint32_t* v2 = s.get<uint32_t>(n);
  vl_box* src = s.get<vl_box>(n);
  void* rt = s.get<char>(radix_plan(n, e->radix_bits).tmp_bytes);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  launch_morton_keys(d, k, v, n, e->st);
  radix_sort_pairs(k, v, k2, v2, n, rt, e->st, e->radix_bits);
  HIPCK(hipMemcpyAsync(src, d, bytes, hipMemcpyDeviceToDevice, e->st));
  launch_box_gather(src, v, d, n, e->st);
  if (!dev) HIPCK(hipMemcpyAsync(boxes, d, bytes, hipMemcpyDeviceToHost, e->st));
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

int vl_morton_sort_indirect(vl_engine* e, const vl_box* boxes, int64_t* idx, int64_t n) {
  if (!e || (n > 0 && (!boxes || !idx)) || n < 0) return fail(VL_EINVAL, "bad argument");
  if (n <= 0) return VL_OK;
  if (n > (int64_t)UINT32_MAX) return fail(VL_EINVAL, "more than 2^32 boxes");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  OpTimer tm(e);
  Scratch s(e);
  const bool devb = is_device_ptr(boxes), devi = is_device_ptr(idx);
  const vl_box* d = boxes;
  if (!devb) {
    void* p = nullptr;
    VLCK(upload_tmp(e, s, boxes, (size_t)n * sizeof(vl_box), &p));
    d = (const vl_box*)p;
  }
  int64_t* di = devi ? idx : s.get<int64_t>(n);
  uint64_t* k = s.get<uint64_t>(n);
  uint32_t* v = s.get<uint32_t>(n);
  uint64_t* k2 = s.get<uint64_t>(n);
  uint32_t* v2 = s.get<uint32_t>(n);
  void* rt = s.get<char>(radix_plan(n, e->radix_bits).tmp_bytes);
  if (s.oom) return fail(VL_ENOMEM, "scratch");
  launch_morton_keys(d, k, v, n, e->st);
  if (n > 1) radix_sort_pairs(k, v, k2, v2, n, rt, e->st, e->radix_bits);
  launch_perm_to_index1(v, di, n, e->st);
  if (!devi) HIPCK(hipMemcpyAsync(idx, di, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, e->st));
  HIPCK(hipStreamSynchronize(e->st));
  return VL_OK;
}

void device_event_sort(vl_tracker* d_arr, int64_t n) {
  if (vl_event_sort(default_engine(), d_arr, n) != VL_OK) {
    fprintf(stderr, "device_event_sort: %s\n", g_err.c_str());
    abort();
  }
}
void thrust_sort_boxes_direct(vl_box* d_boxes, int64_t n) {
  if (vl_morton_sort_direct(default_engine(), d_boxes, n) != VL_OK) {
    fprintf(stderr, "thrust_sort_boxes_direct: %s\n", g_err.c_str());
    abort();
  }
}
void thrust_sort_boxes_indirect(vl_box* d_boxes, int64_t* d_indices, int64_t n) {
  if (vl_morton_sort_indirect(default_engine(), d_boxes, d_indices, n) != VL_OK) {
    fprintf(stderr, "thrust_sort_boxes_indirect: %s\n", g_err.c_str());
    abort();
  }
}

}  // extern "C"

// ---- programs ----------------------------------------------------------------
struct ProgOp {
  int op;
  const vl_layer* a;
  const vl_layer* b;
  vl_layer* c;
  int32_t args[4];
};

struct vl_program {
  vl_engine* e = nullptr;
  std::vector<ProgOp> ops;
  int64_t failures = 0;
  std::vector<vl_engine*> workers;   // worker engines (same device, own streams)
  int max_workers = 16;
  // graph capture (vl_program_set_graph)
  bool graph = false;
  hipGraph_t gr = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<void*> pinned;                                   // sources of captured uploads
  std::vector<std::pair<const vl_layer*, uint64_t>> versions;  // layer versions the graph holds for
  int64_t cap_failures = 0, cap_fused = 0;                     // per replay
  int graph_nw = 0;                                            // worker streams the graph was captured on
  int64_t captures = 0, replays = 0, fallbacks = 0;
};
static int run_prog_op(vl_engine* e, const ProgOp& p, int64_t* failures) {
  switch (p.op) {
    case -1:
      if (p.a->n > 0) (*failures)++;
      return VL_OK;
    case VL_OP_XOR: return vl_xor(e, p.a, p.b, p.c);
    case VL_OP_OR: return vl_or(e, p.a, p.b, p.c);
    case VL_OP_AND: return vl_and(e, p.a, p.b, p.c);
    case VL_OP_NOT: return vl_not(e, p.a, p.b, p.c);
    case VL_OP_MERGE: return vl_merge(e, p.a, p.c);
    case VL_OP_SIZE: return vl_size(e, p.a, p.args[0], p.c);
    case VL_OP_GROW: return vl_grow(e, p.a, p.args[0], p.args[1], p.args[2], p.args[3], p.c);
    default: return fail(VL_EINVAL, "unknown program op %d", p.op);
  }
}

extern "C" {

int vl_program_create(vl_engine* e, vl_program** out) {
  if (!e || !out) return fail(VL_EINVAL, "null argument");
  vl_program* p = new vl_program();
  p->e = e;
  e->refs.fetch_add(1);
  *out = p;
  return VL_OK;
}

static void graph_drop(vl_program* p);

int vl_program_destroy(vl_program* p) {
  if (!p) return VL_OK;
  vl_engine* e = p->e;
  {
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    graph_drop(p);
  }
  for (vl_engine* w : p->workers) engine_release(w);
  for (const ProgOp& o : p->ops)
    for (const vl_layer* l : {o.a, o.b, (const vl_layer*)o.c})
      if (l) layer_release(const_cast<vl_layer*>(l));
  delete p;
  engine_release(e);
  return VL_OK;
}

int vl_program_add(vl_program* p, int op, const vl_layer* a, const vl_layer* b, vl_layer* c, const int32_t* args) {
  if (!p || !a) return fail(VL_EINVAL, "null argument");
  if (op != -1 && !c) return fail(VL_EINVAL, "null output layer");
  if ((op == VL_OP_XOR || op == VL_OP_OR || op == VL_OP_AND || op == VL_OP_NOT) && !b)
    return fail(VL_EINVAL, "binary operator needs two inputs");
  if (op < -1 || op > VL_OP_GROW) return fail(VL_EINVAL, "unknown op %d", op);
  for (const vl_layer* l : {a, b, (const vl_layer*)c})
    if (l && l->e != p->e) return fail(VL_EINVAL, "layer belongs to another engine");
  ProgOp o{op, a, b, c, {0, 0, 0, 0}};
  if (args) {
    const int na = (op == VL_OP_GROW) ? 4 : 1;
    for (int i = 0; i < na; i++) o.args[i] = args[i];
  }
  // the program keeps every layer it records alive until it is destroyed
  for (const vl_layer* l : {a, b, (const vl_layer*)c})
    if (l) const_cast<vl_layer*>(l)->refs.fetch_add(1);
  p->ops.push_back(o);
  return VL_OK;
}

// ---- program compilation: fused groups ----------------------------------------
// A step of a compiled program: one operator, or a fused group (ops are the
// group's operators in program order; run one by one if the group is declined).
struct ProgStep {
  std::vector<int> ops;
  bool fused = false;
  bool check = false;   // a single AND read only by ASSERT_ZERO: and_check
  const vl_layer* src = nullptr;
  const vl_layer* src2 = nullptr;   // the group's second layer (an image of kind MORPH_LAYER)
  FusedSpec fs{};
  vl_layer* outs[FUSE_MAX_OUT] = {};
};

// Greedy grouping of consecutive operators over one source layer S and at
// most one second layer T: the group starts at a GROW, SIZE or MERGE of S, or
// at a Boolean operator of two layers S and T, and takes every following
// operator whose inputs are S, T or earlier group outputs (GROW with
// non-negative arguments and SIZE only on S itself, at most two images in
// all, counting T, at most FUSE_MAX_OUT outputs, each written once and
// never S or T) and every ASSERT_ZERO of such a layer.  Each output is a
// truth table over (S, image 0, image 1); S is 0xAA, image 0 is 0xCC, image 1
// is 0xF0.  The OP1 ring of a layer is one group (S and two morphological
// images); AND, NOT, OR of a layer pair and a GROW of the first (the
// uneven-tile batch) is one group (S, T and one grown image).
struct GroupBuilder {
  const vl_layer* src = nullptr;
  const vl_layer* src2 = nullptr;
  FusedSpec fs{};
  std::vector<std::pair<vl_layer*, uint32_t>> nodes;
  std::vector<int> ops;
  int nprod = 0;

  bool tt_of(const vl_layer* l, uint32_t* tt) const {
    if (l == src) {
      *tt = 0xAAu;
      return true;
    }
    if (l && l == src2) {
      for (int k = 0; k < fs.nmorph; k++)
        if (fs.kind[k] == MORPH_LAYER) {
          *tt = k ? 0xF0u : 0xCCu;
          return true;
        }
    }
    for (const auto& nd : nodes)
      if (nd.first == l) {
        *tt = nd.second;
        return true;
      }
    return false;
  }
  // Image index of a morphological operator of S (added if new), -1 if full.
  int max_images = 2;
  int morph(int32_t kind, const int32_t a[4]) {
    for (int k = 0; k < fs.nmorph; k++)
      if (fs.kind[k] == kind && memcmp(fs.arg[k], a, 4 * sizeof(int32_t)) == 0) return k;
    if (fs.nmorph >= max_images) return -1;
    const int k = fs.nmorph++;
    fs.kind[k] = kind;
    memcpy(fs.arg[k], a, 4 * sizeof(int32_t));
    const int32_t reach = kind == MORPH_GROW ? std::max(std::max(a[0], a[1]), std::max(a[2], a[3])) : a[0];
    fs.halo = std::max(fs.halo, reach);
    return k;
  }
  // The image of second layer l (added if there is none yet and a slot is free).
  bool two_layers = true;
  bool layer_image(const vl_layer* l, uint32_t* tt) {
    if (src2) return tt_of(l, tt);
    if (!two_layers) return false;
    if (!l || l == src || fs.nmorph >= max_images) return false;
    for (const auto& nd : nodes)
      if (nd.first == l) return false;
    const int k = fs.nmorph++;
    fs.kind[k] = MORPH_LAYER;
    memset(fs.arg[k], 0, sizeof(fs.arg[k]));
    src2 = l;
    *tt = k ? 0xF0u : 0xCCu;
    return true;
  }
  bool add(const ProgOp& o) {
    if (!src) return false;
    uint32_t ta = 0, tb = 0, tt = 0;
    if (o.op == -1) return tt_of(o.a, &ta);
    if (!o.c || o.c == src || (src2 && o.c == src2) || (int)nodes.size() == FUSE_MAX_OUT) return false;
    for (const auto& nd : nodes)
      if (nd.first == o.c) return false;
    const FusedSpec saved = fs;
    const vl_layer* const saved2 = src2;
    switch (o.op) {
      case VL_OP_GROW: {
        if (o.a != src || o.args[0] < 0 || o.args[1] < 0 || o.args[2] < 0 || o.args[3] < 0) return false;
        const int k = morph(MORPH_GROW, o.args);
        if (k < 0) return false;
        tt = k ? 0xF0u : 0xCCu;
        break;
      }
      case VL_OP_SIZE: {
        const int32_t d = o.args[0];
        if (o.a != src || d == INT32_MIN) return false;
        if (d == 0) {
          tt = 0xAAu;
          break;
        }
        const int32_t a[4] = {d > 0 ? d : -d, d > 0 ? d : 0, d > 0 ? d : 0, d > 0 ? d : 0};
        const int k = morph(d > 0 ? MORPH_GROW : MORPH_ERODE, a);
        if (k < 0) return false;
        tt = k ? 0xF0u : 0xCCu;
        break;
      }
      case VL_OP_MERGE:
        if (!tt_of(o.a, &ta)) return false;
        tt = ta;
        break;
      case VL_OP_AND:
      case VL_OP_NOT:
      case VL_OP_OR:
      case VL_OP_XOR: {
        const bool ka = tt_of(o.a, &ta), kb = tt_of(o.b, &tb);
        // one operand outside the group: the group's second layer
        const bool ok = (ka && kb) || (ka && !kb && layer_image(o.b, &tb)) || (!ka && kb && layer_image(o.a, &ta));
        tt = o.op == VL_OP_AND ? (ta & tb) : o.op == VL_OP_NOT ? (ta & ~tb) : o.op == VL_OP_OR ? (ta | tb) : (ta ^ tb);
        tt &= 0xFFu;
        // An operator whose result is empty for every S (the OP1 ring's
        // ANDs of a ring with its own source) is a consistency check of the
        // group's outputs: it stays out of the group and runs on the
        // materialised layers, so the check can fail when they are wrong.
        if (!ok || o.c == src2 || tt == 0u) {
          fs = saved;
          src2 = saved2;
          return false;
        }
        break;
      }
      default:
        fs = saved;
        src2 = saved2;
        return false;
    }
    nodes.push_back({o.c, tt});
    nprod++;
    return true;
  }
  bool start(const ProgOp& o) {
    const bool binop = o.op == VL_OP_AND || o.op == VL_OP_NOT || o.op == VL_OP_OR || o.op == VL_OP_XOR;
    if (o.op != VL_OP_GROW && o.op != VL_OP_SIZE && o.op != VL_OP_MERGE && !(two_layers && binop && o.b && o.b != o.a))
      return false;
    src = o.a;
    if (add(o)) return true;
    const int mi = max_images;
    const bool tl = two_layers;
    *this = GroupBuilder();
    max_images = mi;
    two_layers = tl;
    return false;
  }
};

// fuse: 0 no groups, 1 groups of up to two morphological images, 2 groups of
// at most one (smaller compressed grids per tile, one more binning of S).
// fuse2: groups may take a second layer (they run fused even with fuse 0,
// which then only keeps groups without one from fusing).
static std::vector<ProgStep> compile_program(const std::vector<ProgOp>& ops, int fuse, bool fuse2, bool check_ands) {
  std::vector<ProgStep> steps;
  const int max_images = fuse == 2 ? 1 : 2;
  const bool grouping = fuse || fuse2;
  GroupBuilder g;
  g.max_images = max_images;
  g.two_layers = fuse2;
  auto single = [&](int i) {
    ProgStep st;
    st.ops.push_back(i);
    steps.push_back(st);
  };
  auto flush = [&]() {
    if (g.ops.empty()) return;
    if (g.nprod >= 2 && (fuse || g.src2)) {
      ProgStep st;
      st.ops = g.ops;
      st.fused = true;
      st.src = g.src;
      st.src2 = g.src2;
      st.fs = g.fs;
      st.fs.nout = (int32_t)g.nodes.size();
      for (size_t k = 0; k < g.nodes.size(); k++) {
        st.outs[k] = g.nodes[k].first;
        st.fs.tt[k] = g.nodes[k].second;
      }
      steps.push_back(st);
    } else {
      for (int i : g.ops) single(i);
    }
    g = GroupBuilder();
    g.max_images = max_images;
    g.two_layers = fuse2;
  };
  for (int i = 0; i < (int)ops.size(); i++) {
    if (grouping && g.add(ops[i])) {
      g.ops.push_back(i);
      continue;
    }
    flush();
    if (grouping && g.start(ops[i])) {
      g.ops.push_back(i);
      continue;
    }
    single(i);
  }
  flush();
  // ANDs whose result is read only by ASSERT_ZERO until the layer is written
  // again (checks expected to pass) are planned as emptiness tests first
  if (check_ands) {
    for (ProgStep& st : steps) {
      if (st.fused || st.ops.size() != 1 || ops[st.ops[0]].op != VL_OP_AND) continue;
      const int i = st.ops[0];
      const vl_layer* c = ops[i].c;
      bool only_asserts = true, read = false;
      for (int j = i + 1; j < (int)ops.size() && only_asserts; j++) {
        const ProgOp& o = ops[j];
        if (o.a == c || o.b == c) {
          if (o.op == -1) read = true;
          else only_asserts = false;
        }
        if (o.c == c) break;
      }
      st.check = only_asserts && read;
    }
  }
  return steps;
}

// Runs one step on engine e.
static int run_step(vl_engine* e, const std::vector<ProgOp>& ops, const ProgStep& st, int64_t* failures) {
  if (st.fused) {
    int rc;
    {
      std::lock_guard<std::recursive_mutex> lk(e->mu);
      HIPCK(hipSetDevice(e->device));
      rc = run_fused(e, st.src, st.src2, st.fs, st.outs);
    }
    if (rc == VL_OK) {
      e->fused_groups++;
      for (int i : st.ops)
        if (ops[i].op == -1 && ops[i].a->n > 0) (*failures)++;
      return VL_OK;
    }
    // a group whose scratch does not fit (its output streams are sized for
    // every output at once) runs operator by operator, which needs less
    if (rc == VL_ENOMEM && e->cap.mode == CAP_EAGER) {
      if (g_debug) fprintf(stderr, "[vl] fused group out of memory, run by operator: %s\n", g_err.c_str());
      e->pool.trim();
      rc = FUSE_DECLINED;
    }
    if (rc != FUSE_DECLINED) return rc;
    g_err.clear();
  }
  if (st.check) {
    const ProgOp& o = ops[st.ops[0]];
    return and_check(e, o.a, o.b, o.c);
  }
  for (int i : st.ops) VLCK(run_prog_op(e, ops[i], failures));
  return VL_OK;
}

// Independent chains of a compiled program.  Step j depends on an earlier i
// when it reads or writes a layer that i writes, or writes a layer that i
// reads; the connected components of that relation run concurrently, each in
// program order on its own worker engine (stream) and host thread.
static std::vector<std::vector<int>> program_chains(const std::vector<ProgOp>& ops,
                                                    const std::vector<ProgStep>& steps) {
  const int n = (int)steps.size();
  std::vector<int> parent(n);
  for (int i = 0; i < n; i++) parent[i] = i;
  std::function<int(int)> find = [&](int x) { return parent[x] == x ? x : parent[x] = find(parent[x]); };
  std::vector<std::vector<const vl_layer*>> rd(n), wr(n);
  for (int i = 0; i < n; i++)
    for (int k : steps[i].ops) {
      const ProgOp& o = ops[k];
      if (o.a) rd[i].push_back(o.a);
      if (o.b) rd[i].push_back(o.b);
      if (o.c) wr[i].push_back(o.c);
    }
  auto meets = [](const std::vector<const vl_layer*>& x, const std::vector<const vl_layer*>& y) {
    for (const vl_layer* p : x)
      for (const vl_layer* q : y)
        if (p == q) return true;
    return false;
  };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < j; i++) {
      const bool dep = meets(wr[i], rd[j]) || meets(wr[i], wr[j]) || meets(rd[i], wr[j]);
      if (dep) parent[find(j)] = find(i);
    }
  std::map<int, std::vector<int>> groups;
  for (int i = 0; i < n; i++) groups[find(i)].push_back(i);
  std::vector<std::vector<int>> out;
  for (auto& kv : groups) out.push_back(kv.second);
  return out;
}

// Worker engines of a program: at least n, with the root engine's settings.
static int program_workers(vl_program* p, int n) {
  vl_engine* e = p->e;
  while ((int)p->workers.size() < n) {
    vl_engine* w = nullptr;
    VLCK(engine_new(e->device, &w));
    w->root = e;
    w->pool.fallback = &e->pool;
    p->workers.push_back(w);
  }
  for (int k = 0; k < n; k++) {
    vl_engine* w = p->workers[k];
    w->nmax = e->nmax;
    w->target = e->target;
    w->erode_target = e->erode_target;
    w->level_factor = e->level_factor;
    w->erode_kernel = e->erode_kernel;
    w->max_wg_cap = e->max_wg_cap;
    w->fuse = e->fuse;
    w->fuse2 = e->fuse2;
    w->fuse_target = e->fuse_target;
    w->bin_mode = e->bin_mode;
    w->radix_bits = e->radix_bits;
    w->check_target = e->check_target;
    w->tile_shape = e->tile_shape;
    w->region_scale = std::max(w->region_scale, e->region_scale);
    w->fuse_region_scale = std::max(w->fuse_region_scale, e->fuse_region_scale);
    (void)prof_resolve(w);
    w->prof = e->prof;
    w->fused_groups = 0;
    for (auto& st : w->kst) st = vl_engine::KStat();
  }
  return VL_OK;
}

// Runs the chains on workers 0..nw-1 (chain c on worker c % nw, in order):
// each worker on its own host thread when `threads`, else one after another
// from this thread (capture: the graph's branches come from the streams, so
// they still run concurrently when replayed).  With `sync` every worker
// stream is waited for.  Adds assertion failures, fused-group counts and
// kernel statistics to the program and its engine.
static int run_chains(vl_program* p, const std::vector<ProgStep>& steps, const std::vector<std::vector<int>>& chains,
                      int nw, bool threads, bool sync) {
  vl_engine* e = p->e;
  std::vector<int> rcs(nw, VL_OK);
  std::vector<std::string> errs(nw);
  std::vector<int64_t> fails(nw, 0);
  auto work = [&](int k) {
    vl_engine* w = p->workers[k];
    (void)hipSetDevice(w->device);
    for (size_t c = k; c < chains.size(); c += nw) {
      for (int i : chains[c]) {
        rcs[k] = run_step(w, p->ops, steps[i], &fails[k]);
        if (rcs[k] != VL_OK) {
          errs[k] = g_err;
          return;
        }
      }
    }
    if (sync && hipStreamSynchronize(w->st) != hipSuccess) {
      rcs[k] = VL_EHIP;
      errs[k] = "worker stream failed";
    }
  };
  if (threads && nw > 1) {
    std::vector<std::thread> th;
    for (int k = 0; k < nw; k++) th.emplace_back(work, k);
    for (auto& t : th) t.join();
  } else {
    for (int k = 0; k < nw; k++) {
      work(k);
      if (rcs[k] != VL_OK) break;
    }
  }
  int rc = VL_OK;
  for (int k = 0; k < nw; k++) {
    p->failures += fails[k];
    vl_engine* w = p->workers[k];
    e->fused_groups += w->fused_groups;
    w->fused_groups = 0;
    if (rc == VL_OK && rcs[k] == VL_OK && sync) rcs[k] = prof_resolve(w);
    for (int s = 0; s < KS_N; s++) {
      e->kst[s].launches += w->kst[s].launches;
      e->kst[s].ms += w->kst[s].ms;
      e->kst[s].bytes += w->kst[s].bytes;
      w->kst[s] = vl_engine::KStat();
    }
    if (rc == VL_OK && rcs[k] != VL_OK) {
      rc = rcs[k];
      g_err = errs[k];
    }
  }
  return rc;
}

static void graph_drop(vl_program* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->gr) (void)hipGraphDestroy(p->gr);
  p->gexec = nullptr;
  p->gr = nullptr;
  for (void* q : p->pinned) (void)hipHostFree(q);
  p->pinned.clear();
  p->versions.clear();
}

static std::vector<const vl_layer*> program_layers(const vl_program* p) {
  std::vector<const vl_layer*> v;
  for (const ProgOp& o : p->ops)
    for (const vl_layer* l : {o.a, o.b, (const vl_layer*)o.c})
      if (l && std::find(v.begin(), v.end(), l) == v.end()) v.push_back(l);
  return v;
}

static void forget_output_mbrs(vl_program* p) {
  for (const ProgOp& o : p->ops)
    if (o.c) {
      o.c->mbr_ok = false;
      o.c->bound_version = ~0ull;
    }
}

// Launches the captured graph and reports whether every size it was planned
// with held (*ok).
static int graph_replay(vl_program* p, int nw, bool* ok) {
  vl_engine* e = p->e;
  const double tl = g_debug ? now_ms() : 0.0;
  HIPCK(hipGraphLaunch(p->gexec, e->st));
  if (g_debug) fprintf(stderr, "[vl] graph launch returned after %.2f ms\n", now_ms() - tl);
  HIPCK(hipStreamSynchronize(e->st));
  if (g_debug) fprintf(stderr, "[vl] graph replay done after %.2f ms\n", now_ms() - tl);
  uint32_t flag = 0;
  *ok = true;
  for (int k = 0; k < nw; k++) {
    HIPCK(hipMemcpy(&flag, cap_flag(p->workers[k]), 4, hipMemcpyDeviceToHost));
    if (flag) *ok = false;
  }
  return VL_OK;
}

// Record run (eager, sizes logged), then capture of the same work into a
// graph, then its first launch.  Falls back to an eager run when the capture
// fails or the first launch's checks do not hold.
static int graph_record_capture(vl_program* p, const std::vector<ProgStep>& steps,
                                const std::vector<std::vector<int>>& chains, int nw) {
  vl_engine* e = p->e;
  graph_drop(p);
  struct Tune {
    double rs, frs;
  };
  std::vector<Tune> tune(nw);
  for (int k = 0; k < nw; k++) {
    vl_engine* w = p->workers[k];
    tune[k] = Tune{w->region_scale, w->fuse_region_scale};
    w->cap.mode = CAP_RECORD;
    w->cap.log.clear();
    w->cap.pos = 0;
  }
  forget_output_mbrs(p);
  int rc = run_chains(p, steps, chains, nw, true, true);
  for (int k = 0; k < nw; k++) p->workers[k]->cap.mode = CAP_EAGER;
  if (rc != VL_OK) return rc;
  // capture: the same decisions as the record run (same tuning state, same
  // cached bounding boxes, logged sizes)
  const int64_t failures = p->failures, fused = e->fused_groups;
  forget_output_mbrs(p);
  for (int k = 0; k < nw; k++) {
    vl_engine* w = p->workers[k];
    w->region_scale = tune[k].rs;
    w->fuse_region_scale = tune[k].frs;
    w->cap.mode = CAP_CAPTURE;
    w->cap.pos = 0;
    w->cap.pinned.clear();
    w->pool.no_trim = true;
    (void)prof_resolve(w);   // stages of the record run
  }
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins(nw, nullptr);
  bool ev_ok = hipEventCreateWithFlags(&fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; k < nw && ev_ok; k++) ev_ok = hipEventCreateWithFlags(&joins[k], hipEventDisableTiming) == hipSuccess;
  hipGraph_t gr = nullptr;
  int crc = ev_ok ? VL_OK : fail(VL_EHIP, "graph events");
  if (crc == VL_OK && hipStreamBeginCapture(e->st, hipStreamCaptureModeRelaxed) != hipSuccess)
    crc = fail(VL_EHIP, "hipStreamBeginCapture: %s", hipGetErrorString(hipGetLastError()));
  const bool capturing = crc == VL_OK;
  if (capturing) {
    if (hipEventRecord(fork, e->st) != hipSuccess) crc = fail(VL_EHIP, "fork event");
    for (int k = 0; k < nw && crc == VL_OK; k++) {
      vl_engine* w = p->workers[k];
      if (hipStreamWaitEvent(w->st, fork, 0) != hipSuccess) crc = fail(VL_EHIP, "fork wait");
      launch_fill_u32(cap_flag(w), 0u, 1, w->st);
    }
    if (crc == VL_OK) crc = run_chains(p, steps, chains, nw, false, false);
    for (int k = 0; k < nw; k++) {
      vl_engine* w = p->workers[k];
      if (hipEventRecord(joins[k], w->st) != hipSuccess || hipStreamWaitEvent(e->st, joins[k], 0) != hipSuccess)
        if (crc == VL_OK) crc = fail(VL_EHIP, "join");
    }
    const hipError_t er = hipStreamEndCapture(e->st, &gr);
    if (er != hipSuccess && crc == VL_OK) crc = fail(VL_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(er));
  }
  for (int k = 0; k < nw; k++) {
    vl_engine* w = p->workers[k];
    if (crc == VL_OK && w->cap.pos != w->cap.log.size()) crc = fail(VL_EINTERNAL, "capture left recorded sizes unread");
    w->cap.mode = CAP_EAGER;
    w->pool.no_trim = false;
    for (void* q : w->cap.pinned) p->pinned.push_back(q);
    w->cap.pinned.clear();
  }
  if (fork) (void)hipEventDestroy(fork);
  for (hipEvent_t j : joins)
    if (j) (void)hipEventDestroy(j);
  // counted once (the record run counted them)
  p->cap_failures = p->failures - failures;
  p->cap_fused = e->fused_groups - fused;
  e->fused_groups = fused;
  p->failures = failures;
  if (crc == VL_OK) {
    p->gr = gr;
    if (hipGraphInstantiate(&p->gexec, gr, nullptr, nullptr, 0) != hipSuccess) {
      p->gexec = nullptr;
      crc = fail(VL_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(hipGetLastError()));
    }
  } else if (gr) {
    (void)hipGraphDestroy(gr);
  }
  bool ok = false;
  if (crc == VL_OK) crc = graph_replay(p, nw, &ok);
  if (crc == VL_OK && ok) {
    for (const vl_layer* l : program_layers(p)) p->versions.push_back({l, l->version});
    p->captures++;
    p->graph_nw = nw;
    return VL_OK;
  }
  // the capture replaced the outputs' storage: compute them again eagerly
  if (g_debug) fprintf(stderr, "[vl] graph capture not used: %s\n", crc == VL_OK ? "first launch's checks failed" : g_err.c_str());
  graph_drop(p);
  p->fallbacks++;
  g_err.clear();
  p->failures = 0;
  forget_output_mbrs(p);
  return run_chains(p, steps, chains, nw, true, true);
}

int vl_program_run(vl_program* p) {
  if (!p) return fail(VL_EINVAL, "null program");
  vl_engine* e = p->e;
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  p->failures = 0;
  const std::vector<ProgStep> steps = compile_program(p->ops, e->fuse, e->fuse2 != 0, e->and_check != 0);
  const std::vector<std::vector<int>> chains = program_chains(p->ops, steps);
  int nw = (int)std::min<size_t>(chains.size(), (size_t)std::min(p->max_workers, p->e->workers));
  // Chains run side by side only as far as their scratch fits in device
  // memory: an operator holds up to about 300 bytes per input box at its
  // peak (tile entries, strip streams, outputs, pool rounding), and a
  // chain's operators work on layers about as large as the chain's inputs
  // (the layers no operator of the program writes, whose sizes do not change
  // from run to run).  The chains with the largest inputs count first.
  if (nw > 1) {
    auto written = [&](const vl_layer* l) {
      for (const ProgOp& o : p->ops)
        if (o.c == l) return true;
      return false;
    };
    std::vector<double> est;
    for (const auto& ch : chains) {
      double in = 0;
      std::vector<const vl_layer*> seen;
      for (int si : ch)
        for (int k : steps[si].ops)
          for (const vl_layer* l : {p->ops[k].a, p->ops[k].b})
            if (l && !written(l) && std::find(seen.begin(), seen.end(), l) == seen.end()) {
              seen.push_back(l);
              in += (double)l->n;
            }
      est.push_back(300.0 * in);
    }
    std::sort(est.begin(), est.end(), std::greater<double>());
    const double avail = 0.7 * (double)e->pool.available();
    double acc = 0;
    int fit = 0;
    for (int k = 0; k < nw; k++) {
      acc += est[k];
      if (k > 0 && acc > avail) break;
      fit++;
    }
    if (fit < nw && g_debug) fprintf(stderr, "[vl] program: %d of %d worker streams fit in device memory\n", fit, nw);
    nw = std::max(1, fit);
  }
  // the bounding boxes of the program's inputs (layers no operator writes)
  // are cached before concurrent chains read them
  {
    auto written = [&](const vl_layer* l) {
      for (const ProgOp& o : p->ops)
        if (o.c == l) return true;
      return false;
    };
    for (const ProgOp& o : p->ops) {
      Box64 m;
      if (o.a && !written(o.a)) VLCK(layer_mbr(e, o.a, &m));
      if (o.b && !written(o.b)) VLCK(layer_mbr(e, o.b, &m));
    }
  }
  hipEvent_t t0, t1;
  HIPCK(hipEventCreate(&t0));
  HIPCK(hipEventCreate(&t1));
  HIPCK(hipEventRecord(t0, e->st));
  int rc = VL_OK;
  if (p->graph && !chains.empty()) {
    // a captured program runs on worker engines only (their memory pools
    // hold the graph's buffers), one per chain up to max_workers; a replay
    // checks the streams the graph was captured on
    nw = std::max(nw, 1);
    if (p->gexec) nw = p->graph_nw;
    rc = program_workers(p, nw);
    bool replayed = false;
    if (rc == VL_OK && p->gexec) {
      bool same = true;
      for (const auto& lv : p->versions)
        if (lv.first->version != lv.second) same = false;
      if (same) {
        for (int k = 0; k < nw; k++) HIPCK(hipStreamWaitEvent(p->workers[k]->st, t0, 0));
        bool ok = false;
        rc = graph_replay(p, nw, &ok);
        if (rc == VL_OK && ok) {
          p->failures = p->cap_failures;
          e->fused_groups += p->cap_fused;
          p->replays++;
          replayed = true;
        } else if (rc == VL_OK) {
          p->fallbacks++;   // sizes differed: the results are recomputed below
        }
      }
      if (!replayed) graph_drop(p);
    }
    if (rc == VL_OK && !replayed) {
      for (int k = 0; k < nw; k++) HIPCK(hipStreamWaitEvent(p->workers[k]->st, t0, 0));
      rc = graph_record_capture(p, steps, chains, nw);
    }
  } else if (nw <= 1) {
    for (const ProgStep& st : steps) {
      rc = run_step(e, p->ops, st, &p->failures);
      if (rc != VL_OK) break;
    }
  } else {
    rc = program_workers(p, nw);
    for (int k = 0; k < nw && rc == VL_OK; k++) HIPCK(hipStreamWaitEvent(p->workers[k]->st, t0, 0));
    if (rc == VL_OK) rc = run_chains(p, steps, chains, nw, true, true);
  }
  (void)hipEventRecord(t1, e->st);
  (void)hipEventSynchronize(t1);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, t0, t1) == hipSuccess) e->last_ms = ms;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return rc;
}

int vl_program_set_graph(vl_program* p, int enable) {
  if (!p) return fail(VL_EINVAL, "null program");
  std::lock_guard<std::recursive_mutex> lk(p->e->mu);
  p->graph = enable != 0;
  if (!p->graph) graph_drop(p);
  return VL_OK;
}

int vl_program_graph_stats(vl_program* p, int64_t* captures, int64_t* replays, int64_t* fallbacks) {
  if (!p || !captures || !replays || !fallbacks) return fail(VL_EINVAL, "null argument");
  *captures = p->captures;
  *replays = p->replays;
  *fallbacks = p->fallbacks;
  return VL_OK;
}

int vl_program_assert_count(vl_program* p, int64_t* failures) {
  if (!p || !failures) return fail(VL_EINVAL, "null argument");
  *failures = p->failures;
  return VL_OK;
}

}  // extern "C"

// ---- query server --------------------------------------------------------------
namespace {

// A layer binned into a uniform cell grid whose cells are numbered in Morton
// (Z) order, so a window's cells form few long runs of entries and nearby
// cells sit together in memory (row-major numbering when the Morton square
// would be far larger than the grid); cell starts also on the host.
struct CellIndex {
  Grid g{};
  vl_box* entries = nullptr;
  uint64_t* d_start = nullptr;          // 2*cells+1 (layer A keys only)
  std::vector<uint64_t> start;          // host copy
  int64_t nboxes = 0;
  Scratch* keep = nullptr;
};

// Morton-number runs [z0, z1) of the cells of columns c1..c2 and rows r1..r2:
// the quadtree over the 2^level square at (r0, c0), numbers from zbase, is
// taken whole when the window holds it and split into its four quadrants when
// the window cuts it.  Runs come out in increasing order; adjacent ones merge.
void zruns(int level, uint32_t zbase, int r0, int c0, int c1, int c2, int r1, int r2,
           std::vector<std::pair<uint32_t, uint32_t>>& out) {
  const int size = 1 << level;
  if (c0 > c2 || c0 + size - 1 < c1 || r0 > r2 || r0 + size - 1 < r1) return;
  if (c0 >= c1 && c0 + size - 1 <= c2 && r0 >= r1 && r0 + size - 1 <= r2) {
    const uint32_t z1 = zbase + (uint32_t)size * (uint32_t)size;
    if (!out.empty() && out.back().second == zbase) out.back().second = z1;
    else out.push_back({zbase, z1});
    return;
  }
  const int h = size >> 1;
  const uint32_t q = (uint32_t)h * (uint32_t)h;
  zruns(level - 1, zbase, r0, c0, c1, c2, r1, r2, out);
  zruns(level - 1, zbase + q, r0, c0 + h, c1, c2, r1, r2, out);
  zruns(level - 1, zbase + 2 * q, r0 + h, c0, c1, c2, r1, r2, out);
  zruns(level - 1, zbase + 3 * q, r0 + h, c0 + h, c1, c2, r1, r2, out);
}

// Entry ranges of the cells a window covers: one per Morton run (one per cell
// row with row-major numbering).
void window_ranges(const CellIndex& ix, const vl_box& w, std::vector<uint64_t>& rbeg, std::vector<uint64_t>& roff) {
  rbeg.clear();
  roff.assign(1, 0);
  if (ix.nboxes == 0) return;
  const Grid& g = ix.g;
  const int64_t x1 = std::max<int64_t>(w.x1, g.x0), x2 = std::min<int64_t>(w.x2, g.x1);
  const int64_t y1 = std::max<int64_t>(w.y1, g.y0), y2 = std::min<int64_t>(w.y2, g.y1);
  if (x1 >= x2 || y1 >= y2) return;
  const int c1 = (int)((x1 - g.x0) / g.tw), c2 = (int)((x2 - 1 - g.x0) / g.tw);
  const int r1 = (int)((y1 - g.y0) / g.th), r2 = (int)((y2 - 1 - g.y0) / g.th);
  auto add = [&](uint64_t b, uint64_t e) {
    if (e > b) {
      rbeg.push_back(b);
      roff.push_back(roff.back() + (e - b));
    }
  };
  if (g.zb) {
    std::vector<std::pair<uint32_t, uint32_t>> runs;
    zruns(g.zb, 0u, 0, 0, c1, c2, r1, r2, runs);
    for (const auto& z : runs) add(ix.start[2 * (uint64_t)z.first], ix.start[2 * (uint64_t)z.second]);
    return;
  }
  for (int r = r1; r <= r2; r++) {
    const int64_t t1 = (int64_t)r * g.cols + c1, t2 = (int64_t)r * g.cols + c2;
    add(ix.start[2 * t1], ix.start[2 * t2 + 2]);
  }
}

// Boxes of the index inside w, clipped to w, into layer out (engine e's stream).
int window_extract(vl_engine* e, const CellIndex& ix, const vl_box& w, vl_layer* out) {
  std::vector<uint64_t> rbeg, roff;
  window_ranges(ix, w, rbeg, roff);
  const uint64_t n = roff.back();
  const int nr = (int)rbeg.size();
  if (n == 0) return set_empty(e, out);
  Scratch s(e);
  uint64_t* d_rbeg = s.get<uint64_t>(nr);
  uint64_t* d_roff = s.get<uint64_t>(nr + 1);
  vl_box* clip = s.get<vl_box>(n);
  uint32_t* keep = s.get<uint32_t>(n + 1);
  uint32_t* pos = s.get<uint32_t>(n + 1);
  void* stmp = s.get<char>(scan_tmp_bytes(n + 1));
  if (s.oom) return fail(VL_ENOMEM, "window scratch");
  HIPCK(hipMemcpyAsync(d_rbeg, rbeg.data(), nr * sizeof(uint64_t), hipMemcpyHostToDevice, e->st));
  HIPCK(hipMemcpyAsync(d_roff, roff.data(), (nr + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, e->st));
  launch_window_clip(ix.entries, d_rbeg, d_roff, nr, n, w, ix.d_start, ix.g, clip, keep, e->st);
  scan_exclusive_u32(keep, pos, n, pos + n, stmp, e->st);
  uint32_t m = 0;
  VLCK(readback(e, &m, pos + n, sizeof(m)));
  if (m == 0) return set_empty(e, out);
  size_t r = 0;
  vl_box* d = (vl_box*)e->pool.get(m * sizeof(vl_box), &r);
  if (!d) return fail(VL_ENOMEM, "window layer");
  launch_window_compact(clip, keep, pos, n, d, e->st);
  HIPCK(hipGetLastError());
  if (out->d) e->pool.put(out->d, out->bytes);
  out->d = d;
  out->bytes = r;
  out->n = m;
  out->mbr_ok = false;
  out->version++;
  return VL_OK;
}

}  // namespace

struct vl_server {
  vl_engine* e = nullptr;
  std::vector<CellIndex> index;
  std::vector<vl_engine*> slots;
  uint64_t budget = 0;
};

namespace {

void release_layer(vl_engine* e, vl_layer& l) {
  if (l.d) e->pool.put(l.d, l.bytes);
  l.d = nullptr;
  l.n = 0;
  l.bytes = 0;
}

// One query on slot engine w: R = GROW_s(A ∩ W) AND (B ∩ W).
int run_query(vl_server* sv, vl_engine* w, const vl_query& q, vl_query_result& r) {
  std::lock_guard<std::recursive_mutex> lk(w->mu);
  HIPCK(hipSetDevice(w->device));
  if (q.layer_a < 0 || q.layer_a >= (int)sv->index.size() || q.layer_b < 0 || q.layer_b >= (int)sv->index.size())
    return fail(VL_EINVAL, "query layer index out of range");
  if (q.spacing < 0) return fail(VL_EINVAL, "negative spacing");
  if (q.window.x1 >= q.window.x2 || q.window.y1 >= q.window.y2) return fail(VL_EINVAL, "invalid window");
  vl_layer A, B, R;
  A.e = B.e = R.e = sv->e;
  int rc = window_extract(w, sv->index[q.layer_a], q.window, &A);
  if (rc == VL_OK) rc = window_extract(w, sv->index[q.layer_b], q.window, &B);
  r.in_boxes = A.n + B.n;
  const int32_t s = q.spacing;
  // R feeds ASSERT_ZERO: the grown boxes of A are tested against B for an
  // overlap first (GROW of a union is the union of the grown boxes), and R is
  // computed only from the boxes that take part in one
  Box64 ma, mb;
  if (rc == VL_OK && A.n > 0 && B.n > 0) rc = mbr2(w, A.d, A.n, Xform{s, s, s, s}, B.d, B.n, &ma, &mb);
  if (rc == VL_OK) {
    if (A.n == 0 || B.n == 0) rc = set_empty(w, &R);
    else rc = and_check_core(w, Side{A.d, A.n, Xform{s, s, s, s}}, Side{B.d, B.n, Xform{0, 0, 0, 0}}, ma, mb, &R);
  }
  if (rc == VL_OK && R.n == 0) {
    r.count = 0;
    r.hash = 0;   // the hash of no boxes
  } else if (rc == VL_OK) {
    r.count = R.n;
    Scratch hs(w);
    unsigned long long* h = hs.get<unsigned long long>(1);
    if (!h) {
      rc = fail(VL_ENOMEM, "hash scratch");
    } else {
      launch_fill_u64((uint64_t*)h, 0ull, 1, w->st);
      launch_box_hash(R.d, R.n, h, w->st);
      uint64_t hv = 0;
      rc = readback(w, &hv, h, sizeof(hv));
      r.hash = hv;
    }
  }
  release_layer(w, A);
  release_layer(w, B);
  release_layer(w, R);
  return rc;
}

}  // namespace

extern "C" {

int vl_server_create(vl_engine* e, const vl_layer* const* layers, int nlayers, int slots, uint64_t hbm_budget,
                     vl_server** out) {
  if (!e || !out || (nlayers > 0 && !layers) || nlayers < 0) return fail(VL_EINVAL, "bad argument");
  if (slots < 1 || slots > 16) return fail(VL_EINVAL, "slots must be 1..16");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  HIPCK(hipSetDevice(e->device));
  vl_server* sv = new vl_server();
  sv->e = e;
  sv->budget = hbm_budget ? hbm_budget : (uint64_t)64 << 30;
  e->refs.fetch_add(1);
  auto cleanup = [&](int rc) {
    vl_server_destroy(sv);
    return rc;
  };
  sv->index.resize(nlayers);
  for (int i = 0; i < nlayers; i++) {
    const vl_layer* l = layers[i];
    if (!l || l->e != e) return cleanup(fail(VL_EINVAL, "layer %d is null or belongs to another engine", i));
    CellIndex& ix = sv->index[i];
    ix.nboxes = l->n;
    if (l->n == 0) continue;
    Box64 m;
    int rc = layer_mbr(e, l, &m);
    if (rc != VL_OK) return cleanup(rc);
    ix.g = make_grid(m, l->n, 64);   // about 64 boxes per cell
    {
      int zb = 0;
      while ((1 << zb) < std::max(ix.g.cols, ix.g.rows)) zb++;
      // Morton numbering when its square stays within 8x the cells (and 2^30 keys)
      if (zb <= 14 && ((int64_t)1 << (2 * zb)) <= 8 * (int64_t)ix.g.cols * ix.g.rows) ix.g.zb = std::max(zb, 1);
    }
    ix.keep = new Scratch(e);
    uint64_t total = 0;
    rc = bin_entries(e, ix.g, l->d, l->n, Xform{0, 0, 0, 0}, Xform{0, 0, 0, 0}, nullptr, 0, *ix.keep, &ix.entries,
                     &ix.d_start, &total, ix.g.zb == 0);
    if (rc == BIN_TOO_MANY) {   // boxes spanning many cells: row-major cells, coarsened
      g_err.clear();
      delete ix.keep;
      ix.keep = new Scratch(e);
      ix.g = make_grid(m, l->n, 64);
      rc = bin_entries(e, ix.g, l->d, l->n, Xform{0, 0, 0, 0}, Xform{0, 0, 0, 0}, nullptr, 0, *ix.keep, &ix.entries,
                       &ix.d_start, &total);
    }
    if (rc != VL_OK) return cleanup(rc);
    const size_t ns = 2 * (size_t)grid_tiles(ix.g) + 1;
    ix.start.resize(ns);
    if (hipMemcpyAsync(ix.start.data(), ix.d_start, ns * sizeof(uint64_t), hipMemcpyDeviceToHost, e->st) !=
            hipSuccess ||
        hipStreamSynchronize(e->st) != hipSuccess)
      return cleanup(fail(VL_EHIP, "index readback"));
  }
  for (int k = 0; k < slots; k++) {
    vl_engine* w = nullptr;
    int rc = engine_new(e->device, &w);
    if (rc != VL_OK) return cleanup(rc);
    w->root = e;
    w->pool.fallback = &e->pool;
    w->nmax = e->nmax;
    w->target = e->target;
    w->erode_target = e->erode_target;
    w->level_factor = e->level_factor;
    w->erode_kernel = e->erode_kernel;
    w->bin_mode = e->bin_mode;
    w->check_target = e->check_target;
    sv->slots.push_back(w);
  }
  *out = sv;
  return VL_OK;
}

int vl_server_destroy(vl_server* sv) {
  if (!sv) return VL_OK;
  vl_engine* e = sv->e;
  {
    std::lock_guard<std::recursive_mutex> lk(e->mu);
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->st);
    for (CellIndex& ix : sv->index) delete ix.keep;
  }
  for (vl_engine* w : sv->slots) engine_release(w);
  delete sv;
  engine_release(e);
  return VL_OK;
}

int vl_server_run(vl_server* sv, const vl_query* q, int64_t n, vl_query_result* res) {
  if (!sv || (n > 0 && (!q || !res)) || n < 0) return fail(VL_EINVAL, "bad argument");
  using clk = std::chrono::steady_clock;
  // in-flight memory estimate of a query: its candidate boxes times the
  // per-box footprint of extraction, binning and tile scratch
  std::vector<uint64_t> est(n);
  {
    std::vector<uint64_t> rb, ro;
    for (int64_t i = 0; i < n; i++) {
      uint64_t c = 0;
      for (int32_t li : {q[i].layer_a, q[i].layer_b})
        if (li >= 0 && li < (int)sv->index.size()) {
          window_ranges(sv->index[li], q[i].window, rb, ro);
          c += ro.back();
        }
      est[i] = 256 * c + ((uint64_t)8 << 20);
    }
  }
  std::mutex mu;
  std::condition_variable cv;
  std::vector<int64_t> pending;        // released, not yet admitted (arrival order)
  size_t head = 0;
  bool all_released = false;
  uint64_t used = 0;
  int inflight = 0;
  int first_rc = VL_OK;
  const clk::time_point t0 = clk::now();
  auto ms_since = [&](clk::time_point t) { return std::chrono::duration<double, std::milli>(t - t0).count(); };
  std::vector<std::thread> th;
  for (int k = 0; k < (int)sv->slots.size(); k++) {
    th.emplace_back([&, k]() {
      vl_engine* w = sv->slots[k];
      (void)hipSetDevice(w->device);
      for (;;) {
        int64_t i = -1;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&]() {
            if (head < pending.size()) {
              const int64_t c = pending[head];
              if (inflight == 0 || used + est[c] <= sv->budget) return true;
              return false;
            }
            return all_released;
          });
          if (head >= pending.size()) return;   // all released and none left
          i = pending[head++];
          used += est[i];
          inflight++;
        }
        vl_query_result& r = res[i];
        r = vl_query_result{};
        r.slot = k;
        r.start_ms = ms_since(clk::now());
        r.status = run_query(sv, w, q[i], r);
        r.end_ms = ms_since(clk::now());
        {
          std::lock_guard<std::mutex> lk(mu);
          used -= est[i];
          inflight--;
          if (r.status != VL_OK && first_rc == VL_OK) first_rc = r.status;
        }
        cv.notify_all();
      }
    });
  }
  // release requests at their arrival times
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return q[a].arrival_ms < q[b].arrival_ms; });
  for (int64_t i : order) {
    const clk::time_point due = t0 + std::chrono::duration_cast<clk::duration>(
                                         std::chrono::duration<double, std::milli>(std::max(0.0, q[i].arrival_ms)));
    std::this_thread::sleep_until(due);
    {
      std::lock_guard<std::mutex> lk(mu);
      pending.push_back(i);
    }
    cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    all_released = true;
  }
  cv.notify_all();
  for (auto& t : th) t.join();
  return first_rc;
}

}  // extern "C"

// ---- autotuner -------------------------------------------------------------------
extern "C" {

// Picks the tile box budget of the operator grid ("tile_target"), of the
// erosion grid ("erode_target") and of a fused group's grid ("fuse_target") by
// timing candidates on the caller's layers: A NOT B (or the merge of A) plus
// GROW(A; 1) for the operator grid, SIZE(A; -1) for the erosion grid, and the
// fused ring of A (GROW(A; 1), SIZE(A; -1), GROW(A; 1) NOT A, A NOT
// SIZE(A; -1)) for the fused grid, then whether programs fuse groups at all
// ("fuse": the same ring as four single operators against the fused group);
// each candidate is timed twice (device time, HIP events) and the faster run
// counts.  Results never depend on these parameters, only run time.
int vl_engine_autotune(vl_engine* e, const vl_layer* a, const vl_layer* b, int64_t* target, int64_t* erode_target) {
  if (!e || !a) return fail(VL_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(e->mu);
  vl_layer* c = nullptr;
  VLCK(vl_layer_create(e, &c));
  auto best_of = [&](const std::function<int()>& f, double* ms) {
    double m = 1e30;
    for (int rep = 0; rep < 2; rep++) {
      int rc = f();
      if (rc != VL_OK) return rc;
      m = std::min(m, e->last_ms);
    }
    *ms = m;
    return VL_OK;
  };
  int rc = VL_OK;
  const int targets[] = {12, 16, 20, 24, 32};
  const int etargets[] = {8, 12, 16};
  const int t0 = e->target, et0 = e->erode_target;
  int bt = t0, bet = et0;
  double bms = 1e30;
  for (int t : targets) {
    e->target = t;
    double m1 = 0, m2 = 0;
    rc = best_of([&]() { return b ? vl_not(e, a, b, c) : vl_merge(e, a, c); }, &m1);
    if (rc == VL_OK) rc = best_of([&]() { return vl_grow(e, a, 1, 1, 1, 1, c); }, &m2);
    if (rc != VL_OK) break;
    if (m1 + m2 < bms) {
      bms = m1 + m2;
      bt = t;
    }
  }
  e->target = bt;
  bms = 1e30;
  for (int et : etargets) {
    if (rc != VL_OK) break;
    e->erode_target = et;
    double m = 0;
    rc = best_of([&]() { return vl_size(e, a, -1, c); }, &m);
    if (rc == VL_OK && m < bms) {
      bms = m;
      bet = et;
    }
  }
  e->erode_target = bet;
  // fused ring of A: four outputs over S = A and the images GROW 1, erode 1
  // (the current value stays unless a candidate is more than 3 % faster:
  // neighbouring budgets differ by less than the run-to-run noise)
  const int ftargets[] = {4, 5, 6, 7, 8, 10};
  const int ft0 = e->fuse_target, fu0 = e->fuse, fu20 = e->fuse2;
  int bft = ft0;
  double ft0_ms = 1e30;
  {
    FusedSpec fs{};
    fs.nmorph = 2;
    fs.kind[0] = MORPH_GROW;
    fs.kind[1] = MORPH_ERODE;
    for (int k = 0; k < 4; k++) fs.arg[0][k] = 1;
    fs.arg[1][0] = 1;
    fs.halo = 1;
    fs.nout = 4;
    fs.tt[0] = 0xCCu;                  // GROW(A)
    fs.tt[1] = 0xF0u;                  // SIZE(A, -1)
    fs.tt[2] = 0xCCu & ~0xAAu;         // GROW(A) NOT A
    fs.tt[3] = 0xAAu & ~0xF0u;         // A NOT SIZE(A, -1)
    vl_layer* outs[4] = {};
    for (int k = 0; k < 4 && rc == VL_OK; k++) rc = vl_layer_create(e, &outs[k]);
    bms = 1e30;
    for (int ft : ftargets) {
      if (rc != VL_OK) break;
      e->fuse_target = ft;
      double m = 0;
      bool declined = false;
      rc = best_of([&]() {
        HIPCK(hipSetDevice(e->device));
        OpTimer tm(e);
        const int r = run_fused(e, a, nullptr, fs, outs);
        if (r == FUSE_DECLINED) declined = true;
        return r == FUSE_DECLINED ? VL_OK : r;
      }, &m);
      if (declined) m = 1e30;   // a declined candidate times nothing useful
      if (rc == VL_OK && ft == ft0) ft0_ms = m;
      if (rc == VL_OK && m < bms) {
        bms = m;
        bft = ft;
      }
    }
    if (bms > 0.97 * ft0_ms) {
      bft = ft0;
      bms = ft0_ms;
    }
    e->fuse_target = bft;
    // fused groups at all: the same ring as four single operators (GROW,
    // SIZE and the two differences) against the fused group at its best
    // budget; the current mode stays unless the other is more than 3 % faster
    if (rc == VL_OK) {
      double mu = 1e30;
      for (int rep = 0; rep < 2 && rc == VL_OK; rep++) {
        double acc = 0;
        rc = vl_grow(e, a, 1, 1, 1, 1, outs[0]);
        acc += e->last_ms;
        if (rc == VL_OK) rc = vl_size(e, a, -1, outs[1]);
        acc += e->last_ms;
        if (rc == VL_OK) rc = vl_not(e, outs[0], a, outs[2]);
        acc += e->last_ms;
        if (rc == VL_OK) rc = vl_not(e, a, outs[1], outs[3]);
        acc += e->last_ms;
        mu = std::min(mu, acc);
      }
      if (rc == VL_OK) {
        const bool fused_now = e->fuse != 0;
        const double cur = fused_now ? bms : mu, alt = fused_now ? mu : bms;
        if (alt < 0.97 * cur) e->fuse = fused_now ? 0 : 1;
      }
    }
    // two-layer groups: AND, NOT and OR of A and B in one fused pass against
    // the three single operators (the current mode stays unless the other
    // is more than 3 % faster; a declined group counts as slower)
    if (rc == VL_OK && b && b->n > 0) {
      FusedSpec f2{};
      f2.nmorph = 1;
      f2.kind[0] = MORPH_LAYER;
      f2.nout = 3;
      f2.tt[0] = 0xAAu & 0xCCu;
      f2.tt[1] = 0xAAu & ~0xCCu & 0xFFu;
      f2.tt[2] = 0xAAu | 0xCCu;
      double mf = 1e30, ms = 1e30;
      bool declined = false;
      rc = best_of([&]() {
        HIPCK(hipSetDevice(e->device));
        OpTimer tm(e);
        const int r = run_fused(e, a, b, f2, outs);
        if (r == FUSE_DECLINED) declined = true;
        return r == FUSE_DECLINED ? VL_OK : r;
      }, &mf);
      if (declined) mf = 1e30;
      for (int rep = 0; rep < 2 && rc == VL_OK; rep++) {
        double acc = 0;
        rc = vl_and(e, a, b, outs[0]);
        acc += e->last_ms;
        if (rc == VL_OK) rc = vl_not(e, a, b, outs[1]);
        acc += e->last_ms;
        if (rc == VL_OK) rc = vl_or(e, a, b, outs[2]);
        acc += e->last_ms;
        ms = std::min(ms, acc);
      }
      if (rc == VL_OK) {
        const bool now = e->fuse2 != 0;
        const double cur = now ? mf : ms, alt = now ? ms : mf;
        if (alt < 0.97 * cur) e->fuse2 = now ? 0 : 1;
      }
    }
    for (int k = 0; k < 4; k++)
      if (outs[k]) vl_layer_destroy(outs[k]);
  }
  // Binning mode, radix digit width and the tile-kernel occupancy cap, each
  // timed on the operators it affects; the current value stays unless a
  // candidate is more than 3 % faster.
  const int bm0 = e->bin_mode, rb0 = e->radix_bits, mw0 = e->max_wg_cap;
  auto pick = [&](int* field, std::initializer_list<int> cands, const std::function<int(double*)>& f) {
    const int v0 = *field;
    int best = v0;
    double ms0 = 1e30, bms2 = 1e30;
    for (int v : cands) {
      *field = v;
      double m = 1e30;
      for (int rep = 0; rep < 2; rep++) {
        double acc = 0;
        const int r = f(&acc);
        if (r != VL_OK) {
          *field = v0;
          return r;
        }
        m = std::min(m, acc);
      }
      if (v == v0) ms0 = m;
      if (m < bms2) {
        bms2 = m;
        best = v;
      }
    }
    *field = bms2 > 0.97 * ms0 ? v0 : best;
    return VL_OK;
  };
  // binary operators of A and B (or A's merge) plus A's GROW: binning, tile
  // kernels and the x-stitch sort
  auto ops_ab = [&](double* acc) {
    int r = b ? vl_not(e, a, b, c) : vl_merge(e, a, c);
    *acc += e->last_ms;
    if (r == VL_OK && b) {
      r = vl_and(e, a, b, c);
      *acc += e->last_ms;
    }
    if (r == VL_OK) {
      r = vl_grow(e, a, 1, 1, 1, 1, c);
      *acc += e->last_ms;
    }
    return r;
  };
  if (rc == VL_OK) rc = pick(&e->bin_mode, {2, 1}, ops_ab);
  if (rc == VL_OK) rc = pick(&e->radix_bits, {8, 11}, ops_ab);
  if (rc == VL_OK) rc = pick(&e->max_wg_cap, {0, 4096, 2048}, ops_ab);
  if (rc != VL_OK) {
    e->target = t0;
    e->erode_target = et0;
    e->fuse_target = ft0;
    e->bin_mode = bm0;
    e->radix_bits = rb0;
    e->max_wg_cap = mw0;
    e->fuse = fu0;
    e->fuse2 = fu20;
  }
  vl_layer_destroy(c);
  if (target) *target = e->target;
  if (erode_target) *erode_target = e->erode_target;
  return rc;
}

}  // extern "C"
