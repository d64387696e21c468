// vl_noalloc.hip — the event sort for callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc): vl_event_sort_noalloc.
//
// Its scratch is one file-scope device array, `scratch`, of
// VL_NOALLOC_SCRATCH_BYTES: every device the library runs on holds its own
// copy until the process ends.  The host never passes the array's name to a
// kernel or a copy (in host code that name is only a stand-in): at set-up,
// once per device and with that device current, hipGetSymbolAddress gives the
// device address, which is kept per device and used everywhere instead.  Set-
// up runs when an engine is created on the device (vl_engine_create) and, for
// safety, again on an entry point's first call there.
//
// The sort itself is the library's LSD radix sort run for all eight 8-bit
// digits of the 64-bit key (no histogram pre-pass, so no host wait inside),
// then a gather of the 12-byte records.  No step needs zeroed scratch: every
// digit pass writes all of its histograms and offsets.  Calls that overlap on
// one device share the array and race; calls on one engine are serialised
// by the engine's lock.
#include <cstring>
#include <mutex>

#include "vl_internal.h"

namespace vl {

namespace {

constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[VL_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte layout of the scratch for n trackers: keys, values, their alternates,
// a copy of the records, the radix temporaries.
struct Layout {
  size_t keys, vals, keys2, vals2, copy, tmp, end;
};
Layout layout(int64_t n) {
  Layout l;
  size_t o = 0;
  l.keys = o;  o += align256((size_t)n * 8);
  l.vals = o;  o += align256((size_t)n * 4);
  l.keys2 = o; o += align256((size_t)n * 8);
  l.vals2 = o; o += align256((size_t)n * 4);
  l.copy = o;  o += align256((size_t)n * sizeof(vl_tracker));
  l.tmp = o;   o += align256(radix_plan(n).tmp_bytes);
  l.end = o;
  return l;
}

}  // namespace

// Device address of this device's scratch (the current device must be
// `device`).  Idempotent; returns VL_OK or VL_EHIP.
int noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return set_error(VL_EINVAL, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return VL_OK;
  void* p = nullptr;
  const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (r != hipSuccess || !p)
    return set_error(VL_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(r));
  g_ptr[device] = p;
  return VL_OK;
}

}  // namespace vl

using namespace vl;

extern "C" int vl_event_sort_noalloc(vl_engine* e, vl_tracker* d_arr, int64_t n) {
  if (!e || n < 0 || (n > 0 && !d_arr)) return set_error(VL_EINVAL, "bad argument");
  if (n > VL_NOALLOC_MAX_TRACKERS)
    return set_error(VL_EINVAL, "%lld trackers: vl_event_sort_noalloc sorts at most %d (its static scratch is %d bytes)",
                     (long long)n, VL_NOALLOC_MAX_TRACKERS, VL_NOALLOC_SCRATCH_BYTES);
  if (n <= 1) return VL_OK;
  EngineLock el(e);
  int rc = el.begin();
  if (rc != VL_OK) return rc;
  hipPointerAttribute_t at;
  memset(&at, 0, sizeof(at));
  if (hipPointerGetAttributes(&at, d_arr) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return set_error(VL_EINVAL, "vl_event_sort_noalloc takes device memory only");
  }
  rc = noalloc_setup(el.device());
  if (rc != VL_OK) return rc;
  char* base;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = (char*)g_ptr[el.device()];
  }
  const Layout l = layout(n);
  if (l.end > (size_t)VL_NOALLOC_SCRATCH_BYTES)
    return set_error(VL_EINTERNAL, "scratch layout of %zu bytes exceeds the array", l.end);
  const hipStream_t st = el.stream();
  uint64_t* k = (uint64_t*)(base + l.keys);
  uint32_t* v = (uint32_t*)(base + l.vals);
  uint64_t* k2 = (uint64_t*)(base + l.keys2);
  uint32_t* v2 = (uint32_t*)(base + l.vals2);
  vl_tracker* src = (vl_tracker*)(base + l.copy);
  launch_tracker_keys(d_arr, k, v, n, st);
  const bool alt = radix_sort_pairs_bits(k, v, k2, v2, n, 64, base + l.tmp, st);
  if (hipMemcpyAsync(src, d_arr, (size_t)n * sizeof(vl_tracker), hipMemcpyDeviceToDevice, st) != hipSuccess)
    return set_error(VL_EHIP, "copy: %s", hipGetErrorString(hipGetLastError()));
  launch_tracker_gather(src, alt ? v2 : v, d_arr, n, st);
  const hipError_t r = hipGetLastError();
  if (r != hipSuccess) return set_error(VL_EHIP, "launch: %s", hipGetErrorString(r));
  if (hipStreamSynchronize(st) != hipSuccess) return set_error(VL_EHIP, "stream: %s", hipGetErrorString(hipGetLastError()));
  return VL_OK;
}
