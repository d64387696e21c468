// vl_io.hip — layer files: KLBIN (src/kldata.f90:14-97) and the SNAP
// container (src/serialization.f90:15-56, src/box_codec.f90:197-894).
//
// SNAP file (stream, little endian), as SaveCompressedStreamToDisk writes it:
//   int32 codec version (1), int32 layer properties (LAYER_STATE_* bits),
//   int64 total boxes, int64 chunks, int32 compression method
//   (1 Snappy, 2 zlib, 3 zstd), [int64 polygon number per box when the PNUM
//   bit is set], then per chunk int64 boxes, int64 raw bytes, int64 stored
//   bytes and the stored bytes, then the chunks' 1-based file positions
//   (int64 each) and, last, the 1-based position of that table (int64).
// A chunk holds 65536 boxes (the last one the rest), codec-encoded (the
// W/H-dictionary + delta format of box_codec.f90:565-666, encoded here on the
// host exactly as the reference does, so the bytes agree) and then
// compressed.  Reading decompresses the chunks on the host (one thread per
// chunk) and decodes the codec on the GPU (vl_snap.hip) into the layer's
// device buffer.  Snappy is this file's own implementation of the published
// raw Snappy format; zlib is the system library (compress / uncompress, as
// the reference calls them); zstd is loaded at run time from libzstd.so.1
// when a file needs it.
#include <dlfcn.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vl_internal.h"

namespace vl {

namespace {

constexpr int64_t BOXES_PER_CHUNK = 65536;
constexpr int32_t CODEC_VERSION = 1;
constexpr int32_t LAYER_STATE_PNUM_BIT = 4;     // src/design.f90:41
constexpr int32_t LAYER_STATE_RTREE_BIT = 8;    // src/design.f90:42
constexpr int32_t LAYER_STATE_EVERYTHING = 63;  // src/design.f90:47
enum : int32_t { METHOD_NONE = 0, METHOD_SNAPPY = 1, METHOD_ZLIB = 2, METHOD_ZSTD = 3 };

// ---- the box codec (encode), one chunk -------------------------------------
void put_varint(std::vector<uint8_t>& o, uint64_t v) {
  for (;;) {
    const uint8_t b = (uint8_t)(v & 127u);
    v >>= 7;
    if (!v) {
      o.push_back(b);
      return;
    }
    o.push_back((uint8_t)(b | 128u));
  }
}
uint64_t zigzag(int32_t v) {
  const int64_t t = v;
  return (uint64_t)(t < 0 ? (-t) * 2 - 1 : t * 2);
}

// Dictionary of distinct values: 1-based index of the first entry of a value.
struct Dict {
  std::unordered_map<int32_t, int64_t> first;
  int64_t n = 0;
  int64_t find(int32_t v) const {
    auto it = first.find(v);
    return it == first.end() ? -1 : it->second;
  }
  int64_t add(int32_t v) {
    ++n;
    first.emplace(v, n);   // a value added again keeps its first index
    return n;
  }
};

void encode_chunk(const vl_box* b, int64_t n, std::vector<uint8_t>& out) {
  out.clear();
  out.reserve((size_t)n * 6);
  Dict dw, dh;
  int64_t cw = 0, ch = 0;
  int32_t cx = 0, cy = 0;
  for (int64_t i = 0; i < n; i++) {
    const int32_t w = (int32_t)((uint32_t)b[i].x2 - (uint32_t)b[i].x1);
    const int32_t h = (int32_t)((uint32_t)b[i].y2 - (uint32_t)b[i].y1);
    uint8_t info = 0;
    const bool sq = w == h;
    if (sq) info |= 16;
    const int32_t dx = (int32_t)((uint32_t)b[i].x1 - (uint32_t)cx);
    if (dx) {
      info |= 32;
      cx = b[i].x1;
    }
    const int32_t dy = (int32_t)((uint32_t)b[i].y1 - (uint32_t)cy);
    if (dy) {
      info |= 64;
      cy = b[i].y1;
    }
    // An index is written as int16: an entry past 32767 that would have to be
    // named by index is added again instead (the reference writes a wrapped
    // index its own reader rejects; wherever its stream is readable, these
    // bytes equal it).
    int64_t wi = dw.find(w);
    if (wi != cw || dw.n == 0) {
      info |= 1;
      if (wi <= 0 || wi > 32767) {
        info |= 2;
        wi = dw.add(w);
      }
      cw = wi;
    }
    int64_t hi = 0;
    if (!sq) {
      hi = dh.find(h);
      if (hi != ch || dh.n == 0) {
        info |= 4;
        if (hi <= 0 || hi > 32767) {
          info |= 8;
          hi = dh.add(h);
        }
        ch = hi;
      }
    }
    out.push_back(info);
    if (info & 32) put_varint(out, zigzag(dx));
    if (info & 64) put_varint(out, zigzag(dy));
    if (info & 1) {
      if (info & 2) {
        put_varint(out, (uint64_t)(int64_t)w);
      } else {
        out.push_back((uint8_t)(wi & 255));
        out.push_back((uint8_t)((wi >> 8) & 255));
      }
    }
    if (!sq && (info & 4)) {
      if (info & 8) {
        put_varint(out, (uint64_t)(int64_t)h);
      } else {
        out.push_back((uint8_t)(hi & 255));
        out.push_back((uint8_t)((hi >> 8) & 255));
      }
    }
  }
}

// ---- Snappy (raw format) ---------------------------------------------------
// Stream: varint uncompressed length, then elements; tag & 3: 0 literal
// (length-1 in the tag's upper 6 bits, or 60..63 for 1..4 length bytes), 1 copy
// with a 1-byte offset (length 4..11, offset's bits 8..10 in the tag), 2 copy
// with a 2-byte offset, 3 copy with a 4-byte offset (length 1..64).
void snappy_literal(std::vector<uint8_t>& o, const uint8_t* p, size_t n) {
  while (n > 0) {
    const size_t k = std::min<size_t>(n, (size_t)1 << 16);
    const size_t l = k - 1;
    if (l < 60) {
      o.push_back((uint8_t)(l << 2));
    } else if (l < 256) {
      o.push_back((uint8_t)(60 << 2));
      o.push_back((uint8_t)l);
    } else {
      o.push_back((uint8_t)(61 << 2));
      o.push_back((uint8_t)(l & 255));
      o.push_back((uint8_t)(l >> 8));
    }
    o.insert(o.end(), p, p + k);
    p += k;
    n -= k;
  }
}
void snappy_copy(std::vector<uint8_t>& o, size_t offset, size_t len) {
  while (len > 0) {
    // pieces of at most 64; never leave a remainder below 4
    size_t k = std::min<size_t>(len, 64);
    if (len > 64 && len - 64 < 4) k = 60;
    o.push_back((uint8_t)(((k - 1) << 2) | 2));
    o.push_back((uint8_t)(offset & 255));
    o.push_back((uint8_t)(offset >> 8));
    len -= k;
  }
}
void snappy_compress(const uint8_t* in, size_t n, std::vector<uint8_t>& o) {
  o.clear();
  {
    uint64_t v = n;
    for (;;) {
      const uint8_t b = (uint8_t)(v & 127u);
      v >>= 7;
      if (!v) {
        o.push_back(b);
        break;
      }
      o.push_back((uint8_t)(b | 128u));
    }
  }
  constexpr int HB = 14;
  std::vector<int64_t> table((size_t)1 << HB, -1);
  size_t lit = 0, i = 0;
  auto load32 = [&](size_t k) {
    uint32_t v;
    memcpy(&v, in + k, 4);
    return v;
  };
  while (i + 4 <= n) {
    const uint32_t v = load32(i);
    const uint32_t hsh = (v * 0x1E35A7BDu) >> (32 - HB);
    const int64_t cand = table[hsh];
    table[hsh] = (int64_t)i;
    if (cand >= 0 && i - (size_t)cand <= 65535 && load32((size_t)cand) == v) {
      size_t len = 4;
      while (i + len < n && in[(size_t)cand + len] == in[i + len]) len++;
      snappy_literal(o, in + lit, i - lit);
      snappy_copy(o, i - (size_t)cand, len);
      i += len;
      lit = i;
    } else {
      i++;
    }
  }
  snappy_literal(o, in + lit, n - lit);
}
bool snappy_uncompress(const uint8_t* in, size_t n, uint8_t* out, size_t cap, size_t* got) {
  size_t p = 0;
  uint64_t ulen = 0;
  for (int sh = 0;; sh += 7) {
    if (p >= n || sh > 63) return false;
    const uint8_t b = in[p++];
    ulen |= (uint64_t)(b & 127u) << sh;
    if (!(b & 128u)) break;
  }
  if (ulen > cap) return false;
  size_t q = 0;
  while (p < n) {
    const uint8_t tag = in[p++];
    size_t len, off = 0;
    switch (tag & 3) {
      case 0: {
        len = (size_t)(tag >> 2) + 1;
        if (len > 60) {
          const size_t nb = len - 60;
          if (p + nb > n) return false;
          len = 0;
          for (size_t k = 0; k < nb; k++) len |= (size_t)in[p + k] << (8 * k);
          len += 1;
          p += nb;
        }
        if (p + len > n || q + len > ulen) return false;
        memcpy(out + q, in + p, len);
        p += len;
        q += len;
        continue;
      }
      case 1:
        if (p + 1 > n) return false;
        len = (size_t)((tag >> 2) & 7) + 4;
        off = ((size_t)(tag >> 5) << 8) | in[p++];
        break;
      case 2:
        if (p + 2 > n) return false;
        len = (size_t)(tag >> 2) + 1;
        off = (size_t)in[p] | ((size_t)in[p + 1] << 8);
        p += 2;
        break;
      default:
        if (p + 4 > n) return false;
        len = (size_t)(tag >> 2) + 1;
        off = (size_t)in[p] | ((size_t)in[p + 1] << 8) | ((size_t)in[p + 2] << 16) | ((size_t)in[p + 3] << 24);
        p += 4;
        break;
    }
    if (off == 0 || off > q || q + len > ulen) return false;
    for (size_t k = 0; k < len; k++, q++) out[q] = out[q - off];   // copies may overlap
  }
  *got = q;
  return q == ulen;
}

// ---- zstd, loaded when a file needs it ----------------------------------------
struct Zstd {
  size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
  size_t (*decompress)(void*, size_t, const void*, size_t) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
  size_t (*bound)(size_t) = nullptr;
  bool ok = false;
};
const Zstd& zstd() {
  static Zstd z;
  static std::once_flag once;
  std::call_once(once, [] {
    void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    z.compress = (size_t(*)(void*, size_t, const void*, size_t, int))dlsym(h, "ZSTD_compress");
    z.decompress = (size_t(*)(void*, size_t, const void*, size_t))dlsym(h, "ZSTD_decompress");
    z.is_error = (unsigned (*)(size_t))dlsym(h, "ZSTD_isError");
    z.bound = (size_t(*)(size_t))dlsym(h, "ZSTD_compressBound");
    z.ok = z.compress && z.decompress && z.is_error && z.bound;
  });
  return z;
}

bool compress_chunk(int32_t method, const std::vector<uint8_t>& raw, std::vector<uint8_t>& out) {
  switch (method) {
    case METHOD_SNAPPY:
      snappy_compress(raw.data(), raw.size(), out);
      return true;
    case METHOD_ZLIB: {
      uLongf cap = compressBound((uLong)raw.size());
      out.resize(cap);
      if (compress(out.data(), &cap, raw.data(), (uLong)raw.size()) != Z_OK) return false;
      out.resize(cap);
      return true;
    }
    case METHOD_ZSTD: {
      const Zstd& z = zstd();
      if (!z.ok) return false;
      out.resize(z.bound(raw.size()));
      const size_t r = z.compress(out.data(), out.size(), raw.data(), raw.size(), 20);   // the reference's level
      if (z.is_error(r)) return false;
      out.resize(r);
      return true;
    }
    default:
      return false;
  }
}
bool uncompress_chunk(int32_t method, const uint8_t* in, size_t n, uint8_t* out, size_t raw) {
  switch (method) {
    case METHOD_SNAPPY: {
      size_t got = 0;
      return snappy_uncompress(in, n, out, raw, &got) && got == raw;
    }
    case METHOD_ZLIB: {
      uLongf got = (uLongf)raw;
      return uncompress(out, &got, in, (uLong)n) == Z_OK && got == raw;
    }
    case METHOD_ZSTD: {
      const Zstd& z = zstd();
      if (!z.ok) return false;
      const size_t r = z.decompress(out, raw, in, n);
      return !z.is_error(r) && r == raw;
    }
    case METHOD_NONE:
      if (n != raw) return false;
      memcpy(out, in, n);
      return true;
    default:
      return false;
  }
}

// Runs f(k) for k in [0, n) on up to hardware_concurrency threads.
template <typename F>
void parallel_for(int64_t n, F f) {
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(n, std::max(1u, std::thread::hardware_concurrency())));
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&] {
      for (int64_t k; (k = next.fetch_add(1)) < n;) f(k);
    });
  for (auto& t : th) t.join();
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(fopen(path, mode)) {}
  ~File() {
    if (f) fclose(f);
  }
};

template <typename T>
bool rd(FILE* f, T* v, size_t n = 1) {
  return fread(v, sizeof(T), n, f) == n;
}
template <typename T>
bool wr(FILE* f, const T* v, size_t n = 1) {
  return fwrite(v, sizeof(T), n, f) == n;
}

// Device scratch of one decode (freed on every path).
struct DevBufs {
  std::vector<void*> v;
  ~DevBufs() {
    for (void* p : v) (void)hipFree(p);
  }
  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    v.push_back(p);
    return (T*)p;
  }
};

}  // namespace

// Decodes concatenated codec chunks (host bytes) on the GPU into layer l.
int snap_decode_to_layer(vl_engine* e, vl_layer* l, const uint8_t* raw, const std::vector<int64_t>& chunk_bytes,
                         const std::vector<int64_t>& chunk_boxes) {
  const int nch = (int)chunk_bytes.size();
  std::vector<int64_t> coff(nch + 1, 0), roff(nch + 1, 0), boff(nch + 1, 0);
  for (int c = 0; c < nch; c++) {
    if (chunk_bytes[c] < 0 || chunk_boxes[c] < 0) return set_error(VL_EINVAL, "negative chunk size");
    if ((chunk_bytes[c] == 0) != (chunk_boxes[c] == 0))
      return set_error(VL_EINVAL, "chunk %d: %lld boxes in %lld bytes", c, (long long)chunk_boxes[c],
                       (long long)chunk_bytes[c]);
    coff[c + 1] = coff[c] + chunk_bytes[c];
    roff[c + 1] = roff[c] + chunk_boxes[c];
    boff[c + 1] = boff[c] + snap_blocks(chunk_bytes[c]);
  }
  const int64_t total_raw = coff[nch], nrec = roff[nch], nblocks = boff[nch];
  if (nrec == 0) return vl_layer_upload(l, nullptr, 0);
  if (nrec > ((int64_t)1 << 32) - 2) return set_error(VL_EINVAL, "more than 2^32 - 2 boxes");
  EngineLock el(e);
  int rc = el.begin();
  if (rc != VL_OK) return rc;
  const hipStream_t st = el.stream();
  DevBufs d;
  SnapDecode s{};
  uint8_t* draw = d.get<uint8_t>(total_raw);
  int64_t* dcoff = d.get<int64_t>(nch + 1);
  int64_t* droff = d.get<int64_t>(nch + 1);
  int64_t* dboff = d.get<int64_t>(nch + 1);
  s.len = d.get<uint8_t>(total_raw);
  s.exit_e = d.get<uint8_t>(nblocks * 41);
  s.cnt = d.get<uint16_t>(nblocks * 41);
  s.entry = d.get<uint8_t>(nblocks);
  s.rbase = d.get<int64_t>(nblocks);
  s.start = d.get<int64_t>(nrec);
  s.rx = d.get<int32_t>(nrec);
  s.ry = d.get<int32_t>(nrec);
  s.dictw = d.get<int32_t>(nrec);
  s.dicth = d.get<int32_t>(nrec);
  s.widx = d.get<uint32_t>(nrec);
  s.hidx = d.get<uint32_t>(nrec);
  s.out = d.get<vl_box>(nrec);
  s.err = d.get<uint32_t>(1);
  for (void* p : d.v)
    if (!p) return set_error(VL_ENOMEM, "decode scratch for %lld boxes", (long long)nrec);
  if (d.v.size() != 18) return set_error(VL_ENOMEM, "decode scratch for %lld boxes", (long long)nrec);
  if (hipMemcpyAsync(draw, raw, (size_t)total_raw, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dcoff, coff.data(), coff.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(droff, roff.data(), roff.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dboff, boff.data(), boff.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess)
    return set_error(VL_EHIP, "upload: %s", hipGetErrorString(hipGetLastError()));
  launch_fill_u32(s.err, 0u, 1, st);
  s.raw = draw;
  s.total_raw = total_raw;
  s.coff = dcoff;
  s.roff = droff;
  s.boff = dboff;
  s.nchunks = nch;
  s.nblocks = nblocks;
  hipEvent_t pa = engine_prof_mark(e);
  launch_snap_decode(s, st);
  hipEvent_t pb = engine_prof_mark(e);
  // algorithmic bytes: the codec stream read once, 16 B written per box
  engine_prof_add(e, pa, pb, KSTAT_SNAP, 5, (double)total_raw + 16.0 * (double)nrec);
  if (hipGetLastError() != hipSuccess) return set_error(VL_EHIP, "decode launch");
  uint32_t err = 0;
  if (hipMemcpyAsync(&err, s.err, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return set_error(VL_EHIP, "decode: %s", hipGetErrorString(hipGetLastError()));
  if (err) return set_error(VL_EINVAL, "malformed SNAP codec stream (code %u)", err);
  rc = vl_layer_upload_device(l, s.out, nrec);   // validates the boxes
  (void)hipStreamSynchronize(st);                 // before the scratch is freed
  return rc;
}

}  // namespace vl

using namespace vl;

extern "C" {

int vl_layer_load_klbin(vl_engine* e, vl_layer* l, const char* path) {
  if (!e || !l || !path) return set_error(VL_EINVAL, "null argument");
  File f(path, "rb");
  if (!f.f) return set_error(VL_EINVAL, "KLBIN file %s is empty or does not exist", path);
  fseek(f.f, 0, SEEK_END);
  const long bytes = ftell(f.f);
  fseek(f.f, 0, SEEK_SET);
  // LoadKLBin: an empty file is an error; bytes past the last whole box are
  // not read
  if (bytes <= 0) return set_error(VL_EINVAL, "KLBIN file %s is empty", path);
  std::vector<vl_box> b((size_t)bytes / 16);
  if (!b.empty() && !rd(f.f, b.data(), b.size())) return set_error(VL_EINVAL, "KLBIN file %s: short read", path);
  return vl_layer_upload(l, b.data(), (int64_t)b.size());   // invalid boxes are rejected
}

int vl_layer_save_klbin(vl_engine* e, const vl_layer* l, const char* path) {
  if (!e || !l || !path) return set_error(VL_EINVAL, "null argument");
  int64_t n = 0;
  int rc = vl_layer_size(l, &n);
  if (rc != VL_OK) return rc;
  std::vector<vl_box> b((size_t)std::max<int64_t>(n, 1));
  rc = vl_layer_download(l, b.data(), n, &n);
  if (rc != VL_OK) return rc;
  File f(path, "wb");
  if (!f.f) return set_error(VL_EINVAL, "KLBIN file %s cannot be written", path);
  if (n > 0 && !wr(f.f, b.data(), (size_t)n)) return set_error(VL_EINVAL, "KLBIN file %s: short write", path);
  return VL_OK;
}

int vl_snap_write_boxes(const vl_box* boxes, int64_t n, const char* path, int32_t method, int32_t layer_properties,
                        const int64_t* pnum) {
  if (!path || n < 0 || (n > 0 && !boxes)) return set_error(VL_EINVAL, "bad argument");
  if (method != METHOD_SNAPPY && method != METHOD_ZLIB && method != METHOD_ZSTD)
    return set_error(VL_EINVAL, "unknown compression method %d (1 Snappy, 2 zlib, 3 zstd)", method);
  if (method == METHOD_ZSTD && !zstd().ok) return set_error(VL_EINVAL, "zstd requested but libzstd.so.1 is missing");
  for (int64_t i = 0; i < n; i++)
    if (!(boxes[i].x1 < boxes[i].x2 && boxes[i].y1 < boxes[i].y2))
      return set_error(VL_EINVAL, "box %lld is invalid", (long long)i);
  const int64_t nch = (n + BOXES_PER_CHUNK - 1) / BOXES_PER_CHUNK;
  std::vector<std::vector<uint8_t>> stored((size_t)nch);
  std::vector<int64_t> rawsz((size_t)nch);
  std::atomic<bool> ok{true};
  parallel_for(nch, [&](int64_t c) {
    const int64_t s0 = c * BOXES_PER_CHUNK, s1 = std::min(n, s0 + BOXES_PER_CHUNK);
    std::vector<uint8_t> raw;
    encode_chunk(boxes + s0, s1 - s0, raw);
    rawsz[c] = (int64_t)raw.size();
    if (!compress_chunk(method, raw, stored[c])) ok = false;
  });
  if (!ok) return set_error(VL_EINVAL, "compression (method %d) failed", method);
  // SaveLayerToSnap: the R-tree bit is dropped; the polygon numbers follow
  // the header exactly when the PNUM bit is set, which here means when the
  // caller passes them
  const int32_t props =
      (layer_properties & ~(LAYER_STATE_RTREE_BIT | LAYER_STATE_PNUM_BIT)) | (pnum ? LAYER_STATE_PNUM_BIT : 0);
  File f(path, "wb");
  if (!f.f) return set_error(VL_EINVAL, "SNAP file %s cannot be written", path);
  bool w = wr(f.f, &CODEC_VERSION) && wr(f.f, &props) && wr(f.f, &n) && wr(f.f, &nch) && wr(f.f, &method);
  if (w && pnum && n > 0) w = wr(f.f, pnum, (size_t)n);
  std::vector<int64_t> pos((size_t)nch);
  for (int64_t c = 0; c < nch && w; c++) {
    pos[c] = (int64_t)ftell(f.f) + 1;   // 1-based stream positions, as Fortran's inquire(pos=) reports them
    const int64_t nb = std::min(n - c * BOXES_PER_CHUNK, BOXES_PER_CHUNK), cs = (int64_t)stored[c].size();
    w = wr(f.f, &nb) && wr(f.f, &rawsz[c]) && wr(f.f, &cs) && (cs == 0 || wr(f.f, stored[c].data(), (size_t)cs));
  }
  const int64_t table = (int64_t)ftell(f.f) + 1;
  w = w && (nch == 0 || wr(f.f, pos.data(), (size_t)nch)) && wr(f.f, &table);
  if (!w) return set_error(VL_EINVAL, "SNAP file %s: short write", path);
  return VL_OK;
}

int vl_layer_save_snap(vl_engine* e, const vl_layer* l, const char* path, int32_t method, int32_t layer_properties,
                       const int64_t* pnum) {
  if (!e || !l || !path) return set_error(VL_EINVAL, "null argument");
  int64_t n = 0;
  int rc = vl_layer_size(l, &n);
  if (rc != VL_OK) return rc;
  std::vector<vl_box> b((size_t)std::max<int64_t>(n, 1));
  rc = vl_layer_download(l, b.data(), n, &n);
  if (rc != VL_OK) return rc;
  return vl_snap_write_boxes(b.data(), n, path, method, layer_properties, pnum);
}

int vl_snap_info(const char* path, int32_t* layer_properties, int64_t* total_boxes, int64_t* chunks,
                 int32_t* method) {
  if (!path) return set_error(VL_EINVAL, "null argument");
  File f(path, "rb");
  if (!f.f) return set_error(VL_EINVAL, "SNAP file %s cannot be opened", path);
  int32_t version = 0, props = 0, m = 0;
  int64_t total = 0, nch = 0;
  fseek(f.f, 0, SEEK_END);
  const long fsize = ftell(f.f);
  fseek(f.f, 0, SEEK_SET);
  if (fsize > 0 && !(rd(f.f, &version) && rd(f.f, &props) && rd(f.f, &total) && rd(f.f, &nch) && rd(f.f, &m)))
    return set_error(VL_EINVAL, "SNAP file %s: short header", path);
  if (layer_properties) *layer_properties = props;
  if (total_boxes) *total_boxes = total;
  if (chunks) *chunks = nch;
  if (method) *method = m;
  return VL_OK;
}

int vl_layer_load_snap(vl_engine* e, vl_layer* l, const char* path, int32_t* layer_properties, int64_t* pnum,
                       int64_t pnum_cap) {
  if (!e || !l || !path) return set_error(VL_EINVAL, "null argument");
  File f(path, "rb");
  if (!f.f) return set_error(VL_EINVAL, "SNAP file %s cannot be opened", path);
  fseek(f.f, 0, SEEK_END);
  const long fsize = ftell(f.f);
  fseek(f.f, 0, SEEK_SET);
  int32_t version = 0, props = 0, method = 0;
  int64_t total = 0, nch = 0;
  if (fsize == 0) {   // an empty file is an empty layer (RestoreCompressedStreamFromDisk)
    if (layer_properties) *layer_properties = 0;
    return vl_layer_upload(l, nullptr, 0);
  }
  if (!(rd(f.f, &version) && rd(f.f, &props) && rd(f.f, &total) && rd(f.f, &nch) && rd(f.f, &method)))
    return set_error(VL_EINVAL, "SNAP file %s: short header", path);
  // every chunk has a 24-byte header: bounds the chunk table before it is read
  if (total < 0 || nch < 0 || nch > (int64_t)fsize / 24 || (total > 0) != (nch > 0))
    return set_error(VL_EINVAL, "SNAP file %s: bad header (%lld boxes, %lld chunks)", path, (long long)total,
                     (long long)nch);
  if (props & LAYER_STATE_PNUM_BIT) {   // polygon numbers
    if (pnum && pnum_cap < total)
      return set_error(VL_ENOSPC, "SNAP file %s: %lld polygon numbers, room for %lld", path, (long long)total,
                       (long long)pnum_cap);
    if (pnum) {
      if (total > 0 && !rd(f.f, pnum, (size_t)total)) return set_error(VL_EINVAL, "SNAP file %s: short", path);
    } else if (fseek(f.f, (long)(8 * total), SEEK_CUR) != 0) {
      return set_error(VL_EINVAL, "SNAP file %s: short", path);
    }
  }
  std::vector<int64_t> nb((size_t)nch), rawsz((size_t)nch), csz((size_t)nch);
  std::vector<std::vector<uint8_t>> stored((size_t)nch);
  int64_t sum = 0;
  for (int64_t c = 0; c < nch; c++) {
    // a chunk holds at most 65536 boxes (the reference decodes each into a
    // buffer of that many) and a record takes 1 to 40 bytes
    if (!(rd(f.f, &nb[c]) && rd(f.f, &rawsz[c]) && rd(f.f, &csz[c])) || nb[c] < 0 || nb[c] > BOXES_PER_CHUNK ||
        rawsz[c] < nb[c] || rawsz[c] > 40 * nb[c] || csz[c] < 0 || csz[c] > (int64_t)fsize)
      return set_error(VL_EINVAL, "SNAP file %s: bad chunk %lld header", path, (long long)c);
    stored[c].resize((size_t)csz[c]);
    if (csz[c] && !rd(f.f, stored[c].data(), (size_t)csz[c]))
      return set_error(VL_EINVAL, "SNAP file %s: chunk %lld short", path, (long long)c);
    sum += nb[c];
  }
  if (sum != total) return set_error(VL_EINVAL, "SNAP file %s: chunks hold %lld boxes, header %lld", path,
                                     (long long)sum, (long long)total);
  if (method == METHOD_ZSTD && !zstd().ok) return set_error(VL_EINVAL, "SNAP file %s: zstd needs libzstd.so.1", path);
  std::vector<int64_t> roff((size_t)nch + 1, 0);
  for (int64_t c = 0; c < nch; c++) roff[c + 1] = roff[c] + rawsz[c];
  std::vector<uint8_t> raw((size_t)std::max<int64_t>(roff[nch], 1));
  std::atomic<bool> ok{true};
  parallel_for(nch, [&](int64_t c) {
    if (!uncompress_chunk(method, stored[c].data(), stored[c].size(), raw.data() + roff[c], (size_t)rawsz[c]))
      ok = false;
  });
  if (!ok) return set_error(VL_EINVAL, "SNAP file %s: decompression (method %d) failed", path, method);
  if (layer_properties) *layer_properties = props & LAYER_STATE_EVERYTHING;
  return snap_decode_to_layer(e, l, raw.data(), rawsz, nb);
}

// RestoreSnapToLayer's dispatch (src/diskdesign_impl.f90:83-114): a name
// holding ".bin" is KLBIN, anything else SNAP.
int vl_layer_load_file(vl_engine* e, vl_layer* l, const char* path, int32_t* layer_properties) {
  if (!path) return set_error(VL_EINVAL, "null argument");
  if (strstr(path, ".bin")) {
    if (layer_properties) *layer_properties = 0;
    return vl_layer_load_klbin(e, l, path);
  }
  return vl_layer_load_snap(e, l, path, layer_properties, nullptr, 0);
}

int vl_codec_decode(vl_engine* e, const uint8_t* raw, const int64_t* chunk_bytes, const int64_t* chunk_boxes,
                    int64_t nchunks, vl_layer* out) {
  if (!e || !out || nchunks < 0 || (nchunks > 0 && (!raw || !chunk_bytes || !chunk_boxes)))
    return set_error(VL_EINVAL, "bad argument");
  std::vector<int64_t> cb(chunk_bytes, chunk_bytes + nchunks), nb(chunk_boxes, chunk_boxes + nchunks);
  return snap_decode_to_layer(e, out, raw, cb, nb);
}

int vl_codec_encode(const vl_box* boxes, int64_t n, uint8_t* out, int64_t cap, int64_t* out_n) {
  if (n < 0 || (n > 0 && !boxes) || !out_n || n > BOXES_PER_CHUNK) return set_error(VL_EINVAL, "bad argument");
  std::vector<uint8_t> raw;
  encode_chunk(boxes, n, raw);
  *out_n = (int64_t)raw.size();
  if ((int64_t)raw.size() > cap) return set_error(VL_ENOSPC, "%zu bytes needed", raw.size());
  if (!raw.empty()) memcpy(out, raw.data(), raw.size());
  return VL_OK;
}

}  // extern "C"
