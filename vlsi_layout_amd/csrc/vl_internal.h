// vl_internal.h — shared declarations of the gfx950 engine (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "vl_api.h"

namespace vl {

// ---------------------------------------------------------------------------
// Tiles.  A tile is a rectangle of the op's domain plus the boxes (of one or
// two layers) whose interior meets it.  Entries of tile t live at
// entries[off .. off+nA) (layer A) and [off+nA .. off+nA+nB) (layer B).
// Seam flags say which tile borders are internal to the domain: a result run
// touching such a border is a piece of a longer run and goes to the side
// stream for seam resolution instead of the final output.
enum : uint32_t {
  SEAM_B = 1u,   // bottom border y0 is internal
  SEAM_T = 2u,   // top border y1 is internal
  SEAM_L = 4u,   // left border x0 is internal
  SEAM_R = 8u,   // right border x1 is internal
};

struct TileDesc {
  int32_t x0, y0, x1, y1;
  int64_t off;
  int32_t nA, nB;
  uint32_t flags;   // SEAM_*
  int32_t op;       // per-tile operator (batched launches mix operators)
};
static_assert(sizeof(TileDesc) == 40, "TileDesc layout");

// Operators evaluated per cell by the tile kernel.
enum TileOp : int32_t {
  TOP_UNION = 0,   // A only (merge / grow)
  TOP_AND = 1,
  TOP_NOT = 2,     // A \ B
  TOP_OR = 3,
  TOP_XOR = 4,
  TOP_OVERLAP = 5, // cells covered by at least two boxes of A (single-layer AND)
};

// Uniform grid over a domain.  Tile (c, r) covers
// [x0 + c*tw, min(x0+(c+1)*tw, x1)) x [y0 + r*th, min(...)).  Tiles are
// numbered row by row (r * cols + c), or, with zb > 0, in Morton (Z) order
// over a 2^zb x 2^zb square of tiles (the bits of c and r interleaved, c in
// the even bits): then 4^zb numbers, those outside cols x rows unused.
struct Grid {
  int64_t x0, y0, x1, y1;   // domain (int64 so widths never overflow)
  int64_t tw, th;
  int32_t cols, rows;
  int32_t zb = 0;
};

// Tile numbers of grid g (the size of its per-tile arrays).
inline int64_t grid_tiles(const Grid& g) { return g.zb ? (int64_t)1 << (2 * g.zb) : (int64_t)g.cols * g.rows; }

// Morton number of tile (c, r): bit k of c at bit 2k, bit k of r at bit 2k+1.
__host__ __device__ inline uint32_t morton_spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__host__ __device__ inline uint32_t tile_number(int r, int c, const Grid& g) {
  return g.zb ? (morton_spread16((uint32_t)c) | (morton_spread16((uint32_t)r) << 1))
              : (uint32_t)r * (uint32_t)g.cols + (uint32_t)c;
}

// Per-box transform applied while binning (GROW): x1-=w, y1-=s, x2+=e, y2+=n.
struct Xform {
  int32_t e, n, w, s;
};

// Status bits written by kernels into the engine's device error word.
enum : uint32_t {
  ERR_INVALID_BOX = 1u,
  ERR_LOOKBACK = 2u,
  ERR_OVERFLOW_CAP = 4u,
};

// ---------------------------------------------------------------------------
// Engine access for entry points outside vl_engine.hip.

// Sets the calling thread's last error text (vl_last_error) and returns code.
int set_error(int code, const char* fmt, ...);

// Holds an engine's lock for its scope; begin() takes it and makes the
// engine's device current.
class EngineLock {
 public:
  explicit EngineLock(vl_engine* e) : e_(e) {}
  ~EngineLock();
  int begin();
  hipStream_t stream() const;
  int device() const;

 private:
  vl_engine* e_;
  bool locked_ = false;
};

// Kernel statistics of stages outside vl_engine.hip (vl_engine_kernel_stats
// "snap"): marks are events on the engine stream (null when the engine does
// not collect statistics); the caller holds the engine lock.
enum { KSTAT_SNAP = 4 };
hipEvent_t engine_prof_mark(vl_engine* e);
void engine_prof_add(vl_engine* e, hipEvent_t a, hipEvent_t b, int kind, int64_t launches, double bytes);

// vl_snap.hip: GPU decode of SNAP codec chunks (concatenated raw codec bytes;
// coff: nchunks+1 byte offsets, roff: nchunks+1 record offsets, boff:
// nchunks+1 walk-block offsets).  Scratch: len (total_raw), exit_e and cnt
// (nblocks*41), entry and rbase (nblocks), start/rx/ry/widx/hidx/dictw/dicth
// (records).  *err (zeroed by the caller) != 0: malformed stream.
struct SnapDecode {
  const uint8_t* raw;
  int64_t total_raw;
  const int64_t* coff;
  const int64_t* roff;
  const int64_t* boff;
  int nchunks;
  int64_t nblocks;
  uint8_t* len;
  uint8_t* exit_e;
  uint16_t* cnt;
  uint8_t* entry;
  int64_t* rbase;
  int64_t* start;
  int32_t *rx, *ry, *dictw, *dicth;
  uint32_t *widx, *hidx;
  vl_box* out;
  uint32_t* err;
};
int64_t snap_blocks(int64_t raw_bytes);   // walk blocks of one chunk
void launch_snap_decode(const SnapDecode& d, hipStream_t st);

// vl_noalloc.hip: the static scratch's device address on `device` (current).
int noalloc_setup(int device);

// ---------------------------------------------------------------------------
// Host-side launch wrappers (each .hip file implements its own kernels).

// vl_prims.hip
// Graph replays: *flag |= 1 when the bytes (a multiple of 4) at dev differ
// from the host values `expected` (copied into the launch arguments).
struct Expect8 {
  uint32_t w[8];
};
void launch_expect(const void* dev, const void* expected, size_t bytes, uint32_t* flag, hipStream_t st);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t st);
// GRID: rows x cols boxes of the domain (CreateGrid's arithmetic).
void launch_grid(int64_t x_lo, int64_t y_lo, int64_t x_hi, int64_t y_hi, int rows, int cols, int64_t overlap,
                 vl_box* out, hipStream_t st);
void launch_fill_u64(uint64_t* p, uint64_t v, size_t n, hipStream_t st);
// MBR of boxes after a transform; validates boxes (ERR_INVALID_BOX).
// mbr[4] must be pre-set to (INT_MAX, INT_MAX, INT_MIN, INT_MIN).  sums (may
// be null, zeroed by the caller): += summed widths, heights of the boxes.
void launch_mbr(const vl_box* b, int64_t n, Xform xf, int32_t* mbr, uint32_t* err, hipStream_t st,
                unsigned long long* sums = nullptr);
// Exclusive scans.  tmp must hold scan_tmp_bytes(n) bytes.  total may be null.
size_t scan_tmp_bytes(size_t n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total, void* tmp,
                        hipStream_t st);
void scan_exclusive_u64(const uint64_t* in, uint64_t* out, size_t n, uint64_t* total, void* tmp,
                        hipStream_t st);
// Exclusive scan of (cntA[i] + cntB[i]) into off (u64); total into *total.
void scan_pair_counts(const uint32_t* cntA, const uint32_t* cntB, uint64_t* off, size_t n,
                      uint64_t* total, void* tmp, hipStream_t st);

// vl_sort.hip — LSD radix sort of (u64 key, u32 value) pairs.
struct RadixPlan {
  int64_t n;
  int nblocks;
  size_t tmp_bytes;
};
// digit_bits: 8 or 11 (any other value means 8); the plan's scratch size
// depends on it, so a sort takes the digit width its plan was made with.
RadixPlan radix_plan(int64_t n, int digit_bits = 8);
// Sorts keys/vals in place (uses alt buffers of the same size + tmp).  The
// host reads the digit histograms once to skip passes in which all keys share
// a digit (one synchronisation).
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      int64_t n, void* tmp, hipStream_t st, int digit_bits = 8);
// Same sort for keys below 2^key_bits: ceil(key_bits/digit_bits) passes, no
// histogram pre-pass and no host synchronisation.  Returns true when the
// sorted pairs are in keys_alt/vals_alt (odd number of passes), false when in
// keys/vals.
bool radix_sort_pairs_bits(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n,
                           int key_bits, void* tmp, hipStream_t st, int digit_bits = 8);
// The same for 32-bit keys (key_bits <= 32).
bool radix_sort_pairs_bits32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, int64_t n,
                             int key_bits, void* tmp, hipStream_t st, int digit_bits = 8);
void launch_tracker_keys(const vl_tracker* t, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st);
void launch_tracker_gather(const vl_tracker* in, const uint32_t* perm, vl_tracker* out, int64_t n,
                           hipStream_t st);
void launch_morton_keys(const vl_box* b, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st);
void launch_corner_keys(const vl_box* b, uint64_t* keys, uint32_t* vals, int64_t n, hipStream_t st);
void launch_box_gather(const vl_box* in, const uint32_t* perm, vl_box* out, int64_t n, hipStream_t st);
void launch_perm_to_index1(const uint32_t* perm, int64_t* idx, int64_t n, hipStream_t st);

// vl_tile.hip
// Binning by a stable sort of (tile, layer) keys.  Boxes of layer A are
// indices 0..na-1, boxes of layer B na..na+nb-1; a layer-A box is binned by its
// extent grown by xra (the tile range), and stored grown by xsa.
void launch_bin_span(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g, uint32_t* cnt,
                     hipStream_t st);
void launch_bin_emit(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g,
                     const uint64_t* pos, uint32_t* keys, uint32_t* vals, hipStream_t st);
// start[k] = number of keys < k, for k = 0..nkeys.
void launch_bin_bounds(const uint32_t* keys, uint64_t total, int64_t nkeys, uint64_t* start, hipStream_t st);
void launch_bin_gather(const vl_box* a, int64_t na, Xform xsa, const vl_box* b, const uint32_t* vals,
                       uint64_t total, vl_box* entries, hipStream_t st);
// Atomic binning: cnt[2*tile + layer] (zeroed by the caller) += the boxes
// meeting each tile; after an exclusive scan of cnt into start, the scatter
// counts cnt back down to zero while placing every (box, tile) pair in its
// key's range (entries of a key in no particular order).  grouped: one
// atomic per distinct key of a wave instead of one per run of equal keys.
void launch_bin_count(const vl_box* a, int64_t na, Xform xra, const vl_box* b, int64_t nb, Grid g, uint32_t* cnt,
                      bool grouped, hipStream_t st);
void launch_bin_scatter(const vl_box* a, int64_t na, Xform xra, Xform xsa, const vl_box* b, int64_t nb, Grid g,
                        const uint64_t* start, uint32_t* cnt, vl_box* entries, bool grouped, hipStream_t st);
// AND tiles: appends to list (count in *list_n, zeroed by the caller) the
// tiles whose result may be non-empty (a clipped A box overlaps a B box, or
// the tile is too large to test); the others get tab_n[t] = tab_n[ntab+t] = 0.
void launch_and_filter(const TileDesc* desc, int ntiles, const vl_box* entries, uint32_t* list, uint32_t* list_n,
                       uint32_t* tab_n, int ntab, hipStream_t st);
// AND emptiness by semi-join: entries/start hold one input binned alone into g
// (layer-A keys); *hit (cleared by the caller) is set to 1 when some box of
// p[0..n), transformed by xp, overlaps a binned box with positive area; mark
// (may be null): per box of p, 1 when it overlaps one, else 0.
void launch_and_probe(const vl_box* p, int64_t n, Xform xp, Grid g, const vl_box* entries, const uint64_t* start,
                      uint32_t* hit, uint32_t* mark, hipStream_t st);
// out[pos[i]] = p[i] transformed by xp for every i with mark[i] != 0.
void launch_compact_marked(const vl_box* p, int64_t n, Xform xp, const uint32_t* mark, const uint32_t* pos,
                           vl_box* out, hipStream_t st);
// The same scatter writing box indices (idx) instead of boxes (one layer).
void launch_bin_scatter_idx(const vl_box* a, int64_t na, Xform xra, Grid g, const uint64_t* start, uint32_t* cnt,
                            uint32_t* idx, bool grouped, hipStream_t st);
// Tile descriptors from the key bounds (start has 2*ntiles+1 entries).
void launch_build_desc(Grid g, const uint64_t* start, int32_t op, TileDesc* desc, hipStream_t st);
// Tile kernel: result strips of every tile.  Strips touching an internal seam
// go to side[] with a tag byte; the rest go to out[].  Tiles with more boxes
// than the kernel's capacity are listed in ovf[] (count in *ovf_n).
// Workgroup g of a launch takes tiles g, g+G, g+2G, ... (no shared ticket
// counter).  A tile's strips are one range of the scratch stream, taken from
// its workgroup's current chunk; a workgroup whose chunk cannot hold the next
// tile takes a new chunk of max(chunk, strips) from the stream's bump counter
// *next (one atomic per chunk, not per tile), so uneven tiles cost at most
// one partly used chunk per workgroup.  Each tile's range is recorded in
// per-tile tables; each strip carries a tag byte (0 = final, SEAM_* bits =
// seam piece).  A reorder pass then copies the strips in tile order into the
// final and side streams, so results do not depend on scheduling.  Ranges
// past stream_cap are not written and set ERR_OVERFLOW_CAP; *next then holds
// the size to retry with.
struct TileOut {
  vl_box* strips;              // scratch stream
  uint8_t* tag;                // tag per strip
  uint64_t stream_cap;         // strips the scratch stream holds
  uint64_t chunk;              // strips per chunk
  unsigned long long* next;    // bump counter of the stream (zeroed by the host)
  int32_t erode;               // erosion kernels: erosion distance (tile windows grown by it)
  unsigned long long* cnt;     // [0] strips written, [1] seam strips (one add per workgroup)
  uint64_t* tab_off;           // per tile: range start
  uint32_t* tab_n;             // per tile: [t] final count; tab_n[ntab + t] side count
  int ntab;                    // table stride (number of descriptors)
  const uint32_t* list;        // optional: tile = list[i] (null = i)
  const uint32_t* list_n;      // optional: number of tiles, read on the device (null = ntiles)
  uint32_t* ovf;               // tiles above the kernel's capacity
  uint32_t* ovf_n;
  long long* area;             // area mode: accumulated union area (int64)
  uint32_t* err;
};
// Wave-per-tile kernel (capacity nmax = 31, 63 or 127 boxes; erosion: 15 or
// 31), launched with exactly wg workgroups.
// With o.list set it takes the tiles of that list.  Erosion mode
// computes the erosion of layer A by the square of half-side o.erode inside
// each tile from the boxes meeting the tile grown by o.erode.
void launch_tile_wave(int nmax, bool two_layers, bool area_mode, bool erode, const TileDesc* desc, int ntiles,
                      const vl_box* entries, TileOut o, int wg, hipStream_t st);
int tile_wave_max_wg(int nmax, bool two_layers, bool area_mode, bool erode);
// Workgroup-per-tile kernel (capacity 255 boxes; erosion: 127) for tiles above
// the wave kernel's capacity.
void launch_tile_block(bool two_layers, bool area_mode, bool erode, const TileDesc* desc, int ntiles,
                       const vl_box* entries, TileOut o, int wg, hipStream_t st);
int tile_block_max_wg(bool two_layers, bool area_mode, bool erode);
constexpr int TILE_BLOCK_NMAX = 255;
// Fused tile kernel (vl_fused.hip): several operators over ONE source layer S
// evaluated per tile from shared masks.  The base masks are S itself and up to
// two morphological images of S (GROW by (e, n, w, s) or erosion by the
// square of half-side d); every output is a Boolean function of the three,
// given as a truth table: bit (s | m0 << 1 | m1 << 2) is the output value for
// those base values.  Tile windows are grown by `halo` (at least every
// morph's reach), so each output is exact inside the tile.
constexpr int FUSE_MAX_OUT = 8;
// Images of a fused group besides its source S: a GROW of S, an erosion of
// S, or a second input layer (its boxes are the tiles' layer-B entries; at
// most one image is a layer).
enum : int32_t { MORPH_GROW = 1, MORPH_ERODE = 2, MORPH_LAYER = 3 };
struct FusedSpec {
  int32_t nmorph;                 // 0..2
  int32_t kind[2];                // MORPH_GROW / MORPH_ERODE / MORPH_LAYER
  int32_t arg[2][4];              // GROW: e, n, w, s (all >= 0); ERODE: d > 0 in arg[k][0]
  int32_t halo;
  int32_t nout;                   // 1..FUSE_MAX_OUT
  uint32_t tt[FUSE_MAX_OUT];      // truth table per output (bit 0 must be 0)
};
// Per-output scratch streams and tables (same meaning as TileOut's, one set
// per output; a workgroup takes chunks of every output's stream).
struct FusedOut {
  vl_box* strips[FUSE_MAX_OUT];
  uint8_t* tag[FUSE_MAX_OUT];
  uint64_t* tab_off[FUSE_MAX_OUT];
  uint32_t* tab_n[FUSE_MAX_OUT];           // [t] final, [ntab + t] seam count
  unsigned long long* cnt[FUSE_MAX_OUT];   // [0] strips, [1] seam strips, [2] the stream's bump counter
  uint64_t stream_cap[FUSE_MAX_OUT];
  uint64_t chunk;
  int ntab;
  const uint32_t* list;                    // optional tile list (as TileOut)
  const uint32_t* list_n;
  uint32_t* ovf;
  uint32_t* ovf_n;
  uint32_t* err;
};
// Capacities (boxes per tile window) of the fused kernel's tiers (64-, 128-
// and 256-slot grids of 6 coordinates per box and axis).
constexpr int FUSE_NMAX_TINY = 10;
constexpr int FUSE_NMAX_SMALL = 20;
constexpr int FUSE_NMAX_LARGE = 42;
void launch_tile_fused(int nmax, const TileDesc* desc, int ntiles, const vl_box* entries, const FusedSpec& fs,
                       const FusedOut& o, int wg, hipStream_t st);
int tile_fused_max_wg(int nmax);

// Copies the per-tile ranges into tile order: final strips of tile t go to
// out[out_base + dstF[t] ...], seam strips to side[side_base + dstS[t] ...].
void launch_tile_reorder(int ntiles, const uint64_t* tab_off, const uint32_t* tab_n, int ntab,
                         const uint64_t* dstF, const uint64_t* dstS, const vl_box* scr,
                         const uint8_t* scrTag, vl_box* out, uint64_t out_base, vl_box* side,
                         uint8_t* side_tag, uint64_t side_base, hipStream_t st);
void launch_gather_desc(const TileDesc* desc, const uint32_t* idx, int n, TileDesc* out, hipStream_t st);
// Exclusive scan of u32 values into u64 offsets.
void scan_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, uint64_t* total, void* tmp,
                     hipStream_t st);
// Split overflow tiles: per parent one workgroup re-bins its entries into a
// kx x ky grid of sub-tiles (row-major sub-descriptors precomputed on the
// host, which splits in x only (ky = 1); all but the last column / row as
// wide / tall as the first sub-tile).
// halo > 0: windows grown by halo (a box belongs to every sub-tile its extent
// grown by halo meets).
void launch_split_count(const TileDesc* parents, int np, const int32_t* sub_first,
                        const TileDesc* subs, uint32_t* sub_cntA, uint32_t* sub_cntB,
                        const vl_box* entries, int32_t halo, hipStream_t st);
void launch_split_scatter(const TileDesc* parents, int np, const int32_t* sub_first, TileDesc* subs,
                          int ns, const uint64_t* off, const uint32_t* cntA, const uint32_t* cntB,
                          uint32_t* fillA, uint32_t* fillB, const vl_box* entries,
                          vl_box* sub_entries, int32_t halo, hipStream_t st);

// vl_dense.hip — tiles above every capacity-bounded kernel (one workgroup per
// listed tile, global-memory grids).  Per tile i of the list: coordinate
// scratch of dense_coord_slots(n_i) int32 at cs + coff[i]; k_dense_coords
// writes the distinct x / y breakpoint counts to nxy[2i], nxy[2i+1].  Then,
// with ncol = nx - 1, nrows = ny - 1, W = (nrows + 64) / 64: grid words
// (two-layer operators: 2 *) ncol * W at grid + goff[i], ncol column
// counters at colcnt + cofs[i], and at most ncol * (nrows / 2 + 1) strips at
// strips + soff[i]; tab_off / tab_n as TileOut's, indexed by list position
// (stride nlist).  Area mode (area non-null) adds the tile areas to *area.
int64_t dense_coord_slots(int64_t nboxes);
void launch_dense_coords(const TileDesc* desc, const uint32_t* list, int nlist, const vl_box* entries,
                         const int64_t* coff, int32_t* cs, int32_t* nxy, hipStream_t st);
void launch_dense_tile(const TileDesc* desc, const uint32_t* list, int nlist, const vl_box* entries,
                       const int64_t* coff, const int32_t* cs, const int32_t* nxy, uint64_t* grid,
                       const uint64_t* goff, uint32_t* colcnt, const uint64_t* cofs, vl_box* strips, uint8_t* tag,
                       const uint64_t* soff, uint64_t* tab_off, uint32_t* tab_n, bool two,
                       unsigned long long* cnt, long long* area, hipStream_t st);

// vl_pnum.hip — pairs of touching boxes per tile (entries = box indices);
// mode 0: union-find of the pairs (parent, touched) and the sums
// [0] positive intersection areas, [1] zero-area intersection perimeters;
// mode 1: sums[2] = pairs overlapping with positive area.
void launch_uf_init(uint32_t* parent, uint32_t* touched, int64_t n, hipStream_t st);
void launch_pairs(const TileDesc* desc, int ntiles, const uint32_t* idx, const vl_box* boxes, Grid g, int mode,
                  uint32_t* parent, uint32_t* touched, unsigned long long* sums, hipStream_t st);
void launch_uf_roots(uint32_t* parent, const uint32_t* touched, uint32_t* root, uint32_t* is_root, int64_t n,
                     hipStream_t st);
void launch_uf_labels(const uint32_t* root, const uint32_t* touched, const uint32_t* rank, int64_t* labels, int64_t n,
                      hipStream_t st);

// vl_window.hip — window extraction from a cell-indexed layer.  Candidates
// are the concatenation of the entry ranges [rbeg[r], rbeg[r] + roff[r+1] -
// roff[r]); each is clipped to w (clip[i]) and flagged when it keeps area.
// A box binned into several cells is kept only in the cell holding the
// lower-left corner of its part inside w (start: the index's cell starts,
// layer-A keys of grid g), so every box of the window is returned once.
void launch_window_clip(const vl_box* entries, const uint64_t* rbeg, const uint64_t* roff, int nr, uint64_t n,
                        vl_box w, const uint64_t* start, Grid g, vl_box* clip, uint32_t* keep, hipStream_t st);
// out[pos[i]] = clip[i] for the flagged candidates.
void launch_window_compact(const vl_box* clip, const uint32_t* keep, const uint32_t* pos, uint64_t n, vl_box* out,
                           hipStream_t st);
// *out += order-independent 64-bit hash of the box set b[0..n).
void launch_box_hash(const vl_box* b, int64_t n, unsigned long long* out, hipStream_t st);

// vl_stitch.hip
// Partition side strips: tags with SEAM_B|SEAM_T -> next (border), the rest -> stitch.
void launch_partition_side(const vl_box* side, const uint8_t* tag, int64_t n, vl_box* next,
                           vl_box* stitch, uint8_t* stitch_tag, uint64_t* counts, void* tmp,
                           hipStream_t st);
size_t partition_tmp_bytes(int64_t n);
// X-stitch of strips whose left/right ends sit on internal x-seams.
// Links every strip ending on a right seam to the strip starting on a left
// seam at its lower-right corner with the same rows (next[i], has_prev[j]),
// through a hash table of the starts (tk: stitch_table_slots(n) keys, tv as
// many indices) keyed relative to the domain origin (x0, y0; y in the low yb
// bits).
uint64_t stitch_table_slots(int64_t n);
void launch_stitch_link(const vl_box* s, const uint8_t* tag, int64_t n, int64_t x0, int64_t y0, int yb,
                        unsigned long long* tk, uint32_t* tv, int32_t* next, uint8_t* has_prev, hipStream_t st);
void launch_stitch_emit(const vl_box* s, int64_t n, const int32_t* next, const uint8_t* has_prev,
                        vl_box* out, uint64_t out_base, void* tmp, uint64_t* count, hipStream_t st);

}  // namespace vl
