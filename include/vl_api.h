/*
 * vl_api.h — C ABI of the MI355X (gfx950) rectangle-Boolean engine.
 *
 * This is the drop-in boundary for the MAGPARSER `exec run` operator path of
 * skoranne/VLSI_LAYOUT.  Every function uses plain pointers, sizes and
 * status codes; nothing here depends on torch or on any C++ type.  Fortran
 * binds to it with iso_c_binding (vlsi_layout_amd/fortran/vl_layer_ops.f90),
 * Python with ctypes (vlsi_layout_amd/__init__.py).
 *
 * Semantics (point-set, as in the reference): a layer is the union of its
 * boxes; a box is the half-open integer rectangle [x1,x2) x [y1,y2) and is
 * valid iff x1<x2 and y1<y2 (src/geometry.f90:21-22,66-70).  Every operator
 * returns its result as the unique maximal-strip decomposition the reference's
 * heal_boxes produces (src/polygon_fracture.f90:805-1004): at every x the
 * covered set is split into maximal y-intervals, and each interval extends in
 * x as long as it stays identical.  Output order is deterministic (tile order,
 * then a fixed order inside a tile); vl_canonicalize additionally sorts by
 * (x1,y1,x2,y2).  All arithmetic is exact integer (int32 coordinates, int64
 * areas).
 *
 * Threading: an engine owns one HIP stream; calls on one engine are
 * serialised by an internal mutex.  Use one engine per host thread for
 * concurrency.
 */
#ifndef VL_API_H
#define VL_API_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference has none: it error-stops) -------------- */
#define VL_OK         0
#define VL_EINVAL     1  /* invalid argument, or an invalid box (x1>=x2 or y1>=y2):
                            the reference error-stops in LoadKLBin (src/kldata.f90:55-57)
                            and box_grow_directional (src/geometry.f90:187,192) */
#define VL_ENOSPC     2  /* caller capacity too small; the required count is returned */
#define VL_ENOMEM     3  /* device allocation failed */
#define VL_EHIP       4  /* HIP runtime error */
#define VL_EASSERT    5  /* ASSERT_ZERO found shapes and abort was requested */
#define VL_EINTERNAL  7  /* internal consistency check failed */

/* Box: identical layout to the reference's Box (src/geometry.f90:21-22) and
 * c_box of the Boost bridge (src/boost_polygon_api.h). */
typedef struct vl_box { int32_t x1, y1, x2, y2; } vl_box;

/* XYTracker: the 12-byte event record of src/commonmodule.f90:30-34 as laid
 * out by src/event_sorter.cu:15-19. */
typedef struct vl_tracker { int32_t X, Y; int8_t polygonNumber; } vl_tracker;

typedef struct vl_engine vl_engine;   /* device, stream, scratch, tuning */
typedef struct vl_layer  vl_layer;    /* a device-resident box layer       */

/* Boost control codes (src/commonmodule.f90:12-14). */
#define VL_OP_XOR   0
#define VL_OP_OR    1
#define VL_OP_AND   2
#define VL_OP_NOT   3
#define VL_OP_MERGE 4
#define VL_OP_SIZE  5
#define VL_OP_GROW  6   /* not a Boost code: CalculateGROWLayer */

/* ---- engine -------------------------------------------------------------- */
/* Thread-local text of the last error on this thread. */
const char* vl_last_error(void);
int vl_engine_create(int device, vl_engine** out);
int vl_engine_destroy(vl_engine* e);
/* Tuning knobs (results never depend on them):
 *   "tile_nmax"    wave-kernel tile capacity, 63 (default) or 127 boxes
 *   "tile_target"  mean boxes per tile the operator grid aims for (20)
 *   "erode_target" the same for SIZE(-d) erosion grids (12)
 *   "level_factor" row coarsening per seam level (1024)
 *   "erode_kernel" 1: SIZE(-d) by the per-tile erosion kernel (default);
 *                  0: by complement, grow and difference
 *   "max_wg"       cap on the tile-kernel grid (0 = occupancy limit)
 *   "fuse"         programs: 1 (default) evaluates groups of operators over
 *                  one source layer in one fused tile pass (up to two
 *                  morphological images per group); 2 the same with at most
 *                  one image per group; 0 runs every operator on its own
 *   "fuse_target"  mean boxes per tile of a fused group's grid (6)
 *   "bin_mode"     2 (default): tiles binned by per-tile counters and an
 *                  atomic scatter, one atomic per distinct tile of a wave;
 *                  1: the same with one atomic per run of equal tiles;
 *                  0: by a stable radix sort of tile keys
 *   "radix_bits"   digit width of the radix sorts (x-stitch, sort binning,
 *                  box and event sorts): 8 (default) or 11
 *   "workers"      programs: chains run concurrently on up to this many
 *                  worker streams (1 to 16, default 4)
 *   "fuse2"        programs: 1 lets a fused group take a second layer as an
 *                  image (default 0; set by the autotuner)
 *   "and_check"    programs: 1 (default) tests an AND read only by
 *                  ASSERT_ZERO for emptiness instead of computing it
 *   "check_target" mean binned boxes per tile of that test's grid (6)
 *   "tile_shape"   tiles shaped like the boxes, up to this aspect (32;
 *                  0: square tiles)
 *   "profile"      1: collect per-kernel statistics (vl_engine_kernel_stats)
 * Read-only: "fused_groups" = groups run fused by programs on this engine. */
int vl_engine_set_param(vl_engine* e, const char* key, int64_t value);
int vl_engine_get_param(vl_engine* e, const char* key, int64_t* value);
/* Block until all work queued on the engine's stream has finished. */
int vl_engine_sync(vl_engine* e);
/* Autotuner (no counterpart in the reference): sets "tile_target",
 * "erode_target" and "fuse_target" to the fastest of a few candidates, timed
 * on A NOT B (or the merge of A when b is NULL) plus GROW(A; 1), on
 * SIZE(A; -1), and on the fused ring of A (GROW(A; 1), SIZE(A; -1) and their
 * two differences with A), then "fuse" (1 or 0: that ring as one fused group
 * against the same four operators one at a time); then "bin_mode" (2 or 1),
 * "radix_bits" (8 or 11) and "max_wg" (0, 4096 or 2048), each timed on A NOT
 * B, A AND B and GROW(A; 1).  A value changes only when a candidate is more than 3 % faster.  The
 * first two are returned when the pointers are non-null (the others through
 * vl_engine_get_param).  Results of every operator are independent of these
 * parameters. */
int vl_engine_autotune(vl_engine* e, const vl_layer* a, const vl_layer* b, int64_t* target,
                       int64_t* erode_target);

/* ---- layers (the Layer type of src/design.f90:65-82, device-resident) ---- */
int vl_layer_create(vl_engine* e, vl_layer** out);
int vl_layer_destroy(vl_layer* l);
/* Copy n boxes from host memory; every box must be valid (VL_EINVAL). */
int vl_layer_upload(vl_layer* l, const vl_box* host, int64_t n);
/* Copy n boxes already in device memory (same device). */
int vl_layer_upload_device(vl_layer* l, const vl_box* dev, int64_t n);
/* Copy the layer to host memory; VL_ENOSPC with *n set when cap is too small. */
int vl_layer_download(const vl_layer* l, vl_box* host, int64_t cap, int64_t* n);
int vl_layer_size(const vl_layer* l, int64_t* n);
/* Device address of the layer's boxes (valid until the layer is next written). */
int vl_layer_device_ptr(const vl_layer* l, const vl_box** dev);

/* ---- operators: the `exec run` surface (src/controlparser.f90:409-585) ---- */
/* C = A AND B.  Replaces CalculateAND '*' (src/design.f90:1131-1235) and AND2. */
int vl_and(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c);
/* C = A NOT B (A \ B).  Replaces CalculateFrameNOT '~'
 * (src/design_impl.f90:1152-1164), CalculateNOT '@' (src/design.f90:1238-1268)
 * and the Boost NOT '!' (src/boost_polygon_api.cpp:129-131). */
int vl_not(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c);
/* C = A OR B.  Replaces Boost OR '+' and CalculateOR 'OR2' (src/design.f90:1085-1103). */
int vl_or(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c);
/* C = A XOR B.  Replaces Boost XOR '^' and CalculateXOR/CalculateExplicitXOR
 * (src/design_impl.f90:961-991,1166-1177); covers all of B, including the part
 * outside MBR(A) that the reference's tiled XOR drops (src/design_impl.f90:1030). */
int vl_xor(vl_engine* e, const vl_layer* a, const vl_layer* b, vl_layer* c);
/* C = GROW(A; E,N,W,S): x1-=W, y1-=S, x2+=E, y2+=N per box, then the union.
 * Replaces CalculateGROWLayer (src/design_impl.f90:571-586, src/design.f90:242-246).
 * VL_EINVAL when a grown box is invalid (src/geometry.f90:192). */
int vl_grow(vl_engine* e, const vl_layer* a, int32_t east, int32_t north, int32_t west,
            int32_t south, vl_layer* c);
/* C = SIZE(A; d): d>=0 dilation by the square [-d,d]^2, d<0 erosion by it.
 * Replaces CalculateBoostOperation(K_BOOST_CONTROL_SIZE) (src/design_impl.f90:931-951). */
int vl_size(vl_engine* e, const vl_layer* a, int32_t d, vl_layer* c);
/* C = union of A as maximal strips.  Replaces MergeHealLayer
 * (src/design_impl.f90:1318-1341), heal_boxes (src/polygon_fracture.f90:805) and
 * the Boost MERGE (src/boost_polygon_api.cpp:41-70). */
int vl_merge(vl_engine* e, const vl_layer* a, vl_layer* c);
/* C = the points covered by at least two boxes of A (the union of the
 * positive-area intersections of every pair of A's boxes).  Replaces
 * CalculateSingleLayerAND, the `AND` verbose operator with one input
 * (src/design_impl.f90:385-509, src/controlparser.f90:545-549). */
int vl_self_and(vl_engine* e, const vl_layer* a, vl_layer* c);
/* C = rows x cols boxes tiling MBR(A), each padded by `overlap` (overlap/2
 * on the low side, the rest on the high side) and clamped to MBR(A), in
 * row-major order from the bottom-left, exactly as CreateGrid
 * (src/design_impl.f90:159-275; the GRID operator passes 10, 10, 10,
 * src/controlparser.f90:559-562).  The boxes are kept as they are (not
 * merged).  VL_EINVAL when rows or cols < 1 or A is empty. */
int vl_grid(vl_engine* e, const vl_layer* a, int32_t rows, int32_t cols, int32_t overlap, vl_layer* c);
/* C = the one box MBR(A).  Replaces CreateEXTENT (src/design_impl.f90:277-293).
 * VL_EINVAL when A is empty. */
int vl_extent(vl_engine* e, const vl_layer* a, vl_layer* c);
/* Polygon numbering of A's boxes in their stored order: the PNUM step of
 * PreprocessLayer, PerformMerge (src/pnummerge.f90:279-385).  Boxes whose
 * closed extents meet (overlap, shared edge or shared corner) are connected;
 * labels[i] = 0 when box i meets no other box, else the 1-based number of
 * its connected component, components numbered in the order of their
 * smallest box index (the reference's union-find and uf_fullreduce,
 * src/datastructures.f90:90-171).  *ncomp = components of two or more boxes.
 * *overlap_area = the sum over pairs i < k of their positive intersection
 * areas; *touch_perimeter = the sum of the perimeters 2 (w + h) of the
 * zero-area intersections of touching pairs, 0 when overlap_area > 0 (both
 * exact here; the reference accumulates them in real64).  labels: n int64 in
 * host or device memory. */
int vl_pnum(vl_engine* e, const vl_layer* a, int64_t* labels, int64_t* ncomp, int64_t* overlap_area,
            int64_t* touch_perimeter);
/* Pairs i < k of A's boxes that overlap with positive area
 * (ComputeInteractionsGPU, src/interaction.f90:204-310); CertifyHealing
 * (src/design_impl.f90:35-56) holds when it is 0. */
int vl_interaction_count(vl_engine* e, const vl_layer* a, int64_t* count);

/* ---- layer files ------------------------------------------------------------ */
/* KLBIN: a raw array of boxes (x1, y1, x2, y2 as int32, 16 bytes each).
 * LoadKLBin (src/kldata.f90:14-63): a missing or empty file is VL_EINVAL,
 * bytes past the last whole box are not read, every box must be valid. */
int vl_layer_load_klbin(vl_engine* e, vl_layer* l, const char* path);
/* WriteKLBin (src/kldata.f90:64-97): replaces the file; an empty layer writes
 * an empty file. */
int vl_layer_save_klbin(vl_engine* e, const vl_layer* l, const char* path);
/* SNAP container (SaveLayerToSnap + CompressBoxesToStream +
 * SaveCompressedStreamToDisk, src/diskdesign_impl.f90:59-81,
 * src/box_codec.f90:670-764, src/serialization.f90:15-56): the layer's boxes
 * in stored order, 65536 per chunk, codec-encoded and compressed with method
 * 1 (Snappy), 2 (zlib) or 3 (zstd, level 20; needs libzstd.so.1 at run
 * time).  layer_properties are the LAYER_STATE_* bits written to the header;
 * the R-tree bit (8) is cleared and the PNUM bit (4) is set exactly when pnum
 * (n int64 polygon numbers, host memory) is given, which then follow the
 * header.  The reference sorts a layer for its R-tree before saving; the file
 * holds the same boxes in this layer's order. */
int vl_layer_save_snap(vl_engine* e, const vl_layer* l, const char* path, int32_t method, int32_t layer_properties,
                       const int64_t* pnum);
/* The same file from n host boxes (host only: no GPU is used; every box must
 * be valid). */
int vl_snap_write_boxes(const vl_box* boxes, int64_t n, const char* path, int32_t method, int32_t layer_properties,
                        const int64_t* pnum);
/* Reads a SNAP header: properties, box and chunk counts, compression method
 * (an empty file reads as all zeros). */
int vl_snap_info(const char* path, int32_t* layer_properties, int64_t* total_boxes, int64_t* chunks,
                 int32_t* method);
/* RestoreCompressedStreamFromDisk + DecompressStreamToBoxes
 * (src/serialization.f90:58-130, src/box_codec.f90:769-891): chunks are
 * decompressed on the host and the box codec is decoded on the GPU into the
 * layer.  An empty file is an empty layer.  *layer_properties (may be null)
 * receives the header's LAYER_STATE_* bits masked to 63 (RestoreSnapToLayer).
 * When the file holds polygon numbers and pnum is not null they are copied
 * to pnum (VL_ENOSPC when pnum_cap < boxes).  Malformed files (header, chunk
 * sizes, decompression, a codec record past its chunk, a dictionary index
 * out of range, a box count that does not match) are VL_EINVAL; decoded
 * boxes must be valid (VL_EINVAL). */
int vl_layer_load_snap(vl_engine* e, vl_layer* l, const char* path, int32_t* layer_properties, int64_t* pnum,
                       int64_t pnum_cap);
/* RestoreSnapToLayer's dispatch (src/diskdesign_impl.f90:83-114): a path
 * containing ".bin" is KLBIN, anything else SNAP. */
int vl_layer_load_file(vl_engine* e, vl_layer* l, const char* path, int32_t* layer_properties);
/* The box codec alone (CompressBoxesUsingCodec, src/box_codec.f90:565-666):
 * encodes n <= 65536 boxes as one chunk into out (VL_ENOSPC with *out_n set
 * when cap is too small).  Host only. */
int vl_codec_encode(const vl_box* boxes, int64_t n, uint8_t* out, int64_t cap, int64_t* out_n);
/* GPU decode of uncompressed codec chunks (DecompressBoxesUsingCodec,
 * src/box_codec.f90:447-536) into `out`: raw holds the chunks back to back
 * (host memory), chunk_bytes[k] / chunk_boxes[k] their sizes.  `out` must
 * belong to e. */
int vl_codec_decode(vl_engine* e, const uint8_t* raw, const int64_t* chunk_bytes, const int64_t* chunk_boxes,
                    int64_t nchunks, vl_layer* out);
/* Box count of A (the ASSERT_ZERO check, src/controlparser.f90:524-530).
 * Returns VL_EASSERT when *count>0 and abort_on_nonzero!=0. */
int vl_assert_zero(vl_engine* e, const vl_layer* a, int abort_on_nonzero, int64_t* count);
/* Exact int64 union area (calculate_union_area_fast, src/geometry.f90:1840-1927;
 * src/klee.cu:164-210). */
int vl_klee_area(vl_engine* e, const vl_layer* a, int64_t* area);
/* C = merge(A) sorted by (x1,y1,x2,y2): the canonical form used for parity. */
int vl_canonicalize(vl_engine* e, const vl_layer* a, vl_layer* c);
/* *empty = 1 iff A XOR B is empty (the BOOST_XOR_CHECKER exit code,
 * src/boost_xor_checker.cpp:151), computed as the area of A XOR B on the GPU. */
int vl_xor_empty(vl_engine* e, const vl_layer* a, const vl_layer* b, int* empty);

/* One entry for all operators with the Boost bridge's op codes and the host
 * pointers of PerformBoostPolygonOperation (src/boost_polygon_api.cpp:97-105),
 * plus the output capacity the reference lacks: VL_ENOSPC with *out_n set to
 * the required count when cap is too small.  ctrl: VL_OP_*; val: SIZE amount,
 * or for VL_OP_GROW the same amount on all four sides. */
int vl_boolean_op(vl_engine* e, const vl_box* a, uint64_t an, const vl_box* b, uint64_t bn,
                  vl_box* out, uint64_t cap, uint64_t* out_n, uint64_t ctrl, int64_t val);

/* ---- sorts (the Thrust prototypes' roles) --------------------------------- */
/* Stable sort of n trackers by ((X^0x80000000)<<32)|(Y^0x80000000).
 * Replaces device_event_sort (src/event_sorter.cu:48-81).  arr may be host or
 * device memory. */
int vl_event_sort(vl_engine* e, vl_tracker* arr, int64_t n);
/* Stable sort of boxes by the 64-bit Morton code of their centre
 * ((x1+x2)/2 ^ 0x80000000, y in odd bits).  Replaces thrust_sort_boxes_direct
 * (src/gpu_morton.cu:149-164).  Host or device memory. */
int vl_morton_sort_direct(vl_engine* e, vl_box* boxes, int64_t n);
/* 1-based permutation that Morton-sorts boxes (boxes untouched).  Replaces
 * thrust_sort_boxes_indirect (src/gpu_morton.cu:169-187).  Host or device memory. */
int vl_morton_sort_indirect(vl_engine* e, const vl_box* boxes, int64_t* idx, int64_t n);

/* Same-name, same-signature entry points for the reference's Fortran
 * interfaces (src/event_sorter_apif.f90:17-23, src/gpu_morton_sort.f90:23-36).
 * They use a process-wide engine on device 0 and abort the process on error,
 * as the reference's error stop would. */
void device_event_sort(vl_tracker* d_arr, int64_t n);
void thrust_sort_boxes_direct(vl_box* d_boxes, int64_t n);
void thrust_sort_boxes_indirect(vl_box* d_boxes, int64_t* d_indices, int64_t n);

/* No counterpart in the reference: vl_event_sort for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same result as vl_event_sort on device
 * memory; host memory is rejected (VL_EINVAL).  Its scratch is a static
 * device array of VL_NOALLOC_SCRATCH_BYTES per device, held until the process
 * ends, so n is at most VL_NOALLOC_MAX_TRACKERS (VL_EINVAL beyond).  It calls
 * no allocator.  Every call on a device shares that device's array: calls
 * that overlap on one device race (calls on one engine are serialised). */
#define VL_NOALLOC_SCRATCH_BYTES 134217728
#define VL_NOALLOC_MAX_TRACKERS 3500000
int vl_event_sort_noalloc(vl_engine* e, vl_tracker* d_arr, int64_t n);

/* ---- batched form ------------------------------------------------------- */
/* A program is a recorded list of operators over layers (no counterpart in the
 * reference, whose control file runs one operator at a time).  vl_program_run
 * splits the list into independent chains (operator j depends on an earlier i
 * when it reads or writes a layer i writes, or writes a layer i reads) and runs
 * up to four chains concurrently, each in program order on its own worker
 * stream and host thread, after the work already queued on the engine.
 * Before that, consecutive operators that all derive from one source layer S
 * are grouped (a GROW with non-negative arguments, SIZE or MERGE of S starts a
 * group; it takes the following AND / NOT / OR / XOR / MERGE whose inputs are
 * S or group outputs, GROW / SIZE of S itself (at most two distinct ones) and
 * ASSERT_ZERO of group outputs, up to eight outputs): S is binned once and
 * every output is computed per tile from the same masks.  A group that cannot
 * run fused (tiles too dense to split, windows too large for the tiles) runs
 * operator by operator.  Results are the same point sets, in canonical strips,
 * as running the operators one by one (the order of the strips of a fused
 * output may differ; vl_canonicalize sorts).  Layers used by a program must
 * not be touched by other calls while it runs. */
typedef struct vl_program vl_program;
int vl_program_create(vl_engine* e, vl_program** out);
int vl_program_destroy(vl_program* p);
/* op: VL_OP_*; args: GROW E,N,W,S / SIZE d in args[0] / unused.  b may be NULL
 * for unary operators.  VL_OP_MERGE with a=b is the union; ASSERT_ZERO is
 * recorded with op = -1 and reports through vl_program_assert_count. */
int vl_program_add(vl_program* p, int op, const vl_layer* a, const vl_layer* b, vl_layer* c,
                   const int32_t* args);
int vl_program_run(vl_program* p);
/* Number of ASSERT_ZERO checks of the last run that found shapes. */
int vl_program_assert_count(vl_program* p, int64_t* failures);
/* Graph mode (no counterpart in the reference; its control file runs one
 * operator at a time, src/controlparser.f90:308-309 notes that a graph could
 * be built).  With enable != 0 the next vl_program_run runs the program
 * eagerly while logging every size the host reads back, captures the same
 * work on the program's worker streams into a HIP graph (no host waits;
 * each logged size is checked on the device), and launches it; later runs
 * replay the graph.  A replay is used only while every layer the program
 * reads or writes still has the version the graph was captured with (any
 * other write to one of them, or a new upload, makes the next run capture
 * again), and a replay whose device checks fail is recomputed eagerly.
 * Layers written by the program keep their storage across replays.  Results
 * are the same as vl_program_run without graph mode. */
int vl_program_set_graph(vl_program* p, int enable);
/* Captures made, replays used, and replays or captures that fell back to
 * eager execution, over the program's life. */
int vl_program_graph_stats(vl_program* p, int64_t* captures, int64_t* replays, int64_t* fallbacks);

/* ---- serving: windowed queries against resident layers --------------------- */
/* No counterpart in the reference (its control file runs one operator at a
 * time).  A server indexes its layers once (boxes binned into a uniform cell
 * grid whose cells are numbered in Morton order, cell starts kept on the host
 * too) and answers windowed DRC-style queries  R = GROW_s(A ∩ W) AND (B ∩ W)
 * (R is tested for emptiness before it is computed in full, as for an AND
 * read only by ASSERT_ZERO) with continuous batching: requests
 * are released at their arrival times, admitted in arrival order whenever a
 * slot is free and the in-flight memory estimate stays within the HBM budget
 * (a request is always admitted when nothing else is in flight), and run
 * concurrently, one per slot, each slot with its own stream and host thread.
 * Every result is the same alone and in any mix. */
typedef struct vl_server vl_server;
typedef struct vl_query {
  vl_box window;              /* W */
  int32_t layer_a, layer_b;   /* indices into the server's layers */
  int32_t spacing;            /* s >= 0 (GROW on all four sides) */
  int32_t reserved;
  double arrival_ms;          /* release time, relative to the start of vl_server_run */
} vl_query;
typedef struct vl_query_result {
  int32_t status;             /* VL_OK or the failure code */
  int32_t slot;               /* slot that ran it */
  int64_t in_boxes;           /* boxes of A plus boxes of B with positive area inside W, each counted once */
  int64_t count;              /* boxes of R (canonical maximal strips): 0 = ASSERT_ZERO holds */
  uint64_t hash;              /* order-independent hash of R's boxes */
  double start_ms, end_ms;    /* admission and completion, relative to the start of the run */
} vl_query_result;
/* slots: concurrent queries (1..16); hbm_budget: bytes (0 = 64 GiB). */
int vl_server_create(vl_engine* e, const vl_layer* const* layers, int nlayers, int slots, uint64_t hbm_budget,
                     vl_server** out);
/* Runs n queries, replaying their arrival times in real time; res[i] is the
 * result of q[i].  Returns the first failure code (all queries still run). */
int vl_server_run(vl_server* s, const vl_query* q, int64_t n, vl_query_result* res);
int vl_server_destroy(vl_server* s);

/* ---- timing ---------------------------------------------------------------- */
/* Device time (ms) of the last operator / program run on this engine, measured
 * with HIP events on the engine stream from its first to its last kernel. */
int vl_engine_last_time_ms(vl_engine* e, double* ms);
/* Per-kernel statistics collected while the "profile" parameter is 1 (setting
 * it resets them).  Durations come from HIP events on the stream the kernels
 * run on; they are read at the next query, so collecting them adds no host
 * waits.  kernel: "tile" (the single-operator tile stage: the wave-kernel
 * tiers and the workgroup kernel launched back to back), "fused" (the fused
 * tile kernel's tiers of a program's operator group), "bin" (tile binning:
 * span, emit, radix sort, bounds, gather), "sort" (the radix passes inside
 * the binning), "snap" (the GPU box-codec decode of SNAP loads and
 * vl_codec_decode: its five kernels back to back).  *launches counts kernel
 * launches (sort: digit passes) and *bytes is their algorithmic byte count
 * (tile and fused: 16 B per entry read + 16 B per strip written + 48 B per
 * tile of descriptor and range tables; sort: 16 B per entry per pass; bin:
 * 32 B per input box + 8 B per entry emitted + the sort + 32 B per entry
 * gathered; snap: the codec bytes read once + 16 B per box written). */
int vl_engine_kernel_stats(vl_engine* e, const char* kernel, int64_t* launches, double* total_ms,
                           double* bytes);

#ifdef __cplusplus
}
#endif
#endif
