/*
 * vl_synth.h — seeded synthetic layers for the benchmark workloads.
 *
 * Not part of the reference interface: the reference has no benchmark suite
 * and its design archives are Git-LFS pointers, so the workloads are
 * generated.  std::mt19937_64 drives hand-written distributions (no
 * std::*_distribution), so a (config, seed, layer, n) tuple names one exact
 * layer on every platform with the same libm.
 *
 *   VL_SYN_UNIFORM  200k-style boxes: lower-left uniform in [0,2^20)^2,
 *                   width/height log-uniform in [1,2000], a fraction of boxes
 *                   cloned into an earlier box so that about 30% overlap.
 *   VL_SYN_STDCELL  standard-cell-like rows (pitch 2720 DBU, 2000 cells per
 *                   row, 5-DBU grid); layer 0 poly (vertical 10x200), 1 li1
 *                   (pads 6..60), 2 met1 (horizontal 8xL, L log-uniform
 *                   20..5000), 3 met2 (vertical 8xL).
 *   VL_SYN_TILES    one tile of the uneven-tile batch: coordinates uniform in
 *                   [0,2^16)^2, sizes log-uniform [1,500]; layer = 2*tile+side.
 *   VL_SYN_CHIP     full-chip routing: [0,2^26)^2, 70% tracks (w=8, length
 *                   log-uniform 10..1e5, 5-DBU grid; layer 0 vertical, 1
 *                   horizontal), 30% 6x6 vias.
 *   VL_SYN_ROUTE    spacing-check design: layers 0/1 the wires (w=8, length
 *                   log-uniform 20..5000) of the even/odd horizontal tracks of
 *                   a 14-DBU pitch (6 DBU between the two layers), 2/3 the
 *                   same vertically, over [0,2^20); one wire in 4096 jogs 2 DBU
 *                   towards its neighbour track (a 5-DBU spacing violation).
 */
#ifndef VL_SYNTH_H
#define VL_SYNTH_H
#include <stdint.h>
#include "vl_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VL_SYN_UNIFORM 1
#define VL_SYN_STDCELL 2
#define VL_SYN_TILES   3
#define VL_SYN_CHIP    4
#define VL_SYN_ROUTE   5

/* Writes exactly n boxes of the given layer into out. */
int vl_synth_layer(int config, uint64_t seed, int layer, int64_t n, vl_box* out);
/* Box count of tile `tile` in the uneven-tile batch: LogNormal(ln 5000, 1.34)
 * clipped to [100, 200000]. */
int64_t vl_synth_tile_count(uint64_t seed, int tile);

#ifdef __cplusplus
}
#endif
#endif
