// vl_synth.cpp — seeded synthetic layers for the benchmark workloads (host code).
#include <cmath>
#include <cstdint>
#include <random>

#include "vl_synth.h"

namespace {

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  double u() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
  int64_t below(int64_t n) { return (int64_t)(u() * (double)n); }          // [0,n)
  // floor(exp(U[ln lo, ln(hi+1)))) in [lo, hi]
  int64_t logu(int64_t lo, int64_t hi) {
    double a = std::log((double)lo), b = std::log((double)hi + 1.0);
    int64_t v = (int64_t)std::floor(std::exp(a + u() * (b - a)));
    return v < lo ? lo : (v > hi ? hi : v);
  }
  double normal() {   // Box-Muller
    double u1 = u(), u2 = u();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

uint64_t mix(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t z = seed ^ (a * 0x9E3779B97F4A7C15ull) ^ (b * 0xC2B2AE3D27D4EB4Full);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Fraction of boxes placed inside a random earlier box (tuned so that ~30% of
// all boxes intersect another box at 200k boxes in [0,2^20)^2).
constexpr double CLONE_FRACTION = 0.16;

void gen_uniform(uint64_t seed, int64_t n, vl_box* out) {
  Rng r(mix(seed, 1, 0));
  for (int64_t i = 0; i < n; i++) {
    int64_t w = r.logu(1, 2000), h = r.logu(1, 2000);
    int64_t x, y;
    if (i > 0 && r.u() < CLONE_FRACTION) {
      const vl_box& p = out[r.below(i)];
      x = p.x1 + r.below((int64_t)p.x2 - p.x1);
      y = p.y1 + r.below((int64_t)p.y2 - p.y1);
    } else {
      x = r.below(1 << 20);
      y = r.below(1 << 20);
    }
    out[i] = vl_box{(int32_t)x, (int32_t)y, (int32_t)(x + w), (int32_t)(y + h)};
  }
}

// Standard-cell rows: 2000 cells per row, cell widths 460*k (k in [2,12]),
// 4 boxes of the layer per cell; rows are added until n boxes exist.
void gen_stdcell(uint64_t seed, int layer, int64_t n, vl_box* out) {
  const int64_t ROW_PITCH = 2720, CELLS = 2000, PER_CELL = 4;
  int64_t i = 0;
  for (int64_t row = 0; i < n; row++) {
    Rng cellr(mix(seed, 2, (uint64_t)row));          // cell widths shared by all layers
    Rng r(mix(seed, 3 + (uint64_t)layer, (uint64_t)row));
    const int64_t y0 = row * ROW_PITCH;
    int64_t x0 = 0;
    for (int64_t c = 0; c < CELLS && i < n; c++) {
      const int64_t cw = 460 * (2 + cellr.below(11));
      for (int k = 0; k < PER_CELL && i < n; k++) {
        int64_t bx1, by1, bx2, by2;
        switch (layer) {
          case 0: {   // poly: vertical 10x200
            bx1 = x0 + 5 * r.below((cw - 10) / 5 + 1);
            by1 = y0 + 5 * r.below((ROW_PITCH - 200) / 5 + 1);
            bx2 = bx1 + 10;
            by2 = by1 + 200;
            break;
          }
          case 1: {   // li1 pads 6..60
            int64_t w = 6 + r.below(55), h = 6 + r.below(55);
            bx1 = x0 + 5 * r.below(cw / 5);
            by1 = y0 + 5 * r.below(ROW_PITCH / 5);
            bx2 = bx1 + w;
            by2 = by1 + h;
            break;
          }
          case 2: {   // met1: horizontal 8 x L
            int64_t L = r.logu(20, 5000);
            bx1 = x0 + 5 * r.below(cw / 5);
            by1 = y0 + 5 * r.below(ROW_PITCH / 5);
            bx2 = bx1 + L;
            by2 = by1 + 8;
            break;
          }
          default: {  // met2: vertical 8 x L
            int64_t L = r.logu(20, 5000);
            bx1 = x0 + 5 * r.below(cw / 5);
            by1 = y0 + 5 * r.below(ROW_PITCH / 5);
            bx2 = bx1 + 8;
            by2 = by1 + L;
            break;
          }
        }
        out[i++] = vl_box{(int32_t)bx1, (int32_t)by1, (int32_t)bx2, (int32_t)by2};
      }
      x0 += cw;
    }
  }
}

void gen_tile(uint64_t seed, int layer, int64_t n, vl_box* out) {
  Rng r(mix(seed, 10, (uint64_t)layer));
  for (int64_t i = 0; i < n; i++) {
    int64_t w = r.logu(1, 500), h = r.logu(1, 500);
    int64_t x = r.below(1 << 16), y = r.below(1 << 16);
    out[i] = vl_box{(int32_t)x, (int32_t)y, (int32_t)(x + w), (int32_t)(y + h)};
  }
}

void gen_chip(uint64_t seed, int layer, int64_t n, vl_box* out) {
  const int64_t SPAN = (int64_t)1 << 26;
  for (int64_t i0 = 0; i0 < n; i0 += 65536) {
    Rng r(mix(seed, 20 + (uint64_t)layer, (uint64_t)(i0 >> 16)));
    const int64_t i1 = (i0 + 65536 < n) ? i0 + 65536 : n;
    for (int64_t i = i0; i < i1; i++) {
      int64_t x, y, w, h;
      if (r.u() < 0.7) {
        int64_t L = r.logu(10, 100000);
        x = 5 * r.below(SPAN / 5);
        y = 5 * r.below(SPAN / 5);
        if (layer & 1) { w = L; h = 8; } else { w = 8; h = L; }
      } else {
        x = 5 * r.below(SPAN / 5);
        y = 5 * r.below(SPAN / 5);
        w = h = 6;
      }
      out[i] = vl_box{(int32_t)x, (int32_t)y, (int32_t)(x + w), (int32_t)(y + h)};
    }
  }
}

// Routing tracks for DRC spacing checks: layers 0 and 1 are the met1 wires of
// the even and odd horizontal tracks, layers 2 and 3 the met2 wires of the
// even and odd vertical tracks.  Track pitch 14 DBU, wire width 8, so wires of
// neighbouring tracks (one layer of a pair against the other) are 6 DBU apart;
// along a track, wires of length log-uniform 20..5000 with gaps of 10..100,
// over [0, 2^20).  One wire in 4096 is a jog moved 2 DBU towards the next
// track (4 DBU from it): a spacing check of the pair at 5 DBU fails there, at
// 1 and 2 DBU nowhere.
void gen_route(uint64_t seed, int layer, int64_t n, vl_box* out) {
  const int64_t PITCH = 14, WIDTH = 8, SPAN = (int64_t)1 << 20;
  const bool vertical = layer >= 2;
  int64_t i = 0;
  for (int64_t k = 0; i < n; k++) {
    const int64_t track = 2 * k + (layer & 1);
    Rng r(mix(seed, 40 + (uint64_t)layer, (uint64_t)k));
    for (int64_t pos = r.below(100); pos < SPAN && i < n;) {
      const int64_t L = r.logu(20, 5000);
      const int64_t off = track * PITCH + (r.below(4096) == 0 ? 2 : 0);
      out[i++] = vertical ? vl_box{(int32_t)off, (int32_t)pos, (int32_t)(off + WIDTH), (int32_t)(pos + L)}
                          : vl_box{(int32_t)pos, (int32_t)off, (int32_t)(pos + L), (int32_t)(off + WIDTH)};
      pos += L + 10 + r.below(91);
    }
  }
}

}  // namespace

extern "C" int vl_synth_layer(int config, uint64_t seed, int layer, int64_t n, vl_box* out) {
  if (n < 0 || (n > 0 && !out) || layer < 0) return VL_EINVAL;
  switch (config) {
    case VL_SYN_UNIFORM: gen_uniform(seed + (uint64_t)layer * 0x1000193ull, n, out); return VL_OK;
    case VL_SYN_STDCELL: gen_stdcell(seed, layer, n, out); return VL_OK;
    case VL_SYN_TILES: gen_tile(seed, layer, n, out); return VL_OK;
    case VL_SYN_CHIP: gen_chip(seed, layer, n, out); return VL_OK;
    case VL_SYN_ROUTE:
      if (layer > 3) return VL_EINVAL;
      gen_route(seed, layer, n, out);
      return VL_OK;
    default: return VL_EINVAL;
  }
}

extern "C" int64_t vl_synth_tile_count(uint64_t seed, int tile) {
  Rng r(mix(seed, 30, (uint64_t)tile));
  double v = std::exp(std::log(5000.0) + 1.34 * r.normal());
  int64_t c = (int64_t)v;
  return c < 100 ? 100 : (c > 200000 ? 200000 : c);
}
