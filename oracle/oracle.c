/*
 * oracle.c — CPU restatement of the reference rectangle-Boolean path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for parity tests and the CPU baseline
 * of bench.py.  The product never links or calls it.
 *
 * Reference files restated (paths relative to the reference repo root):
 *   src/geometry.f90          Box semantics (:21-35,:66-70,:184-215), Klee area (:1840-1972)
 *   src/polygon_fracture.f90  generate_trackers (:767-803), heal_boxes (:805-1004),
 *                             skip list (:125-250), tracker order (:254-352)
 *   src/design.f90            CalculateOR (:1085-1103), CalculateAND (:1131-1235),
 *                             CalculateNOT (:1238-1268)
 *   src/design_impl.f90       CalculateGROWLayer (:571-586), CreateGrid (:159-275),
 *                             frame functions (:993-1129, :1193-1316), XOR (:961-991)
 *   src/boost_polygon_api.cpp op codes and SIZE (:97-156); Boost.Polygon itself is not
 *                             vendored: SIZE restates Boost's documented 90-degree
 *                             resize (bloat = Minkowski sum with a square, shrink =
 *                             frame \ bloat(frame \ A)).
 *   src/event_sorter.cu, src/gpu_morton.cu  sort keys.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   or_heal / or_canonical  flang build of the reference heal_boxes (oracle/_ref) on the
 *                           reference's own SED/square fixtures and random sets; known
 *                           answer (0,0,2,2)+(2,0,4,2)->(0,0,4,2); brute-force raster
 *                           canonical form on tiny grids.
 *   or_complement           flang build of the reference scanline_fracture; raster.
 *   or_klee_area            flang build of calculate_union_area_fast; known answer 175
 *                           (src/klee.cu:214-224); raster count.
 *   or_and/not/or/xor       raster set algebra on tiny grids; known answer
 *                           (0,0,10,10)\(2,2,8,8) (src/test_rectboolean.f90); area identities.
 *   or_grow / or_size       raster dilation/erosion by a square; single-box closed forms.
 *   or_self_and             raster count of boxes per cell (>= 2); identical boxes.
 *   or_pnum                 brute-force connected components of the closed-touch
 *                           graph (all pairs, breadth-first search) with the label
 *                           rule; known answers for edge, corner and no contact;
 *                           areas by brute force over all pairs.
 *   or_interactions         brute force over all pairs.
 *   or_codec_encode/decode  known answers derived by hand from the format
 *                           (every record form: new value, index, current,
 *                           square, zero and negative deltas); decode of
 *                           encode is the identity, also across int32 wraps.
 *   or_grid                 a hand-evaluated case of CreateGrid's arithmetic; area and
 *                           tiling invariants.
 *   sorts                   library equivalence (Python sorted() with the same keys).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void or_free(void* p) { free(p); }

/* ------------------------------------------------------------------------- */
/* growable box vector                                                        */
typedef struct { ob_box* v; int64_t n, cap; } bvec;
static void bv_push(bvec* b, ob_box x) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 64;
    b->v = (ob_box*)realloc(b->v, (size_t)b->cap * sizeof(ob_box));
  }
  b->v[b->n++] = x;
}
static int64_t bv_release(bvec* b, ob_box** out) {
  if (b->n == 0) { free(b->v); *out = NULL; return 0; }
  *out = b->v;
  return b->n;
}
static inline int box_valid(ob_box b) { return b.x1 < b.x2 && b.y1 < b.y2; }   /* geometry.f90:66-70 */
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------- */
/* Ordered y -> lap_change map: the reference's arena skip list with its      */
/* deterministic hash levels (polygon_fracture.f90:125-250).  Nodes whose     */
/* accumulated lap_change returns to zero are removed.                        */
#define SL_MAXLVL 24
typedef struct { int32_t y; int64_t lap; int32_t fwd[SL_MAXLVL]; } sl_node;
typedef struct { sl_node* a; int32_t cap, freeh, head, level; } skiplist;

static int sl_level(int32_t y) {                  /* get_skip_level, :125-141 */
  uint32_t h = (uint32_t)y;
  h ^= h >> 16; h *= 212519137u; h ^= h >> 15; h *= 401662913u;
  int lvl = 1;
  while ((h & 1u) == 0u && lvl < SL_MAXLVL) { lvl++; h >>= 1; }
  return lvl;
}
static int32_t sl_alloc(skiplist* s) {
  if (s->freeh < 0) {
    int32_t old = s->cap;
    s->cap = s->cap ? s->cap * 2 : 256;
    s->a = (sl_node*)realloc(s->a, (size_t)s->cap * sizeof(sl_node));
    for (int32_t i = old; i < s->cap; i++) s->a[i].fwd[0] = (i + 1 < s->cap) ? i + 1 : -1;
    s->freeh = old;
  }
  int32_t k = s->freeh;
  s->freeh = s->a[k].fwd[0];
  for (int i = 0; i < SL_MAXLVL; i++) s->a[k].fwd[i] = -1;
  return k;
}
static void sl_init(skiplist* s) {
  memset(s, 0, sizeof(*s));
  s->freeh = -1;
  s->head = sl_alloc(s);
  s->a[s->head].y = INT32_MIN;
  s->a[s->head].lap = 0;
  s->level = 1;
}
static void sl_destroy(skiplist* s) { free(s->a); memset(s, 0, sizeof(*s)); }
static void sl_update(skiplist* s, int32_t y, int64_t d) {   /* update_active_edges, :146-250 */
  int32_t upd[SL_MAXLVL];
  int32_t cur = s->head;
  for (int i = s->level - 1; i >= 0; i--) {
    for (;;) {
      int32_t nx = s->a[cur].fwd[i];
      if (nx < 0 || s->a[nx].y >= y) break;
      cur = nx;
    }
    upd[i] = cur;
  }
  int32_t t = s->a[cur].fwd[0];
  if (t >= 0 && s->a[t].y == y) {
    s->a[t].lap += d;
    if (s->a[t].lap == 0) {
      for (int i = 0; i < s->level; i++) {
        if (s->a[upd[i]].fwd[i] != t) break;
        s->a[upd[i]].fwd[i] = s->a[t].fwd[i];
      }
      while (s->level > 1 && s->a[s->head].fwd[s->level - 1] < 0) s->level--;
      s->a[t].fwd[0] = s->freeh;
      s->freeh = t;
    }
    return;
  }
  if (d == 0) return;
  int lvl = sl_level(y);
  if (lvl > s->level) {
    for (int i = s->level; i < lvl; i++) upd[i] = s->head;
    s->level = lvl;
  }
  int32_t k = sl_alloc(s);       /* may realloc: re-read nothing cached by pointer */
  s->a[k].y = y;
  s->a[k].lap = d;
  for (int i = 0; i < lvl; i++) {
    s->a[k].fwd[i] = s->a[upd[i]].fwd[i];
    s->a[upd[i]].fwd[i] = k;
  }
}

/* ------------------------------------------------------------------------- */
/* Tracker order by (X, Y) (is_less_than, polygon_fracture.f90:345-352).      */
static int trk_cmp(const void* pa, const void* pb) {
  const ob_tracker* a = (const ob_tracker*)pa;
  const ob_tracker* b = (const ob_tracker*)pb;
  if (a->X != b->X) return a->X < b->X ? -1 : 1;
  if (a->Y != b->Y) return a->Y < b->Y ? -1 : 1;
  return 0;
}

/* The heal_boxes sweep body (polygon_fracture.f90:866-967) over sorted tokens.
 * At every distinct X the maximal y-intervals with positive winding are
 * listed; intervals identical to the previous slice continue, the others are
 * emitted as [x_start, X) x [y1, y2). */
typedef struct { int32_t y1, y2, xs; } region;
static int64_t heal_sweep(const ob_tracker* ev, int64_t ne, ob_box** out) {
  bvec o = {0};
  skiplist s;
  sl_init(&s);
  region* prev = NULL; region* curr = NULL;
  int64_t rcap = 0, nprev = 0, ncurr = 0;
  for (int64_t i = 0; i < ne; i++) {
    sl_update(&s, ev[i].Y, ev[i].w >= 0 ? 1 : -1);   /* sign(1_int8, polygonNumber) */
    if (!(i == ne - 1 || ev[i].X < ev[i + 1].X)) continue;
    int32_t X = ev[i].X;
    ncurr = 0;
    int64_t lap = 0;
    int32_t ys = 0;
    for (int32_t nd = s.a[s.head].fwd[0]; nd >= 0; nd = s.a[nd].fwd[0]) {
      int64_t d = s.a[nd].lap;
      if (lap <= 0 && lap + d > 0) {
        ys = s.a[nd].y;
      } else if (lap > 0 && lap + d <= 0) {
        if (ncurr == rcap || nprev == rcap) {
          rcap = rcap ? rcap * 2 : 64;
          prev = (region*)realloc(prev, (size_t)rcap * sizeof(region));
          curr = (region*)realloc(curr, (size_t)rcap * sizeof(region));
        }
        curr[ncurr].y1 = ys; curr[ncurr].y2 = s.a[nd].y; curr[ncurr].xs = X;
        ncurr++;
      }
      lap += d;
    }
    int64_t p = 0, c = 0;
    while (p < nprev && c < ncurr) {
      if (prev[p].y1 == curr[c].y1 && prev[p].y2 == curr[c].y2) {
        curr[c].xs = prev[p].xs; p++; c++;
      } else if (prev[p].y1 < curr[c].y1) {
        if (X > prev[p].xs) { ob_box b = {prev[p].xs, prev[p].y1, X, prev[p].y2}; bv_push(&o, b); }
        p++;
      } else if (prev[p].y1 > curr[c].y1) {
        c++;
      } else {
        if (X > prev[p].xs) { ob_box b = {prev[p].xs, prev[p].y1, X, prev[p].y2}; bv_push(&o, b); }
        p++;
      }
    }
    for (; p < nprev; p++)
      if (X > prev[p].xs) { ob_box b = {prev[p].xs, prev[p].y1, X, prev[p].y2}; bv_push(&o, b); }
    region* t = prev; prev = curr; curr = t;
    nprev = ncurr;
  }
  free(prev); free(curr);
  sl_destroy(&s);
  return bv_release(&o, out);
}

/* heal_boxes (polygon_fracture.f90:805-1004): four corner tokens per box,
 * (x1,y1)+1 (x1,y2)-1 (x2,y1)-1 (x2,y2)+1, sorted by (X,Y), then the sweep. */
int64_t or_heal(const ob_box* in, int64_t n, ob_box** out) {
  *out = NULL;
  if (n <= 0) return 0;
  if (n == 1) {                                     /* :830-833 returns the box as is */
    *out = (ob_box*)malloc(sizeof(ob_box));
    (*out)[0] = in[0];
    return 1;
  }
  ob_tracker* ev = (ob_tracker*)malloc((size_t)(4 * n) * sizeof(ob_tracker));
  for (int64_t i = 0; i < n; i++) {
    int32_t x1 = imin(in[i].x1, in[i].x2), x2 = imax(in[i].x1, in[i].x2);
    int32_t y1 = imin(in[i].y1, in[i].y2), y2 = imax(in[i].y1, in[i].y2);
    ev[4 * i + 0] = (ob_tracker){x1, y1, 1};
    ev[4 * i + 1] = (ob_tracker){x1, y2, -1};
    ev[4 * i + 2] = (ob_tracker){x2, y1, -1};
    ev[4 * i + 3] = (ob_tracker){x2, y2, 1};
  }
  qsort(ev, (size_t)(4 * n), sizeof(ob_tracker), trk_cmp);
  int64_t r = heal_sweep(ev, 4 * n, out);
  free(ev);
  return r;
}

static int box_cmp(const void* pa, const void* pb) {
  const ob_box* a = (const ob_box*)pa;
  const ob_box* b = (const ob_box*)pb;
  if (a->x1 != b->x1) return a->x1 < b->x1 ? -1 : 1;
  if (a->y1 != b->y1) return a->y1 < b->y1 ? -1 : 1;
  if (a->x2 != b->x2) return a->x2 < b->x2 ? -1 : 1;
  if (a->y2 != b->y2) return a->y2 < b->y2 ? -1 : 1;
  return 0;
}

/* Canonical form: heal (unique maximal-strip decomposition of the point set),
 * sorted by (x1,y1,x2,y2).  Invalid (empty) boxes contribute no points. */
int64_t or_canonical(const ob_box* in, int64_t n, ob_box** out) {
  ob_box* v = (ob_box*)malloc((size_t)(n > 0 ? n : 1) * sizeof(ob_box));
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++) if (box_valid(in[i])) v[m++] = in[i];
  int64_t r = or_heal(v, m, out);
  free(v);
  if (r > 1) qsort(*out, (size_t)r, sizeof(ob_box), box_cmp);
  return r;
}

/* generate_trackers (polygon_fracture.f90:767-803): bbox tokens with the box
 * winding, every input box with the opposite winding; the sweep then yields
 * bbox minus the union of the boxes (CalculateNOT step 1, design.f90:1259-1262). */
int64_t or_complement(const ob_box* b, int64_t n, ob_box bbox, ob_box** out) {
  *out = NULL;
  if (!box_valid(bbox)) return 0;
  ob_tracker* ev = (ob_tracker*)malloc((size_t)(4 * (n + 1)) * sizeof(ob_tracker));
  ev[0] = (ob_tracker){bbox.x1, bbox.y1, 1};
  ev[1] = (ob_tracker){bbox.x1, bbox.y2, -1};
  ev[2] = (ob_tracker){bbox.x2, bbox.y1, -1};
  ev[3] = (ob_tracker){bbox.x2, bbox.y2, 1};
  for (int64_t i = 0; i < n; i++) {
    ev[4 * i + 4] = (ob_tracker){b[i].x1, b[i].y1, -1};
    ev[4 * i + 5] = (ob_tracker){b[i].x1, b[i].y2, 1};
    ev[4 * i + 6] = (ob_tracker){b[i].x2, b[i].y1, 1};
    ev[4 * i + 7] = (ob_tracker){b[i].x2, b[i].y2, -1};
  }
  qsort(ev, (size_t)(4 * (n + 1)), sizeof(ob_tracker), trk_cmp);
  int64_t r = heal_sweep(ev, 4 * (n + 1), out);
  free(ev);
  return r;
}

/* ------------------------------------------------------------------------- */
/* Klee area: calculate_union_area_fast + update_tree (geometry.f90:1840-1972) */
typedef struct { int32_t x, y1, y2, lap; } kev;
static int kev_cmp(const void* pa, const void* pb) {
  const kev* a = (const kev*)pa; const kev* b = (const kev*)pb;
  if (a->x != b->x) return a->x < b->x ? -1 : 1;
  if (a->lap != b->lap) return a->lap > b->lap ? -1 : 1;  /* +1 before -1 (klee.cu:35-39) */
  return 0;
}
static int i32_cmp(const void* pa, const void* pb) {
  int32_t a = *(const int32_t*)pa, b = *(const int32_t*)pb;
  return a < b ? -1 : (a > b);
}
static void seg_update(int node, int s, int e, int l, int r, int val, const int32_t* uy,
                       int32_t* cnt, int64_t* len) {
  if (l <= s && e <= r) {
    cnt[node] += val;
  } else {
    int mid = s + (e - s) / 2;
    if (l <= mid) seg_update(2 * node, s, mid, l, r, val, uy, cnt, len);
    if (r > mid) seg_update(2 * node + 1, mid + 1, e, l, r, val, uy, cnt, len);
  }
  if (cnt[node] > 0) len[node] = (int64_t)uy[e + 1] - uy[s];
  else if (s == e) len[node] = 0;
  else len[node] = len[2 * node] + len[2 * node + 1];
}
int64_t or_klee_area(const ob_box* in, int64_t n) {
  if (n <= 0) return 0;
  int32_t* yv = (int32_t*)malloc((size_t)(2 * n) * sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) {
    yv[2 * i] = imin(in[i].y1, in[i].y2);
    yv[2 * i + 1] = imax(in[i].y1, in[i].y2);
  }
  qsort(yv, (size_t)(2 * n), sizeof(int32_t), i32_cmp);
  int64_t ny = 1;
  for (int64_t i = 1; i < 2 * n; i++) if (yv[i] != yv[ny - 1]) yv[ny++] = yv[i];
  kev* ev = (kev*)malloc((size_t)(2 * n) * sizeof(kev));
  for (int64_t i = 0; i < n; i++) {
    int32_t y1 = imin(in[i].y1, in[i].y2), y2 = imax(in[i].y1, in[i].y2);
    ev[2 * i] = (kev){imin(in[i].x1, in[i].x2), y1, y2, 1};
    ev[2 * i + 1] = (kev){imax(in[i].x1, in[i].x2), y1, y2, -1};
  }
  qsort(ev, (size_t)(2 * n), sizeof(kev), kev_cmp);
  int64_t tsz = 4 * ny + 4;
  int32_t* cnt = (int32_t*)calloc((size_t)tsz, sizeof(int32_t));
  int64_t* len = (int64_t*)calloc((size_t)tsz, sizeof(int64_t));
  int64_t area = 0, cx = ev[0].x;
  for (int64_t i = 0; i < 2 * n; i++) {
    int64_t dx = (int64_t)ev[i].x - cx;
    if (dx > 0) { area += dx * len[1]; cx = ev[i].x; }
    int32_t* p1 = (int32_t*)bsearch(&ev[i].y1, yv, (size_t)ny, sizeof(int32_t), i32_cmp);
    int32_t* p2 = (int32_t*)bsearch(&ev[i].y2, yv, (size_t)ny, sizeof(int32_t), i32_cmp);
    int j1 = (int)(p1 - yv), j2 = (int)(p2 - yv);
    if (j1 < j2 && ny >= 2) seg_update(1, 0, (int)ny - 2, j1, j2 - 1, ev[i].lap, yv, cnt, len);
  }
  free(yv); free(ev); free(cnt); free(len);
  return area;
}

/* ------------------------------------------------------------------------- */
/* Pairwise positive-area intersections (CalculateAND, design.f90:1168-1194:  */
/* box_interact then A(i)*B(k) kept when its area is positive).  A uniform    */
/* grid over MBR(B) replaces the RTree query; each pair is reported once, in  */
/* the cell holding the intersection's lower-left corner.                     */
int64_t or_and_fragments(const ob_box* a, int64_t na, const ob_box* b, int64_t nb, ob_box** out) {
  bvec o = {0};
  *out = NULL;
  if (na <= 0 || nb <= 0) return 0;
  int64_t bx1 = INT32_MAX, by1 = INT32_MAX, bx2 = INT32_MIN, by2 = INT32_MIN;
  for (int64_t i = 0; i < nb; i++) {
    if (b[i].x1 < bx1) bx1 = b[i].x1;
    if (b[i].y1 < by1) by1 = b[i].y1;
    if (b[i].x2 > bx2) bx2 = b[i].x2;
    if (b[i].y2 > by2) by2 = b[i].y2;
  }
  int64_t G = 1;
  while (G * G * 4 < nb && G < 4096) G *= 2;
  int64_t cw = (bx2 - bx1 + G - 1) / G; if (cw < 1) cw = 1;
  int64_t ch = (by2 - by1 + G - 1) / G; if (ch < 1) ch = 1;
  int64_t* cntc = (int64_t*)calloc((size_t)(G * G + 1), sizeof(int64_t));
#define CX(v) ((int64_t)(((int64_t)(v) - bx1) / cw))
#define CY(v) ((int64_t)(((int64_t)(v) - by1) / ch))
#define CLAMPG(v) ((v) < 0 ? 0 : ((v) >= G ? G - 1 : (v)))
  for (int64_t i = 0; i < nb; i++) {
    if (!box_valid(b[i])) continue;
    int64_t c1 = CLAMPG(CX(b[i].x1)), c2 = CLAMPG(CX((int64_t)b[i].x2 - 1));
    int64_t r1 = CLAMPG(CY(b[i].y1)), r2 = CLAMPG(CY((int64_t)b[i].y2 - 1));
    for (int64_t r = r1; r <= r2; r++) for (int64_t c = c1; c <= c2; c++) cntc[r * G + c + 1]++;
  }
  for (int64_t k = 0; k < G * G; k++) cntc[k + 1] += cntc[k];
  int64_t* cur = (int64_t*)malloc((size_t)(G * G) * sizeof(int64_t));
  memcpy(cur, cntc, (size_t)(G * G) * sizeof(int64_t));
  int64_t* lst = (int64_t*)malloc((size_t)(cntc[G * G] > 0 ? cntc[G * G] : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < nb; i++) {
    if (!box_valid(b[i])) continue;
    int64_t c1 = CLAMPG(CX(b[i].x1)), c2 = CLAMPG(CX((int64_t)b[i].x2 - 1));
    int64_t r1 = CLAMPG(CY(b[i].y1)), r2 = CLAMPG(CY((int64_t)b[i].y2 - 1));
    for (int64_t r = r1; r <= r2; r++) for (int64_t c = c1; c <= c2; c++) lst[cur[r * G + c]++] = i;
  }
  for (int64_t i = 0; i < na; i++) {
    ob_box A = a[i];
    if (!box_valid(A)) continue;
    if (A.x2 <= bx1 || A.x1 >= bx2 || A.y2 <= by1 || A.y1 >= by2) continue;
    int64_t c1 = CLAMPG(CX(A.x1)), c2 = CLAMPG(CX((int64_t)A.x2 - 1));
    int64_t r1 = CLAMPG(CY(A.y1)), r2 = CLAMPG(CY((int64_t)A.y2 - 1));
    for (int64_t r = r1; r <= r2; r++)
      for (int64_t c = c1; c <= c2; c++)
        for (int64_t k = cntc[r * G + c]; k < cntc[r * G + c + 1]; k++) {
          ob_box B = b[lst[k]];
          ob_box I = {imax(A.x1, B.x1), imax(A.y1, B.y1), imin(A.x2, B.x2), imin(A.y2, B.y2)};
          if (!box_valid(I)) continue;                   /* positive area only */
          if (CLAMPG(CX(I.x1)) != c || CLAMPG(CY(I.y1)) != r) continue;
          bv_push(&o, I);
        }
  }
#undef CX
#undef CY
#undef CLAMPG
  free(cntc); free(cur); free(lst);
  return bv_release(&o, out);
}

/* ------------------------------------------------------------------------- */
/* Operators                                                                  */
static ob_box mbr_of(const ob_box* v, int64_t n) {          /* mbr_of_array, geometry.f90:137-148 */
  ob_box m = {INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN};
  for (int64_t i = 0; i < n; i++) {
    m.x1 = imin(m.x1, v[i].x1); m.y1 = imin(m.y1, v[i].y1);
    m.x2 = imax(m.x2, v[i].x2); m.y2 = imax(m.y2, v[i].y2);
  }
  return m;
}

/* CalculateAND: heal both inputs (PreprocessLayer), all positive-area pairwise
 * intersections, then the union of the fragments. */
int64_t or_and(const ob_box* a, int64_t na, const ob_box* b, int64_t nb, ob_box** out) {
  ob_box *ha = NULL, *hb = NULL, *fr = NULL;
  int64_t nha = or_heal(a, na, &ha), nhb = or_heal(b, nb, &hb);
  int64_t nf = or_and_fragments(ha, nha, hb, nhb, &fr);
  int64_t r = or_canonical(fr, nf, out);
  free(ha); free(hb); free(fr);
  return r;
}

/* CalculateNOT ('@', design.f90:1238-1268): complement of B inside MBR(A), then
 * AND with A.  FrameNOT ('~') computes the same point set per tile with Boost. */
int64_t or_not(const ob_box* a, int64_t na, const ob_box* b, int64_t nb, ob_box** out) {
  *out = NULL;
  if (na <= 0) return 0;
  ob_box* cb = NULL;
  int64_t ncb = or_complement(b, nb, mbr_of(a, na), &cb);
  int64_t r = or_and(a, na, cb, ncb, out);
  free(cb);
  return r;
}

/* CalculateOR (design.f90:1085-1103): concatenation, then heal. */
int64_t or_or(const ob_box* a, int64_t na, const ob_box* b, int64_t nb, ob_box** out) {
  ob_box* v = (ob_box*)malloc((size_t)(na + nb + 1) * sizeof(ob_box));
  if (na > 0) memcpy(v, a, (size_t)na * sizeof(ob_box));
  if (nb > 0) memcpy(v + na, b, (size_t)nb * sizeof(ob_box));
  int64_t r = or_canonical(v, na + nb, out);
  free(v);
  return r;
}

/* CalculateExplicitXOR (design_impl.f90:961-991): (A\B) OR (B\A). */
int64_t or_xor(const ob_box* a, int64_t na, const ob_box* b, int64_t nb, ob_box** out) {
  ob_box *p = NULL, *q = NULL;
  int64_t np = or_not(a, na, b, nb, &p), nq = or_not(b, nb, a, na, &q);
  int64_t r = or_or(p, np, q, nq, out);
  free(p); free(q);
  return r;
}

/* CalculateGROWLayer + box_grow_directional (design_impl.f90:571-586,
 * geometry.f90:184-193): x1-=W, x2+=E, y1-=S, y2+=N per box, then merge. */
int64_t or_grow(const ob_box* a, int64_t na, int32_t e, int32_t n, int32_t w, int32_t s, ob_box** out) {
  *out = NULL;
  ob_box* v = (ob_box*)malloc((size_t)(na > 0 ? na : 1) * sizeof(ob_box));
  for (int64_t i = 0; i < na; i++) {
    int64_t x1 = (int64_t)a[i].x1 - w, x2 = (int64_t)a[i].x2 + e;
    int64_t y1 = (int64_t)a[i].y1 - s, y2 = (int64_t)a[i].y2 + n;
    if (!(x1 < x2 && y1 < y2) || x1 < INT32_MIN || x2 > INT32_MAX || y1 < INT32_MIN || y2 > INT32_MAX) {
      free(v);
      return -1;
    }
    v[i] = (ob_box){(int32_t)x1, (int32_t)y1, (int32_t)x2, (int32_t)y2};
  }
  int64_t r = or_canonical(v, na, out);
  free(v);
  return r;
}

/* SIZE d: Boost resize on a 90-degree set.  d >= 0: bloat = Minkowski sum with
 * the square [-d,d]^2, i.e. GROW(d,d,d,d).  d < 0: shrink by e=-d = erosion:
 * A \ bloat_e(F \ A) with the frame F = MBR(A) grown by 1. */
int64_t or_size(const ob_box* a, int64_t na, int32_t d, ob_box** out) {
  *out = NULL;
  if (na <= 0) return 0;
  if (d >= 0) return or_grow(a, na, d, d, d, d, out);
  int32_t e = -d;
  ob_box m = mbr_of(a, na);
  ob_box F = {m.x1 - 1, m.y1 - 1, m.x2 + 1, m.y2 + 1};
  ob_box *c = NULL, *g = NULL;
  int64_t nc = or_complement(a, na, F, &c);
  int64_t ng = or_grow(c, nc, e, e, e, e, &g);
  int64_t r = or_not(a, na, g, ng < 0 ? 0 : ng, out);
  free(c); free(g);
  return r;
}

/* ------------------------------------------------------------------------- */
/* Sorts                                                                      */
typedef struct { uint64_t k; int64_t i; } kidx;
static int kidx_cmp(const void* pa, const void* pb) {
  const kidx* a = (const kidx*)pa; const kidx* b = (const kidx*)pb;
  if (a->k != b->k) return a->k < b->k ? -1 : 1;
  return a->i < b->i ? -1 : (a->i > b->i);       /* index tie-break = stable */
}
static inline uint64_t tracker_key(ob_tracker t) {  /* PackKeys, event_sorter.cu:48-58 */
  uint32_t ux = (uint32_t)t.X ^ 0x80000000u, uy = (uint32_t)t.Y ^ 0x80000000u;
  return ((uint64_t)ux << 32) | uy;
}
void or_event_sort(ob_tracker* t, int64_t n) {
  if (n <= 1) return;
  kidx* v = (kidx*)malloc((size_t)n * sizeof(kidx));
  for (int64_t i = 0; i < n; i++) { v[i].k = tracker_key(t[i]); v[i].i = i; }
  qsort(v, (size_t)n, sizeof(kidx), kidx_cmp);
  ob_tracker* c = (ob_tracker*)malloc((size_t)n * sizeof(ob_tracker));
  memcpy(c, t, (size_t)n * sizeof(ob_tracker));
  for (int64_t i = 0; i < n; i++) t[i] = c[v[i].i];
  free(c); free(v);
}
static inline uint64_t spread32(uint32_t c) {
  uint64_t m = c;
  m = (m | (m << 16)) & 0x0000FFFF0000FFFFull;
  m = (m | (m << 8)) & 0x00FF00FF00FF00FFull;
  m = (m | (m << 4)) & 0x0F0F0F0F0F0F0F0Full;
  m = (m | (m << 2)) & 0x3333333333333333ull;
  m = (m | (m << 1)) & 0x5555555555555555ull;
  return m;
}
uint64_t or_morton_key(ob_box b) {                    /* ComputeMorton, gpu_morton.cu:113-144 */
  uint32_t cx = (uint32_t)(((int64_t)b.x1 + b.x2) / 2) ^ 0x80000000u;
  uint32_t cy = (uint32_t)(((int64_t)b.y1 + b.y2) / 2) ^ 0x80000000u;
  return (spread32(cy) << 1) | spread32(cx);
}
void or_morton_sort_direct(ob_box* b, int64_t n) {
  if (n <= 1) return;
  kidx* v = (kidx*)malloc((size_t)n * sizeof(kidx));
  for (int64_t i = 0; i < n; i++) { v[i].k = or_morton_key(b[i]); v[i].i = i; }
  qsort(v, (size_t)n, sizeof(kidx), kidx_cmp);
  ob_box* c = (ob_box*)malloc((size_t)n * sizeof(ob_box));
  memcpy(c, b, (size_t)n * sizeof(ob_box));
  for (int64_t i = 0; i < n; i++) b[i] = c[v[i].i];
  free(c); free(v);
}
void or_morton_sort_indirect(const ob_box* b, int64_t* idx, int64_t n) {
  if (n <= 0) return;
  if (n == 1) { idx[0] = 1; return; }
  kidx* v = (kidx*)malloc((size_t)n * sizeof(kidx));
  for (int64_t i = 0; i < n; i++) { v[i].k = or_morton_key(b[i]); v[i].i = i; }
  qsort(v, (size_t)n, sizeof(kidx), kidx_cmp);
  for (int64_t i = 0; i < n; i++) idx[i] = v[i].i + 1;   /* 1-based */
  free(v);
}

/* CalculateSingleLayerAND (design_impl.f90:385-509): for every pair i < k of */
/* A's boxes that interact, A(i)*A(k) is kept when its area is positive       */
/* (:459-470); the union of the fragments, canonical.  The RTree query is     */
/* replaced by a sweep over the boxes sorted by x1.                           */
static const ob_box* g_sort_base;
static int idx_x1_cmp(const void* pa, const void* pb) {
  const ob_box a = g_sort_base[*(const int64_t*)pa], b = g_sort_base[*(const int64_t*)pb];
  if (a.x1 != b.x1) return a.x1 < b.x1 ? -1 : 1;
  return (*(const int64_t*)pa < *(const int64_t*)pb) ? -1 : 1;
}
int64_t or_self_and(const ob_box* a, int64_t na, ob_box** out) {
  bvec o = {0};
  *out = NULL;
  if (na <= 1) return 0;
  int64_t* id = (int64_t*)malloc((size_t)na * sizeof(int64_t));
  for (int64_t i = 0; i < na; i++) id[i] = i;
  g_sort_base = a;
  qsort(id, (size_t)na, sizeof(int64_t), idx_x1_cmp);
  for (int64_t p = 0; p < na; p++) {
    const ob_box A = a[id[p]];
    for (int64_t q = p + 1; q < na; q++) {
      const ob_box B = a[id[q]];
      if (B.x1 >= A.x2) break;                     /* sorted by x1: no later box meets A */
      ob_box I = {imax(A.x1, B.x1), imax(A.y1, B.y1), imin(A.x2, B.x2), imin(A.y2, B.y2)};
      if (box_valid(I)) bv_push(&o, I);
    }
  }
  free(id);
  ob_box* fr = NULL;
  const int64_t nf = bv_release(&o, &fr);
  const int64_t r = or_canonical(fr, nf, out);
  free(fr);
  return r;
}

/* PerformMerge (pnummerge.f90:279-385): every pair i < k whose closed       */
/* extents meet (:336-341, <= on both axes) is an edge; process_edges         */
/* (:99-110) inserts both ends and merges them in the union-find of           */
/* datastructures.f90: the larger root is hooked under the smaller (:90-112); */
/* uf_fullreduce (:143-171) renumbers the roots 1.. in index order and leaves */
/* never-inserted boxes at 0.  The overlap area sums the positive             */
/* intersection areas, the perimeter the zero-area intersections' 2 (w + h),  */
/* reset to 0 when the area is not 0 (:377-379).  A sweep over the boxes      */
/* sorted by x1 replaces the RTree query.                                     */
static int64_t uf_root(int64_t* p, int64_t x) {
  while (p[x] != x) x = p[x];
  return x;
}
int64_t or_pnum(const ob_box* a, int64_t n, int64_t* labels, int64_t* area, int64_t* perim) {
  *area = 0;
  *perim = 0;
  if (n <= 0) return 0;
  int64_t* par = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  unsigned char* ins = (unsigned char*)calloc((size_t)n, 1);
  for (int64_t i = 0; i < n; i++) par[i] = i;
  int64_t* id = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) id[i] = i;
  g_sort_base = a;
  qsort(id, (size_t)n, sizeof(int64_t), idx_x1_cmp);
  int64_t ar = 0, pe = 0;
  for (int64_t p = 0; p < n; p++) {
    const ob_box A = a[id[p]];
    for (int64_t q = p + 1; q < n; q++) {
      const ob_box B = a[id[q]];
      if (B.x1 > A.x2) break;                          /* closed: x1 == x2 still meets */
      const int32_t ix1 = imax(A.x1, B.x1), iy1 = imax(A.y1, B.y1), ix2 = imin(A.x2, B.x2), iy2 = imin(A.y2, B.y2);
      if (ix1 > ix2 || iy1 > iy2) continue;
      const int64_t i = id[p], k = id[q];
      ins[i] = ins[k] = 1;
      const int64_t ri = uf_root(par, i), rk = uf_root(par, k);
      if (ri < rk) par[rk] = ri; else if (rk < ri) par[ri] = rk;
      if (ix1 < ix2 && iy1 < iy2) ar += (int64_t)(ix2 - ix1) * (iy2 - iy1);
      else pe += 2 * ((int64_t)(ix2 - ix1) + (iy2 - iy1));
    }
  }
  int64_t next = 0;
  int64_t* map = (int64_t*)calloc((size_t)n, sizeof(int64_t));
  for (int64_t i = 0; i < n; i++)
    if (ins[i] && uf_root(par, i) == i) map[i] = ++next;
  for (int64_t i = 0; i < n; i++) labels[i] = ins[i] ? map[uf_root(par, i)] : 0;
  free(par); free(ins); free(id); free(map);
  *area = ar;
  *perim = ar ? 0 : pe;
  return next;
}

/* ComputeInteractionsGPU (interaction.f90:204-310): pairs i < j with          */
/* max(x1) < min(x2) and max(y1) < min(y2) (strict, :268-273).                */
int64_t or_interactions(const ob_box* a, int64_t n) {
  if (n < 2) return 0;
  int64_t* id = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) id[i] = i;
  g_sort_base = a;
  qsort(id, (size_t)n, sizeof(int64_t), idx_x1_cmp);
  int64_t c = 0;
  for (int64_t p = 0; p < n; p++) {
    const ob_box A = a[id[p]];
    for (int64_t q = p + 1; q < n; q++) {
      const ob_box B = a[id[q]];
      if (B.x1 >= A.x2) break;
      if (imax(A.y1, B.y1) < imin(A.y2, B.y2)) c++;
    }
  }
  free(id);
  return c;
}

/* The SNAP box codec, one chunk (box_codec.f90:447-666).  Per box an info   */
/* byte: bit 4 square (h = w), bit 5/6 x1/y1 moved (zigzag varint of the     */
/* int32 delta from the previous box), bit 0 width not the current           */
/* dictionary entry (then bit 1: new value as a varint, appended to the      */
/* dictionary; else its 1-based int16 index), bits 2/3 the same for the      */
/* height (not present for squares).  Dictionaries, cursor and current       */
/* indices start empty / 0 per chunk; varints are 7-bit little-endian groups */
/* with a continuation bit (box_codec.f90:342-442).  Arithmetic on x, y, w   */
/* and h is int32 with wrap-around, as the reference's int32 code does.      */
static int64_t zz_enc(int32_t v) { int64_t t = v; return t < 0 ? (-t) * 2 - 1 : t * 2; }
static int32_t zz_dec(uint64_t v) { return (int32_t)(uint32_t)((v >> 1) ^ (0 - (v & 1))); }
static int64_t put_varint(uint8_t* o, int64_t p, uint64_t v) {
  for (;;) {
    const uint8_t b = (uint8_t)(v & 127u);
    v >>= 7;
    if (!v) { o[p++] = b; return p; }
    o[p++] = (uint8_t)(b | 128u);
  }
}
static int64_t get_varint(const uint8_t* in, int64_t p, int64_t len, uint64_t* v) {
  uint64_t r = 0;
  int sh = 0;
  for (;;) {
    if (p >= len || sh > 63) return -1;
    const uint8_t b = in[p++];
    r |= (uint64_t)(b & 127u) << sh;
    if (!(b & 128u)) break;
    sh += 7;
  }
  *v = r;
  return p;
}
static int64_t dict_find(const int32_t* d, int64_t n, int32_t v) {
  for (int64_t i = 0; i < n; i++)
    if (d[i] == v) return i + 1;
  return -1;
}
int64_t or_codec_encode(const ob_box* b, int64_t n, uint8_t* out, int64_t cap) {
  int32_t* uw = (int32_t*)malloc((size_t)(n + 1) * sizeof(int32_t));
  int32_t* uh = (int32_t*)malloc((size_t)(n + 1) * sizeof(int32_t));
  int64_t nw = 0, nh = 0, cw = 0, chh = 0, p = 0;
  int32_t cx = 0, cy = 0;
  for (int64_t i = 0; i < n; i++) {
    if (p + 48 > cap) { p = -1; break; }
    const int32_t w = (int32_t)((uint32_t)b[i].x2 - (uint32_t)b[i].x1);
    const int32_t h = (int32_t)((uint32_t)b[i].y2 - (uint32_t)b[i].y1);
    uint8_t info = 0;
    const int sq = w == h;
    if (sq) info |= 16;
    const int32_t dx = (int32_t)((uint32_t)b[i].x1 - (uint32_t)cx);
    if (dx) { info |= 32; cx = b[i].x1; }
    const int32_t dy = (int32_t)((uint32_t)b[i].y1 - (uint32_t)cy);
    if (dy) { info |= 64; cy = b[i].y1; }
    int64_t wi = dict_find(uw, nw, w);
    if (wi != cw || nw == 0) {
      info |= 1;
      if (wi <= 0) { info |= 2; uw[nw++] = w; wi = nw; }
      cw = wi;
    }
    int64_t hi = 0;
    if (!sq) {
      hi = dict_find(uh, nh, h);
      if (hi != chh || nh == 0) {
        info |= 4;
        if (hi <= 0) { info |= 8; uh[nh++] = h; hi = nh; }
        chh = hi;
      }
    }
    out[p++] = info;
    if (info & 32) p = put_varint(out, p, (uint64_t)zz_enc(dx));
    if (info & 64) p = put_varint(out, p, (uint64_t)zz_enc(dy));
    if (info & 1) {
      if (info & 2) p = put_varint(out, p, (uint64_t)(int64_t)w);
      else { out[p++] = (uint8_t)(wi & 255); out[p++] = (uint8_t)((wi >> 8) & 255); }
    }
    if (!sq && (info & 4)) {
      if (info & 8) p = put_varint(out, p, (uint64_t)(int64_t)h);
      else { out[p++] = (uint8_t)(hi & 255); out[p++] = (uint8_t)((hi >> 8) & 255); }
    }
  }
  free(uw); free(uh);
  return p;
}
int or_codec_decode(const uint8_t* in, int64_t len, int64_t n, ob_box* out) {
  int32_t* uw = (int32_t*)malloc((size_t)(n + 1) * sizeof(int32_t));
  int32_t* uh = (int32_t*)malloc((size_t)(n + 1) * sizeof(int32_t));
  int64_t nw = 0, nh = 0, cw = 0, chh = 0, p = 0;
  int32_t cx = 0, cy = 0;
  int rc = 0;
  for (int64_t i = 0; i < n && !rc; i++) {
    if (p >= len) { rc = 1; break; }
    const uint8_t info = in[p++];
    const int sq = (info >> 4) & 1;
    uint64_t v;
    if (info & 32) { if ((p = get_varint(in, p, len, &v)) < 0) { rc = 1; break; } cx = (int32_t)((uint32_t)cx + (uint32_t)zz_dec(v)); }
    if (info & 64) { if ((p = get_varint(in, p, len, &v)) < 0) { rc = 1; break; } cy = (int32_t)((uint32_t)cy + (uint32_t)zz_dec(v)); }
    int32_t w, h;
    if (info & 1) {
      if (info & 2) {
        if ((p = get_varint(in, p, len, &v)) < 0) { rc = 1; break; }
        uw[nw++] = (int32_t)(uint32_t)v;
        cw = nw;
      } else {
        if (p + 2 > len) { rc = 1; break; }
        const int16_t k = (int16_t)(in[p] | (in[p + 1] << 8));
        p += 2;
        if (k < 1 || k > nw) { rc = 1; break; }
        cw = k;
      }
    }
    if (cw < 1) { rc = 1; break; }
    w = uw[cw - 1];
    if (sq) {
      h = w;
    } else {
      if (info & 4) {
        if (info & 8) {
          if ((p = get_varint(in, p, len, &v)) < 0) { rc = 1; break; }
          uh[nh++] = (int32_t)(uint32_t)v;
          chh = nh;
        } else {
          if (p + 2 > len) { rc = 1; break; }
          const int16_t k = (int16_t)(in[p] | (in[p + 1] << 8));
          p += 2;
          if (k < 1 || k > nh) { rc = 1; break; }
          chh = k;
        }
      }
      if (chh < 1) { rc = 1; break; }
      h = uh[chh - 1];
    }
    out[i].x1 = cx; out[i].y1 = cy;
    out[i].x2 = (int32_t)((uint32_t)cx + (uint32_t)w);
    out[i].y2 = (int32_t)((uint32_t)cy + (uint32_t)h);
  }
  if (!rc && p != len) rc = 2;   /* bytes left over */
  free(uw); free(uh);
  return rc;
}

/* CreateGrid (design_impl.f90:159-275): rows x cols cells of MBR m, base    */
/* size plus one unit for the first (extent mod count) cells (:204-214),      */
/* padded by overlap/2 below / left and the rest above / right (:219-223),    */
/* clamped to m (:253-257); row-major, rows outer (:236-262).                 */
int64_t or_grid(ob_box m, int32_t rows, int32_t cols, int32_t overlap, ob_box** out) {
  *out = NULL;
  if (rows < 1 || cols < 1 || !box_valid(m)) return -1;
  const int64_t n = (int64_t)rows * cols;
  ob_box* v = (ob_box*)malloc((size_t)n * sizeof(ob_box));
  const int64_t fw = (int64_t)m.x2 - m.x1, fh = (int64_t)m.y2 - m.y1;
  const int64_t cw = fw / cols, ew = fw % cols, rh = fh / rows, eh = fh % rows;
  const int64_t pl = overlap / 2, pr = overlap - pl;
  int64_t k = 0;
  for (int64_t i = 1; i <= rows; i++) {
    const int64_t y0 = m.y1 + (i - 1) * rh + (i - 1 < eh ? i - 1 : eh);
    const int64_t y1 = m.y1 + i * rh + (i < eh ? i : eh);
    for (int64_t j = 1; j <= cols; j++) {
      const int64_t x0 = m.x1 + (j - 1) * cw + (j - 1 < ew ? j - 1 : ew);
      const int64_t x1 = m.x1 + j * cw + (j < ew ? j : ew);
      int64_t bx1 = x0 - pl, bx2 = x1 + pr, by1 = y0 - pl, by2 = y1 + pr;
      if (bx1 < m.x1) bx1 = m.x1;
      if (bx2 > m.x2) bx2 = m.x2;
      if (by1 < m.y1) by1 = m.y1;
      if (by2 > m.y2) by2 = m.y2;
      v[k++] = (ob_box){(int32_t)bx1, (int32_t)by1, (int32_t)bx2, (int32_t)by2};
    }
  }
  *out = v;
  return n;
}

/* ------------------------------------------------------------------------- */
/* CPU baseline: frame-tiled operators with one OpenMP task per tile.         */
/* K_GRID = floor(CalculateTotalNodes(max(NA,NB),16)^(1/4)) (design_impl.f90: */
/* 1013-1015, rtree.f90:19-35); tiles from CreateGrid on the frame MBR        */
/* (design_impl.f90:159-275); per tile FilterLayer clips the inputs           */
/* (:99-143); the per-tile results are concatenated (:1110-1120).             */
static int64_t total_nodes(int64_t n, int64_t cap) {
  if (n == 0) return 0;
  int64_t lvl = (n + cap - 1) / cap, t = lvl;
  while (lvl > 1) { lvl = (lvl + cap - 1) / cap; t += lvl; }
  return t;
}
static int k_grid(int64_t n) {
  double t = (double)total_nodes(n, 16);
  int k = 1;
  while ((double)(k + 1) * (k + 1) * (k + 1) * (k + 1) <= t) k++;
  return k;
}
static ob_box grid_cell(ob_box m, int K, int r, int c) {
  int32_t fw = m.x2 - m.x1, fh = m.y2 - m.y1;
  int32_t cw = fw / K, ew = fw % K, rh = fh / K, eh = fh % K;
  ob_box g;
  g.x1 = m.x1 + c * cw + imin(c, ew);
  g.x2 = m.x1 + (c + 1) * cw + imin(c + 1, ew);
  g.y1 = m.y1 + r * rh + imin(r, eh);
  g.y2 = m.y1 + (r + 1) * rh + imin(r + 1, eh);
  return g;
}
static int64_t clip_to(const ob_box* v, int64_t n, ob_box t, ob_box** out) {
  bvec o = {0};
  for (int64_t i = 0; i < n; i++) {
    ob_box c = {imax(v[i].x1, t.x1), imax(v[i].y1, t.y1), imin(v[i].x2, t.x2), imin(v[i].y2, t.y2)};
    if (box_valid(c)) bv_push(&o, c);
  }
  return bv_release(&o, out);
}
static int64_t tile_op(int op, const ob_box* a, int64_t na, const ob_box* b, int64_t nb,
                       const int32_t* args, ob_box** out) {
  switch (op) {
    case 0: return or_xor(a, na, b, nb, out);
    case 1: return or_or(a, na, b, nb, out);
    case 2: { ob_box* f = NULL; int64_t nf = or_and_fragments(a, na, b, nb, &f); *out = f; return nf; }
    case 3: return or_not(a, na, b, nb, out);
    case 4: return or_heal(a, na, out);
    default: *out = NULL; return 0;
  }
}
static int64_t par_frame(int op, const ob_box* a, int64_t na, const ob_box* b, int64_t nb,
                         ob_box frame, int nthreads, ob_box** out) {
  *out = NULL;
  if (!box_valid(frame)) return 0;
  int K = k_grid(na > nb ? na : nb);
  int T = K * K;
  ob_box** res = (ob_box**)calloc((size_t)T, sizeof(ob_box*));
  int64_t* cnt = (int64_t*)calloc((size_t)T, sizeof(int64_t));
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int t = 0; t < T; t++) {
    ob_box g = grid_cell(frame, K, t / K, t % K);
    ob_box *ca = NULL, *cb = NULL;
    int64_t nca = clip_to(a, na, g, &ca), ncb = clip_to(b, nb, g, &cb);
    cnt[t] = tile_op(op, ca, nca, cb, ncb, NULL, &res[t]);
    free(ca); free(cb);
  }
  int64_t tot = 0;
  for (int t = 0; t < T; t++) tot += cnt[t];
  ob_box* o = (ob_box*)malloc((size_t)(tot > 0 ? tot : 1) * sizeof(ob_box));
  int64_t k = 0;
  for (int t = 0; t < T; t++) {
    if (cnt[t] > 0) memcpy(o + k, res[t], (size_t)cnt[t] * sizeof(ob_box));
    k += cnt[t];
    free(res[t]);
  }
  free(res); free(cnt);
  if (tot == 0) { free(o); o = NULL; }
  *out = o;
  return tot;
}
/* MergeHealLayer (design_impl.f90:1318-1341): heal below 1024 boxes, else per tile. */
static int64_t par_merge(const ob_box* a, int64_t na, int nthreads, ob_box** out) {
  if (na < 1024) return or_heal(a, na, out);
  return par_frame(4, a, na, NULL, 0, mbr_of(a, na), nthreads, out);
}
int64_t or_par_op(int op, const ob_box* a, int64_t na, const ob_box* b, int64_t nb,
                  const int32_t* args, int nthreads, ob_box** out) {
  *out = NULL;
  ob_box m = mbr_of(a, na);
  switch (op) {
    case 0: case 3: {                   /* frame XOR / FrameNOT, then merge */
      /* FrameNOT tiles MBR(A).  The reference tiles XOR over MBR(A) too and so
       * drops B outside it (design_impl.f90:1030); the frame here is
       * MBR(A) u MBR(B), which keeps the XOR result exact. */
      ob_box f = m;
      if (op == 0) {
        ob_box mb = mbr_of(b, nb);
        f = (ob_box){imin(m.x1, mb.x1), imin(m.y1, mb.y1), imax(m.x2, mb.x2), imax(m.y2, mb.y2)};
      }
      ob_box *t = NULL;
      int64_t nt = par_frame(op, a, na, b, nb, f, nthreads, &t);
      int64_t r = par_merge(t, nt, nthreads, out);
      free(t);
      return r;
    }
    case 1: {                           /* OR2: concatenate + merge */
      ob_box* v = (ob_box*)malloc((size_t)(na + nb + 1) * sizeof(ob_box));
      if (na) memcpy(v, a, (size_t)na * sizeof(ob_box));
      if (nb) memcpy(v + na, b, (size_t)nb * sizeof(ob_box));
      int64_t r = par_merge(v, na + nb, nthreads, out);
      free(v);
      return r;
    }
    case 2: {                           /* AND: parallel pairwise fragments (output marked healed) */
      ob_box mb = mbr_of(b, nb);
      ob_box f = {imax(m.x1, mb.x1), imax(m.y1, mb.y1), imin(m.x2, mb.x2), imin(m.y2, mb.y2)};
      return par_frame(2, a, na, b, nb, f, nthreads, out);
    }
    case 4: return par_merge(a, na, nthreads, out);
    case 5: case 6: {
      int32_t e, n_, w, s;
      if (op == 5) { e = n_ = w = s = args[0]; } else { e = args[0]; n_ = args[1]; w = args[2]; s = args[3]; }
      if (op == 5 && args[0] < 0) {     /* shrink: A \ grow(frame \ A) */
        int32_t d = -args[0];
        ob_box F = {m.x1 - 1, m.y1 - 1, m.x2 + 1, m.y2 + 1};
        ob_box *c = NULL, *g = NULL;
        int64_t nc = par_frame(3, &F, 1, a, na, F, nthreads, &c);
        int32_t gargs[4] = {d, d, d, d};
        int64_t ng = or_par_op(6, c, nc, NULL, 0, gargs, nthreads, &g);
        int64_t r = or_par_op(3, a, na, g, ng < 0 ? 0 : ng, NULL, nthreads, out);
        free(c); free(g);
        return r;
      }
      ob_box* v = (ob_box*)malloc((size_t)(na > 0 ? na : 1) * sizeof(ob_box));
      for (int64_t i = 0; i < na; i++) {
        v[i] = (ob_box){a[i].x1 - w, a[i].y1 - s, a[i].x2 + e, a[i].y2 + n_};
        if (!box_valid(v[i])) { free(v); return -1; }
      }
      int64_t r = par_merge(v, na, nthreads, out);
      free(v);
      return r;
    }
    default: return 0;
  }
}
