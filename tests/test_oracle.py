"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Checks, per oracle function:
  * golden vectors written by the reference's own heal_boxes / scanline_fracture /
    calculate_union_area_fast (flang build, tests/golden/make_golden.py) on the
    reference's SEDFXTP1 and square KLBIN fixtures and seeded random layers —
    heal output must match element by element, in the reference's order;
  * known answers quoted in the reference (src/klee.cu:214-224,
    src/thrust_boolean.cu:546, src/test_rectboolean.f90, src/test_boxmerge.f90);
  * brute-force rasterisation on small grids for every operator;
  * algebraic identities on larger random layers;
  * library equivalence for the sorts.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from raster import dilate, erode, raster, strips_from_raster

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLD, "golden.json")))


def load(path):
    return np.fromfile(path, dtype=np.int32).reshape(-1, 4)


@pytest.mark.parametrize("name", sorted(META))
def test_heal_matches_reference_golden(name):
    a = load(os.path.join(GOLD, "inputs", name + ".bin"))
    exp = load(os.path.join(GOLD, "expected", name + ".heal.bin"))
    got = O.heal(a)
    assert np.array_equal(got, exp), name          # same boxes, same emission order


@pytest.mark.parametrize("name", sorted(META))
def test_area_matches_reference_golden(name):
    a = load(os.path.join(GOLD, "inputs", name + ".bin"))
    assert O.klee_area(a) == META[name]["area"]
    assert O.klee_area(O.canonical(a)) == META[name]["area"]


@pytest.mark.parametrize("name", [k for k in sorted(META) if META[k]["n"] > 0])
def test_complement_matches_reference_golden(name):
    a = load(os.path.join(GOLD, "inputs", name + ".bin"))
    exp = load(os.path.join(GOLD, "expected", name + ".complement.bin"))
    bb = META[name]["complement_bbox"]
    assert np.array_equal(O.canonical(O.complement(a, bb)), O.canonical(exp))


def test_known_answers():
    assert O.klee_area([[0, 0, 10, 10], [5, 5, 15, 15]]) == 175                    # src/klee.cu:214-224
    assert O.op_and([[0, 0, 10, 10]], [[5, 5, 20, 20]]).tolist() == [[5, 5, 10, 10]]  # thrust_boolean.cu:546
    # (0,0,10,10) minus the hole (2,2,8,8): src/test_rectboolean.f90:12,44 (fragments of
    # src/rectboolean.f90:164-170; the canonical strips of the same point set)
    got = O.op_not([[0, 0, 10, 10]], [[2, 2, 8, 8]])
    assert got.tolist() == [[0, 0, 2, 10], [2, 0, 8, 2], [2, 8, 8, 10], [8, 0, 10, 10]]
    assert O.klee_area(got) == 100 - 36
    # touching boxes merge: src/test_boxmerge.f90:15-17
    assert O.heal([[0, 0, 2, 2], [2, 0, 4, 2]]).tolist() == [[0, 0, 4, 2]]
    # a box inside another: src/test_boxmerge.f90:19-21
    assert O.canonical([[10, 10, 20, 20], [12, 12, 15, 15]]).tolist() == [[10, 10, 20, 20]]
    # L-shape: src/test_boxmerge.f90:23-25
    assert O.canonical([[30, 30, 40, 40], [35, 35, 45, 45]]).tolist() == [
        [30, 30, 35, 40], [35, 30, 40, 45], [40, 35, 45, 45]]
    # corner-only touch does not merge, edge touch of unequal spans does not extend
    assert O.canonical([[0, 0, 1, 1], [1, 1, 2, 2]]).tolist() == [[0, 0, 1, 1], [1, 1, 2, 2]]


def rand_boxes(rng, n, span, wmax, off=0):
    x = rng.integers(off, off + span, n)
    y = rng.integers(off, off + span, n)
    w = rng.integers(1, wmax + 1, n)
    h = rng.integers(1, wmax + 1, n)
    return np.stack([x, y, x + w, y + h], 1).astype(np.int32)


@pytest.mark.parametrize("seed", range(40))
def test_ops_against_raster(seed):
    rng = np.random.default_rng(seed)
    span, wmax = 24, 9
    a = rand_boxes(rng, int(rng.integers(0, 14)), span, wmax)
    b = rand_boxes(rng, int(rng.integers(0, 14)), span, wmax)
    M = 8                              # margin for GROW/SIZE
    ox = oy = -M
    W = H = span + wmax + 2 * M
    ga, gb = raster(a, ox, oy, W, H), raster(b, ox, oy, W, H)
    exp = {
        "and": ga & gb, "not": ga & ~gb, "or": ga | gb, "xor": ga ^ gb,
    }
    for op, g in exp.items():
        got = getattr(O, "op_" + op)(a, b)
        assert np.array_equal(got, strips_from_raster(g, ox, oy)), (op, seed)
    assert np.array_equal(O.canonical(a), strips_from_raster(ga, ox, oy))
    assert O.klee_area(a) == int(ga.sum())
    e, n, w, s = (int(v) for v in rng.integers(0, 4, 4))
    assert np.array_equal(O.grow(a, e, n, w, s), strips_from_raster(dilate(ga, e, n, w, s), ox, oy))
    d = int(rng.integers(1, 4))
    assert np.array_equal(O.size(a, d), strips_from_raster(dilate(ga, d, d, d, d), ox, oy))
    assert np.array_equal(O.size(a, -d), strips_from_raster(erode(ga, d), ox, oy))
    assert np.array_equal(O.size(a, 0), strips_from_raster(ga, ox, oy))


def test_size_single_box_closed_form():
    assert O.size([[0, 0, 10, 10]], -1).tolist() == [[1, 1, 9, 9]]
    assert O.size([[0, 0, 10, 10]], -5).tolist() == []
    assert O.size([[0, 0, 10, 4]], -2).tolist() == []
    assert O.size([[0, 0, 10, 10]], 3).tolist() == [[-3, -3, 13, 13]]
    with pytest.raises(ValueError):
        O.grow([[0, 0, 2, 2]], -2, 0, 0, 0)      # box becomes invalid (geometry.f90:192)


@pytest.mark.parametrize("seed", range(5))
def test_identities_medium(seed):
    rng = np.random.default_rng(100 + seed)
    a = rand_boxes(rng, 3000, 20000, 400, off=-10000)
    b = rand_boxes(rng, 3000, 20000, 400, off=-10000)
    A, B = O.klee_area(a), O.klee_area(b)
    AND, NOT, OR = O.op_and(a, b), O.op_not(a, b), O.op_or(a, b)
    assert O.klee_area(AND) + O.klee_area(NOT) == A
    assert O.klee_area(OR) + O.klee_area(AND) == A + B
    assert len(O.op_and(NOT, b)) == 0
    assert len(O.op_xor(a, a)) == 0
    assert np.array_equal(O.op_xor(a, b), O.op_or(O.op_not(a, b), O.op_not(b, a)))
    # OP1 ring invariants (src/op_template.m4:14-22)
    L = O.canonical(a)
    L1, L2 = O.grow(L, 1, 1, 1, 1), O.size(L, -1)
    outer, inner = O.op_not(L1, L), O.op_not(L, L2)
    assert len(O.op_and(outer, L)) == 0
    assert len(O.op_and(outer, inner)) == 0


@pytest.mark.parametrize("seed", range(10))
def test_oracle_against_reference_build(seed):
    # the reference build is loaded only here, when this test runs (never at
    # collection, so a GPU-only test run does not map it)
    if not os.path.exists(O.REF_SO):
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(1000 + seed)
    a = rand_boxes(rng, int(rng.integers(2, 800)), 3000, 200, off=-1500)
    assert np.array_equal(O.heal(a), O.ref_heal(a))
    assert O.klee_area(a) == O.ref_union_area(a)
    bb = (-2000, -2000, 2000, 2000)
    assert np.array_equal(O.canonical(O.complement(a, bb)), O.canonical(O.ref_complement(a, bb)))


def test_event_sort_is_stable_key_sort():
    rng = np.random.default_rng(7)
    n = 5000
    t = np.zeros(n, dtype=[("X", "<i4"), ("Y", "<i4"), ("w", "i1"), ("pad", "V3")])
    t["X"] = rng.integers(-50, 50, n)
    t["Y"] = rng.integers(-50, 50, n)
    t["w"] = np.arange(n) % 100                      # payload to check stability
    ref = sorted(range(n), key=lambda i: ((int(t["X"][i]) ^ -0x80000000) & 0xFFFFFFFF,
                                          (int(t["Y"][i]) ^ -0x80000000) & 0xFFFFFFFF, i))
    got = O.event_sort(t.copy())
    assert np.array_equal(got, t[ref])


def test_morton_key_and_sorts():
    # centre (0,0) -> cx = cy = 0x80000000: only the top bits of both set
    assert O.morton_key([-1, -1, 1, 1]) == (1 << 63) | (1 << 62)
    # centre (1, 0): bit 0 from x
    assert O.morton_key([0, -1, 2, 1]) == (1 << 63) | (1 << 62) | 1
    # negative odd sums truncate toward zero: (-3 + 0) / 2 = -1
    assert O.morton_key([-3, 0, 0, 0]) & 1 == 1
    rng = np.random.default_rng(3)
    b = rand_boxes(rng, 4000, 1 << 12, 50, off=-(1 << 11))
    keys = np.array([O.morton_key(x) for x in b], dtype=np.uint64)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(O.morton_sort_direct(b), b[order])
    assert np.array_equal(O.morton_sort_indirect(b), order + 1)


@pytest.mark.parametrize("op", ["and", "not", "or", "xor", "merge", "grow", "size"])
def test_cpu_baseline_equals_oracle(op):
    rng = np.random.default_rng(55)
    a = rand_boxes(rng, 6000, 40000, 500)
    b = rand_boxes(rng, 6000, 40000, 500)
    if op in ("and", "not", "or", "xor"):
        got = O.par_op(op, a, b)
        exp = getattr(O, "op_" + op)(a, b)
    elif op == "merge":
        got, exp = O.par_op("merge", a), O.canonical(a)
    elif op == "grow":
        got, exp = O.par_op("grow", a, args=(2, 3, 1, 4)), O.grow(a, 2, 3, 1, 4)
    else:
        got, exp = O.par_op("size", a, args=(-2,)), O.size(a, -2)
    assert np.array_equal(O.canonical(got), exp)


def count_raster(boxes, ox, oy, W, H):
    g = np.zeros((W, H), np.int32)
    for x1, y1, x2, y2 in np.asarray(boxes, np.int64).reshape(-1, 4):
        g[x1 - ox:x2 - ox, y1 - oy:y2 - oy] += 1
    return g


@pytest.mark.parametrize("seed", range(8))
def test_self_and_against_cell_counts(seed):
    """or_self_and = the cells two or more boxes cover (brute force on a grid)."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(1, 60))
    x = rng.integers(0, 40, n)
    y = rng.integers(0, 40, n)
    a = np.stack([x, y, x + rng.integers(1, 12, n), y + rng.integers(1, 12, n)], 1).astype(np.int32)
    if seed % 2:   # exact duplicates count twice
        a = np.concatenate([a, a[: n // 3]])
    g = count_raster(a, 0, 0, 60, 60) >= 2
    assert np.array_equal(O.self_and(a), strips_from_raster(g, 0, 0))


def test_self_and_known_answers():
    assert len(O.self_and([[0, 0, 10, 10]])) == 0
    assert np.array_equal(O.self_and([[0, 0, 10, 10], [0, 0, 10, 10]]), [[0, 0, 10, 10]])
    # touching boxes share no area
    assert len(O.self_and([[0, 0, 10, 10], [10, 0, 20, 10]])) == 0
    assert np.array_equal(O.self_and([[0, 0, 10, 10], [5, 5, 15, 15]]), [[5, 5, 10, 10]])


def test_grid_hand_evaluated_case():
    """CreateGrid (src/design_impl.f90:159-275) on MBR (0,0,100,50), 2 rows,
    3 columns, overlap 10: widths 34/33/33 (100 = 3*33 + 1, the first column
    takes the extra unit), heights 25/25, padded 5 on each side, clamped."""
    got = O.grid((0, 0, 100, 50), 2, 3, 10)
    exp = [[0, 0, 39, 30], [29, 0, 72, 30], [62, 0, 100, 30],
           [0, 20, 39, 50], [29, 20, 72, 50], [62, 20, 100, 50]]
    assert np.array_equal(got, np.array(exp, np.int32))


@pytest.mark.parametrize("rows,cols,overlap", [(1, 1, 0), (10, 10, 10), (7, 3, 0), (4, 9, 3), (10, 10, 1)])
def test_grid_invariants(rows, cols, overlap):
    m = (-37, 11, 1000, 515)
    g = O.grid(m, rows, cols, overlap)
    assert len(g) == rows * cols
    assert (g[:, 0] >= m[0]).all() and (g[:, 2] <= m[2]).all() and (g[:, 1] >= m[1]).all() and (g[:, 3] <= m[3]).all()
    assert O.klee_area(g) == (m[2] - m[0]) * (m[3] - m[1])       # the cells cover the domain
    if overlap == 0:                                              # and tile it exactly
        assert int(((g[:, 2] - g[:, 0]).astype(np.int64) * (g[:, 3] - g[:, 1])).sum()) == O.klee_area(g)


def pnum_brute(a):
    """PerformMerge's result by another route: every pair tested (closed
    extents), components by breadth-first search, numbered in the order of
    their smallest box index; boxes that touch nothing keep 0."""
    a = np.asarray(a, np.int64).reshape(-1, 4)
    n = len(a)
    adj = [[] for _ in range(n)]
    area = perim = 0
    for i in range(n):
        for k in range(i + 1, n):
            ix1, iy1 = max(a[i, 0], a[k, 0]), max(a[i, 1], a[k, 1])
            ix2, iy2 = min(a[i, 2], a[k, 2]), min(a[i, 3], a[k, 3])
            if ix1 <= ix2 and iy1 <= iy2:
                adj[i].append(k)
                adj[k].append(i)
                if ix1 < ix2 and iy1 < iy2:
                    area += int((ix2 - ix1) * (iy2 - iy1))
                else:
                    perim += int(2 * ((ix2 - ix1) + (iy2 - iy1)))
    labels = np.zeros(n, np.int64)
    nc = 0
    for s0 in range(n):
        if labels[s0] or not adj[s0]:
            continue
        nc += 1
        labels[s0] = nc
        todo = [s0]
        while todo:
            u = todo.pop()
            for v in adj[u]:
                if not labels[v]:
                    labels[v] = nc
                    todo.append(v)
    return labels, nc, area, 0 if area else perim


def small_layout(rng, n, span, wmax):
    x = rng.integers(0, span, n)
    y = rng.integers(0, span, n)
    return np.stack([x, y, x + rng.integers(1, wmax + 1, n), y + rng.integers(1, wmax + 1, n)], 1).astype(np.int32)


@pytest.mark.parametrize("seed", range(10))
def test_pnum_against_brute_force(seed):
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(1, 300))
    a = small_layout(rng, n, 200 if seed % 3 else 60, 12)
    if seed % 4 == 1:            # snap to a coarse grid: many shared edges and corners
        a = (a // 4 * 4).astype(np.int32)
        a[:, 2:] = np.maximum(a[:, 2:], a[:, :2] + 4)
    labels, nc, area, perim = O.pnum(a)
    bl, bnc, barea, bperim = pnum_brute(a)
    assert np.array_equal(labels, bl) and nc == bnc and area == barea and perim == bperim
    pos = sum(1 for i in range(n) for k in range(i + 1, n)
              if max(a[i, 0], a[k, 0]) < min(a[i, 2], a[k, 2]) and max(a[i, 1], a[k, 1]) < min(a[i, 3], a[k, 3]))
    assert O.interactions(a) == pos


def test_pnum_known_answers():
    # an edge, a corner, a gap of one unit
    labels, nc, area, perim = O.pnum([[0, 0, 10, 10], [10, 0, 20, 10], [20, 10, 30, 20], [31, 0, 40, 5]])
    assert labels.tolist() == [1, 1, 1, 0] and nc == 1 and area == 0 and perim == 2 * 10 + 0
    # overlap: the area counts and the touch perimeter is reported as 0
    labels, nc, area, perim = O.pnum([[0, 0, 10, 10], [5, 5, 15, 15], [15, 0, 20, 5], [100, 100, 110, 110]])
    assert labels.tolist() == [1, 1, 1, 0] and nc == 1 and area == 25 and perim == 0
    # components numbered by their smallest index, not by position
    labels, nc, _, _ = O.pnum([[100, 0, 110, 10], [0, 0, 10, 10], [105, 5, 120, 20], [5, 5, 12, 12]])
    assert labels.tolist() == [1, 2, 1, 2] and nc == 2
    assert O.pnum([[0, 0, 1, 1]])[0].tolist() == [0]
    assert O.interactions([[0, 0, 10, 10], [10, 0, 20, 10], [5, 5, 15, 15]]) == 2


@pytest.mark.parametrize("seed", range(6))
def test_pnum_against_reference_build(seed):
    """Labels, area and the touch perimeter equal the reference's own
    BuildRTree + PerformMerge, and the interaction count its
    ComputeInteractionsGPU (flang build, oracle/_ref)."""
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(700 + seed)
    n = int(rng.integers(1, 3000))
    a = small_layout(rng, n, 4000 if seed % 2 else 1500, 40)
    if seed == 5:
        a = (a // 8 * 8).astype(np.int32)
        a[:, 2:] = np.maximum(a[:, 2:], a[:, :2] + 8)
    labels, nc, area, perim = O.pnum(a)
    rl, rarea, rperim = O.ref_pnum(a)
    assert np.array_equal(labels, rl)
    assert nc == int(rl.max(initial=0))
    assert area == rarea and perim == rperim
    assert O.interactions(a) == O.ref_interactions(a)


# (0,0,10,10): square, no move, new width 10           -> 13 0a
# (5,0,8,4): x moved +5 (zigzag 10), new w 3, new h 4  -> 2f 0a 03 04
# (5,-1,15,9): y moved -1 (zigzag 1), square, w 10 is
#              entry 1 (not the current entry 2)        -> 51 01 01 00
# (7,-1,17,5): x moved +2 (zigzag 4), w 10 current,
#              new h 6                                  -> 2c 04 06
CODEC_KNOWN = ([[0, 0, 10, 10], [5, 0, 8, 4], [5, -1, 15, 9], [7, -1, 17, 5]],
               bytes([0x13, 0x0A, 0x2F, 0x0A, 0x03, 0x04, 0x51, 0x01, 0x01, 0x00, 0x2C, 0x04, 0x06]))


def test_codec_known_answer():
    b, data = CODEC_KNOWN
    assert O.codec_encode(b) == data
    assert O.codec_decode(data, 4).tolist() == b
    assert O.codec_decode(data[:-1], 4) is None          # truncated
    assert O.codec_decode(data + b"\x00", 4) is None    # trailing byte
    # an index past the dictionary is rejected
    assert O.codec_decode(bytes([0x13, 0x0A, 0x21, 0x02, 0x05, 0x00]), 2) is None


@pytest.mark.parametrize("seed", range(5))
def test_codec_round_trip(seed):
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(1, 5000))
    if seed == 4:    # extreme coordinates: deltas and widths wrap int32
        x = rng.integers(-2**31, 2**31 - 1, n)
        w = rng.integers(1, 2**31 - 1, n)
        x1 = np.minimum(x, 2**31 - 1 - w)
        a = np.stack([x1, x1 // 3, x1 + w, x1 // 3 + 1], 1).astype(np.int32)
    else:            # a few widths/heights (dictionary hits), squares, moves of every sign
        x = rng.integers(-5000, 5000, n)
        y = rng.integers(-5000, 5000, n)
        w = rng.choice([1, 2, 5, 8, 40, 333], n)
        h = np.where(rng.random(n) < 0.3, w, rng.choice([1, 3, 8, 600], n))
        a = np.stack([x, y, x + w, y + h], 1).astype(np.int32)
        a[rng.random(n) < 0.2] = a[0]          # repeats: zero deltas
    data = O.codec_encode(a)
    assert np.array_equal(O.codec_decode(data, n), a)

