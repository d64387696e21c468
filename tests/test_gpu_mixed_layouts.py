"""Seeded layouts of mixed families against the oracle: wires of both
orientations, buses of parallel wires in a sparse layer, dense clusters,
standard-cell and routing-track layers, on engines with tiles shaped like
the boxes or square, and several fused, check and operator tile budgets.

Each case checks AND, NOT, OR and XOR of two such layers, GROW and SIZE of
one, and, as one program run eagerly and as a replayed HIP graph, the OP1
ring (src/op_template.m4:14-22) of both layers with an AND of the two read
only by ASSERT_ZERO (the emptiness-test path), whose failure count must say
whether the oracle's AND is empty.  Results are compared with the oracle as
canonical strip sets (sorted)."""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu

KINDS = ["rand", "hwires", "vwires", "mixwires", "cluster", "bus", "stdcell", "route"]


def layout(rng, kind, n):
    span = int(rng.integers(200, 20000))
    if kind == "rand":
        wmax = int(rng.integers(1, 400))
        x, y = rng.integers(0, span, n), rng.integers(0, span, n)
        w, h = rng.integers(1, wmax + 1, n), rng.integers(1, wmax + 1, n)
    elif kind == "hwires":
        x, y = rng.integers(0, span, n), rng.integers(0, span, n)
        w, h = rng.integers(20, 2000, n), rng.integers(1, 6, n)
    elif kind == "vwires":
        x, y = rng.integers(0, span, n), rng.integers(0, span, n)
        w, h = rng.integers(1, 6, n), rng.integers(20, 2000, n)
    elif kind == "mixwires":
        return np.concatenate([layout(rng, "hwires", n // 2), layout(rng, "vwires", n - n // 2)])
    elif kind == "cluster":
        cx, cy = rng.integers(0, span, 2)
        r = int(rng.integers(5, 60))
        x, y = cx + rng.integers(-r, r, n), cy + rng.integers(-r, r, n)
        w, h = rng.integers(1, 3 * r, n), rng.integers(1, 3 * r, n)
    elif kind == "bus":
        k = min(n, int(rng.integers(50, 600)))
        pitch = int(rng.integers(2, 8))
        x0, length = int(rng.integers(0, span)), int(rng.integers(50, 3000))
        rows = np.arange(k) * pitch
        bus = np.stack([np.full(k, x0), rows, np.full(k, x0 + length), rows + pitch // 2 + 1], 1)
        return np.concatenate([bus.astype(np.int32), layout(rng, "rand", max(1, n - k))])
    elif kind == "stdcell":
        return V.synth_layer(V.SYN_STDCELL, int(rng.integers(1, 1 << 30)), int(rng.integers(0, 4)), n)
    else:
        return V.synth_layer(V.SYN_ROUTE, int(rng.integers(1, 1 << 30)), int(rng.integers(0, 4)), n)
    return np.stack([x, y, x + w, y + h], 1).astype(np.int32)


def same(got, want):
    return np.array_equal(O.sort_boxes(got), O.sort_boxes(want))


def ring_program(eng, layers, d):
    prog = V.Program(eng)
    outs = []
    for L in layers:
        t = [V.Layer(eng) for _ in range(6)]
        L1, L2, outer, inner, and1, ring = t
        prog.add(V.VL_OP_GROW, L, None, L1, (d, d, d, d))
        prog.add(V.K_BOOST_CONTROL_SIZE, L, None, L2, (-d,))
        prog.add(V.K_BOOST_CONTROL_NOT, L1, L, outer)
        prog.add(V.K_BOOST_CONTROL_NOT, L, L2, inner)
        prog.add(V.K_BOOST_CONTROL_AND, outer, L, and1)
        prog.add(V.K_BOOST_CONTROL_AND, outer, inner, ring)
        prog.add(-1, and1)
        prog.add(-1, ring)
        outs.append(t)
    return prog, outs


@pytest.mark.parametrize("seed", [11, 23, 37, 41, 53, 67, 79, 97])
def test_mixed_layouts_match_oracle(gpu_engine, seed):
    rng = np.random.default_rng(seed)
    ka, kb = KINDS[seed % len(KINDS)], KINDS[int(rng.integers(len(KINDS)))]
    a = layout(rng, ka, int(rng.integers(2000, 15000)))
    b = layout(rng, kb, int(rng.integers(2000, 15000)))
    eng = V.Engine(0)
    params = {"tile_shape": [0, 4, 32][seed % 3], "fuse_target": [3, 6, 9][seed % 3 - 1],
              "check_target": int(rng.choice([2, 6, 12])), "tile_target": int(rng.choice([6, 20, 40])),
              "erode_target": int(rng.choice([6, 12]))}
    for k, v in params.items():
        eng.set(k, v)
    case = f"seed {seed}: {ka}({len(a)}) x {kb}({len(b)}), {params}"
    print(case, flush=True)
    A, B = V.Layer(eng, a), V.Layer(eng, b)
    for name, fn, ofn in (("AND", V.CalculateAND, O.op_and), ("NOT", V.CalculateFrameNOT, O.op_not),
                          ("OR", V.CalculateOR, O.op_or), ("XOR", V.CalculateXOR, O.op_xor)):
        c = V.Layer(eng)
        fn(A, B, c)
        assert same(c.boxes(), ofn(a, b)), f"{name}, {case}"
    g = tuple(int(v) for v in rng.integers(0, 4, 4))
    c = V.Layer(eng)
    V.CalculateGROWLayer(A, c, g)
    assert same(c.boxes(), O.grow(a, *g)), f"GROW{g}, {case}"
    d = int(rng.choice([-3, -2, -1, 1, 2]))
    c = V.Layer(eng)
    V.CalculateBoostOperation(A, A, c, V.K_BOOST_CONTROL_SIZE, d)
    assert same(c.boxes(), O.size(a, d)), f"SIZE({d}), {case}"

    dd = int(rng.choice([1, 2]))
    want_fail = int(len(O.op_and(a, b)) > 0)
    want = []
    for src in (a, b):
        o1, o2 = O.grow(src, dd, dd, dd, dd), O.size(src, -dd)
        want.append([o1, o2, O.op_not(o1, src), O.op_not(src, o2)])
    for graph in (False, True):
        prog, outs = ring_program(eng, (A, B), dd)
        X = V.Layer(eng)
        prog.add(V.K_BOOST_CONTROL_AND, A, B, X)
        prog.add(-1, X)
        if graph:
            prog.set_graph(True)
        for _ in range(2 if graph else 1):   # capture, then a replay
            prog.run()
            # ASSERT_ZERO failures of the run: the two ring checks never
            # fail, the AND of the two layers fails when it is not empty
            assert prog.assert_failures() == want_fail, f"ASSERT_ZERO count, graph={graph}, {case}"
        if graph:
            assert prog.graph_stats()["replays"] == 1, case
        for t, w in zip(outs, want):
            for lname, lay, wb in zip(("L1", "L2", "outer", "inner"), t[:4], w):
                assert same(lay.boxes(), wb), f"ring {lname}, graph={graph}, {case}"
        prog.close()
