"""Randomized parity sweep on a GPU (run by hand; pytest does not collect it):

    python tests/fuzz_parity.py SECONDS SEED MAX_BOXES

Draws cases until SECONDS have passed: two layers of random families (the
generators of test_gpu_mixed_layouts: wires of both orientations, buses,
clusters, standard-cell and routing layers) of up to MAX_BOXES boxes each,
random tile, fused, check and erosion budgets, then the same checks as that
test (AND, NOT, OR, XOR, GROW, SIZE, and the ring program with a checked AND,
eager and as a replayed graph) against the oracle.  Prints each case's seed
before running it; exits 1 if any case differs."""
import os
import sys
import time

import numpy as np

TESTS = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(TESTS))
sys.path.insert(0, TESTS)
import oracle_lib as O  # noqa: E402
import vlsi_layout_amd as V  # noqa: E402
from test_gpu_mixed_layouts import KINDS, layout, ring_program, same  # noqa: E402


def run_case(eng, s, nmax):
    r = np.random.default_rng(s)
    ka, kb = KINDS[r.integers(len(KINDS))], KINDS[r.integers(len(KINDS))]
    a = layout(r, ka, int(r.integers(50, nmax)))
    b = layout(r, kb, int(r.integers(50, nmax)))
    params = {"tile_shape": int(r.choice([0, 4, 32])), "fuse_target": int(r.choice([3, 6, 9])),
              "check_target": int(r.choice([2, 6, 12])), "tile_target": int(r.choice([6, 20, 40])),
              "erode_target": int(r.choice([6, 12]))}
    for k, v in params.items():
        eng.set(k, v)
    print(f"seed {s}: {ka}({len(a)}) x {kb}({len(b)}), {params}", flush=True)
    A, B = V.Layer(eng, a), V.Layer(eng, b)
    bad = []
    for name, fn, ofn in (("AND", V.CalculateAND, O.op_and), ("NOT", V.CalculateFrameNOT, O.op_not),
                          ("OR", V.CalculateOR, O.op_or), ("XOR", V.CalculateXOR, O.op_xor)):
        c = V.Layer(eng)
        fn(A, B, c)
        if not same(c.boxes(), ofn(a, b)):
            bad.append(name)
    g = tuple(int(v) for v in r.integers(0, 4, 4))
    c = V.Layer(eng)
    V.CalculateGROWLayer(A, c, g)
    if not same(c.boxes(), O.grow(a, *g)):
        bad.append(f"GROW{g}")
    d = int(r.choice([-3, -2, -1, 1, 2]))
    c = V.Layer(eng)
    V.CalculateBoostOperation(A, A, c, V.K_BOOST_CONTROL_SIZE, d)
    if not same(c.boxes(), O.size(a, d)):
        bad.append(f"SIZE({d})")
    dd = int(r.choice([1, 2]))
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
        for _ in range(2 if graph else 1):
            prog.run()
            if prog.assert_failures() != want_fail:
                bad.append(f"ASSERT_ZERO count, graph={graph}")
        for t, w in zip(outs, want):
            for lname, lay, wb in zip(("L1", "L2", "outer", "inner"), t[:4], w):
                if not same(lay.boxes(), wb):
                    bad.append(f"ring {lname}, graph={graph}")
        prog.close()
    return bad


def main():
    seconds = float(sys.argv[1]) if len(sys.argv) > 1 else 240.0
    seed = int(sys.argv[2]) if len(sys.argv) > 2 else 1
    nmax = int(sys.argv[3]) if len(sys.argv) > 3 else 30000
    end = time.time() + seconds
    rng = np.random.default_rng(seed)
    eng = V.Engine(0)
    cases = fails = 0
    while time.time() < end:
        cases += 1
        bad = run_case(eng, int(rng.integers(1, 1 << 31)), nmax)
        if bad:
            fails += 1
            print("  MISMATCH:", bad, flush=True)
    print(f"done: {cases} cases, {fails} with mismatches", flush=True)
    return 1 if fails else 0


if __name__ == "__main__":
    sys.exit(main())
