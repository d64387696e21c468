"""The headline configurations at full size.

C2: the OP1 ring (src/op_template.m4:14-22) with d = 1 DBU on the 4-layer
standard-cell design, 10M boxes per layer:
    L1 = GROW(L; 1,1,1,1)   L2 = SIZE(L; -1)   outer = L1 ~ L   inner = L ~ L2
    and1 = outer * L        ring = outer * inner
run three ways: operator by operator, as one recorded program (the fused tile
pass), and as replays of the program's captured HIP graph.  Checked:
  - exact Klee areas: area(outer) = area(L1) - area(L), area(inner) =
    area(L) - area(L2) (L2 is inside L, L inside L1), and both ANDs empty;
  - the program's and every graph replay's outputs equal the operator-by-
    operator ones (vl_xor_empty), with other engine work between replays;
  - the oracle on two windows W of two layers: each output clipped to W equals
    the oracle's ring of the boxes meeting W grown by 2 DBU (GROW and SIZE by
    1 DBU see no farther), clipped to W.

C5: the 50M-box serving design (4 layers of 12.5M boxes): a sample of
queries run together equals each run alone, and the oracle on small windows.
"""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu

C2_SEED = 0x5EED0002
C2_PER_LAYER = 10_000_000


def clip(a, w):
    c = np.stack([np.maximum(a[:, 0], w[0]), np.maximum(a[:, 1], w[1]),
                  np.minimum(a[:, 2], w[2]), np.minimum(a[:, 3], w[3])], 1).astype(np.int32)
    return c[(c[:, 0] < c[:, 2]) & (c[:, 1] < c[:, 3])]


def meets(a, w):
    return a[(a[:, 0] < w[2]) & (a[:, 2] > w[0]) & (a[:, 1] < w[3]) & (a[:, 3] > w[1])]


class Ring:
    NAMES = ("L1", "L2", "outer", "inner", "and1", "ring")

    def __init__(self, eng, L):
        self.L = L
        self.out = {k: V.Layer(eng) for k in self.NAMES}

    def run_plain(self):
        o = self.out
        V.CalculateGROWLayer(self.L, o["L1"], (1, 1, 1, 1))
        V.CalculateBoostOperation(self.L, self.L, o["L2"], V.K_BOOST_CONTROL_SIZE, -1)
        V.CalculateFrameNOT(o["L1"], self.L, o["outer"])
        V.CalculateFrameNOT(self.L, o["L2"], o["inner"])
        V.CalculateAND(o["outer"], self.L, o["and1"])
        V.CalculateAND(o["outer"], o["inner"], o["ring"])

    def record(self, prog):
        o = self.out
        prog.add(V.VL_OP_GROW, self.L, None, o["L1"], (1, 1, 1, 1))
        prog.add(V.K_BOOST_CONTROL_SIZE, self.L, None, o["L2"], (-1,))
        prog.add(V.K_BOOST_CONTROL_NOT, o["L1"], self.L, o["outer"])
        prog.add(V.K_BOOST_CONTROL_NOT, self.L, o["L2"], o["inner"])
        prog.add(V.K_BOOST_CONTROL_AND, o["outer"], self.L, o["and1"])
        prog.add(V.K_BOOST_CONTROL_AND, o["outer"], o["inner"], o["ring"])
        prog.add(-1, o["and1"])
        prog.add(-1, o["ring"])


def ring_window(boxes, w):
    """The oracle's ring outputs clipped to w, from the boxes meeting w grown by 2."""
    g = (w[0] - 2, w[1] - 2, w[2] + 2, w[3] + 2)
    near = meets(boxes, g)
    Lw = O.canonical(clip(near, w))
    L1w = O.canonical(clip(O.grow(near, 1, 1, 1, 1), w))
    L2w = O.canonical(clip(O.size(near, -1), w))
    outer = O.op_not(L1w, Lw)
    inner = O.op_not(Lw, L2w)
    return {"L1": L1w, "L2": L2w, "outer": outer, "inner": inner}


@pytest.mark.timeout(900)
def test_c2_op1_ring_full_size(gpu_engine):
    eng = V.Engine(0)
    boxes = [V.synth_layer(V.SYN_STDCELL, C2_SEED, k, C2_PER_LAYER) for k in range(4)]
    inputs = [V.Layer(eng, b) for b in boxes]
    plain = [Ring(eng, L) for L in inputs]
    fused = [Ring(eng, L) for L in inputs]
    for r in plain:
        r.run_plain()
    for k, r in enumerate(plain):
        o = r.out
        aL, a1, a2 = V.klee_area(r.L), V.klee_area(o["L1"]), V.klee_area(o["L2"])
        assert V.klee_area(o["outer"]) == a1 - aL, k
        assert V.klee_area(o["inner"]) == aL - a2, k
        assert a1 > aL > a2 > 0, k
        assert len(o["and1"]) == 0 and len(o["ring"]) == 0, k
        assert len(o["outer"]) > C2_PER_LAYER and len(o["inner"]) > C2_PER_LAYER, k

    # the oracle on two windows of two layers
    for k in (0, 2):
        b = boxes[k]
        m = (int(b[:, 0].min()), int(b[:, 1].min()), int(b[:, 2].max()), int(b[:, 3].max()))
        for fx, fy in ((0.25, 0.6), (0.7, 0.15)):
            x0 = m[0] + int(fx * (m[2] - m[0]))
            y0 = m[1] + int(fy * (m[3] - m[1]))
            w = (x0, y0, x0 + 20_000, y0 + 20_000)
            exp = ring_window(b, w)
            assert len(exp["outer"]) > 100, (k, w)
            for name, e in exp.items():
                got = plain[k].out[name].boxes()
                assert np.array_equal(O.canonical(clip(meets(got, w), w)), e), (k, w, name)
                del got

    # the recorded program (fused tile pass) and its graph replays
    prog = V.Program(eng)
    for r in fused:
        r.record(prog)
    prog.run()
    assert prog.assert_failures() == 0

    def same_as_plain(tag):
        for k in range(4):
            for name in Ring.NAMES:
                assert V.xor_empty(fused[k].out[name], plain[k].out[name]), (tag, k, name)

    same_as_plain("program")
    prog.set_graph(True)
    side = V.Layer(eng, boxes[1][:100_000])
    side_out = V.Layer(eng)
    for rep in range(4):
        prog.run()                       # run 0 records and captures; runs 1.. replay the graph
        assert prog.assert_failures() == 0
        st = prog.graph_stats()
        assert st["captures"] == 1 and st["replays"] == rep and st["fallbacks"] == 0, (rep, st)
        same_as_plain(f"graph run {rep}")
        V.CalculateGROWLayer(side, side_out, (3, 3, 3, 3))        # eager work between replays
        eng.sync()
    prog.close()
    eng.close()


def box_hash(b):
    m64 = (1 << 64) - 1

    def mix64(z):
        z = (z + 0x9E3779B97F4A7C15) & m64
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m64
        return z ^ (z >> 31)

    h = 0
    for x1, y1, x2, y2 in np.asarray(b, np.int64):
        lo = ((int(x1) & 0xFFFFFFFF) << 32) | (int(y1) & 0xFFFFFFFF)
        hi = ((int(x2) & 0xFFFFFFFF) << 32) | (int(y2) & 0xFFFFFFFF)
        h = (h + mix64(lo ^ mix64(hi))) & m64
    return h


@pytest.mark.timeout(900)
def test_c5_serving_full_design(gpu_engine):
    per = 12_500_000
    boxes = [V.synth_layer(V.SYN_STDCELL, 0x5EED0005, k, per) for k in range(4)]
    eng = V.Engine(0)
    layers = [V.Layer(eng, b) for b in boxes]
    srv = V.Server(eng, layers, slots=4, hbm_budget=64 << 30)
    mbr = [(int(b[:, 0].min()), int(b[:, 1].min()), int(b[:, 2].max()), int(b[:, 3].max())) for b in boxes]
    dens = [len(b) / float((m[2] - m[0]) * (m[3] - m[1])) for b, m in zip(boxes, mbr)]
    rng = np.random.default_rng(0xC5)
    n = 24
    q = np.zeros(n, V.QUERY_DTYPE)
    t = 0.0
    for i in range(n):
        la, lb = (int(v) for v in rng.choice(4, 2, replace=False))
        target = float(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        side = max(int(np.sqrt(target / (dens[la] + dens[lb]))), 1)
        m = mbr[la]
        x = int(rng.integers(m[0], max(m[0] + 1, m[2] - side)))
        y = int(rng.integers(m[1], max(m[1] + 1, m[3] - side)))
        q[i] = (x, y, x + side, y + side, la, lb, int(rng.choice([1, 2, 5])), 0, t)
        t += rng.exponential(0.5)
    rc, mix = srv.run(q)
    assert rc == 0 and (mix["status"] == 0).all()
    for i in range(n):
        rc1, alone = srv.run(q[i:i + 1])
        assert rc1 == 0 and alone["count"][0] == mix["count"][i] and alone["hash"][0] == mix["hash"][i], i
    for i in range(0, n, 3):
        w = (int(q["x1"][i]), int(q["y1"][i]), int(q["x2"][i]), int(q["y2"][i]))
        a = clip(meets(boxes[q["layer_a"][i]], w), w)
        b = clip(meets(boxes[q["layer_b"][i]], w), w)
        s = int(q["spacing"][i])
        r = O.op_and(O.grow(a, s, s, s, s), b) if len(a) and len(b) else np.zeros((0, 4), np.int32)
        assert mix["count"][i] == len(r) and mix["hash"][i] == box_hash(r), i
    srv.close()
    eng.close()
