"""Query server (vl_server_*): windowed DRC-style queries
R = GROW_s(A & W) AND (B & W) under continuous batching.

Every result must be the same alone and in the mix (count and the
order-independent hash of the canonical result), equal to the oracle's, and an
HBM budget too small for two queries must serialise them."""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def box_hash(b):
    h = 0
    for x1, y1, x2, y2 in np.asarray(b, np.int64):
        lo = ((int(x1) & 0xFFFFFFFF) << 32) | (int(y1) & 0xFFFFFFFF)
        hi = ((int(x2) & 0xFFFFFFFF) << 32) | (int(y2) & 0xFFFFFFFF)
        h = (h + mix64(lo ^ mix64(hi))) & M64
    return h


def clip(a, w):
    c = np.stack([np.maximum(a[:, 0], w[0]), np.maximum(a[:, 1], w[1]),
                  np.minimum(a[:, 2], w[2]), np.minimum(a[:, 3], w[3])], 1).astype(np.int32)
    return c[(c[:, 0] < c[:, 2]) & (c[:, 1] < c[:, 3])]


def expected(layers, q):
    w = (q["x1"], q["y1"], q["x2"], q["y2"])
    a = clip(layers[q["layer_a"]], w)
    b = clip(layers[q["layer_b"]], w)
    s = int(q["spacing"])
    r = O.op_and(O.grow(a, s, s, s, s), b) if len(a) and len(b) else np.zeros((0, 4), np.int32)
    return len(r), box_hash(r)


def make_queries(rng, layers, n, mean_gap_ms):
    allb = np.concatenate(layers)
    X0, Y0, X1, Y1 = allb[:, 0].min(), allb[:, 1].min(), allb[:, 2].max(), allb[:, 3].max()
    q = np.zeros(n, V.QUERY_DTYPE)
    t = 0.0
    for i in range(n):
        side_x = int(rng.uniform(0.02, 0.5) * (X1 - X0))
        side_y = int(rng.uniform(0.02, 0.5) * (Y1 - Y0))
        x = int(rng.integers(X0 - side_x // 4, X1 - side_x // 2))
        y = int(rng.integers(Y0 - side_y // 4, Y1 - side_y // 2))
        q[i] = (x, y, x + max(side_x, 1), y + max(side_y, 1), int(rng.integers(0, len(layers))),
                int(rng.integers(0, len(layers))), int(rng.choice([1, 2, 5])), 0, t)
        t += rng.exponential(mean_gap_ms)
    return q


@pytest.fixture(scope="module")
def design(gpu_engine):
    eng = V.Engine(0)
    boxes = [V.synth_layer(V.SYN_STDCELL, 0x5EED0005, k, 40000) for k in range(4)]
    layers = [V.Layer(eng, b) for b in boxes]
    yield eng, boxes, layers
    eng.close()


def test_mix_equals_alone_and_oracle(design):
    eng, boxes, layers = design
    rng = np.random.default_rng(5)
    q = make_queries(rng, boxes, 40, mean_gap_ms=0.5)
    srv = V.Server(eng, layers, slots=4)
    rc, mix = srv.run(q)
    assert rc == 0 and (mix["status"] == 0).all()
    assert len(set(mix["slot"].tolist())) > 1                # work did spread over slots
    for i in range(len(q)):
        rc1, alone = srv.run(q[i:i + 1])
        assert rc1 == 0
        assert alone["count"][0] == mix["count"][i] and alone["hash"][0] == mix["hash"][i]
    for i in range(0, len(q), 4):                            # oracle on a subset
        cnt, h = expected(boxes, q[i])
        assert mix["count"][i] == cnt and mix["hash"][i] == h, i
    srv.close()


def test_budget_serialises_admission(design):
    eng, boxes, layers = design
    rng = np.random.default_rng(6)
    q = make_queries(rng, boxes, 12, mean_gap_ms=0.0)
    big = V.Server(eng, layers, slots=4)
    rc, ref = big.run(q)
    assert rc == 0
    big.close()
    tiny = V.Server(eng, layers, slots=4, hbm_budget=1)      # never two in flight
    rc, res = tiny.run(q)
    assert rc == 0
    assert np.array_equal(res["count"], ref["count"]) and np.array_equal(res["hash"], ref["hash"])
    iv = sorted(zip(res["start_ms"], res["end_ms"]))
    for (s0, e0), (s1, e1) in zip(iv, iv[1:]):
        assert s1 >= e0 - 1e-6
    tiny.close()


def test_bad_query_reports_status(design):
    eng, boxes, layers = design
    q = np.zeros(3, V.QUERY_DTYPE)
    q[0] = (0, 0, 20000, 20000, 0, 1, 1, 0, 0.0)
    q[1] = (10, 10, 5, 50, 0, 1, 1, 0, 0.0)                  # x1 > x2
    q[2] = (0, 0, 20000, 20000, 2, 9, 1, 0, 0.0)             # layer out of range
    srv = V.Server(eng, layers, slots=2)
    rc, res = srv.run(q)
    assert rc == V.VL_EINVAL
    assert res["status"].tolist() == [0, V.VL_EINVAL, V.VL_EINVAL]
    assert (res["count"][0], res["hash"][0]) == expected(boxes, q[0])
    srv.close()


@pytest.fixture(scope="module")
def route_design(gpu_engine):
    eng = V.Engine(0)
    boxes = [V.synth_layer(V.SYN_ROUTE, 0x5EED0006, k, 60000) for k in range(4)]
    layers = [V.Layer(eng, b) for b in boxes]
    yield eng, boxes, layers
    eng.close()


def test_spacing_queries_on_routing_design(route_design):
    """DRC spacing queries between the two layers of one metal (neighbouring
    tracks 6 DBU apart, a few wires jogged to 4 DBU): every result equals the
    oracle's, in_boxes counts each box with area inside the window once (the
    index keeps one copy of a box that spans several cells), checks at 1 and
    2 DBU always pass, and some at 5 DBU fail (the jogs)."""
    eng, boxes, layers = route_design
    rng = np.random.default_rng(11)
    allb = np.concatenate(boxes)
    X0, Y0, X1, Y1 = allb[:, 0].min(), allb[:, 1].min(), allb[:, 2].max(), allb[:, 3].max()
    n = 48
    q = np.zeros(n, V.QUERY_DTYPE)
    for i in range(n):
        sx = int(rng.uniform(0.05, 0.6) * (X1 - X0))
        sy = int(rng.uniform(0.05, 0.6) * (Y1 - Y0))
        x = int(rng.integers(X0 - sx // 4, X1 - sx // 2))
        y = int(rng.integers(Y0 - sy // 4, Y1 - sy // 2))
        la = int(rng.integers(0, 4))
        q[i] = (x, y, x + sx, y + sy, la, la ^ 1, (1, 2, 5)[i % 3], 0, 0.3 * i)
    srv = V.Server(eng, layers, slots=4)
    rc, res = srv.run(q)
    assert rc == 0 and (res["status"] == 0).all()
    for i in range(n):
        w = (q["x1"][i], q["y1"][i], q["x2"][i], q["y2"][i])
        assert res["in_boxes"][i] == len(clip(boxes[q["layer_a"][i]], w)) + len(clip(boxes[q["layer_b"][i]], w)), i
        cnt, h = expected(boxes, q[i])
        assert res["count"][i] == cnt and res["hash"][i] == h, i
    assert (res["count"][q["spacing"] < 5] == 0).all()
    assert (res["count"][q["spacing"] == 5] > 0).any()
    srv.close()
