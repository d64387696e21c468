"""Programs captured as HIP graphs (vl_program_set_graph): the first run
records the sizes, captures the work and launches the graph; later runs
replay it.  Every replay must give exactly the plain one-call results; a
changed input layer, or an output written by another call, makes the next
run capture again.  Replays are interleaved with other eager work and host
synchronisations on the same engine."""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu


def op1_record(prog, L, t):
    L1, L2, outer, inner, and1, ring = t
    prog.add(V.VL_OP_GROW, L, None, L1, (1, 1, 1, 1))
    prog.add(V.K_BOOST_CONTROL_SIZE, L, None, L2, (-1,))
    prog.add(V.K_BOOST_CONTROL_NOT, L1, L, outer)
    prog.add(V.K_BOOST_CONTROL_NOT, L, L2, inner)
    prog.add(V.K_BOOST_CONTROL_AND, outer, L, and1)
    prog.add(V.K_BOOST_CONTROL_AND, outer, inner, ring)
    prog.add(-1, and1)
    prog.add(-1, ring)


def op1_expected(a):
    La = O.canonical(a)
    e1, e2 = O.grow(a, 1, 1, 1, 1), O.size(a, -1)
    return [e1, e2, O.op_not(e1, La), O.op_not(La, e2), np.zeros((0, 4), np.int32), np.zeros((0, 4), np.int32)]


def check(temps, exps, tag):
    for k, (g, x) in enumerate(zip(temps, exps)):
        assert np.array_equal(O.sort_boxes(g.boxes()), x), (tag, k)


@pytest.mark.parametrize("fuse", [1, 0])
def test_graph_replays_equal_the_eager_results(gpu_engine, fuse):
    eng = V.Engine(0)
    eng.set("fuse", fuse)
    data = [V.synth_layer(V.SYN_STDCELL, 0x5EED0002, k, 15000 + 4000 * k) for k in range(3)]
    layers = [V.Layer(eng, d) for d in data]
    temps = [[V.Layer(eng) for _ in range(6)] for _ in data]
    prog = V.Program(eng)
    for L, t in zip(layers, temps):
        op1_record(prog, L, t)
    prog.set_graph(True)
    exps = [op1_expected(d) for d in data]
    other_a, other_b, other_c = V.Layer(eng, data[0][:3000]), V.Layer(eng, data[1][:3000]), V.Layer(eng)
    for rep in range(5):
        prog.run()
        assert prog.assert_failures() == 0
        for t, x in zip(temps, exps):
            check(t, x, (fuse, rep))
        # eager work and a synchronisation between replays
        V.CalculateOR(other_a, other_b, other_c)
        eng.sync()
        assert np.array_equal(O.sort_boxes(other_c.boxes()), O.op_or(data[0][:3000], data[1][:3000]))
    st = prog.graph_stats()
    assert st == {"captures": 1, "replays": 4, "fallbacks": 0}, st
    prog.close()
    eng.close()


def test_graph_recaptures_after_an_input_or_output_changes(gpu_engine):
    eng = V.Engine(0)
    a0 = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 0, 20000)
    a1 = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 1, 25000)
    L = V.Layer(eng, a0)
    t = [V.Layer(eng) for _ in range(6)]
    prog = V.Program(eng)
    op1_record(prog, L, t)
    prog.set_graph(True)
    prog.run()
    prog.run()
    check(t, op1_expected(a0), "a0")
    L.upload(a1)                       # new input: the next run captures again
    prog.run()
    check(t, op1_expected(a1), "a1")
    prog.run()
    check(t, op1_expected(a1), "a1 replay")
    V.MergeHealLayer(t[2])             # an output written by another call
    prog.run()
    check(t, op1_expected(a1), "after outside write")
    st = prog.graph_stats()
    assert st["captures"] == 3 and st["replays"] == 2 and st["fallbacks"] == 0, st
    prog.set_graph(False)              # back to eager runs
    prog.run()
    check(t, op1_expected(a1), "eager")
    prog.close()
    eng.close()


def test_graph_with_dense_tiles_and_size_composition(gpu_engine):
    """A captured program whose operators take the dense-tile kernel, tile
    splits and the SIZE(-d) complement composition (host-computed uploads
    inside the graph)."""
    eng = V.Engine(0)
    rng = np.random.default_rng(3)
    i = np.arange(300)
    bus = np.stack([400_000 + (i * 37) % 997, 500_000 + 8 * i, 450_000 + (i * 37) % 997, 500_003 + 8 * i], 1)
    x = rng.integers(0, 1_000_000, (2000, 2))
    back = np.concatenate([x, x + rng.integers(1, 2000, (2000, 2))], 1)
    a = np.concatenate([back, bus]).astype(np.int32)
    b = (a[::2] + np.array([3, 5, 3, 5])).astype(np.int32)
    A, B = V.Layer(eng, a), V.Layer(eng, b)
    outs = [V.Layer(eng) for _ in range(5)]
    prog = V.Program(eng)
    prog.add(V.K_BOOST_CONTROL_SIZE, A, None, outs[0], (-2,))
    prog.add(V.K_BOOST_CONTROL_NOT, A, B, outs[1])
    prog.add(V.K_BOOST_CONTROL_XOR, outs[0], B, outs[2])
    prog.add(V.VL_OP_GROW, B, None, outs[3], (2, 0, 7, 1))
    prog.add(V.K_BOOST_CONTROL_AND, outs[3], A, outs[4])
    prog.set_graph(True)
    e0 = O.size(a, -2)
    e3 = O.grow(b, 2, 0, 7, 1)
    exps = [e0, O.op_not(a, b), O.op_xor(e0, b), e3, O.op_and(e3, a)]
    for rep in range(3):
        prog.run()
        check(outs, exps, rep)
    st = prog.graph_stats()
    assert st["captures"] == 1 and st["replays"] == 2, st
    prog.close()
    eng.close()
