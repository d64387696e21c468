"""Handle lifetimes: layers, programs and servers keep their engine alive, and
programs keep the layers they record alive, whatever order the caller drops
or destroys them in."""
import gc

import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu


def test_engine_destroyed_before_its_layers_and_programs(gpu_engine):
    eng = V.Engine(0)
    a = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 0, 3000)
    L, out = V.Layer(eng, a), V.Layer(eng)
    prog = V.Program(eng)
    prog.add(V.VL_OP_GROW, L, None, out, (1, 1, 1, 1))
    prog.run()
    assert np.array_equal(O.sort_boxes(out.boxes()), O.grow(a, 1, 1, 1, 1))
    eng.close()          # the handle goes first; the engine lives on for its layers and program
    prog.run()           # still usable
    assert np.array_equal(O.sort_boxes(out.boxes()), O.grow(a, 1, 1, 1, 1))
    L.close()
    prog.close()
    out.close()


def test_program_keeps_dropped_layers_alive(gpu_engine):
    """Layers recorded by a program and destroyed by the caller before run()
    stay valid: the program holds a reference to each."""
    eng = V.Engine(0)
    rng = np.random.default_rng(5)
    a = rng.integers(0, 5000, (800, 2))
    a = np.concatenate([a, a + rng.integers(1, 300, (800, 2))], 1).astype(np.int32)
    b = rng.integers(0, 5000, (800, 2))
    b = np.concatenate([b, b + rng.integers(1, 300, (800, 2))], 1).astype(np.int32)
    out = V.Layer(eng)
    prog = V.Program(eng)
    A, B = V.Layer(eng, a), V.Layer(eng, b)
    prog.add(V.K_BOOST_CONTROL_NOT, A, B, out)
    A.close()          # the handles go; the program's own references keep the layers
    B.close()
    del A, B
    gc.collect()
    for _ in range(2):
        prog.run()
        assert np.array_equal(O.sort_boxes(out.boxes()), O.op_not(a, b))
    prog.close()
    eng.close()
