"""Tiles that no capacity-bounded tile kernel holds: more boxes over one
x-position than the workgroup kernel's 255, which no x-split separates.
They run on the dense kernel (vl_dense.hip); SIZE(-d) falls back to the
complement composition; a fused group declines and runs operator by operator.
Every operator is checked against the oracle (bit-exact canonical strips,
exact area)."""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu


def sparse(rng, n, span, wmax):
    x = rng.integers(0, span, n)
    y = rng.integers(0, span, n)
    w = rng.integers(1, wmax + 1, n)
    h = rng.integers(1, wmax + 1, n)
    return np.stack([x, y, x + w, y + h], 1).astype(np.int32)


def bus(nwires, x0=400_000, y0=500_000, length=50_000, pitch=8, width=3, stagger=0):
    """nwires horizontal wires; stagger > 0 shifts each wire's ends."""
    i = np.arange(nwires)
    x1 = x0 + (i * stagger) % 997
    y = y0 + pitch * i
    return np.stack([x1, y, x1 + length, y + width], 1).astype(np.int32)


def via_array(k, pitch=10, size=6, x0=300_000, y0=300_000):
    gx, gy = np.meshgrid(np.arange(k), np.arange(k))
    x = x0 + pitch * gx.ravel()
    y = y0 + pitch * gy.ravel()
    return np.stack([x, y, x + size, y + size], 1).astype(np.int32)


def layouts():
    rng = np.random.default_rng(77)
    back = sparse(rng, 2000, 1_000_000, 2000)
    ident = np.tile(np.array([[600_000, 600_000, 600_040, 600_025]], np.int32), (300, 1))
    return {
        "bus300": np.concatenate([back, bus(300)]),
        "bus300_staggered": np.concatenate([back, bus(300, stagger=37)]),
        "identical300": np.concatenate([back, ident]),
        "vias40": np.concatenate([back, via_array(40)]),
    }


LAYOUTS = layouts()


def check(got, exp, tag):
    assert np.array_equal(O.sort_boxes(got.boxes()), exp), tag


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_dense_unary_operators(gpu_engine, name):
    eng = V.Engine(0)
    a = LAYOUTS[name]
    A = V.Layer(eng, a)
    out = V.Layer(eng)
    V.CalculateBoostOperation(A, A, out, V.K_BOOST_CONTROL_MERGE)
    check(out, O.canonical(a), (name, "merge"))
    V.CalculateGROWLayer(A, out, (1, 2, 3, 4))
    check(out, O.grow(a, 1, 2, 3, 4), (name, "grow"))
    V.CalculateGROWLayer(A, out, (40, 40, 40, 40))   # grown far beyond the via pitch
    check(out, O.grow(a, 40, 40, 40, 40), (name, "grow40"))
    for d in (2, -1, -2):
        V.CalculateBoostOperation(A, A, out, V.K_BOOST_CONTROL_SIZE, d)
        check(out, O.size(a, d), (name, "size", d))
    assert V.klee_area(A) == O.klee_area(a)
    eng.close()


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_dense_binary_operators(gpu_engine, name):
    eng = V.Engine(0)
    a = LAYOUTS[name]
    rng = np.random.default_rng(len(name))
    # B overlaps the dense part: the same structure shifted, plus its own background
    b = np.concatenate([a[2000:] + np.array([3, 5, 3, 5], np.int32), sparse(rng, 1500, 1_000_000, 3000)])
    A, B, out = V.Layer(eng, a), V.Layer(eng, b), V.Layer(eng)
    for code, fn in ((V.K_BOOST_CONTROL_AND, O.op_and), (V.K_BOOST_CONTROL_NOT, O.op_not),
                     (V.K_BOOST_CONTROL_OR, O.op_or), (V.K_BOOST_CONTROL_XOR, O.op_xor)):
        V.CalculateBoostOperation(A, B, out, code)
        check(out, fn(a, b), (name, code))
    assert V.xor_empty(A, A)
    assert not V.xor_empty(A, B)
    eng.close()


@pytest.mark.parametrize("target", [4, 20, 200])
def test_dense_tiles_under_other_grids(gpu_engine, target):
    """Many dense tiles at once (a small target makes every bus tile narrow)."""
    eng = V.Engine(0)
    eng.set("tile_target", target)
    a = np.concatenate([bus(280, x0=1000, y0=1000, length=30_000, stagger=13),
                        bus(290, x0=20_000, y0=900, length=25_000, pitch=7, width=5)])
    A, out = V.Layer(eng, a), V.Layer(eng)
    V.CalculateBoostOperation(A, A, out, V.K_BOOST_CONTROL_MERGE)
    check(out, O.canonical(a), target)
    V.CalculateGROWLayer(A, out, (2, 2, 2, 2))
    check(out, O.grow(a, 2, 2, 2, 2), target)
    assert V.klee_area(A) == O.klee_area(a)
    eng.close()


def test_dense_fused_group_declines_to_plain(gpu_engine):
    """The OP1 ring over a layer with a 300-wire bus as one program."""
    eng = V.Engine(0)
    a = LAYOUTS["bus300"]
    L = V.Layer(eng, a)
    t = [V.Layer(eng) for _ in range(6)]
    L1, L2, outer, inner, and1, ring = t
    prog = V.Program(eng)
    prog.add(V.VL_OP_GROW, L, None, L1, (1, 1, 1, 1))
    prog.add(V.K_BOOST_CONTROL_SIZE, L, None, L2, (-1,))
    prog.add(V.K_BOOST_CONTROL_NOT, L1, L, outer)
    prog.add(V.K_BOOST_CONTROL_NOT, L, L2, inner)
    prog.add(V.K_BOOST_CONTROL_AND, outer, L, and1)
    prog.add(V.K_BOOST_CONTROL_AND, outer, inner, ring)
    prog.add(-1, and1)
    prog.add(-1, ring)
    prog.run()
    assert prog.assert_failures() == 0
    e1, e2 = O.grow(a, 1, 1, 1, 1), O.size(a, -1)
    check(L1, e1, "L1")
    check(L2, e2, "L2")
    check(outer, O.op_not(e1, a), "outer")
    check(inner, O.op_not(O.canonical(a), e2), "inner")
    assert len(and1) == 0 and len(ring) == 0
    prog.close()
    eng.close()


def long_strips_layout(seed=77, nlong=2_000, nsmall=20_000, span_x=1_000_000, span_y=8_000):
    """Long horizontal strips over a field of small boxes: the grid sized by the
    box count is fine in x, so every long strip spans hundreds of tiles (the
    binning coarsens the grid) and the coarse tiles stack more strips over one
    x-position than a split can separate (they run on the dense kernel)."""
    rng = np.random.default_rng(seed)
    i = np.arange(nlong)
    y = (i * 4) % span_y
    x1 = rng.integers(0, span_x // 8, nlong)
    longs = np.stack([x1, y, x1 + rng.integers(span_x // 2, span_x, nlong), y + 2], 1)
    small = sparse(rng, nsmall, span_x, 40)
    small[:, 1] %= span_y
    small[:, 3] = small[:, 1] + rng.integers(1, 40, nsmall)
    return np.concatenate([longs, small]).astype(np.int32)


def test_long_strips_coarsen_the_grid(gpu_engine):
    eng = V.Engine(0)
    a = long_strips_layout()
    rng = np.random.default_rng(3)
    b = sparse(rng, 1_500, 1_000_000, 300)
    b[:, 1] %= 8_000
    b[:, 3] = b[:, 1] + rng.integers(1, 300, len(b))
    A, B, out = V.Layer(eng, a), V.Layer(eng, b), V.Layer(eng)
    V.CalculateBoostOperation(A, A, out, V.K_BOOST_CONTROL_MERGE)
    check(out, O.canonical(a), "merge")
    V.CalculateGROWLayer(A, out, (1, 1, 1, 1))
    check(out, O.grow(a, 1, 1, 1, 1), "grow")
    V.CalculateBoostOperation(A, A, out, V.K_BOOST_CONTROL_SIZE, -1)
    check(out, O.size(a, -1), "size-1")
    for code, fn in ((V.K_BOOST_CONTROL_AND, O.op_and), (V.K_BOOST_CONTROL_NOT, O.op_not)):
        V.CalculateBoostOperation(A, B, out, code)
        check(out, fn(a, b), ("binary", code))
    assert V.klee_area(A) == O.klee_area(a)
    eng.close()
