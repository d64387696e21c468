"""Single-layer AND (vl_self_and: the points two or more boxes cover,
src/design_impl.f90:385-509), GRID (vl_grid, CreateGrid
src/design_impl.f90:159-275) and EXTENT (vl_extent, CreateEXTENT :277-293)
against the oracle: canonical strips bit-exact for the single-layer AND (on
the wave, workgroup, split and dense tile paths), the grid's boxes bit-exact
and in order."""
import numpy as np
import pytest

import oracle_lib as O
import vlsi_layout_amd as V

pytestmark = pytest.mark.gpu


def rand_boxes(rng, n, span, wmax):
    x = rng.integers(0, span, n)
    y = rng.integers(0, span, n)
    return np.stack([x, y, x + rng.integers(1, wmax + 1, n), y + rng.integers(1, wmax + 1, n)], 1).astype(np.int32)


def layouts():
    rng = np.random.default_rng(11)
    r = rand_boxes(rng, 4000, 30000, 800)
    dup = np.concatenate([r[:1500], r[:700]])                                  # exact duplicates
    dense5 = (rand_boxes(rng, 5000, 600, 6) * 5).astype(np.int32)              # 5-DBU grid, touching
    i = np.arange(300)
    bus = np.stack([100 + (i * 37) % 997, 1000 + 8 * i, 60_000 + (i * 37) % 997, 1010 + 8 * i], 1).astype(np.int32)
    ident = np.tile(np.array([[500, 500, 540, 525]], np.int32), (300, 1))
    return {"random": r, "duplicates": dup, "dense5": dense5, "bus_overlapping": bus, "identical300": ident,
            "stdcell": V.synth_layer(V.SYN_STDCELL, 0x5EED0002, 0, 20000)}


LAYOUTS = layouts()


@pytest.mark.parametrize("target", [20, 4])
@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_self_and_against_oracle(gpu_engine, name, target):
    eng = V.Engine(0)
    eng.set("tile_target", target)
    a = LAYOUTS[name]
    A, out = V.Layer(eng, a), V.Layer(eng)
    V.CalculateSingleLayerAND(A, out)
    assert np.array_equal(O.sort_boxes(out.boxes()), O.self_and(a)), (name, target)
    eng.close()


def test_self_and_small_cases(gpu_engine):
    eng = V.Engine(0)
    out = V.Layer(eng)
    for a, exp in (([[0, 0, 10, 10]], []), ([[0, 0, 10, 10], [0, 0, 10, 10]], [[0, 0, 10, 10]]),
                   ([[0, 0, 10, 10], [10, 0, 20, 10]], []), ([[0, 0, 10, 10], [5, 5, 15, 15]], [[5, 5, 10, 10]])):
        V.CalculateSingleLayerAND(V.Layer(eng, a), out)
        assert O.sort_boxes(out.boxes()).tolist() == exp
    V.CalculateSingleLayerAND(V.Layer(eng), out)
    assert len(out) == 0
    eng.close()


@pytest.mark.parametrize("rows,cols,overlap", [(10, 10, 10), (1, 1, 0), (3, 7, 0), (5, 2, 3), (64, 33, 1)])
def test_grid_against_oracle(gpu_engine, rows, cols, overlap):
    eng = V.Engine(0)
    a = LAYOUTS["random"]
    A, out = V.Layer(eng, a), V.Layer(eng)
    V.CreateGrid(A, out, rows, cols, overlap)
    m = (a[:, 0].min(), a[:, 1].min(), a[:, 2].max(), a[:, 3].max())
    assert np.array_equal(out.boxes(), O.grid(m, rows, cols, overlap))
    V.CreateEXTENT(A, out)
    assert out.boxes().tolist() == [list(m)]
    eng.close()


def test_grid_and_extent_errors(gpu_engine):
    eng = V.Engine(0)
    A, E, out = V.Layer(eng, [[0, 0, 5, 5]]), V.Layer(eng), V.Layer(eng)
    with pytest.raises(V.VLError) as ei:
        V.CreateGrid(A, out, 0, 3, 0)
    assert ei.value.code == V.VL_EINVAL
    with pytest.raises(V.VLError) as ei:
        V.CreateEXTENT(E, out)
    assert ei.value.code == V.VL_EINVAL
    V.CreateEXTENT(A, A)                 # in place
    assert A.boxes().tolist() == [[0, 0, 5, 5]]
    eng.close()
