"""The C-ABI library loads and exports every function include/*.h declares (no
GPU needed: nothing here launches a kernel)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "vlsi_layout_amd", "libvlayout.so")

DECL = re.compile(r"^\s*(?:const\s+)?(?:int|void|int64_t|char)\s*\*?\s+(\w+)\s*\(", re.M)


def declared():
    names = []
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += DECL.findall(txt)
    return sorted(set(names))


def test_headers_declare_the_boundary():
    names = declared()
    for must in ("vl_and", "vl_not", "vl_or", "vl_xor", "vl_grow", "vl_size", "vl_merge",
                 "vl_assert_zero", "vl_klee_area", "vl_canonicalize", "vl_xor_empty",
                 "vl_boolean_op", "vl_event_sort", "vl_morton_sort_direct",
                 "vl_morton_sort_indirect", "device_event_sort", "thrust_sort_boxes_direct",
                 "thrust_sort_boxes_indirect", "vl_program_run", "vl_synth_layer"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvlayout.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvlayout.so not built")
def test_synth_is_deterministic_host_code():
    import numpy as np
    import vlsi_layout_amd as V
    a = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 0, 5000)
    b = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 0, 5000)
    assert np.array_equal(a, b)
    assert (a[:, 0] < a[:, 2]).all() and (a[:, 1] < a[:, 3]).all()
    for L in range(4):
        c = V.synth_layer(V.SYN_STDCELL, 0x5EED0002, L, 3000)
        assert (c[:, 0] < c[:, 2]).all() and (c[:, 1] < c[:, 3]).all()
    t = V.synth_tile_count(0x5EED0003, 0)
    assert 100 <= t <= 200000


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvlayout.so not built")
def test_uniform_overlap_fraction_near_30_percent():
    """The 200k-box uniform layer is tuned so ~30% of boxes intersect another."""
    import numpy as np
    import vlsi_layout_amd as V
    a = V.synth_layer(V.SYN_UNIFORM, 0x5EED0001, 0, 200000)
    # grid hash of boxes -> count boxes with a positive-area intersection partner
    cell = 2048
    from collections import defaultdict
    buckets = defaultdict(list)
    for i, (x1, y1, x2, y2) in enumerate(a.tolist()):
        for cx in range(x1 // cell, (x2 - 1) // cell + 1):
            for cy in range(y1 // cell, (y2 - 1) // cell + 1):
                buckets[(cx, cy)].append(i)
    hit = np.zeros(len(a), bool)
    for ids in buckets.values():
        if len(ids) < 2:
            continue
        b = a[ids]
        ix = (np.minimum(b[:, None, 2], b[None, :, 2]) > np.maximum(b[:, None, 0], b[None, :, 0]))
        iy = (np.minimum(b[:, None, 3], b[None, :, 3]) > np.maximum(b[:, None, 1], b[None, :, 1]))
        m = ix & iy
        np.fill_diagonal(m, False)
        hit[np.asarray(ids)[m.any(1)]] = True
    frac = hit.mean()
    print("overlap fraction", frac)
    assert 0.28 <= frac <= 0.32, frac
