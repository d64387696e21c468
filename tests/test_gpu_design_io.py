"""Designs through the `exec run` driver: `decl design` loads a .mag file whose
sections hold rect lines or name KLBIN / SNAP files (parseMagicLayoutFile,
src/magparser.f90:108-418), `decl output` names the files `flush` writes
(writeMagicLayoutFile, :510-564), `decl memory` with temporary_layers 2
spills to SNAP files and reads them back (src/controlparser.f90:613-626,
src/diskdesign_impl.f90:83-114).  Every written file is read back
independently (raw KLBIN, the test-side SNAP reader with the oracle codec)
and must equal the oracle's result for its operator."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import snap_file as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vlsi_layout_amd", "vl_magparser")
GOLDEN_IN = os.path.join(ROOT, "tests", "golden", "inputs")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (flang missing)")

# overlapping rectangles: the design load heals them (src/magparser.f90:358-418)
DIFF = [[-10, -50, 20, 25], [0, -50, 70, 25], [40, 0, 90, 40], [-5, 90, 20, 165], [10, 100, 70, 165]]


def read_klbin(path):
    return np.fromfile(path, dtype="<i4").reshape(-1, 4)


def run_driver(tmp_path, script):
    ctl = tmp_path / "run.ctr"
    ctl.write_text(script)
    return subprocess.run([DRIVER, str(ctl)], capture_output=True, text=True, timeout=300, cwd=tmp_path)


def setup_design(tmp_path, met1_method=S.ZLIB):
    poly = read_klbin(os.path.join(GOLDEN_IN, "SED_L66_D20.bin"))
    met1 = read_klbin(os.path.join(GOLDEN_IN, "SED_L67_D20.bin"))
    shutil.copy(os.path.join(GOLDEN_IN, "SED_L66_D20.bin"), tmp_path / "poly.bin")
    (tmp_path / "met1.snap").write_bytes(S.build(met1, met1_method, props=1))
    rects = "".join(f"rect {x1} {y1} {x2} {y2}\n" for x1, y1, x2, y2 in DIFF)
    (tmp_path / "in.mag").write_text(
        "magic\ntech sky130A\ntimestamp 1\nmagscale 1 2\n"
        f"<< diff >>\n{rects}<< poly >>\nKLBIN poly.bin\n<< met1 >>\nSNAP met1.snap\n"
        "<< labels >>\nrlabel metal1 -30 -45 102 -30 1 VGND\n<< end >>\n")
    (tmp_path / "out.mag").write_text(
        "magic\ntech sky130A\ntimestamp 1\n"
        "<< outer >>\nKLBIN outer.bin\n<< inner >>\nSNAP inner.snap\n<< diff >>\nKLBIN diff.bin\n"
        "<< m1 >>\nSNAP m1.snap\n<< end >>\n")
    return poly, met1


DESIGN_SCRIPT = """program design_io
systeminfo
var temp_folder tmp_
var abort_on_assert_zero 1
var temporary_layers 2
decl design d1 = in.mag
decl output o1 = out.mag
decl memory t1 = nothing
info
exec init t1:nothing
exec run t1:poly_1 = d1:poly GROW t1:nothing 0.001 0.001 0.001 0.001
exec run t1:poly_2 = d1:poly SIZE t1:nothing -0.001
exec run t1:poly_outer_ring = t1:poly_1 ~ d1:poly
exec run t1:poly_inner_ring = d1:poly ~ t1:poly_2
exec run t1:poly_ring_and1 = t1:poly_outer_ring * d1:poly
exec run t1:nothing = t1:poly_ring_and1 ASSERT_ZERO t1:nothing
exec run o1:outer = t1:poly_outer_ring COPY nothing
exec run o1:inner = t1:poly_inner_ring COPY nothing
exec run o1:diff = d1:diff COPY nothing
exec run o1:m1 = d1:met1 * d1:poly
dump d1:diff d1_diff.bin
flush
dump t1:poly_1 g_after_flush.bin
dump o1:m1 m1_after_flush.bin
area d1:diff
end
"""


@needs_driver
@pytest.mark.parametrize("method", [S.SNAPPY, S.ZLIB, S.ZSTD])
def test_design_load_operators_and_flush(tmp_path, gpu_engine, method):
    poly, met1 = setup_design(tmp_path, method)
    r = run_driver(tmp_path, DESIGN_SCRIPT)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ASSERT_ZERO failures 0" in r.stdout and "--- Starting Program: design_io" in r.stdout
    g = O.grow(poly, 1, 1, 1, 1)
    s = O.size(poly, -1)
    outer, inner = O.op_not(g, poly), O.op_not(poly, s)
    diff = np.array(DIFF, np.int32)
    # output design: KLBIN and SNAP files named by out.mag
    assert np.array_equal(O.sort_boxes(read_klbin(tmp_path / "outer.bin")), outer)
    assert np.array_equal(O.sort_boxes(S.boxes_of((tmp_path / "inner.snap").read_bytes())), inner)
    assert np.array_equal(O.sort_boxes(read_klbin(tmp_path / "diff.bin")), O.canonical(diff))
    # the overlapping rect layer was healed when the design was loaded
    assert np.array_equal(O.sort_boxes(read_klbin(tmp_path / "d1_diff.bin")), O.canonical(diff))
    m1 = S.boxes_of((tmp_path / "m1.snap").read_bytes())
    assert np.array_equal(O.sort_boxes(m1), O.op_and(met1, poly))
    f = S.parse((tmp_path / "m1.snap").read_bytes())
    assert f["method"] == S.SNAPPY and f["props"] == 1        # flush writes Snappy; results are healed
    # memory design spilled at flush (temporary_layers 2), read back when used
    spill = tmp_path / "tmp_t1_MTL_poly_1"
    assert spill.exists()
    assert np.array_equal(O.sort_boxes(S.boxes_of(spill.read_bytes())), g)
    assert np.array_equal(O.sort_boxes(read_klbin(tmp_path / "g_after_flush.bin")), g)
    assert np.array_equal(O.sort_boxes(read_klbin(tmp_path / "m1_after_flush.bin")), O.op_and(met1, poly))
    assert f"AREA d1:diff = {O.klee_area(diff)}" in r.stdout
    assert not (tmp_path / "MTL_NOTHING").exists()


@needs_driver
def test_design_statement_forms(tmp_path, gpu_engine):
    """exec modedesign loads like decl design; unknown statements are reported
    and skipped; without temporary_layers 2 memory designs stay in memory."""
    poly, met1 = setup_design(tmp_path)
    script = """exec modedesign d1 = in.mag
decl output o1 = out.mag
decl memory t1 = nothing
frobnicate 1 2 3
exec run t1:x = d1:met1 + d1:poly
exec run o1:m1 = t1:x COPY nothing
flush
end
"""
    r = run_driver(tmp_path, script)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Unknown command: frobnicate" in r.stdout
    assert np.array_equal(O.sort_boxes(S.boxes_of((tmp_path / "m1.snap").read_bytes())), O.op_or(met1, poly))
    assert not (tmp_path / "t1_MTL_x").exists()


@needs_driver
@pytest.mark.parametrize("case", ["unlisted_output", "deleted_design", "memory_backing", "missing_file",
                                  "bad_rect", "duplicate_design"])
def test_design_errors(tmp_path, gpu_engine, case):
    setup_design(tmp_path)
    head = "decl design d1 = in.mag\ndecl output o1 = out.mag\ndecl memory t1 = nothing\n"
    body = {
        # a layer out.mag does not list has a temporary backing name: flush refuses it
        "unlisted_output": "exec run o1:zz = d1:poly COPY nothing\nflush\n",
        "deleted_design": "delete d1\nexec run t1:x = d1:poly COPY nothing\ndump t1:x x.bin\n",
        "memory_backing": "decl memory t2 = foo\n",
        "missing_file": "decl design d2 = nosuch.mag\n",
        "bad_rect": None,
        "duplicate_design": "decl design d1 = in.mag\n",
    }[case]
    if case == "bad_rect":
        (tmp_path / "bad.mag").write_text("<< x >>\nrect 5 5 5 9\n<< end >>\n")
        body = "decl design d2 = bad.mag\n"
    r = run_driver(tmp_path, head + body + "end\n")
    assert r.returncode != 0, r.stdout + r.stderr


@needs_driver
def test_missing_layer_file_is_empty_with_temporary_layers(tmp_path, gpu_engine):
    """With temporary_layers 2 a layer whose file is missing is empty
    (src/diskdesign_impl.f90:87-97); without it the load is an error."""
    (tmp_path / "in.mag").write_text("<< a >>\nSNAP gone.snap\n<< b >>\nrect 0 0 4 4\n<< end >>\n")
    script = "var temporary_layers 2\ndecl design d1 = in.mag\nexec run t1:c = d1:b + d1:a\ndump t1:c c.bin\nend\n"
    r = run_driver(tmp_path, script)
    assert r.returncode == 0, r.stdout + r.stderr
    assert read_klbin(tmp_path / "c.bin").tolist() == [[0, 0, 4, 4]]
    r = run_driver(tmp_path, script.replace("var temporary_layers 2\n", ""))
    assert r.returncode != 0
