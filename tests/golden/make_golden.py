"""Writes the golden fixtures under tests/golden/ (run in a container that has
the reference sources and oracle/_ref built: make -C oracle).

Inputs are the reference's own KLBIN fixtures (raw little-endian int32
x1,y1,x2,y2 records, src/kldata.f90:14-97): the SKY130 SEDFXTP1 cell layers
(src/SEDFXTP1_DATA/SED_L*_D*.bin) and src/square_L68_D20.bin.  Expected
outputs come from the reference's own routines compiled with flang
(oracle/_ref/libvlref.so):
  <name>.heal.bin        heal_boxes            (src/polygon_fracture.f90:805-1004)
  <name>.complement.bin  generate_trackers + scanline_fracture on MBR+10
                         (src/polygon_fracture.f90:767-803, 361-610)
  golden.json            calculate_union_area_fast (src/geometry.f90:1840-1927)
                         plus counts.
"""
import glob
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

REF = "/root/reference/src"


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libvlref.so is not built (make -C oracle)")
    inp = os.path.join(HERE, "inputs")
    exp = os.path.join(HERE, "expected")
    os.makedirs(inp, exist_ok=True)
    os.makedirs(exp, exist_ok=True)
    files = sorted(glob.glob(os.path.join(REF, "SEDFXTP1_DATA", "SED_L*.bin")))
    files.append(os.path.join(REF, "square_L68_D20.bin"))
    # Seeded random layers (overlapping boxes) exercise the sweep beyond the
    # small cell fixtures.
    rng = np.random.default_rng(0x5EED0000)
    randoms = {}
    for name, n, span, wmax in (("rand_500", 500, 1000, 80), ("rand_2000", 2000, 5000, 300),
                                ("rand_dense_200", 200, 100, 30)):
        x = rng.integers(-span // 2, span // 2, n)
        y = rng.integers(-span // 2, span // 2, n)
        w = rng.integers(1, wmax, n)
        h = rng.integers(1, wmax, n)
        randoms[name] = np.stack([x, y, x + w, y + h], 1).astype(np.int32)
    meta = {}
    for f in files + sorted(randoms):
        if f in randoms:
            name = f
            randoms[name].tofile(os.path.join(inp, name + ".bin"))
        else:
            name = os.path.basename(f)[:-4]
            shutil.copyfile(f, os.path.join(inp, name + ".bin"))
        a = np.fromfile(os.path.join(inp, name + ".bin"), dtype=np.int32).reshape(-1, 4)
        heal = O.ref_heal(a) if len(a) else np.zeros((0, 4), np.int32)
        heal.astype(np.int32).tofile(os.path.join(exp, name + ".heal.bin"))
        rec = {"n": int(len(a)), "heal_n": int(len(heal)),
               "area": int(O.ref_union_area(a)) if len(a) else 0}
        if len(a):
            bb = (int(a[:, 0].min()) - 10, int(a[:, 1].min()) - 10, int(a[:, 2].max()) + 10,
                  int(a[:, 3].max()) + 10)
            comp = O.ref_complement(a, bb)
            comp.astype(np.int32).tofile(os.path.join(exp, name + ".complement.bin"))
            rec["complement_bbox"] = list(bb)
            rec["complement_n"] = int(len(comp))
        meta[name] = rec
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print(f"wrote {len(meta)} fixtures")


if __name__ == "__main__":
    main()
