#!/usr/bin/env python3
"""OP1-ring throughput of the gfx950 rectangle-Boolean engine.

Workload: the standard-cell-like 4-layer design (poly, li1, met1, met2; rows of
2000 cells on a 5-DBU grid, seed 0x5EED0002) at 10M boxes per layer, resident
in HBM.  One step is the OP1 ring (src/op_template.m4:14-22) with d = 1 DBU on
every layer:
    L1 = GROW(L; 1,1,1,1)   L2 = SIZE(L; -1)   outer = L1 ~ L   inner = L ~ L2
    and1 = outer * L        ring = outer * inner   ASSERT_ZERO(and1), ASSERT_ZERO(ring)
Throughput = input boxes of all operators of a step (N_A+N_B for binary ops,
N for GROW/SIZE) / wall time of the step, with the engine stream synchronised
on both sides of the timed region.  "plain" runs one operator per call with
the default tile budgets; "batched" runs the same six operators per layer as
one recorded program after the autotuner has picked the tile budgets and
whether groups fuse: GROW, SIZE and the two NOTs of a layer are one fused
group (the layer binned once, the four outputs computed per tile from shared
masks) unless the autotuner finds the four single operators faster, the two
ANDs run on the materialised outer / inner layers (so the ASSERT_ZEROs check
what the group produced), and the four layer chains run concurrently on
worker streams.  "graph" (the headline) is that program captured as a HIP graph
(vl_program_set_graph): the first warm-up run records the sizes, captures
the work and launches it, and the timed steps are replays of the graph
(no host round trips; every replay checks its sizes on the device).  All
forms evaluate every operator on the GPU and are timed with the same
settings; the box count of each form is taken from its own run.

The roofline figure is for the kernel stage with the largest device time in
the timed step: algorithmic bytes of its launches (tile / fused stage: 16 B
per entry read, 16 B per strip written, 48 B per tile; sort: 16 B per entry
per pass) / their summed duration from HIP events on the stream they run on,
against the MI355X's 8 TB/s HBM peak.  For the batched and graph forms the
durations come from one more untimed step of the eager program with its
chains on one worker stream (the kernel alone on the GPU: on four streams
every launch shares the GPU with the other chains' kernels, and its events
time the sharing); events recorded by graph replays give no usable elapsed
times, and the eager program launches the same kernels with the same sizes.
"compulsory_io_frac" is 16 B per input and output box of every operator /
step time, against the same peak.

cpu_baseline: the same OP1 ring on the first 2M healed boxes of met1 (a band
of rows), run by the C/OpenMP restatement of the reference CPU path
(oracle/liboracle.so, frame-tiled like src/design_impl.f90:993-1316) on the
host's cores.  Its time grows faster than the box count (the frame grid has
only nodes^(1/4) tiles per side), so a whole 10M-box layer is not sampled.

Also reported in the same line: "c2_10M_total", the same ring with 10M boxes
over the 4 layers (2.5M per layer), plain and batched; "c3_uneven_tiles",
AND/NOT/OR/GROW+2 over 4096 packed tiles of uneven size (~50M boxes), plain,
as one recorded program and as its graph, autotuned; "c4_full_chip", AND,
NOT, GROW+1 and the exact Klee area on a 2 x 180M-box layer pair;
"design_io", loading the 10M-box met1 layer from a SNAP file (host Snappy +
GPU codec decode) and from KLBIN; and "serving", windowed query latency
(p50/p99) and throughput of the query server on a resident 50M-box 4-layer
design under Poisson arrivals.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Boolean-op boxes/sec at 10M boxes/layer, batched vs plain; query p99 latency"
SEED = 0x5EED0002
LAYERS = ["poly", "li1", "met1", "met2"]
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--boxes-per-layer", type=int, default=10_000_000)
    p.add_argument("--cpu-sample", type=int, default=2_000_000,
                   help="healed boxes of met1 (a band of rows) in the CPU baseline's ring")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--mode", choices=["plain", "batched", "graph", "all"], default="all",
                   help="plain: one call per operator; batched: the recorded program run eagerly; "
                        "graph: the same program captured as a HIP graph and replayed (the headline)")
    p.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                   help="engine parameter for every form (e.g. the autotuned values, with --no-autotune)")
    p.add_argument("--no-autotune", action="store_true", help="batched form with the default tile budgets")
    p.add_argument("--c2-total", type=int, default=10_000_000,
                   help="also time the same ring at this many boxes over the 4 layers (0 = skip)")
    p.add_argument("--serving-boxes", type=int, default=50_000_000,
                   help="serving design size over 4 layers (0 = skip the serving measurement)")
    p.add_argument("--serving-requests", type=int, default=400)
    p.add_argument("--no-c3", action="store_true", help="skip the 4096-uneven-tile measurement")
    p.add_argument("--c4-boxes", type=int, default=360_000_000,
                   help="full-chip layer pair size for the C4 measurement (0 = skip)")
    p.add_argument("--design-io-boxes", type=int, default=10_000_000,
                   help="boxes of the layer-file measurement (0 skips it)")
    p.add_argument("--calibrate-copy-mib", type=int, default=0,
                   help="first copy this many MiB device-to-device (a copy kernel of known size for "
                        "calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    return p.parse_args()


class Ring:
    """Layers and temporaries of the OP1 ring for one input layer."""

    def __init__(self, V, eng, boxes):
        self.L = V.Layer(eng, boxes)
        self.L1, self.L2, self.outer, self.inner, self.and1, self.ring = (V.Layer(eng) for _ in range(6))

    def run_plain(self, V):
        V.CalculateGROWLayer(self.L, self.L1, (1, 1, 1, 1))
        V.CalculateBoostOperation(self.L, self.L, self.L2, V.K_BOOST_CONTROL_SIZE, -1)
        V.CalculateFrameNOT(self.L1, self.L, self.outer)
        V.CalculateFrameNOT(self.L, self.L2, self.inner)
        V.CalculateAND(self.outer, self.L, self.and1)
        V.CalculateAND(self.outer, self.inner, self.ring)
        fails = int(V.ASSERT_ZERO(self.and1) > 0) + int(V.ASSERT_ZERO(self.ring) > 0)
        return self.counts(), fails

    def counts(self):
        """(input boxes, output boxes) of the ring's six operators as last run:
        N for GROW and SIZE, N_A + N_B for NOT and AND.  Both forms evaluate all
        six on the GPU (the batched form fuses GROW, SIZE and the two NOTs into
        one tile pass and runs the two ANDs on the materialised outer and inner
        layers), so the same count holds for both."""
        n, l1, l2, o, i = (x.n_used for x in (self.L, self.L1, self.L2, self.outer, self.inner))
        a1, r = self.and1.n_used, self.ring.n_used
        ins = n + n + (l1 + n) + (n + l2) + (o + n) + (o + i)
        outs = l1 + l2 + o + i + a1 + r
        return ins, outs

    def record(self, V, prog):
        prog.add(V.VL_OP_GROW, self.L, None, self.L1, (1, 1, 1, 1))
        prog.add(V.K_BOOST_CONTROL_SIZE, self.L, None, self.L2, (-1,))
        prog.add(V.K_BOOST_CONTROL_NOT, self.L1, self.L, self.outer)
        prog.add(V.K_BOOST_CONTROL_NOT, self.L, self.L2, self.inner)
        prog.add(V.K_BOOST_CONTROL_AND, self.outer, self.L, self.and1)
        prog.add(V.K_BOOST_CONTROL_AND, self.outer, self.inner, self.ring)
        prog.add(-1, self.and1)
        prog.add(-1, self.ring)


def time_steps(fn, eng, steps, warmup):
    """W untimed warm-up steps, then K timed steps with the engine stream
    synchronised on both sides; per-stage kernel statistics (HIP events on
    the streams the kernels run on, read after the timed region) are reset at
    the start of the timed region."""
    for _ in range(warmup):
        fn()
    eng.sync()
    eng.set("profile", 1)
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    eng.sync()
    secs = time.perf_counter() - t0
    stats = stage_stats(eng, steps)
    eng.set("profile", 0)
    return secs, out, stats


STAGES = {
    "tile": "single-operator tile stage (k_tile_w tiers + k_tile_b)",
    "fused": "fused tile kernel of an operator group (k_tile_f tiers)",
    "sort": "binning radix sort (k_rs_upsweep + k_rs_downsweep passes)",
    "bin": "tile binning (span, emit, radix sort, bounds, gather)",
}


def stage_stats(eng, steps):
    """Per-step launches, device ms and algorithmic bytes of each stage."""
    out = {}
    for k in STAGES:
        n, ms, by = eng.kernel_stats(k)
        out[k] = {"launches_per_step": n / steps, "ms_per_step": ms / steps, "bytes_per_step": by / steps,
                  "gbs": (by / (ms * 1e-3) / 1e9) if ms > 0 else 0.0}
    return out


def profile_figure(name, key):
    """A figure recorded in a committed rocprofv3 summary (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith(key + " = "):
            return float(ln.split("=")[1])
    return None


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def serving(V, args):
    """Windowed DRC spacing queries R = GROW_s(A & W) AND (B & W), ASSERT_ZERO(R),
    against a resident 4-layer routing design (met1 wires of the even and odd
    horizontal tracks, met2 of the even and odd vertical ones; neighbouring
    tracks 6 DBU apart, one wire in 4096 jogged to 4 DBU): A and B are the two
    layers of one metal, so a query fails exactly where a jog lies in its
    window at s = 5.  Poisson arrivals at 200 req/s, window box counts
    log-uniform in [1k, 5M], s in {1, 2, 5} DBU, 64 GiB admission budget,
    4 concurrent slots."""
    rng = np.random.default_rng(0x5EED0005)
    per = args.serving_boxes // 4
    boxes = [V.synth_layer(V.SYN_ROUTE, 0x5EED0005, k, per) for k in range(4)]
    eng = V.Engine(0)
    layers = [V.Layer(eng, b) for b in boxes]
    t = time.perf_counter()
    srv = V.Server(eng, layers, slots=4, hbm_budget=64 << 30)
    t_index = time.perf_counter() - t
    mbr = [(int(b[:, 0].min()), int(b[:, 1].min()), int(b[:, 2].max()), int(b[:, 3].max())) for b in boxes]
    dens = [len(b) / float((m[2] - m[0]) * (m[3] - m[1])) for b, m in zip(boxes, mbr)]
    n = args.serving_requests
    q = np.zeros(n, V.QUERY_DTYPE)
    arrival = 0.0
    for i in range(n):
        la = int(rng.integers(0, 4))
        lb = la ^ 1                                                     # the other layer of the metal
        target = float(np.exp(rng.uniform(np.log(1e3), np.log(5e6))))     # boxes of A and B in W
        side = int(np.sqrt(target / (dens[la] + dens[lb])))
        m = mbr[la]
        x = int(rng.integers(m[0], max(m[0] + 1, m[2] - side)))
        y = int(rng.integers(m[1], max(m[1] + 1, m[3] - side)))
        q[i] = (x, y, x + max(side, 1), y + max(side, 1), la, lb, int(rng.choice([1, 2, 5])), 0, arrival)
        arrival += rng.exponential(1000.0 / 200.0)
    srv.run(q[:8])                                                           # warm the slots
    rc, res = srv.run(q)
    if rc != 0:
        raise SystemExit(f"serving run failed with status {rc}")
    lat = res["end_ms"] - q["arrival_ms"]
    wall = (res["end_ms"].max() - q["arrival_ms"].min()) / 1000.0
    events = sorted([(a, 1) for a in q["arrival_ms"]] + [(s_, -1) for s_ in res["start_ms"]])
    depth = peak = 0
    for _, d in events:
        depth += d
        peak = max(peak, depth)
    srv.close()
    eng.close()
    return {"requests": n, "design_boxes": per * 4, "arrival_rate_per_s": 200, "slots": 4,
            "hbm_budget_gib": 64, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "boxes_per_s": float(res["in_boxes"].sum() / wall),
            "window_boxes_mean": float(res["in_boxes"].mean()), "max_queue_depth": int(peak),
            "assert_zero_failures": int((res["count"] > 0).sum()),
            "assert_zero_failures_by_spacing": {str(sp): int(((res["count"] > 0) & (q["spacing"] == sp)).sum())
                                                for sp in (1, 2, 5)},
            "queries_by_spacing": {str(sp): int((q["spacing"] == sp).sum()) for sp in (1, 2, 5)},
            "design": "routing tracks (met1/met2 even and odd tracks, 6 DBU apart, 1 in 4096 wires jogged to 4 DBU)",
            "index_build_s": t_index}


def c3_uneven_tiles(V, args):
    """4096 independent tiles (box counts LogNormal(ln 5000, 1.34) clipped to
    [100, 200000], split over two layers, tile-local coordinates in [0, 2^16)^2)
    packed into one layer pair at a 2^17 pitch; one step = AND, NOT, OR and
    GROW +2 of the pair, operator by operator (plain), as one recorded program
    (batched) and as its captured graph, autotuned first."""
    seed, ntiles, pitch = 0x5EED0003, 4096, 1 << 17
    counts = np.array([V.synth_tile_count(seed, t) for t in range(ntiles)])
    A_t, B_t = [], []
    for t in range(ntiles):
        na = int(counts[t]) // 2
        off = np.array([(t % 64) * pitch, (t // 64) * pitch] * 2, np.int32)
        A_t.append(V.synth_layer(V.SYN_TILES, seed, 2 * t, na) + off)
        B_t.append(V.synth_layer(V.SYN_TILES, seed, 2 * t + 1, int(counts[t]) - na) + off)
    A, B = np.concatenate(A_t), np.concatenate(B_t)
    del A_t, B_t
    eng = V.Engine(0)
    LA, LB = V.Layer(eng, A), V.Layer(eng, B)
    outs = [V.Layer(eng) for _ in range(4)]
    tuned = eng.autotune(LA, LB)

    def plain():
        V.CalculateAND(LA, LB, outs[0])
        V.CalculateFrameNOT(LA, LB, outs[1])
        V.CalculateOR(LA, LB, outs[2])
        V.CalculateGROWLayer(LA, outs[3], (2, 2, 2, 2))

    prog = V.Program(eng)
    prog.add(V.K_BOOST_CONTROL_AND, LA, LB, outs[0])
    prog.add(V.K_BOOST_CONTROL_NOT, LA, LB, outs[1])
    prog.add(V.K_BOOST_CONTROL_OR, LA, LB, outs[2])
    prog.add(V.VL_OP_GROW, LA, None, outs[3], (2, 2, 2, 2))
    plain()
    ref = [o.n_used for o in outs]
    ins = 3 * (len(A) + len(B)) + len(A)
    res = {"tiles": ntiles, "boxes": int(len(A) + len(B)), "input_boxes_per_step": int(ins), "autotuned": tuned}
    for name, fn, w in (("plain", plain, args.warmup), ("batched", prog.run, args.warmup), ("graph", prog.run, 1)):
        if name == "graph":
            prog.set_graph(True)
        secs, _, _ = time_steps(fn, eng, args.steps, max(w, 1))
        if [o.n_used for o in outs] != ref:
            raise SystemExit(f"c3 {name}: output sizes differ from the plain form's")
        res[f"{name}_ms_per_step"] = 1000.0 * secs / args.steps
        res[f"{name}_boxes_per_s"] = ins * args.steps / secs
    prog.close()
    eng.close()
    return res


def c4_full_chip(V, args):
    """One layer pair of 2 x (boxes/2) on [0, 2^26)^2 (70 % routing tracks on
    a 5-DBU grid, 30 % 6x6 vias); one step = AND, NOT, GROW +1 and the exact
    Klee area of A, operator by operator, and with AND, NOT and GROW as one
    recorded program (independent chains on worker streams) before the area."""
    seed, n = 0x5EED0004, args.c4_boxes // 2
    a = V.synth_layer(V.SYN_CHIP, seed, 0, n)
    b = V.synth_layer(V.SYN_CHIP, seed, 1, n)
    eng = V.Engine(0)
    LA, LB = V.Layer(eng, a), V.Layer(eng, b)
    del a, b
    AND, NOT, GROW = V.Layer(eng), V.Layer(eng), V.Layer(eng)
    area = {}

    def step():
        V.CalculateAND(LA, LB, AND)
        V.CalculateFrameNOT(LA, LB, NOT)
        V.CalculateGROWLayer(LA, GROW, (1, 1, 1, 1))
        area["A"] = V.klee_area(LA)

    step()
    if V.klee_area(AND) + V.klee_area(NOT) != area["A"]:
        raise SystemExit("c4: area(A AND B) + area(A NOT B) != area(A)")
    ins = 2 * (2 * n) + n + n
    secs, _, _ = time_steps(step, eng, args.steps, args.warmup)
    ref = [AND.n_used, NOT.n_used, GROW.n_used]
    prog = V.Program(eng)
    prog.add(V.K_BOOST_CONTROL_AND, LA, LB, AND)
    prog.add(V.K_BOOST_CONTROL_NOT, LA, LB, NOT)
    prog.add(V.VL_OP_GROW, LA, None, GROW, (1, 1, 1, 1))

    def batched():
        prog.run()
        area["A"] = V.klee_area(LA)

    bsecs, _, _ = time_steps(batched, eng, args.steps, max(args.warmup, 1))
    if [AND.n_used, NOT.n_used, GROW.n_used] != ref:
        raise SystemExit("c4: the program's outputs differ from the operator-by-operator ones")
    prog.close()
    eng.close()
    return {"boxes": 2 * n, "input_boxes_per_step": int(ins), "ms_per_step": 1000.0 * secs / args.steps,
            "boxes_per_s": ins * args.steps / secs, "batched_ms_per_step": 1000.0 * bsecs / args.steps,
            "batched_boxes_per_s": ins * args.steps / bsecs, "klee_area_A": int(area["A"])}


def design_io(V, args):
    """Layer files: the benchmark design's met1 layer (10M boxes by default)
    written as a Snappy SNAP file (the reference's flush format) and as KLBIN,
    then read back into HBM.  One SNAP step = one RestoreSnapToLayer: file
    read, Snappy decompression of the 65536-box chunks on the host cores,
    upload of the codec bytes, GPU codec decode straight into the layer and
    box validation.  The GPU decode alone is timed by its HIP events
    (kernel statistic "snap"; algorithmic bytes: the codec stream read once
    plus 16 B per box written).  cpu: the oracle's sequential codec decode
    (the reference's DecompressBoxesUsingCodec loop) of the first 16 chunks
    on one core."""
    import shutil
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n = args.design_io_boxes
    boxes = V.synth_layer(V.SYN_STDCELL, SEED, 2, n)
    eng = V.Engine(0)
    L = V.Layer(eng, boxes)
    d = tempfile.mkdtemp(prefix="vl_io_")
    try:
        snap, kl = os.path.join(d, "met1.snap"), os.path.join(d, "met1.bin")
        t0 = time.perf_counter()
        V.SaveLayerToSnap(L, snap, V.COMPRESSION_METHOD_SNAPPY)
        t_save = time.perf_counter() - t0
        V.WriteKLBin(kl, L)
        L.close()
        R = V.Layer(eng)

        def timed(fn):
            for _ in range(args.warmup):
                fn()
            eng.sync()
            eng.set("profile", 1)
            t = time.perf_counter()
            for _ in range(args.steps):
                fn()
            eng.sync()
            secs = time.perf_counter() - t
            st = eng.kernel_stats("snap")
            eng.set("profile", 0)
            return secs / args.steps, st

        s_snap, (nl, ms, by) = timed(lambda: V.RestoreSnapToLayer(R, snap))
        if not np.array_equal(R.boxes(), boxes):
            raise SystemExit("design_io: the SNAP file did not load back to the layer")
        s_kl, _ = timed(lambda: V.LoadKLBin(kl, R))
        _, total, nch, _ = V.SnapInfo(snap)
        dec_ms = ms / args.steps
        dec_gbs = (by / args.steps) / (dec_ms * 1e-3) / 1e9
        chunks = [boxes[k * 65536:(k + 1) * 65536] for k in range(min(16, nch))]
        raws = [V.codec_encode(c) for c in chunks]
        t = time.perf_counter()
        for c, raw in zip(chunks, raws):
            if O.codec_decode(raw, len(c)) is None:
                raise SystemExit("design_io: oracle decode failed")
        cpu_s = time.perf_counter() - t
        res = {"boxes": int(total), "chunks": int(nch), "snap_bytes": os.path.getsize(snap),
               "codec_bytes_per_box": (by / args.steps - 16.0 * total) / total,
               "save_snap_s": t_save,
               "load_snap_ms": 1000.0 * s_snap, "load_snap_boxes_per_s": total / s_snap,
               "load_klbin_ms": 1000.0 * s_kl, "load_klbin_boxes_per_s": total / s_kl,
               "gpu_decode_ms": dec_ms, "gpu_decode_boxes_per_s": total / (dec_ms * 1e-3),
               "gpu_decode_gbs": dec_gbs, "gpu_decode_hbm_frac": dec_gbs / HBM_PEAK_GBS,
               "gpu_decode_launches": nl / args.steps,
               "cpu_decode_boxes_per_s": sum(len(c) for c in chunks) / cpu_s, "cpu_decode_cores": 1}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    eng.close()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import vlsi_layout_amd as V

    npl = args.boxes_per_layer
    t_gen = time.perf_counter()
    layers = [V.synth_layer(V.SYN_STDCELL, SEED, k, npl) for k in range(4)]
    t_gen = time.perf_counter() - t_gen

    print(f"[bench] generated 4 x {npl} boxes in {t_gen:.1f} s", file=sys.stderr, flush=True)
    eng = V.Engine(0)
    for kv in args.set:
        k, v = kv.split("=")
        eng.set(k, int(v))
    rings = [Ring(V, eng, lay) for lay in layers]
    if args.calibrate_copy_mib:
        cal = V.Layer(eng).copy_from(rings[0].L, args.calibrate_copy_mib * (1 << 20) // 16)
        eng.sync()
        del cal
    result = {}

    def step_counts():
        ins = outs = 0
        for r in rings:
            i, o = r.counts()
            ins += i
            outs += o
        return ins, outs

    def plain_step():
        fails = 0
        for r in rings:
            fails += r.run_plain(V)[1]
        return fails

    prog = V.Program(eng)
    for r in rings:
        r.record(V, prog)

    def batched_step():
        prog.run()
        return prog.assert_failures()

    modes = ["plain", "batched", "graph"] if args.mode == "all" else [args.mode]
    if modes[0] == "plain":
        if plain_step():
            raise SystemExit("OP1 ring assertion failed")
    else:
        if batched_step():
            raise SystemExit("OP1 ring assertion failed")
    plain_counts = step_counts()
    print(f"[bench] first step ok: {plain_counts[0]} input boxes per step", file=sys.stderr, flush=True)

    tuned = None
    for m in modes:
        if m == "graph":
            prog.set_graph(True)     # the first warm-up run records, captures and launches the graph
        if m in ("batched", "graph") and tuned is None and not args.no_autotune:
            # the batched form runs with autotuned tile budgets (timed on met1's ring:
            # met1 NOT SIZE(met1, -1), GROW, SIZE and the fused ring of met1)
            tuned = eng.autotune(rings[2].L, rings[2].L2)
            print(f"[bench] autotuned: {tuned}", file=sys.stderr, flush=True)
        # Both forms are timed the same way: per-stage statistics on, read
        # after the timed region (their HIP events add no waits).
        secs, fails, stats = time_steps(plain_step if m == "plain" else batched_step, eng, args.steps,
                                        max(args.warmup, 1 if m == "graph" else 0))
        if fails:
            raise SystemExit(f"OP1 ring assertion failed in {m} mode")
        ins, outs = step_counts()
        if ins != plain_counts[0]:
            raise SystemExit(f"{m}: {ins} input boxes per step, the plain form had {plain_counts[0]}")
        ms = 1000.0 * secs / args.steps
        result[m] = {
            "boxes_per_s": ins * args.steps / secs,
            "ms_per_step": ms,
            "input_boxes_per_step": ins,
            "output_boxes_per_step": outs,
            "compulsory_io_gbs": 16.0 * (ins + outs) / (ms * 1e-3) / 1e9,
            "stages": stats,
        }
        if m == "graph":
            gs = prog.graph_stats()
            result[m]["graph"] = gs
            if gs["fallbacks"] or gs["replays"] < args.steps:
                raise SystemExit(f"graph form: the timed steps were not all replays ({gs})")
        if m == "batched":
            # Kernel figures of the program with its chains on one worker
            # stream (one more, untimed step): on four streams each launch
            # shares the GPU with the other chains' kernels, and its events
            # time that sharing, not the kernel.
            workers = eng.get("workers")
            eng.set("workers", 1)
            eng.set("profile", 1)
            if batched_step():
                raise SystemExit("OP1 ring assertion failed in the one-stream batched step")
            eng.sync()
            result[m]["stages_one_stream"] = stage_stats(eng, 1)
            eng.set("profile", 0)
            eng.set("workers", workers)
        print(f"[bench] {m}: {ms:.1f} ms/step", file=sys.stderr, flush=True)

    head = "graph" if "graph" in result else "batched" if "batched" in result else "plain"
    # Stage timings cannot be taken inside graph replays (events recorded by
    # a replay give no usable elapsed times): the graph form's roofline is
    # read from the eager batched form's timed steps, which launch the same
    # kernels with the same sizes.
    rform = "batched" if head == "graph" and "batched" in result else head
    hs = result[rform]["stages"]
    # the roofline names the kernel stage with the largest device time in the
    # timed step (the fused kernel in the batched form, the tile stage in the
    # plain one); its achieved rate = algorithmic bytes / summed duration of
    # its launches, for the batched form from the one-stream step (the kernel
    # alone on the GPU)
    dom = max(("fused", "tile", "sort"), key=lambda k: hs[k]["ms_per_step"])
    d = result[rform].get("stages_one_stream", hs)[dom]
    achieved = d["gbs"]
    per_launch = d["bytes_per_step"] / max(d["launches_per_step"], 1e-9)
    # HBM traffic per launch of the same stage on this workload, from the
    # committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counters cannot be
    # read inside the timed run)
    traffic = None
    traffic_src = f"profiles/{dom}_traffic_stdcell_10M_per_layer.md"
    if npl == 10_000_000:
        traffic = profile_figure(f"{dom}_traffic_stdcell_10M_per_layer.md", "traffic_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": result[head]["boxes_per_s"],
        "unit": "boxes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": result[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded standard-cell generator (seed 0x5EED0002), layers resident in HBM",
        "config": {"workload": "stdcell-4layer-OP1-ring", "boxes_per_layer": npl, "layers": LAYERS,
                   "ring_distance_dbu": 1, "operators_per_step": 24,
                   "input_boxes_per_step": int(plain_counts[0])},
        "plain": result.get("plain"),
        "batched": result.get("batched"),
        "graph": result.get("graph"),
        "autotuned": tuned,
        "batched_over_plain": (result[head]["boxes_per_s"] / result["plain"]["boxes_per_s"]
                               if head != "plain" and "plain" in result else None),
        "roofline": {"bound": "hbm", "kernel": STAGES[dom], "stage": dom,
                     "form": rform + (" (one worker stream)" if rform == "batched" else ""),
                     "ms_per_launch": d["ms_per_step"] / max(d["launches_per_step"], 1e-9),
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": per_launch, "traffic": traffic,
                     "traffic_source": traffic_src if traffic is not None else None,
                     "compulsory_io_frac": result[head]["compulsory_io_gbs"] / HBM_PEAK_GBS},
        "generation_s": t_gen,
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        # the ring of a band of met1 rows, healed first (by the engine's
        # merge, bit-identical to the oracle's canonical form)
        k = 2
        M = V.Layer(eng)
        V.CalculateBoostOperation(rings[k].L, rings[k].L, M, V.K_BOOST_CONTROL_MERGE)
        L = M.boxes()[: args.cpu_sample]
        M.close()
        n = len(L)
        t0 = time.perf_counter()
        L1 = O.par_op("grow", L, args=(1, 1, 1, 1), nthreads=nthreads)
        L2 = O.par_op("size", L, args=(-1,), nthreads=nthreads)
        outer = O.par_op("not", L1, L, nthreads=nthreads)
        inner = O.par_op("not", L, L2, nthreads=nthreads)
        and1 = O.par_op("and", outer, L, nthreads=nthreads)
        ring = O.par_op("and", outer, inner, nthreads=nthreads)
        secs = time.perf_counter() - t0
        if len(and1) or len(ring):
            raise SystemExit("CPU baseline OP1 assertion failed")
        boxes = n + n + (len(L1) + n) + (n + len(L2)) + (len(outer) + n) + (len(outer) + len(inner))
        line["cpu_baseline"] = {"value": boxes / secs, "unit": "boxes/s", "cores": nthreads, "kind": "port",
                                "cpu": cpu_model(),
                                "sample": f"OP1 ring on the first {n} healed boxes of {LAYERS[k]} "
                                          f"(a band of rows of its {npl}), "
                                          f"{secs:.1f} s of CPU work on {nthreads} threads"}
        print(f"[bench] cpu baseline: {secs:.1f} s", file=sys.stderr, flush=True)
    # the configurations below run on engines of their own: the ring's
    # program, layers and engine (with its memory pools) go first
    prog.close()
    for r in rings:
        for lay in (r.L, r.L1, r.L2, r.outer, r.inner, r.and1, r.ring):
            lay.close()
    eng.close()
    del rings
    if args.c2_total:
        # the same ring on the C2 design as 10M boxes in total (2.5M per layer)
        per = args.c2_total // 4
        eng2 = V.Engine(0)
        rings2 = [Ring(V, eng2, lay[:per]) for lay in layers]
        prog2 = V.Program(eng2)
        for r in rings2:
            r.record(V, prog2)
        for r in rings2:
            r.run_plain(V)
        tot2 = sum(r.counts()[0] for r in rings2)
        secs_p, _, _ = time_steps(lambda: [r.run_plain(V) for r in rings2], eng2, args.steps, args.warmup)
        secs_b, _, _ = time_steps(lambda: prog2.run(), eng2, args.steps, args.warmup)
        prog2.set_graph(True)
        secs_g, _, _ = time_steps(lambda: prog2.run(), eng2, args.steps, max(args.warmup, 1))
        if sum(r.counts()[0] for r in rings2) != tot2:
            raise SystemExit("c2: the batched form ran different operators")
        line["c2_10M_total"] = {"boxes_per_layer": per, "input_boxes_per_step": int(tot2),
                                "plain_boxes_per_s": tot2 * args.steps / secs_p,
                                "batched_boxes_per_s": tot2 * args.steps / secs_b,
                                "plain_ms_per_step": 1000.0 * secs_p / args.steps,
                                "batched_ms_per_step": 1000.0 * secs_b / args.steps,
                                "graph_boxes_per_s": tot2 * args.steps / secs_g,
                                "graph_ms_per_step": 1000.0 * secs_g / args.steps}
        print(f"[bench] c2 (2.5M/layer): plain {line['c2_10M_total']['plain_ms_per_step']:.1f} ms, "
              f"batched {line['c2_10M_total']['batched_ms_per_step']:.1f} ms, "
              f"graph {line['c2_10M_total']['graph_ms_per_step']:.1f} ms", file=sys.stderr, flush=True)
        prog2.close()
        eng2.close()
    if not args.no_c3:
        line["c3_uneven_tiles"] = c3_uneven_tiles(V, args)
        c3 = line["c3_uneven_tiles"]
        print(f"[bench] c3: plain {c3['plain_ms_per_step']:.1f} ms, batched {c3['batched_ms_per_step']:.1f} ms, "
              f"graph {c3['graph_ms_per_step']:.1f} ms", file=sys.stderr, flush=True)
    if args.c4_boxes:
        line["c4_full_chip"] = c4_full_chip(V, args)
        print(f"[bench] c4: {line['c4_full_chip']['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)
    if args.design_io_boxes:
        line["design_io"] = design_io(V, args)
        dio = line["design_io"]
        print(f"[bench] design_io: SNAP load {dio['load_snap_ms']:.1f} ms "
              f"(GPU decode {dio['gpu_decode_ms']:.2f} ms, {dio['gpu_decode_gbs']:.0f} GB/s), "
              f"KLBIN load {dio['load_klbin_ms']:.1f} ms", file=sys.stderr, flush=True)
    if args.serving_boxes:
        line["serving"] = serving(V, args)
        print(f"[bench] serving: p50 {line['serving']['p50_ms']:.2f} ms, p99 {line['serving']['p99_ms']:.2f} ms",
              file=sys.stderr, flush=True)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
