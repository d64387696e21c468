"""vlsi_layout_amd — MI355X (gfx950) engine for the rectangle-Boolean `exec run`
operators of skoranne/VLSI_LAYOUT, driven from Python through its C ABI.

The compute runs in the HIP library ``libvlayout.so`` built from ``csrc/``
(``python -m vlsi_layout_amd.build``).  This module is a thin ctypes layer over
``include/vl_api.h``; it mirrors the reference's operator surface
(src/design.f90:20-31, src/design_impl.f90, src/controlparser.f90:409-585):

    CalculateAND(A, B, C)            '*'   src/design.f90:1131-1235
    CalculateFrameNOT(A, B, C)       '~'   src/design_impl.f90:1152-1164
    CalculateNOT(A, B, C)            '@'   src/design.f90:1238-1268
    CalculateOR(A, B, C)             OR2   src/design.f90:1085-1103
    CalculateXOR(A, B, C)            XOR   src/design_impl.f90:1166-1177
    CalculateGROWLayer(A, C, ivar)   GROW  src/design_impl.f90:571-586 (ivar = E,N,W,S)
    CalculateBoostOperation(A, B, C, ctrl, val)   src/design_impl.f90:931-951
    MergeHealLayer(L)                      src/design_impl.f90:1318-1341
    heal_boxes(boxes)                      src/polygon_fracture.f90:805-1004
    device_event_sort / thrust_sort_boxes_direct / thrust_sort_boxes_indirect

Layers live in device memory; ``Layer.boxes()`` copies them to the host as an
(n, 4) int32 numpy array of (x1, y1, x2, y2).  There is no CPU fallback: if
the library is missing every call raises.
"""
import ctypes as C
import os
import threading

import numpy as np

__all__ = [
    "VLError", "Engine", "Layer", "lib", "engine",
    "CalculateAND", "CalculateFrameNOT", "CalculateNOT", "CalculateOR", "CalculateXOR",
    "CalculateGROWLayer", "CalculateBoostOperation", "MergeHealLayer", "heal_boxes",
    "CalculateFrameAND", "CalculateSingleLayerAND", "CreateGrid", "CreateEXTENT",
    "PerformMerge", "InteractionCount", "CertifyHealing",
    "ASSERT_ZERO", "klee_area", "canonicalize", "xor_empty",
    "device_event_sort", "thrust_sort_boxes_direct", "thrust_sort_boxes_indirect",
    "synth_layer", "synth_tile_count",
    "K_BOOST_CONTROL_XOR", "K_BOOST_CONTROL_OR", "K_BOOST_CONTROL_AND", "K_BOOST_CONTROL_NOT",
    "K_BOOST_CONTROL_MERGE", "K_BOOST_CONTROL_SIZE",
]

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libvlayout.so")

# Boost control codes (src/commonmodule.f90:12-14)
K_BOOST_CONTROL_XOR, K_BOOST_CONTROL_OR, K_BOOST_CONTROL_AND = 0, 1, 2
K_BOOST_CONTROL_NOT, K_BOOST_CONTROL_MERGE, K_BOOST_CONTROL_SIZE = 3, 4, 5
VL_OP_GROW = 6

VL_OK, VL_EINVAL, VL_ENOSPC, VL_ENOMEM, VL_EHIP, VL_EASSERT, VL_EINTERNAL = 0, 1, 2, 3, 4, 5, 7

SYN_UNIFORM, SYN_STDCELL, SYN_TILES, SYN_CHIP, SYN_ROUTE = 1, 2, 3, 4, 5

# vl_event_sort_noalloc's static scratch bound (include/vl_api.h)
VL_NOALLOC_MAX_TRACKERS = 3500000


class VLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_lib_lock = threading.Lock()


class _Tracker(C.Structure):
    _fields_ = [("X", C.c_int32), ("Y", C.c_int32), ("polygonNumber", C.c_int8)]


TRACKER_DTYPE = np.dtype({"names": ["X", "Y", "polygonNumber"], "formats": [np.int32, np.int32, np.int8],
                          "offsets": [0, 4, 8], "itemsize": 12})


def lib():
    """The loaded engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise VLError(VL_EINVAL, f"{LIB_PATH} is not built; run python -m vlsi_layout_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        sig = {
            "vl_last_error": (C.c_char_p, []),
            "vl_engine_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "vl_engine_destroy": (C.c_int, [vp]),
            "vl_engine_set_param": (C.c_int, [vp, C.c_char_p, i64]),
            "vl_engine_get_param": (C.c_int, [vp, C.c_char_p, C.POINTER(i64)]),
            "vl_engine_sync": (C.c_int, [vp]),
            "vl_engine_last_time_ms": (C.c_int, [vp, C.POINTER(C.c_double)]),
            "vl_engine_kernel_stats": (C.c_int, [vp, C.c_char_p, C.POINTER(i64), C.POINTER(C.c_double),
                                                 C.POINTER(C.c_double)]),
            "vl_layer_create": (C.c_int, [vp, C.POINTER(vp)]),
            "vl_layer_destroy": (C.c_int, [vp]),
            "vl_layer_upload": (C.c_int, [vp, vp, i64]),
            "vl_layer_upload_device": (C.c_int, [vp, vp, i64]),
            "vl_layer_download": (C.c_int, [vp, vp, i64, C.POINTER(i64)]),
            "vl_layer_size": (C.c_int, [vp, C.POINTER(i64)]),
            "vl_layer_device_ptr": (C.c_int, [vp, C.POINTER(vp)]),
            "vl_and": (C.c_int, [vp, vp, vp, vp]),
            "vl_not": (C.c_int, [vp, vp, vp, vp]),
            "vl_or": (C.c_int, [vp, vp, vp, vp]),
            "vl_xor": (C.c_int, [vp, vp, vp, vp]),
            "vl_grow": (C.c_int, [vp, vp, i32, i32, i32, i32, vp]),
            "vl_size": (C.c_int, [vp, vp, i32, vp]),
            "vl_merge": (C.c_int, [vp, vp, vp]),
            "vl_self_and": (C.c_int, [vp, vp, vp]),
            "vl_grid": (C.c_int, [vp, vp, i32, i32, i32, vp]),
            "vl_extent": (C.c_int, [vp, vp, vp]),
            "vl_pnum": (C.c_int, [vp, vp, vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
            "vl_interaction_count": (C.c_int, [vp, vp, C.POINTER(i64)]),
            "vl_layer_load_klbin": (C.c_int, [vp, vp, C.c_char_p]),
            "vl_layer_save_klbin": (C.c_int, [vp, vp, C.c_char_p]),
            "vl_layer_save_snap": (C.c_int, [vp, vp, C.c_char_p, i32, i32, vp]),
            "vl_snap_write_boxes": (C.c_int, [vp, i64, C.c_char_p, i32, i32, vp]),
            "vl_snap_info": (C.c_int, [C.c_char_p, C.POINTER(i32), C.POINTER(i64), C.POINTER(i64), C.POINTER(i32)]),
            "vl_layer_load_snap": (C.c_int, [vp, vp, C.c_char_p, C.POINTER(i32), vp, i64]),
            "vl_layer_load_file": (C.c_int, [vp, vp, C.c_char_p, C.POINTER(i32)]),
            "vl_codec_encode": (C.c_int, [vp, i64, vp, i64, C.POINTER(i64)]),
            "vl_codec_decode": (C.c_int, [vp, vp, vp, vp, i64, vp]),
            "vl_assert_zero": (C.c_int, [vp, vp, C.c_int, C.POINTER(i64)]),
            "vl_klee_area": (C.c_int, [vp, vp, C.POINTER(i64)]),
            "vl_canonicalize": (C.c_int, [vp, vp, vp]),
            "vl_xor_empty": (C.c_int, [vp, vp, vp, C.POINTER(C.c_int)]),
            "vl_boolean_op": (C.c_int, [vp, vp, C.c_uint64, vp, C.c_uint64, vp, C.c_uint64,
                                        C.POINTER(C.c_uint64), C.c_uint64, i64]),
            "vl_event_sort": (C.c_int, [vp, vp, i64]),
            "vl_event_sort_noalloc": (C.c_int, [vp, vp, i64]),
            "vl_morton_sort_direct": (C.c_int, [vp, vp, i64]),
            "vl_morton_sort_indirect": (C.c_int, [vp, vp, vp, i64]),
            "device_event_sort": (None, [vp, i64]),
            "thrust_sort_boxes_direct": (None, [vp, i64]),
            "thrust_sort_boxes_indirect": (None, [vp, vp, i64]),
            "vl_program_create": (C.c_int, [vp, C.POINTER(vp)]),
            "vl_program_destroy": (C.c_int, [vp]),
            "vl_program_add": (C.c_int, [vp, C.c_int, vp, vp, vp, vp]),
            "vl_program_run": (C.c_int, [vp]),
            "vl_program_assert_count": (C.c_int, [vp, C.POINTER(i64)]),
            "vl_program_set_graph": (C.c_int, [vp, C.c_int]),
            "vl_program_graph_stats": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
            "vl_engine_autotune": (C.c_int, [vp, vp, vp, C.POINTER(i64), C.POINTER(i64)]),
            "vl_server_create": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
            "vl_server_run": (C.c_int, [vp, vp, i64, vp]),
            "vl_server_destroy": (C.c_int, [vp]),
            "vl_synth_layer": (C.c_int, [C.c_int, C.c_uint64, C.c_int, i64, vp]),
            "vl_synth_tile_count": (i64, [C.c_uint64, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def _check(rc):
    if rc != VL_OK:
        msg = lib().vl_last_error()
        raise VLError(rc, msg.decode() if msg else "")


def as_boxes(a):
    """(n, 4) int32 C-contiguous array of (x1, y1, x2, y2)."""
    a = np.asarray(a, dtype=np.int32)
    return np.ascontiguousarray(a.reshape(-1, 4))


class Engine:
    """One device, one HIP stream, a scratch pool and tuning parameters."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().vl_engine_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().vl_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, key, value):
        _check(lib().vl_engine_set_param(self._h, key.encode(), int(value)))

    def get(self, key):
        v = C.c_int64()
        _check(lib().vl_engine_get_param(self._h, key.encode(), C.byref(v)))
        return v.value

    def sync(self):
        _check(lib().vl_engine_sync(self._h))

    def last_ms(self):
        v = C.c_double()
        _check(lib().vl_engine_last_time_ms(self._h, C.byref(v)))
        return v.value

    def kernel_stats(self, kernel):
        """(launches, total_ms, algorithmic_bytes) of a kernel while profile=1."""
        n, ms, by = C.c_int64(), C.c_double(), C.c_double()
        _check(lib().vl_engine_kernel_stats(self._h, kernel.encode(), C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value

    def layer(self, boxes=None):
        return Layer(self, boxes)

    def autotune(self, a, b=None):
        """Picks tile_target, erode_target, fuse_target, fuse, fuse2 (with b), bin_mode, radix_bits
        and max_wg by timing candidates on layers a (and b); returns the chosen values."""
        t, et = C.c_int64(), C.c_int64()
        _check(lib().vl_engine_autotune(self._h, a.handle, b.handle if b is not None else None, C.byref(t),
                                        C.byref(et)))
        return {"tile_target": t.value, "erode_target": et.value, "fuse_target": self.get("fuse_target"),
                "fuse": self.get("fuse"), "fuse2": self.get("fuse2"), "bin_mode": self.get("bin_mode"),
                "radix_bits": self.get("radix_bits"),
                "max_wg": self.get("max_wg")}


class Layer:
    """A device-resident box layer (the reference's type(Layer), src/design.f90:65-82)."""

    def __init__(self, eng, boxes=None):
        self.engine = eng
        h = C.c_void_p()
        _check(lib().vl_layer_create(eng.handle, C.byref(h)))
        self._h = h
        if boxes is not None:
            self.upload(boxes)

    @property
    def handle(self):
        return self._h

    def upload(self, boxes):
        b = as_boxes(boxes)
        _check(lib().vl_layer_upload(self._h, b.ctypes.data_as(C.c_void_p), len(b)))
        return self

    @property
    def n_used(self):
        v = C.c_int64()
        _check(lib().vl_layer_size(self._h, C.byref(v)))
        return v.value

    def __len__(self):
        return self.n_used

    def boxes(self):
        n = self.n_used
        out = np.zeros((n, 4), np.int32)
        got = C.c_int64()
        _check(lib().vl_layer_download(self._h, out.ctypes.data_as(C.c_void_p), n, C.byref(got)))
        return out

    def copy_from(self, other, n=None):
        """Device-to-device copy of the first n boxes of another layer of the same device."""
        n = other.n_used if n is None else min(int(n), other.n_used)
        _check(lib().vl_layer_upload_device(self._h, other.device_ptr(), n))
        return self

    def device_ptr(self):
        p = C.c_void_p()
        _check(lib().vl_layer_device_ptr(self._h, C.byref(p)))
        return p.value

    def close(self):
        if getattr(self, "_h", None):
            lib().vl_layer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def engine():
    """Process-wide engine on device 0."""
    global _default
    if _default is None:
        _default = Engine(0)
    return _default


def _binop(fn, A, B, Cl):
    _check(getattr(lib(), fn)(A.engine.handle, A.handle, B.handle, Cl.handle))
    return Cl


def CalculateAND(A, B, C_out):
    """C = A AND B  ('*', src/design.f90:1131-1235)."""
    return _binop("vl_and", A, B, C_out)


def CalculateFrameNOT(A, B, C_out):
    """C = A NOT B  ('~', src/design_impl.f90:1152-1164)."""
    return _binop("vl_not", A, B, C_out)


def CalculateNOT(A, B, C_out):
    """C = A NOT B  ('@', src/design.f90:1238-1268): same point set as '~'."""
    return _binop("vl_not", A, B, C_out)


def CalculateOR(A, B, C_out):
    """C = A OR B  (OR2, src/design.f90:1085-1103; '+' Boost OR)."""
    return _binop("vl_or", A, B, C_out)


def CalculateXOR(A, B, C_out):
    """C = A XOR B  (XOR / '^', src/design_impl.f90:1166-1177)."""
    return _binop("vl_xor", A, B, C_out)


def CalculateGROWLayer(A, C_out, ivar):
    """C = GROW(A; E, N, W, S) in DBU (src/design_impl.f90:571-586)."""
    e, n, w, s = (int(v) for v in ivar)
    _check(lib().vl_grow(A.engine.handle, A.handle, e, n, w, s, C_out.handle))
    return C_out


def CalculateBoostOperation(A, B, C_out, control_parameter, control_value=0):
    """Boost-bridge operator codes (src/design_impl.f90:931-951)."""
    e = A.engine.handle
    cp = int(control_parameter)
    if cp == K_BOOST_CONTROL_XOR:
        return CalculateXOR(A, B, C_out)
    if cp == K_BOOST_CONTROL_OR:
        return CalculateOR(A, B, C_out)
    if cp == K_BOOST_CONTROL_AND:
        return CalculateAND(A, B, C_out)
    if cp == K_BOOST_CONTROL_NOT:
        return CalculateFrameNOT(A, B, C_out)
    if cp == K_BOOST_CONTROL_MERGE:
        _check(lib().vl_merge(e, A.handle, C_out.handle))
        return C_out
    if cp == K_BOOST_CONTROL_SIZE:
        _check(lib().vl_size(e, A.handle, int(control_value), C_out.handle))
        return C_out
    raise VLError(VL_EINVAL, f"unknown control parameter {control_parameter}")


def CalculateFrameAND(A, B, C_out):
    """C = A AND B  (AND2, src/design.f90:296-299): same point set as '*'."""
    return _binop("vl_and", A, B, C_out)


def CalculateSingleLayerAND(A, C_out):
    """C = the points covered by at least two boxes of A  (the one-input `AND`,
    src/design_impl.f90:385-509)."""
    _check(lib().vl_self_and(A.engine.handle, A.handle, C_out.handle))
    return C_out


def CreateGrid(A, C_out, rows, cols, grid_overlap):
    """C = rows x cols padded sub-boxes of MBR(A)  (src/design_impl.f90:159-275)."""
    _check(lib().vl_grid(A.engine.handle, A.handle, int(rows), int(cols), int(grid_overlap), C_out.handle))
    return C_out


def CreateEXTENT(A, C_out):
    """C = MBR(A) as one box  (src/design_impl.f90:277-293)."""
    _check(lib().vl_extent(A.engine.handle, A.handle, C_out.handle))
    return C_out


def PerformMerge(L):
    """Polygon numbering of L's boxes in their stored order (src/pnummerge.f90:279-385):
    returns (labels int64[n], components, overlap_area, touch_perimeter); labels
    are 0 for boxes that meet no other box, else 1..components."""
    n = L.n_used
    labels = np.zeros(max(n, 1), np.int64)
    nc, ar, pe = C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib().vl_pnum(L.engine.handle, L.handle, labels.ctypes.data_as(C.c_void_p), C.byref(nc), C.byref(ar),
                         C.byref(pe)))
    return labels[:n], nc.value, ar.value, pe.value


def InteractionCount(L):
    """Pairs of L's boxes overlapping with positive area (src/interaction.f90:204-310)."""
    v = C.c_int64()
    _check(lib().vl_interaction_count(L.engine.handle, L.handle, C.byref(v)))
    return v.value


def CertifyHealing(L):
    """True when no two boxes of L overlap with positive area (src/design_impl.f90:35-56)."""
    return L.n_used < 2 or InteractionCount(L) == 0


COMPRESSION_METHOD_SNAPPY, COMPRESSION_METHOD_ZLIB, COMPRESSION_METHOD_ZSTD = 1, 2, 3


def _path(p):
    return os.fsencode(os.fspath(p))


def LoadKLBin(path, L):
    """Replace L by the boxes of a KLBIN file (src/kldata.f90:14-63)."""
    _check(lib().vl_layer_load_klbin(L.engine.handle, L.handle, _path(path)))
    return L


def WriteKLBin(path, L):
    """Write L's boxes as a KLBIN file (src/kldata.f90:64-97)."""
    _check(lib().vl_layer_save_klbin(L.engine.handle, L.handle, _path(path)))


def SaveLayerToSnap(L, path, method=COMPRESSION_METHOD_SNAPPY, layer_properties=0, pnum=None):
    """Write L as a SNAP file (src/diskdesign_impl.f90:59-81); pnum: optional
    int64 polygon numbers, one per box, stored after the header."""
    ptr = None
    if pnum is not None:
        pnum = np.ascontiguousarray(pnum, dtype=np.int64)
        if pnum.shape != (L.n_used,):
            raise VLError(VL_EINVAL, f"pnum must hold {L.n_used} values")
        ptr = pnum.ctypes.data_as(C.c_void_p)
    _check(lib().vl_layer_save_snap(L.engine.handle, L.handle, _path(path), int(method), int(layer_properties),
                                    ptr))


def snap_write_boxes(boxes, path, method=COMPRESSION_METHOD_SNAPPY, layer_properties=0, pnum=None):
    """SNAP file from host boxes (no GPU), as SaveLayerToSnap writes it."""
    b = as_boxes(boxes)
    ptr = None
    if pnum is not None:
        pnum = np.ascontiguousarray(pnum, dtype=np.int64)
        if pnum.shape != (len(b),):
            raise VLError(VL_EINVAL, f"pnum must hold {len(b)} values")
        ptr = pnum.ctypes.data_as(C.c_void_p)
    _check(lib().vl_snap_write_boxes(b.ctypes.data_as(C.c_void_p), len(b), _path(path), int(method),
                                     int(layer_properties), ptr))


def SnapInfo(path):
    """(layer_properties, total_boxes, chunks, method) of a SNAP file's header."""
    pr, nb, nc, m = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32()
    _check(lib().vl_snap_info(_path(path), C.byref(pr), C.byref(nb), C.byref(nc), C.byref(m)))
    return pr.value, nb.value, nc.value, m.value


def RestoreSnapToLayer(L, path, with_pnum=False):
    """Replace L by a SNAP or KLBIN file's boxes (src/diskdesign_impl.f90:83-114;
    a path containing ".bin" is KLBIN).  Returns the layer properties, or
    (properties, pnum) with with_pnum (pnum None when the file has none)."""
    pr = C.c_int32()
    if not with_pnum:
        _check(lib().vl_layer_load_file(L.engine.handle, L.handle, _path(path), C.byref(pr)))
        return pr.value
    props, total, _, _ = SnapInfo(path)
    pn = np.zeros(max(total, 1), np.int64) if props & 4 else None
    _check(lib().vl_layer_load_snap(L.engine.handle, L.handle, _path(path), C.byref(pr),
                                    None if pn is None else pn.ctypes.data_as(C.c_void_p),
                                    0 if pn is None else len(pn)))
    return pr.value, (None if pn is None else pn[:total])


def codec_encode(boxes):
    """Box codec bytes of one chunk (<= 65536 boxes; src/box_codec.f90:565-666)."""
    b = as_boxes(boxes)
    cap = 41 * len(b) + 16
    out = np.zeros(cap, np.uint8)
    n = C.c_int64()
    _check(lib().vl_codec_encode(b.ctypes.data_as(C.c_void_p), len(b), out.ctypes.data_as(C.c_void_p), cap,
                                 C.byref(n)))
    return out[:n.value].tobytes()


def codec_decode(chunks, counts, out):
    """GPU decode of uncompressed codec chunks (bytes objects) holding
    counts[k] boxes each into layer `out` (src/box_codec.f90:447-536)."""
    raw = np.frombuffer(b"".join(chunks) or b"\0", np.uint8)
    cb = np.array([len(c) for c in chunks] or [0], np.int64)
    nb = np.array(list(counts) or [0], np.int64)
    _check(lib().vl_codec_decode(out.engine.handle, raw.ctypes.data_as(C.c_void_p), cb.ctypes.data_as(C.c_void_p),
                                 nb.ctypes.data_as(C.c_void_p), len(chunks), out.handle))
    return out


def MergeHealLayer(L):
    """Replace L by the maximal-strip union of its boxes (src/design_impl.f90:1318-1341)."""
    _check(lib().vl_merge(L.engine.handle, L.handle, L.handle))
    return L


def ASSERT_ZERO(L, abort=False):
    """Box count of L; raises VLError(VL_EASSERT) when abort and non-empty
    (src/controlparser.f90:524-530)."""
    v = C.c_int64()
    _check(lib().vl_assert_zero(L.engine.handle, L.handle, 1 if abort else 0, C.byref(v)))
    return v.value


def klee_area(L):
    v = C.c_int64()
    _check(lib().vl_klee_area(L.engine.handle, L.handle, C.byref(v)))
    return v.value


def canonicalize(L, out=None):
    out = out if out is not None else Layer(L.engine)
    _check(lib().vl_canonicalize(L.engine.handle, L.handle, out.handle))
    return out


def xor_empty(A, B):
    v = C.c_int()
    _check(lib().vl_xor_empty(A.engine.handle, A.handle, B.handle, C.byref(v)))
    return bool(v.value)


def heal_boxes(boxes, eng=None):
    """heal_boxes on host boxes (src/polygon_fracture.f90:805-1004): returns the
    maximal-strip union as an (m, 4) array."""
    eng = eng or engine()
    L = Layer(eng, boxes)
    MergeHealLayer(L)
    return L.boxes()


def device_event_sort(trackers, eng=None):
    """Stable sort of XYTracker records by (X, Y) (src/event_sorter.cu:48-81).
    trackers: numpy array with TRACKER_DTYPE (12-byte records); sorted in place."""
    t = np.ascontiguousarray(trackers)
    if t.dtype != TRACKER_DTYPE:
        raise VLError(VL_EINVAL, "trackers must use TRACKER_DTYPE")
    eng = eng or engine()
    _check(lib().vl_event_sort(eng.handle, t.ctypes.data_as(C.c_void_p), len(t)))
    return t


def event_sort_noalloc(dev_ptr, n, eng=None):
    """vl_event_sort_noalloc on n trackers already in device memory at dev_ptr
    (sorted in place; no device allocation; n <= VL_NOALLOC_MAX_TRACKERS)."""
    eng = eng or engine()
    _check(lib().vl_event_sort_noalloc(eng.handle, C.c_void_p(dev_ptr), int(n)))


def thrust_sort_boxes_direct(boxes, eng=None):
    """Morton-sort boxes by centre (src/gpu_morton.cu:149-164); returns a sorted copy."""
    b = as_boxes(boxes).copy()
    eng = eng or engine()
    _check(lib().vl_morton_sort_direct(eng.handle, b.ctypes.data_as(C.c_void_p), len(b)))
    return b


def thrust_sort_boxes_indirect(boxes, eng=None):
    """1-based Morton permutation (src/gpu_morton.cu:169-187)."""
    b = as_boxes(boxes)
    idx = np.zeros(len(b), np.int64)
    eng = eng or engine()
    _check(lib().vl_morton_sort_indirect(eng.handle, b.ctypes.data_as(C.c_void_p),
                                         idx.ctypes.data_as(C.c_void_p), len(b)))
    return idx


class Program:
    """Recorded operator list run as one batch (vl_program_*)."""

    def __init__(self, eng):
        self.engine = eng
        h = C.c_void_p()
        _check(lib().vl_program_create(eng.handle, C.byref(h)))
        self._h = h
        self._keep = []

    def add(self, op, a, b=None, c=None, args=None):
        arr = None
        if args is not None:
            arr = (C.c_int32 * 4)(*([int(v) for v in args] + [0] * (4 - len(args))))
            self._keep.append(arr)
        _check(lib().vl_program_add(self._h, int(op), a.handle, b.handle if b is not None else None,
                                    c.handle if c is not None else None, arr))
        # the library holds its own reference to each layer; the Python
        # objects are kept too, so their handles stay usable by the caller
        self._keep.extend(x for x in (a, b, c) if x is not None)

    def run(self):
        _check(lib().vl_program_run(self._h))

    def assert_failures(self):
        v = C.c_int64()
        _check(lib().vl_program_assert_count(self._h, C.byref(v)))
        return v.value

    def set_graph(self, enable=True):
        """Capture the program as a HIP graph on its next run and replay it afterwards."""
        _check(lib().vl_program_set_graph(self._h, 1 if enable else 0))

    def graph_stats(self):
        """{'captures', 'replays', 'fallbacks'} over the program's life."""
        c, r, f = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().vl_program_graph_stats(self._h, C.byref(c), C.byref(r), C.byref(f)))
        return {"captures": c.value, "replays": r.value, "fallbacks": f.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().vl_program_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


QUERY_DTYPE = np.dtype([("x1", "<i4"), ("y1", "<i4"), ("x2", "<i4"), ("y2", "<i4"), ("layer_a", "<i4"),
                        ("layer_b", "<i4"), ("spacing", "<i4"), ("reserved", "<i4"), ("arrival_ms", "<f8")])
RESULT_DTYPE = np.dtype([("status", "<i4"), ("slot", "<i4"), ("in_boxes", "<i8"), ("count", "<i8"),
                         ("hash", "<u8"), ("start_ms", "<f8"), ("end_ms", "<f8")])


class Server:
    """Windowed-query server over resident layers (vl_server_*): answers
    R = GROW_s(A & W) AND (B & W) with continuous batching over `slots`
    concurrent queries and an HBM admission budget."""

    def __init__(self, eng, layers, slots=4, hbm_budget=0):
        self.engine = eng
        arr = (C.c_void_p * len(layers))(*[l.handle.value for l in layers])
        h = C.c_void_p()
        _check(lib().vl_server_create(eng.handle, arr, len(layers), int(slots), int(hbm_budget), C.byref(h)))
        self._h = h

    def run(self, queries):
        """queries: QUERY_DTYPE array; returns (status code, RESULT_DTYPE array)."""
        q = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        res = np.zeros(len(q), RESULT_DTYPE)
        rc = lib().vl_server_run(self._h, q.ctypes.data_as(C.c_void_p), len(q), res.ctypes.data_as(C.c_void_p))
        return rc, res

    def close(self):
        if getattr(self, "_h", None):
            lib().vl_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_layer(config, seed, layer, n):
    out = np.zeros((int(n), 4), np.int32)
    rc = lib().vl_synth_layer(int(config), int(seed), int(layer), int(n), out.ctypes.data_as(C.c_void_p))
    if rc != VL_OK:
        raise VLError(rc, "vl_synth_layer: bad arguments")
    return out


def synth_tile_count(seed, tile):
    return int(lib().vl_synth_tile_count(int(seed), int(tile)))
