"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the flang build of
the reference's own sweep routines (oracle/_ref/libvlref.so).

Test infrastructure only.  Boxes are numpy int32 arrays of shape (n, 4) holding
(x1, y1, x2, y2), the reference Box layout (src/geometry.f90:21-22).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libvlref.so")

_lib = None
_ref = None


class OBox(C.Structure):
    _fields_ = [("x1", C.c_int32), ("y1", C.c_int32), ("x2", C.c_int32), ("y2", C.c_int32)]


def boxes(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1, 4))
    return a


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        PB = C.POINTER(C.c_void_p)
        for name in ("or_heal", "or_canonical"):
            f = getattr(L, name)
            f.restype = C.c_int64
            f.argtypes = [C.c_void_p, C.c_int64, PB]
        L.or_complement.restype = C.c_int64
        L.or_complement.argtypes = [C.c_void_p, C.c_int64, OBox, PB]
        L.or_klee_area.restype = C.c_int64
        L.or_klee_area.argtypes = [C.c_void_p, C.c_int64]
        for name in ("or_and", "or_not", "or_or", "or_xor", "or_and_fragments"):
            f = getattr(L, name)
            f.restype = C.c_int64
            f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, PB]
        L.or_grow.restype = C.c_int64
        L.or_grow.argtypes = [C.c_void_p, C.c_int64] + [C.c_int32] * 4 + [PB]
        L.or_size.restype = C.c_int64
        L.or_size.argtypes = [C.c_void_p, C.c_int64, C.c_int32, PB]
        L.or_self_and.restype = C.c_int64
        L.or_self_and.argtypes = [C.c_void_p, C.c_int64, PB]
        L.or_grid.restype = C.c_int64
        L.or_grid.argtypes = [OBox, C.c_int32, C.c_int32, C.c_int32, PB]
        L.or_pnum.restype = C.c_int64
        L.or_pnum.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.or_interactions.restype = C.c_int64
        L.or_interactions.argtypes = [C.c_void_p, C.c_int64]
        L.or_codec_encode.restype = C.c_int64
        L.or_codec_encode.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.or_codec_decode.restype = C.c_int
        L.or_codec_decode.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
        L.or_free.argtypes = [C.c_void_p]
        L.or_event_sort.argtypes = [C.c_void_p, C.c_int64]
        L.or_morton_key.restype = C.c_uint64
        L.or_morton_key.argtypes = [OBox]
        L.or_morton_sort_direct.argtypes = [C.c_void_p, C.c_int64]
        L.or_morton_sort_indirect.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.or_par_op.restype = C.c_int64
        L.or_par_op.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                C.c_void_p, C.c_int, PB]
        _lib = L
    return _lib


def ref():
    """The reference's own routines, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO)
        P64 = C.POINTER(C.c_int64)
        L.ref_heal_boxes.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, P64]
        L.ref_complement.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, P64]
        L.ref_union_area.restype = C.c_int64
        L.ref_union_area.argtypes = [C.c_int64, C.c_void_p]
        L.ref_pnum.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ref_interactions.restype = C.c_int64
        L.ref_interactions.argtypes = [C.c_int64, C.c_void_p]
        _ref = L
    return _ref


def _take(n, p):
    if n <= 0:
        if p.value:
            lib().or_free(p)
        return np.zeros((0, 4), np.int32)
    buf = (C.c_int32 * (4 * n)).from_address(p.value)
    out = np.frombuffer(buf, dtype=np.int32).reshape(n, 4).copy()
    lib().or_free(p)
    return out


def _unary(name, a, *extra):
    a = boxes(a)
    p = C.c_void_p()
    n = getattr(lib(), name)(_ptr(a), len(a), *extra, C.byref(p))
    if n < 0:
        raise ValueError(f"{name}: invalid box produced")
    return _take(n, p)


def _binary(name, a, b):
    a, b = boxes(a), boxes(b)
    p = C.c_void_p()
    n = getattr(lib(), name)(_ptr(a), len(a), _ptr(b), len(b), C.byref(p))
    return _take(n, p)


def heal(a):
    return _unary("or_heal", a)


def canonical(a):
    return _unary("or_canonical", a)


def complement(b, bbox):
    b = boxes(b)
    p = C.c_void_p()
    n = lib().or_complement(_ptr(b), len(b), OBox(*[int(v) for v in bbox]), C.byref(p))
    return _take(n, p)


def klee_area(a):
    a = boxes(a)
    return int(lib().or_klee_area(_ptr(a), len(a)))


def op_and(a, b):
    return _binary("or_and", a, b)


def op_not(a, b):
    return _binary("or_not", a, b)


def op_or(a, b):
    return _binary("or_or", a, b)


def op_xor(a, b):
    return _binary("or_xor", a, b)


def and_fragments(a, b):
    return _binary("or_and_fragments", a, b)


def self_and(a):
    return _unary("or_self_and", a)


def pnum(a):
    """(labels, components, overlap_area, touch_perimeter) of PerformMerge."""
    a = boxes(a)
    labels = np.zeros(max(len(a), 1), np.int64)
    ar, pe = C.c_int64(), C.c_int64()
    nc = lib().or_pnum(_ptr(a), len(a), _ptr(labels), C.byref(ar), C.byref(pe))
    return labels[: len(a)], int(nc), ar.value, pe.value


def codec_encode(a):
    """SNAP codec bytes of one chunk of boxes (src/box_codec.f90:565-666)."""
    a = boxes(a)
    out = np.zeros(48 * len(a) + 16, np.uint8)
    n = lib().or_codec_encode(_ptr(a), len(a), _ptr(out), len(out))
    assert n >= 0
    return out[:n].tobytes()


def codec_decode(data, n):
    """Boxes of one SNAP codec chunk of n boxes, or None when malformed."""
    buf = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros((max(n, 1), 4), np.int32)
    rc = lib().or_codec_decode(_ptr(buf), len(data), n, _ptr(out))
    return None if rc else out[:n]


def interactions(a):
    a = boxes(a)
    return int(lib().or_interactions(_ptr(a), len(a)))


def grid(mbr, rows, cols, overlap):
    p = C.c_void_p()
    n = lib().or_grid(OBox(*[int(v) for v in mbr]), int(rows), int(cols), int(overlap), C.byref(p))
    if n < 0:
        raise ValueError("grid: bad arguments")
    return _take(n, p)


def grow(a, e, n, w, s):
    return _unary("or_grow", a, int(e), int(n), int(w), int(s))


def size(a, d):
    return _unary("or_size", a, int(d))


def event_sort(trackers):
    """trackers: structured array with fields X, Y (int32), w (int8), 12-byte records."""
    t = np.ascontiguousarray(trackers)
    lib().or_event_sort(_ptr(t), len(t))
    return t


def morton_key(b):
    return int(lib().or_morton_key(OBox(*[int(v) for v in b])))


def morton_sort_direct(a):
    a = boxes(a).copy()
    lib().or_morton_sort_direct(_ptr(a), len(a))
    return a


def morton_sort_indirect(a):
    a = boxes(a)
    idx = np.zeros(len(a), np.int64)
    lib().or_morton_sort_indirect(_ptr(a), _ptr(idx), len(a))
    return idx


OPCODE = {"xor": 0, "or": 1, "and": 2, "not": 3, "merge": 4, "size": 5, "grow": 6}


def par_op(op, a, b=None, args=(0, 0, 0, 0), nthreads=0):
    """CPU baseline operator (tile-parallel restatement of the reference frame functions)."""
    a = boxes(a)
    b = boxes(b if b is not None else np.zeros((0, 4), np.int32))
    ar = np.asarray(list(args) + [0] * (4 - len(args)), np.int32)
    p = C.c_void_p()
    n = lib().or_par_op(OPCODE[op], _ptr(a), len(a), _ptr(b), len(b), _ptr(ar), int(nthreads), C.byref(p))
    if n < 0:
        raise ValueError("par_op: invalid box produced")
    return _take(n, p)


# ---- the reference's own routines (flang build) -------------------------------------------
def ref_heal(a):
    L = ref()
    a = boxes(a)
    cap = max(16, 8 * len(a) + 16)
    out = np.zeros((cap, 4), np.int32)
    nout = C.c_int64()
    L.ref_heal_boxes(len(a), _ptr(a), cap, _ptr(out), C.byref(nout))
    assert nout.value <= cap
    return out[: nout.value].copy()


def ref_complement(b, bbox):
    L = ref()
    b = boxes(b)
    bb = np.asarray(bbox, np.int32)
    cap = max(64, 16 * len(b) + 64)
    out = np.zeros((cap, 4), np.int32)
    nout = C.c_int64()
    L.ref_complement(len(b), _ptr(b), _ptr(bb), cap, _ptr(out), C.byref(nout))
    assert nout.value <= cap
    return out[: nout.value].copy()


def ref_union_area(a):
    a = boxes(a)
    return int(ref().ref_union_area(len(a), _ptr(a)))


def sort_boxes(a):
    """Sort rows by (x1, y1, x2, y2)."""
    a = boxes(a)
    if len(a) == 0:
        return a
    order = np.lexsort((a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    return a[order]


def ref_pnum(a):
    """PerformMerge of the reference (n >= 1): (labels, overlap_area, overlap_perimeter) as it
    returns them (real64 sums)."""
    a = boxes(a)
    labels = np.zeros(len(a), np.int64)
    ar, pe = C.c_double(), C.c_double()
    ref().ref_pnum(len(a), _ptr(a), _ptr(labels), C.byref(ar), C.byref(pe))
    return labels, ar.value, pe.value


def ref_interactions(a):
    a = boxes(a)
    return int(ref().ref_interactions(len(a), _ptr(a)))
