"""Test-side writer and reader of the SNAP layer file, independent of the
library: the container of SaveCompressedStreamToDisk /
RestoreCompressedStreamFromDisk (src/serialization.f90:15-130), chunks of
CompressBoxesToStream (src/box_codec.f90:670-764) encoded by the oracle's
box codec and compressed by Python's zlib and pyarrow's Snappy and zstd
(the Google Snappy and Facebook zstd libraries).

Layout (little endian): int32 codec version 1, int32 layer properties, int64
boxes, int64 chunks, int32 method; int64 polygon numbers when properties has
bit 4; per chunk int64 boxes, int64 raw bytes, int64 stored bytes, stored
bytes; int64 1-based positions of the chunk headers; int64 1-based position
of that table, last."""
import struct
import zlib

import numpy as np
import pyarrow as pa

import oracle_lib as O

BOXES_PER_CHUNK = 65536
SNAPPY, ZLIB, ZSTD = 1, 2, 3


def compress(method, raw):
    if method == SNAPPY:
        return pa.Codec("snappy").compress(raw, asbytes=True)
    if method == ZLIB:
        return zlib.compress(raw)   # level 6, the default of zlib's compress()
    if method == ZSTD:
        return pa.Codec("zstd", compression_level=20).compress(raw, asbytes=True)
    raise ValueError(method)


def decompress(method, data, raw_size):
    if method == SNAPPY:
        return pa.Codec("snappy").decompress(data, decompressed_size=raw_size, asbytes=True)
    if method == ZLIB:
        return zlib.decompress(data)
    if method == ZSTD:
        return pa.Codec("zstd").decompress(data, decompressed_size=raw_size, asbytes=True)
    raise ValueError(method)


def build(boxes, method, props=0, pnum=None, encode=O.codec_encode, chunk=BOXES_PER_CHUNK):
    """SNAP file bytes of boxes ((n, 4) int32)."""
    b = np.ascontiguousarray(boxes, np.int32).reshape(-1, 4)
    n = len(b)
    nch = (n + chunk - 1) // chunk
    if pnum is not None:
        props |= 4
    out = bytearray(struct.pack("<iiqqi", 1, props, n, nch, method))
    if pnum is not None:
        out += np.ascontiguousarray(pnum, np.int64).tobytes()
    pos = []
    for c in range(nch):
        part = b[c * chunk:(c + 1) * chunk]
        raw = encode(part)
        data = compress(method, raw)
        pos.append(len(out) + 1)
        out += struct.pack("<qqq", len(part), len(raw), len(data)) + data
    table = len(out) + 1
    out += struct.pack(f"<{nch}q", *pos) + struct.pack("<q", table)
    return bytes(out)


def parse(data):
    """Fields of a SNAP file; raises on any inconsistency of the container."""
    if len(data) == 0:
        return {"props": 0, "total": 0, "chunks": [], "method": 0, "pnum": None}
    version, props, total, nch, method = struct.unpack_from("<iiqqi", data, 0)
    assert version == 1
    p = 28
    pnum = None
    if props & 4:
        pnum = np.frombuffer(data, np.int64, total, p).copy()
        p += 8 * total
    chunks = []
    for _ in range(nch):
        nb, raw, cs = struct.unpack_from("<qqq", data, p)
        chunks.append({"pos": p + 1, "boxes": nb, "raw": raw, "stored": data[p + 24:p + 24 + cs]})
        p += 24 + cs
    offs = struct.unpack_from(f"<{nch}q", data, p)
    assert list(offs) == [c["pos"] for c in chunks]
    (table,) = struct.unpack_from("<q", data, p + 8 * nch)
    assert table == p + 1 and p + 8 * nch + 8 == len(data)
    assert sum(c["boxes"] for c in chunks) == total
    return {"props": props, "total": total, "chunks": chunks, "method": method, "pnum": pnum}


def raw_chunks(f):
    return [decompress(f["method"], c["stored"], c["raw"]) for c in f["chunks"]]


def boxes_of(data):
    """Boxes of a SNAP file, decoded by the oracle."""
    f = parse(data)
    parts = []
    for c, raw in zip(f["chunks"], raw_chunks(f)):
        assert len(raw) == c["raw"]
        d = O.codec_decode(raw, c["boxes"])
        assert d is not None
        parts.append(d)
    return np.concatenate(parts) if parts else np.zeros((0, 4), np.int32)
