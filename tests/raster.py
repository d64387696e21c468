"""Brute-force rasterisation on small integer grids: an independent check of the
oracle.  A box [x1,x2) x [y1,y2) covers the unit cells (i, j) with x1<=i<x2 and
y1<=j<y2; set operations are numpy boolean operations on cell grids."""
import numpy as np


def raster(boxes, ox, oy, W, H):
    g = np.zeros((W, H), bool)   # g[x, y]
    for x1, y1, x2, y2 in np.asarray(boxes, np.int64).reshape(-1, 4):
        if x1 < x2 and y1 < y2:
            g[x1 - ox:x2 - ox, y1 - oy:y2 - oy] = True
    return g


def strips_from_raster(g, ox, oy):
    """Maximal-strip form of a cell set: per unit x-column the maximal y-runs,
    each run extended along x while it stays identical; sorted (x1,y1,x2,y2)."""
    W, H = g.shape
    open_runs = {}
    out = []

    def runs(col):
        r, y = [], 0
        while y < H:
            if col[y]:
                s = y
                while y < H and col[y]:
                    y += 1
                r.append((s, y))
            else:
                y += 1
        return r

    for x in range(W + 1):
        cur = set(runs(g[x])) if x < W else set()
        for run in list(open_runs):
            if run not in cur:
                out.append((open_runs.pop(run) + ox, run[0] + oy, x + ox, run[1] + oy))
        for run in cur:
            if run not in open_runs:
                open_runs[run] = x
    out.sort()
    return np.asarray(out, np.int32).reshape(-1, 4)


def dilate(g, e, n, w, s):
    """Cells of GROW(E,N,W,S): cell (i,j) is set when some set cell (p,q) has
    p-W <= i < p+1+E and q-S <= j < q+1+N.  The grid must have margin."""
    W_, H_ = g.shape
    out = np.zeros_like(g)
    xs, ys = np.nonzero(g)
    for p, q in zip(xs, ys):
        out[max(0, p - w):min(W_, p + 1 + e), max(0, q - s):min(H_, q + 1 + n)] = True
    return out


def erode(g, d):
    """Cells whose whole (2d+1)^2 neighbourhood is set (outside the grid = unset)."""
    W_, H_ = g.shape
    pad = np.zeros((W_ + 2 * d, H_ + 2 * d), bool)
    pad[d:d + W_, d:d + H_] = g
    out = np.ones_like(g)
    for dx in range(-d, d + 1):
        for dy in range(-d, d + 1):
            out &= pad[d + dx:d + dx + W_, d + dy:d + dy + H_]
    return out
