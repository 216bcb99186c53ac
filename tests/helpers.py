"""Shared test helpers: matrices with prescribed window shapes, random
canonical CSR, and the FP64 error check."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def kat() -> dict:
    return json.loads((GOLDEN / "kat.json").read_text())


def csr_from_window_cols(cols_per_window, rw: int = 8):
    """CSR whose window w has `cols_per_window[w]` distinct columns, all in the
    window's first row (so the window has ceil(n/8) tiles)."""
    W = len(cols_per_window)
    rows = W * rw
    ncols = max(max(cols_per_window, default=1), 1)
    rp, ci, va = [0], [], []
    for w in range(W):
        for r in range(rw):
            if r == 0:
                cs = list(range(cols_per_window[w]))
                ci += cs
                va += [1.0] * len(cs)
            rp.append(len(ci))
    return rows, ncols, np.array(rp, np.uint32), np.array(ci, np.uint32), \
        np.array(va, np.float32)


def random_csr(rng: np.random.Generator, rows: int, cols: int, density: float,
               empty_rows: float = 0.0, heavy_rows: int = 0):
    """Random canonical CSR (sorted, duplicate-free rows)."""
    rp = [0]
    ci, va = [], []
    heavy = set(rng.choice(rows, size=min(heavy_rows, rows), replace=False).tolist()) \
        if heavy_rows else set()
    for r in range(rows):
        if rng.random() < empty_rows:
            rp.append(len(ci))
            continue
        if r in heavy:
            n = int(min(cols, max(1, cols * 0.5)))
        else:
            n = int(rng.binomial(cols, density))
        cs = np.sort(rng.choice(cols, size=n, replace=False)) if n else np.array([], int)
        ci += cs.tolist()
        va += rng.uniform(-1, 1, size=n).astype(np.float32).tolist()
        rp.append(len(ci))
    return (np.array(rp, np.uint32), np.array(ci, np.uint32),
            np.array(va, np.float32))


def canonical_np(rows: int, cols: int, row_ptr, col_idx, val):
    """Canonical form of a CSR, restated in numpy: entries ordered by (row,
    column) keeping input order among equal columns (a stable sort), each run
    of equal columns summed left to right in float32."""
    rp = np.asarray(row_ptr, np.int64)
    r = np.repeat(np.arange(rows, dtype=np.int64), np.diff(rp))
    c = np.asarray(col_idx, np.int64)
    v = np.asarray(val, np.float32)
    order = np.lexsort((c, r))          # stable: ties keep input order
    key = r[order] * cols + c[order]
    v = v[order]
    head = np.r_[True, key[1:] != key[:-1]] if key.size else np.zeros(0, bool)
    starts = np.flatnonzero(head)
    group = np.cumsum(head) - 1
    rank = np.arange(key.size) - starts[group] if key.size else np.zeros(0, np.int64)
    sums = v[starts].copy()
    for k in range(1, int(rank.max()) + 1 if rank.size else 1):
        sel = rank == k                 # the k-th member of every run, in order
        sums[group[sel]] = (sums[group[sel]] + v[sel]).astype(np.float32)
    ukey = key[head]
    out_rp = np.zeros(rows + 1, np.int64)
    np.add.at(out_rp, ukey // cols + 1, 1)
    return (np.cumsum(out_rp).astype(np.uint32), (ukey % cols).astype(np.uint32),
            sums.astype(np.float32))


def scrambled_csr(rng: np.random.Generator, rows: int, cols: int, per_row_max: int,
                  dup_frac: float = 0.2, heavy=(), empty_frac: float = 0.1):
    """CSR with rows in random column order and repeated columns (not
    canonical); `heavy` lists (row, length) pairs of long rows."""
    lens = rng.integers(0, per_row_max + 1, rows)
    lens[rng.random(rows) < empty_frac] = 0
    for r_, n_ in heavy:
        lens[r_] = n_
    rp = np.r_[0, np.cumsum(lens)].astype(np.uint32)
    nnz = int(rp[-1])
    ci = rng.integers(0, cols, nnz).astype(np.uint32)
    ndup = int(dup_frac * nnz)
    if ndup and nnz > 1:
        # copy a column from elsewhere in the same row
        r_of = np.repeat(np.arange(rows), lens)
        idx = rng.integers(0, nnz, ndup)
        src = rp[r_of[idx]] + (rng.random(ndup) * lens[r_of[idx]]).astype(np.int64)
        ci[idx] = ci[src]
    va = rng.uniform(-1, 1, nnz).astype(np.float32)
    return rp, ci, va


def dense_block_csr(rng: np.random.Generator, rows: int, cols: int, block_cols: int,
                    fill: float):
    """Canonical CSR whose 16-row blocks each draw block_cols random columns
    and keep each (row, column) of the block with probability fill: tiles of
    16-row windows near `fill` full (groups of up to 512 values), every
    third block empty of values in its second half of rows."""
    rp, ci, va = [0], [], []
    for b0 in range(0, rows, 16):
        cs = np.sort(rng.choice(cols, size=min(block_cols, cols), replace=False))
        for r in range(b0, min(rows, b0 + 16)):
            if (b0 // 16) % 3 == 2 and r - b0 >= 8:
                rp.append(len(ci))
                continue
            keep = cs[rng.random(len(cs)) < fill]
            ci += keep.tolist()
            va += rng.uniform(-1, 1, size=len(keep)).astype(np.float32).tolist()
            rp.append(len(ci))
    return (np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va, np.float32))


def max_group_values(row_window: int, rows: int, row_ptr, col_idx, val) -> int:
    """Values of the fullest 4-tile group of the oracle's METCF (groups start
    at each window's first tile and every 4 tiles after it)."""
    from oracle import oracle as O
    m = O.metcf_build(row_window, rows, row_ptr, col_idx, val)
    rwo = np.asarray(m["rowWindowOffset"], np.int64)
    tco = np.asarray(m["tcOffset"], np.int64)
    best = 0
    for w in range(len(rwo) - 1):
        s = np.arange(rwo[w], rwo[w + 1], 4)
        if len(s):
            best = max(best, int((tco[np.minimum(s + 4, rwo[w + 1])] - tco[s]).max()))
    return best
