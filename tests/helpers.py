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
