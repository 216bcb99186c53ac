"""ctypes wrapper of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / the timed CPU baseline.  Each function
restates a reference routine; the file:line citations are in oracle.h.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class _Metcf(C.Structure):
    _fields_ = [("rw", C.c_uint32), ("num_windows", C.c_uint32),
                ("num_tiles", C.c_uint32), ("nnz", C.c_uint32),
                ("rowWindowOffset", C.POINTER(C.c_uint32)),
                ("tcOffset", C.POINTER(C.c_uint32)),
                ("tcLocalBit", C.POINTER(C.c_uint64)),
                ("sparseA2B", C.POINTER(C.c_uint32)),
                ("data", C.POINTER(C.c_float)),
                ("num_groups", C.c_uint32),
                ("groupOffset", C.POINTER(C.c_uint32)),
                ("rowIndices", C.POINTER(C.c_uint32))]


_L = None
_u32p = C.POINTER(C.c_uint32)


def _lib():
    global _L
    if _L is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(LIB_PATH))
        L.orc_read_mtx.restype = C.c_int
        L.orc_read_mtx.argtypes = [C.c_char_p, _u32p, _u32p, _u32p,
                                   C.POINTER(_u32p), C.POINTER(_u32p),
                                   C.POINTER(C.POINTER(C.c_float))]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_coo_to_csr.argtypes = [C.c_uint32, C.c_uint32] + [C.c_void_p] * 6
        L.orc_metcf_build.restype = C.POINTER(_Metcf)
        L.orc_metcf_build.argtypes = [C.c_uint32, C.c_uint32] + [C.c_void_p] * 3
        L.orc_adpbme_build.restype = C.POINTER(_Metcf)
        L.orc_adpbme_build.argtypes = [C.c_uint32, C.c_uint32] + [C.c_void_p] * 3 + \
            [C.c_uint32] * 4
        L.orc_metcf_free.argtypes = [C.POINTER(_Metcf)]
        L.orc_ibd.restype = C.c_uint32
        L.orc_ibd.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                              C.POINTER(C.c_int)]
        L.orc_metcf_decode.restype = C.c_uint32
        L.orc_metcf_decode.argtypes = [C.POINTER(_Metcf)] + [C.c_void_p] * 3
        L.orc_spmm_f64.argtypes = [C.c_uint32] + [C.c_void_p] * 4 + [C.c_uint32,
                                                                     C.c_uint32, C.c_void_p]
        L.orc_spmm_abs_f64.argtypes = L.orc_spmm_f64.argtypes
        L.orc_spmm_tf32_f64.argtypes = L.orc_spmm_f64.argtypes
        L.orc_tf32_rna.restype = C.c_float
        L.orc_tf32_rna.argtypes = [C.c_float]
        L.orc_errors.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p, C.c_uint32,
                                 C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.orc_spmm_f32_omp.restype = C.c_int
        L.orc_spmm_f32_omp.argtypes = [C.c_uint32] + [C.c_void_p] * 4 + \
            [C.c_uint32, C.c_uint32, C.c_void_p, C.c_int]
        _L = L
    return _L


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def read_mtx(path: str):
    """Returns (code, rows, cols, r, c, v) — COO in the reference's order."""
    L = _lib()
    m, n, nz = C.c_uint32(), C.c_uint32(), C.c_uint32()
    r, c, v = _u32p(), _u32p(), C.POINTER(C.c_float)()
    code = L.orc_read_mtx(str(path).encode(), C.byref(m), C.byref(n), C.byref(nz),
                          C.byref(r), C.byref(c), C.byref(v))
    if code != 0:
        return code, 0, 0, None, None, None
    k = nz.value
    R = np.ctypeslib.as_array(r, shape=(max(k, 1),))[:k].copy()
    Cc = np.ctypeslib.as_array(c, shape=(max(k, 1),))[:k].copy()
    V = np.ctypeslib.as_array(v, shape=(max(k, 1),))[:k].copy()
    for p in (r, c, v):
        L.orc_free(C.cast(p, C.c_void_p))
    return 0, m.value, n.value, R, Cc, V


def coo_to_csr(rows: int, r, c, v):
    r = np.ascontiguousarray(r, np.uint32)
    c = np.ascontiguousarray(c, np.uint32)
    v = np.ascontiguousarray(v, np.float32)
    rp = np.zeros(rows + 1, np.uint32)
    ci = np.zeros(r.size, np.uint32)
    va = np.zeros(r.size, np.float32)
    _lib().orc_coo_to_csr(rows, r.size, _p(r), _p(c), _p(v), _p(rp), _p(ci), _p(va))
    return rp, ci, va


def _unpack(m, adp: bool) -> dict:
    s = m.contents
    T, W, words = s.num_tiles, s.num_windows, s.rw // 8

    def arr(ptr, n, dt):
        if n == 0 or not ptr:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

    out = {
        "rowWindowOffset": arr(s.rowWindowOffset, W + 1, np.uint32),
        "tcOffset": arr(s.tcOffset, T + 1, np.uint32),
        "tcLocalBit": arr(s.tcLocalBit, T * words, np.uint64),
        "sparseA2B": arr(s.sparseA2B, 8 * T, np.uint32),
        "data": arr(s.data, s.nnz, np.float32),
    }
    if adp:
        out["groupOffset"] = arr(s.groupOffset, s.num_groups + 1, np.uint32)
        out["rowIndices"] = arr(s.rowIndices, T, np.uint32)
    return out


def metcf_build(rw: int, rows: int, row_ptr, col_idx, val) -> dict:
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    L = _lib()
    m = L.orc_metcf_build(rw, rows, _p(rp), _p(ci), _p(va))
    try:
        return _unpack(m, False)
    finally:
        L.orc_metcf_free(m)


def adpbme_build(rw: int, rows: int, row_ptr, col_idx, val, t_load=3, t_comp=1,
                 t_write=2, target_num=16) -> dict:
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    L = _lib()
    m = L.orc_adpbme_build(rw, rows, _p(rp), _p(ci), _p(va), t_load, t_comp,
                           t_write, target_num)
    try:
        return _unpack(m, True)
    finally:
        L.orc_metcf_free(m)


def ibd(row_window_offset, threshold: int = 8):
    rwo = np.ascontiguousarray(row_window_offset, np.uint32)
    W = rwo.size - 1
    T = int(rwo[-1])
    bal = C.c_int()
    v = _lib().orc_ibd(_p(rwo), W, T, threshold, C.byref(bal))
    return int(v), bool(bal.value)


def metcf_decode(m: dict, rw: int):
    """METCF -> (rows, cols, vals) triples in tile order."""
    s = _Metcf()
    s.rw = rw
    s.num_windows = m["rowWindowOffset"].size - 1
    s.num_tiles = m["tcOffset"].size - 1
    s.nnz = m["data"].size
    keep = {k: np.ascontiguousarray(v) for k, v in m.items()}
    s.rowWindowOffset = keep["rowWindowOffset"].ctypes.data_as(_u32p)
    s.tcOffset = keep["tcOffset"].ctypes.data_as(_u32p)
    s.tcLocalBit = keep["tcLocalBit"].ctypes.data_as(C.POINTER(C.c_uint64))
    s.sparseA2B = keep["sparseA2B"].ctypes.data_as(_u32p)
    s.data = keep["data"].ctypes.data_as(C.POINTER(C.c_float))
    n = s.nnz
    r = np.zeros(max(n, 1), np.uint32)
    c = np.zeros(max(n, 1), np.uint32)
    v = np.zeros(max(n, 1), np.float32)
    k = _lib().orc_metcf_decode(C.byref(s), _p(r), _p(c), _p(v))
    return r[:k], c[:k], v[:k]


def spmm_f64(row_ptr, col_idx, val, B: np.ndarray) -> np.ndarray:
    rows = np.asarray(row_ptr).size - 1
    B = np.ascontiguousarray(B, np.float32)
    K = B.shape[1]
    out = np.empty((rows, K), np.float64)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    _lib().orc_spmm_f64(rows, _p(rp), _p(ci), _p(va), _p(B), K, K, _p(out))
    return out


def tf32_rna(x: float) -> float:
    return _lib().orc_tf32_rna(x)


def spmm_tf32_f64(row_ptr, col_idx, val, B: np.ndarray) -> np.ndarray:
    """tf32(A) * tf32(B), FP64 accumulation (the reference's precision class)."""
    rows = np.asarray(row_ptr).size - 1
    B = np.ascontiguousarray(B, np.float32)
    K = B.shape[1]
    out = np.empty((rows, K), np.float64)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    _lib().orc_spmm_tf32_f64(rows, _p(rp), _p(ci), _p(va), _p(B), K, K, _p(out))
    return out


def spmm_abs_f64(row_ptr, col_idx, val, B: np.ndarray) -> np.ndarray:
    rows = np.asarray(row_ptr).size - 1
    B = np.ascontiguousarray(B, np.float32)
    K = B.shape[1]
    out = np.empty((rows, K), np.float64)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    _lib().orc_spmm_abs_f64(rows, _p(rp), _p(ci), _p(va), _p(B), K, K, _p(out))
    return out


def errors(Cmat: np.ndarray, C64: np.ndarray, absAB: np.ndarray):
    """(max_rel componentwise vs |A||B|, Frobenius relative)."""
    Cmat = np.ascontiguousarray(Cmat, np.float32)
    mx, fr = C.c_double(), C.c_double()
    _lib().orc_errors(Cmat.shape[0], Cmat.shape[1], _p(Cmat), Cmat.shape[1],
                      _p(np.ascontiguousarray(C64)), _p(np.ascontiguousarray(absAB)),
                      C.byref(mx), C.byref(fr))
    return mx.value, fr.value


def spmm_f32_omp(row_ptr, col_idx, val, B: np.ndarray, out: np.ndarray | None = None,
                 threads: int = 0):
    rows = np.asarray(row_ptr).size - 1
    B = np.ascontiguousarray(B, np.float32)
    K = B.shape[1]
    if out is None:
        out = np.empty((rows, K), np.float32)
    rp = np.ascontiguousarray(row_ptr, np.uint32)
    ci = np.ascontiguousarray(col_idx, np.uint32)
    va = np.ascontiguousarray(val, np.float32)
    used = _lib().orc_spmm_f32_omp(rows, _p(rp), _p(ci), _p(va), _p(B), K, K, _p(out),
                                   threads)
    return out, used
