"""accspmm_amd — MI355X (gfx950) METCF SpMM.

Python mirror of AccSpMM's host interface over the C ABI of
``libaccspmm.so`` (declared in ``include/accspmm.h``):

* :class:`CSR` / :func:`read_mtx` — ``CSR<T>(const COO<T>*)`` and
  ``read_from_mtx`` (include/format.hpp:110-142, include/mmio.hpp:171-210)
* :class:`METCFBit` — ``METCFBit<T>::convertFromCSR`` (include/format.hpp:215-274),
  built on the GPU and exported byte-exact
* :class:`AdpBME` — ``AdpBME<T>::CSR2AdpBME`` (include/format.hpp:340-429)
* :func:`balanceStrategy` — include/utils.hpp:75-85
* :class:`DeviceGraph` — device-resident METCF handle with :meth:`DeviceGraph.spmm`
  (replaces ``TF32Compute::compute`` / ``adpBalanceCompute``,
  include/tf32_comp.hpp:52-113)

Device buffers are torch tensors (torch is only the allocator and stream
provider); all arithmetic runs in the library's HIP kernels.  There is no CPU
fallback: if the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

__all__ = [
    "lib", "Status", "AccSpMMError", "FORM_FP32", "FORM_SPLIT_BF16", "FORM_BF16",
    "BuildOpts", "Tile", "CSR", "read_mtx", "canonicalize_device", "autotune_layout", "METCFBit", "AdpBME",
    "balanceStrategy", "DeviceGraph", "Batch", "Server", "gen_chung_lu", "gen_stencil27",
    "gen_dense", "MetcfInfo", "order_rcm", "gather_ceiling",
]

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "lib" / "libaccspmm.so"

FORM_FP32, FORM_SPLIT_BF16, FORM_BF16 = 0, 1, 2
FORM_NAMES = {FORM_FP32: "fp32", FORM_SPLIT_BF16: "split_bf16", FORM_BF16: "bf16"}


class Status:
    OK = 0
    ERR_INVALID_ARG = 1
    ERR_NOT_CANONICAL = 2
    ERR_HIP = 3
    ERR_OUT_OF_MEMORY = 4
    ERR_UNSUPPORTED = 5
    ERR_BUDGET = 6
    ERR_NOT_FOUND = 7
    ERR_QUEUE_FULL = 8
    ERR_MTX = 9
    ERR_NOT_READY = 10
    ERR_SHAPE = 11
    ERR_FILE = 12
    ERR_SCRATCH = 13


class AccSpMMError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = lib().accspmm_status_string(status).decode() if _LIB else str(status)
        super().__init__(f"{what}: {msg} (status {status})")


# ---------------------------------------------------------------- structs
class CsrView(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("nnz", C.c_uint32),
                ("row_ptr", C.c_void_p), ("col_idx", C.c_void_p), ("val", C.c_void_p)]


class BuildOpts(C.Structure):
    _fields_ = [("row_window", C.c_uint32), ("col_window", C.c_uint32),
                ("t_load", C.c_uint32), ("t_comp", C.c_uint32),
                ("t_write", C.c_uint32), ("target_num", C.c_uint32),
                ("imbalance_threshold", C.c_uint32), ("build_adpbme", C.c_uint32),
                ("reorder_rows", C.c_uint32), ("symmetric_order", C.c_uint32)]


class MetcfInfo(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("nnz", C.c_uint32),
                ("row_window", C.c_uint32), ("num_windows", C.c_uint32),
                ("num_tiles", C.c_uint32), ("num_groups", C.c_uint32),
                ("ibd", C.c_uint32), ("balanced", C.c_uint32),
                ("max_window_tiles", C.c_uint32), ("device_bytes", C.c_uint64),
                ("reordered", C.c_uint32), ("max_group_values", C.c_uint32),
                ("symmetric_order", C.c_uint32)]


class Tile(C.Structure):
    _fields_ = [("feat_per_lane", C.c_uint32), ("waves_per_block", C.c_uint32),
                ("split_tiles", C.c_uint32), ("xcd_slice", C.c_uint32),
                ("pack_tiles", C.c_uint32), ("col_panels", C.c_uint32),
                ("sweep", C.c_uint32), ("b_split", C.c_uint32)]

    def as_tuple(self):
        return (self.feat_per_lane, self.waves_per_block, self.split_tiles, self.xcd_slice,
                self.pack_tiles, self.col_panels, self.sweep, self.b_split)


class HostCsr(C.Structure):
    _fields_ = [("rows", C.c_uint32), ("cols", C.c_uint32), ("nnz", C.c_uint32),
                ("row_ptr", C.POINTER(C.c_uint32)), ("col_idx", C.POINTER(C.c_uint32)),
                ("val", C.POINTER(C.c_float))]


class Problem(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("B", C.c_void_p), ("ldb", C.c_uint32),
                ("C", C.c_void_p), ("ldc", C.c_uint32), ("K", C.c_uint32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch", C.c_uint32),
                ("max_queue", C.c_uint32), ("max_inflight", C.c_uint32),
                ("form", C.c_uint32), ("one_per_launch", C.c_uint32),
                ("large_work", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64),
                ("launches", C.c_uint64), ("rejected_budget", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("peak_inflight_bytes", C.c_uint64),
                ("busy_s", C.c_double)]


_LIB = None

_SIGS = {
    "accspmm_status_string": (C.c_char_p, [C.c_int]),
    "accspmm_version": (C.c_char_p, []),
    "accspmm_build_opts_default": (None, [C.POINTER(BuildOpts)]),
    "accspmm_metcf_build": (C.c_int, [C.POINTER(CsrView), C.POINTER(BuildOpts),
                                      C.POINTER(C.c_void_p), C.c_void_p]),
    "accspmm_metcf_build_host": (C.c_int, [C.POINTER(CsrView), C.POINTER(BuildOpts),
                                           C.POINTER(C.c_void_p)]),
    "accspmm_csr_canonicalize": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                           C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32),
                                           C.c_void_p]),
    "accspmm_metcf_destroy": (C.c_int, [C.c_void_p]),
    "accspmm_metcf_get_info": (C.c_int, [C.c_void_p, C.POINTER(MetcfInfo)]),
    "accspmm_metcf_export": (C.c_int, [C.c_void_p] + [C.c_void_p] * 7),
    "accspmm_metcf_row_order": (C.c_int, [C.c_void_p, C.c_void_p]),
    "accspmm_metcf_col_order": (C.c_int, [C.c_void_p, C.c_void_p]),
    "accspmm_order_rcm": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "accspmm_gather_probe": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]),
    "accspmm_autotune_layout": (C.c_int, [C.POINTER(CsrView), C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_uint32, C.c_uint32, C.c_int,
                                          C.POINTER(BuildOpts), C.POINTER(C.c_float),
                                          C.c_void_p]),
    "accspmm_metcf_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "accspmm_metcf_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.c_void_p]),
    "accspmm_spmm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                               C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Tile),
                               C.c_void_p]),
    "accspmm_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "accspmm_spmm_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                       C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Tile),
                                       C.c_void_p]),
    "accspmm_spmm_prepare_b": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                         C.c_int, C.POINTER(Tile), C.c_void_p]),
    "accspmm_spmm_prepared": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                        C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Tile),
                                        C.c_void_p]),
    "accspmm_tile_select": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(Tile)]),
    "accspmm_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                   C.c_uint32, C.c_uint32, C.c_int, C.c_char_p,
                                   C.POINTER(Tile), C.POINTER(C.c_float), C.c_void_p]),
    "accspmm_tune_key": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_char_p, C.c_size_t]),
    "accspmm_batch_create": (C.c_int, [C.POINTER(Problem), C.c_uint32, C.c_int,
                                       C.POINTER(C.c_void_p), C.c_void_p]),
    "accspmm_batch_rebind": (C.c_int, [C.c_void_p, C.POINTER(Problem), C.c_uint32, C.c_void_p]),
    "accspmm_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "accspmm_batch_run_graph": (C.c_int, [C.c_void_p, C.c_void_p]),
    "accspmm_batch_destroy": (C.c_int, [C.c_void_p]),
    "accspmm_batch_num_tasks": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "accspmm_host_csr_free": (None, [C.POINTER(HostCsr)]),
    "accspmm_read_mtx": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(HostCsr),
                                   C.POINTER(C.c_int)]),
    "accspmm_csr_is_canonical": (C.c_int, [C.POINTER(HostCsr)]),
    "accspmm_read_mtx_threads_used": (C.c_int, []),
    "accspmm_write_mtx": (C.c_int, [C.POINTER(HostCsr), C.c_char_p, C.c_int]),
    "accspmm_gen_chung_lu": (C.c_int, [C.c_uint32, C.c_double, C.c_double, C.c_uint64,
                                       C.c_int, C.POINTER(HostCsr)]),
    "accspmm_gen_stencil27": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(HostCsr)]),
    "accspmm_gen_dense": (None, [C.c_uint64, C.c_uint64, C.c_void_p]),
    "accspmm_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "accspmm_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "accspmm_server_register": (C.c_int, [C.c_void_p, C.POINTER(CsrView),
                                          C.POINTER(C.c_uint32)]),
    "accspmm_server_register_opts": (C.c_int, [C.c_void_p, C.POINTER(CsrView),
                                               C.POINTER(BuildOpts), C.POINTER(C.c_uint32)]),
    "accspmm_server_register_file": (C.c_int, [C.c_void_p, C.c_char_p,
                                               C.POINTER(C.c_uint32)]),
    "accspmm_server_unregister": (C.c_int, [C.c_void_p, C.c_uint32]),
    "accspmm_server_submit": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                        C.c_void_p, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_uint64)]),
    "accspmm_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    "accspmm_server_drain": (C.c_int, [C.c_void_p]),
    "accspmm_server_pause": (C.c_int, [C.c_void_p, C.c_int]),
    "accspmm_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "accspmm_server_destroy": (C.c_int, [C.c_void_p]),
}


def lib() -> C.CDLL:
    """Loads libaccspmm.so (after torch, so both share one HIP runtime)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m accspmm_amd.build` "
            "(there is no CPU fallback)")
    try:
        import torch  # noqa: F401  (loads the HIP runtime the library binds to)
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue  # reported by the ABI export test
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _check(status: int, what: str) -> None:
    if status != Status.OK:
        raise AccSpMMError(status, what)


def _ptr(a) -> int:
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()  # torch tensor


def _dense(t, rows: int, what: str, name: str, K: int | None = None) -> int:
    """Checks a dense operand: a 2-D float32 CUDA tensor of `rows` rows (and
    K columns when given) with unit column stride.  The library receives
    bare pointers and strides, so a smaller buffer would be read or written
    out of bounds on the GPU; this raises instead.  Returns its K."""
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_cuda \
            or t.dim() != 2:
        raise TypeError(f"{what}: {name} must be a 2-D float32 CUDA tensor")
    if t.shape[0] != rows or t.shape[1] < 1 or (K is not None and t.shape[1] != K):
        want = f"({rows}, {K if K is not None else 'K >= 1'})"
        raise ValueError(f"{what}: {name} has shape {tuple(t.shape)}, expected {want}")
    if t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError(f"{what}: {name} needs unit column stride (row-major rows)")
    if t.shape[0] > 1 and t.stride(0) < t.shape[1]:
        raise ValueError(f"{what}: {name} rows overlap (row stride {t.stride(0)})")
    return int(t.shape[1])


def _operands(B, C_out, rows: int, cols: int, what: str) -> int:
    """B (cols, K) and C (rows, K) on one device; returns K."""
    K = _dense(B, cols, what, "B")
    _dense(C_out, rows, what, "C", K)
    if B.device != C_out.device:
        raise ValueError(f"{what}: B and C are on different devices")
    return K


def _stream(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


# ---------------------------------------------------------------- host CSR
class CSR:
    """Host CSR with the reference's member names (include/format.hpp:96-142)."""

    def __init__(self, rows: int, cols: int, row_ptr, col_idx, data):
        self.rows, self.cols = int(rows), int(cols)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        self.data = np.ascontiguousarray(data, dtype=np.float32)
        self.nnz = int(self.col_idx.size)

    @classmethod
    def from_coo(cls, rows: int, cols: int, r, c, v) -> "CSR":
        """Stable counting sort by row (include/format.hpp:110-142)."""
        r = np.asarray(r, dtype=np.uint32)
        order = np.argsort(r, kind="stable")
        counts = np.bincount(r, minlength=rows).astype(np.uint32)
        rp = np.zeros(rows + 1, dtype=np.uint32)
        np.cumsum(counts, out=rp[1:])
        return cls(rows, cols, rp, np.asarray(c, np.uint32)[order],
                   np.asarray(v, np.float32)[order])

    @classmethod
    def _from_host(cls, h: HostCsr) -> "CSR":
        rp = np.ctypeslib.as_array(h.row_ptr, shape=(h.rows + 1,)).copy()
        ci = np.ctypeslib.as_array(h.col_idx, shape=(max(h.nnz, 1),))[: h.nnz].copy()
        va = np.ctypeslib.as_array(h.val, shape=(max(h.nnz, 1),))[: h.nnz].copy()
        out = cls(h.rows, h.cols, rp, ci, va)
        lib().accspmm_host_csr_free(C.byref(h))
        return out

    def _host_view(self) -> HostCsr:
        return HostCsr(self.rows, self.cols, self.nnz,
                       self.row_ptr.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.col_idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                       self.data.ctypes.data_as(C.POINTER(C.c_float)))

    def is_canonical(self) -> bool:
        return bool(lib().accspmm_csr_is_canonical(C.byref(self._host_view())))

    def write_mtx(self, path, symmetric: bool = False) -> None:
        """Writes a Matrix Market file (accspmm_write_mtx)."""
        _check(lib().accspmm_write_mtx(C.byref(self._host_view()), os.fsencode(path),
                                       int(symmetric)), "write_mtx")

    def to_device(self, device="cuda"):
        import torch
        rp = torch.from_numpy(self.row_ptr.view(np.int32)).to(device)
        ci = torch.from_numpy(self.col_idx.view(np.int32)).to(device)
        va = torch.from_numpy(self.data).to(device)
        return rp, ci, va

    def printSize(self, print_: bool = True) -> int:
        size = 4 * (self.row_ptr.size + self.col_idx.size + self.data.size)
        if print_:
            print(f"\n========== size of CSR =========\nTotal size: {size} bytes")
        return size


def read_mtx(path: str, canonicalize: bool = False) -> CSR:
    """Matrix Market -> CSR (read_from_mtx + CSR ctor semantics)."""
    h = HostCsr()
    code = C.c_int(0)
    st = lib().accspmm_read_mtx(str(path).encode(), int(canonicalize), C.byref(h),
                                C.byref(code))
    if st != Status.OK:
        err = AccSpMMError(st, f"read_mtx({path})")
        err.mtx_code = code.value
        raise err
    return CSR._from_host(h)


def gen_chung_lu(n: int, mean_degree: float, exponent: float, seed: int,
                 value_mode: int = 0) -> CSR:
    h = HostCsr()
    _check(lib().accspmm_gen_chung_lu(n, mean_degree, exponent, seed, value_mode,
                                      C.byref(h)), "gen_chung_lu")
    return CSR._from_host(h)


def gen_stencil27(g: int, seed: int) -> CSR:
    h = HostCsr()
    _check(lib().accspmm_gen_stencil27(g, seed, C.byref(h)), "gen_stencil27")
    return CSR._from_host(h)


def gen_dense(seed: int, rows: int, cols: int) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float32)
    lib().accspmm_gen_dense(seed, rows * cols, out.ctypes.data)
    return out


# ---------------------------------------------------------------- device graph
def default_build_opts(row_window: int = 8, **kw) -> BuildOpts:
    o = BuildOpts()
    lib().accspmm_build_opts_default(C.byref(o))
    o.row_window = row_window
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def order_rcm(row_ptr, col_idx, n: int | None = None) -> np.ndarray:
    """Reverse Cuthill-McKee order of an n x n pattern (host arrays):
    perm[i] = the node placed at position i (accspmm_order_rcm)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    ci = np.ascontiguousarray(col_idx, dtype=np.uint32)
    n = len(rp) - 1 if n is None else n
    if len(rp) != n + 1:
        raise ValueError("row_ptr must have n + 1 entries")
    perm = np.empty(n, np.uint32)
    _check(lib().accspmm_order_rcm(n, _ptr(rp), _ptr(ci) if len(ci) else 0, _ptr(perm)),
           "order_rcm")
    return perm


def gather_ceiling(region_rows: int, row_bytes: int = 512, waves: int = 4096,
                   rows_per_wave: int = 2048, reps: int = 5, seed: int = 7) -> float:
    """GB/s at which the GPU serves random row gathers from a region of
    region_rows rows of row_bytes bytes in the SpMM kernels' load pattern
    (accspmm_gather_probe): median of `reps` timed launches after two
    warm-up launches, uniform random row indices."""
    import torch
    region = torch.zeros(region_rows * row_bytes // 4, dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    idx = torch.randint(0, region_rows, (waves * rows_per_wave,), dtype=torch.int32,
                        device="cuda", generator=g)
    out = torch.empty(waves * 64, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    ms = []
    for i in range(reps + 2):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        _check(lib().accspmm_gather_probe(region.data_ptr(), region_rows, row_bytes,
                                          idx.data_ptr(), waves, rows_per_wave,
                                          out.data_ptr(), s.cuda_stream), "gather_probe")
        e1.record(s)
        e1.synchronize()
        if i >= 2:
            ms.append(e0.elapsed_time(e1))
    ms.sort()
    return waves * rows_per_wave * row_bytes / (ms[len(ms) // 2] * 1e-3) / 1e9


def autotune_layout(csr: CSR, B, C_out, form: int = FORM_FP32, stream=None):
    """Fastest layout (row window 8/16 x input rows, rows by length, and
    reverse Cuthill-McKee for square matrices) for (csr, B, form): returns
    (BuildOpts, ms) (accspmm_autotune_layout)."""
    rp, ci, va = csr.to_device()
    view = CsrView(csr.rows, csr.cols, csr.nnz, rp.data_ptr(), ci.data_ptr(), va.data_ptr())
    o = BuildOpts()
    ms = C.c_float()
    _check(lib().accspmm_autotune_layout(C.byref(view), B.data_ptr(), B.stride(0),
                                         C_out.data_ptr(), C_out.stride(0), B.shape[1], form,
                                         C.byref(o), C.byref(ms), _stream(stream)),
           "autotune_layout")
    return o, ms.value


def canonicalize_device(rows: int, cols: int, row_ptr, col_idx, val, stream=None) -> int:
    """Makes a device CSR (torch int32/float32 tensors) canonical in place
    (accspmm_csr_canonicalize); returns the new nonzero count."""
    nnz = C.c_uint32(0)
    _check(lib().accspmm_csr_canonicalize(rows, cols, col_idx.numel(), row_ptr.data_ptr(),
                                          col_idx.data_ptr(), val.data_ptr(), C.byref(nnz),
                                          _stream(stream)), "csr_canonicalize")
    return nnz.value


class DeviceGraph:
    """A METCF(+AdpBME) graph resident on the GPU (opaque C handle).
    canonicalize=True sorts and de-duplicates the rows on the device first
    (input rows may then be in any order, with repeated columns)."""

    def __init__(self, csr: CSR, opts: BuildOpts | None = None, stream=None,
                 device="cuda", canonicalize: bool = False):
        import torch
        self.csr = csr
        self._dev = csr.to_device(device)
        rp, ci, va = self._dev
        nnz = csr.nnz
        if canonicalize:
            nnz = canonicalize_device(csr.rows, csr.cols, rp, ci, va, stream)
        view = CsrView(csr.rows, csr.cols, nnz, rp.data_ptr(), ci.data_ptr(),
                       va.data_ptr())
        self.opts = opts or default_build_opts()
        h = C.c_void_p()
        st = lib().accspmm_metcf_build(C.byref(view), C.byref(self.opts), C.byref(h),
                                       _stream(stream))
        _check(st, "metcf_build")
        self.handle = h
        self._dev = None  # the handle does not keep the CSR
        torch.cuda.synchronize()

    def save(self, path) -> None:
        """Writes the graph to a binary file (accspmm_metcf_save)."""
        _check(lib().accspmm_metcf_save(self.handle, os.fsencode(path)), "metcf_save")

    @classmethod
    def load(cls, path, stream=None) -> "DeviceGraph":
        """A graph from a file written by save() (accspmm_metcf_load)."""
        import torch
        h = C.c_void_p()
        _check(lib().accspmm_metcf_load(os.fsencode(path), C.byref(h), _stream(stream)),
               "metcf_load")
        g = cls.__new__(cls)
        g.handle = h
        g.csr = None
        g._dev = None
        inf = g.info
        g.opts = default_build_opts(inf.row_window, build_adpbme=1 if inf.num_groups else 0,
                                    reorder_rows=inf.reordered)
        torch.cuda.synchronize()
        return g

    def close(self):
        if getattr(self, "handle", None):
            lib().accspmm_metcf_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def info(self) -> MetcfInfo:
        inf = MetcfInfo()
        _check(lib().accspmm_metcf_get_info(self.handle, C.byref(inf)), "get_info")
        return inf

    def export(self) -> dict:
        inf = self.info
        words = inf.row_window // 8
        out = {
            "rowWindowOffset": np.empty(inf.num_windows + 1, np.uint32),
            "tcOffset": np.empty(inf.num_tiles + 1, np.uint32),
            "tcLocalBit": np.empty(inf.num_tiles * words, np.uint64),
            "sparseA2B": np.empty(inf.num_tiles * 8, np.uint32),
            "data": np.empty(inf.nnz, np.float32),
        }
        grp = rix = None
        if self.opts.build_adpbme:
            grp = np.empty(inf.num_groups + 1, np.uint32)
            rix = np.empty(inf.num_tiles, np.uint32)
            out["groupOffset"], out["rowIndices"] = grp, rix
        _check(lib().accspmm_metcf_export(
            self.handle, _ptr(out["rowWindowOffset"]), _ptr(out["tcOffset"]),
            _ptr(out["tcLocalBit"]), _ptr(out["sparseA2B"]), _ptr(out["data"]),
            _ptr(grp), _ptr(rix)), "metcf_export")
        return out

    def row_order(self) -> np.ndarray:
        """Input row of each format row (identity unless reorder_rows or
        symmetric_order)."""
        out = np.empty(self.info.rows, np.uint32)
        _check(lib().accspmm_metcf_row_order(self.handle, _ptr(out)), "metcf_row_order")
        return out

    def col_order(self) -> np.ndarray:
        """Input column (B row) of each format column (identity unless
        symmetric_order)."""
        out = np.empty(self.info.cols, np.uint32)
        _check(lib().accspmm_metcf_col_order(self.handle, _ptr(out)), "metcf_col_order")
        return out

    def tile_select(self, K: int, form: int = FORM_FP32) -> Tile:
        t = Tile()
        _check(lib().accspmm_tile_select(self.handle, K, form, C.byref(t)), "tile_select")
        return t

    def tune_key(self, K: int, form: int = FORM_FP32) -> str:
        buf = C.create_string_buffer(512)
        _check(lib().accspmm_tune_key(self.handle, K, form, buf, 512), "tune_key")
        return buf.value.decode()

    def spmm(self, B, C_out=None, form: int = FORM_FP32, tile: Tile | None = None,
             stream=None):
        """C = A @ B on the GPU.  B: (cols, K) float32 CUDA tensor."""
        import torch
        inf = self.info
        K = _dense(B, inf.cols, "spmm", "B")
        if C_out is None:
            C_out = torch.empty((inf.rows, K), dtype=torch.float32, device=B.device)
        _operands(B, C_out, inf.rows, inf.cols, "spmm")
        st = lib().accspmm_spmm(self.handle, B.data_ptr(), B.stride(0),
                                C_out.data_ptr(), C_out.stride(0), K, form,
                                C.byref(tile) if tile is not None else None,
                                _stream(stream))
        _check(st, "spmm")
        return C_out

    def prepare_b(self, B, form: int, tile: Tile | None = None, stream=None):
        """First half of :meth:`spmm`: converts B for the BF16 forms."""
        _dense(B, self.info.cols, "spmm_prepare_b", "B")
        st = lib().accspmm_spmm_prepare_b(self.handle, B.data_ptr(), B.stride(0),
                                          B.shape[1], form,
                                          C.byref(tile) if tile is not None else None,
                                          _stream(stream))
        _check(st, "spmm_prepare_b")

    def spmm_prepared(self, B, C_out, form: int, tile: Tile | None = None, stream=None):
        """Second half of :meth:`spmm`: the SpMM kernel on the prepared B."""
        _operands(B, C_out, self.info.rows, self.info.cols, "spmm_prepared")
        st = lib().accspmm_spmm_prepared(self.handle, B.data_ptr(), B.stride(0),
                                         C_out.data_ptr(), C_out.stride(0), B.shape[1],
                                         form, C.byref(tile) if tile is not None else None,
                                         _stream(stream))
        _check(st, "spmm_prepared")
        return C_out

    def autotune(self, B, C_out, form: int = FORM_FP32, cache_path: str | None = None,
                 stream=None):
        _operands(B, C_out, self.info.rows, self.info.cols, "autotune")
        t = Tile()
        ms = C.c_float()
        st = lib().accspmm_autotune(self.handle, B.data_ptr(), B.stride(0),
                                    C_out.data_ptr(), C_out.stride(0), B.shape[1], form,
                                    cache_path.encode() if cache_path else None,
                                    C.byref(t), C.byref(ms), _stream(stream))
        _check(st, "autotune")
        return t, ms.value


# ---------------------------------------------------------------- reference-named API
class METCFBit:
    """``METCFBit<float>`` (include/format.hpp:202-325): the public vectors
    are filled from the GPU builder's byte-exact export."""

    def __init__(self, row_window: int = 8):
        self.row_window = row_window
        self.graph: DeviceGraph | None = None

    def convertFromCSR(self, csr: CSR) -> None:
        self.graph = DeviceGraph(csr, default_build_opts(self.row_window, build_adpbme=0))
        ex = self.graph.export()
        self.rowWindowOffset = ex["rowWindowOffset"]
        self.tcOffset = ex["tcOffset"]
        self.tcLocalBit = ex["tcLocalBit"]
        self.sparseA2B = ex["sparseA2B"]
        self.data = ex["data"]

    def printSize(self, print_: bool = True) -> int:
        s = 4 * (self.rowWindowOffset.size + self.tcOffset.size + self.sparseA2B.size) \
            + 8 * self.tcLocalBit.size
        if print_:
            print(f"\n========== size of ME-CTF-bit =========\ntotal size(without data): {s} bytes")
        return s


class AdpBME:
    """``AdpBME<float>`` (include/format.hpp:327-503)."""

    def __init__(self, row_window: int = 8):
        self.row_window = row_window
        self.graph: DeviceGraph | None = None

    def CSR2AdpBME(self, csr: CSR, t_load: int, t_comp: int, t_write: int,
                   target_num: int = 16) -> None:
        opts = default_build_opts(self.row_window, t_load=t_load, t_comp=t_comp,
                                  t_write=t_write, target_num=target_num, build_adpbme=1)
        self.graph = DeviceGraph(csr, opts)
        ex = self.graph.export()
        self.groupOffset = ex["groupOffset"]
        self.tcOffset = ex["tcOffset"]
        self.rowIndices = ex["rowIndices"]
        self.tcLocalBit = ex["tcLocalBit"]
        self.sparseA2B = ex["sparseA2B"]
        self.data = ex["data"]


def balanceStrategy(metcf: METCFBit, numTcBlocks: int, numBlocks: int) -> bool:
    """include/utils.hpp:75-85, computed on the GPU at build time (the
    arguments are accepted for signature parity and must match the graph)."""
    inf = metcf.graph.info
    if numTcBlocks != inf.num_tiles or numBlocks != inf.num_windows:
        raise ValueError("balanceStrategy arguments do not match the built graph")
    return bool(inf.balanced)


# ---------------------------------------------------------------- batch
class Batch:
    """Many (graph, B, C, K) problems in one persistent launch."""

    def __init__(self, problems, form: int = FORM_SPLIT_BF16, stream=None):
        self._keep = problems
        arr = (Problem * len(problems))()
        for i, (g, B, Cout) in enumerate(problems):
            inf = g.info
            _operands(B, Cout, inf.rows, inf.cols, f"batch problem {i}")
            arr[i] = Problem(g.handle.value, B.data_ptr(), B.stride(0), Cout.data_ptr(),
                             Cout.stride(0), B.shape[1])
        h = C.c_void_p()
        _check(lib().accspmm_batch_create(arr, len(problems), form, C.byref(h),
                                          _stream(stream)), "batch_create")
        self.handle = h

    def run(self, stream=None, graph: bool = False):
        fn = lib().accspmm_batch_run_graph if graph else lib().accspmm_batch_run
        _check(fn(self.handle, _stream(stream)), "batch_run")

    def autotune(self, stream=None) -> float:
        """Times batch-wide choices (NB cap, short-window packing) and keeps
        the fastest; returns its ms per run.  Runs overwrite the C buffers."""
        ms = C.c_float()
        _check(lib().accspmm_batch_autotune(self.handle, _stream(stream), C.byref(ms)),
               "batch_autotune")
        return ms.value

    @property
    def num_tasks(self) -> int:
        n = C.c_uint64()
        _check(lib().accspmm_batch_num_tasks(self.handle, C.byref(n)), "batch_num_tasks")
        return n.value

    def close(self):
        if getattr(self, "handle", None):
            lib().accspmm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving
class Server:
    """Continuous-batching SpMM server over resident METCF graphs."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_batch: int = 64,
                 max_queue: int = 1 << 16, max_inflight: int = 2,
                 form: int = FORM_SPLIT_BF16, one_per_launch: bool = False,
                 large_work: int = 0):
        o = ServerOpts()
        lib().accspmm_server_opts_default(C.byref(o))
        o.hbm_budget_bytes = hbm_budget_bytes
        o.max_batch = max_batch
        o.max_queue = max_queue
        o.max_inflight = max_inflight
        o.form = form
        o.one_per_launch = int(one_per_launch)
        if large_work:
            o.large_work = large_work
        h = C.c_void_p()
        _check(lib().accspmm_server_create(C.byref(o), C.byref(h)), "server_create")
        self.handle = h
        self._keep = {}
        self._shapes = {}  # graph id -> (rows, cols), for operand checks

    def register(self, csr: CSR, opts: BuildOpts | None = None) -> int:
        """Builds and registers a resident graph (build options: row window,
        reorder_rows; None = the library's automatic layout)."""
        import torch
        rp, ci, va = csr.to_device()
        view = CsrView(csr.rows, csr.cols, csr.nnz, rp.data_ptr(), ci.data_ptr(),
                       va.data_ptr())
        gid = C.c_uint32()
        torch.cuda.synchronize()
        _check(lib().accspmm_server_register_opts(self.handle, C.byref(view),
                                                  C.byref(opts) if opts is not None else None,
                                                  C.byref(gid)), "server_register")
        self._shapes[gid.value] = (csr.rows, csr.cols)
        return gid.value

    def register_file(self, path) -> int:
        """Registers a graph saved by DeviceGraph.save (warm start)."""
        gid = C.c_uint32()
        _check(lib().accspmm_server_register_file(self.handle, os.fsencode(path),
                                                  C.byref(gid)), "server_register_file")
        # rows and cols: the first two header words after magic and version
        with open(path, "rb") as f:
            rows, cols = np.frombuffer(f.read(20)[12:20], "<u4").tolist()
        self._shapes[gid.value] = (rows, cols)
        return gid.value

    def unregister(self, gid: int) -> None:
        _check(lib().accspmm_server_unregister(self.handle, gid), "server_unregister")
        self._shapes.pop(gid, None)

    def submit(self, gid: int, B, C_out) -> int:
        if gid in self._shapes:
            rows, cols = self._shapes[gid]
            _operands(B, C_out, rows, cols, "server_submit")
        rid = C.c_uint64()
        st = lib().accspmm_server_submit(self.handle, gid, B.data_ptr(), B.stride(0),
                                         C_out.data_ptr(), C_out.stride(0), B.shape[1],
                                         C.byref(rid))
        _check(st, "server_submit")
        self._keep[rid.value] = (B, C_out)
        return rid.value

    def poll(self, rid: int):
        """Latency in seconds if done, else None."""
        lat = C.c_double()
        st = lib().accspmm_server_poll(self.handle, rid, C.byref(lat))
        if st == Status.ERR_NOT_READY:
            return None
        self._keep.pop(rid, None)
        _check(st, "server_poll")
        return lat.value

    def drain(self) -> None:
        _check(lib().accspmm_server_drain(self.handle), "server_drain")

    def pause(self, paused: bool = True) -> None:
        _check(lib().accspmm_server_pause(self.handle, int(paused)), "server_pause")

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().accspmm_server_get_stats(self.handle, C.byref(st)), "server_stats")
        return st

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().accspmm_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
