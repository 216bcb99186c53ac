"""Device canonicalisation (accspmm_csr_canonicalize) against the numpy
restatement (pinned by brute force in test_host), bit for bit, and graphs
built from scrambled device CSR against graphs built from canonical CSR."""
import ctypes

import numpy as np
import pytest

import accspmm_amd as acc
from helpers import canonical_np, scrambled_csr

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _device(rp, ci, va):
    torch = _torch()
    return (torch.from_numpy(rp.view(np.int32).copy()).cuda(),
            torch.from_numpy(ci.view(np.int32).copy()).cuda(),
            torch.from_numpy(va.copy()).cuda())


def _canon_check(rows, cols, rp, ci, va):
    d_rp, d_ci, d_va = _device(rp, ci, va)
    n = acc.canonicalize_device(rows, cols, d_rp, d_ci, d_va)
    _torch().cuda.synchronize()
    erp, eci, eva = canonical_np(rows, cols, rp, ci, va)
    assert n == eci.size
    assert np.array_equal(d_rp.cpu().numpy().view(np.uint32), erp)
    assert np.array_equal(d_ci.cpu().numpy()[:n].view(np.uint32), eci)
    assert np.array_equal(d_va.cpu().numpy()[:n].view(np.uint32), eva.view(np.uint32))


@pytest.mark.parametrize("case", ["tiny", "medium_heavy", "wide_64bit_keys", "one_by_one",
                                  "all_duplicates"])
def test_canonicalize_matches_restatement(case):
    rng = np.random.default_rng(len(case))
    if case == "tiny":
        for _ in range(20):
            rows, cols = int(rng.integers(1, 12)), int(rng.integers(1, 12))
            _canon_check(rows, cols, *scrambled_csr(rng, rows, cols, 8, 0.3))
        return
    if case == "medium_heavy":
        rows, cols = 5000, 3000
        rp, ci, va = scrambled_csr(rng, rows, cols, 40, 0.2, heavy=[(17, 70000), (4000, 3000)])
    elif case == "wide_64bit_keys":     # 17 row bits + 17 column bits
        rows, cols = 70000, 100000
        rp, ci, va = scrambled_csr(rng, rows, cols, 12, 0.1)
    elif case == "one_by_one":
        rows, cols = 1, 1
        rp, ci, va = np.array([0, 5], np.uint32), np.zeros(5, np.uint32), \
            rng.uniform(-1, 1, 5).astype(np.float32)
    else:                               # every row one column, many times
        rows, cols = 300, 50
        lens = rng.integers(0, 30, rows)
        rp = np.r_[0, np.cumsum(lens)].astype(np.uint32)
        ci = np.repeat(rng.integers(0, cols, rows), lens).astype(np.uint32)
        va = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    _canon_check(rows, cols, rp, ci, va)


def test_canonical_input_is_unchanged():
    csr = acc.gen_chung_lu(20000, 12.0, 2.1, 4, 0)
    d_rp, d_ci, d_va = _device(csr.row_ptr, csr.col_idx, csr.data)
    n = acc.canonicalize_device(csr.rows, csr.cols, d_rp, d_ci, d_va)
    assert n == csr.nnz
    assert np.array_equal(d_ci.cpu().numpy().view(np.uint32), csr.col_idx)
    assert np.array_equal(d_va.cpu().numpy(), csr.data)


def test_invalid_inputs():
    torch = _torch()
    rp = np.array([0, 2, 3], np.uint32)
    ci = np.array([1, 0, 2], np.uint32)
    va = np.ones(3, np.float32)
    L = acc.lib()
    out = ctypes.c_uint32(0)
    cases = {
        "column_out_of_range": (rp, np.array([1, 0, 3], np.uint32)),
        "row_ptr_decreasing": (np.array([0, 3, 2], np.uint32), ci),
        "row_ptr_end_not_nnz": (np.array([0, 1, 2], np.uint32), ci),
        "row_ptr_start_not_zero": (np.array([1, 2, 3], np.uint32), ci),
    }
    for name, (r_, c_) in cases.items():
        d_rp, d_ci, d_va = _device(r_, c_, va)
        st = L.accspmm_csr_canonicalize(2, 3, 3, d_rp.data_ptr(), d_ci.data_ptr(),
                                        d_va.data_ptr(), ctypes.byref(out),
                                        torch.cuda.current_stream().cuda_stream)
        assert st == acc.Status.ERR_INVALID_ARG, name
    assert L.accspmm_csr_canonicalize(0, 3, 0, 0, 0, 0, ctypes.byref(out), 0) == \
        acc.Status.ERR_INVALID_ARG


@pytest.mark.parametrize("form", [acc.FORM_FP32, acc.FORM_SPLIT_BF16])
def test_graph_from_scrambled_device_csr(form):
    torch = _torch()
    rng = np.random.default_rng(9)
    rows, cols = 3000, 2500
    rp, ci, va = scrambled_csr(rng, rows, cols, 30, 0.15, heavy=[(5, 2000)])
    scrambled = acc.CSR(rows, cols, rp, ci, va)
    g1 = acc.DeviceGraph(scrambled, canonicalize=True)
    crp, cci, cva = canonical_np(rows, cols, rp, ci, va)
    g2 = acc.DeviceGraph(acc.CSR(rows, cols, crp, cci, cva))
    e1, e2 = g1.export(), g2.export()
    for k in e2:
        assert np.array_equal(np.asarray(e1[k]), np.asarray(e2[k])), k
    B = torch.from_numpy(acc.gen_dense(2, cols, 64)).cuda()
    C1 = g1.spmm(B, None, form).clone()
    C2 = g2.spmm(B, None, form).clone()
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
