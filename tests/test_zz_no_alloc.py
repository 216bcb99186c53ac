"""accspmm_spmm_noalloc, called through the C ABI, against accspmm_spmm: the
same C bits on two calls in a row (the second on a smaller graph), both on
handles without a device schedule yet (host plan copied into the scratch
array) and with one (after accspmm_spmm built it); calls that need more than
the scratch array are refused with ACCSPMM_ERR_SCRATCH and leave C alone."""
import ctypes

import numpy as np
import pytest

import accspmm_amd as acc
from helpers import random_csr

pytestmark = pytest.mark.gpu

FORMS = [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16]


def _noalloc(g, B, C, form, tile):
    import torch
    return acc.lib().accspmm_spmm_noalloc(
        g.handle, B.data_ptr(), B.stride(0), C.data_ptr(), C.stride(0), B.shape[1], form,
        ctypes.byref(tile) if tile is not None else None,
        torch.cuda.current_stream().cuda_stream)


def _graph(seed, rows, cols, dens, rw):
    rng = np.random.default_rng(seed)
    rp, ci, va = random_csr(rng, rows, cols, dens, empty_rows=0.05, heavy_rows=4)
    return acc.DeviceGraph(acc.CSR(rows, cols, rp, ci, va), acc.default_build_opts(rw))


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("rw", [8, 16])
def test_noalloc_equals_spmm(form, rw):
    import torch
    # split windows and column panels (partials in the scratch array), and
    # the library's own tile
    for tile in (acc.Tile(2, 4, 16, 0, 64, 2), None):
        big = _graph(11 + rw, 4000, 5000, 0.02, rw)
        small = _graph(12 + rw, 1500, 2000, 0.01, rw)
        cases = [(big, 96, 1), (small, 40, 2)]
        Bs = [torch.from_numpy(acc.gen_dense(s, g.info.cols, K)).cuda() for g, K, s in cases]
        Cs = [torch.full((g.info.rows, K), float("nan"), device="cuda") for g, K, _ in cases]
        # two calls in a row, the second smaller, before either handle has a
        # device schedule for the tile
        for (g, _, _), B, C in zip(cases, Bs, Cs):
            assert _noalloc(g, B, C, form, tile) == acc.Status.OK
        torch.cuda.synchronize()
        refs = [g.spmm(B, None, form, tile).clone() for (g, _, _), B in zip(cases, Bs)]
        torch.cuda.synchronize()
        for C, R in zip(Cs, refs):
            assert torch.equal(C, R)
        # again, now on the device schedules accspmm_spmm cached
        for (g, _, _), B, C in zip(cases, Bs, Cs):
            C.fill_(float("nan"))
            assert _noalloc(g, B, C, form, tile) == acc.Status.OK
        torch.cuda.synchronize()
        for C, R in zip(Cs, refs):
            assert torch.equal(C, R)


def test_noalloc_refuses_oversized_and_sweep():
    import torch
    # ~5,000 windows of ~100 tiles, split every 8 tiles, at K = 512: about
    # 60,000 partial slots of 16 KiB, far above the 64 MiB scratch array
    g = _graph(21, 40000, 20000, 0.005, 8)
    B = torch.from_numpy(acc.gen_dense(3, 20000, 512)).cuda()
    C = torch.full((40000, 512), float("nan"), device="cuda")
    tile = acc.Tile(2, 4, 8, 0, 0)
    assert _noalloc(g, B, C, acc.FORM_SPLIT_BF16, tile) == acc.Status.ERR_SCRATCH
    torch.cuda.synchronize()
    assert torch.isnan(C).all()
    assert b"scratch" in acc.lib().accspmm_status_string(acc.Status.ERR_SCRATCH)
    sweep = acc.Tile(2, 4, 0, 0, 64, 2, 1)
    assert _noalloc(g, B, C, acc.FORM_FP32, sweep) == acc.Status.ERR_UNSUPPORTED
    # the same graph within the limit (K = 16, no split) runs and matches
    B16 = torch.from_numpy(acc.gen_dense(4, 20000, 16)).cuda()
    C16 = torch.full((40000, 16), float("nan"), device="cuda")
    small_tile = acc.Tile(1, 4, 0, 0, 64)
    assert _noalloc(g, B16, C16, acc.FORM_SPLIT_BF16, small_tile) == acc.Status.OK
    ref = g.spmm(B16, None, acc.FORM_SPLIT_BF16, small_tile)
    torch.cuda.synchronize()
    assert torch.equal(C16, ref)
