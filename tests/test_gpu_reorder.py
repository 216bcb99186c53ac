"""Graphs built with reorder_rows = 1 (rows by decreasing length): the
format equals the format of the explicitly permuted CSR (oracle builder),
C comes back in input row order equal, bit for bit, to the permuted graph's
C scattered through the permutation, and within each form's tolerance of
the FP64 oracle — on the single launch, sweep, batched and saved paths."""
import numpy as np
import pytest

import accspmm_amd as acc
from helpers import random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {acc.FORM_FP32: 1e-5, acc.FORM_SPLIT_BF16: 1e-5, acc.FORM_BF16: 1e-2}
FORMS = [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16]


def _torch():
    import torch
    return torch


def _permuted(csr):
    """Rows in order of decreasing length, ties in input order."""
    lens = np.diff(csr.row_ptr.astype(np.int64))
    order = np.argsort(-lens, kind="stable")
    rp = np.r_[0, np.cumsum(lens[order])].astype(np.uint32)
    idx = np.concatenate([np.arange(csr.row_ptr[r], csr.row_ptr[r + 1]) for r in order]
                         + [np.zeros(0, np.int64)]).astype(np.int64)
    return order.astype(np.uint32), acc.CSR(csr.rows, csr.cols, rp, csr.col_idx[idx],
                                            csr.data[idx])


def _graphs():
    rng = np.random.default_rng(5)
    rp, ci, va = random_csr(rng, 1203, 900, 0.02, empty_rows=0.15, heavy_rows=3)
    return {"random": acc.CSR(1203, 900, rp, ci, va),
            "chung_lu": acc.gen_chung_lu(9000, 20.0, 2.1, 8, 1)}


@pytest.mark.parametrize("rw", [8, 16])
@pytest.mark.parametrize("name", ["random", "chung_lu"])
def test_format_and_results(name, rw):
    torch = _torch()
    csr = _graphs()[name]
    order, pcsr = _permuted(csr)
    g = acc.DeviceGraph(csr, acc.default_build_opts(rw, reorder_rows=1))
    gp = acc.DeviceGraph(pcsr, acc.default_build_opts(rw))
    assert g.info.reordered == 1 and gp.info.reordered == 0
    assert np.array_equal(g.row_order(), order)
    assert np.array_equal(gp.row_order(), np.arange(csr.rows))
    ea, eb = g.export(), gp.export()
    for k in eb:
        assert np.array_equal(ea[k], eb[k]), k
    ref = O.metcf_build(rw, pcsr.rows, pcsr.row_ptr, pcsr.col_idx, pcsr.data)
    for k in ("rowWindowOffset", "tcOffset", "tcLocalBit", "sparseA2B"):
        assert np.array_equal(ea[k], ref[k]), k
    for K in (32, 100):
        Bh = acc.gen_dense(K, csr.cols, K)
        B = torch.from_numpy(Bh).cuda()
        C64 = O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        ab = O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        for form in FORMS:
            for tile in (None, acc.Tile(2, 4, 12, 0, 64, 4), acc.Tile(4, 4, 0, 1, 64, 0)):
                C = torch.full((csr.rows, K), float("nan"), device="cuda")
                g.spmm(B, C, form, tile)
                Cp = gp.spmm(B, None, form, tile)
                torch.cuda.synchronize()
                Cn = C.cpu().numpy()
                assert np.array_equal(Cn[order], Cp.cpu().numpy()), (form, K)
                mx, fr = O.errors(Cn, C64, ab)
                assert mx <= TOL[form], (form, K, mx)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_sweep_and_batch_and_saved(form, tmp_path):
    torch = _torch()
    graphs = _graphs()
    built = []
    for i, (name, csr) in enumerate(graphs.items()):
        g = acc.DeviceGraph(csr, acc.default_build_opts(8, reorder_rows=1))
        B = torch.from_numpy(acc.gen_dense(40 + i, csr.cols, 64)).cuda()
        built.append((csr, g, B))
    # a graph in input order in the same batch
    csr0 = graphs["random"]
    g0 = acc.DeviceGraph(csr0)
    B0 = torch.from_numpy(acc.gen_dense(7, csr0.cols, 64)).cuda()
    built.append((csr0, g0, B0))
    alone = [g.spmm(B, None, form).clone() for _, g, B in built]
    # sweep mode equals work units
    for (_, g, B), C0 in zip(built, alone):
        Cs = g.spmm(B, None, form, acc.Tile(2, 4, 0, 0, 64, 4, 1)).clone()
        Cw = g.spmm(B, None, form, acc.Tile(2, 4, 0, 0, 64, 4, 0)).clone()
        torch.cuda.synchronize()
        assert torch.equal(Cs, Cw)
    # batched launch equals standalone
    Cs = [torch.full_like(C0, float("nan")) for C0 in alone]
    batch = acc.Batch([(g, B, C) for (_, g, B), C in zip(built, Cs)], form)
    batch.run()
    torch.cuda.synchronize()
    for C1, C0 in zip(Cs, alone):
        assert torch.equal(C1, C0)
    batch.close()
    # save / load keeps the row order
    csr, g, B = built[1]
    path = tmp_path / "r.metcf"
    g.save(path)
    h = acc.DeviceGraph.load(path)
    assert h.info.reordered == 1
    assert np.array_equal(h.row_order(), g.row_order())
    assert torch.equal(h.spmm(B, None, form), alone[1])


def test_reorder_rejects_bad_row_ptr():
    torch = _torch()
    import ctypes
    rp = torch.tensor([0, 3, 2, 4], dtype=torch.int32, device="cuda")   # decreasing
    ci = torch.tensor([0, 1, 2, 0], dtype=torch.int32, device="cuda")
    va = torch.ones(4, device="cuda")
    view = acc.CsrView(3, 3, 4, rp.data_ptr(), ci.data_ptr(), va.data_ptr())
    h = ctypes.c_void_p()
    st = acc.lib().accspmm_metcf_build(ctypes.byref(view),
                                       ctypes.byref(acc.default_build_opts(8, reorder_rows=1)),
                                       ctypes.byref(h), None)
    assert st == acc.Status.ERR_INVALID_ARG
    for rw, ro in ((8, 3), (12, 0)):
        st = acc.lib().accspmm_metcf_build(ctypes.byref(view),
                                           ctypes.byref(acc.default_build_opts(rw, reorder_rows=ro)),
                                           ctypes.byref(h), None)
        assert st == acc.Status.ERR_UNSUPPORTED


def test_automatic_choices():
    # row_window 0 and reorder_rows 2: 16-row windows and reordered rows for
    # rows averaging >= 128 nonzeros; for shorter rows, 16-row windows in
    # input order when a graph of >= 64 * 8192 8-row tiles has at most 92%
    # as many 16-row ones (banded: the 27-point stencil), the reference's
    # layout otherwise; the result
    # is the same as asking for that layout explicitly
    torch = _torch()
    rng = np.random.default_rng(12)
    long_rp, long_ci, long_va = random_csr(rng, 300, 2000, 0.08)      # ~160 per row
    short = acc.gen_chung_lu(5000, 20.0, 2.1, 9, 0)
    banded = acc.gen_stencil27(82, 13)      # 810,286 8-row tiles: large enough
    cases = ((acc.CSR(300, 2000, long_rp, long_ci, long_va), (16, 1)), (short, (8, 0)),
             (banded, (16, 0)))
    for csr, want in cases:
        assert (csr.nnz >= 128 * csr.rows) == (want == (16, 1))
        if want != (16, 1):
            # the tile ratio decides, as the oracle's builder counts it
            t8, t16 = (len(O.metcf_build(rw, csr.rows, csr.row_ptr, csr.col_idx,
                                         csr.data)["tcOffset"]) - 1 for rw in (8, 16))
            assert (t16 * 100 <= t8 * 92 and t8 >= 64 * 8192) == (want == (16, 0)), (t8, t16)
        g = acc.DeviceGraph(csr, acc.default_build_opts(0, reorder_rows=2))
        assert (g.info.row_window, g.info.reordered) == want
        ge = acc.DeviceGraph(csr, acc.default_build_opts(want[0], reorder_rows=want[1]))
        ea, eb = g.export(), ge.export()
        for k in eb:
            assert np.array_equal(ea[k], eb[k]), k
        B = torch.from_numpy(acc.gen_dense(4, csr.cols, 64)).cuda()
        assert torch.equal(g.spmm(B, None, acc.FORM_SPLIT_BF16),
                           ge.spmm(B, None, acc.FORM_SPLIT_BF16))


@pytest.mark.parametrize("form", [acc.FORM_SPLIT_BF16, acc.FORM_FP32])
def test_layout_autotuner(form):
    # the returned layout is one of the candidates (row window 8/16 x input
    # rows, rows by length, or reverse Cuthill-McKee for square matrices),
    # and a graph built with it computes C within the form's tolerance; bad
    # input reports its status
    torch = _torch()
    for name, csr in _graphs().items():
        B = torch.from_numpy(acc.gen_dense(8, csr.cols, 64)).cuda()
        Cs = torch.empty((csr.rows, 64), device="cuda")
        opts, ms = acc.autotune_layout(csr, B, Cs, form)
        assert opts.row_window in (8, 16) and opts.reorder_rows in (0, 1) and ms > 0
        assert opts.symmetric_order in (0, 1)
        assert not (opts.symmetric_order and opts.reorder_rows)
        assert not opts.symmetric_order or csr.rows == csr.cols
        g = acc.DeviceGraph(csr, opts)
        Cg = g.spmm(B, None, form).cpu().numpy()
        Bh = B.cpu().numpy()
        mx, _ = O.errors(Cg, O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, Bh),
                         O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, Bh))
        assert mx <= TOL[form], (name, mx)
    bad = acc.CSR(3, 3, np.array([0, 2, 3, 3], np.uint32), np.array([2, 1, 0], np.uint32),
                  np.ones(3, np.float32))                      # row 0 not sorted
    B = torch.zeros((3, 16), device="cuda")
    with pytest.raises(acc.AccSpMMError) as e:
        acc.autotune_layout(bad, B, torch.empty((3, 16), device="cuda"), form)
    assert e.value.status == acc.Status.ERR_NOT_CANONICAL
