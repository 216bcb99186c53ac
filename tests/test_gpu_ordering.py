"""Graphs built with a symmetric ordering (symmetric_order 1: reverse
Cuthill-McKee; 2: rows and columns by decreasing length).  The format
equals the format of the explicitly permuted matrix P A P^T; C comes back
in input row order equal, bit for bit, to P A P^T's C (with B permuted the
same way) scattered through P, and within each form's tolerance of the
FP64 oracle — on the single launch (work units and sweep), the batched,
saved and allocation-free paths."""
import ctypes

import numpy as np
import pytest

import accspmm_amd as acc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {acc.FORM_FP32: 1e-5, acc.FORM_SPLIT_BF16: 1e-5, acc.FORM_BF16: 1e-2}
FORMS = [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16]


def _torch():
    import torch
    return torch


def _sym_permuted(csr, perm):
    """P A P^T as a canonical CSR: row i is input row perm[i], columns
    relabelled by the inverse permutation and sorted."""
    n = csr.rows
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    rp, ci, va = [0], [], []
    for r in perm:
        a, b = int(csr.row_ptr[r]), int(csr.row_ptr[r + 1])
        c = inv[csr.col_idx[a:b]]
        o = np.argsort(c, kind="stable")
        ci.append(c[o])
        va.append(csr.data[a:b][o])
        rp.append(rp[-1] + (b - a))
    return acc.CSR(n, n, np.array(rp, np.uint32), np.concatenate(ci).astype(np.uint32),
                   np.concatenate(va).astype(np.float32))


def _scrambled_stencil(g, seed):
    c = acc.gen_stencil27(g, seed)
    p = np.random.default_rng(seed).permutation(c.rows)
    # relabel node v -> position of v in p, i.e. P A P^T with P = p
    return _sym_permuted(c, p)


def _graphs():
    return {"stencil_scrambled": _scrambled_stencil(14, 3),
            "chung_lu": acc.gen_chung_lu(6000, 18.0, 2.1, 12, 0)}


@pytest.mark.parametrize("rw", [8, 16])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("name", ["stencil_scrambled", "chung_lu"])
def test_format_and_results(name, order, rw):
    torch = _torch()
    csr = _graphs()[name]
    g = acc.DeviceGraph(csr, acc.default_build_opts(rw, symmetric_order=order))
    perm = g.row_order()
    assert g.info.symmetric_order == order and g.info.reordered == 1
    assert np.array_equal(g.col_order(), perm)
    assert np.array_equal(np.sort(perm), np.arange(csr.rows))
    if order == 1:
        assert np.array_equal(perm, acc.order_rcm(csr.row_ptr, csr.col_idx))
    else:
        lens = np.diff(csr.row_ptr.astype(np.int64))
        assert np.array_equal(perm, np.argsort(-lens, kind="stable"))
    pcsr = _sym_permuted(csr, perm)
    gp = acc.DeviceGraph(pcsr, acc.default_build_opts(rw))
    ea, eb = g.export(), gp.export()
    for k in eb:
        assert np.array_equal(ea[k], eb[k]), k
    ref = O.metcf_build(rw, pcsr.rows, pcsr.row_ptr, pcsr.col_idx, pcsr.data)
    for k in ("rowWindowOffset", "tcOffset", "tcLocalBit", "sparseA2B"):
        assert np.array_equal(ea[k], ref[k]), k
    for K in (32, 100):
        Bh = acc.gen_dense(K, csr.cols, K)
        B = torch.from_numpy(Bh).cuda()
        Bp = torch.from_numpy(np.ascontiguousarray(Bh[perm])).cuda()
        C64 = O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        ab = O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        for form in FORMS:
            for tile in (None, acc.Tile(2, 4, 12, 0, 64, 4), acc.Tile(4, 4, 0, 1, 64, 0),
                         acc.Tile(2, 4, 0, 0, 64, 4, 1)):
                C = torch.full((csr.rows, K), float("nan"), device="cuda")
                g.spmm(B, C, form, tile)
                Cp = gp.spmm(Bp, None, form, tile)
                torch.cuda.synchronize()
                Cn = C.cpu().numpy()
                assert np.array_equal(Cn[perm], Cp.cpu().numpy()), (form, K)
                mx, _ = O.errors(Cn, C64, ab)
                assert mx <= TOL[form], (form, K, mx)


def _noalloc(g, B, C, form, tile=None):
    import torch
    return acc.lib().accspmm_spmm_noalloc(
        g.handle, B.data_ptr(), B.stride(0), C.data_ptr(), C.stride(0), B.shape[1], form,
        ctypes.byref(tile) if tile is not None else None,
        torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_batch_saved_noalloc(form, tmp_path):
    torch = _torch()
    graphs = _graphs()
    built = []
    for i, (order, (name, csr)) in enumerate(zip((1, 2), graphs.items())):
        g = acc.DeviceGraph(csr, acc.default_build_opts(8, symmetric_order=order))
        B = torch.from_numpy(acc.gen_dense(60 + i, csr.cols, 64)).cuda()
        built.append((csr, g, B))
    csr0 = graphs["chung_lu"]
    g0 = acc.DeviceGraph(csr0)
    B0 = torch.from_numpy(acc.gen_dense(9, csr0.cols, 64)).cuda()
    built.append((csr0, g0, B0))
    alone = [g.spmm(B, None, form).clone() for _, g, B in built]
    torch.cuda.synchronize()
    # batched launch equals standalone, bit for bit
    Cs = [torch.full_like(C0, float("nan")) for C0 in alone]
    batch = acc.Batch([(g, B, C) for (_, g, B), C in zip(built, Cs)], form)
    batch.run()
    torch.cuda.synchronize()
    for C1, C0 in zip(Cs, alone):
        assert torch.equal(C1, C0)
    batch.close()
    # allocation-free entry equals accspmm_spmm
    for (_, g, B), C0 in zip(built, alone):
        C = torch.full_like(C0, float("nan"))
        assert _noalloc(g, B, C, form) == acc.Status.OK
        torch.cuda.synchronize()
        assert torch.equal(C, C0)
    # save / load keeps both orders and the result
    for j in (0, 1):
        csr, g, B = built[j]
        path = tmp_path / f"s{j}.metcf"
        g.save(path)
        h = acc.DeviceGraph.load(path)
        assert h.info.symmetric_order == g.info.symmetric_order
        assert np.array_equal(h.row_order(), g.row_order())
        assert np.array_equal(h.col_order(), g.col_order())
        assert torch.equal(h.spmm(B, None, form), alone[j])


def test_rejects_non_square_and_non_canonical():
    torch = _torch()
    rp = torch.tensor([0, 2, 3], dtype=torch.int32, device="cuda")
    ci = torch.tensor([0, 2, 1], dtype=torch.int32, device="cuda")
    va = torch.ones(3, device="cuda")
    h = ctypes.c_void_p()
    view = acc.CsrView(2, 3, 3, rp.data_ptr(), ci.data_ptr(), va.data_ptr())
    st = acc.lib().accspmm_metcf_build(ctypes.byref(view),
                                       ctypes.byref(acc.default_build_opts(8, symmetric_order=1)),
                                       ctypes.byref(h), None)
    assert st == acc.Status.ERR_UNSUPPORTED
    ci2 = torch.tensor([1, 0, 1], dtype=torch.int32, device="cuda")   # row 0 unsorted
    view = acc.CsrView(2, 2, 3, rp.data_ptr(), ci2.data_ptr(), va.data_ptr())
    for order in (1, 2):
        st = acc.lib().accspmm_metcf_build(
            ctypes.byref(view), ctypes.byref(acc.default_build_opts(8, symmetric_order=order)),
            ctypes.byref(h), None)
        assert st == acc.Status.ERR_NOT_CANONICAL
    st = acc.lib().accspmm_metcf_build(ctypes.byref(view),
                                       ctypes.byref(acc.default_build_opts(8, symmetric_order=3)),
                                       ctypes.byref(h), None)
    assert st == acc.Status.ERR_UNSUPPORTED
