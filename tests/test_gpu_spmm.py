"""GPU SpMM per form against the host FP64 oracle.

Tolerances (componentwise max-rel against |A||B|, FP64 oracle):
  FP32-MFMA: 1e-5;  split-BF16: 1e-5;  single-BF16: 1e-2.
Split-BF16 splits A exactly into three bf16 terms and B into two (remainder
at most 2^-17 |b|) and accumulates the five products A_i B_j, j <= 1, so one
product is off by at most 2^-17 + 2^-25 relative (7.7e-6) before FP32
rounding: rows with one nonzero (test_split_single_products) expose that
bound directly, since no other term hides it.  The full-size workloads are in
test_gpu_large.py.
"""
import os
from pathlib import Path

import numpy as np
import pytest

import accspmm_amd as acc
from helpers import random_csr
from oracle import oracle as O

REPO_TESTS = Path(__file__).resolve().parent

pytestmark = pytest.mark.gpu

TOL = {acc.FORM_FP32: 1e-5, acc.FORM_SPLIT_BF16: 1e-5, acc.FORM_BF16: 1e-2}
FORMS = [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16]


def _torch():
    import torch
    return torch


def run_case(csr: acc.CSR, K: int, form: int, tile=None, rw: int = 8, seed: int = 0,
             graph=None, B=None):
    torch = _torch()
    g = graph or acc.DeviceGraph(csr, acc.default_build_opts(rw))
    if B is None:
        B = acc.gen_dense(seed, csr.cols, K)
    Bd = torch.from_numpy(B).cuda()
    Cd = torch.full((csr.rows, K), float("nan"), device="cuda")
    g.spmm(Bd, Cd, form, tile)
    torch.cuda.synchronize()
    Cg = Cd.cpu().numpy()
    C64 = O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, B)
    ab = O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, B)
    mx, fr = O.errors(Cg, C64, ab)
    return mx, fr, Cg


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("K", [1, 16, 17, 32, 64, 100, 128, 256])
def test_forms_random(form, K):
    rng = np.random.default_rng(K)
    rows, cols = 1000, 1200
    rp, ci, va = random_csr(rng, rows, cols, 0.01, empty_rows=0.1, heavy_rows=3)
    csr = acc.CSR(rows, cols, rp, ci, va)
    mx, fr, _ = run_case(csr, K, form, seed=K)
    assert mx <= TOL[form], (mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("nb", [1, 2, 4, 8])
@pytest.mark.parametrize("split", [0, 4])
@pytest.mark.parametrize("xcd", [0, 1])
@pytest.mark.parametrize("pack", [0, 64])
def test_tile_variants(form, nb, split, xcd, pack):
    rng = np.random.default_rng(nb * 10 + split)
    rows, cols = 777, 900
    rp, ci, va = random_csr(rng, rows, cols, 0.02, empty_rows=0.05, heavy_rows=4)
    csr = acc.CSR(rows, cols, rp, ci, va)
    tile = acc.Tile(nb, 4, split, xcd, pack)
    mx, fr, _ = run_case(csr, 128, form, tile=tile, seed=3)
    assert mx <= TOL[form], (mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("panels,split", [(2, 0), (5, 0), (5, 20), (32, 0), (300, 12)])
def test_column_panels(form, panels, split):
    # ~200 tiles per window: every window is cut at the panel boundaries
    # (300 panels leave pieces below the minimum length, which are merged)
    rng = np.random.default_rng(panels + split)
    rows, cols = 603, 5000
    rp, ci, va = random_csr(rng, rows, cols, 0.05, empty_rows=0.05, heavy_rows=3)
    csr = acc.CSR(rows, cols, rp, ci, va)
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        for K in (32, 100):
            tile = acc.Tile(2, 4, split, 0, 64, panels)
            mx, fr, Cg = run_case(csr, K, form, tile=tile, rw=rw, seed=K, graph=g)
            assert mx <= TOL[form], (rw, K, mx, fr)
            # a second launch (counters reset by the last piece) is bit-equal
            _, _, Cg2 = run_case(csr, K, form, tile=tile, rw=rw, seed=K, graph=g)
            assert np.array_equal(Cg, Cg2)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("panels,split,nb,wpb", [(8, 0, 2, 4), (16, 20, 4, 1), (3, 0, 8, 1),
                                                 (300, 12, 2, 2)])
def test_panels_pinned_to_xcds(form, panels, split, nb, wpb):
    # xcd_slice 3 runs panel p on XCD p % 8 (per-XCD unit lists padded with
    # no-op units): the same pieces and summation order as round-robin
    # dispatch of that schedule, so C is bit-equal to it and within the
    # form's tolerance of FP64; with one panel the mode falls back to 0
    rng = np.random.default_rng(panels * 7 + nb)
    rows, cols = 603, 5000
    rp, ci, va = random_csr(rng, rows, cols, 0.05, empty_rows=0.05, heavy_rows=3)
    csr = acc.CSR(rows, cols, rp, ci, va)
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        for K in (32, 100, 128):
            pinned = acc.Tile(nb, wpb, split, 3, 64, panels)
            plain = acc.Tile(nb, wpb, split, 0, 64, panels)
            mx, fr, Cp = run_case(csr, K, form, tile=pinned, rw=rw, seed=K, graph=g)
            assert mx <= TOL[form], (rw, K, mx, fr)
            _, _, Cr = run_case(csr, K, form, tile=plain, rw=rw, seed=K, graph=g)
            assert np.array_equal(Cp, Cr), (rw, K)
    one = acc.Tile(2, 4, 0, 3, 64, 1)
    mx, _, _ = run_case(csr, 64, form, tile=one, seed=1)
    assert mx <= TOL[form]


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("case", ["long_windows", "many_windows"])
def test_sweep_equals_work_units(form, case):
    # The resident-window sweep runs the work-unit schedule's pieces in
    # another order with the same sums: C must be bit-identical.
    # "many_windows" has more windows than one round of the sweep holds.
    torch = _torch()
    rng = np.random.default_rng(77 if case == "long_windows" else 78)
    if case == "long_windows":
        rows, cols, dens = 803, 6000, 0.04
    else:
        rows, cols, dens = 180001, 9000, 0.0006
    rp, ci, va = random_csr(rng, rows, cols, dens, empty_rows=0.05, heavy_rows=3)
    csr = acc.CSR(rows, cols, rp, ci, va)
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        for K, nb, split, panels in ((100, 2, 0, 4), (128, 4, 24, 8), (32, 1, 0, 0)):
            B = torch.from_numpy(acc.gen_dense(K, cols, K)).cuda()
            outs = []
            for sweep in (0, 1):
                tile = acc.Tile(nb, 4, split, 0, 64, panels, sweep)
                outs.append(g.spmm(B, None, form, tile).clone())
            torch.cuda.synchronize()
            assert torch.equal(outs[0], outs[1]), (rw, K, nb, split, panels)
        if case == "long_windows":
            mx, fr, _ = run_case(csr, 128, form, tile=acc.Tile(4, 8, 0, 0, 64, 8, 1), rw=rw,
                                 seed=5, graph=g)
            assert mx <= TOL[form], (rw, mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_row_window_16(form):
    rng = np.random.default_rng(16)
    rp, ci, va = random_csr(rng, 501, 640, 0.02, empty_rows=0.1, heavy_rows=2)
    csr = acc.CSR(501, 640, rp, ci, va)
    for tile in (None, acc.Tile(4, 2, 3, 0)):
        mx, fr, _ = run_case(csr, 64, form, tile=tile, rw=16, seed=5)
        assert mx <= TOL[form], (mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_chung_lu_c0(form):
    # 16,384-node Chung-Lu (exponent 2.1, mean degree 16), values U(0,1], K=32
    csr = acc.gen_chung_lu(16384, 16.0, 2.1, 1, 0)
    mx, fr, _ = run_case(csr, 32, form, seed=2)
    assert mx <= TOL[form], (mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_stencil(form):
    csr = acc.gen_stencil27(24, 13)
    mx, fr, _ = run_case(csr, 128, form, seed=14)
    assert mx <= TOL[form], (mx, fr)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("pack", [8, 64, 1000])
def test_packed_units_with_empty_windows(form, pack):
    # short windows, runs of empty windows (rows 16..95 = windows 2..11) and
    # a ragged last window
    rows, cols = 1003, 700
    rp0, ci0, va0 = random_csr(np.random.default_rng(pack), rows, cols, 0.004,
                               empty_rows=0.6)
    cnt = np.diff(rp0.astype(np.int64))
    starts = rp0[:-1].astype(np.int64)
    sel = np.concatenate([np.arange(s0, s0 + c) for r, (s0, c) in enumerate(zip(starts, cnt))
                          if not 16 <= r < 96])
    cnt[16:96] = 0
    rp = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32)
    csr = acc.CSR(rows, cols, rp, ci0[sel], va0[sel])
    for K in (16, 100):
        mx, fr, Cg = run_case(csr, K, form, tile=acc.Tile(2, 4, 0, 0, pack), seed=K)
        assert mx <= TOL[form], (mx, fr)
        assert np.all(Cg[16:96] == 0)


def test_rectangular_and_strided():
    torch = _torch()
    rng = np.random.default_rng(9)
    rp, ci, va = random_csr(rng, 300, 1700, 0.01)
    csr = acc.CSR(300, 1700, rp, ci, va)
    g = acc.DeviceGraph(csr)
    B = acc.gen_dense(4, 1700, 48)
    Bp = np.zeros((1700, 52), np.float32)
    Bp[:, :48] = B
    Bd = torch.from_numpy(Bp).cuda()[:, :48]
    Cd = torch.full((300, 56), 7.0, device="cuda")
    g.spmm(Bd, Cd[:, :48], acc.FORM_FP32)
    torch.cuda.synchronize()
    out = Cd.cpu().numpy()
    assert np.all(out[:, 48:] == 7.0)  # padding columns untouched
    C64 = O.spmm_f64(rp, ci, va, B)
    ab = O.spmm_abs_f64(rp, ci, va, B)
    assert O.errors(out[:, :48], C64, ab)[0] <= 1e-5


def test_deterministic_and_split_consistent():
    torch = _torch()
    csr = acc.gen_chung_lu(8000, 40.0, 2.0, 21, 0)
    g = acc.DeviceGraph(csr)
    B = torch.from_numpy(acc.gen_dense(5, csr.cols, 64)).cuda()
    outs = []
    for _ in range(3):
        outs.append(g.spmm(B, None, acc.FORM_SPLIT_BF16, acc.Tile(4, 4, 8, 0)).cpu())
    assert all(torch.equal(outs[0], o) for o in outs[1:])


def test_invalid_arguments():
    torch = _torch()
    rng = np.random.default_rng(1)
    rp, ci, va = random_csr(rng, 50, 50, 0.1)
    g = acc.DeviceGraph(acc.CSR(50, 50, rp, ci, va))
    B = torch.zeros((50, 8), device="cuda")
    Cm = torch.zeros((50, 8), device="cuda")
    L = acc.lib()
    assert L.accspmm_spmm(g.handle, B.data_ptr(), 8, Cm.data_ptr(), 8, 0, 0, None,
                          None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_spmm(g.handle, B.data_ptr(), 4, Cm.data_ptr(), 8, 8, 0, None,
                          None) == acc.Status.ERR_SHAPE
    assert L.accspmm_spmm(g.handle, B.data_ptr(), 8, Cm.data_ptr(), 8, 8, 7, None,
                          None) == acc.Status.ERR_INVALID_ARG


@pytest.mark.parametrize("rw", [8, 16])
@pytest.mark.parametrize("b_split", [1, 2])
def test_split_single_products(rw, b_split):
    # rows of one or two nonzeros: C is one or two products, so the split's
    # representation error is not averaged away.  Two-term splits of both A
    # and B (or a missing A2 B0 product) are off by up to ~1.5e-5 on such
    # rows; three-term A times two-term B stays within 2^-17 + 2^-25 plus
    # the FP32 rounding of a five-term sum (< 8.5e-6).
    torch = _torch()
    rng = np.random.default_rng(70 + rw)
    rows = cols = 40000
    deg = rng.integers(1, 3, rows)
    rp = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(cols, d, replace=False)) for d in deg]).astype(np.uint32)
    # values spread over many binades, mantissas uniform
    va = (rng.uniform(1, 2, rp[-1]) * np.exp2(rng.integers(-20, 20, rp[-1]))).astype(np.float32)
    va *= np.where(rng.random(rp[-1]) < 0.5, -1, 1).astype(np.float32)
    csr = acc.CSR(rows, cols, rp, ci, va)
    B = (rng.uniform(1, 2, (cols, 64)) * np.exp2(rng.integers(-20, 20, (cols, 64)))).astype(np.float32)
    B *= np.where(rng.random((cols, 64)) < 0.5, -1, 1).astype(np.float32)
    for tile in (acc.Tile(1, 4, 0, 0, 0, 0, 0, b_split),
                 acc.Tile(4, 1, 8, 1, 64, 2, 0, b_split), acc.Tile(8, 1, 0, 0, 64, 0, 0, b_split)):
        mx, fr, _ = run_case(csr, 64, acc.FORM_SPLIT_BF16, tile=tile, rw=rw, B=B)
        assert mx <= 8.5e-6, (rw, tile.as_tuple(), mx)


@pytest.mark.parametrize("form,b_split", [(acc.FORM_FP32, 0), (acc.FORM_SPLIT_BF16, 1),
                                          (acc.FORM_SPLIT_BF16, 2)])
def test_row0_nonfinite_masked(form, b_split):
    # padded tile columns gather B's row 0 (their A value is zero); an
    # infinity or NaN there must not reach C.  Column 0 holds no nonzero, so
    # C equals the product with row 0 zeroed, bit for bit, on every schedule
    torch = _torch()
    rng = np.random.default_rng(62)
    rows, cols = 1501, 2300
    rp, ci, va = random_csr(rng, rows, cols, 0.03, empty_rows=0.1, heavy_rows=3)
    keep = ci != 0
    ck = np.concatenate([[0], np.cumsum(keep)])
    rp2 = ck[rp.astype(np.int64)].astype(np.uint32)
    csr = acc.CSR(rows, cols, rp2, ci[keep], va[keep])
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        for K in (8, 17, 64, 128):
            Bh = acc.gen_dense(K, cols, K)
            Bz = Bh.copy()
            Bz[0, :] = 0.0
            Bh[0, :] = np.inf
            Bh[0, ::3] = np.nan
            for tile in (acc.Tile(1, 4, 0, 0, 0, 0, 0, b_split),
                         acc.Tile(2, 4, 16, 0, 64, 0, 0, b_split),
                         acc.Tile(4, 4, 0, 0, 64, 4, 0, b_split),
                         acc.Tile(8, 2, 24, 0, 64, 2, 0, b_split),
                         acc.Tile(4, 4, 0, 0, 64, 4, 1, b_split)):
                ref = g.spmm(torch.from_numpy(Bz).cuda(), None, form, tile).clone()
                got = g.spmm(torch.from_numpy(Bh).cuda(), None, form, tile).clone()
                torch.cuda.synchronize()
                assert torch.equal(ref, got), (rw, K, tile.as_tuple())
        C64 = O.spmm_f64(rp2, ci[keep], va[keep], Bz)
        ab = O.spmm_abs_f64(rp2, ci[keep], va[keep], Bz)
        assert O.errors(got.cpu().numpy(), C64, ab)[0] <= TOL[form]


@pytest.mark.parametrize("form", [acc.FORM_SPLIT_BF16, acc.FORM_BF16], ids=acc.FORM_NAMES.get)
def test_two_streams_share_a_handle(form):
    # one handle used from two streams at once with different B: each
    # stream's calls keep their own workspace (split-window partials, the
    # BF16 form's converted B), so every C equals its standalone result
    torch = _torch()
    rng = np.random.default_rng(64)
    rows, cols, K = 3000, 4000, 96
    rp, ci, va = random_csr(rng, rows, cols, 0.02, empty_rows=0.05, heavy_rows=6)
    g = acc.DeviceGraph(acc.CSR(rows, cols, rp, ci, va))
    tile = acc.Tile(2, 4, 8, 0, 64, 2)  # split windows and column panels
    Bs = [torch.from_numpy(acc.gen_dense(s, cols, K)).cuda() for s in (1, 2)]
    refs = [g.spmm(B, None, form, tile).clone() for B in Bs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [[torch.full((rows, K), float("nan"), device="cuda") for _ in range(8)]
            for _ in Bs]
    for rep in range(8):
        for s, B, Cl in zip(streams, Bs, outs):
            g.spmm(B, Cl[rep], form, tile, stream=s)
    torch.cuda.synchronize()
    for Cl, R in zip(outs, refs):
        assert all(torch.equal(C, R) for C in Cl)


@pytest.mark.parametrize("rw", [8, 16])
def test_split_in_registers_equals_conversion(rw):
    # b_split 2 (the kernel splits the FP32 B) and 1 (a conversion pass)
    # round B the same way: C must be bit-identical, K tails included
    torch = _torch()
    rng = np.random.default_rng(60 + rw)
    rows, cols = 1501, 2300
    rp, ci, va = random_csr(rng, rows, cols, 0.03, empty_rows=0.1, heavy_rows=3)
    g = acc.DeviceGraph(acc.CSR(rows, cols, rp, ci, va), acc.default_build_opts(rw))
    for K in (8, 17, 64, 100, 128):
        Bh = acc.gen_dense(K, cols, K)
        B = torch.from_numpy(Bh).cuda()
        for nb, split, panels, sweep in ((1, 0, 0, 0), (2, 16, 0, 0), (4, 0, 4, 0),
                                         (8, 24, 2, 0), (4, 0, 4, 1)):
            outs = []
            for bs in (1, 2):
                tile = acc.Tile(nb, 4, split, 0, 64, panels, sweep, bs)
                outs.append(g.spmm(B, None, acc.FORM_SPLIT_BF16, tile).clone())
            torch.cuda.synchronize()
            assert torch.equal(outs[0], outs[1]), (K, nb, split, panels, sweep)
        C64 = O.spmm_f64(rp, ci, va, Bh)
        ab = O.spmm_abs_f64(rp, ci, va, Bh)
        mx, fr = O.errors(outs[1].cpu().numpy(), C64, ab)
        assert mx <= TOL[acc.FORM_SPLIT_BF16], (K, mx)


def test_unaligned_b_view_narrows_nb():
    # a B view with ldb % NB != 0 cannot take NB-wide loads: the library
    # narrows NB (ldb = 66 -> NB 2), within tolerance of the oracle
    torch = _torch()
    rng = np.random.default_rng(61)
    rp, ci, va = random_csr(rng, 400, 500, 0.05)
    g = acc.DeviceGraph(acc.CSR(400, 500, rp, ci, va))
    Bh = acc.gen_dense(3, 500, 66)
    Bfull = torch.from_numpy(Bh).cuda()
    Bv = Bfull[:, :64]                                  # ldb = 66
    for form in FORMS:
        got = g.spmm(Bv, None, form, acc.Tile(4, 4, 0, 0, 64)).cpu().numpy()
        B64 = np.ascontiguousarray(Bh[:, :64])
        C64 = O.spmm_f64(rp, ci, va, B64)
        ab = O.spmm_abs_f64(rp, ci, va, B64)
        assert O.errors(got, C64, ab)[0] <= TOL[form], form


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_edge_shapes(form):
    # an all-empty matrix writes exact zeros over a NaN-filled C; a 1x1
    # matrix; one row of 3,000 nonzeros (one window, split into pieces);
    # explicit zero values; K far above the reference's 512
    torch = _torch()
    empty = acc.CSR(19, 7, np.zeros(20, np.uint32), np.zeros(0, np.uint32),
                    np.zeros(0, np.float32))
    g = acc.DeviceGraph(empty)
    for K in (1, 16, 33):
        B = torch.from_numpy(acc.gen_dense(K, 7, K)).cuda()
        C = torch.full((19, K), float("nan"), device="cuda")
        g.spmm(B, C, form)
        torch.cuda.synchronize()
        assert torch.count_nonzero(C).item() == 0 and not torch.isnan(C).any()
    one = acc.CSR(1, 1, np.array([0, 1], np.uint32), np.array([0], np.uint32),
                  np.array([2.5], np.float32))
    for K in (1, 8):
        mx, _, Cg = run_case(one, K, form, seed=K)
        assert mx <= TOL[form], (K, mx)
    rng = np.random.default_rng(95)
    ci = np.sort(rng.choice(5000, 3000, replace=False)).astype(np.uint32)
    va = rng.uniform(-1, 1, 3000).astype(np.float32)
    va[::7] = 0.0                                            # stored zeros
    row = acc.CSR(1, 5000, np.array([0, 3000], np.uint32), ci, va)
    for tile in (None, acc.Tile(4, 4, 64, 0, 64)):
        mx, _, _ = run_case(row, 64, form, tile=tile, seed=4)
        assert mx <= TOL[form], (tile, mx)
    rp, ci, va = random_csr(rng, 300, 400, 0.03, empty_rows=0.1, heavy_rows=2)
    wide = acc.CSR(300, 400, rp, ci, va)
    for K in (1024, 2000):
        mx, _, _ = run_case(wide, K, form, seed=K)
        assert mx <= TOL[form], (K, mx)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("K", [24, 32, 40, 64, 96, 100])
def test_two_row_gathers_match_row_loads(form, K):
    # 128-byte chunk rows (NB 2 with 4-byte B, NB 4 with BF16): the work-unit
    # kernel loads two rows per 16-byte buffer load and lanes n, n ^ 8 swap
    # halves; the sweep kernel loads each lane's features of each row on
    # their own.  Same C bits (chunk tails included), and within the form's
    # tolerance of the FP64 oracle
    torch = _torch()
    rng = np.random.default_rng(1000 + K)
    rows, cols = 700, 1500
    rp, ci, va = random_csr(rng, rows, cols, 0.02, empty_rows=0.1, heavy_rows=2)
    csr = acc.CSR(rows, cols, rp, ci, va)
    nb = 4 if form == acc.FORM_BF16 else 2
    bs = (1, 2) if form == acc.FORM_SPLIT_BF16 else (0,)
    Bh = acc.gen_dense(K, cols, K)
    B = torch.from_numpy(Bh).cuda()
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        for b in bs:
            outs = [g.spmm(B, None, form, acc.Tile(nb, 2, 16, 0, 64, 2, sw, b)).clone()
                    for sw in (0, 1)]
            torch.cuda.synchronize()
            assert torch.equal(outs[0], outs[1]), (rw, b)
            mx, fr, _ = run_case(csr, K, form, tile=acc.Tile(nb, 2, 16, 0, 64, 2, 0, b), rw=rw,
                                 graph=g, B=Bh)
            assert mx <= TOL[form], (rw, b, mx)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_dense_groups(form):
    # 16-row tiles about 90% full: 4-tile groups of up to 512 values, more
    # than the 256 one span load brings (and full 8-row groups of 256).
    # Every schedule shape against the oracle; the sweep bit-equal to the
    # work units; the fullest group reported as the oracle counts it
    torch = _torch()
    from helpers import dense_block_csr, max_group_values
    rng = np.random.default_rng(97)
    rp, ci, va = dense_block_csr(rng, 300, 900, 120, 0.9)
    csr = acc.CSR(300, 900, rp, ci, va)
    full = acc.CSR(64, 64, *dense_block_csr(rng, 64, 64, 64, 1.0))
    bs = (0,) if form != acc.FORM_SPLIT_BF16 else (1, 2)
    for rw in (16, 8):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw))
        assert g.info.max_group_values == max_group_values(rw, 300, rp, ci, va)
        assert (g.info.max_group_values > 256) == (rw == 16)
        for K in (16, 100):
            for b in bs:
                for tile in (acc.Tile(2, 4, 0, 0, 64, 0, 0, b), acc.Tile(4, 4, 8, 0, 0, 0, 0, b),
                             acc.Tile(1, 8, 0, 1, 64, 3, 0, b), acc.Tile(8, 2, 4, 2, 64, 0, 0, b)):
                    mx, fr, _ = run_case(csr, K, form, tile=tile, rw=rw, seed=K, graph=g)
                    assert mx <= TOL[form], (rw, K, tile.as_tuple(), mx)
        B = torch.from_numpy(acc.gen_dense(7, 900, 64)).cuda()
        outs = [g.spmm(B, None, form, acc.Tile(2, 4, 8, 0, 64, 2, sw)).clone() for sw in (0, 1)]
        torch.cuda.synchronize()
        assert torch.equal(outs[0], outs[1]), rw
        gf = acc.DeviceGraph(full, acc.default_build_opts(rw))
        assert gf.info.max_group_values == (512 if rw == 16 else 256)
        mx, fr, _ = run_case(full, 32, form, rw=rw, seed=5, graph=gf)
        assert mx <= TOL[form], (rw, mx)


def test_operand_checks():
    # the wrappers refuse operands that would make the kernels read or write
    # outside the caller's buffers, before any library call
    torch = _torch()
    rng = np.random.default_rng(96)
    rp, ci, va = random_csr(rng, 50, 70, 0.1)
    g = acc.DeviceGraph(acc.CSR(50, 70, rp, ci, va))
    B = torch.zeros((70, 16), device="cuda")
    Cm = torch.zeros((50, 16), device="cuda")
    bad = [
        (torch.zeros((69, 16), device="cuda"), Cm, ValueError),      # B short of rows
        (B, torch.zeros((49, 16), device="cuda"), ValueError),       # C short of rows
        (B, torch.zeros((50, 8), device="cuda"), ValueError),        # K mismatch
        (torch.zeros((16, 70), device="cuda").t(), Cm, ValueError),  # column stride 16
        (B.cpu(), Cm, TypeError),                                    # host memory
        (B.double(), Cm, TypeError),                                 # float64
        (B, Cm.cpu(), TypeError),
    ]
    for Bx, Cx, err in bad:
        with pytest.raises(err):
            g.spmm(Bx, Cx, acc.FORM_FP32)
        with pytest.raises(err):
            acc.Batch([(g, Bx, Cx)], acc.FORM_FP32)
    srv = acc.Server(max_inflight=1)
    gid = srv.register(acc.CSR(50, 70, rp, ci, va))
    with pytest.raises(ValueError):
        srv.submit(gid, torch.zeros((69, 16), device="cuda"), Cm)
    srv.close()
    # the checked call itself still runs
    assert torch.equal(g.spmm(B, Cm, acc.FORM_FP32), torch.zeros((50, 16), device="cuda"))


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_prepare_then_prepared_equals_spmm(form):
    # the two halves (B conversion, then the kernel on the converted B) give
    # spmm's C bits, with the library's tile and with explicit ones
    torch = _torch()
    rng = np.random.default_rng(93)
    rp, ci, va = random_csr(rng, 800, 1100, 0.02, empty_rows=0.1, heavy_rows=3)
    g = acc.DeviceGraph(acc.CSR(800, 1100, rp, ci, va))
    bs = (0,) if form != acc.FORM_SPLIT_BF16 else (1, 2)
    for K in (24, 128):
        B = torch.from_numpy(acc.gen_dense(K + 1, 1100, K)).cuda()
        tiles = [None] + [acc.Tile(2, 4, 16, 0, 64, 2, 0, b) for b in bs]
        for tile in tiles:
            ref = g.spmm(B, None, form, tile).clone()
            Cm = torch.full((800, K), float("nan"), device="cuda")
            g.prepare_b(B, form, tile)
            g.spmm_prepared(B, Cm, form, tile)
            torch.cuda.synchronize()
            assert torch.equal(Cm, ref), (K, tile and tile.as_tuple())


def test_autotune_cache_file(tmp_path):
    # the tuner's pick runs within tolerance, becomes the library's tile for
    # that key, and is written to the cache file beside the file's other
    # entries; a fresh process reading the file through ACCSPMM_TUNE_CACHE
    # takes the valid entry and ignores a malformed one
    import json
    import subprocess
    import sys
    torch = _torch()
    rng = np.random.default_rng(91)
    rp, ci, va = random_csr(rng, 900, 1200, 0.02, empty_rows=0.05, heavy_rows=2)
    csr = acc.CSR(900, 1200, rp, ci, va)
    g = acc.DeviceGraph(csr)
    form, K = acc.FORM_SPLIT_BF16, 64
    path = tmp_path / "tune.json"
    path.write_text('{\n  "f9-k1-foreign": [2, 4, 0, 0, 64, 0, 0, 0]\n}\n')
    Bh = acc.gen_dense(5, 1200, K)
    B = torch.from_numpy(Bh).cuda()
    Cm = torch.empty((900, K), device="cuda")
    tile, ms = g.autotune(B, Cm, form, str(path))
    assert ms > 0
    assert g.tile_select(K, form).as_tuple() == tile.as_tuple()
    mx, _, _ = run_case(csr, K, form, tile=tile, seed=5, graph=g, B=Bh)
    assert mx <= TOL[form]
    saved = json.loads(path.read_text())
    key = g.tune_key(K, form)
    assert saved[key] == list(tile.as_tuple()) and "f9-k1-foreign" in saved
    # a malformed entry (NB = 3) for another K is ignored by a fresh process
    saved[g.tune_key(32, form)] = [3, 4, 0, 0, 64, 0, 0, 0]
    # the library's layout: one entry per line
    path.write_text("{\n" + ",\n".join(f"  {json.dumps(k)}: {json.dumps(v)}"
                                         for k, v in saved.items()) + "\n}\n")
    code = (
        "import sys; sys.path.insert(0, %r); import accspmm_amd as acc, numpy as np\n"
        "from helpers import random_csr\n"
        "rng = np.random.default_rng(91)\n"
        "rp, ci, va = random_csr(rng, 900, 1200, 0.02, empty_rows=0.05, heavy_rows=2)\n"
        "g = acc.DeviceGraph(acc.CSR(900, 1200, rp, ci, va))\n"
        "print(list(g.tile_select(64, 1).as_tuple()), list(g.tile_select(32, 1).as_tuple()))\n"
        % str(REPO_TESTS))
    env = dict(os.environ, ACCSPMM_TUNE_CACHE=str(path))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=120, env=env, cwd=str(REPO_TESTS.parent))
    assert r.returncode == 0, r.stderr
    got64, got32 = json.loads("[" + r.stdout.strip().splitlines()[-1].replace("] [", "], [")
                              + "]")
    assert got64 == list(tile.as_tuple())
    assert got32[0] in (1, 2, 4, 8) and got32 != [3, 4, 0, 0, 64, 0, 0, 0]
