"""GPU METCF / AdpBME / IBD builder: bit-exact against the CPU oracle and the
hand-worked known answers."""
import numpy as np
import pytest

import accspmm_amd as acc
from helpers import csr_from_window_cols, kat, random_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu
K = kat()


def _build(rows, cols, rp, ci, va, rw=8, **kw):
    # AdpBME groups are built on request (they are exported, not computed on)
    csr = acc.CSR(rows, cols, rp, ci, va)
    kw.setdefault("build_adpbme", 1)
    g = acc.DeviceGraph(csr, acc.default_build_opts(rw, **kw))
    return g, g.export()


def _assert_same(got, exp, keys):
    for k in keys:
        assert got[k].dtype == exp[k].dtype, k
        assert np.array_equal(got[k].view(np.uint8), exp[k].view(np.uint8)), k


ALL = ("rowWindowOffset", "tcOffset", "tcLocalBit", "sparseA2B", "data",
       "groupOffset", "rowIndices")


@pytest.mark.parametrize("rw", [8, 16])
def test_known_answer(rw):
    a = K["a10"]
    g, got = _build(a["rows"], a["cols"], a["row_ptr"], a["col_idx"], a["val"], rw)
    exp = K[f"metcf_rw{rw}"]
    for key in ("rowWindowOffset", "sparseA2B", "tcOffset", "tcLocalBit",
                "groupOffset", "rowIndices"):
        assert got[key].tolist() == exp[key], key
    assert got["data"].tolist() == [float(x) for x in exp["data"]]
    inf = g.info
    assert (inf.ibd, bool(inf.balanced)) == (exp["ibd"], exp["balanced"])


@pytest.mark.parametrize("case", K["adpbme_groups"] + [{"windows": [0, 0, 5]}],
                         ids=lambda c: str(c["windows"]))
def test_adpbme_groups(case):
    rows, cols, rp, ci, va = csr_from_window_cols(case["windows"])
    g, got = _build(rows, cols, rp, ci, va)
    exp = O.adpbme_build(8, rows, rp, ci, va)
    _assert_same(got, exp, ALL)
    if "groupOffset" in case:
        assert got["groupOffset"].tolist() == case["groupOffset"]


@pytest.mark.parametrize("case", K["ibd_cases"], ids=lambda c: str(c["tiles"]))
def test_ibd(case):
    rows, cols, rp, ci, va = csr_from_window_cols([8 * t for t in case["tiles"]])
    g, _ = _build(rows, cols, rp, ci, va)
    assert (g.info.ibd, bool(g.info.balanced)) == (case["ibd"], case["balanced"])


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("rw", [8, 16])
def test_random_matches_oracle(seed, rw):
    rng = np.random.default_rng(100 + seed)
    rows = int(rng.integers(1, 2000))
    cols = int(rng.integers(1, 3000))
    rp, ci, va = random_csr(rng, rows, cols, float(rng.uniform(0.0005, 0.02)),
                            empty_rows=0.15, heavy_rows=3)
    g, got = _build(rows, cols, rp, ci, va, rw)
    exp = O.adpbme_build(rw, rows, rp, ci, va)
    _assert_same(got, exp, ALL)
    ibd, bal = O.ibd(exp["rowWindowOffset"])
    assert (g.info.ibd, bool(g.info.balanced)) == (ibd, bal)
    from helpers import max_group_values
    assert g.info.max_group_values == max_group_values(rw, rows, rp, ci, va)


@pytest.mark.parametrize("t", [(1, 1, 1, 4), (2, 2, 0, 8), (5, 0, 3, 10)])
def test_adpbme_cost_params(t):
    rng = np.random.default_rng(5)
    rp, ci, va = random_csr(rng, 500, 800, 0.01, heavy_rows=2)
    g, got = _build(500, 800, rp, ci, va, 8, t_load=t[0], t_comp=t[1], t_write=t[2],
                    target_num=t[3])
    exp = O.adpbme_build(8, 500, rp, ci, va, *t)
    _assert_same(got, exp, ("groupOffset", "rowIndices"))


def test_chung_lu_matches_oracle():
    csr = acc.gen_chung_lu(16384, 16.0, 2.1, 1, 0)
    g = acc.DeviceGraph(csr, acc.default_build_opts(8, build_adpbme=1))
    got = g.export()
    exp = O.adpbme_build(8, csr.rows, csr.row_ptr, csr.col_idx, csr.data)
    _assert_same(got, exp, ALL)


def test_stencil_matches_oracle():
    csr = acc.gen_stencil27(20, 13)
    for rw in (8, 16):
        g = acc.DeviceGraph(csr, acc.default_build_opts(rw, build_adpbme=1))
        exp = O.adpbme_build(rw, csr.rows, csr.row_ptr, csr.col_idx, csr.data)
        _assert_same(g.export(), exp, ALL)


def test_reference_named_api():
    a = K["a10"]
    csr = acc.CSR(a["rows"], a["cols"], a["row_ptr"], a["col_idx"], a["val"])
    m = acc.METCFBit()
    m.convertFromCSR(csr)
    assert m.rowWindowOffset.tolist() == K["metcf_rw8"]["rowWindowOffset"]
    assert acc.balanceStrategy(m, m.tcOffset.size - 1, m.rowWindowOffset.size - 1) is False
    adp = acc.AdpBME()
    adp.CSR2AdpBME(csr, 3, 1, 2)
    assert adp.groupOffset.tolist() == K["metcf_rw8"]["groupOffset"]


def test_rejects_non_canonical():
    rp = np.array([0, 2, 3], np.uint32)
    for ci in ([3, 1, 0], [1, 1, 0], [0, 9, 0]):  # unsorted, duplicate, out of range
        csr = acc.CSR(2, 5, rp, np.array(ci, np.uint32), np.ones(3, np.float32))
        with pytest.raises(acc.AccSpMMError) as e:
            acc.DeviceGraph(csr)
        assert e.value.status in (acc.Status.ERR_NOT_CANONICAL, acc.Status.ERR_INVALID_ARG)
    bad_rp = acc.CSR(2, 5, np.array([0, 3, 2], np.uint32), np.array([0, 1], np.uint32),
                     np.ones(2, np.float32))
    with pytest.raises(acc.AccSpMMError):
        acc.DeviceGraph(bad_rp)


def test_empty_matrix_rows_without_nonzeros():
    rp = np.zeros(20, np.uint32)
    csr = acc.CSR(19, 7, rp, np.zeros(0, np.uint32), np.zeros(0, np.float32))
    g = acc.DeviceGraph(csr, acc.default_build_opts(8, build_adpbme=1))
    got = g.export()
    exp = O.adpbme_build(8, 19, rp, [], [])
    _assert_same(got, exp, ALL)
    assert g.info.num_tiles == 0


@pytest.mark.parametrize("rw,adpbme", [(8, 1), (8, 0), (16, 1)])
def test_save_load_round_trip(tmp_path, rw, adpbme):
    # a saved graph loads back to the same arrays and the same SpMM bits
    import torch
    rng = np.random.default_rng(40 + rw)
    rows, cols = 2003, 1500
    rp, ci, va = random_csr(rng, rows, cols, 0.01, empty_rows=0.1, heavy_rows=2)
    g = acc.DeviceGraph(acc.CSR(rows, cols, rp, ci, va),
                        acc.default_build_opts(rw, build_adpbme=adpbme))
    path = tmp_path / "g.metcf"
    g.save(path)
    h = acc.DeviceGraph.load(path)
    a, b = g.export(), h.export()
    assert sorted(a) == sorted(b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    ia, ib = g.info, h.info
    for f in ("rows", "cols", "nnz", "row_window", "num_windows", "num_tiles", "num_groups",
              "ibd", "balanced", "max_window_tiles", "max_group_values"):
        assert getattr(ia, f) == getattr(ib, f), f
    assert g.tune_key(64) == h.tune_key(64)
    B = torch.from_numpy(acc.gen_dense(3, cols, 64)).cuda()
    for form in (acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16):
        assert torch.equal(g.spmm(B, None, form), h.spmm(B, None, form))


def test_load_rejects_bad_files(tmp_path):
    rng = np.random.default_rng(41)
    rp, ci, va = random_csr(rng, 300, 300, 0.02)
    g = acc.DeviceGraph(acc.CSR(300, 300, rp, ci, va))
    good = tmp_path / "g.metcf"
    g.save(good)
    raw = bytearray(good.read_bytes())
    cases = {
        "flipped": raw[:100] + bytes([raw[100] ^ 1]) + raw[101:],
        "truncated": raw[:-9],
        "extended": raw + b"\0",
        "magic": b"XCCSPMTC" + raw[8:],
    }
    for name, data in cases.items():
        bad = tmp_path / f"{name}.metcf"
        bad.write_bytes(bytes(data))
        with pytest.raises(acc.AccSpMMError) as e:
            acc.DeviceGraph.load(bad)
        assert e.value.status == acc.Status.ERR_FILE, name
    with pytest.raises(acc.AccSpMMError):
        acc.DeviceGraph.load(tmp_path / "missing.metcf")
