"""CPU oracle pinned against hand-worked known answers, invariants, brute force
and scipy.  No GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import csr_from_window_cols, kat, random_csr
from oracle import oracle as O

K = kat()


def _a10():
    a = K["a10"]
    return a["rows"], a["row_ptr"], a["col_idx"], a["val"]


@pytest.mark.parametrize("rw", [8, 16])
def test_metcf_known_answer(rw):
    rows, rp, ci, va = _a10()
    exp = K[f"metcf_rw{rw}"]
    got = O.metcf_build(rw, rows, rp, ci, va)
    for key in ("rowWindowOffset", "sparseA2B", "tcOffset", "tcLocalBit"):
        assert got[key].tolist() == exp[key], key
    assert got["data"].tolist() == [float(x) for x in exp["data"]]
    adp = O.adpbme_build(rw, rows, rp, ci, va)
    assert adp["groupOffset"].tolist() == exp["groupOffset"]
    assert adp["rowIndices"].tolist() == exp["rowIndices"]
    # AdpBME tiles are the METCF tiles
    for key in ("sparseA2B", "tcOffset", "tcLocalBit", "data"):
        assert np.array_equal(adp[key], got[key])
    assert O.ibd(got["rowWindowOffset"]) == (exp["ibd"], exp["balanced"])


@pytest.mark.parametrize("case", K["adpbme_groups"], ids=lambda c: str(c["windows"]))
def test_adpbme_group_known_answers(case):
    rows, cols, rp, ci, va = csr_from_window_cols(case["windows"])
    got = O.adpbme_build(8, rows, rp, ci, va)
    assert got["groupOffset"].tolist() == case["groupOffset"]


@pytest.mark.parametrize("case", K["ibd_cases"], ids=lambda c: str(c["tiles"]))
def test_ibd_known_answers(case):
    rwo = np.concatenate([[0], np.cumsum(case["tiles"])]).astype(np.uint32)
    assert O.ibd(rwo) == (case["ibd"], case["balanced"])
    # the same through a built METCF
    rows, cols, rp, ci, va = csr_from_window_cols([8 * t for t in case["tiles"]])
    m = O.metcf_build(8, rows, rp, ci, va)
    assert m["rowWindowOffset"].tolist() == rwo.tolist()


def test_ibd_excludes_last_window():
    # a plausible mistake (summing every window) flips [1,1,1,40] to True
    rwo = np.array([0, 1, 2, 3, 43], np.uint32)
    W, T = 4, 43
    avg = T // W
    all_windows = sum(abs(int(rwo[i + 1] - rwo[i]) - avg) for i in range(W)) // W
    assert all_windows > 8
    assert O.ibd(rwo) == (6, False)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


@pytest.mark.parametrize("name", ["symmetric_3x3", "general_pattern"])
def test_mtx_known_answers(tmp_path, name):
    case = K["mtx"][name]
    p = _write(tmp_path, name + ".mtx", case["text"])
    code, m, n, r, c, v = O.read_mtx(str(p))
    assert code == 0
    assert [[int(a), int(b), float(x)] for a, b, x in zip(r, c, v)] == case["coo"]
    rp, ci, va = O.coo_to_csr(m, r, c, v)
    assert rp.tolist() == case["row_ptr"]
    assert ci.tolist() == case["col_idx"]
    assert va.tolist() == case["val"]


def test_mtx_bad_banner_and_missing(tmp_path):
    for i, text in enumerate(K["mtx"]["bad_banners"]):
        p = _write(tmp_path, f"bad{i}.mtx", text)
        assert O.read_mtx(str(p))[0] == -2, text
    assert O.read_mtx(str(tmp_path / "missing.mtx"))[0] == -1


def test_coo_to_csr_is_stable_counting_sort():
    rng = np.random.default_rng(3)
    r = rng.integers(0, 50, 2000).astype(np.uint32)
    c = rng.integers(0, 70, 2000).astype(np.uint32)
    v = rng.standard_normal(2000).astype(np.float32)
    rp, ci, va = O.coo_to_csr(50, r, c, v)
    order = np.argsort(r, kind="stable")
    assert np.array_equal(ci, c[order]) and np.array_equal(va, v[order])
    assert np.array_equal(rp[1:], np.cumsum(np.bincount(r, minlength=50)))


def _check_invariants(m, rp, ci, va, rows, rw):
    words = rw // 8
    T = m["tcOffset"].size - 1
    bits = m["tcLocalBit"].reshape(T, words) if T else np.zeros((0, words), np.uint64)
    pc = np.array([sum(bin(int(x)).count("1") for x in row) for row in bits], np.int64)
    assert np.array_equal(np.diff(m["tcOffset"].astype(np.int64)), pc)
    assert int(m["tcOffset"][-1]) == len(ci)
    rwo = m["rowWindowOffset"]
    for w in range(rwo.size - 1):
        seg = m["sparseA2B"][8 * rwo[w]: 8 * rwo[w + 1]].astype(np.int64)
        real = seg[seg != 0xFFFFFFFF]
        assert np.all(np.diff(real) > 0)
        assert np.all(seg[real.size:] == 0xFFFFFFFF)
        assert seg.size - real.size < 8
        lo, hi = rp[w * rw], rp[min(w * rw + rw, rows)]
        assert np.array_equal(real, np.unique(ci[lo:hi]))
    # decode round trip reproduces the CSR triples
    r, c, v = O.metcf_decode(m, rw)
    got = sorted(zip(r.tolist(), c.tolist(), v.tolist()))
    rows_of = np.repeat(np.arange(rows), np.diff(rp.astype(np.int64)))
    exp = sorted(zip(rows_of.tolist(), ci.tolist(), va.tolist()))
    assert got == exp


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("rw", [8, 16])
def test_metcf_invariants_random(seed, rw):
    rng = np.random.default_rng(seed)
    rows = int(rng.integers(1, 300))
    cols = int(rng.integers(1, 400))
    rp, ci, va = random_csr(rng, rows, cols, float(rng.uniform(0.001, 0.1)),
                            empty_rows=0.2, heavy_rows=2)
    m = O.metcf_build(rw, rows, rp, ci, va)
    _check_invariants(m, rp, ci, va, rows, rw)
    adp = O.adpbme_build(rw, rows, rp, ci, va)
    g = adp["groupOffset"].astype(np.int64)
    T = m["tcOffset"].size - 1
    first = np.zeros(T, bool)
    first[m["rowWindowOffset"][:-1][np.diff(m["rowWindowOffset"]) > 0]] = True
    cost = np.where(first, 6, 4)
    # each group fits the budget (or is a single tile) and is maximal
    for a, b in zip(g[:-1], g[1:]):
        assert b > a
        assert cost[a:b].sum() <= 66 or b - a == 1
        if b < T:
            assert cost[a:b + 1].sum() > 66
    assert (g[-1] == T) if T else g.tolist() == [0]


def test_spmm_f64_brute_force_and_scipy():
    rng = np.random.default_rng(7)
    rows, cols = 37, 53
    rp, ci, va = random_csr(rng, rows, cols, 0.15, empty_rows=0.1)
    B = rng.uniform(-1, 1, (cols, 9)).astype(np.float32)
    dense = np.zeros((rows, cols))
    for r in range(rows):
        for e in range(rp[r], rp[r + 1]):
            dense[r, ci[e]] += va[e]
    ref = dense @ B.astype(np.float64)
    got = O.spmm_f64(rp, ci, va, B)
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)
    A = sp.csr_matrix((va.astype(np.float64), ci, rp), shape=(rows, cols))
    assert np.allclose(got, A @ B.astype(np.float64), rtol=1e-13, atol=1e-13)
    absab = O.spmm_abs_f64(rp, ci, va, B)
    assert np.allclose(absab, np.abs(dense) @ np.abs(B.astype(np.float64)))


def test_errors_metric_and_cpu_baseline():
    rng = np.random.default_rng(11)
    rows, cols = 500, 400
    rp, ci, va = random_csr(rng, rows, cols, 0.05)
    B = rng.uniform(-1, 1, (cols, 33)).astype(np.float32)
    C64 = O.spmm_f64(rp, ci, va, B)
    ab = O.spmm_abs_f64(rp, ci, va, B)
    C32, used = O.spmm_f32_omp(rp, ci, va, B)
    assert used >= 1
    mx, fr = O.errors(C32, C64, ab)
    assert mx < 1e-5 and fr < 1e-6
    # a planted error is seen by both metrics
    bad = C32.copy()
    bad[3, 4] += 1e-3
    mx2, fr2 = O.errors(bad, C64, ab)
    assert mx2 > 1e-5 and fr2 > fr
    # an entry whose |A||B| is 0 must be exactly 0
    z = np.zeros_like(C32)
    zero_ab = np.zeros_like(ab)
    assert O.errors(z, np.zeros_like(C64), zero_ab)[0] == 0.0
    z[0, 0] = 1e-30
    assert O.errors(z, np.zeros_like(C64), zero_ab)[0] == float("inf")


def test_tf32_rna_known_values():
    # TF32 keeps 10 explicit mantissa bits; ties (exactly half an ulp of
    # 2^-10 at 1.0) round away from zero, as the PTX ISA's .rna
    import math
    ulp = 2.0 ** -10
    cases = [
        (1.0, 1.0),
        (1.0 + 2.0 ** -11, 1.0 + ulp),          # tie -> away from zero
        (-(1.0 + 2.0 ** -11), -(1.0 + ulp)),
        (1.0 + 2.0 ** -12, 1.0),                 # below half: down
        (1.0 + 3 * 2.0 ** -12, 1.0 + ulp),       # above half: up
        (2.0 - 2.0 ** -12, 2.0),                 # carry into the exponent
        (0.0, 0.0),
        (1.5 * 2.0 ** 100, 1.5 * 2.0 ** 100),   # exactly representable: kept
    ]
    for x, want in cases:
        assert O.tf32_rna(float(np.float32(x))) == want, (x, O.tf32_rna(x), want)
    assert math.isinf(O.tf32_rna(float("inf"))) and math.isnan(O.tf32_rna(float("nan")))
    assert math.isinf(O.tf32_rna(float(np.finfo(np.float32).max)))  # rounds past the top
    # every result has its 13 dropped bits clear
    xs = np.random.default_rng(0).standard_normal(5000).astype(np.float32)
    ys = np.array([O.tf32_rna(float(x)) for x in xs], np.float32)
    assert np.all(ys.view(np.uint32) & 0x1FFF == 0)
    assert np.all(np.abs(ys - xs) <= 2.0 ** -11 * np.abs(xs) * (1 + 1e-6))


def test_tf32_precision_class():
    # the reference's input rounding gives errors of order 2^-11..2^-10
    # relative to |A||B|: far above the split-BF16 form's 1e-5 bound
    rng = np.random.default_rng(5)
    from helpers import random_csr
    rp, ci, va = random_csr(rng, 300, 400, 0.05)
    B = rng.uniform(-1, 1, (400, 32)).astype(np.float32)
    C64 = O.spmm_f64(rp, ci, va, B)
    ab = O.spmm_abs_f64(rp, ci, va, B)
    Ct = O.spmm_tf32_f64(rp, ci, va, B).astype(np.float32)
    mx, fr = O.errors(Ct, C64, ab)
    assert 1e-5 < mx <= 2.0 ** -10 * (1 + 1e-3), mx
