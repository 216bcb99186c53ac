"""Host utilities of the product library (MTX ingest, CSR, generators)
against the oracle and closed forms.  No GPU."""
import numpy as np
import pytest

import accspmm_amd as acc
from helpers import kat
from oracle import oracle as O

K = kat()


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


@pytest.mark.parametrize("name", ["symmetric_3x3", "general_pattern"])
def test_read_mtx_known_answers(tmp_path, name):
    case = K["mtx"][name]
    csr = acc.read_mtx(_write(tmp_path, name + ".mtx", case["text"]))
    assert csr.row_ptr.tolist() == case["row_ptr"]
    assert csr.col_idx.tolist() == case["col_idx"]
    assert csr.data.tolist() == case["val"]


def test_read_mtx_bad_banners(tmp_path):
    for i, text in enumerate(K["mtx"]["bad_banners"]):
        with pytest.raises(acc.AccSpMMError) as e:
            acc.read_mtx(_write(tmp_path, f"b{i}.mtx", text))
        assert e.value.mtx_code == -2
    with pytest.raises(acc.AccSpMMError) as e:
        acc.read_mtx(tmp_path / "missing.mtx")
    assert e.value.mtx_code == -1


@pytest.mark.parametrize("field", ["real", "pattern", "integer"])
@pytest.mark.parametrize("sym", ["general", "symmetric"])
def test_read_mtx_matches_oracle(tmp_path, field, sym):
    rng = np.random.default_rng(hash((field, sym)) % 1000)
    n = 60
    ents = set()
    while len(ents) < 300:
        r, c = rng.integers(1, n + 1, 2)
        if sym == "symmetric" and c > r:
            r, c = c, r
        ents.add((int(r), int(c)))
    lines = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% comment",
             f"{n} {n} {len(ents)}"]
    for r, c in sorted(ents, key=lambda x: (x[1], x[0])):
        if field == "real":
            lines.append(f"{r} {c} {rng.uniform(-5, 5):.9g}")
        elif field == "integer":
            lines.append(f"{r} {c} {int(rng.integers(-9, 10))}")
        else:
            lines.append(f"{r} {c}")
    p = _write(tmp_path, "m.mtx", "\n".join(lines) + "\n")
    code, m, nn, rr, cc, vv = O.read_mtx(str(p))
    assert code == 0
    rp, ci, va = O.coo_to_csr(m, rr, cc, vv)
    got = acc.read_mtx(p)
    assert np.array_equal(got.row_ptr, rp)
    assert np.array_equal(got.col_idx, ci)
    assert np.array_equal(got.data, va)


def test_canonicalize_sorts_and_sums(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real general\n3 4 5\n"
            "1 3 1.0\n1 1 2.0\n1 3 0.5\n3 2 4.0\n2 4 1.0\n")
    p = _write(tmp_path, "u.mtx", text)
    raw = acc.read_mtx(p)
    assert not raw.is_canonical()
    can = acc.read_mtx(p, canonicalize=True)
    assert can.is_canonical()
    assert can.row_ptr.tolist() == [0, 2, 3, 4]
    assert can.col_idx.tolist() == [0, 2, 3, 1]
    assert can.data.tolist() == [2.0, 1.5, 1.0, 4.0]


def test_chung_lu_properties_and_determinism():
    a = acc.gen_chung_lu(3000, 20.0, 2.1, 5, 0)
    b = acc.gen_chung_lu(3000, 20.0, 2.1, 5, 0)
    assert np.array_equal(a.col_idx, b.col_idx) and np.array_equal(a.data, b.data)
    assert a.is_canonical()
    import scipy.sparse as sp
    A = sp.csr_matrix((a.data, a.col_idx, a.row_ptr), shape=(a.rows, a.cols))
    assert (A != A.T).nnz == 0            # symmetric pattern and values
    assert np.all(A.diagonal() > 0)       # a self loop on every node
    assert np.all((a.data > 0) & (a.data <= 1))
    deg = np.diff(a.row_ptr.astype(np.int64))
    assert 0.85 * 21 < deg.mean() < 1.05 * 21
    assert deg.max() > 8 * deg.mean()     # heavy tail
    c = acc.gen_chung_lu(3000, 20.0, 2.1, 6, 1)
    rows = np.repeat(np.arange(c.rows), np.diff(c.row_ptr.astype(np.int64)))
    assert np.allclose(c.data, (1.0 / np.diff(c.row_ptr.astype(np.float64)))[rows])


def test_stencil_closed_form():
    for g in (1, 2, 5):
        s = acc.gen_stencil27(g, 3)
        assert s.rows == g ** 3 and s.nnz == (3 * g - 2) ** 3
        assert s.is_canonical()
    s = acc.gen_stencil27(4, 3)
    # interior node (1,1,1) has 27 neighbours, corner (0,0,0) has 8
    deg = np.diff(s.row_ptr.astype(np.int64))
    assert deg[1 + 4 + 16] == 27 and deg[0] == 8
    assert np.all((s.data >= -1) & (s.data < 1))


def test_dense_generator_range():
    d = acc.gen_dense(1, 100, 33)
    assert d.shape == (100, 33) and np.all((d >= -1) & (d < 1))
    assert np.array_equal(d, acc.gen_dense(1, 100, 33))


def test_mma_driver_exit_codes(tmp_path):
    import subprocess
    mma = str(acc.LIB_PATH.parent.parent / "bin" / "mma")
    r = subprocess.run([mma], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 4 and "Usage" in r.stdout
    r = subprocess.run([mma, str(tmp_path / "none.mtx"), "32"], capture_output=True,
                       text=True, cwd=tmp_path)
    assert r.returncode == 3
    r = subprocess.run([mma, str(tmp_path / "none.mtx"), "32", "--layout", "fast"],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 4 and "unknown layout" in r.stdout
    bad = _write(tmp_path, "bad.mtx", K["mtx"]["bad_banners"][0])
    r = subprocess.run([mma, str(bad), "32"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 3 and "banner" in r.stdout


def test_canonical_restatement_brute_force():
    # the numpy restatement the canonicalisation tests check against, pinned
    # by a per-row dictionary walk on small scrambled matrices
    from helpers import canonical_np, scrambled_csr
    rng = np.random.default_rng(0)
    for _ in range(300):
        rows, cols = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        rp, ci, va = scrambled_csr(rng, rows, cols, 6, 0.3)
        orp, oci, ova = canonical_np(rows, cols, rp, ci, va)
        exp_rp, exp_c, exp_v = [0], [], []
        for r in range(rows):
            runs = {}
            for c, v in zip(ci[rp[r]:rp[r + 1]].tolist(), va[rp[r]:rp[r + 1]].tolist()):
                runs.setdefault(c, []).append(np.float32(v))
            for c in sorted(runs):
                s = runs[c][0]
                for v in runs[c][1:]:
                    s = np.float32(s + v)
                exp_c.append(c)
                exp_v.append(s)
            exp_rp.append(len(exp_c))
        assert orp.tolist() == exp_rp and oci.tolist() == exp_c
        assert np.array_equal(ova.view(np.uint32), np.array(exp_v, np.float32).view(np.uint32))


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("n", [50, 60000])
def test_write_mtx_round_trip(tmp_path, symmetric, n):
    # written files read back to the same CSR bit for bit (9-digit values),
    # and match the oracle's fscanf reader
    csr = acc.gen_chung_lu(n, 12.0, 2.1, 3, 0)
    p = tmp_path / "w.mtx"
    csr.write_mtx(p, symmetric=symmetric)
    head = p.read_text()[:80].splitlines()
    assert head[0] == "%%MatrixMarket matrix coordinate real " + \
        ("symmetric" if symmetric else "general")
    back = acc.read_mtx(p, canonicalize=True)
    assert np.array_equal(back.row_ptr, csr.row_ptr)
    assert np.array_equal(back.col_idx, csr.col_idx)
    assert np.array_equal(back.data.view(np.uint32), csr.data.view(np.uint32))
    if n <= 50:
        code, m, nn, rr, cc, vv = O.read_mtx(str(p))
        rp, ci, va = O.coo_to_csr(m, rr, cc, vv)
        assert code == 0 and np.array_equal(rp, csr.row_ptr) and np.array_equal(ci, csr.col_idx)
        assert np.array_equal(va.view(np.uint32), csr.data.view(np.uint32))


def test_write_mtx_rejects_non_square_symmetric(tmp_path):
    rng = np.random.default_rng(1)
    rp, ci, va = random_csr_small(rng)
    csr = acc.CSR(4, 6, rp, ci, va)
    with pytest.raises(acc.AccSpMMError) as e:
        csr.write_mtx(tmp_path / "x.mtx", symmetric=True)
    assert e.value.status == acc.Status.ERR_SHAPE
    csr.write_mtx(tmp_path / "x.mtx", symmetric=False)
    back = acc.read_mtx(tmp_path / "x.mtx")
    assert np.array_equal(back.col_idx, ci) and np.array_equal(back.data, va)


def random_csr_small(rng):
    rp = np.array([0, 2, 2, 5, 6], np.uint32)
    ci = np.array([1, 4, 0, 3, 5, 2], np.uint32)
    va = rng.uniform(-1, 1, 6).astype(np.float32)
    return rp, ci, va


@pytest.mark.parametrize("sym", ["Hermitian", "skew-symmetric"])
def test_read_mtx_hermitian_and_skew_as_the_reference(tmp_path, sym):
    # the reference's reading (include/mmio.hpp:123,133-158,201): complex
    # values keep their real part; Hermitian entries are mirrored like
    # symmetric ones (without conjugation, the imaginary part being gone);
    # skew-symmetric entries are neither mirrored nor negated
    text = (f"%%MatrixMarket matrix coordinate complex {sym}\n3 3 3\n"
            "1 1 2.5 9.0\n3 1 -1.5 4.0\n3 2 0.25 -7.0\n")
    csr = acc.read_mtx(_write(tmp_path, "h.mtx", text))
    if sym == "Hermitian":
        dense = np.array([[2.5, 0, -1.5], [0, 0, 0.25], [-1.5, 0.25, 0]], np.float32)
    else:
        dense = np.array([[2.5, 0, 0], [0, 0, 0], [-1.5, 0.25, 0]], np.float32)
    got = np.zeros((3, 3), np.float32)
    for r in range(3):
        for e in range(csr.row_ptr[r], csr.row_ptr[r + 1]):
            got[r, csr.col_idx[e]] += csr.data[e]
    assert np.array_equal(got, dense)
    code, m, nn, rr, cc, vv = O.read_mtx(str(tmp_path / "h.mtx"))
    assert code == 0
    rp, ci, va = O.coo_to_csr(m, rr, cc, vv)
    assert np.array_equal(csr.row_ptr, rp) and np.array_equal(csr.col_idx, ci)
    assert np.array_equal(csr.data, va)


@pytest.mark.parametrize("sym", ["symmetric", "Hermitian"])
def test_read_mtx_rejects_non_square_mirrored(tmp_path, sym):
    # a mirrored entry (c, r) of a non-square matrix falls outside it (a
    # 4 x 2 symmetric file with entry (3, 1) would need column 3): an error,
    # not a CSR with a column index past cols
    field = "complex" if sym == "Hermitian" else "real"
    vals = " 1.5 0.5" if sym == "Hermitian" else " 1.5"
    text = f"%%MatrixMarket matrix coordinate {field} {sym}\n4 2 1\n3 1{vals}\n"
    with pytest.raises(acc.AccSpMMError) as e:
        acc.read_mtx(_write(tmp_path, "ns.mtx", text))
    assert e.value.mtx_code == 3


@pytest.mark.parametrize("size_line", ["4294967296 2 1", "2 -1 1", "2 2 99999999999999999999",
                                       "2 2 x1"])
def test_read_mtx_rejects_bad_size_line(tmp_path, size_line):
    # sizes are unsigned 32-bit integers: an overflowing, signed or other
    # token is an error (not a wrapped or saturated size)
    text = f"%%MatrixMarket matrix coordinate real general\n{size_line}\n1 1 1.0\n"
    with pytest.raises(acc.AccSpMMError) as e:
        acc.read_mtx(_write(tmp_path, "sz.mtx", text))
    assert e.value.mtx_code == 3


def test_read_mtx_size_line_split_over_lines(tmp_path):
    # as the reference's reader, the three sizes may continue on later lines
    text = "%%MatrixMarket matrix coordinate real general\n3 2\n2\n1 1 1.0\n3 2 -2.0\n"
    csr = acc.read_mtx(_write(tmp_path, "split.mtx", text))
    assert (csr.rows, csr.cols, csr.nnz) == (3, 2, 2)
    assert csr.row_ptr.tolist() == [0, 1, 1, 2] and csr.col_idx.tolist() == [0, 1]
