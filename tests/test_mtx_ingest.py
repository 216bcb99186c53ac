"""Multi-threaded Matrix Market ingest (files with at least 1 MiB of entries)
against the oracle's fscanf reader, against Python's correctly rounded
float(), and against the sequential tokenizer (ACCSPMM_HOST_THREADS=1) on
files the line parser hands back to it.  No GPU."""
import zlib

import numpy as np
import pytest

import accspmm_amd as acc
from helpers import canonical_np
from oracle import oracle as O

PAR = "4"   # threads for the parallel path
MIN_BYTES = 1 << 20


def _read(path, monkeypatch, threads, canonicalize=False, parsed_with=None):
    monkeypatch.setenv("ACCSPMM_HOST_THREADS", threads)
    csr = acc.read_mtx(path, canonicalize=canonicalize)
    used = acc.lib().accspmm_read_mtx_threads_used()
    assert used == (int(threads) if parsed_with is None else parsed_with)
    return csr


def _read_err(path, monkeypatch, threads):
    monkeypatch.setenv("ACCSPMM_HOST_THREADS", threads)
    with pytest.raises(acc.AccSpMMError) as e:
        acc.read_mtx(path)
    return e.value.status, e.value.mtx_code


def _same(a: acc.CSR, b_rp, b_ci, b_va):
    assert np.array_equal(a.row_ptr, b_rp)
    assert np.array_equal(a.col_idx, b_ci)
    assert np.array_equal(a.data.view(np.uint32), np.asarray(b_va, np.float32).view(np.uint32))


def _value_token(rng, k):
    # formats that take the exact fast path and formats that need strtod
    x = float(rng.uniform(-10, 10)) * 10.0 ** int(rng.integers(-30, 30))
    fmt = k % 9
    if fmt == 0:
        return f"{x:.9g}"
    if fmt == 1:
        return f"{x:.17g}"            # 17 digits: above 2^53, strtod
    if fmt == 2:
        return f"{x:e}"
    if fmt == 3:
        return f"{rng.uniform(-1, 1):.6f}"
    if fmt == 4:
        return str(int(rng.integers(-10**6, 10**6)))
    if fmt == 5:
        return f"{rng.uniform(0, 1):.25f}"  # long fraction
    if fmt == 6:
        return ["-0", "0.", ".5", "+3.25", "1E5", "-2e-3", "0e0", "7.0E+01"][k % 8]
    if fmt == 7:
        return f"{x:.3E}"
    return f"{rng.uniform(-1e-38, 1e-38):.8e}"  # float32 subnormal range


def _entries_text(rng, n_rows, n_cols, count, field, sym, messy):
    lines = []
    for k in range(count):
        r = int(rng.integers(1, n_rows + 1))
        c = int(rng.integers(1, n_cols + 1))
        if sym and c > r:
            r, c = c, r
        sep = ("\t" if (messy and k % 7 == 0) else " ")
        toks = [str(r), str(c)]
        if field == "real":
            toks.append(_value_token(rng, k))
        elif field == "complex":
            toks += [_value_token(rng, k), _value_token(rng, k + 1)]
        elif field == "integer":
            toks.append(str(int(rng.integers(-10**9, 10**9))))
        line = sep.join(toks)
        if messy:
            if k % 11 == 0:
                line = "  " + line + " \t"
            if k % 13 == 0:
                line += "\r"
            if k % 17 == 0:
                lines.append("")
        lines.append(line)
    return lines


@pytest.mark.parametrize("field", ["real", "integer", "pattern", "complex"])
@pytest.mark.parametrize("sym", ["general", "symmetric"])
@pytest.mark.parametrize("messy", [False, True])
def test_parallel_matches_oracle(tmp_path, monkeypatch, field, sym, messy):
    rng = np.random.default_rng(zlib.crc32(f"{field}/{sym}/{messy}".encode()))
    n = 5000
    count = 140000 if field == "pattern" else 90000
    body = _entries_text(rng, n, n, count, field, sym == "symmetric", messy)
    text = "\n".join([f"%%MatrixMarket matrix coordinate {field} {sym}", "% c",
                      f"{n} {n} {count}"] + body) + "\n"
    assert len(text) > MIN_BYTES
    p = tmp_path / "m.mtx"
    p.write_text(text)
    code, m, nn, rr, cc, vv = O.read_mtx(str(p))
    assert code == 0
    rp, ci, va = O.coo_to_csr(m, rr, cc, vv)
    _same(_read(p, monkeypatch, PAR), rp, ci, va)
    _same(_read(p, monkeypatch, "1"), rp, ci, va)


def test_parallel_values_equal_python_float(tmp_path, monkeypatch):
    # Python's float() rounds correctly, as strtod does: every parsed value
    # must be float32(float(token)) bit for bit
    rng = np.random.default_rng(5)
    count = 120000
    toks = [_value_token(rng, k) for k in range(count)]
    lines = [f"{k % 1000 + 1} {k // 1000 + 1} {t}" for k, t in enumerate(toks)]
    text = "\n".join(["%%MatrixMarket matrix coordinate real general",
                      f"1000 {count // 1000 + 1} {count}"] + lines) + "\n"
    p = tmp_path / "v.mtx"
    p.write_text(text)
    got = _read(p, monkeypatch, PAR)
    want = np.array([float(t) for t in toks], np.float64).astype(np.float32)
    rows = np.array([k % 1000 for k in range(count)])
    order = np.argsort(rows, kind="stable")
    assert np.array_equal(got.data.view(np.uint32), want[order].view(np.uint32))


def test_parallel_canonicalize_sorts_and_sums(tmp_path, monkeypatch):
    rng = np.random.default_rng(11)
    n, count = 3000, 110000
    r = rng.integers(1, n + 1, count)
    c = rng.integers(1, n + 1, count)
    dup = rng.integers(0, count, 20000)   # repeated coordinates, new values
    r = np.concatenate([r, r[dup]])
    c = np.concatenate([c, c[dup]])
    v = rng.uniform(-1, 1, r.size).astype(np.float32)
    lines = [f"{a} {b} {x:.9g}" for a, b, x in zip(r, c, v)]
    text = "\n".join(["%%MatrixMarket matrix coordinate real general",
                      f"{n} {n} {r.size}"] + lines) + "\n"
    p = tmp_path / "d.mtx"
    p.write_text(text)
    par = _read(p, monkeypatch, PAR, canonicalize=True)
    seq = _read(p, monkeypatch, "1", canonicalize=True)
    assert par.is_canonical()
    _same(par, seq.row_ptr, seq.col_idx, seq.data)
    # independent restatement (numpy): stable (row, col) order, duplicates
    # summed left to right in float32
    vals = np.array([np.float32(float(f"{x:.9g}")) for x in v])
    rr = np.sort(r - 1, kind="stable")
    order = np.argsort(r, kind="stable")
    raw_rp = np.r_[0, np.cumsum(np.bincount(rr, minlength=n))]
    crp, cci, cva = canonical_np(n, n, raw_rp, c[order] - 1, vals[order])
    _same(par, crp, cci, cva)


def _big_general(n_entries, extra_lines=(), split_entry=None, nz=None):
    rng = np.random.default_rng(3)
    lines = []
    for k in range(n_entries):
        a, b = int(rng.integers(1, 4001)), int(rng.integers(1, 4001))
        lines.append(f"{a} {b} {rng.uniform(-1, 1):.9g}")
    if split_entry is not None:
        a, b, x = lines[split_entry].split()
        lines[split_entry] = f"{a}\n{b} {x}"
    lines[n_entries // 2:n_entries // 2] = list(extra_lines)
    hdr = ["%%MatrixMarket matrix coordinate real general",
           f"4000 4000 {n_entries if nz is None else nz}"]
    return "\n".join(hdr + lines) + "\n"


def test_irregular_files_fall_back(tmp_path, monkeypatch):
    # entry split over two lines and trailing text after the last entry: the
    # sequential tokenizer reads them (as the oracle's fscanf does)
    n = 60000
    cases = {
        "split": _big_general(n, split_entry=1234),
        "trailing": _big_general(n) + "junk after the entries\n",
    }
    for name, text in cases.items():
        p = tmp_path / f"{name}.mtx"
        p.write_text(text)
        code, m, nn, rr, cc, vv = O.read_mtx(str(p))
        assert code == 0
        rp, ci, va = O.coo_to_csr(m, rr, cc, vv)
        _same(_read(p, monkeypatch, PAR, parsed_with=1), rp, ci, va)
    # fewer entries than the size line names, a comment inside the entries,
    # an index beyond the size line: the same error on both paths
    bad = {
        "short": _big_general(n, nz=n + 5),
        "comment": _big_general(n, extra_lines=["% inside"], nz=n + 1),
        "range": _big_general(n, extra_lines=["4001 1 1.0"], nz=n + 1),
        "token": _big_general(n, extra_lines=["1 2 1.5x"], nz=n + 1),
    }
    for name, text in bad.items():
        p = tmp_path / f"{name}.mtx"
        p.write_text(text)
        e_par = _read_err(p, monkeypatch, PAR)
        e_seq = _read_err(p, monkeypatch, "1")
        assert e_par == e_seq == (acc.Status.ERR_MTX, 3), name


def test_threads_and_sizes_agree(tmp_path, monkeypatch):
    # more threads than lines in a chunk, and a file just above the threshold
    text = _big_general(50000)
    assert MIN_BYTES < len(text) < 1.3 * MIN_BYTES
    p = tmp_path / "t.mtx"
    p.write_text(text)
    ref = _read(p, monkeypatch, "1")
    for th in ("2", "3", "16", "64"):
        got = _read(p, monkeypatch, th)
        _same(got, ref.row_ptr, ref.col_idx, ref.data)


def _outcome(path, monkeypatch, threads):
    monkeypatch.setenv("ACCSPMM_HOST_THREADS", threads)
    try:
        c = acc.read_mtx(path)
        return ("ok", c.rows, c.cols, c.row_ptr.tobytes(), c.col_idx.tobytes(),
                c.data.tobytes())
    except acc.AccSpMMError as e:
        return ("err", e.status, e.mtx_code)


def test_random_mutations_match_sequential(tmp_path, monkeypatch):
    # random single-character edits anywhere in the entry section: whatever
    # the edit does (a new token, a split or merged line, a comment, a bad
    # number), the parallel reader returns what the sequential tokenizer does
    rng = np.random.default_rng(2024)
    base = _big_general(52000)
    assert len(base) > MIN_BYTES
    head = base.index("\n", base.index("\n") + 1) + 1
    alphabet = [" ", "\n", "\t", "%", "-", "+", "e", "E", ".", "0", "7", "x", "\r", "\n\n"]
    agree = fallback = 0
    for case in range(24):
        text = list(base)
        for _ in range(int(rng.integers(1, 4))):
            pos = int(rng.integers(head, len(text)))
            op = int(rng.integers(3))
            ch = alphabet[int(rng.integers(len(alphabet)))]
            if op == 0:
                text.insert(pos, ch)
            elif op == 1:
                text[pos] = ch
            else:
                del text[pos]
        p = tmp_path / f"m{case}.mtx"
        p.write_text("".join(text))
        seq = _outcome(p, monkeypatch, "1")
        par = _outcome(p, monkeypatch, PAR)
        assert par == seq, (case, seq[:3], par[:3])
        used = acc.lib().accspmm_read_mtx_threads_used()
        fallback += used == 1
        agree += 1
    assert agree == 24 and fallback < 24   # some edits keep the parallel path
