"""The reference's interfaces on the GPU: a C++ program written against the
format.hpp / mmio.hpp API (compiled unchanged against the drop-in headers)
and the `./mma <matrix> <feature_dim>` driver."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import accspmm_amd as acc
from helpers import kat
from oracle import oracle as O

pytestmark = pytest.mark.gpu
REPO = Path(__file__).resolve().parent.parent
K = kat()


def _mtx_of_a10(path: Path):
    a = K["a10"]
    lines = ["%%MatrixMarket matrix coordinate real general",
             f"{a['rows']} {a['cols']} {len(a['col_idx'])}"]
    for r in range(a["rows"]):
        for e in range(a["row_ptr"][r], a["row_ptr"][r + 1]):
            lines.append(f"{r + 1} {a['col_idx'][e] + 1} {a['val'][e]}")
    path.write_text("\n".join(lines) + "\n")


def test_cpp_format_api_known_answer(tmp_path):
    mtx = tmp_path / "a10.mtx"
    _mtx_of_a10(mtx)
    r = subprocess.run([str(REPO / "tests" / "cpp" / "format_dropin"), str(mtx)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout[r.stdout.index("{"):])
    exp = K["metcf_rw8"]
    for key in ("rowWindowOffset", "tcOffset", "tcLocalBit", "sparseA2B",
                "groupOffset", "rowIndices"):
        assert out[key] == exp[key], key
    got = np.array(out["data_bits"], np.uint32).view(np.float32)
    assert got.tolist() == [float(x) for x in exp["data"]]
    assert out["balanced"] == 0


def test_cpp_format_api_random_graph(tmp_path):
    csr = acc.gen_chung_lu(3000, 12.0, 2.1, 77, 0)
    mtx = tmp_path / "g.mtx"
    rows = np.repeat(np.arange(csr.rows), np.diff(csr.row_ptr.astype(np.int64)))
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{csr.rows} {csr.cols} {csr.nnz}\n")
        for r, c, v in zip(rows, csr.col_idx, csr.data):
            f.write(f"{r + 1} {c + 1} {v:.9g}\n")
    r = subprocess.run([str(REPO / "tests" / "cpp" / "format_dropin"), str(mtx)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout[r.stdout.index("{"):])
    back = acc.read_mtx(mtx)
    exp = O.adpbme_build(8, back.rows, back.row_ptr, back.col_idx, back.data)
    for key in ("rowWindowOffset", "tcOffset", "tcLocalBit", "sparseA2B",
                "groupOffset", "rowIndices"):
        assert out[key] == exp[key].tolist(), key
    assert np.array_equal(np.array(out["data_bits"], np.uint32).view(np.float32),
                          exp["data"])


@pytest.mark.parametrize("form", ["fp32", "split_bf16", "bf16"])
def test_mma_driver_appends_result_row(tmp_path, form):
    mtx = tmp_path / "g.mtx"
    csr = acc.gen_chung_lu(2000, 10.0, 2.2, 5, 0)
    rows = np.repeat(np.arange(csr.rows), np.diff(csr.row_ptr.astype(np.int64)))
    sym = [(r, c, v) for r, c, v in zip(rows, csr.col_idx, csr.data) if r >= c]
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{csr.rows} {csr.cols} {len(sym)}\n")
        for r, c, v in sym:
            f.write(f"{r + 1} {c + 1} {v:.9g}\n")
    mma = str(REPO / "accspmm_amd" / "bin" / "mma")
    for K_ in (32, 48):
        r = subprocess.run([mma, str(mtx), str(K_), "--form", form, "--random-b", "3",
                            "--warmup", "2", "--exe", "3", "--check"],
                           capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "m_tmp: 2000" in r.stdout
        assert "pipelining" in r.stdout
        # the driver's own FP64 check of C (random B) passes its form's bound
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("check:")]
        assert line and line[0].endswith("ok"), r.stdout
    rows_csv = (tmp_path / "result.csv").read_text().strip().splitlines()
    assert len(rows_csv) == 2
    name, k, ms, gf = rows_csv[0].split(",")
    assert name == "g.mtx" and k == "32" and float(ms) > 0 and float(gf) > 0


@pytest.mark.parametrize("form", ["fp32", "split_bf16"])
def test_mma_driver_auto_layout(tmp_path, form):
    # --layout auto launches on the library's automatic layout (here 16-row
    # windows and reordered rows: rows average >= 128 nonzeros); the driver's
    # FP64 check of C in input row order passes
    rng = np.random.default_rng(33)
    rows, cols = 1200, 1500
    from helpers import random_csr
    rp, ci, va = random_csr(rng, rows, cols, 0.12, empty_rows=0.05, heavy_rows=2)
    csr = acc.CSR(rows, cols, rp, ci, va)
    assert csr.nnz >= 128 * rows
    mtx = tmp_path / "long.mtx"
    csr.write_mtx(mtx)
    mma = str(REPO / "accspmm_amd" / "bin" / "mma")
    r = subprocess.run([mma, str(mtx), "64", "--form", form, "--layout", "auto",
                        "--warmup", "2", "--exe", "3", "--check"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("check:")]
    assert line and line[0].endswith("ok"), r.stdout
    assert (tmp_path / "result.csv").read_text().startswith("long.mtx,64,")


def test_mma_driver_path_selection(tmp_path):
    # the driver picks the balanced schedule exactly when the oracle's
    # balanceStrategy restatement says IBD > 8 (include/utils.hpp:75-85) and
    # prints the reference's line for the chosen path
    from helpers import random_csr
    rng = np.random.default_rng(90)
    even = acc.gen_stencil27(12, 3)                         # every window alike
    rp, ci, va = random_csr(rng, 900, 4000, 0.004, heavy_rows=6)
    skew = acc.CSR(900, 4000, rp, ci, va)                   # a few dense rows
    mma = str(REPO / "accspmm_amd" / "bin" / "mma")
    seen = set()
    for name, csr in (("even", even), ("skew", skew)):
        m = O.metcf_build(8, csr.rows, csr.row_ptr, csr.col_idx, csr.data)
        _, balanced = O.ibd(m["rowWindowOffset"])
        seen.add(balanced)
        mtx = tmp_path / f"{name}.mtx"
        csr.write_mtx(mtx)
        r = subprocess.run([mma, str(mtx), "32", "--warmup", "1", "--exe", "2"],
                           capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        want = "doing AdpBalance pipelining ..." if balanced else "doing pipelining mma ..."
        assert want in r.stdout, (name, balanced, r.stdout)
    assert seen == {True, False}, "the two matrices must take both paths"
