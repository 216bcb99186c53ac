"""The C-ABI library loads and exports every symbol declared in include/*.h;
argument checks that fire before any device call (no compute calls: this
runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np

import accspmm_amd as acc

REPO = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(accspmm_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(str(acc.LIB_PATH))
    decl = _declared()
    assert len(decl) > 30
    missing = sorted(n for n in decl if not hasattr(L, n))
    assert not missing, missing


def test_python_bindings_cover_the_abi():
    assert _declared() <= set(acc._SIGS), sorted(_declared() - set(acc._SIGS))


def test_status_strings_and_version():
    L = acc.lib()
    assert L.accspmm_status_string(0) == b"ok"
    assert b"increasing" in L.accspmm_status_string(2)
    assert b"gfx950" in L.accspmm_version()
    o = acc.BuildOpts()
    L.accspmm_build_opts_default(ctypes.byref(o))
    assert (o.row_window, o.col_window, o.t_load, o.t_comp, o.t_write, o.target_num,
            o.imbalance_threshold) == (8, 8, 3, 1, 2, 16, 8)


def test_null_arguments_rejected_without_gpu():
    L = acc.lib()
    assert L.accspmm_metcf_build(None, None, None, None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_spmm(None, None, 0, None, 0, 0, 0, None, None) == \
        acc.Status.ERR_INVALID_ARG
    assert L.accspmm_metcf_destroy(None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_batch_run(None, None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_server_poll(None, 0, None) == acc.Status.ERR_INVALID_ARG


def test_oversized_nnz_rejected_before_device_work():
    # value offsets are 32-bit: more than 2^32 - 1024 nonzeros is refused
    # before any device call (the pointers are never read)
    L = acc.lib()
    C = ctypes
    dummy = (C.c_uint32 * 4)()
    view = acc.CsrView(4, 4, 0xFFFFFFFF - 1023 + 1, C.addressof(dummy), C.addressof(dummy),
                       C.addressof(dummy))
    h = C.c_void_p()
    assert L.accspmm_metcf_build(C.byref(view), None, C.byref(h), None) == \
        acc.Status.ERR_UNSUPPORTED
    assert not h.value


def test_metcf_load_checks_length_before_allocating(tmp_path):
    # a METCF file whose header claims ~4 GB of tiles but holds a few bytes
    # is refused from its length, before any array is sized from the header
    # or any device call (no GPU here)
    import struct
    L = acc.lib()
    hdr = [10, 10, 0, 8, 2, 0x1FFFFFFF, 0, 0, 0, 0, 0, 8, 8, 1, 1, 1, 1, 1, 0, 0, 0]
    body = b"ACCSPMTC" + struct.pack("<I", 3) + struct.pack("<21I", *hdr) + bytes(64)
    for name, data in (("huge.metcf", body), ("short.metcf", body[:20]), ("empty.metcf", b"")):
        path = tmp_path / name
        path.write_bytes(data)
        h = ctypes.c_void_p()
        assert L.accspmm_metcf_load(str(path).encode(), ctypes.byref(h), None) == \
            acc.Status.ERR_FILE, name
        assert not h.value


def test_operand_checks_refuse_host_tensors():
    # no GPU needed: host tensors and wrong shapes are refused in Python
    import pytest
    import torch
    with pytest.raises(TypeError):
        acc._operands(torch.zeros((4, 2)), torch.zeros((3, 2)), 3, 4, "spmm")
    with pytest.raises(TypeError):
        acc._dense(np.zeros((4, 2), np.float32), 4, "spmm", "B")
