"""METCF file loader checks, on the CPU: a file written by hand in the
documented layout (header, arrays, FNV-1a 64 checksum) passes every format
check (the load then stops at the device step: no GPU here), and each kind
of corruption is rejected with ACCSPMM_ERR_FILE before any device call."""
import struct

import numpy as np

import accspmm_amd as acc
from helpers import random_csr
from oracle import oracle as O


def _fnv(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _file(rows=61, cols=50, seed=3, version=1, order=None):
    """Version 1: 19 header words.  Version 2: a 20th word (reorder_rows)
    and, when it is 1, the row order after the arrays."""
    rng = np.random.default_rng(seed)
    rp, ci, va = random_csr(rng, rows, cols, 0.08, empty_rows=0.1)
    m = O.metcf_build(8, rows, rp, ci, va)
    W, T, nnz = len(m["rowWindowOffset"]) - 1, len(m["tcOffset"]) - 1, len(va)
    key = b"rw8-test"
    hdr = [rows, cols, nnz, 8, W, T, 0, 0, 0, 0, 0, 8, 8, 3, 1, 2, 16, 8, len(key)]
    if version == 2:
        hdr.append(0 if order is None else 1)
    body = (b"ACCSPMTC" + struct.pack("<I", version) + struct.pack(f"<{len(hdr)}I", *hdr) +
            key +
            m["rowWindowOffset"].astype("<u4").tobytes() +
            m["tcOffset"].astype("<u4").tobytes() +
            m["tcLocalBit"].astype("<u8").tobytes() +
            m["sparseA2B"].astype("<u4").tobytes() +
            np.asarray(m["data"], "<f4").tobytes())
    if order is not None:
        body += np.asarray(order, "<u4").tobytes()
    return body + struct.pack("<Q", _fnv(body))


def _load(tmp_path, name, data) -> int:
    p = tmp_path / name
    p.write_bytes(data)
    h = acc.C.c_void_p()
    return acc.lib().accspmm_metcf_load(str(p).encode(), acc.C.byref(h), None)


def test_valid_file_passes_format_checks(tmp_path):
    st = _load(tmp_path, "ok.metcf", _file())
    assert st in (acc.Status.OK, acc.Status.ERR_HIP), st
    for order in (None, np.random.default_rng(4).permutation(61)):
        st = _load(tmp_path, "ok2.metcf", _file(version=2, order=order))
        assert st in (acc.Status.OK, acc.Status.ERR_HIP), st


def test_row_order_must_be_a_permutation(tmp_path):
    order = np.arange(61)
    order[5] = 7                        # row 7 twice, row 5 never
    assert _load(tmp_path, "dup.metcf", _file(version=2, order=order)) == acc.Status.ERR_FILE
    order = np.arange(61)
    order[0] = 61                       # out of range
    assert _load(tmp_path, "oob.metcf", _file(version=2, order=order)) == acc.Status.ERR_FILE


def test_corruptions_rejected(tmp_path):
    good = _file()
    body = good[:-8]
    hdr_at = 12  # magic + version

    def resum(b):
        return b + struct.pack("<Q", _fnv(b))

    def set_word(b, i, v):
        return b[:hdr_at + 4 * i] + struct.pack("<I", v) + b[hdr_at + 4 * i + 4:]

    cases = {
        "checksum": good[:-1] + bytes([good[-1] ^ 0xFF]),
        "payload": good[:200] + bytes([good[200] ^ 1]) + good[201:],
        "truncated": good[:-20],
        "trailing": good + b"x",
        "magic": b"ACCSPMTX" + good[8:],
        "version": good[:8] + struct.pack("<I", 3) + good[12:],
        "v2_header_on_v1_body": good[:8] + struct.pack("<I", 2) + good[12:],
        # header inconsistent with the arrays (re-summed, so only the
        # structural checks can catch it)
        "window_count": resum(set_word(body, 4, 3)),
        "row_window": resum(set_word(body, 3, 12)),
        "nnz": resum(set_word(body, 2, 5)),
    }
    for name, data in cases.items():
        assert _load(tmp_path, name + ".metcf", data) == acc.Status.ERR_FILE, name


def test_invariant_violations_rejected(tmp_path):
    # well-formed and correctly summed files whose arrays break the format
    rows, cols = 61, 50
    good = _file(rows, cols)
    body = bytearray(good[:-8])
    key_len = struct.unpack_from("<I", body, 12 + 4 * 18)[0]
    W = struct.unpack_from("<I", body, 12 + 4 * 4)[0]
    T = struct.unpack_from("<I", body, 12 + 4 * 5)[0]
    rwo_at = 12 + 4 * 19 + key_len
    tco_at = rwo_at + 4 * (W + 1)
    bits_at = tco_at + 4 * (T + 1)
    a2b_at = bits_at + 8 * T

    def case(mutate):
        b = bytearray(body)
        mutate(b)
        return bytes(b) + struct.pack("<Q", _fnv(bytes(b)))

    bad = {
        # tcOffset no longer monotone
        "tco": case(lambda b: struct.pack_into("<I", b, tco_at + 4, 0xFFFF)),
        # a tile's bitmap popcount differs from its tcOffset span
        "bits": case(lambda b: struct.pack_into("<Q", b, bits_at, 0)),
        # a condensed column outside the matrix
        "a2b": case(lambda b: struct.pack_into("<I", b, a2b_at, cols + 7)),
        # rowWindowOffset not ending at T
        "rwo": case(lambda b: struct.pack_into("<I", b, rwo_at + 4 * W, T + 1)),
    }
    for name, data in bad.items():
        assert _load(tmp_path, name + ".metcf", data) == acc.Status.ERR_FILE, name
