"""accspmm_gather_probe, the gather-ceiling measurement: every lane's folded
word equals the XOR of the words it is meant to read (rows of the index
list in the kernels' quarter-wave pattern; indices past the region read
zeros), checked against a host restatement of that pattern."""
import ctypes

import numpy as np
import pytest

import accspmm_amd as acc

pytestmark = pytest.mark.gpu


def _expected(region, row_words, idx, waves, rows_per_wave):
    out = np.zeros(waves * 64, np.uint32)
    halves = row_words // 64
    for w in range(waves):
        rows = idx[w * rows_per_wave:(w + 1) * rows_per_wave]
        for lane in range(64):
            q, n = lane >> 4, lane & 15
            acc_ = np.uint32(0)
            for base in range(0, rows_per_wave, 64):
                for g in range(16):
                    r = int(rows[base + 4 * g + q])
                    for h in range(halves):
                        if r < region.shape[0]:
                            for x in region[r, 64 * h + 4 * n: 64 * h + 4 * n + 4]:
                                acc_ ^= x
            out[w * 64 + lane] = acc_
    return out


@pytest.mark.parametrize("row_bytes", [256, 512])
def test_probe_reads_the_pattern(row_bytes):
    import torch
    rng = np.random.default_rng(row_bytes)
    rows, waves, rpw = 300, 4, 128
    words = row_bytes // 4
    region = rng.integers(0, 2**32, size=(rows, words), dtype=np.uint32)
    idx = rng.integers(0, rows + 40, size=waves * rpw).astype(np.uint32)  # some past the end
    dr = torch.from_numpy(region.view(np.int32)).cuda()
    di = torch.from_numpy(idx.view(np.int32)).cuda()
    out = torch.zeros(waves * 64, dtype=torch.int32, device="cuda")
    st = acc.lib().accspmm_gather_probe(dr.data_ptr(), rows, row_bytes, di.data_ptr(), waves, rpw,
                                        out.data_ptr(), None)
    assert st == acc.Status.OK
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, _expected(region, words, idx, waves, rpw))


def test_probe_rejects_bad_shapes():
    import torch
    r = torch.zeros(1024, device="cuda")
    i = torch.zeros(4 * 64, dtype=torch.int32, device="cuda")
    o = torch.zeros(4 * 64, dtype=torch.int32, device="cuda")
    L = acc.lib()
    assert L.accspmm_gather_probe(r.data_ptr(), 8, 512, i.data_ptr(), 3, 64, o.data_ptr(),
                                  None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_gather_probe(r.data_ptr(), 8, 512, i.data_ptr(), 4, 32, o.data_ptr(),
                                  None) == acc.Status.ERR_INVALID_ARG
    assert L.accspmm_gather_probe(r.data_ptr(), 8, 128, i.data_ptr(), 4, 64, o.data_ptr(),
                                  None) == acc.Status.ERR_UNSUPPORTED
    assert acc.gather_ceiling(64, 512, waves=8, rows_per_wave=64, reps=1) > 0
