"""Serving loop: results under continuous batching equal the standalone
results bit for bit, admission honours the HBM budget, and the error paths
(unknown graph, full queue, budget) are reported."""
import time

import numpy as np
import pytest

import accspmm_amd as acc

pytestmark = pytest.mark.gpu


def _pool(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nodes = int(round(np.exp(rng.uniform(np.log(500), np.log(8000)))))
        out.append(acc.gen_chung_lu(nodes, float(rng.uniform(4, 64)),
                                    float(rng.uniform(2.0, 2.5)), 3000 + i, 0))
    return out


@pytest.mark.parametrize("large_work,layout", [(0, "input"), (3_000_000, "input"),
                                               (0, "rw16_reordered"), (0, "mixed")])
def test_mix_equals_alone(large_work, layout):
    # large_work = 3e6 makes the requests of the bigger graphs (nnz * K above
    # it) run in launches of their own beside batches of the small ones;
    # rw16_reordered registers graphs with 16-row windows and reordered rows;
    # mixed alternates the two layouts (a launch serves one row window)
    import torch
    pool = _pool(8, 1)
    srv = acc.Server(max_batch=8, max_inflight=3, form=acc.FORM_SPLIT_BF16,
                     large_work=large_work)
    rw16 = acc.default_build_opts(16, reorder_rows=1)
    opts = [None if layout == "input" or (layout == "mixed" and i % 2) else rw16
            for i in range(len(pool))]
    gids = [srv.register(c, o) for c, o in zip(pool, opts)]
    graphs = [acc.DeviceGraph(c, o) for c, o in zip(pool, opts)]
    rng = np.random.default_rng(2)
    reqs = []
    for i in range(40):
        j = int(rng.integers(len(pool)))
        K = int(rng.choice([32, 128]))
        B = torch.from_numpy(acc.gen_dense(100 + i, pool[j].cols, K)).cuda()
        Cm = torch.full((pool[j].rows, K), float("nan"), device="cuda")
        reqs.append((j, B, Cm))
    torch.cuda.synchronize()
    jobs = []
    # the first 20 queue up while admission is paused, so they must share
    # launches; the rest arrive while those launches are on the GPU
    srv.pause(True)
    for j, B, Cm in reqs[:20]:
        jobs.append((j, B, Cm, srv.submit(gids[j], B, Cm)))
    srv.pause(False)
    for i, (j, B, Cm) in enumerate(reqs[20:]):
        jobs.append((j, B, Cm, srv.submit(gids[j], B, Cm)))
        if i % 7 == 0:
            time.sleep(0.001)
    srv.drain()
    for j, B, Cm, rid in jobs:
        lat = srv.poll(rid)
        assert lat is not None and lat > 0
        alone = graphs[j].spmm(B, None, acc.FORM_SPLIT_BF16)
        torch.cuda.synchronize()
        assert torch.equal(Cm, alone)
    st = srv.stats()
    assert st.completed == 40 and st.launches < 40  # requests were batched
    srv.close()


def test_one_per_launch_baseline_matches():
    import torch
    pool = _pool(3, 3)
    srv = acc.Server(form=acc.FORM_SPLIT_BF16, one_per_launch=True)
    gids = [srv.register(c) for c in pool]
    outs = []
    for i, (c, g) in enumerate(zip(pool, gids)):
        B = torch.from_numpy(acc.gen_dense(7 + i, c.cols, 64)).cuda()
        Cm = torch.zeros((c.rows, 64), device="cuda")
        outs.append((c, B, Cm, srv.submit(g, B, Cm)))
    srv.drain()
    for c, B, Cm, rid in outs:
        assert srv.poll(rid) is not None
        assert torch.equal(Cm, acc.DeviceGraph(c).spmm(B, None, acc.FORM_SPLIT_BF16))
    assert srv.stats().launches == 3
    srv.close()


def test_budget_queue_and_unknown_graph():
    import torch
    c = acc.gen_chung_lu(4000, 16.0, 2.2, 9, 0)
    probe = acc.Server()
    probe.register(c)
    gbytes = probe.stats().resident_bytes
    probe.close()
    # a budget that holds the graph but not a second copy (BF16 form: its
    # requests always need a converted copy of B in the workspace)
    srv = acc.Server(hbm_budget_bytes=int(gbytes * 1.5), max_queue=2,
                     max_inflight=1, form=acc.FORM_BF16)
    gid = srv.register(c)
    with pytest.raises(acc.AccSpMMError) as e:
        srv.register(c)
    assert e.value.status == acc.Status.ERR_BUDGET
    B = torch.zeros((c.cols, 16), device="cuda")
    Cm = torch.zeros((c.rows, 16), device="cuda")
    with pytest.raises(acc.AccSpMMError) as e:
        srv.submit(gid + 99, B, Cm)
    assert e.value.status == acc.Status.ERR_NOT_FOUND
    # a request whose workspace can never fit is refused up front
    Bbig = torch.zeros((c.cols, 4096), device="cuda")
    Cbig = torch.zeros((c.rows, 4096), device="cuda")
    with pytest.raises(acc.AccSpMMError) as e:
        srv.submit(gid, Bbig, Cbig)
    assert e.value.status == acc.Status.ERR_BUDGET
    # queue full while paused
    srv.pause(True)
    r1 = srv.submit(gid, B, Cm)
    r2 = srv.submit(gid, B, Cm)
    with pytest.raises(acc.AccSpMMError) as e:
        srv.submit(gid, B, Cm)
    assert e.value.status == acc.Status.ERR_QUEUE_FULL
    assert srv.poll(r1) is None  # still queued
    with pytest.raises(acc.AccSpMMError) as e:
        srv.unregister(gid)      # requests still reference it
    assert e.value.status == acc.Status.ERR_NOT_READY
    srv.pause(False)
    srv.drain()
    assert srv.poll(r1) is not None and srv.poll(r2) is not None
    srv.unregister(gid)
    assert srv.stats().resident_bytes == 0
    srv.close()


def test_register_file_warm_start(tmp_path):
    # a graph saved once is served from the file without a rebuild
    import torch
    c = _pool(1, 9)[0]
    g = acc.DeviceGraph(c)
    path = tmp_path / "g.metcf"
    g.save(path)
    srv = acc.Server(form=acc.FORM_SPLIT_BF16)
    gid = srv.register_file(path)
    B = torch.from_numpy(acc.gen_dense(11, c.cols, 64)).cuda()
    Cm = torch.zeros((c.rows, 64), device="cuda")
    rid = srv.submit(gid, B, Cm)
    srv.drain()
    assert srv.poll(rid) is not None
    assert torch.equal(Cm, g.spmm(B, None, acc.FORM_SPLIT_BF16))
    with pytest.raises(acc.AccSpMMError):
        srv.register_file(tmp_path / "missing.metcf")
    srv.close()


def test_plain_register_is_the_automatic_layout():
    # accspmm_server_register (no build options) serves the graph in the
    # library's automatic layout: here rows average >= 128 nonzeros, so
    # 16-row windows and reordered rows; C still equals that layout's
    # standalone C, in input row order
    import torch
    from helpers import random_csr
    rng = np.random.default_rng(92)
    rp, ci, va = random_csr(rng, 600, 1500, 0.12, empty_rows=0.05, heavy_rows=2)
    c = acc.CSR(600, 1500, rp, ci, va)
    assert c.nnz >= 128 * c.rows
    srv = acc.Server(form=acc.FORM_SPLIT_BF16)
    rp_d, ci_d, va_d = c.to_device()
    view = acc.CsrView(c.rows, c.cols, c.nnz, rp_d.data_ptr(), ci_d.data_ptr(),
                       va_d.data_ptr())
    gid = acc.C.c_uint32()
    torch.cuda.synchronize()
    assert acc.lib().accspmm_server_register(srv.handle, acc.C.byref(view),
                                             acc.C.byref(gid)) == acc.Status.OK
    B = torch.from_numpy(acc.gen_dense(13, c.cols, 32)).cuda()
    Cm = torch.full((c.rows, 32), float("nan"), device="cuda")
    rid = srv.submit(gid.value, B, Cm)
    srv.drain()
    assert srv.poll(rid) is not None
    auto = acc.DeviceGraph(c, acc.default_build_opts(0, reorder_rows=2))
    assert auto.info.row_window == 16 and auto.info.reordered == 1
    assert torch.equal(Cm, auto.spmm(B, None, acc.FORM_SPLIT_BF16))
    srv.close()
