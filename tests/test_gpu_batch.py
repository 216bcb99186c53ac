"""Batched persistent launch: every problem's C equals its standalone C bit for
bit, matches the FP64 oracle within its form's tolerance, and hipGraph
replays reproduce the eager result."""
import numpy as np
import pytest

import accspmm_amd as acc
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {acc.FORM_FP32: 1e-5, acc.FORM_SPLIT_BF16: 1e-5, acc.FORM_BF16: 1e-2}


def _pool(n, seed, nmin=500, nmax=6000):
    """Chung-Lu graphs: node counts log-uniform, mean degree U[4,64],
    exponent U[2.0,2.5], U(0,1] values; K from {16, 64, 128}."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nodes = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        deg = float(rng.uniform(4, 64))
        gam = float(rng.uniform(2.0, 2.5))
        K = int(rng.choice([16, 64, 128]))
        csr = acc.gen_chung_lu(nodes, deg, gam, 1000 + i + 100 * seed, 0)
        out.append((csr, K, 2000 + i + 100 * seed))
    return out


@pytest.mark.parametrize("form", [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16],
                         ids=acc.FORM_NAMES.get)
@pytest.mark.parametrize("layouts", ["rw8", "mixed"])
def test_batch_equals_alone(form, layouts):
    # mixed: every other graph has 16-row windows and reordered rows, so the
    # batch runs one persistent kernel per row-window size
    import torch
    pool = _pool(12, form)
    graphs, Bs, Cs, alone, hostB = [], [], [], [], []
    for i, (csr, K, bseed) in enumerate(pool):
        opts = acc.default_build_opts(16, reorder_rows=1) if layouts == "mixed" and i % 2 \
            else None
        g = acc.DeviceGraph(csr, opts)
        Bh = acc.gen_dense(bseed, csr.cols, K)
        B = torch.from_numpy(Bh).cuda()
        graphs.append(g)
        Bs.append(B)
        hostB.append(Bh)
        Cs.append(torch.full((csr.rows, K), float("nan"), device="cuda"))
        alone.append(g.spmm(B, None, form).clone())
    batch = acc.Batch(list(zip(graphs, Bs, Cs)), form)
    assert batch.num_tasks > 0
    batch.run()
    torch.cuda.synchronize()
    for (csr, K, _), Bh, C1, C0 in zip(pool, hostB, Cs, alone):
        assert torch.equal(C1, C0), "batched result differs from the standalone one"
        C64 = O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        ab = O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, Bh)
        mx, _ = O.errors(C1.cpu().numpy(), C64, ab)
        assert mx <= TOL[form]
    batch.close()


def test_batch_graph_replay_and_rebind():
    import torch
    pool = _pool(6, 7)
    graphs = [acc.DeviceGraph(c) for c, _, _ in pool]
    Bs = [torch.from_numpy(acc.gen_dense(s, c.cols, K)).cuda() for c, K, s in pool]
    Cs = [torch.zeros((c.rows, K), device="cuda") for c, K, _ in pool]
    form = acc.FORM_SPLIT_BF16
    batch = acc.Batch(list(zip(graphs, Bs, Cs)), form)
    batch.run()
    torch.cuda.synchronize()
    eager = [c.clone() for c in Cs]
    for rep in range(3):
        for c in Cs:
            c.fill_(float(rep + 1))  # eager work between replays
        torch.cuda.synchronize()
        batch.run(graph=True)
        torch.cuda.synchronize()
        for c, e in zip(Cs, eager):
            assert torch.equal(c, e), f"graph replay {rep} differs from eager run"
    # rebind to new B and C buffers
    Bs2 = [torch.from_numpy(acc.gen_dense(s + 50, c.cols, K)).cuda() for c, K, s in pool]
    Cs2 = [torch.zeros((c.rows, K), device="cuda") for c, K, _ in pool]
    arr = (acc.Problem * len(pool))()
    for i, (g, B, Cm) in enumerate(zip(graphs, Bs2, Cs2)):
        arr[i] = acc.Problem(g.handle.value, B.data_ptr(), B.stride(0), Cm.data_ptr(),
                             Cm.stride(0), B.shape[1])
    assert acc.lib().accspmm_batch_rebind(batch.handle, arr, len(pool), None) == 0
    batch.run(graph=True)
    torch.cuda.synchronize()
    for g, B, Cm in zip(graphs, Bs2, Cs2):
        assert torch.equal(Cm, g.spmm(B, None, form))
    batch.close()


def test_batch_mixed_row_windows_and_replay():
    # 8- and 16-row graphs in one batch: one persistent kernel per row-window
    # size, eager and hipGraph replays equal to the standalone calls
    import torch
    a = acc.gen_chung_lu(1000, 8.0, 2.2, 1, 0)
    b = acc.gen_chung_lu(3000, 20.0, 2.1, 2, 0)
    gs = [acc.DeviceGraph(a), acc.DeviceGraph(a, acc.default_build_opts(16)),
          acc.DeviceGraph(b, acc.default_build_opts(16, reorder_rows=1)), acc.DeviceGraph(b)]
    Bs = [torch.from_numpy(acc.gen_dense(i, g.info.cols, K)).cuda()
          for i, (g, K) in enumerate(zip(gs, (16, 64, 128, 32)))]
    Cs = [torch.full((g.info.rows, B.shape[1]), float("nan"), device="cuda")
          for g, B in zip(gs, Bs)]
    for form in (acc.FORM_FP32, acc.FORM_SPLIT_BF16):
        batch = acc.Batch(list(zip(gs, Bs, Cs)), form)
        for graph in (False, True, True):
            for c in Cs:
                c.fill_(float("nan"))
            batch.run(graph=graph)
            torch.cuda.synchronize()
            for g, B, Cm in zip(gs, Bs, Cs):
                assert torch.equal(Cm, g.spmm(B, None, form)), (form, graph)
        batch.close()


def test_single_problem_batches():
    # a batch of one split-form problem whose B is split in registers runs
    # the single-graph kernel with the batch's workspace; eager and replayed
    # results equal the standalone call, with split windows and without
    import torch
    from helpers import random_csr
    rng = np.random.default_rng(44)
    short = acc.gen_chung_lu(20000, 6.0, 2.1, 3, 0)             # few gathers per B row
    rp, ci, va = random_csr(rng, 700, 900, 0.05, heavy_rows=4)   # heavy rows: split windows
    for csr in (short, acc.CSR(700, 900, rp, ci, va)):
        g = acc.DeviceGraph(csr)
        for K in (16, 100):
            B = torch.from_numpy(acc.gen_dense(K, csr.cols, K)).cuda()
            Cm = torch.full((csr.rows, K), float("nan"), device="cuda")
            alone = g.spmm(B, None, acc.FORM_SPLIT_BF16).clone()
            batch = acc.Batch([(g, B, Cm)], acc.FORM_SPLIT_BF16)
            for graph in (False, True, True):
                Cm.fill_(float("nan"))
                batch.run(graph=graph)
                torch.cuda.synchronize()
                assert torch.equal(Cm, alone), (csr.rows, K, graph)
            batch.close()


@pytest.mark.parametrize("form", [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16],
                         ids=acc.FORM_NAMES.get)
def test_batch_dense_groups(form):
    # 16-row graphs whose 4-tile groups hold more than 256 values, batched
    # with sparse ones: each C bit-equal to its standalone call, eager and
    # replayed, in a batch of several problems and in a batch of one
    import torch
    from helpers import dense_block_csr
    rng = np.random.default_rng(98)
    dense = acc.CSR(200, 700, *dense_block_csr(rng, 200, 700, 100, 0.85))
    sparse = acc.gen_chung_lu(3000, 12.0, 2.2, 9, 0)
    gs = [acc.DeviceGraph(dense, acc.default_build_opts(16)), acc.DeviceGraph(sparse),
          acc.DeviceGraph(dense, acc.default_build_opts(16, reorder_rows=1)),
          acc.DeviceGraph(sparse, acc.default_build_opts(16))]
    assert gs[0].info.max_group_values > 256
    Ks = (64, 16, 128, 32)
    Bs = [torch.from_numpy(acc.gen_dense(i + 3, g.info.cols, K)).cuda()
          for i, (g, K) in enumerate(zip(gs, Ks))]
    Cs = [torch.full((g.info.rows, B.shape[1]), float("nan"), device="cuda")
          for g, B in zip(gs, Bs)]
    alone = [g.spmm(B, None, form).clone() for g, B in zip(gs, Bs)]
    for sel in (slice(0, 4), slice(0, 1)):
        batch = acc.Batch(list(zip(gs[sel], Bs[sel], Cs[sel])), form)
        for graph in (False, True, True):
            for c in Cs[sel]:
                c.fill_(float("nan"))
            batch.run(graph=graph)
            torch.cuda.synchronize()
            for Cm, A in zip(Cs[sel], alone[sel]):
                assert torch.equal(Cm, A), (sel, graph)
        batch.close()


@pytest.mark.parametrize("form", [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16],
                         ids=acc.FORM_NAMES.get)
def test_batch_autotune_keeps_bits(form):
    # the batch autotuner's choices (NB cap, short-window packing) leave
    # every C bit-equal to its standalone call, eager and under a graph
    # captured after the tuning
    import torch
    pool = _pool(10, 30 + form)
    graphs = [acc.DeviceGraph(c) for c, _, _ in pool]
    Bs = [torch.from_numpy(acc.gen_dense(s, c.cols, K)).cuda() for c, K, s in pool]
    Cs = [torch.full((c.rows, K), float("nan"), device="cuda") for c, K, _ in pool]
    alone = [g.spmm(B, None, form).clone() for g, B in zip(graphs, Bs)]
    batch = acc.Batch(list(zip(graphs, Bs, Cs)), form)
    batch.run(graph=True)
    ms = batch.autotune()
    assert ms > 0
    for graph in (False, True, True):
        for c in Cs:
            c.fill_(float("nan"))
        batch.run(graph=graph)
        torch.cuda.synchronize()
        for Cm, A in zip(Cs, alone):
            assert torch.equal(Cm, A), graph
    batch.close()
