"""The measured workloads at full size, on the GPU, against the FP64 oracle:

* the 232,965-node power-law graph (114.8M nnz, 1/deg values) at
  K = 32, 64, 128, 256, 512 in the FP32 form;
* the precision forms at K = 128 on that graph and on the 100^3 27-point
  stencil (FP32 and split-BF16 within 1e-5, single-BF16 within 1e-2);
* the 256-graph batch (500-50,000 nodes, K in {16, 64, 128}): one
  persistent launch, every C equal to its standalone C bit for bit and
  within 1e-5 of the oracle;
* the serving pool (64 graphs plus two 2,000,000-node graphs): a few hundred
  requests at Poisson arrivals through the continuous-batching server, every
  served C equal to the same request run alone bit for bit and within 1e-5
  of the oracle.

The inputs are the ones bench.py measures (accspmm_amd.workloads).  The
oracle's FP64 product of each distinct (graph, B) is computed once.
"""
import time

import numpy as np
import pytest

import accspmm_amd as acc
from accspmm_amd.workloads import (WORKLOADS, batch_specs, make_graph, poisson_gaps,
                                   serving_pool, serving_requests)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {acc.FORM_FP32: 1e-5, acc.FORM_SPLIT_BF16: 1e-5, acc.FORM_BF16: 1e-2}
FORMS = [acc.FORM_FP32, acc.FORM_SPLIT_BF16, acc.FORM_BF16]


class _Matrix:
    """A workload graph built with the bench's automatic layout, its B per K
    and the oracle's FP64 C and |A||B| per K (computed once)."""

    def __init__(self, name):
        spec = WORKLOADS[name]
        self.spec = spec
        t0 = time.perf_counter()
        self.csr = make_graph(spec)
        self.graph = acc.DeviceGraph(self.csr, acc.default_build_opts(0, reorder_rows=2))
        print(f"{name}: {self.csr.nnz} nnz, generated and built in "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        self._ref = {}

    def B(self, K):
        return acc.gen_dense(self.spec["bseed"] + K, self.csr.cols, K)

    def ref(self, K, Bh):
        if K not in self._ref:
            c = self.csr
            self._ref = {K: (O.spmm_f64(c.row_ptr, c.col_idx, c.data, Bh),
                             O.spmm_abs_f64(c.row_ptr, c.col_idx, c.data, Bh))}
        return self._ref[K]

    def run(self, K, form):
        import torch
        Bh = self.B(K)
        C = torch.full((self.csr.rows, K), float("nan"), device="cuda")
        self.graph.spmm(torch.from_numpy(Bh).cuda(), C, form)
        torch.cuda.synchronize()
        C64, ab = self.ref(K, Bh)
        return O.errors(C.cpu().numpy(), C64, ab)


@pytest.fixture(scope="module")
def headline():
    m = _Matrix("chung_lu_233k")
    yield m
    m.graph.close()


@pytest.fixture(scope="module")
def stencil():
    m = _Matrix("stencil27_100")
    yield m
    m.graph.close()


@pytest.mark.parametrize("K", [32, 64, 128, 256, 512])
def test_headline_graph_k_sweep_fp32(headline, K):
    mx, fr = headline.run(K, acc.FORM_FP32)
    print(f"233k K={K} fp32 max_rel={mx:.3e} fro_rel={fr:.3e}", flush=True)
    assert mx <= TOL[acc.FORM_FP32], (K, mx)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_headline_graph_forms_k128(headline, form):
    mx, fr = headline.run(128, form)
    print(f"233k K=128 {acc.FORM_NAMES[form]} max_rel={mx:.3e} fro_rel={fr:.3e}", flush=True)
    assert mx <= TOL[form], (form, mx)


@pytest.mark.parametrize("form", FORMS, ids=acc.FORM_NAMES.get)
def test_stencil_100_forms_k128(stencil, form):
    assert stencil.csr.nnz == 298 ** 3
    mx, fr = stencil.run(128, form)
    print(f"stencil K=128 {acc.FORM_NAMES[form]} max_rel={mx:.3e} fro_rel={fr:.3e}", flush=True)
    assert mx <= TOL[form], (form, mx)


def test_batch_256_graphs():
    import torch
    form = acc.FORM_SPLIT_BF16
    probs = []
    for sp in batch_specs(256):
        csr = acc.gen_chung_lu(sp["n"], sp["deg"], sp["gamma"], sp["seed"], 0)
        g = acc.DeviceGraph(csr, acc.default_build_opts(8, build_adpbme=0))
        Bh = acc.gen_dense(sp["bseed"], csr.cols, sp["K"])
        B = torch.from_numpy(Bh).cuda()
        probs.append((csr, g, Bh, B, torch.full((csr.rows, sp["K"]), float("nan"),
                                                device="cuda")))
    alone = [g.spmm(B, None, form).clone() for _, g, _, B, _ in probs]
    batch = acc.Batch([(g, B, C) for _, g, _, B, C in probs], form)
    batch.run()
    torch.cuda.synchronize()
    worst = 0.0
    for (csr, g, Bh, B, C), C0 in zip(probs, alone):
        assert torch.equal(C, C0), "batched C differs from the standalone C"
        mx, _ = O.errors(C.cpu().numpy(), O.spmm_f64(csr.row_ptr, csr.col_idx, csr.data, Bh),
                         O.spmm_abs_f64(csr.row_ptr, csr.col_idx, csr.data, Bh))
        worst = max(worst, mx)
    # a hipGraph replay gives the same bits
    for _, _, _, _, C in probs:
        C.fill_(float("nan"))
    batch.run(graph=True)
    torch.cuda.synchronize()
    assert all(torch.equal(C, C0) for (_, _, _, _, C), C0 in zip(probs, alone))
    batch.close()
    print(f"batch: 256 graphs, {sum(p[0].nnz for p in probs)} nnz, worst max_rel {worst:.3e}",
          flush=True)
    assert worst <= TOL[form]


def test_serving_pool_poisson():
    import torch
    form = acc.FORM_SPLIT_BF16
    t0 = time.perf_counter()
    pool = serving_pool()
    assert len(pool) == 66 and pool[-1].rows == 2_000_000
    print(f"serving pool: {sum(c.nnz for c in pool)} nnz, generated in "
          f"{time.perf_counter() - t0:.1f} s", flush=True)
    requests = 300
    gsel, ksel = serving_requests(len(pool), requests)
    # both 2M-node graphs get requests of both K
    big = len(pool) - 2
    gsel[:4] = [big, big, big + 1, big + 1]
    ksel[:4] = [32, 128, 128, 32]
    opts = acc.default_build_opts(0, reorder_rows=2, build_adpbme=0)
    # the server builds its graphs with these options too (automatic layout)
    graphs = [acc.DeviceGraph(c, opts) for c in pool]
    resident = sum(int(g.info.device_bytes) for g in graphs)
    # a budget that holds the pool plus 3 GiB of in-flight workspace (a
    # K = 128 request on a 2M-node graph reserves 1 GiB for its converted B)
    ws_budget = 3 << 30
    srv = acc.Server(hbm_budget_bytes=resident + ws_budget, max_batch=64, max_inflight=3,
                     form=form)
    gids = [srv.register(c) for c in pool]
    Bs, Bh = {}, {}
    for g, K in sorted(set(zip(gsel.tolist(), ksel.tolist()))):
        Bh[(g, K)] = acc.gen_dense(7000 + 1000 * g + K, pool[g].cols, K)
        Bs[(g, K)] = torch.from_numpy(Bh[(g, K)]).cuda()
    Cs = [torch.full((pool[g].rows, int(K)), float("nan"), device="cuda")
          for g, K in zip(gsel, ksel)]
    torch.cuda.synchronize()
    gaps = poisson_gaps(requests, 2000.0)
    rids = []
    t_next = time.perf_counter()
    for i in range(requests):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        key = (int(gsel[i]), int(ksel[i]))
        rids.append(srv.submit(gids[key[0]], Bs[key], Cs[i]))
    srv.drain()
    st = srv.stats()
    assert st.resident_bytes == resident
    lats = np.array([srv.poll(r) for r in rids])
    assert np.all(lats > 0)
    assert st.completed == requests and st.launches < requests
    assert st.peak_inflight_bytes <= ws_budget
    srv.close()
    print(f"serving: {requests} requests in {st.launches} launches, p50 "
          f"{np.percentile(lats, 50) * 1e3:.3f} ms, p99 {np.percentile(lats, 99) * 1e3:.3f} ms",
          flush=True)
    alone = {}
    for i in range(requests):
        key = (int(gsel[i]), int(ksel[i]))
        if key not in alone:
            alone[key] = graphs[key[0]].spmm(Bs[key], None, form).clone()
        assert torch.equal(Cs[i], alone[key]), (i, key)
    worst = 0.0
    for key, C0 in alone.items():
        c = pool[key[0]]
        mx, _ = O.errors(C0.cpu().numpy(), O.spmm_f64(c.row_ptr, c.col_idx, c.data, Bh[key]),
                         O.spmm_abs_f64(c.row_ptr, c.col_idx, c.data, Bh[key]))
        worst = max(worst, mx)
    print(f"serving: {len(alone)} distinct (graph, B), worst max_rel {worst:.3e}", flush=True)
    assert worst <= TOL[form]
