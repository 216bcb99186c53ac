"""The measured workloads as seeded inputs, shared by bench.py and the
full-size GPU tests so that both run exactly the same matrices:

* chung_lu_233k: 232,965-node symmetric Chung-Lu power-law graph (exponent
  2.1, mean degree 492, self loops, seeded relabel, ~114.8M nnz), values
  1/deg of the row; B seed 12 + K.
* stencil27_100: 27-point stencil on a 100^3 grid (26,463,592 nnz), values
  U[-1,1]; B seed 14 + K.
* chung_lu_16k: 16,384-node Chung-Lu, exponent 2.1, mean degree 16, U(0,1]
  symmetric values; B seed 2 (+ K).
* chung_lu_2m: one of the serving pool's two 2,000,000-node graphs (mean
  degree 32, exponent 2.1, U(0,1] values, graph seed 3100).
* stencil27_100_scrambled: the 100^3 stencil with its nodes relabelled by a
  seeded random permutation (P A P^T), the case a locality ordering must
  undo; B seed 14 + K.
* the batch: graph i has 500-50,000 nodes (log-uniform), mean degree
  U[4,64], exponent U[2.0,2.5], U(0,1] values, graph seed 1000 + i, K from
  {16, 64, 128}, B seed 2000 + i (numpy Generator seeded 1000 draws the
  sizes).
* the serving pool: 64 graphs drawn like the batch's (generator seeded
  3000, graph seed 3000 + i) plus two 2,000,000-node mean-degree-32 graphs
  (seeds 3100, 3101); requests draw a graph uniformly and K from {32, 128}
  (generator seeded 4000); arrivals are Poisson (generator seeded 5000).
"""
from __future__ import annotations

import math

import numpy as np

WORKLOADS = {
    "chung_lu_233k": dict(kind="chung_lu", n=232965, deg=492.0, gamma=2.1, seed=11,
                          values=1, bseed=12),
    "stencil27_100": dict(kind="stencil", g=100, seed=13, bseed=14),
    "chung_lu_16k": dict(kind="chung_lu", n=16384, deg=16.0, gamma=2.1, seed=1,
                         values=0, bseed=2),
    "chung_lu_2m": dict(kind="chung_lu", n=2000000, deg=32.0, gamma=2.1, seed=3100,
                        values=0, bseed=3101),
}


def make_graph(spec: dict):
    """The CSR of a WORKLOADS entry (seeded, deterministic)."""
    from . import gen_chung_lu, gen_stencil27
    if spec["kind"] == "chung_lu":
        return gen_chung_lu(spec["n"], spec["deg"], spec["gamma"], spec["seed"],
                            spec["values"])
    return gen_stencil27(spec["g"], spec["seed"])


def symmetric_permuted(csr, perm):
    """P A P^T as a canonical CSR: row i is input row perm[i], columns
    relabelled by the inverse permutation and sorted within each row."""
    from . import CSR
    n = csr.rows
    perm = np.asarray(perm, np.int64)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    rp = csr.row_ptr.astype(np.int64)
    lens = np.diff(rp)[perm]
    nrp = np.concatenate([[0], np.cumsum(lens)])
    src = np.repeat(rp[perm] - nrp[:-1], lens) + np.arange(nrp[-1])
    rows = np.repeat(np.arange(n), lens)
    cols = inv[csr.col_idx.astype(np.int64)[src]]
    order = np.lexsort((cols, rows))
    return CSR(n, n, nrp.astype(np.uint32), cols[order].astype(np.uint32),
               csr.data[src][order].astype(np.float32))


def make_graph_by_name(name: str):
    """(CSR, B seed) of a WORKLOADS entry or of stencil27_100_scrambled."""
    if name == "stencil27_100_scrambled":
        base = WORKLOADS["stencil27_100"]
        c = make_graph(base)
        p = np.random.default_rng(17).permutation(c.rows)
        return symmetric_permuted(c, p), base["bseed"]
    spec = WORKLOADS[name]
    return make_graph(spec), spec["bseed"]


def batch_specs(ngraphs: int = 256) -> list[dict]:
    """Sizes, K and seeds of the batch's graphs, in batch order."""
    rng = np.random.default_rng(1000)
    out = []
    for i in range(ngraphs):
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        deg = float(rng.uniform(4, 64))
        gam = float(rng.uniform(2.0, 2.5))
        K = int(rng.choice([16, 64, 128]))
        out.append(dict(n=n, deg=deg, gamma=gam, K=K, seed=1000 + i, bseed=2000 + i))
    return out


def serving_pool(small: int = 64, big: int = 2, big_nodes: int = 2_000_000) -> list:
    """The serving pool's CSRs: `small` graphs like the batch's, then `big`
    big_nodes-node graphs of mean degree 32."""
    from . import gen_chung_lu
    rng = np.random.default_rng(3000)
    pool = []
    for i in range(small):
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        pool.append(gen_chung_lu(n, float(rng.uniform(4, 64)),
                                 float(rng.uniform(2.0, 2.5)), 3000 + i, 0))
    for i in range(big):
        pool.append(gen_chung_lu(big_nodes, 32.0, 2.1, 3100 + i, 0))
    return pool


SERVING_KS = (32, 128)


def serving_requests(npool: int, requests: int, seed: int = 4000):
    """(graph index, K) of each request."""
    rng = np.random.default_rng(seed)
    gsel = rng.integers(0, npool, size=requests)
    ksel = rng.choice(SERVING_KS, size=requests)
    return gsel, ksel


def poisson_gaps(requests: int, qps: float) -> np.ndarray:
    """Inter-arrival gaps (seconds) of a Poisson stream at qps (zeros if qps
    is 0: all requests at once)."""
    if not qps:
        return np.zeros(requests)
    return np.random.default_rng(5000).exponential(1.0 / qps, size=requests)
