"""Reverse Cuthill-McKee order (accspmm_order_rcm), the builder's
symmetric_order 1, on the host: a permutation of every pattern, the
Cuthill-McKee structure (a breadth-first order: parents' positions never
decrease, siblings by increasing degree, each component from one of its
nodes), and bandwidths against the natural order of grids (RCM recovers a band
close to it from a scrambled labelling) and against scipy's
reverse_cuthill_mckee."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

import accspmm_amd as acc


def _sym_pattern(rng, n, p):
    m = sp.random(n, n, density=p, random_state=rng, format="csr")
    m = ((m + m.T) > 0).astype(np.float32).tocsr()
    m.sort_indices()
    return m


def _bandwidth(m, perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    coo = m.tocoo()
    if coo.nnz == 0:
        return 0
    return int(np.max(np.abs(inv[coo.row].astype(np.int64) - inv[coo.col].astype(np.int64))))


def _grid27(g):
    c = acc.gen_stencil27(g, 1)
    return sp.csr_matrix((c.data, c.col_idx, c.row_ptr), shape=(c.rows, c.cols))


def _scramble(m, seed):
    p = np.random.default_rng(seed).permutation(m.shape[0])
    return m[p][:, p].tocsr(), p


def _rcm(m):
    m = m.tocsr()
    m.sort_indices()
    return acc.order_rcm(m.indptr, m.indices, m.shape[0])


def _check_cm_structure(m, perm):
    """perm reversed is a Cuthill-McKee order: breadth first over each
    component, the nodes first reached from one parent consecutive and in
    increasing degree (ties: index)."""
    n = m.shape[0]
    cm = perm[::-1]
    pos = np.empty(n, np.int64)
    pos[cm] = np.arange(n)
    deg = np.diff(m.indptr)
    last_parent = -1
    prev = None
    for i in range(n):
        v = cm[i]
        nb = m.indices[m.indptr[v]:m.indptr[v + 1]]
        earlier = pos[nb][pos[nb] < i]
        if len(earlier) == 0:
            # a component's start: every earlier node's neighbours are placed
            last_parent = -1
            prev = None
            continue
        parent = int(earlier.min())
        assert parent >= last_parent, i
        if parent == last_parent and prev is not None:
            assert (deg[prev], prev) <= (deg[v], v) or deg[prev] < deg[v], i
        last_parent, prev = parent, v


def test_permutation_and_structure():
    rng = np.random.default_rng(3)
    for n, p in ((1, 0.0), (7, 0.3), (200, 0.02), (500, 0.004), (1000, 0.001)):
        m = _sym_pattern(rng, n, p)
        perm = _rcm(m)
        assert sorted(perm.tolist()) == list(range(n))
        _check_cm_structure(m, perm)


def test_disconnected_and_empty_rows():
    # two grids side by side plus isolated nodes
    a = _grid27(4)
    b = _grid27(3)
    m = sp.block_diag([a, sp.csr_matrix((5, 5)), b]).tocsr()
    m.sort_indices()
    perm = _rcm(m)
    assert sorted(perm.tolist()) == list(range(m.shape[0]))
    _check_cm_structure(m, perm)
    # a component never interleaves with another in the order
    comp = np.r_[np.zeros(64), np.arange(1, 6), np.full(27, 6)]
    seq = comp[perm]
    changes = np.count_nonzero(np.diff(seq))
    assert changes == 6


@pytest.mark.parametrize("g", [6, 10, 14])
def test_scrambled_grid_band_recovered(g):
    m = _grid27(g)
    natural = _bandwidth(m, np.arange(m.shape[0]))
    s, _ = _scramble(m, g)
    scrambled = _bandwidth(s, np.arange(s.shape[0]))
    perm = _rcm(s)
    got = _bandwidth(s, perm)
    ref = _bandwidth(s, reverse_cuthill_mckee(s, symmetric_mode=True).astype(np.int64))
    # breadth-first levels of a 27-point grid from a corner are the shells
    # max(x, y, z) = k of up to 3g^2 nodes, so RCM's band is about three
    # times the lexicographic g^2 + g + 1, far below a scrambled order's
    assert got <= 3.5 * natural, (got, natural)
    assert got <= 1.25 * ref, (got, ref)
    assert got * 2 < scrambled, (got, scrambled)


def test_power_law_band_not_worse_than_scipy():
    c = acc.gen_chung_lu(3000, 12.0, 2.1, 4, 0)
    m = sp.csr_matrix((c.data, c.col_idx, c.row_ptr), shape=(c.rows, c.cols))
    perm = _rcm(m)
    ref = reverse_cuthill_mckee(m, symmetric_mode=True).astype(np.int64)
    assert _bandwidth(m, perm) <= 1.25 * _bandwidth(m, ref)


def test_deterministic_and_bad_input():
    m = _sym_pattern(np.random.default_rng(9), 300, 0.02)
    assert np.array_equal(_rcm(m), _rcm(m))
    rp = np.array([0, 1, 2], np.uint32)
    with pytest.raises(acc.AccSpMMError):
        acc.order_rcm(rp, np.array([0, 5], np.uint32))      # column >= n
    with pytest.raises(acc.AccSpMMError):
        acc.order_rcm(np.array([0, 2, 1], np.uint32), np.array([0, 1], np.uint32))


def test_unsymmetric_pattern_is_a_permutation():
    # entries (i, j) without (j, i) reach nodes of components already
    # placed; the order must still hold every node exactly once
    rng = np.random.default_rng(21)
    for n, p in ((50, 0.05), (400, 0.006), (2000, 0.001)):
        m = sp.random(n, n, density=p, random_state=rng, format="csr")
        m = (m > 0).astype(np.float32).tocsr()
        m.sort_indices()
        perm = _rcm(m)
        assert sorted(perm.tolist()) == list(range(n))
    # a chain 0 -> 1 -> 2 with edges only one way, then an isolated node
    rp = np.array([0, 1, 2, 2, 2], np.uint32)
    ci = np.array([1, 2], np.uint32)
    perm = acc.order_rcm(rp, ci, 4)
    assert sorted(perm.tolist()) == [0, 1, 2, 3]
