"""The oracle's multigrid hierarchy and inter-grid transfers, pinned without a GPU.

- or_coarsen (coarsenOctree, src/coarsen.cpp:64-104): the output is sorted,
  linear and complete, and every output octant is an input octant or the
  parent of 8 consecutive input siblings;
- or_restrict_entries (restrictMatVecType1 with its touched flags,
  src/omg/restrictMatVec.cpp:71-619) on the reference's own stencil and
  vertex-map files: P = R^T must be the finite-element interpolation from the
  coarse to the fine mesh, so it reproduces every linear function exactly at
  every fine node (a dropped, doubled or misplaced entry breaks this), each
  (coarse, fine) pair appears once, and suppressed nodes drop out.
"""
import hashlib
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN

RMAT = os.path.join(GOLDEN, "rmat_stencils.npz")


@pytest.fixture(scope="module")
def tables():
    return dict(np.load(RMAT))


def test_fixture_hash(tables):
    with open(os.path.join(GOLDEN, "rmat_stencils.sha256")) as f:
        want = f.read().split()[0]
    h = hashlib.sha256(b"".join(tables[k].tobytes() for k in ("rmat1", "rmat2", "map1", "map2", "map3", "map4")))
    assert h.hexdigest() == want


def test_stencil_index_order(tables):
    """Element types 0 (no hanging vertex): the tables hold trilinear weights,
    rmat[..][cCtr][fCtr] = weight of coarse vertex cCtr at fine vertex fCtr."""
    r1, r2 = tables["rmat1"], tables["rmat2"]
    for cf in range(8):
        for v in range(8):
            p = [((cf >> a) & 1) * 0.5 + ((v >> a) & 1) * 0.5 for a in range(3)]
            for c in range(8):
                w = np.prod([p[a] if (c >> a) & 1 else 1 - p[a] for a in range(3)])
                assert r1[3, cf, 0, c, v] == w
    assert np.array_equal(r2[:, 0], np.broadcast_to(np.eye(8), (8, 8, 8)))


def _cloud(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        return rng.random((n, 3))
    if kind == "lognormal":
        return np.clip(np.exp(rng.normal(-1.0, 0.5, (n, 3))), 0.0, np.nextafter(1.0, 0.0))
    c = rng.uniform(0.3, 0.7, (4, 3))
    p = c[rng.integers(0, 4, n)] + rng.normal(0.0, 0.03, (n, 3))
    return np.clip(p, 0.0, np.nextafter(1.0, 0.0))


def _domain_boundary(da):
    X = da.node_coords()
    D = 1 << (int(da.octs[0, 6]) - 1)
    return ((X == 0) | (X == D)).any(axis=1).astype(np.uint8)


@pytest.mark.parametrize("hilbert", [True, False])
@pytest.mark.parametrize("kind,n,m", [("uniform", 6000, 1), ("lognormal", 12000, 2), ("gauss", 12000, 4)])
def test_coarsen_and_restriction(oracle, tables, hilbert, kind, n, m):
    oracle.set_ordering(hilbert)
    try:
        fine = oracle.balance(oracle.points2octree(_cloud(kind, n, 7), 30, m))
        c0 = oracle.coarsen(fine)
        assert oracle.is_sorted_unique(c0) and oracle.is_linear(c0) and oracle.is_complete(c0)
        fi = {tuple(t) for t in fine[:, :4]}
        pos = {tuple(t[:4]): i for i, t in enumerate(fine)}
        n_par = 0
        for t in c0:
            if tuple(t[:4]) in fi:
                continue
            # a parent: its first child in curve order and the 7 octants after it
            lev = int(t[3] & 31) + 1
            kids = [i for i in range(len(fine)) if int(fine[i, 3] & 31) == lev and all(
                fine[i, a] >> (30 - lev + 1) == t[a] >> (30 - lev + 1) for a in range(3))] if n_par < 3 else None
            if kids is not None:
                assert len(kids) == 8 and kids == list(range(kids[0], kids[0] + 8))
            n_par += 1
        assert n_par > 0 and len(c0) == len(fine) - 7 * n_par
        coarse = oracle.balance(c0)
        df, dc = oracle.DA(fine), oracle.DA(coarse)
        r, c, v = df.restriction(dc, tables)
        R = sp.coo_matrix((v, (r, c)), shape=(dc.n_nodes, df.n_nodes)).tocsr()
        assert R.nnz == len(v), "a (coarse, fine) pair contributes more than once"
        P = R.T.tocsr()
        Xc = dc.node_coords().astype(np.float64)
        Xf = df.node_coords().astype(np.float64)
        for coef in ([1.0, 0, 0, 0], [0.25, 2.0, -3.0, 0.5], [-1.0, 0.0, 1.5, -2.5]):
            lin = lambda X: coef[0] + (X @ np.array(coef[1:])) / float(1 << 29)
            assert np.abs(P @ lin(Xc) - lin(Xf)).max() <= 1e-13 * max(1.0, np.abs(lin(Xf)).max())
        # the coarse node at a fine node's position carries weight 1 there
        key = {tuple(x): i for i, x in enumerate(dc.node_coords())}
        hits = [(key[tuple(x)], j) for j, x in enumerate(df.node_coords()) if tuple(x) in key]
        assert len(hits) == dc.n_nodes
        assert all(P[j, i] == 1.0 for i, j in hits[:2000])
        # suppressed (Dirichlet) nodes drop out of rows and columns
        sf, sc = _domain_boundary(df), _domain_boundary(dc)
        r2, c2, v2 = df.restriction(dc, tables, sf, sc)
        assert not sc[r2].any() and not sf[c2].any()
        keep = (sc[r] == 0) & (sf[c] == 0)
        assert np.array_equal(r2, r[keep]) and np.array_equal(c2, c[keep]) and np.array_equal(v2, v[keep])
    finally:
        oracle.set_ordering(True)
