"""Multigrid hierarchy and inter-grid transfers on the GPU against the oracle.

Each level's DA (coarsenOctree + balanceOctree of the level below,
src/omg/blockDiag.cpp:1384-1425) must be bit-exact with the oracle's, and
the library's R and P (dgx_mg_restrict / dgx_mg_prolong) must equal the
oracle's restatement of restrictMatVecType1 / prolongMatVecType1 with the
touched flags and the reference's stencil and vertex-map files
(src/omg/restrictMatVec.cpp:71-619, prolongMatVec.cpp:154-292), the boundary
nodes of both levels suppressed, to 1e-14 relative (max-norm; FP64 sums in a
different order)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RTOL = 1e-14
NONE = 0xFFFFFFFF


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def geometric_boundary(da):
    X = da.node_coords()
    D = 1 << (int(da.octs[0, 6]) - 1)
    return ((X == 0) | (X == D)).any(axis=1).astype(np.uint8)


@pytest.mark.parametrize("curve,dist,n,m,seed", [(0, 1, 30000, 1, 21), (0, 2, 40000, 4, 22), (1, 0, 20000, 1, 23),
                                                 (1, 2, 30000, 2, 24)])
def test_hierarchy_and_transfers(dgx, oracle, curve, dist, n, m, seed):
    kw = dict(n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.02) if dist == 2 else {}
    pts = dgx.gen_points(dist, seed, n, **kw)
    print(f"transfers curve={curve} dist={dist} n={n} m={m}", flush=True)
    tabs = dict(np.load(os.path.join(GOLDEN, "rmat_stencils.npz")))
    oracle.set_ordering(curve == 0)
    try:
        ob = dgx.balance_octree(dgx.points2octree(pts, 30, m, curve))
        mg = dgx.MG(ob, nlevels=3)
        assert mg.nlevels >= 3
        o = oracle.balance(oracle.points2octree(pts, 30, m))
        assert np.array_equal(o, ob.treenodes())
        levels = []
        for k in range(mg.nlevels):
            if k:
                o = oracle.balance(oracle.coarsen(o))
            rda = oracle.DA(o)
            gda = mg.level_da(k)
            nl, mk, ol, l2n = gda.maps()
            assert gda.element_size == rda.n_elem and gda.local_buffer_size == rda.n_local, f"level {k} sizes"
            assert np.array_equal(nl, rda.nlist), f"level {k} nlist"
            assert np.array_equal(mk, rda.mask), f"level {k} masks"
            assert np.array_equal(ol, rda.oct_levels), f"level {k} oct levels"
            assert np.array_equal(l2n, rda.local_to_node), f"level {k} node compaction"
            bf = gda.boundary_flags()
            sel = l2n != NONE
            gb = np.zeros(rda.n_nodes, np.uint8)
            gb[l2n[sel]] = bf[sel] != 0
            sup = geometric_boundary(rda)
            assert np.array_equal(gb, sup), f"level {k} boundary nodes"
            levels.append((rda, sup))
        rng = np.random.default_rng(seed)
        for k in range(mg.nlevels - 1):
            (df, sf), (dc, sc) = levels[k], levels[k + 1]
            r, c, v = df.restriction(dc, tabs, sf, sc)
            R = sp.csr_matrix((v, (r, c)), shape=(dc.n_nodes, df.n_nodes))
            f = rng.standard_normal(df.n_nodes)
            fd = dgx.DeviceArray.from_host(f)
            cd = dgx.DeviceArray((dc.n_nodes,), np.float64)
            mg.restrict(k, fd, cd)
            assert rel_err(cd.to_host(), R @ f) < RTOL, f"R at level {k}"
            u = rng.standard_normal(dc.n_nodes)
            ud = dgx.DeviceArray.from_host(u)
            pd = dgx.DeviceArray.from_host(rng.standard_normal(df.n_nodes))  # overwritten
            mg.prolong(k, ud, pd)
            assert rel_err(pd.to_host(), R.T @ u) < RTOL, f"P at level {k}"
        mg.free()
    finally:
        oracle.set_ordering(True)
