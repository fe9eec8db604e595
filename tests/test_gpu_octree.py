"""GPU parity tests: the HIP path (through the C ABI) against the oracle on the
same seeded inputs.  Octree and DA results must be bit-exact; the FP64 matvec
must agree to 1e-12 relative (max-norm)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

MATVEC_RTOL = 1e-12


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def check_pipeline(dgx, O, pts, md, m, L=None, M=None, curve=0):
    print(f"pipeline n={len(pts)} md={md} m={m} curve={curve}", flush=True)
    O.set_ordering(curve == 0)
    try:
        ou = dgx.points2octree(pts, md, m, curve)
        tu = ou.treenodes()
        ru = O.points2octree(pts, md, m)
        assert np.array_equal(tu, ru), "points2Octree"
        ob = dgx.balance_octree(ou)
        tb = ob.treenodes()
        rb = O.balance(ru)
        assert np.array_equal(tb, rb), "balanceOctree"
        if len(tb) <= 8:
            return ou, ob, None
        da = dgx.DA(ob)
        rda = O.DA(rb)
        nl, mk, ol, l2n = da.maps()
        assert da.element_size == rda.n_elem and da.local_buffer_size == rda.n_local
        assert da.node_size == rda.node_size and da.boundary_node_size == rda.bnd_node_size
        assert np.array_equal(nl, rda.nlist), "DA nlist"
        assert np.array_equal(mk, rda.mask), "DA hanging masks"
        assert np.array_equal(ol, rda.oct_levels), "DA oct levels"
        assert np.array_equal(l2n, rda.local_to_node), "DA node compaction"
        if L is not None:
            x = np.random.default_rng(len(pts)).standard_normal(da.n_nodes)
            y = da.matvec_host(x)
            yr = rda.matvec(x, L, M)
            assert rel_err(y, yr) < MATVEC_RTOL
            d = da.diagonal_host()
            conn, cn, lev, mk2 = rda.element_view()
            # diagonal from the reference tables: sum over elements of fac1 L_kk + fac2 M_kk
            h = 2.0 ** -(lev.astype(np.float64) - 1)
            dr = np.zeros(da.n_nodes)
            tmap = O.etype_masks()
            for e in range(rda.n_elem):
                t = 0 if mk2[e] == 0 else int(np.nonzero(tmap[cn[e]] == mk2[e])[0][0])
                dd = np.diag(L[cn[e], t]) * h[e] / 2 + np.diag(M[cn[e], t]) * h[e] ** 3 / 8
                np.add.at(dr, conn[e], dd)
            assert rel_err(d, dr) < MATVEC_RTOL
        return ou, ob, da
    finally:
        O.set_ordering(True)


def test_smoke_small_uniform(dgx, oracle, stencils):
    pts = dgx.gen_points(0, 11, 3000)
    check_pipeline(dgx, oracle, pts, 30, 1, *stencils)


@pytest.mark.parametrize("dist,n,m", [(0, 20000, 1), (1, 50000, 8), (2, 30000, 4), (1, 20000, 1), (0, 5000, 4)])
def test_pipeline_parity(dgx, oracle, stencils, dist, n, m):
    kw = dict(n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.02) if dist == 2 else {}
    pts = dgx.gen_points(dist, 100 + dist, n, **kw)
    check_pipeline(dgx, oracle, pts, 30, m, *stencils)


def test_pipeline_morton_curve(dgx, oracle, stencils):
    pts = dgx.gen_points(1, 7, 20000)
    check_pipeline(dgx, oracle, pts, 30, 4, *stencils, curve=1)


@pytest.mark.parametrize("dist,seed,n,md,m", [(0, 3, 20000, 6, 1), (0, 5, 5000, 5, 2), (1, 4, 20000, 8, 2),
                                             (1, 4, 20000, 10, 1), (2, 11, 20000, 16, 4)])
def test_max_depth_variants(dgx, oracle, stencils, dist, seed, n, md, m):
    """Shallow trees: most level-maxDepth cells hold more than maxNumPts
    points, so p2oLocal's saturated-cell path runs all over the domain."""
    pts = (dgx.gen_points(dist, seed, n, n_gauss=3, lo=0.3, hi=0.7, sig_lo=0.002) if dist == 2
           else dgx.gen_points(dist, seed, n))
    check_pipeline(dgx, oracle, pts, md, m, *stencils)


@pytest.mark.parametrize("md,m", [(4, 1), (8, 2)])
def test_stale_cursor_at_domain_end_is_an_error_in_both(dgx, oracle, md, m):
    """Clustered points at a shallow maxDepth leave the p2oLocal cursor stale
    at the end of the domain, where the reference compares against the getNext
    sentinel; the oracle and the library both report it."""
    pts = dgx.gen_points(2, 11, 20000, n_gauss=3, lo=0.3, hi=0.7, sig_lo=0.002)
    with pytest.raises(ValueError):
        oracle.points2octree(pts, md, m)
    with pytest.raises(dgx.DgxError):
        dgx.points2octree(pts, md, m)


def test_domain_lengths(dgx, oracle, stencils):
    """points2Octree over [0, gLens): px = (uint)(x 2^D / gLen).  Scaling by
    2^D is exact, so this equals the unit-cube quantisation of x / gLen."""
    g = np.array([2.0, 1.0, 0.5])
    pts = dgx.gen_points(1, 5, 20000) * g
    ou = dgx.points2octree(pts, 30, 4, glens=g)
    ru = oracle.points2octree(pts / g, 30, 4)
    assert np.array_equal(ou.treenodes(), ru)
    ob = dgx.balance_octree(ou)
    assert np.array_equal(ob.treenodes(), oracle.balance(ru))
    with pytest.raises(dgx.DgxError):  # a point at x = gLen is outside
        dgx.points2octree(np.vstack([pts, [[2.0, 0.5, 0.25]]]), 30, 4, glens=g)


def test_degenerate_inputs(dgx, oracle):
    pts = np.array([[0.25, 0.5, 0.75]])
    assert dgx.points2octree(pts, 30, 1).treenodes().tolist() == [[0, 0, 0, 0, 1, 3, 30]]
    pts = dgx.gen_points(0, 3, 8)
    assert len(dgx.points2octree(pts, 30, 8)) == 1
    assert np.array_equal(dgx.points2octree(pts, 30, 1).treenodes(), oracle.points2octree(pts, 30, 1))
    with pytest.raises(dgx.DgxError):
        dgx.points2octree(np.array([[1.0, 0.5, 0.5], [0.1, 0.1, 0.1]]), 30, 1)  # outside [0,1)
    with pytest.raises(dgx.DgxError):
        dgx.points2octree(pts, 30, 0)
    with pytest.raises(dgx.DgxError):  # DA needs more than 8 octants
        dgx.DA(dgx.points2octree(pts[:2], 30, 1))


def test_saturated_cell_replay_hand_traced(dgx, oracle):
    pts = (np.array([[2, 2, 1]] * 3 + [[3, 0, 0]] * 2) + 0.5) / 4
    o = dgx.points2octree(pts, 2, 1)
    assert np.array_equal(o.treenodes(), oracle.points2octree(pts, 2, 1))
    st = o.stats()
    assert st[1] >= 1 and st[2] >= 1  # a trigger was found and replayed


def test_saturated_cell_replay_hand_traced_morton(dgx, oracle):
    """The Morton-curve hand trace of the oracle tests: O1 stays unrefined."""
    pts = (np.array([[1, 1, 1]] * 3 + [[2, 0, 0], [3, 0, 0]]) + 0.5) / 4
    o = dgx.points2octree(pts, 2, 1, 1)
    got = o.treenodes()[:, :4].tolist()
    assert got == [[0, 0, 0, 2], [1, 0, 0, 2], [0, 1, 0, 2], [1, 1, 0, 2],
                   [0, 0, 1, 2], [1, 0, 1, 2], [0, 1, 1, 2], [1, 1, 1, 2],
                   [2, 0, 0, 1], [0, 2, 0, 1], [2, 2, 0, 1], [0, 0, 2, 1], [2, 0, 2, 1], [0, 2, 2, 1], [2, 2, 2, 1]]
    oracle.set_ordering(False)
    try:
        assert np.array_equal(o.treenodes(), oracle.points2octree(pts, 2, 1))
    finally:
        oracle.set_ordering(True)


def test_saturated_cells_random(dgx, oracle):
    """Clusters of identical points at maxDepth: the replay of the reference's
    cursor must reproduce the oracle exactly (or both refuse the sentinel case)."""
    rng = np.random.default_rng(9)
    ok = refused = 0
    for trial in range(60):
        md = int(rng.integers(2, 5))
        m = int(rng.integers(1, 4))
        k = int(rng.integers(2, 8))
        cells = rng.integers(0, 1 << md, (k, 3))
        reps = rng.integers(1, 3 * m + 2, k)
        pts = (np.repeat(cells, reps, axis=0) + 0.5) / (1 << md)
        pts = np.vstack([pts, rng.random((int(rng.integers(0, 40)), 3))])
        try:
            ref = oracle.points2octree(pts, md, m)
        except ValueError:
            with pytest.raises(dgx.DgxError):
                dgx.points2octree(pts, md, m)
            refused += 1
            continue
        got = dgx.points2octree(pts, md, m).treenodes()
        assert np.array_equal(got, ref), (trial, md, m)
        ok += 1
    assert ok > 20


def test_lognormal_corner_cluster(dgx, oracle, stencils):
    """Config-2 distribution: ~2% of coordinates clip to 1-eps, stacking points
    in the last cell at maxDepth; maxNumPts=8."""
    pts = dgx.workload_points("lognormal-4M", 400000)
    assert (pts == np.nextafter(1.0, 0)).all(1).sum() > 0
    ou, ob, da = check_pipeline(dgx, oracle, pts, 30, 8, *stencils)
    print("stats", ou.stats(), flush=True)


@pytest.mark.parametrize("layout,curve,m", [("long", 0, 1), ("long", 1, 1), ("long", 0, 3), ("short", 0, 1),
                                             ("short", 1, 2)])
def test_points_sharing_level21_cells(dgx, oracle, layout, curve, m):
    """Points that share level-21 cells (2^-21 wide), in shuffled order: their
    keys agree in the high word and differ only in levels 22-30.  "short": 400
    cells with 4 points each, which points2Octree orders run by run after the
    high-word sort; "long": 3000 points in a 3e-7 cube, runs too long for that,
    so the keys are sorted again on the full key."""
    rng = np.random.default_rng(21)
    uni = rng.random((3000, 3))
    if layout == "long":
        c = rng.random(3) * 0.9
        tight = c + rng.random((3000, 3)) * 3e-7
    else:
        corner = np.floor(rng.random((400, 3)) * 2.0 ** 21) / 2.0 ** 21
        tight = np.repeat(corner, 4, axis=0) + rng.random((1600, 3)) * 4e-7
    dup = np.repeat(tight[:5], 3, axis=0)
    pts = np.vstack([uni, tight, dup])
    pts = pts[rng.permutation(len(pts))]
    cells = np.floor(pts * 2.0 ** 21)
    assert len(np.unique(cells, axis=0)) < len(pts) - 15
    check_pipeline(dgx, oracle, pts, 30, m, curve=curve)


def test_sample_sort(dgx, oracle):
    pts = dgx.gen_points(0, 5, 4000)
    t = oracle.points2octree(pts, 30, 1)
    perm = np.random.default_rng(0).permutation(len(t))
    assert np.array_equal(dgx.sample_sort(t[perm]), t)


def random_octants(rng, m, lmin, lmax, md=30):
    lev = rng.integers(lmin, lmax + 1, m).astype(np.uint32)
    xyz = rng.integers(0, 1 << md, (m, 3), dtype=np.uint64).astype(np.uint32)
    xyz &= ~((np.uint32(1) << (md - lev)) - np.uint32(1))[:, None]
    tn = np.zeros((m, 7), np.uint32)
    tn[:, :3], tn[:, 3], tn[:, 4], tn[:, 5], tn[:, 6] = xyz, lev, 1, 3, md
    return tn


@pytest.mark.parametrize("case", ["random", "duplicates", "one_subtree", "coarse"])
def test_sort_many_tiles(dgx, oracle, case):
    """The device radix sort over hundreds of tiles against the oracle's sort:
    octants of all levels (every digit position varies), each octant repeated
    many times (long runs of equal keys across tiles), octants inside one
    level-7 subtree (the leading digit positions are constant and skipped), and
    coarse octants (the trailing positions are constant)."""
    rng = np.random.default_rng(7)
    if case == "random":
        tn = random_octants(rng, 700001, 0, 30)
    elif case == "duplicates":
        tn = np.repeat(random_octants(rng, 40000, 3, 25), 23, axis=0)
    elif case == "one_subtree":
        tn = random_octants(rng, 500000, 8, 30)
        tn[:, :3] = (tn[:, :3] & ((1 << 23) - 1)) | (np.uint32(5) << 23)
    else:
        tn = random_octants(rng, 300000, 0, 7)
    tn = tn[rng.permutation(len(tn))]
    want = oracle.sort(tn)
    print(case, len(tn), flush=True)
    assert np.array_equal(dgx.sample_sort(tn), want)


def test_uniform_256k_full_parity(dgx, oracle, stencils):
    """Config 1 at its full size: 262,144 uniform points, maxNumPts=1."""
    pts = dgx.workload_points("uniform-256K")
    check_pipeline(dgx, oracle, pts, 30, 1, *stencils)


@pytest.mark.timeout(600)
def test_lognormal_4m_full_parity(dgx, oracle, stencils):
    """Config 2 at its full size (4,194,304 log-normal points, maxNumPts=8):
    the octree, the balanced octree and the DA maps (nlist, hanging masks,
    oct levels, node compaction) bit-exact against the oracle, and 100 FP64
    -lap+I matvecs on independent N(0,1) vectors within 1e-12 of the oracle's
    element loop with the reference's type-2 tables (odaJac.C:275-292)."""
    L, M = stencils
    pts = dgx.workload_points("lognormal-4M")
    ou = dgx.points2octree(pts, 30, 8)
    ob = dgx.balance_octree(ou)
    da = dgx.DA(ob)
    tu, tb = ou.treenodes(), ob.treenodes()
    nl, mk, ol, l2n = da.maps()
    print("lognormal-4M: GPU octree, balance and DA built", len(tu), len(tb), da.n_nodes, flush=True)
    ru = oracle.points2octree(pts, 30, 8)
    assert np.array_equal(tu, ru), "points2Octree"
    print("points2Octree bit-exact", flush=True)
    rb = oracle.balance(ru)
    assert np.array_equal(tb, rb), "balanceOctree"
    print("balanceOctree bit-exact", flush=True)
    rda = oracle.DA(rb)
    assert da.element_size == rda.n_elem and da.local_buffer_size == rda.n_local
    assert da.node_size == rda.node_size and da.boundary_node_size == rda.bnd_node_size
    assert np.array_equal(nl, rda.nlist), "DA nlist"
    assert np.array_equal(mk, rda.mask), "DA hanging masks"
    assert np.array_equal(ol, rda.oct_levels), "DA oct levels"
    assert np.array_equal(l2n, rda.local_to_node), "DA node compaction"
    print("DA maps bit-exact", flush=True)
    rng = np.random.default_rng(2)
    worst = 0.0
    for k in range(100):
        x = rng.standard_normal(da.n_nodes)
        worst = max(worst, rel_err(da.matvec_host(x), rda.matvec(x, L, M)))
        if k % 25 == 24:
            print(f"{k + 1} matvecs, worst relative error {worst:.2e}", flush=True)
    assert worst < MATVEC_RTOL
    one = np.ones(da.n_nodes)
    assert abs(da.matvec_host(one, 0.0, 1.0).sum() - 1.0) < 1e-12
    assert np.abs(da.matvec_host(one, 1.0, 0.0)).max() < 1e-12
