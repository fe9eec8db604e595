"""GPU tests of the compressed element-to-node maps in the reference's form
(compressLut; odaBuildNlist.cpp:3120-3262 encode, oda.tcc:1304-3950 decode):
the device encoding equals the oracle's byte for byte (remainders, quotient
masks, sort orders, quotients) on both curves; the device decode equals the
oracle's decode, and with Morton order it reproduces the DA's node list."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("curve", [0, 1])
@pytest.mark.parametrize("dist,n,m", [(1, 50000, 8), (2, 30000, 4), (0, 20000, 1)])
def test_compressed_maps_match_oracle(dgx, oracle, curve, dist, n, m):
    kw = dict(n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.02) if dist == 2 else {}
    pts = dgx.gen_points(dist, 13, n, **kw)
    oracle.set_ordering(curve == 0)
    try:
        ob = dgx.balance_octree(dgx.points2octree(pts, 30, m, curve))
        da = dgx.DA(ob)
        rda = oracle.DA(ob.treenodes())
        nl = da.maps()[0]
        assert np.array_equal(nl, rda.nlist)
        rem, qm, order, quot = da.compressed_maps()
        rr, rq, ro, rquot = rda.compress()
        assert np.array_equal(rem, rr) and np.array_equal(qm, rq) and np.array_equal(order, ro)
        assert np.array_equal(quot, rquot)
        dec = da.decode_maps(rem, qm, order, quot)
        assert np.array_equal(dec, rda.decode(rr, rq, ro, rquot))
        if curve == 1:
            assert np.array_equal(dec, nl)
        nb = 10 * da.element_size + 2 * len(quot)
        print(f"curve={curve} n={n} elements={da.element_size} compressed {nb / da.element_size:.2f} B/element "
              f"(nlist 32), decode differs from nlist at {(dec != nl).any(1).sum()} elements", flush=True)
    finally:
        oracle.set_ordering(True)


def test_compressed_maps_full_gauss4_16m(dgx):
    """Config 3's octree (16,777,216 points, 4 Gaussians) on the Morton curve:
    bytes of the compressed maps per element, and the decode reproduces the
    node list.  (The log-normal cloud of config 2 cannot be built on the Morton
    curve: its clipped corner sends the reference's p2oLocal cursor past the
    end of the domain, which this build reports as an error.)"""
    pts = dgx.workload_points("gauss4-16M")
    da = dgx.DA(dgx.balance_octree(dgx.points2octree(pts, 30, 4, 1)))
    del pts
    nl = da.maps()[0]
    rem, qm, order, quot = da.compressed_maps()
    dec = da.decode_maps(rem, qm, order, quot)
    assert np.array_equal(dec, nl)
    nb = 10 * da.element_size + 2 * len(quot)
    print(f"gauss4-16M (Morton): {da.element_size} elements, {nb / da.element_size:.2f} B/element", flush=True)
    assert nb / da.element_size < 14.0
