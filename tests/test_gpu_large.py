"""Config 5 at its full size on the GPU: 134,217,728 points from 16 Gaussians
(sigma log-uniform in [0.005, 0.05], centres U[0.1,0.9]^3, Dirichlet(1)
weights, seed 4), maxDepth 30, maxNumPts 4.  The octree is too large for the
oracle to rebuild here, so it is checked by the reference's own invariants
(testUtils.C isLinear / isComplete on every octant, isBalanced on a random
sample of a million octants), the DA by conservation (the mass of the unit
cube, K 1 = 0), and the 7-level MG solves by their FP64 residual and the FP32
V-cycle's agreement with the FP64 solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_gauss16_128m_full_size(dgx, oracle):
    pts = dgx.workload_points("gauss16-128M")
    print("points generated", len(pts), flush=True)
    ou = dgx.points2octree(pts, 30, 4)
    del pts
    ob = dgx.balance_octree(ou)
    del ou
    tb = ob.treenodes()
    print("balanced octants", len(tb), flush=True)
    assert oracle.is_linear(tb) and oracle.is_complete(tb)
    sample = np.random.default_rng(0).choice(len(tb), size=min(len(tb), 1000000), replace=False)
    assert oracle.is_balanced_at(tb, np.sort(sample))
    del tb
    print("octree invariants hold", flush=True)
    mg = dgx.MG(ob, 7)
    print("levels", mg.nlevels, mg.level_octants, mg.level_nodes, mg.coarse, flush=True)
    assert mg.nlevels == 7
    da = mg.da
    one = np.ones(da.n_nodes)
    assert abs(da.matvec_host(one, 0.0, 1.0).sum() - 1.0) < 1e-10
    assert np.abs(da.matvec_host(one, 1.0, 0.0)).max() < 1e-12
    n = mg.n_nodes
    df = dgx.DeviceArray.from_host(np.random.default_rng(4).standard_normal(n))
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    b = db.to_host()
    u = {}
    for fp32 in (False, True):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=fp32, rtol=1e-8, maxit=100)
        u[fp32] = du.to_host()
        rel = np.linalg.norm(b - da.matvec_dirichlet_host(u[fp32])) / np.linalg.norm(b)
        print(f"fp32={fp32} it={st['iterations']} res={rel:.2e}", flush=True)
        assert st["rel_residual"] <= 1e-8 and rel <= 1.5e-8
    assert np.linalg.norm(u[True] - u[False]) / np.linalg.norm(u[False]) <= 1e-7
