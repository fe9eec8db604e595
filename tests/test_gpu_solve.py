"""GPU tests of the Dirichlet operator and the multigrid-preconditioned CG
solve.  The reference operator is assembled on the CPU from the oracle's DA
and the reference's type-2 stencil tables (Jacobian1MatMult element loop,
odaJac.C:275-292; Dirichlet rows/columns as in omgJac.C:958-1057) and solved
with scipy's sparse direct solver."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def assemble(rda, L, M, lap=1.0, mass=1.0):
    conn, cn, lev, mk = rda.element_view()
    from oracle import pyoracle as O
    tmap = O.etype_masks()
    E = rda.n_elem
    t = np.zeros(E, np.int64)
    hang = np.nonzero(mk)[0]
    for e in hang:
        t[e] = int(np.nonzero(tmap[cn[e]] == mk[e])[0][0])
    h = 2.0 ** -(lev.astype(np.float64) - 1)
    Ae = (lap * h / 2)[:, None, None] * L[cn, t] + (mass * h ** 3 / 8)[:, None, None] * M[cn, t]
    rows = np.repeat(conn[:, :, None], 8, axis=2).ravel()
    cols = np.repeat(conn[:, None, :], 8, axis=1).ravel()
    A = sp.coo_matrix((Ae.ravel(), (rows, cols)), shape=(rda.n_nodes, rda.n_nodes)).tocsr()
    return A


def boundary_nodes(rda):
    l2n = rda.local_to_node
    idx = np.nonzero(l2n != 0xFFFFFFFF)[0]
    pos = rda.octs[idx, :3].astype(np.int64)
    D = 1 << 30
    bd = np.zeros(rda.n_nodes, bool)
    bd[l2n[idx]] = ((pos == 0) | (pos == D)).any(1)
    return bd


def dirichlet(A, bd):
    d = sp.diags((~bd).astype(np.float64))
    return (d @ A @ d + sp.diags(bd.astype(np.float64))).tocsr()


@pytest.fixture(scope="module")
def problem(dgx, oracle, stencils):
    L, M = stencils
    pts = dgx.gen_points(1, 21, 30000)
    ou = dgx.points2octree(pts, 30, 8)
    ob = dgx.balance_octree(ou)
    rda = oracle.DA(ob.treenodes())
    A = assemble(rda, L, M)
    bd = boundary_nodes(rda)
    return ob, rda, A, bd


def test_dirichlet_matvec_parity(dgx, problem):
    ob, rda, A, bd = problem
    da = dgx.DA(ob)
    x = np.random.default_rng(0).standard_normal(da.n_nodes)
    AD = dirichlet(A, bd)
    y = da.matvec_dirichlet_host(x)
    yr = AD @ x
    assert np.abs(y - yr).max() / np.abs(yr).max() < 1e-12
    yn = da.matvec_host(x)
    assert np.abs(yn - A @ x).max() / np.abs(A @ x).max() < 1e-12


def test_dropped_hierarchy_returns_its_memory(dgx, problem):
    """An MG the caller drops is freed at once (no reference cycle keeps it):
    the library's live bytes go back to where they were, and a hierarchy
    built again takes about the same bytes."""
    ob = problem[0]
    dgx.synchronize()
    base = dgx.memory_info()["library_live"]
    mg = dgx.MG(ob, 5)
    _ = mg.da.n_nodes  # the borrowed finest DA must not keep the MG alive
    held = dgx.memory_info()["library_live"] - base
    assert held > 0
    mg = None
    dgx.synchronize()
    assert dgx.memory_info()["library_live"] == base
    mg = dgx.MG(ob, 5)
    # the same build again (blocks taken from the cache may be up to a quarter larger than asked)
    again = dgx.memory_info()["library_live"] - base
    assert held * 0.9 <= again <= held * 1.25
    mg = None
    dgx.synchronize()
    assert dgx.memory_info()["library_live"] == base


def test_trim_cache_empties_the_block_cache(dgx, problem):
    """dgx_trim_cache returns the cached blocks and the pool's unused memory:
    the cache is empty afterwards, the device holds no more than before, and
    the library allocates and solves as before."""
    ob = problem[0]
    mg = dgx.MG(ob, 3)
    mg = None
    dgx.synchronize()
    before = dgx.memory_info()
    assert before["cached"] > 0
    dgx.trim_cache()
    after = dgx.memory_info()
    assert after["cached"] == 0
    assert after["device_used"] <= before["device_used"]
    assert after["library_live"] == before["library_live"]
    mg = dgx.MG(ob, 3)
    n = mg.n_nodes
    df = dgx.DeviceArray.from_host(np.random.default_rng(5).standard_normal(n))
    db = dgx.DeviceArray((n,), np.float64)
    du = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    st = mg.solve(db, du, fp32=False)
    assert st["rel_residual"] < 1e-8
    mg = None
    dgx.synchronize()


@pytest.mark.parametrize("nlevels", [1, 3, 5])
def test_mg_solve_matches_direct_solve(dgx, problem, nlevels):
    ob, rda, A, bd = problem
    mg = dgx.MG(ob, nlevels)
    print("levels", mg.nlevels, "octants", mg.level_octants, "nodes", mg.level_nodes, flush=True)
    assert mg.level_nodes[0] == rda.n_nodes
    assert all(a > b for a, b in zip(mg.level_octants, mg.level_octants[1:]))
    n = mg.n_nodes
    f = np.random.default_rng(1).standard_normal(n)
    df = dgx.DeviceArray.from_host(f)
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    b = db.to_host()
    M = assemble(rda, *[np.load("tests/golden/type2_stencils.npz")[k] for k in ("lap", "mass")], lap=0.0, mass=1.0)
    br = M @ f
    br[bd] = 0.0
    assert np.abs(b - br).max() / np.abs(br).max() < 1e-12
    AD = dirichlet(A, bd)
    uref = spla.spsolve(AD.tocsc(), br)
    res = {}
    for fp32 in (False, True):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=fp32, pre=2, post=2, omega=0.8, rtol=1e-8, maxit=300)
        u = du.to_host()
        true_res = np.linalg.norm(br - AD @ u) / np.linalg.norm(br)
        err = np.linalg.norm(u - uref) / np.linalg.norm(uref)
        print(f"nlevels={mg.nlevels} fp32={fp32} it={st['iterations']} relres={st['rel_residual']:.2e} "
              f"true={true_res:.2e} err={err:.2e} t={st['seconds']*1e3:.1f}ms", flush=True)
        assert st["rel_residual"] <= 1e-8 and true_res <= 1.5e-8
        assert err <= 1e-7
        res[fp32] = u
    assert np.linalg.norm(res[True] - res[False]) / np.linalg.norm(res[False]) <= 1e-7


def test_graph_replay_matches_eager(dgx, problem):
    """The captured CG step and V-cycle replay the same kernels in the same
    order as the eager path, and every sum in them runs in a fixed order (the
    element loop's node sums are fixed-point integers), so every replay gives
    the eager solve's bits.  A second hierarchy built from the same octree
    (level DAs, diagonals, transfers, coarse factorisation) gives them too."""
    ob, rda, A, bd = problem
    mg = dgx.MG(ob, 4)
    n = mg.n_nodes
    df = dgx.DeviceArray.from_host(np.random.default_rng(3).standard_normal(n))
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    mg2 = dgx.MG(ob, 4)
    for fp32 in (True, False):
        du = dgx.DeviceArray((n,), np.float64)
        st0 = mg.solve(db, du, fp32=fp32, graph=False)
        u0 = du.to_host()
        for rep in range(3):
            du.zero()
            st = mg.solve(db, du, fp32=fp32, graph=True)
            assert st["iterations"] == st0["iterations"]
            assert np.array_equal(du.to_host(), u0), (fp32, rep)
            assert st["rel_residual"] <= 1e-8
        du.zero()
        st = mg2.solve(db, du, fp32=fp32, graph=False)
        assert st["iterations"] == st0["iterations"] and np.array_equal(du.to_host(), u0), fp32
    mg2.free()


def test_operators_are_deterministic(dgx, problem):
    """Node sums are fixed-point integers per element block, added in any order
    with the same result: the FP64, FP32 and Dirichlet matvecs and the
    diagonal give the same bits on every call and on a second DA built from
    the same octree.  A non-finite x value makes the rows of its blocks NaN
    and leaves every other row as it was."""
    ob, rda, A, bd = problem
    da1, da2 = dgx.DA(ob), dgx.DA(ob)
    n = da1.n_nodes
    x = np.random.default_rng(8).standard_normal(n)
    y0 = da1.matvec_host(x)
    for rep in range(4):
        assert np.array_equal(da1.matvec_host(x), y0), rep
    assert np.array_equal(da2.matvec_host(x), y0)
    assert np.array_equal(da1.matvec_dirichlet_host(x), da2.matvec_dirichlet_host(x))
    assert np.array_equal(da1.diagonal_host(), da2.diagonal_host())
    dx = dgx.DeviceArray.from_host(x.astype(np.float32))
    dy = dgx.DeviceArray((n,), np.float32)
    da1.matvec(dx, dy)
    y32 = dy.to_host()
    da2.matvec(dx, dy)
    assert np.array_equal(dy.to_host(), y32)
    i = int(np.nonzero(~bd)[0][n // 7 % max(1, int((~bd).sum()))])
    for bad in (np.nan, np.inf):
        xb = x.copy()
        xb[i] = bad
        yb = da1.matvec_host(xb)
        nan = np.isnan(yb)
        assert nan[i] and 8 <= nan.sum() <= 8 * 1024, nan.sum()
        assert np.array_equal(yb[~nan], y0[~nan])


def test_fp32_matvec_matches_fp64_to_single_precision(dgx, problem):
    """The FP32 operator of the V-cycle (FP32 LDS slots and accumulation)
    against the FP64 operator on the same DA, both plain and Dirichlet."""
    ob, rda, A, bd = problem
    da = dgx.DA(ob)
    x = np.random.default_rng(2).standard_normal(da.n_nodes)
    y64 = da.matvec_host(x)
    dx = dgx.DeviceArray.from_host(x.astype(np.float32))
    dy = dgx.DeviceArray((da.n_nodes,), np.float32)
    da.matvec(dx, dy)
    y32 = dy.to_host().astype(np.float64)
    assert np.abs(y32 - y64).max() / np.abs(y64).max() < 1e-5
    assert np.abs(y64 - A @ x).max() / np.abs(A @ x).max() < 1e-12


def solve_against_direct(dgx, ob, rda, bd, L, M, lap, mass, nlevels=4):
    mg = dgx.MG(ob, nlevels, lap_fac=lap, mass_fac=mass)
    n = mg.n_nodes
    f = np.random.default_rng(5).standard_normal(n)
    df = dgx.DeviceArray.from_host(f)
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    br = db.to_host()
    AD = dirichlet(assemble(rda, L, M, lap=lap, mass=mass), bd)
    uref = spla.spsolve(AD.tocsc(), br)
    for fp32 in (False, True):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=fp32, rtol=1e-8, maxit=300)
        u = du.to_host()
        err = np.linalg.norm(u - uref) / np.linalg.norm(uref)
        true_res = np.linalg.norm(br - AD @ u) / np.linalg.norm(br)
        print(f"lap={lap} mass={mass} fp32={fp32} it={st['iterations']} err={err:.2e} res={true_res:.2e}", flush=True)
        assert st["rel_residual"] <= 1e-8 and true_res <= 1.5e-8 and err <= 1e-7


@pytest.mark.parametrize("lap,mass", [(2.0, 0.5), (1.0, 0.0), (0.25, 4.0)])
def test_mg_solve_operator_factors(dgx, problem, stencils, lap, mass):
    """DAMGSetKSP's lapFac / massFac: -lapFac lap + massFac I, including the
    pure Dirichlet Laplacian (massFac 0)."""
    ob, rda, A, bd = problem
    solve_against_direct(dgx, ob, rda, bd, *stencils, lap, mass)


def test_matvec_path_for_large_blocks(dgx, problem, stencils, monkeypatch):
    """The matvec path taken by DAs whose element blocks touch more than 1024
    nodes (one workgroup per block, then separate smoother / residual
    updates), forced on this DA: matvecs and MG solves against the references."""
    monkeypatch.setenv("DGX_MATVEC_BLOCKED", "1")
    ob, rda, A, bd = problem
    da = dgx.DA(ob)
    x = np.random.default_rng(0).standard_normal(da.n_nodes)
    AD = dirichlet(A, bd)
    yr = AD @ x
    assert np.abs(da.matvec_dirichlet_host(x) - yr).max() / np.abs(yr).max() < 1e-12
    assert np.abs(da.matvec_host(x) - A @ x).max() / np.abs(A @ x).max() < 1e-12
    solve_against_direct(dgx, ob, rda, bd, *stencils, 1.0, 1.0)


def test_coarse_direct_solve_is_exact(dgx, problem):
    """One level: the V-cycle is the sparse direct solve of the whole Dirichlet
    matrix (nested dissection, many fronts), so preconditioned CG converges at
    once, to the direct solution.  Repeated solves give the same bits: the
    direct solve and the matvec sum in a fixed order."""
    ob, rda, A, bd = problem
    mg = dgx.MG(ob, 1)
    assert mg.nlevels == 1 and mg.coarse["fronts"] > 8 and mg.coarse["tree_levels"] >= 3, mg.coarse
    n = mg.n_nodes
    f = np.random.default_rng(6).standard_normal(n)
    df = dgx.DeviceArray.from_host(f)
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    br = db.to_host()
    AD = dirichlet(A, bd)
    uref = spla.spsolve(AD.tocsc(), br)
    outs = []
    for rep in range(2):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=False, rtol=1e-12, maxit=5)
        u = du.to_host()
        assert st["iterations"] <= 2 and st["rel_residual"] <= 1e-12, st
        assert np.linalg.norm(u - uref) / np.linalg.norm(uref) <= 1e-11
        outs.append(u)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.timeout(300)
def test_gauss4_5_levels_match_direct_solve(dgx, oracle, stencils):
    """Config 3's problem at a size scipy solves directly: 100,000 points of the
    4-Gaussian cloud (sigma 0.02, seed 2), maxNumPts 4, 5-level V(2,2) Jacobi
    (0.8) PCG to 1e-8; FP64 and FP32 V-cycles within 1e-7 of the direct solve
    of the operator assembled from the oracle's DA and the reference tables.
    The hierarchy has exactly the 5 requested levels (blockDiag.cpp:1384-1425)."""
    L, M = stencils
    pts = dgx.workload_points("gauss4-16M", 100000)
    ob = dgx.balance_octree(dgx.points2octree(pts, 30, 4))
    rda = oracle.DA(ob.treenodes())
    A = assemble(rda, L, M)
    bd = boundary_nodes(rda)
    AD = dirichlet(A, bd)
    mg = dgx.MG(ob, 5)
    print("levels", mg.nlevels, mg.level_octants, mg.level_nodes, mg.coarse, flush=True)
    assert mg.nlevels == 5
    n = mg.n_nodes
    f = np.random.default_rng(7).standard_normal(n)
    df = dgx.DeviceArray.from_host(f)
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    br = db.to_host()
    lu = spla.splu(AD.tocsc(), permc_spec="MMD_AT_PLUS_A", options=dict(SymmetricMode=True))
    uref = lu.solve(br)
    print("direct solve done", flush=True)
    res = {}
    for fp32 in (False, True):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=fp32, pre=2, post=2, omega=0.8, rtol=1e-8, maxit=300)
        u = du.to_host()
        true_res = np.linalg.norm(br - AD @ u) / np.linalg.norm(br)
        err = np.linalg.norm(u - uref) / np.linalg.norm(uref)
        print(f"fp32={fp32} it={st['iterations']} res={true_res:.2e} err={err:.2e}", flush=True)
        assert st["rel_residual"] <= 1e-8 and true_res <= 1.5e-8 and err <= 1e-7
        res[fp32] = u
    assert np.linalg.norm(res[True] - res[False]) / np.linalg.norm(res[False]) <= 1e-7


@pytest.mark.timeout(600)
def test_gauss4_16m_full_size_solve(dgx):
    """Config 3 at its full size (16,777,216 points, 4 Gaussians, maxNumPts 4):
    5-level hierarchy as requested; the FP64 solve reaches 1e-8 in the FP64
    residual recomputed with the Dirichlet operator, and the FP32 V-cycle
    solution agrees with it to 1e-7."""
    pts = dgx.workload_points("gauss4-16M")
    ob = dgx.balance_octree(dgx.points2octree(pts, 30, 4))
    del pts
    mg = dgx.MG(ob, 5)
    print("levels", mg.nlevels, mg.level_octants, mg.level_nodes, mg.coarse, flush=True)
    assert mg.nlevels == 5
    n = mg.n_nodes
    df = dgx.DeviceArray.from_host(np.random.default_rng(2).standard_normal(n))
    db = dgx.DeviceArray((n,), np.float64)
    mg.rhs(df, db)
    b = db.to_host()
    u = {}
    for fp32 in (False, True):
        du = dgx.DeviceArray((n,), np.float64)
        st = mg.solve(db, du, fp32=fp32, rtol=1e-8, maxit=100)
        u[fp32] = du.to_host()
        r = b - mg.da.matvec_dirichlet_host(u[fp32])
        rel = np.linalg.norm(r) / np.linalg.norm(b)
        print(f"fp32={fp32} it={st['iterations']} res={rel:.2e}", flush=True)
        assert st["rel_residual"] <= 1e-8 and rel <= 1.5e-8
    assert np.linalg.norm(u[True] - u[False]) / np.linalg.norm(u[False]) <= 1e-7
