"""GPU tests of the request server (dgx_server_*): each request's result is
the same alone and in a mix of concurrent requests (bit for bit: every sum
in the build and the solve runs in a fixed order), resident problems are
reused, the HBM budget gates admission, and concurrent requests overlap on
the device."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_requests(dgx, count, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(2 ** rng.uniform(12, 15))
        req = {"n": n, "dist": int(rng.integers(0, 3)), "max_num_pts": int(rng.choice([1, 4, 8])),
               "seed": 1000 + i}
        out.append((req, dgx.serve_points(req)))
    return out


def run_all(dgx, reqs, slots, fp32=True):
    sv = dgx.Server(slots=slots, hbm_budget=64e9, nlevels=4, fp32=fp32, keep_resident=False, return_solution=True)
    ids = [sv.submit(c, 30, r["max_num_pts"], r["seed"]) for r, c in reqs]
    res = [sv.wait(i, solution=True) for i in ids]
    sv.close()
    return res


def test_result_is_the_same_alone_and_in_the_mix(dgx):
    reqs = small_requests(dgx, 12)
    mixed = run_all(dgx, reqs, slots=4)
    alone = []
    for rc in reqs:  # one at a time, nothing else on the device
        alone += run_all(dgx, [rc], slots=1)
    for a, m in zip(alone, mixed):
        assert a["status"] == 0 and m["status"] == 0
        assert a["nodes"] == m["nodes"] and a["octants"] == m["octants"] and a["levels"] == m["levels"]
        assert a["rel_residual"] <= 1e-8 and m["rel_residual"] <= 1e-8
        assert a["iterations"] == m["iterations"]
        assert np.array_equal(a["x"], m["x"])


def test_resident_problem_is_reused(dgx):
    (req, pts), = small_requests(dgx, 1, seed=5)
    sv = dgx.Server(slots=2, hbm_budget=64e9, nlevels=4, keep_resident=True, return_solution=True)
    a = sv.wait(sv.submit(pts, 30, req["max_num_pts"], 7), solution=True)
    b = sv.wait(sv.submit(pts, 30, req["max_num_pts"], 7), solution=True)
    c = sv.wait(sv.submit(pts, 30, req["max_num_pts"], 8), solution=True)
    sv.close()
    assert a["reused"] == 0 and b["reused"] == 1 and c["reused"] == 1
    assert np.array_equal(a["x"], b["x"])  # eager first solve, graph replay after
    assert np.linalg.norm(a["x"] - c["x"]) > 1e-3 * np.linalg.norm(a["x"])  # another right-hand side


def test_budget_gates_admission(dgx):
    reqs = small_requests(dgx, 6, seed=3)
    # every request's estimate (>= 64 MiB + 3000 B per point) exceeds this
    # budget, so a request is admitted only when nothing else holds memory
    budget = (64 << 20) + min(r["n"] for r, _ in reqs) * 3000
    sv = dgx.Server(slots=4, hbm_budget=budget, nlevels=4, keep_resident=False)
    ids = [sv.submit(c, 30, r["max_num_pts"], r["seed"]) for r, c in reqs]
    res = sorted((sv.wait(i) for i in ids), key=lambda x: x["t_start"])
    sv.close()
    for a, b in zip(res, res[1:]):  # one request at a time fits the budget
        assert b["t_start"] >= a["t_done"] - 1e-6


def test_concurrent_requests_overlap(dgx):
    reqs = small_requests(dgx, 8, seed=4)
    sv = dgx.Server(slots=4, hbm_budget=64e9, nlevels=4, keep_resident=False)
    ids = [sv.submit(c, 30, r["max_num_pts"], r["seed"]) for r, c in reqs]
    res = [sv.wait(i) for i in ids]
    sv.close()
    overlaps = sum(1 for i in range(len(res)) for j in range(i + 1, len(res))
                   if res[i]["t_start"] < res[j]["t_done"] and res[j]["t_start"] < res[i]["t_done"])
    assert overlaps >= 1
    assert all(r["status"] == 0 and r["rel_residual"] <= 1e-8 for r in res)


def test_bad_requests_fail_loudly(dgx):
    sv = dgx.Server(slots=1, hbm_budget=8e9, nlevels=3)
    pts = np.full((100, 3), 1.5)  # outside the unit cube
    rid = sv.submit(pts, 30, 1, 0)
    with pytest.raises(dgx.DgxError):
        sv.wait(rid)
    sv.close()


def test_wait_twice_is_an_error(dgx):
    (req, pts), = small_requests(dgx, 1, seed=21)
    sv = dgx.Server(slots=1, hbm_budget=8e9, nlevels=3, keep_resident=False)
    rid = sv.submit(pts, 30, req["max_num_pts"], 1)
    assert sv.wait(rid)["status"] == 0
    with pytest.raises(dgx.DgxError, match="already waited"):
        sv.wait(rid)
    with pytest.raises(dgx.DgxError, match="already waited"):
        sv.wait(rid, solution=True)
    sv.close()


def test_failed_resident_resolve_releases_its_entry(dgx):
    """A resident problem whose re-solve fails is dropped and its bytes go
    back to the budget; a later request that needs that room is admitted."""
    (req, pts), (req2, pts2) = small_requests(dgx, 2, seed=22)
    sv = dgx.Server(slots=2, hbm_budget=8e9, nlevels=3, keep_resident=True)
    a = sv.wait(sv.submit(pts, 30, req["max_num_pts"], 1))
    st = sv.stats()
    assert st["resident"] == 1 and st["reserved"] == st["resident_bytes"] == a["resident_bytes"] > 0
    sv.inject_fault()
    with pytest.raises(dgx.DgxError, match="injected"):
        sv.wait(sv.submit(pts, 30, req["max_num_pts"], 1))
    st = sv.stats()
    assert st["resident"] == 0 and st["reserved"] == 0 and st["running"] == 0
    # the same problem is built again (not reused), and another one runs
    b = sv.wait(sv.submit(pts, 30, req["max_num_pts"], 1))
    c = sv.wait(sv.submit(pts2, 30, req2["max_num_pts"], 2))
    assert b["reused"] == 0 and b["status"] == 0 and c["status"] == 0
    assert b["rel_residual"] <= 1e-8 and c["rel_residual"] <= 1e-8
    sv.close()


def test_eviction_accounting_across_slots(dgx):
    """Resident bytes are counted per request, whichever slot frees them: after
    a stream with evictions on two slots, the reserved bytes equal the
    measured bytes of the resident problems and stay within the budget."""
    reqs = small_requests(dgx, 20, seed=23)
    budget = 400 << 20
    sv = dgx.Server(slots=2, hbm_budget=budget, nlevels=3, keep_resident=True)
    ids = [sv.submit(c, 30, r["max_num_pts"], r["seed"]) for r, c in reqs]
    res = [sv.wait(i) for i in ids]
    st = sv.stats()
    sv.close()
    assert all(r["status"] == 0 and r["rel_residual"] <= 1e-8 for r in res)
    assert st["running"] == 0 and st["queued"] == 0
    assert st["reserved"] == st["resident_bytes"] <= st["budget"]
    assert st["resident"] >= 1
    kept = sum(r["resident_bytes"] for r in res)
    assert kept >= st["resident_bytes"]
    if kept > budget:
        assert st["evictions"] >= 1


def test_resident_resolves_on_two_slots_with_graphs(dgx):
    """Resident problems re-solved at once on two slots (graph capture on both
    threads) give the first solve's result."""
    reqs = small_requests(dgx, 4, seed=24)
    sv = dgx.Server(slots=2, hbm_budget=64e9, nlevels=4, keep_resident=True, return_solution=True)
    first = [sv.wait(sv.submit(c, 30, r["max_num_pts"], r["seed"]), solution=True) for r, c in reqs]
    ids = [sv.submit(c, 30, r["max_num_pts"], r["seed"]) for r, c in reqs]
    again = [sv.wait(i, solution=True) for i in ids]
    sv.close()
    for a, b in zip(first, again):
        assert b["status"] == 0 and b["reused"] == 1
        assert np.array_equal(a["x"], b["x"])


@pytest.mark.timeout(300)
def test_mix_with_a_2_22_point_request(dgx):
    """Config 4's size range: a mix holding a 2^22-point request (the top of
    the stream's log-uniform range) with small ones on two slots.  Every
    result reaches the residual bar and matches its solve alone."""
    big = {"n": 1 << 22, "dist": 0, "max_num_pts": 4, "seed": 4242}
    reqs = [(big, dgx.serve_points(big))] + small_requests(dgx, 5, seed=25)
    print("requests", [r["n"] for r, _ in reqs], flush=True)
    mixed = run_all(dgx, reqs, slots=2)
    for r, m in zip(reqs, mixed):
        assert m["status"] == 0 and m["rel_residual"] <= 1e-8, (r, m)
    assert mixed[0]["octants"] > 1000000
    print("mix done: big request", mixed[0]["octants"], "octants", mixed[0]["levels"], "levels", flush=True)
    for rc, m in zip(reqs, mixed):
        a, = run_all(dgx, [rc], slots=1)
        assert a["nodes"] == m["nodes"] and a["levels"] == m["levels"]
        assert a["iterations"] == m["iterations"]
        assert np.array_equal(a["x"], m["x"])
