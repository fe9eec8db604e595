"""GPU test of dgx_da_matvec_noalloc, the matvec entry point that allocates
nothing (shared-node partial sums in a static device array): called through
the C ABI, it matches dgx_da_matvec to 1e-12 on two calls in a row, the second
on a smaller DA, and refuses a null argument."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def run_both(dgx, da, seed):
    L = dgx.load_library()
    x = np.random.default_rng(seed).standard_normal(da.n_nodes)
    dx = dgx.DeviceArray.from_host(x)
    y_ref = dgx.DeviceArray((da.n_nodes,), np.float64)
    y_na = dgx.DeviceArray((da.n_nodes,), np.float64)
    assert L.dgx_da_matvec(da._h, dx.ptr, y_ref.ptr, C.c_double(1.0), C.c_double(1.0)) == 0
    rc = L.dgx_da_matvec_noalloc(da._h, dx.ptr, y_na.ptr, C.c_double(1.0), C.c_double(1.0))
    assert rc == 0, L.dgx_last_error()
    return y_na.to_host(), y_ref.to_host()


def test_noalloc_matvec_matches_matvec(dgx):
    big = dgx.DA(dgx.balance_octree(dgx.points2octree(dgx.gen_points(1, 31, 200000), 30, 1)))
    small = dgx.DA(dgx.balance_octree(dgx.points2octree(dgx.gen_points(0, 32, 5000), 30, 1)))
    assert small.n_nodes < big.n_nodes
    print("DAs", big.n_nodes, small.n_nodes, big.block_info, flush=True)
    for da, seed in ((big, 1), (small, 2)):  # the second call on a smaller input
        y, yr = run_both(dgx, da, seed)
        assert rel_err(y, yr) < 1e-12
    L = dgx.load_library()
    assert L.dgx_da_matvec_noalloc(big._h, None, None, C.c_double(1.0), C.c_double(1.0)) != 0
    assert b"null" in L.dgx_last_error()
