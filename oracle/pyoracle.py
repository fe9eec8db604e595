"""ctypes binding of the CPU oracle (oracle/dendro_oracle.c).

Test infrastructure only: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``.  The product library never uses it.
TreeNode arrays are numpy ``uint32`` arrays of shape (n, 7) with the
ot::TreeNode field order x, y, z, level|flags, weight, dim, maxDepth.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class _TN(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("x", "y", "z", "lev", "w", "dim", "md")]


class _DA(C.Structure):
    _fields_ = [
        ("octs", C.POINTER(_TN)),
        ("n_local", C.c_size_t),
        ("n_elem", C.c_size_t),
        ("nlist", C.POINTER(C.c_uint32)),
        ("mask", C.POINTER(C.c_uint8)),
        ("node_size", C.c_uint32),
        ("bnd_node_size", C.c_uint32),
        ("local_to_node", C.POINTER(C.c_uint32)),
        ("n_nodes", C.c_uint32),
        ("n_unresolved", C.c_size_t),
        ("n_find_ghost", C.c_size_t),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.or_set_ordering.argtypes = [C.c_int]
        L.or_get_tables.argtypes = [vp, vp]
        L.or_get_etype_masks.argtypes = [vp]
        L.or_tn_less.argtypes = [vp, vp]
        L.or_sort.argtypes = [vp, C.c_size_t]
        L.or_free.argtypes = [vp]
        L.or_quantise.argtypes = [vp, C.c_size_t, C.c_uint32, vp]
        L.or_points2octree.argtypes = [vp, C.c_size_t, C.c_uint32, C.c_uint32, C.POINTER(vp), C.POINTER(C.c_size_t)]
        L.or_balance.argtypes = [vp, C.c_size_t, C.c_int, C.c_int, C.POINTER(vp), C.POINTER(C.c_size_t)]
        for f in ("or_is_sorted_unique", "or_is_linear", "or_is_complete", "or_is_balanced"):
            getattr(L, f).argtypes = [vp, C.c_size_t]
        L.or_is_balanced_at.argtypes = [vp, C.c_size_t, vp, C.c_size_t]
        L.or_num_threads.restype = C.c_int
        L.or_da_build.argtypes = [vp, C.c_size_t, C.POINTER(_DA)]
        L.or_da_free.argtypes = [C.POINTER(_DA)]
        L.or_elem_type.argtypes = [C.c_uint32, C.c_uint32]
        L.or_matvec.argtypes = [C.POINTER(_DA), vp, vp, C.c_double, C.c_double, vp, vp]
        L.or_da_element_view.argtypes = [C.POINTER(_DA), vp, vp, vp, vp]
        L.or_da_compress.argtypes = [C.POINTER(_DA), vp, vp, vp, vp]
        L.or_da_compress.restype = C.c_size_t
        L.or_da_decode.argtypes = [C.POINTER(_DA), vp, vp, vp, vp, vp]
        L.or_coarsen.argtypes = [vp, C.c_size_t, C.POINTER(vp), C.POINTER(C.c_size_t)]
        L.or_restrict_entries.argtypes = [C.POINTER(_DA), C.POINTER(_DA)] + [vp] * 11 + [C.POINTER(C.c_size_t)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def set_ordering(hilbert=True):
    lib().or_set_ordering(1 if hilbert else 0)


def tables():
    rot = np.zeros(384, np.uint8)
    ht = np.zeros(192, np.uint8)
    lib().or_get_tables(_p(rot), _p(ht))
    return rot.reshape(24, 16), ht.reshape(24, 8)


def etype_masks():
    m = np.zeros(144, np.uint8)
    lib().or_get_etype_masks(_p(m))
    return m.reshape(8, 18)


def _take(ptr, n):
    if n == 0:
        return np.zeros((0, 7), np.uint32)
    buf = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint32)), shape=(n * 7,)).copy()
    lib().or_free(ptr)
    return buf.reshape(n, 7)


def tn_less(a, b):
    a = np.ascontiguousarray(a, np.uint32)
    b = np.ascontiguousarray(b, np.uint32)
    return bool(lib().or_tn_less(_p(a), _p(b)))


def sort(tn):
    tn = np.ascontiguousarray(tn, np.uint32).copy()
    lib().or_sort(_p(tn), len(tn))
    return tn


def quantise(pts, max_depth=30):
    pts = np.ascontiguousarray(pts, np.float64)
    out = np.zeros((len(pts), 7), np.uint32)
    lib().or_quantise(_p(pts), len(pts), max_depth, _p(out))
    return out


def points2octree(pts, max_depth=30, max_num_pts=1):
    pts = np.ascontiguousarray(pts, np.float64)
    ptr = C.c_void_p()
    n = C.c_size_t()
    rc = lib().or_points2octree(_p(pts), len(pts), max_depth, max_num_pts, C.byref(ptr), C.byref(n))
    if rc == -2:
        raise ValueError("points2Octree: stale cursor reached the end of the domain (reference compares with "
                         "the getNext sentinel)")
    return _take(ptr, n.value)


def balance(tn, inc_corner=True, emulate_limit=False):
    """balanceOctree on one rank.  emulate_limit=True reproduces comboRipple's
    empty result above maxNum = 1e7 input octants; an integer > 1 is taken as
    maxNum instead."""
    tn = np.ascontiguousarray(tn, np.uint32)
    ptr = C.c_void_p()
    n = C.c_size_t()
    lim = 0 if emulate_limit is False else (1 if emulate_limit is True else int(emulate_limit))
    lib().or_balance(_p(tn), len(tn), 1 if inc_corner else 0, lim, C.byref(ptr), C.byref(n))
    return _take(ptr, n.value)


def coarsen(tn):
    """coarsenOctree on one rank (src/coarsen.cpp:64-104)."""
    tn = np.ascontiguousarray(tn, np.uint32)
    out, n = C.c_void_p(), C.c_size_t()
    if lib().or_coarsen(_p(tn), len(tn), C.byref(out), C.byref(n)) < 0:
        raise ValueError("coarsen: the octree does not start with a first child")
    return _take(out, n.value)


def is_sorted_unique(tn):
    tn = np.ascontiguousarray(tn, np.uint32)
    return bool(lib().or_is_sorted_unique(_p(tn), len(tn)))


def is_linear(tn):
    tn = np.ascontiguousarray(tn, np.uint32)
    return bool(lib().or_is_linear(_p(tn), len(tn)))


def is_complete(tn):
    tn = np.ascontiguousarray(tn, np.uint32)
    return bool(lib().or_is_complete(_p(tn), len(tn)))


def is_balanced(tn):
    tn = np.ascontiguousarray(tn, np.uint32)
    return bool(lib().or_is_balanced(_p(tn), len(tn)))


def is_balanced_at(tn, idx):
    """The balance check of is_balanced for the leaves tn[idx] only."""
    tn = np.ascontiguousarray(tn, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint64)
    return bool(lib().or_is_balanced_at(_p(tn), len(tn), _p(idx), len(idx)))


def num_threads():
    """OpenMP threads of the oracle's sorts and checkers."""
    return int(lib().or_num_threads())


class DA:
    """One-rank ot::DA restated on the CPU."""

    def __init__(self, octree):
        octree = np.ascontiguousarray(octree, np.uint32)
        self._da = _DA()
        rc = lib().or_da_build(_p(octree), len(octree), C.byref(self._da))
        if rc < 0:
            raise ValueError("DA needs more than 8 octants")
        d = self._da
        self.n_local = d.n_local
        self.n_elem = d.n_elem
        self.node_size = d.node_size
        self.bnd_node_size = d.bnd_node_size
        self.n_nodes = d.n_nodes
        self.n_unresolved = d.n_unresolved
        self.n_find_ghost = d.n_find_ghost
        self.octs = np.ctypeslib.as_array(C.cast(d.octs, C.POINTER(C.c_uint32)), shape=(d.n_local * 7,)).reshape(
            d.n_local, 7).copy()
        self.nlist = np.ctypeslib.as_array(d.nlist, shape=(d.n_elem * 8,)).reshape(d.n_elem, 8).copy()
        self.mask = np.ctypeslib.as_array(d.mask, shape=(d.n_elem,)).copy()
        self.local_to_node = np.ctypeslib.as_array(d.local_to_node, shape=(d.n_local,)).copy()
        self.oct_levels = (self.octs[:, 3] & 0xFF).astype(np.uint8)

    def element_view(self):
        n = self.n_elem
        conn = np.zeros((n, 8), np.uint32)
        cn = np.zeros(n, np.uint8)
        lev = np.zeros(n, np.uint8)
        mk = np.zeros(n, np.uint8)
        lib().or_da_element_view(C.byref(self._da), _p(conn), _p(cn), _p(lev), _p(mk))
        return conn, cn, lev, mk

    def compress(self):
        """compressLut maps (odaBuildNlist.cpp:3120-3262): remainders (E, 8) u8,
        quotient masks (E,) u8, sort orders (E,) u8, quotients u16."""
        n = self.n_elem
        rem = np.zeros((n, 8), np.uint8)
        msk = np.zeros(n, np.uint8)
        order = np.zeros(n, np.uint8)
        nq = lib().or_da_compress(C.byref(self._da), _p(rem), _p(msk), _p(order), None)
        quot = np.zeros(max(nq, 1), np.uint16)
        lib().or_da_compress(C.byref(self._da), _p(rem), _p(msk), _p(order), _p(quot))
        return rem, msk, order, quot[:nq]

    def decode(self, rem, msk, order, quot):
        """getNodeIndices on compressed maps (oda.tcc:1304-3950): nlist (E, 8)."""
        out = np.zeros((self.n_elem, 8), np.uint32)
        arrs = [np.ascontiguousarray(a) for a in (rem, msk, order, quot)]
        lib().or_da_decode(C.byref(self._da), *[_p(a) for a in arrs], _p(out))
        return out

    def matvec(self, x, L, M, lap_fac=1.0, mass_fac=1.0):
        x = np.ascontiguousarray(x, np.float64)
        L = np.ascontiguousarray(L, np.float64)
        M = np.ascontiguousarray(M, np.float64)
        y = np.zeros(self.n_nodes, np.float64)
        rc = lib().or_matvec(C.byref(self._da), _p(L), _p(M), lap_fac, mass_fac, _p(x), _p(y))
        if rc:
            raise RuntimeError("unknown element type")
        return y

    def restriction(self, coarse, tables, sup_fine=None, sup_coarse=None):
        """The entries of restrictMatVecType1 from this (fine) DA to `coarse`
        (src/omg/restrictMatVec.cpp:71-619): rows (coarse compact nodes), cols
        (fine compact nodes) and values, in the reference's accumulation order.
        Duplicates add.  tables: the arrays of tests/golden/rmat_stencils.npz.
        sup_fine / sup_coarse: suppressed nodes (uint8 per compact node)."""
        t = {k: np.ascontiguousarray(tables[k]) for k in ("rmat1", "rmat2", "map1", "map2", "map3", "map4")}
        cap = 64 * self.n_elem
        rows = np.zeros(cap, np.uint32)
        cols = np.zeros(cap, np.uint32)
        vals = np.zeros(cap, np.float64)
        sf = None if sup_fine is None else np.ascontiguousarray(sup_fine, np.uint8)
        sc = None if sup_coarse is None else np.ascontiguousarray(sup_coarse, np.uint8)
        nnz = C.c_size_t(cap)
        rc = lib().or_restrict_entries(C.byref(self._da), C.byref(coarse._da), _p(t["rmat1"]), _p(t["rmat2"]),
                                       _p(t["map1"]), _p(t["map2"]), _p(t["map3"]), _p(t["map4"]),
                                       None if sf is None else _p(sf), None if sc is None else _p(sc),
                                       _p(rows), _p(cols), _p(vals), C.byref(nnz))
        if rc:
            raise RuntimeError(f"restriction entries failed ({rc})")
        n = nnz.value
        return rows[:n].copy(), cols[:n].copy(), vals[:n].copy()

    def node_coords(self):
        """Anchor of each compact node in the DA's embedded coordinates, (n_nodes, 3)."""
        sel = self.local_to_node != 0xFFFFFFFF
        out = np.zeros((self.n_nodes, 3), np.int64)
        out[self.local_to_node[sel]] = self.octs[sel, :3]
        return out

    def __del__(self):
        try:
            lib().or_da_free(C.byref(self._da))
        except Exception:
            pass
