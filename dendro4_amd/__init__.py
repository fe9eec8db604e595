"""dendro4_amd -- MI355X-native octree-FEM hot path of Dendro4.

Python view of the C ABI in include/dendro_mi355x.h (libdendro_mi355x.so,
built for gfx950 by dendro4_amd.build).  The names mirror the reference's C++
API on one rank:

    ot::points2Octree        -> points2octree(pts, max_depth, max_num_pts)
    ot::balanceOctree        -> balance_octree(octree, inc_corner=True)
    par::sampleSort          -> sample_sort(treenodes)
    ot::DA                   -> DA(octree)
    Jacobian1MatMult         -> DA.matvec(x, y)

There is no CPU fallback: importing works without a GPU (for the build check),
but every call needs the library and a device, and raises otherwise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libdendro_mi355x.so")

CURVE_HILBERT = 0
CURVE_MORTON = 1
# ot::TreeNode flags (include/TreeNode.h:79-81, include/oda/odaUtils.h:116-118)
FLAG_BOUNDARY = 64
FLAG_NODE = 128

EXPORTED = [
    "dgx_init", "dgx_last_error", "dgx_stream", "dgx_synchronize", "dgx_device_malloc", "dgx_device_free",
    "dgx_memcpy_htod", "dgx_memcpy_dtoh", "dgx_memset_device", "dgx_points_to_octree", "dgx_balance_octree",
    "dgx_octree_import", "dgx_octree_size", "dgx_octree_max_depth", "dgx_octree_export", "dgx_octree_stats",
    "dgx_octree_free", "dgx_da_create", "dgx_da_get_info", "dgx_da_export", "dgx_da_matvec", "dgx_da_matvec_f32",
    "dgx_da_diagonal", "dgx_da_free", "dgx_gen_points", "dgx_da_matvec_dirichlet", "dgx_mg_create", "dgx_mg_info",
    "dgx_mg_finest_da", "dgx_mg_rhs", "dgx_mg_solve", "dgx_mg_free", "dgx_points_to_octree_glens",
    "dgx_sfc_tables", "dgx_da_export_anchors", "dgx_mg_set_operator", "dgx_server_create", "dgx_server_submit",
    "dgx_server_wait", "dgx_server_clock", "dgx_server_free", "dgx_server_peek", "dgx_read_pts", "dgx_write_pts",
    "dgx_read_data_pts", "dgx_write_data_pts", "dgx_read_nodes", "dgx_write_nodes", "dgx_read_nodes_binary",
    "dgx_write_nodes_binary", "dgx_write_vtk", "dgx_server_get_stats", "dgx_server_inject_fault",
    "dgx_mg_coarse_info", "dgx_direct_plan_host", "dgx_direct_plan_export", "dgx_direct_plan_free",
    "dgx_da_matvec_noalloc", "dgx_memory_info", "dgx_trim_cache", "dgx_mg_level_da", "dgx_da_boundary_flags", "dgx_mg_restrict", "dgx_mg_prolong",
    "dgx_da_compress_maps", "dgx_da_decode_maps",
]


class DgxError(RuntimeError):
    pass


class _DaInfo(C.Structure):
    _fields_ = [
        ("local_buffer_size", C.c_uint64),
        ("element_size", C.c_uint64),
        ("node_size", C.c_uint64),
        ("boundary_node_size", C.c_uint64),
        ("n_nodes", C.c_uint64),
        ("max_depth", C.c_uint32),
        ("ghost_scans", C.c_uint32),
        ("matvec_blocks", C.c_uint64),
        ("block_node_pairs", C.c_uint64),
        ("shared_nodes", C.c_uint64),
        ("shared_pairs", C.c_uint64),
        ("max_block_nodes", C.c_uint64),
    ]


class SolveOpts(C.Structure):
    _fields_ = [("fp32_vcycle", C.c_int), ("pre", C.c_int), ("post", C.c_int), ("omega", C.c_double),
                ("rtol", C.c_double), ("maxit", C.c_int), ("use_graph", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_uint), ("hbm_budget", C.c_uint64), ("nlevels", C.c_int), ("fp32_vcycle", C.c_int),
                ("rtol", C.c_double), ("maxit", C.c_int), ("keep_resident", C.c_int), ("return_solution", C.c_int)]


class Served(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("iterations", C.c_int), ("rel_residual", C.c_double),
                ("octants", C.c_uint64), ("nodes", C.c_uint64), ("levels", C.c_int), ("reused", C.c_int),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("peak_bytes", C.c_uint64), ("resident_bytes", C.c_uint64), ("s_octree", C.c_double),
                ("s_hierarchy", C.c_double), ("s_solve", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("budget", "reserved", "resident", "resident_bytes", "queued", "running",
                                          "evictions")]


SERVER_FAULT_RESOLVE = 1


class DirectSizes(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("interior", "fronts", "links", "levels", "entries", "idx_len", "inv_len",
                                          "fmap_len", "a_nnz", "a_src_len", "upd_len")] + \
               [("max_m", C.c_uint32), ("max_s", C.c_uint32)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rel_residual", C.c_double), ("seconds", C.c_double)]


_lib = None
_ready = False


def load_library():
    """Load libdendro_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DgxError(f"{LIB_PATH} is missing: run `python -m dendro4_amd.build` first")
    L = C.CDLL(LIB_PATH)
    vp, sz, u32, i32, u64, d = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int, C.c_uint64, C.c_double
    sig = {
        "dgx_init": ([i32], i32),
        "dgx_last_error": ([], C.c_char_p),
        "dgx_stream": ([], vp),
        "dgx_synchronize": ([], i32),
        "dgx_device_malloc": ([C.POINTER(vp), sz], i32),
        "dgx_device_free": ([vp], i32),
        "dgx_memcpy_htod": ([vp, vp, sz], i32),
        "dgx_memcpy_dtoh": ([vp, vp, sz], i32),
        "dgx_memset_device": ([vp, i32, sz], i32),
        "dgx_points_to_octree": ([vp, sz, u32, u32, i32, i32, C.POINTER(vp)], i32),
        "dgx_points_to_octree_glens": ([vp, sz, vp, u32, u32, i32, i32, C.POINTER(vp)], i32),
        "dgx_server_create": ([C.POINTER(ServerOpts), C.POINTER(vp)], i32),
        "dgx_server_submit": ([vp, vp, sz, u32, u32, C.c_uint64, C.POINTER(C.c_uint64)], i32),
        "dgx_server_wait": ([vp, C.c_uint64, C.POINTER(Served), vp, sz], i32),
        "dgx_server_clock": ([vp], C.c_double),
        "dgx_server_peek": ([vp, C.c_uint64, C.POINTER(Served)], i32),
        "dgx_server_free": ([vp], i32),
        "dgx_server_get_stats": ([vp, C.POINTER(ServerStats)], i32),
        "dgx_mg_coarse_info": ([vp, vp], i32),
        "dgx_memory_info": ([vp, i32], i32),
        "dgx_trim_cache": ([], i32),
        "dgx_mg_level_da": ([vp, i32, C.POINTER(vp)], i32),
        "dgx_mg_restrict": ([vp, i32, vp, vp], i32),
        "dgx_mg_prolong": ([vp, i32, vp, vp], i32),
        "dgx_da_boundary_flags": ([vp, vp], i32),
        "dgx_da_compress_maps": ([vp, vp, vp, vp, vp, sz, C.POINTER(sz)], i32),
        "dgx_da_decode_maps": ([vp, vp, vp, vp, vp, sz, vp], i32),
        "dgx_direct_plan_host": ([u32, u64, vp, vp, vp, vp, u32, u32, C.POINTER(vp), C.POINTER(DirectSizes)], i32),
        "dgx_direct_plan_export": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
        "dgx_direct_plan_free": ([vp], i32),
        "dgx_server_inject_fault": ([vp, i32], i32),
        "dgx_balance_octree": ([vp, i32, C.POINTER(vp)], i32),
        "dgx_octree_import": ([vp, sz, u32, i32, i32, i32, C.POINTER(vp)], i32),
        "dgx_octree_size": ([vp, C.POINTER(sz)], i32),
        "dgx_octree_max_depth": ([vp, C.POINTER(u32)], i32),
        "dgx_octree_export": ([vp, vp, i32], i32),
        "dgx_octree_stats": ([vp, C.POINTER(u64)], i32),
        "dgx_octree_free": ([vp], i32),
        "dgx_da_create": ([vp, C.POINTER(vp)], i32),
        "dgx_da_get_info": ([vp, C.POINTER(_DaInfo)], i32),
        "dgx_da_export": ([vp, vp, vp, vp, vp], i32),
        "dgx_da_matvec": ([vp, vp, vp, d, d], i32),
        "dgx_da_matvec_f32": ([vp, vp, vp, C.c_float, C.c_float], i32),
        "dgx_da_matvec_noalloc": ([vp, vp, vp, d, d], i32),
        "dgx_da_diagonal": ([vp, vp, d, d], i32),
        "dgx_da_free": ([vp], i32),
        "dgx_gen_points": ([i32, u64, sz, i32, d, d, d, d, i32, vp], i32),
        "dgx_sfc_tables": ([i32, vp, vp, vp], i32),
        "dgx_da_matvec_dirichlet": ([vp, vp, vp, d, d], i32),
        "dgx_mg_create": ([vp, i32, d, d, C.POINTER(vp)], i32),
        "dgx_mg_info": ([vp, C.POINTER(i32), vp, vp], i32),
        "dgx_mg_finest_da": ([vp, C.POINTER(vp)], i32),
        "dgx_mg_rhs": ([vp, vp, vp], i32),
        "dgx_mg_solve": ([vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)], i32),
        "dgx_mg_free": ([vp], i32),
        "dgx_read_pts": ([C.c_char_p, vp, sz, C.POINTER(sz)], i32),
        "dgx_write_pts": ([C.c_char_p, vp, sz], i32),
        "dgx_read_data_pts": ([C.c_char_p, vp, vp, sz, C.POINTER(sz)], i32),
        "dgx_write_data_pts": ([C.c_char_p, vp, vp, sz], i32),
        "dgx_read_nodes": ([C.c_char_p, vp, sz, C.POINTER(sz)], i32),
        "dgx_write_nodes": ([C.c_char_p, vp, sz], i32),
        "dgx_read_nodes_binary": ([C.c_char_p, vp, sz, C.POINTER(sz)], i32),
        "dgx_write_nodes_binary": ([C.c_char_p, vp, sz], i32),
        "dgx_write_vtk": ([vp, sz, i32, C.c_char_p, i32, i32], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise DgxError(load_library().dgx_last_error().decode())


def init(device=0):
    """dgx_init: select the device, create the stream and memory pool."""
    global _ready
    L = load_library()
    _check(L.dgx_init(device))
    _ready = True


def _lib_ready():
    if not _ready:
        init(0)
    return _lib


def stream():
    return _lib_ready().dgx_stream()


def memory_info(reset_peak=False):
    """Device memory: total, used on the device, the library's live bytes, their
    peak (reset after reading when reset_peak), bytes in the block cache."""
    out = (C.c_uint64 * 5)()
    _check(_lib_ready().dgx_memory_info(out, 1 if reset_peak else 0))
    return dict(zip(("total", "device_used", "library_live", "library_peak", "cached"), list(out)))


def trim_cache():
    """Return the library's cached blocks and the pool's unused memory to the
    device (waits for the device)."""
    _check(_lib_ready().dgx_trim_cache())


def synchronize():
    _check(_lib_ready().dgx_synchronize())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device allocation with numpy-style dtype/shape, owned by this object."""

    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _check(_lib_ready().dgx_device_malloc(C.byref(self.ptr), max(self.nbytes, 1)))

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            _check(_lib.dgx_memcpy_htod(d.ptr, _ptr(a), a.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(_lib.dgx_memcpy_dtoh(_ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        _check(_lib.dgx_memset_device(self.ptr, 0, self.nbytes))

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr and self.ptr.value:
            _check(_lib.dgx_device_free(self.ptr))
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Octree:
    """Device-resident sorted linear octree (keys on the chosen curve)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle

    def __len__(self):
        n = C.c_size_t()
        _check(_lib.dgx_octree_size(self._h, C.byref(n)))
        return n.value

    @property
    def max_depth(self):
        m = C.c_uint32()
        _check(_lib.dgx_octree_max_depth(self._h, C.byref(m)))
        return m.value

    def stats(self):
        s = (C.c_uint64 * 5)()
        _check(_lib.dgx_octree_stats(self._h, s))
        return list(s)

    def treenodes(self):
        """ot::TreeNode array (n, 7) uint32: x, y, z, level, weight, dim, maxDepth."""
        n = len(self)
        out = np.zeros((n, 7), np.uint32)
        _check(_lib.dgx_octree_export(self._h, _ptr(out), 0))
        return out

    def free(self):
        if self._h and self._h.value:
            _check(_lib.dgx_octree_free(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def points2octree(pts, max_depth=30, max_num_pts=1, curve=CURVE_HILBERT, glens=None):
    """ot::points2Octree on one rank.  pts: (n,3) float64 host array or a
    DeviceArray; glens: domain lengths (default: the unit cube)."""
    L = _lib_ready()
    h = C.c_void_p()
    g = None if glens is None else np.ascontiguousarray(glens, np.float64).reshape(3)
    gp = None if g is None else _ptr(g)
    if isinstance(pts, DeviceArray):
        n = pts.shape[0]
        _check(L.dgx_points_to_octree_glens(pts.ptr, n, gp, max_depth, max_num_pts, curve, 1, C.byref(h)))
    else:
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        _check(L.dgx_points_to_octree_glens(_ptr(pts), len(pts), gp, max_depth, max_num_pts, curve, 0, C.byref(h)))
    return Octree(h)


def balance_octree(octree, inc_corner=True):
    """ot::balanceOctree on one rank (2:1 across faces, edges and corners)."""
    h = C.c_void_p()
    _check(_lib_ready().dgx_balance_octree(octree._h, 1 if inc_corner else 0, C.byref(h)))
    return Octree(h)


def import_treenodes(tn, max_depth=30, curve=CURVE_HILBERT, sort=True):
    tn = np.ascontiguousarray(tn, np.uint32).reshape(-1, 7)
    h = C.c_void_p()
    _check(_lib_ready().dgx_octree_import(_ptr(tn), len(tn), max_depth, curve, 0, 1 if sort else 0, C.byref(h)))
    return Octree(h)


def sample_sort(tn, max_depth=30, curve=CURVE_HILBERT):
    """par::sampleSort on one rank: the TreeNodes sorted on the curve."""
    return import_treenodes(tn, max_depth, curve, sort=True).treenodes()


class DA:
    """ot::DA on one rank: element->node maps and the -lap+I operator."""

    def __init__(self, octree):
        self._h = C.c_void_p()
        _check(_lib_ready().dgx_da_create(octree._h, C.byref(self._h)))
        info = _DaInfo()
        _check(_lib.dgx_da_get_info(self._h, C.byref(info)))
        self.local_buffer_size = info.local_buffer_size
        self.element_size = info.element_size
        self.node_size = info.node_size
        self.boundary_node_size = info.boundary_node_size
        self.n_nodes = info.n_nodes
        self.max_depth = info.max_depth
        self.ghost_scans = info.ghost_scans
        self.block_info = {k: getattr(info, k) for k in ("matvec_blocks", "block_node_pairs", "shared_nodes",
                                                          "shared_pairs", "max_block_nodes")}

    def maps(self):
        """Host copies: nlist (E,8), masks (E,), oct_levels (N,), local_to_node (N,)."""
        E, N = self.element_size, self.local_buffer_size
        nl = np.zeros((E, 8), np.uint32)
        mk = np.zeros(E, np.uint8)
        ol = np.zeros(N, np.uint8)
        l2n = np.zeros(N, np.uint32)
        _check(_lib.dgx_da_export(self._h, _ptr(nl), _ptr(mk), _ptr(ol), _ptr(l2n)))
        return nl, mk, ol, l2n

    def matvec(self, x, y, lap_fac=1.0, mass_fac=1.0):
        """y = (lap_fac K + mass_fac M) x; x, y DeviceArrays of n_nodes float64 (or float32)."""
        if x.dtype == np.float32:
            _check(_lib.dgx_da_matvec_f32(self._h, x.ptr, y.ptr, lap_fac, mass_fac))
        else:
            _check(_lib.dgx_da_matvec(self._h, x.ptr, y.ptr, lap_fac, mass_fac))

    def matvec_dirichlet(self, x, y, lap_fac=1.0, mass_fac=1.0):
        """DirichletJacobianMatMult: identity rows / dropped columns on boundary nodes."""
        _check(_lib.dgx_da_matvec_dirichlet(self._h, x.ptr, y.ptr, lap_fac, mass_fac))

    def matvec_dirichlet_host(self, x, lap_fac=1.0, mass_fac=1.0):
        dx = DeviceArray.from_host(np.asarray(x, np.float64))
        dy = DeviceArray((self.n_nodes,), np.float64)
        self.matvec_dirichlet(dx, dy, lap_fac, mass_fac)
        return dy.to_host()

    def matvec_host(self, x, lap_fac=1.0, mass_fac=1.0):
        dt = np.float32 if np.asarray(x).dtype == np.float32 else np.float64
        dx = DeviceArray.from_host(np.asarray(x, dt))
        dy = DeviceArray((self.n_nodes,), dt)
        self.matvec(dx, dy, lap_fac, mass_fac)
        return dy.to_host()

    def compressed_maps(self):
        """compressLut maps encoded on the device (dgx_da_compress_maps):
        remainders (E, 8) u8, quotient masks (E,), sort orders (E,), quotients u16."""
        E = self.element_size
        nq = C.c_size_t()
        _check(_lib.dgx_da_compress_maps(self._h, None, None, None, None, 0, C.byref(nq)))
        rem = np.zeros((E, 8), np.uint8)
        qm = np.zeros(E, np.uint8)
        order = np.zeros(E, np.uint8)
        quot = np.zeros(max(nq.value, 1), np.uint16)
        _check(_lib.dgx_da_compress_maps(self._h, _ptr(rem), _ptr(qm), _ptr(order), _ptr(quot), len(quot),
                                         C.byref(nq)))
        return rem, qm, order, quot[: nq.value]

    def decode_maps(self, rem, qmask, order, quot):
        """DA::getNodeIndices over compressed maps, on the device (dgx_da_decode_maps)."""
        arrs = [np.ascontiguousarray(a) for a in (rem, qmask, order, np.asarray(quot, np.uint16))]
        out = np.zeros((self.element_size, 8), np.uint32)
        _check(_lib.dgx_da_decode_maps(self._h, _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]), _ptr(arrs[3]),
                                       len(arrs[3]), _ptr(out)))
        return out

    def boundary_flags(self):
        """Dirichlet flags per local-buffer octant (dgx_da_boundary_flags)."""
        out = np.zeros(self.local_buffer_size, np.uint8)
        _check(_lib.dgx_da_boundary_flags(self._h, _ptr(out)))
        return out

    def diagonal_host(self, lap_fac=1.0, mass_fac=1.0):
        dd = DeviceArray((self.n_nodes,), np.float64)
        _check(_lib.dgx_da_diagonal(self._h, dd.ptr, lap_fac, mass_fac))
        return dd.to_host()

    def free(self):
        if getattr(self, "_owner", None) is not None:
            self._h = C.c_void_p()
            return
        if self._h and self._h.value:
            _check(_lib.dgx_da_free(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @classmethod
    def _borrow(cls, handle, owner):
        self = cls.__new__(cls)
        self._h = handle
        self._owner = owner
        info = _DaInfo()
        _check(_lib.dgx_da_get_info(self._h, C.byref(info)))
        for k in ("local_buffer_size", "element_size", "node_size", "boundary_node_size", "n_nodes", "max_depth",
                  "ghost_scans"):
            setattr(self, k, getattr(info, k))
        self.block_info = {}
        return self


class MG:
    """DAMG on one rank: octree hierarchy, per-level DA, V-cycle-preconditioned CG."""

    def __init__(self, octree, nlevels=5, lap_fac=1.0, mass_fac=1.0):
        self._h = C.c_void_p()
        _check(_lib_ready().dgx_mg_create(octree._h, nlevels, lap_fac, mass_fac, C.byref(self._h)))
        n = C.c_int()
        octs = (C.c_uint64 * 32)()
        nodes = (C.c_uint64 * 32)()
        _check(_lib.dgx_mg_info(self._h, C.byref(n), octs, nodes))
        self.nlevels = n.value
        self.level_octants = list(octs)[: n.value]
        self.level_nodes = list(nodes)[: n.value]
        self.n_nodes = self.level_nodes[0]
        ci = (C.c_uint64 * 6)()
        _check(_lib.dgx_mg_coarse_info(self._h, ci))
        self.coarse = dict(zip(("interior", "fronts", "tree_levels", "max_front", "max_separator", "front_entries"),
                               list(ci)))

    def rhs(self, f, b):
        _check(_lib.dgx_mg_rhs(self._h, f.ptr, b.ptr))

    @property
    def da(self):
        """The finest level's DA, owned by this MG.  Made on each access, so
        that the MG holds no reference to it and is freed as soon as the
        caller drops it."""
        h = C.c_void_p()
        _check(_lib.dgx_mg_finest_da(self._h, C.byref(h)))
        return DA._borrow(h, self)

    def level_da(self, level):
        """The DA of level `level` (0 = finest), owned by this MG."""
        h = C.c_void_p()
        _check(_lib.dgx_mg_level_da(self._h, level, C.byref(h)))
        return DA._borrow(h, self)

    def restrict(self, level, fine, coarse):
        """coarse = R fine from level `level` to level + 1 (device arrays, FP64)."""
        _check(_lib.dgx_mg_restrict(self._h, level, fine.ptr, coarse.ptr))

    def prolong(self, level, coarse, fine):
        """fine = P coarse from level + 1 to level `level` (device arrays, FP64)."""
        _check(_lib.dgx_mg_prolong(self._h, level, coarse.ptr, fine.ptr))

    def solve(self, b, u, fp32=True, pre=2, post=2, omega=0.8, rtol=1e-8, maxit=200, graph=True):
        o = SolveOpts(1 if fp32 else 0, pre, post, omega, rtol, maxit, 1 if graph else 0)
        st = SolveStats()
        _check(_lib.dgx_mg_solve(self._h, b.ptr, u.ptr, C.byref(o), C.byref(st)))
        return {"iterations": st.iterations, "rel_residual": st.rel_residual, "seconds": st.seconds}

    def free(self):
        if self._h and self._h.value:
            _check(_lib.dgx_mg_free(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gen_points(dist, seed, n, n_gauss=0, lo=0.0, hi=1.0, sig_lo=0.0, sig_hi=0.0, dirichlet=False):
    """Deterministic synthetic points (host, (n,3) float64); see dgx_gen_points."""
    L = load_library()
    out = np.empty((n, 3), np.float64)
    rc = L.dgx_gen_points(dist, seed, n, n_gauss, lo, hi, sig_lo, sig_hi, 1 if dirichlet else 0, _ptr(out))
    if rc != 0:
        raise DgxError("dgx_gen_points: bad arguments")
    return out


def sfc_tables(curve=CURVE_HILBERT):
    L = load_library()
    r = np.zeros(192, np.uint8)
    m = np.zeros(192, np.uint8)
    c = np.zeros(192, np.uint8)
    L.dgx_sfc_tables(curve, _ptr(r), _ptr(m), _ptr(c))
    return r.reshape(24, 8), m.reshape(24, 8), c.reshape(24, 8)


def direct_plan(conn, meta, bdy, xyz, leaf=0, dense=2048):
    """Symbolic analysis of the coarsest-level direct solver on host arrays
    (dgx_direct_plan_host; no device): conn (E, 8) node ids, meta (E,) as the
    DA's element meta (mask | child << 8 | level << 11), bdy (n,) boundary
    flags, xyz (n, 3) node coordinates.  Returns the plan's arrays."""
    L = load_library()
    conn = np.ascontiguousarray(np.asarray(conn, np.uint32).T)  # [8][E]
    meta = np.ascontiguousarray(meta, np.uint16)
    bdy = np.ascontiguousarray(bdy, np.uint8)
    xyz = np.ascontiguousarray(xyz, np.uint32)
    n, E = len(bdy), conn.shape[1]
    h = C.c_void_p()
    sz = DirectSizes()
    _check(L.dgx_direct_plan_host(n, E, _ptr(conn), _ptr(meta), _ptr(bdy), _ptr(xyz), leaf, dense, C.byref(h),
                                  C.byref(sz)))
    try:
        out = {k: getattr(sz, k) for k, _ in DirectSizes._fields_}
        a = {"front9": np.zeros((sz.fronts, 9), np.uint64), "link3": np.zeros((sz.links, 3), np.uint64),
             "inv": np.zeros(sz.inv_len, np.int32), "fmap": np.zeros(sz.fmap_len, np.uint32),
             "idx": np.zeros(sz.idx_len, np.uint32), "a_off": np.zeros(sz.a_nnz + 1, np.uint32),
             "a_src": np.zeros(sz.a_src_len, np.uint32), "a_dst": np.zeros(sz.a_nnz, np.uint64),
             "inode": np.zeros(sz.interior, np.uint32)}
        _check(L.dgx_direct_plan_export(h, *[_ptr(a[k]) for k in ("front9", "link3", "inv", "fmap", "idx", "a_off",
                                                                    "a_src", "a_dst", "inode")]))
        out.update(a)
        return out
    finally:
        L.dgx_direct_plan_free(h)


# On-disk formats (host only, no init needed) -------------------------------
def _path(p):
    return os.fsencode(p)


def _read_two_call(fn, path, width, dtype, extra=False):
    L = load_library()
    n = C.c_size_t()
    if extra:
        _check(fn(_path(path), None, None, 0, C.byref(n)))
    else:
        _check(fn(_path(path), None, 0, C.byref(n)))
    out = np.empty((n.value, width), dtype)
    if not extra:
        _check(fn(_path(path), _ptr(out), n.value, C.byref(n)))
        return out
    vals = np.empty(n.value, np.float64)
    _check(fn(_path(path), _ptr(out), _ptr(vals), n.value, C.byref(n)))
    return out, vals


def read_pts(path):
    """ot::readPtsFromFile: (N,3) float64 points of a .pts file."""
    return _read_two_call(load_library().dgx_read_pts, path, 3, np.float64)


def write_pts(path, pts):
    """ot::writePtsToFile: u32 N then the coordinates; an empty file for N = 0."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    _check(load_library().dgx_write_pts(_path(path), _ptr(pts), len(pts)))


def read_data_pts(path):
    """ot::readDataPtsFromFile: ((N,3) points, (N,) values)."""
    return _read_two_call(load_library().dgx_read_data_pts, path, 3, np.float64, extra=True)


def write_data_pts(path, pts, vals=None):
    """ot::writeDataPtsToFile: like the reference, writes N and the points only."""
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    _check(load_library().dgx_write_data_pts(_path(path), _ptr(pts), None, len(pts)))


def read_nodes(path):
    """ot::readNodesFromFile: (N,7) uint32 TreeNode rows of a .ot text file."""
    return _read_two_call(load_library().dgx_read_nodes, path, 7, np.uint32)


def write_nodes(path, tn):
    """ot::writeNodesToFile: 'dim maxDepth', N, then 'x y z level' lines."""
    tn = np.ascontiguousarray(tn, np.uint32).reshape(-1, 7)
    _check(load_library().dgx_write_nodes(_path(path), _ptr(tn), len(tn)))


def read_nodes_binary(path):
    """ot::readNodesFromFile_binary: (N,7) uint32 rows of a raw 28-byte dump."""
    return _read_two_call(load_library().dgx_read_nodes_binary, path, 7, np.uint32)


def write_nodes_binary(path, tn):
    """ot::writeNodesToFile_binary: raw 28-byte TreeNode records."""
    tn = np.ascontiguousarray(tn, np.uint32).reshape(-1, 7)
    _check(load_library().dgx_write_nodes_binary(_path(path), _ptr(tn), len(tn)))


def write_vtk(tn, name, rank=0, hsorted=False, curve=CURVE_HILBERT):
    """treeNodesTovtk: sorts a copy of tn along the curve and writes
    <name>_H_<rank>.vtk (or _M_ for Morton); returns the sorted rows."""
    tn = np.array(tn, np.uint32).reshape(-1, 7)
    _check(load_library().dgx_write_vtk(_ptr(tn), len(tn), rank, _path(name), 1 if hsorted else 0, curve))
    return tn


# Benchmark workloads (point clouds) ----------------------------------------
# name -> (generator arguments, maxNumPts); every cloud is deterministic.
WORKLOADS = {
    # 262,144 i.i.d. uniform points in [0,1)^3, seed 0, maxNumPts 1
    "uniform-256K": (dict(dist=0, seed=0, n=262144), 1),
    # 4,194,304 points, coordinates exp(N(-1, 0.5)) clipped to [0,1), seed 1, maxNumPts 8
    "lognormal-4M": (dict(dist=1, seed=1, n=4194304), 8),
    # 16,777,216 points, 4 Gaussians (centres U[0.2,0.8]^3, sigma 0.02), seed 2, maxNumPts 4
    "gauss4-16M": (dict(dist=2, seed=2, n=16777216, n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.02), 4),
    # 134,217,728 points, 16 Gaussians (sigma log-U[0.005,0.05], centres U[0.1,0.9]^3,
    # Dirichlet(1) weights), seed 4, maxNumPts 4
    "gauss16-128M": (dict(dist=2, seed=4, n=134217728, n_gauss=16, lo=0.1, hi=0.9, sig_lo=0.005, sig_hi=0.05,
                          dirichlet=True), 4),
}


def workload_points(name, n=None):
    """Point cloud of a named workload; n overrides the size (same stream, first n points)."""
    a = dict(WORKLOADS[name][0])
    if n is not None:
        a["n"] = n
    return gen_points(a.pop("dist"), a.pop("seed"), a.pop("n"), **a)


def workload_max_num_pts(name):
    return WORKLOADS[name][1]


class Server:
    """Request server (dgx_server_*): streams of Poisson problems on point
    clouds, `slots` at a time, admitted against an HBM budget, problems kept
    resident for reuse."""

    def __init__(self, slots=4, hbm_budget=200e9, nlevels=5, fp32=True, rtol=1e-8, maxit=200, keep_resident=True,
                 return_solution=False):
        self._o = ServerOpts(slots, int(hbm_budget), nlevels, 1 if fp32 else 0, rtol, maxit,
                             1 if keep_resident else 0, 1 if return_solution else 0)
        self._h = C.c_void_p()
        _check(_lib_ready().dgx_server_create(C.byref(self._o), C.byref(self._h)))
        self._sizes = {}

    def submit(self, pts, max_depth=30, max_num_pts=1, seed=0):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        rid = C.c_uint64()
        _check(_lib.dgx_server_submit(self._h, _ptr(pts), len(pts), max_depth, max_num_pts, seed, C.byref(rid)))
        return rid.value

    def wait(self, rid, solution=False):
        r = Served()
        x = None
        if solution:
            _check(_lib.dgx_server_peek(self._h, rid, C.byref(r)))
            x = np.empty(max(int(r.nodes), 1), np.float64)
            rc = _lib.dgx_server_wait(self._h, rid, C.byref(r), _ptr(x), len(x))
        else:
            rc = _lib.dgx_server_wait(self._h, rid, C.byref(r), None, 0)
        _check(rc)
        out = {k: getattr(r, k) for k, _ in Served._fields_}
        if solution:
            out["x"] = x[: r.nodes].copy()
        return out

    def clock(self):
        return _lib.dgx_server_clock(self._h)

    def stats(self):
        st = ServerStats()
        _check(_lib.dgx_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def inject_fault(self, kind=SERVER_FAULT_RESOLVE):
        _check(_lib.dgx_server_inject_fault(self._h, kind))

    def close(self):
        if self._h:
            _check(_lib.dgx_server_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Served stream of Poisson problems -------------------------------------------
SERVE_DISTS = ("uniform", "lognormal", "gauss4")


def serve_workload(count=2000, seed=3, rate=200.0):
    """Request list of the served stream: per request the point count is
    log-uniform in [2^12, 2^22], the distribution uniform over {uniform [0,1)^3,
    exp(N(-1,0.5)) clipped, 4 Gaussians sigma 0.02}, maxNumPts in {1,4,8};
    arrivals are Poisson at `rate` per second.  Returns dicts with n, dist,
    max_num_pts, arrival (seconds from the start) and the cloud's seed."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(round(2.0 ** rng.uniform(12.0, 22.0)))
        dist = int(rng.integers(0, 3))
        mnp = int(rng.choice([1, 4, 8]))
        out.append({"n": n, "dist": dist, "max_num_pts": mnp, "arrival": t, "seed": (seed << 32) + i})
    return out


def serve_points(req):
    """Point cloud of one served request."""
    if req["dist"] == 2:
        return gen_points(2, req["seed"], req["n"], n_gauss=4, lo=0.2, hi=0.8, sig_lo=0.02)
    return gen_points(req["dist"], req["seed"], req["n"])
