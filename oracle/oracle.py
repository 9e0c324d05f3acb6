"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.  Each wrapper names the reference function it restates
(reference paths relative to Old-BVE-code/).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

T_PP, T_PC, T_CP, T_CC = 0, 1, 2, 3
EVAL_PARITY, EVAL_BLTC, EVAL_REFCC = 0, 1, 2
ASSIGN_PARITY, ASSIGN_ROBUST = 0, 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    _LIB = load(path)
    return _LIB


def load(path):
    """A CDLL of an oracle build at `path` with its signatures declared (a
    separately built variant can be checked side by side with the default)."""
    L = C.CDLL(path)
    vp, i64, i32, f64 = C.c_void_p, C.c_int64, C.c_int, C.c_double
    L.orc_level_offset.restype = i64
    L.orc_level_offset.argtypes = [i32]
    L.orc_icos_nverts.restype = i64
    L.orc_icos_nverts.argtypes = [i32]
    L.orc_icos_init.restype = i64
    L.orc_icos_init.argtypes = [i32, f64, _dp, _ip, _dp]
    L.orc_lu_factor.restype = i32
    L.orc_lu_factor.argtypes = [i32, _dp, _ip]
    L.orc_lu_solve.restype = None
    L.orc_lu_solve.argtypes = [i32, _dp, _ip, _dp, i32]
    L.orc_barycoords.restype = i32
    L.orc_barycoords.argtypes = [_dp, _dp, _dp, _dp, _dp]
    L.orc_points_assign.restype = None
    L.orc_points_assign.argtypes = [i32, _dp, _ip, _dp, _dp, _dp, i64, i32, _ip, _bp]
    L.orc_interp_count.restype = i32
    L.orc_interp_count.argtypes = [i32]
    L.orc_interp_nodes.restype = None
    L.orc_interp_nodes.argtypes = [i32, i32, _dp]
    L.orc_interp_init.restype = i32
    L.orc_interp_init.argtypes = [i32, i32, _dp, _dp, _ip]
    L.orc_tree_create.restype = vp
    L.orc_tree_create.argtypes = [i32, f64, f64, i32, i32, i32, i32, i32, _dp, _dp, _dp, i64]
    L.orc_tree_destroy.restype = None
    L.orc_tree_destroy.argtypes = [vp]
    for nm in ("orc_tree_ninter", "orc_tree_nverts"):
        getattr(L, nm).restype = i64
        getattr(L, nm).argtypes = [vp]
    L.orc_tree_margin.restype = f64
    L.orc_tree_margin.argtypes = [vp]
    for nm, ty in (("orc_tree_inter", C.POINTER(C.c_int32)), ("orc_tree_loc", C.POINTER(C.c_int32)),
                   ("orc_tree_found", C.POINTER(C.c_uint8)), ("orc_tree_counts", C.POINTER(C.c_int32)),
                   ("orc_tree_offsets", C.POINTER(C.c_int64)), ("orc_tree_ids", C.POINTER(C.c_int32)),
                   ("orc_tree_level_base", C.POINTER(C.c_int64)),
                   ("orc_tree_tri_info", C.POINTER(C.c_double)),
                   ("orc_tree_tri_verts", C.POINTER(C.c_int32)),
                   ("orc_tree_verts", C.POINTER(C.c_double)), ("orc_tree_pts", C.POINTER(C.c_double)),
                   ("orc_tree_lu", C.POINTER(C.c_double))):
        getattr(L, nm).restype = ty
        getattr(L, nm).argtypes = [vp]
    L.orc_tree_rhs.restype = None
    L.orc_tree_rhs.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, f64, i32, _dp, _dp, _dp, _dp]
    L.orc_tree_rhs_select.restype = None
    L.orc_tree_rhs_select.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, f64, i32, _bp, _dp, _dp, _dp, _dp]
    L.orc_rhs_direct.restype = None
    L.orc_rhs_direct.argtypes = [i64, _dp, _dp, _dp, _dp, _dp, f64, C.c_void_p, i64, i32,
                                 _dp, _dp, _dp, _dp]
    L.orc_project_points.restype = None
    L.orc_project_points.argtypes = [i64, _dp, _dp, _dp, _dp, f64]
    L.orc_area_init.restype = None
    L.orc_area_init.argtypes = [i64, _dp, _dp, _dp, _ip, i64, _dp]
    L.orc_remesh.restype = i64
    L.orc_remesh.argtypes = [i64, _dp, _dp, _dp, _dp, vp, _ip, i64, _ip, _ip, _dp, _dp, _dp, f64, _dp, vp,
                             _ip]
    L.orc_amr.restype = i64
    L.orc_amr.argtypes = [C.POINTER(i64), C.POINTER(i64), i64, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip,
                          f64, f64]
    L.orc_rk4_step_unprojected.restype = None
    L.orc_rk4_step_unprojected.argtypes = [vp, i64, _dp, _dp, _dp, _dp, _dp, f64, f64, i32]
    L.orc_rk4_step.restype = None
    L.orc_rk4_step.argtypes = [vp, i64, _dp, _dp, _dp, _dp, _dp, f64, f64, i32]
    return L


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def level_offset(l):
    return lib().orc_level_offset(l)


def icos_init(levels, radius=1.0):
    """icos_utils.h:15-133.  Returns (verts[V,3], tri_verts[T,3], tri_info[T,4])."""
    L = lib()
    nv = L.orc_icos_nverts(levels)
    T = L.orc_level_offset(levels + 1)
    verts = np.zeros(3 * nv)
    tv = np.zeros(3 * T, dtype=np.int32)
    ti = np.zeros(4 * T)
    got = L.orc_icos_init(levels, radius, verts, tv, ti)
    assert got == nv
    return verts.reshape(-1, 3), tv.reshape(-1, 3), ti.reshape(-1, 4)


def lu_factor(a):
    """LAPACK dgetrf in netlib order; a is a square array (row-major input)."""
    n = a.shape[0]
    m = _c(np.asarray(a, dtype=np.float64).T.copy().ravel())  # column-major
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_lu_factor(n, m, ipiv)
    return m, ipiv, info


def lu_solve(lu, ipiv, b):
    n = ipiv.shape[0]
    b = np.asarray(b, dtype=np.float64)
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    bb = _c(b.reshape(n, nrhs).T.copy().ravel())
    lib().orc_lu_solve(n, lu, ipiv, bb, nrhs)
    out = bb.reshape(nrhs, n).T
    return out[:, 0] if b.ndim == 1 else out


def barycoords(p1, p2, p3, p):
    """helpers.h:137-150 (dgesv of [p1 p2 p3] c = p)."""
    c = np.zeros(3)
    lib().orc_barycoords(_c(p1), _c(p2), _c(p3), _c(p), c)
    return c


def interp_nodes(degree, rule=0):
    p = lib().orc_interp_count(degree)
    pts = np.zeros(3 * p)
    lib().orc_interp_nodes(degree, rule, pts)
    return pts.reshape(p, 3)


def interp_init(degree, rule=0):
    """interp_utils.h:61-75: nodes, LU of the column-major Vandermonde, ipiv, info."""
    p = lib().orc_interp_count(degree)
    pts = np.zeros(3 * p)
    lu = np.zeros(p * p)
    ipiv = np.zeros(p, dtype=np.int32)
    info = lib().orc_interp_init(degree, rule, pts, lu, ipiv)
    return pts.reshape(p, 3), lu, ipiv, info


class Tree:
    """icosahedron_init + interp_init + points_assign + tree_traverse
    (new_dynamics.cpp:76-94 call order) on particles x, y, z."""

    def __init__(self, x, y, z, tree_levels, theta=0.7, many_count=10, degree=2,
                 assign=ASSIGN_PARITY, eval_mode=EVAL_PARITY, node_rule=0, library=None, theta_decay=1.0):
        L = self._L = library if library is not None else lib()
        self.x, self.y, self.z = _c(x), _c(y), _c(z)
        self.n = self.x.shape[0]
        self.L = tree_levels
        self.degree = degree
        self.h = L.orc_tree_create(tree_levels, theta, theta_decay, many_count, degree, assign, eval_mode,
                                   node_rule, self.x, self.y, self.z, self.n)
        if not self.h:
            raise RuntimeError("orc_tree_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_tree_destroy(self.h)
            self.h = None

    @property
    def ninter(self):
        return self._L.orc_tree_ninter(self.h)

    @property
    def mac_margin(self):
        return self._L.orc_tree_margin(self.h)

    def interactions(self):
        """int32 [M, 7]: lev_t, lev_s, idx_t, idx_s, n_t, n_s, type (0 pp,1 pc,2 cp,3 cc)."""
        m = self.ninter
        p = self._L.orc_tree_inter(self.h)
        return np.ctypeslib.as_array(p, shape=(m * 7,)).reshape(m, 7).copy()

    def loc(self):
        p = self._L.orc_tree_loc(self.h)
        return np.ctypeslib.as_array(p, shape=(self.L * self.n,)).reshape(self.L, self.n).copy()

    def found(self):
        p = self._L.orc_tree_found(self.h)
        return np.ctypeslib.as_array(p, shape=(self.L * self.n,)).reshape(self.L, self.n).astype(bool)

    def counts(self):
        T = level_offset(self.L)
        return np.ctypeslib.as_array(self._L.orc_tree_counts(self.h), shape=(T,)).copy()

    def offsets(self):
        T = level_offset(self.L)
        return np.ctypeslib.as_array(self._L.orc_tree_offsets(self.h), shape=(T,)).copy()

    def level_base(self):
        return np.ctypeslib.as_array(self._L.orc_tree_level_base(self.h), shape=(self.L + 1,)).copy()

    def ids(self):
        tot = int(self.level_base()[-1])
        if tot == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(self._L.orc_tree_ids(self.h), shape=(tot,)).copy()

    def tri_info(self):
        T = level_offset(self.L + 1)
        return np.ctypeslib.as_array(self._L.orc_tree_tri_info(self.h), shape=(4 * T,)).reshape(T, 4).copy()

    def tri_verts(self):
        T = level_offset(self.L + 1)
        return np.ctypeslib.as_array(self._L.orc_tree_tri_verts(self.h), shape=(3 * T,)).reshape(T, 3).copy()

    def verts(self):
        nv = self._L.orc_tree_nverts(self.h)
        return np.ctypeslib.as_array(self._L.orc_tree_verts(self.h), shape=(3 * nv,)).reshape(nv, 3).copy()

    def node_list(self, lev, idx):
        g = level_offset(lev) + idx
        c = self.counts()[g]
        o = self.offsets()[g]
        return self.ids()[o:o + c]

    def rhs(self, vort, area, omega=2 * np.pi, nthreads=1, select=None):
        """rhs_func with use_fast (rhs_eval.h:47-56) -> (u, v, w, dvort).
        select (optional): boolean per interaction (rows of interactions());
        only the selected interactions contribute."""
        n = self.n
        out = [np.zeros(n) for _ in range(4)]
        if select is None:
            self._L.orc_tree_rhs(self.h, self.x, self.y, self.z, _c(vort), _c(area), omega, nthreads, *out)
        else:
            take = _c(np.asarray(select, dtype=bool), np.uint8)
            assert take.shape == (self.ninter,)
            self._L.orc_tree_rhs_select(self.h, self.x, self.y, self.z, _c(vort), _c(area), omega, nthreads,
                                      take, *out)
        return tuple(out)


def rhs_direct(x, y, z, vort, area, omega=2 * np.pi, idx=None, nthreads=None):
    """rhs_func with direct sum (rhs_eval.h:15-32, 47-56); targets idx (default all)."""
    x, y, z, vort, area = (_c(a) for a in (x, y, z, vort, area))
    n = x.shape[0]
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    if idx is None:
        m = n
        ip = None
    else:
        idx = _c(idx, np.int64)
        m = idx.shape[0]
        ip = idx.ctypes.data_as(C.c_void_p)
    out = [np.zeros(m) for _ in range(4)]
    lib().orc_rhs_direct(n, x, y, z, vort, area, omega, ip, m, nthreads, *out)
    return tuple(out)


def project_points(x, y, z, vort, omega=2 * np.pi):
    """helpers.h:362-372 on copies; returns (x, y, z, vort)."""
    a = [_c(v).copy() for v in (x, y, z, vort)]
    lib().orc_project_points(a[0].shape[0], *a, omega)
    return tuple(a)


def area_init(x, y, z, tris):
    """helpers.h:344-360: one third of each adjacent spherical triangle."""
    x, y, z = _c(x), _c(y), _c(z)
    tris = _c(tris, np.int32).ravel()
    area = np.zeros(x.shape[0])
    lib().orc_area_init(x.shape[0], x, y, z, tris, tris.shape[0] // 3, area)
    return area


def rk4_step(x, y, z, vort, area, dt, omega=2 * np.pi, tree=None, nthreads=None, project=True):
    """One RK4 step with projection (new_dynamics.cpp:142-238 intent); returns new arrays.
    project=False leaves out the final projection (the state AMR sees)."""
    a = [_c(v).copy() for v in (x, y, z, vort)]
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    f = lib().orc_rk4_step if project else lib().orc_rk4_step_unprojected
    f(tree.h if tree is not None else None, a[0].shape[0], *a, _c(area), omega, dt, nthreads)
    return tuple(a)


def vertex_triangles(n, tris):
    """Triangles adjacent to every vertex, ascending triangle index (the
    vert_tris lists read by read_points helpers.h:299-342), as CSR."""
    tris = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    cnt = np.bincount(tris.ravel(), minlength=n)
    off = np.zeros(n + 1, dtype=np.int32)
    off[1:] = np.cumsum(cnt)
    order = np.argsort(tris.ravel(), kind="stable")
    idx = (order // 3).astype(np.int32)
    return off, idx


def remesh(ax, ay, az, avort, tris, vt_off, vt_idx, x, y, z, omega=2 * np.pi, atracer=None):
    """regrid_points2 remesh.h:109-178 with tri_interp helpers.h:251-265:
    returns (vort at the targets, tracer or None, triangle per target,
    number of targets that needed the full search)."""
    n = len(x)
    tris = _c(tris, np.int32).ravel()
    vort = np.zeros(n)
    tr = np.zeros(n) if atracer is not None else None
    tri_out = np.zeros(n, dtype=np.int32)
    atr = _c(atracer) if atracer is not None else None
    nf = lib().orc_remesh(n, _c(ax), _c(ay), _c(az), _c(avort),
                          atr.ctypes.data_as(C.c_void_p) if atr is not None else None, tris,
                          tris.shape[0] // 3, _c(vt_off, np.int32), _c(vt_idx, np.int32), _c(x), _c(y), _c(z),
                          omega, vort, tr.ctypes.data_as(C.c_void_p) if tr is not None else None, tri_out)
    return vort, tr, tri_out, int(nf)


def amr(state, area, tris, parent, vert_tris, vert_tri_count, n, ntri, max_count,
        circulation_tol=0.0025, vorticity_tol=0.2):
    """amr.h:13-178 on host arrays edited in place (vert_tris stride 12);
    returns (n, ntri, refined)."""
    nn, nt = C.c_int64(n), C.c_int64(ntri)
    r = lib().orc_amr(C.byref(nn), C.byref(nt), max_count, *state, area, tris.reshape(-1), parent.reshape(-1),
                      vert_tris.reshape(-1), vert_tri_count, circulation_tol, vorticity_tol)
    return nn.value, nt.value, int(r)
