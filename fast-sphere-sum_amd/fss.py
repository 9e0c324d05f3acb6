"""Python front end of libfss.so (ctypes over the C ABI in include/fss.h).

Device arrays are torch tensors on a HIP device (float64, contiguous); the
library runs on torch's current stream.  There is no CPU fallback: if the
HIP library is missing, `lib()` raises.

Reference functions replaced (Old-BVE-code/):
  Context.tree_build        icosahedron_init + interp_init + points_assign +
                            tree_traverse (new_dynamics.cpp:84-88)
  Context.velocity_tree     rhs_func with use_fast (rhs_eval.h:47-56, 34-45)
  Context.velocity_direct   rhs_func direct (rhs_eval.h:47-56, 15-32)
  Context.rk4_step          RK4 loop body (new_dynamics.cpp:142-238)
  grid_icosahedral etc.     read_points / area_init / vorticity_init inputs
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfss.so")

FSS_OK = 0
FP64, MIXED = 0, 1
ASSIGN_PARITY, ASSIGN_ROBUST = 0, 1
EVAL_PARITY, EVAL_BLTC = 0, 1
NODES_REFERENCE, NODES_SIN2, NODES_WARP_BLEND = 0, 1, 2

_STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of memory",
           4: "invalid call order", 5: "numerical failure", 6: "no HIP device"}


class FssError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"fss status {status} ({_STATUS.get(status, '?')}): {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("theta", C.c_double), ("degree", C.c_int), ("many_count", C.c_int),
                ("tree_levels", C.c_int), ("radius", C.c_double), ("omega", C.c_double),
                ("precision", C.c_int), ("assign_mode", C.c_int), ("eval_mode", C.c_int),
                ("node_rule", C.c_int), ("device", C.c_int), ("theta_decay", C.c_double)]


class Particles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("vort", C.c_void_p),
                ("area", C.c_void_p), ("n", C.c_int64)]


class Velocity(C.Structure):
    _fields_ = [("u", C.c_void_p), ("v", C.c_void_p), ("w", C.c_void_p), ("dvort", C.c_void_p)]


class TunePoint(C.Structure):
    _fields_ = [("degree", C.c_int), ("theta", C.c_double), ("tree_levels", C.c_int), ("many_count", C.c_int),
                ("near_tile", C.c_int), ("node_rule", C.c_int), ("theta_decay", C.c_double)]


class TreeStats(C.Structure):
    _fields_ = [("n", C.c_int64), ("tree_levels", C.c_int), ("n_nodes", C.c_int64),
                ("n_interactions", C.c_int64), ("n_type", C.c_int64 * 4),
                ("dropped", C.c_int64 * 32), ("pairs_near", C.c_int64), ("pairs_far", C.c_int64),
                ("mac_margin", C.c_double), ("n_generations", C.c_int),
                ("n_records", C.c_int64), ("n_near_items", C.c_int64), ("n_proxy_items", C.c_int64),
                ("n_interpolants", C.c_int64), ("n_charge_nodes", C.c_int64)]


# exported symbol -> (restype, argtypes); the list is the ABI of include/fss.h
_vp, _i64, _int, _dbl, _u64 = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_uint64
SIGNATURES = {
    "fss_version": (_int, []),
    "fss_last_error": (C.c_char_p, []),
    "fss_status_string": (C.c_char_p, [_int]),
    "fss_config_default": (_int, [C.POINTER(Config)]),
    "fss_create": (_int, [C.POINTER(_vp), C.POINTER(Config)]),
    "fss_destroy": (_int, [_vp]),
    "fss_set_stream": (_int, [_vp, _vp]),
    "fss_sync": (_int, [_vp]),
    "fss_tree_build": (_int, [_vp, _vp, _vp, _vp, _i64]),
    "fss_tree_stats_get": (_int, [_vp, C.POINTER(TreeStats)]),
    "fss_batch_build": (_int, [_vp, _int, _vp, _vp, _vp, _vp, _vp]),
    "fss_autotune_levels": (_int, [_vp, C.POINTER(Particles), _int, _int, _int, C.POINTER(_int), _vp]),
    "fss_levels_for": (_int, [C.c_int64, _dbl]),
    "fss_autotune": (_int, [_vp, C.POINTER(Particles), C.POINTER(TunePoint), _int, _vp, _i64, _dbl, _int,
                            C.POINTER(_int), _vp, _vp]),
    "fss_set_near_tile": (_int, [_vp, _int]),
    "fss_tree_export": (_int, [_vp] + [_vp] * 8),
    "fss_velocity_direct": (_int, [_vp, C.POINTER(Particles), C.POINTER(Velocity)]),
    "fss_velocity_tree": (_int, [_vp, C.POINTER(Particles), C.POINTER(Velocity)]),
    "fss_velocity_direct_targets": (_int, [_vp, C.POINTER(Particles), _vp, _i64,
                                           C.POINTER(Velocity)]),
    "fss_velocity_direct_noalloc": (_int, [_vp, C.POINTER(Particles), C.POINTER(Velocity)]),
    "fss_rk4_step": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _int]),
    "fss_rk4_run_graph": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _dbl, _int, _int]),
    "fss_project_points": (_int, [_vp, _vp, _vp, _vp, _vp, _i64]),
    "fss_last_timing": (_int, [_vp, C.POINTER(_dbl), C.POINTER(_dbl)]),
    "fss_timing_reset": (_int, [_vp]),
    "fss_timing_mean": (_int, [_vp, C.POINTER(_i64), C.POINTER(_dbl), C.POINTER(_dbl)]),
    "fss_timing_far_mean": (_int, [_vp, C.POINTER(_dbl)]),
    "fss_grid_icosahedral": (_int, [_int, _vp, _vp, _vp, _vp]),
    "fss_grid_voronoi_areas": (_int, [_i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    "fss_grid_third_areas": (_int, [_i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    "fss_grid_vertex_triangles": (_int, [_i64, _vp, _i64, _vp, _vp]),
    "fss_amr": (_int, [C.POINTER(_i64), C.POINTER(_i64), _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                       _vp, _vp, _int, _dbl, _dbl, C.POINTER(_i64)]),
    "fss_remesh": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                          _i64, C.POINTER(_i64), C.POINTER(_i64)]),
    "fss_grid_random": (_int, [_i64, _u64, _vp, _vp, _vp]),
    "fss_field_rossby_haurwitz": (_int, [_i64, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp]),
    "fss_field_gauss_vortices": (_int, [_i64, _vp, _vp, _vp, _vp, _int, _dbl, _dbl, _dbl, _u64,
                                        _vp]),
    "fss_sample_targets": (_int, [_i64, _i64, _u64, _vp]),
    "fss_sal_love_numbers": (_int, [_int, _u64, _vp, _vp]),
    "fss_sal_coefficients": (_int, [_int, _vp, _vp, _dbl, _int, _vp]),
    "fss_sal_set_series": (_int, [_vp, _vp, _int, _i64, _dbl]),
    "fss_sal_kernel": (_int, [_vp, _int, _vp, _i64, _vp]),
    "fss_sal_direct": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _int, _vp]),
    "fss_sal_mlp_nparam": (_int, []),
    "fss_sal_tree": (_int, [_vp, C.POINTER(Particles), _int, _vp]),
    "fss_sal_mlp_init": (_int, [_vp, _u64]),
    "fss_sal_mlp_train": (_int, [_vp, _int, _i64, _dbl, _dbl, _u64, C.POINTER(_dbl)]),
    "fss_sal_mlp_params": (_int, [_vp, _vp, _int]),
    "fss_sal_mlp_norm": (_int, [_vp, C.POINTER(_dbl), C.POINTER(_dbl), C.POINTER(_dbl), _int]),
    "fss_sal_mlp_grad": (_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
}

_LIB = None


def lib():
    """Load libfss.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _LIB
    if _LIB is None:
        # torch (if installed) ships its own HIP runtime under the same soname;
        # loading it first makes libfss bind to that one runtime instead of a
        # second copy from /opt/rocm, so torch tensors and libfss share a device
        # context.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise FssError(6, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(status):
    if status != FSS_OK:
        raise FssError(status, lib().fss_last_error().decode())


def _np_ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _dev_ptr(t):
    import torch
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous(), "need a contiguous HIP tensor"
    return C.c_void_p(t.data_ptr())


def _check_state(device, *arrays):
    """Particle state arrays: float64, one length, all on the context's device.
    The kernels read n doubles from every pointer, so a shorter or narrower
    array would be read past its end."""
    import torch
    n = arrays[0].numel()
    for t in arrays:
        if not (isinstance(t, torch.Tensor) and t.dtype == torch.float64 and t.numel() == n
                and t.device == device and t.is_contiguous()):
            raise FssError(1, f"particle arrays must be contiguous float64 tensors of one length on "
                              f"{device} (got {getattr(t, 'dtype', type(t))}, {getattr(t, 'numel', lambda: '?')()} "
                              f"elements on {getattr(t, 'device', '?')}, expected {n})")


# ---------------- grids and fields (host) ----------------

def grid_icosahedral(level):
    """Icosahedral vertices (icos_init order) and faces after `level` refinements."""
    nv = 10 * 4 ** level + 2
    nf = 20 * 4 ** level
    x, y, z = (np.zeros(nv) for _ in range(3))
    tris = np.zeros(3 * nf, dtype=np.int32)
    _check(lib().fss_grid_icosahedral(level, _np_ptr(x), _np_ptr(y), _np_ptr(z), _np_ptr(tris)))
    return x, y, z, tris.reshape(nf, 3)


def voronoi_areas(x, y, z, tris):
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    a = np.zeros(len(x))
    _check(lib().fss_grid_voronoi_areas(len(x), _np_ptr(x), _np_ptr(y), _np_ptr(z), _np_ptr(tris),
                                        tris.shape[0], _np_ptr(a)))
    return a


def third_areas(x, y, z, tris):
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    a = np.zeros(len(x))
    _check(lib().fss_grid_third_areas(len(x), _np_ptr(x), _np_ptr(y), _np_ptr(z), _np_ptr(tris),
                                      tris.shape[0], _np_ptr(a)))
    return a


def vertex_triangles(n, tris):
    """CSR (vt_off[n+1], vt_idx) of the triangles adjacent to every vertex."""
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    off = np.zeros(n + 1, dtype=np.int32)
    idx = np.zeros(tris.size, dtype=np.int32)
    _check(lib().fss_grid_vertex_triangles(n, _np_ptr(tris), tris.shape[0], _np_ptr(off), _np_ptr(idx)))
    return off, idx


def amr(state, area, tris, parent, vert_tris, vert_tri_count, n, ntri, max_count, tri_cap,
        circulation_tol=0.0025, vorticity_tol=0.2):
    """One refinement pass (amr.h:13-178) on host arrays, edited in place:
    state = [x, y, z, vort, tracer] (capacity max_count each), area,
    tris [tri_cap, 3] int32, parent [max_count, 2] int32, vert_tris
    [max_count, stride] int32, vert_tri_count [max_count] int32.  Returns
    (n, ntri, refined)."""
    nn, nt, nr = _i64(n), _i64(ntri), _i64()
    arrs = [np.ascontiguousarray(a) for a in state]
    for a, b in zip(arrs, state):
        assert a is b, "state arrays must be contiguous"
    _check(lib().fss_amr(C.byref(nn), C.byref(nt), max_count, tri_cap, *[_np_ptr(a) for a in state],
                         _np_ptr(area), _np_ptr(tris), _np_ptr(parent), _np_ptr(vert_tris),
                         _np_ptr(vert_tri_count), vert_tris.shape[1], circulation_tol, vorticity_tol,
                         C.byref(nr)))
    return nn.value, nt.value, nr.value


def grid_random(n, seed):
    x, y, z = (np.zeros(n) for _ in range(3))
    _check(lib().fss_grid_random(n, seed, _np_ptr(x), _np_ptr(y), _np_ptr(z)))
    return x, y, z


def rossby_haurwitz(x, y, z, w0=0.5, K=0.5):
    v = np.zeros(len(x))
    tr = np.zeros(len(x))
    _check(lib().fss_field_rossby_haurwitz(len(x), _np_ptr(x), _np_ptr(y), _np_ptr(z), w0, K,
                                           _np_ptr(v), _np_ptr(tr)))
    return v, tr


def gauss_vortices(x, y, z, area, nv=12, amp_std=2.0, rmin=0.05, rmax=0.3, seed=0):
    v = np.zeros(len(x))
    _check(lib().fss_field_gauss_vortices(len(x), _np_ptr(x), _np_ptr(y), _np_ptr(z), _np_ptr(area),
                                          nv, amp_std, rmin, rmax, seed, _np_ptr(v)))
    return v


def sample_targets(n, m, seed):
    idx = np.zeros(m, dtype=np.int64)
    _check(lib().fss_sample_targets(n, m, seed, _np_ptr(idx)))
    return idx


def sal_love_numbers(nmax, seed):
    """Synthetic load Love numbers h'_n, k'_n, n = 0..nmax (entry 0 unused)."""
    h = np.zeros(nmax + 1)
    k = np.zeros(nmax + 1)
    _check(lib().fss_sal_love_numbers(nmax, seed, _np_ptr(h), _np_ptr(k)))
    return h, k


def sal_coefficients(h, k, rho_ratio=1 / 5.5, cesaro=False):
    """c_n = 3 rho (1 + k'_n - h'_n)/(4 pi rho_e) (GFnewtest.ipynb:83-94), n = 0..nmax."""
    nmax = len(h) - 1
    c = np.zeros(nmax + 1)
    _check(lib().fss_sal_coefficients(nmax, _np_ptr(np.ascontiguousarray(h)),
                                      _np_ptr(np.ascontiguousarray(k)), rho_ratio, int(cesaro),
                                      _np_ptr(c)))
    return c


def levels_for(n, leaf_occupancy):
    """Tree depth whose leaves hold leaf_occupancy particles on average."""
    return lib().fss_levels_for(int(n), float(leaf_occupancy))


def sal_mlp_nparam():
    return lib().fss_sal_mlp_nparam()


# ---------------- device context ----------------

class Context:
    """One fss_ctx: configuration (run_information fields), device tree, workspace."""

    def __init__(self, theta=0.7, degree=2, many_count=10, tree_levels=1, omega=2 * np.pi,
                 precision=FP64, assign_mode=ASSIGN_PARITY, eval_mode=EVAL_PARITY,
                 node_rule=NODES_REFERENCE, device=0, stream=None, theta_decay=1.0):
        import torch
        L = lib()
        cfg = Config()
        _check(L.fss_config_default(C.byref(cfg)))
        cfg.theta, cfg.degree, cfg.many_count, cfg.tree_levels = theta, degree, many_count, tree_levels
        cfg.omega, cfg.precision, cfg.assign_mode = omega, precision, assign_mode
        cfg.eval_mode, cfg.node_rule, cfg.device = eval_mode, node_rule, device
        cfg.theta_decay = theta_decay
        self.cfg = cfg
        self.device = torch.device("cuda", device)
        h = C.c_void_p()
        _check(L.fss_create(C.byref(h), C.byref(cfg)))
        self.h = h
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        _check(L.fss_set_stream(self.h, C.c_void_p(s.cuda_stream)))
        self.n = None

    def close(self):
        if getattr(self, "h", None):
            lib().fss_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().fss_sync(self.h))

    def _particles(self, x, y, z, vort, area):
        _check_state(self.device, x, y, z, vort, area)
        return Particles(_dev_ptr(x).value, _dev_ptr(y).value, _dev_ptr(z).value,
                         _dev_ptr(vort).value, _dev_ptr(area).value, x.numel())

    def _out(self, m, out=None):
        import torch
        if out is None:
            out = [torch.empty(m, dtype=torch.float64, device=self.device) for _ in range(4)]
        for t in out:
            assert t.numel() == m and t.dtype == torch.float64 and t.is_contiguous()
        return out, Velocity(*(t.data_ptr() for t in out))

    def tree_build(self, x, y, z):
        _check_state(self.device, x, y, z)
        _check(lib().fss_tree_build(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), x.numel()))
        self.n = x.numel()

    def batch_build(self, x, y, z, sizes, tree_levels=None):
        """Trees for len(sizes) independent instances whose particles are
        concatenated in x, y, z; later calls act on the concatenated state."""
        n = np.ascontiguousarray(sizes, dtype=np.int64)
        assert int(n.sum()) == x.numel()
        _check_state(self.device, x, y, z)
        lv = None if tree_levels is None else np.ascontiguousarray(tree_levels, dtype=np.int32)
        _check(lib().fss_batch_build(self.h, len(n), _np_ptr(n), None if lv is None else _np_ptr(lv),
                                     _dev_ptr(x), _dev_ptr(y), _dev_ptr(z)))
        self.n = x.numel()

    def autotune_levels(self, x, y, z, vort, area, lmin, lmax, reps=3):
        """Pick the fastest tree depth in [lmin, lmax] on these particles;
        the context keeps it.  Returns (best_levels, {levels: ms})."""
        p = self._particles(x, y, z, vort, area)
        ms = np.zeros(lmax - lmin + 1)
        best = C.c_int(0)
        _check(lib().fss_autotune_levels(self.h, C.byref(p), lmin, lmax, reps, C.byref(best), _np_ptr(ms)))
        self.n = x.numel()
        self.cfg.tree_levels = best.value
        return best.value, {lmin + k: float(v) for k, v in enumerate(ms)}

    def autotune(self, x, y, z, vort, area, candidates, idx, max_rel_l2, reps=3):
        """Fastest of `candidates` (dicts with degree, theta, tree_levels,
        many_count, near_tile, optional node_rule and theta_decay) whose rel-L2 against the FP64 direct sum at
        the targets idx is <= max_rel_l2; the context keeps it.  Returns
        (index or -1, per-candidate ms, per-candidate errors)."""
        import torch
        p = self._particles(x, y, z, vort, area)
        arr = (TunePoint * len(candidates))()
        for k, c in enumerate(candidates):
            arr[k] = TunePoint(int(c["degree"]), float(c["theta"]), int(c["tree_levels"]),
                               int(c["many_count"]), int(c.get("near_tile", 0)), int(c.get("node_rule", -1)),
                               float(c.get("theta_decay", 0.0)))
        idx = idx.to(device=self.device, dtype=torch.int64).contiguous()
        ms = np.zeros(len(candidates))
        err = np.zeros(len(candidates))
        best = C.c_int(-2)
        _check(lib().fss_autotune(self.h, C.byref(p), arr, len(candidates), _dev_ptr(idx), idx.numel(),
                                  float(max_rel_l2), int(reps), C.byref(best), _np_ptr(ms), _np_ptr(err)))
        b = best.value
        c = candidates[b if b >= 0 else int(np.argmin(err))]
        self.cfg.degree, self.cfg.theta = int(c["degree"]), float(c["theta"])
        self.cfg.tree_levels, self.cfg.many_count = int(c["tree_levels"]), int(c["many_count"])
        if int(c.get("node_rule", -1)) >= 0:
            self.cfg.node_rule = int(c["node_rule"])
        if float(c.get("theta_decay", 0.0)) > 0:
            self.cfg.theta_decay = float(c["theta_decay"])
        self.n = x.numel()
        return b, ms, err

    def set_near_tile(self, targets_per_lane):
        _check(lib().fss_set_near_tile(self.h, int(targets_per_lane)))

    def stats(self):
        s = TreeStats()
        _check(lib().fss_tree_stats_get(self.h, C.byref(s)))
        return s

    def tree_export(self):
        s = self.stats()
        L, n, nn = s.tree_levels, s.n, s.n_nodes
        T_all = 20 * (4 ** (L + 1) - 1) // 3
        loc = np.zeros(L * n, dtype=np.int32)
        found = np.zeros(L * n, dtype=np.uint8)
        counts = np.zeros(nn, dtype=np.int32)
        offsets = np.zeros(nn, dtype=np.int32)
        _check(lib().fss_tree_export(self.h, None, None, _np_ptr(counts), None, None, None, None, None))
        ids = np.zeros(max(1, int(counts.sum())), dtype=np.int32)
        inter = np.zeros(max(1, 7 * s.n_interactions), dtype=np.int32)
        tri_info = np.zeros(4 * T_all)
        tri_xyz = np.zeros(9 * T_all)
        _check(lib().fss_tree_export(self.h, _np_ptr(loc), _np_ptr(found), _np_ptr(counts),
                                     _np_ptr(offsets), _np_ptr(ids), _np_ptr(inter),
                                     _np_ptr(tri_info), _np_ptr(tri_xyz)))
        return dict(loc=loc.reshape(L, n), found=found.reshape(L, n).astype(bool), counts=counts,
                    offsets=offsets, ids=ids[:int(counts.sum())],
                    inter=inter[:7 * s.n_interactions].reshape(-1, 7),
                    tri_info=tri_info.reshape(T_all, 4), tri_xyz=tri_xyz.reshape(T_all, 3, 3))

    def velocity_direct(self, x, y, z, vort, area, out=None):
        o, ov = self._out(x.numel(), out)
        p = self._particles(x, y, z, vort, area)
        _check(lib().fss_velocity_direct(self.h, C.byref(p), C.byref(ov)))
        return tuple(o)

    def velocity_direct_noalloc(self, x, y, z, vort, area, out=None):
        """velocity_direct with the records in the library's per-GPU scratch
        (no allocation in the call; n at most 2^20 in FP64, 2^21 mixed)."""
        o, ov = self._out(x.numel(), out)
        p = self._particles(x, y, z, vort, area)
        _check(lib().fss_velocity_direct_noalloc(self.h, C.byref(p), C.byref(ov)))
        return tuple(o)

    def velocity_tree(self, x, y, z, vort, area, out=None):
        o, ov = self._out(x.numel(), out)
        p = self._particles(x, y, z, vort, area)
        _check(lib().fss_velocity_tree(self.h, C.byref(p), C.byref(ov)))
        return tuple(o)

    def velocity_direct_targets(self, x, y, z, vort, area, idx):
        import torch
        idx = idx.to(device=self.device, dtype=torch.int64).contiguous()
        o, ov = self._out(idx.numel())
        p = self._particles(x, y, z, vort, area)
        _check(lib().fss_velocity_direct_targets(self.h, C.byref(p), _dev_ptr(idx), idx.numel(),
                                                 C.byref(ov)))
        return tuple(o)

    def rk4_step(self, x, y, z, vort, area, dt, use_tree=False):
        _check_state(self.device, x, y, z, vort, area)
        _check(lib().fss_rk4_step(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), _dev_ptr(vort),
                                  _dev_ptr(area), x.numel(), dt, int(use_tree)))

    def rk4_run_graph(self, x, y, z, vort, area, dt, steps, use_tree=False):
        _check_state(self.device, x, y, z, vort, area)
        _check(lib().fss_rk4_run_graph(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), _dev_ptr(vort),
                                       _dev_ptr(area), x.numel(), dt, int(use_tree), steps))

    def remesh(self, ax, ay, az, avort, tris, vt_off, vt_idx, x, y, z, atracer=None, tri_out=False):
        """Remesh the advected state (ax, ay, az, avort) onto the grid points
        (x, y, z): returns (vort, tracer or None, triangles or None,
        n_fallback).  tris / vt_off / vt_idx are device int32 tensors."""
        import torch
        n = x.numel()
        vort = torch.empty(n, dtype=torch.float64, device=x.device)
        tr = torch.empty_like(vort) if atracer is not None else None
        to = torch.empty(n, dtype=torch.int32, device=x.device) if tri_out else None
        nf, nm = _i64(), _i64()
        _check(lib().fss_remesh(self.h, _dev_ptr(ax), _dev_ptr(ay), _dev_ptr(az), _dev_ptr(avort),
                                _dev_ptr(atracer) if atracer is not None else None, _dev_ptr(tris),
                                tris.numel() // 3, _dev_ptr(vt_off), _dev_ptr(vt_idx), _dev_ptr(x), _dev_ptr(y),
                                _dev_ptr(z), _dev_ptr(vort), _dev_ptr(tr) if tr is not None else None,
                                _dev_ptr(to) if to is not None else None, n, C.byref(nf), C.byref(nm)))
        return vort, tr, to, nf.value

    def project_points(self, x, y, z, vort):
        _check_state(self.device, x, y, z, vort)
        _check(lib().fss_project_points(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), _dev_ptr(vort),
                                        x.numel()))

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().fss_last_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- SAL ----
    def sal_set_series(self, c, table_points=1 << 22, A=0.0):
        c = np.ascontiguousarray(c, dtype=np.float64)
        _check(lib().fss_sal_set_series(self.h, _np_ptr(c), len(c) - 1, table_points, A))

    def sal_kernel(self, s, kind=0):
        import torch
        s = s.to(device=self.device, dtype=torch.float64).contiguous()
        g = torch.empty_like(s)
        _check(lib().fss_sal_kernel(self.h, kind, _dev_ptr(s), s.numel(), _dev_ptr(g)))
        return g

    def sal_direct(self, x, y, z, eta, area, idx=None, kind=0):
        import torch
        n = x.numel()
        if idx is not None:
            idx = idx.to(device=self.device, dtype=torch.int64).contiguous()
            m, ip = idx.numel(), _dev_ptr(idx)
        else:
            m, ip = n, None
        out = torch.empty(m, dtype=torch.float64, device=self.device)
        _check(lib().fss_sal_direct(self.h, _dev_ptr(x), _dev_ptr(y), _dev_ptr(z), _dev_ptr(eta),
                                    _dev_ptr(area), n, ip, m, kind, _dev_ptr(out)))
        return out

    def sal_tree(self, x, y, z, eta, area, kind=0, out=None):
        """SAL convolution at every particle by the treecode (tree built
        with eval_mode BLTC on these particles)."""
        import torch
        p = self._particles(x, y, z, eta, area)
        if out is None:
            out = torch.empty(x.numel(), dtype=torch.float64, device=x.device)
        _check(lib().fss_sal_tree(self.h, C.byref(p), int(kind), _dev_ptr(out)))
        return out

    def sal_mlp_init(self, seed=0):
        _check(lib().fss_sal_mlp_init(self.h, seed))

    def sal_mlp_train(self, steps, samples=1 << 16, lr=1e-3, s_min=1e-3, seed=0):
        loss = C.c_double()
        _check(lib().fss_sal_mlp_train(self.h, steps, samples, lr, s_min, seed, C.byref(loss)))
        return loss.value

    def sal_mlp_params(self, params=None):
        n = sal_mlp_nparam()
        if params is None:
            p = np.zeros(n, dtype=np.float32)
            _check(lib().fss_sal_mlp_params(self.h, _np_ptr(p), 0))
            return p
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.shape == (n,)
        _check(lib().fss_sal_mlp_params(self.h, _np_ptr(p), 1))

    def sal_mlp_norm(self, mu=None, sigma=None, s_min=None):
        """Get (no arguments) or set the learned remainder's normalisation
        (mu, sigma) and the input range's lower end s_min."""
        a, b, c = C.c_double(mu or 0.0), C.c_double(sigma or 1.0), C.c_double(s_min or 1e-3)
        _check(lib().fss_sal_mlp_norm(self.h, C.byref(a), C.byref(b), C.byref(c), int(mu is not None)))
        return a.value, b.value, c.value

    def sal_mlp_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(sal_mlp_nparam(), dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().fss_sal_mlp_grad(self.h, _np_ptr(x), _np_ptr(y), len(x), _np_ptr(g), _np_ptr(loss)))
        return g, float(loss[0])

    def timing_reset(self):
        _check(lib().fss_timing_reset(self.h))

    def timing_mean(self):
        """(calls, mean stage ms, mean main-kernel ms) since timing_reset()."""
        n, a, b = C.c_int64(), C.c_double(), C.c_double()
        _check(lib().fss_timing_mean(self.h, C.byref(n), C.byref(a), C.byref(b)))
        return n.value, a.value, b.value

    def timing_far_mean(self):
        """Mean far-field (proxy value) kernel ms since timing_reset()."""
        f = C.c_double()
        _check(lib().fss_timing_far_mean(self.h, C.byref(f)))
        return f.value
