"""ctypes binding of libpdipc.so (include/pdipc.h).

`Simulator` mirrors the reference's PD_IPC::PdIpcSimulator life cycle
(Simulator/PD-IPC/PdIpcSimulator.h): add models, initialization(), then run()
per frame, postRun() to read positions back, saveSimulator() to dump a frame.
It raises when the HIP library is missing or a call fails; there is no
fallback path.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpdipc.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pdipc.h")

FP64 = 0
FP32_SPLIT = 1


class Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("gravity", ctypes.c_double), ("dhat", ctypes.c_double),
                ("kappa", ctypes.c_double), ("lg_iters", ctypes.c_int), ("cheb_iters", ctypes.c_int),
                ("max_outer", ctypes.c_int), ("residual_tol", ctypes.c_double), ("omega", ctypes.c_double),
                ("rho", ctypes.c_double), ("sigma_lo", ctypes.c_double), ("sigma_hi", ctypes.c_double),
                ("ccd_filter", ctypes.c_double), ("toi_scale", ctypes.c_double), ("ccd_iters", ctypes.c_int),
                ("ccd_step", ctypes.c_double), ("ccd_gap", ctypes.c_double), ("static_mass", ctypes.c_double),
                ("inflation", ctypes.c_double), ("skip_static_pairs", ctypes.c_int), ("touch_policy", ctypes.c_int),
                ("precision", ctypes.c_int), ("sell_lanes", ctypes.c_int), ("use_graphs", ctypes.c_int),
                ("sell_stage", ctypes.c_int), ("capacity_hint", ctypes.c_int), ("track_toi_argmin", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [("outer_iters", ctypes.c_int), ("lg_iters", ctypes.c_int), ("patch_pairs", ctypes.c_int),
                ("candidates", ctypes.c_int), ("active_pairs", ctypes.c_int), ("toi_min", ctypes.c_double),
                ("residual", ctypes.c_double), ("nnz_r1", ctypes.c_longlong), ("nnz_r2", ctypes.c_longlong),
                ("n_2paths", ctypes.c_longlong), ("sell_entries", ctypes.c_longlong), ("sim_points", ctypes.c_int),
                ("total_points", ctypes.c_int), ("cloth_points", ctypes.c_int), ("tets", ctypes.c_int),
                ("patches", ctypes.c_int), ("boundary_faces", ctypes.c_int), ("boundary_edges", ctypes.c_int),
                ("toi_argmins", ctypes.c_longlong), ("active_sets", ctypes.c_longlong)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_H = ctypes.c_void_p

# every entry point declared in include/pdipc.h: name -> (restype, argtypes)
SIGNATURES = {
    "pdipc_default_params": (None, [ctypes.POINTER(Params)]),
    "pdipc_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_H)]),
    "pdipc_destroy": (ctypes.c_int, [_H]),
    "pdipc_last_error": (ctypes.c_char_p, []),
    "pdipc_add_cloth": (ctypes.c_int, [_H, _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, _I, ctypes.c_int]),
    "pdipc_add_tet": (ctypes.c_int, [_H, _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int, _I, ctypes.c_int]),
    "pdipc_add_static": (ctypes.c_int, [_H, _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_int]),
    "pdipc_host_setup": (ctypes.c_int, [_H]),
    "pdipc_init": (ctypes.c_int, [_H]),
    "pdipc_step": (ctypes.c_int, [_H, ctypes.c_int]),
    "pdipc_set_precision": (ctypes.c_int, [_H, ctypes.c_int]),
    "pdipc_get_x": (ctypes.c_int, [_H, _D, ctypes.c_int]),
    "pdipc_set_x": (ctypes.c_int, [_H, _D, ctypes.c_int]),
    "pdipc_get_v": (ctypes.c_int, [_H, _D, ctypes.c_int]),
    "pdipc_set_v": (ctypes.c_int, [_H, _D, ctypes.c_int]),
    "pdipc_get_counters": (ctypes.c_int, [_H, ctypes.POINTER(Counters)]),
    "pdipc_save_frame": (ctypes.c_int, [_H, ctypes.c_char_p, ctypes.c_int]),
    "pdipc_save_state": (ctypes.c_int, [_H, ctypes.c_char_p]),
    "pdipc_load_state": (ctypes.c_int, [_H, ctypes.c_char_p]),
    "pdipc_get_candidates": (ctypes.c_int, [_H, _I, ctypes.c_int, _I]),
    "pdipc_get_patch_pairs": (ctypes.c_int, [_H, _I, ctypes.c_int, _I]),
    "pdipc_get_model_dat": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, _I]),
    "pdipc_debug_broad_phase": (ctypes.c_int, [_H, _D, _D, _I]),
    "pdipc_debug_toi": (ctypes.c_int, [_H, _D, _D, _D]),
    "pdipc_kernel_timing": (ctypes.c_int, [_H, ctypes.c_int]),
    "pdipc_get_kernel_times": (ctypes.c_int, [_H, _D, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]),
    "pdipc_toi_noalloc": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pdipc_debug_collision": (ctypes.c_int, [_H, _D, _D, _D, _I]),
    "pdipc_debug_local": (ctypes.c_int, [_H, _D, _D, _D]),
    "pdipc_debug_solve": (ctypes.c_int, [_H, _D, _D, _D, _D]),
    "pdipc_debug_lbvh": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_ulonglong), _I, _I, _I, _I]),
    "pdipc_debug_stage_layout": (ctypes.c_int, [_H, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _I, _I,
                                                ctypes.POINTER(ctypes.c_ushort), ctypes.POINTER(ctypes.c_ushort)]),
    "pdipc_debug_narrow_stage_layout": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_longlong), _I, _I,
                                                       ctypes.POINTER(ctypes.c_ushort),
                                                       ctypes.POINTER(ctypes.c_ubyte)]),
    "pdipc_sweep_staging": (ctypes.c_int, [_H, _I]),
    "pdipc_debug_sell_layout": (ctypes.c_int, [_H, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                               ctypes.POINTER(ctypes.c_longlong), _I, _I, _I,
                                               ctypes.POINTER(ctypes.c_short), _I]),
    "pdipc_count_intersections": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_longlong)]),
    "pdipc_time_kernel": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, _D, _D]),
    "pdipc_set_cloth_stiffness": (ctypes.c_int, [_H, ctypes.c_int, _D, ctypes.c_int]),
    "pdipc_rollout": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int]),
    "pdipc_backward": (ctypes.c_int, [_H, _D, _D, _D, _D, _D, ctypes.c_int]),
    "pdipc_adjoint_stats": (ctypes.c_int, [_H, _D]),
    "pdipc_debug_solve_vjp": (ctypes.c_int, [_H, _D, _D, _D, _D, _D, _D]),
    "pdipc_debug_local_vjp": (ctypes.c_int, [_H, _D, _D, _D]),
    "pdipc_debug_tape": (ctypes.c_int, [_H, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
}

KERNEL_IDS = {"sweep_fp64": 0, "sweep_fp32": 1, "local": 2, "ccd": 3, "operators": 4}
# classes of pdipc_get_kernel_times, in order
TIMER_CLASSES = ["sweep_fp64", "sweep_fp32", "rhs", "project", "first_residual", "operators", "ccd", "barrier",
                 "broad_phase"]

_lib = None


class PdipcError(RuntimeError):
    pass


def lib():
    """Loads the in-tree HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PdipcError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().pdipc_last_error()
        raise PdipcError(f"pdipc error {rc}: {msg.decode() if msg else ''}")


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def default_params(**overrides):
    p = Params()
    lib().pdipc_default_params(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def model_order(scene):
    """The simulated models in add order: scene["model_order"] or cloths then tets."""
    order = scene.get("model_order")
    if order is None:
        order = [("cloth", i) for i in range(len(scene["cloths"]))] + [("tet", i) for i in range(len(scene["tets"]))]
    return list(order)


def params_from_scene(scene, **extra):
    p = default_params()
    for k, v in scene["params"].items():
        if hasattr(p, k):
            setattr(p, k, type(getattr(p, k))(v))
    for k, v in extra.items():
        setattr(p, k, v)
    return p


class Simulator:
    """One simulator handle (PdIpcSimulator)."""

    def __init__(self, params=None):
        self.h = _H()
        self.params = params if params is not None else default_params()
        _check(lib().pdipc_create(ctypes.byref(self.params), ctypes.byref(self.h)))
        self._keep = []
        self.total = self.sim = 0

    @classmethod
    def from_scene(cls, scene, init=True, **extra):
        """Simulated models in scene["model_order"] (("cloth" | "tet", index)
        pairs, the reference's add order) when given, else cloths then tets;
        static models last."""
        s = cls(params_from_scene(scene, **extra))
        for mi, (kind, i) in enumerate(model_order(scene)):
            if kind == "cloth":
                c = scene["cloths"][i]
                s.add_cloth(c["v"], c["f"], c["s0"], c["s1"], c["s2"], c["density"], c["patch_dim"], c["fixed"])
                if c.get("tri_k") is not None:
                    s.set_cloth_stiffness(mi, c["tri_k"])
            else:
                t = scene["tets"][i]
                s.add_tet(t["v"], t["t"], t["s0"], t["s1"], t["s2"], t["density"], t["patch_dim"], t["fixed"])
        for st in scene["statics"]:
            s.add_static(st["v"], st["f"], st["patch_dim"])
        if init:
            s.initialization()
        return s

    def close(self):
        if self.h:
            lib().pdipc_destroy(self.h)
            self.h = _H()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- addModelFromConfigFile / addStaticModelFromConfigFile
    def add_cloth(self, v, f, strain, bend, pos, density, patch_dim, fixed=()):
        v = np.ascontiguousarray(v, np.float64)
        f = np.ascontiguousarray(f, np.int32)
        fx = np.ascontiguousarray(fixed, np.int32)
        self._keep += [v, f, fx]
        _check(lib().pdipc_add_cloth(self.h, _dp(v), len(v), _ip(f), len(f), strain, bend, pos, density,
                                     int(patch_dim), _ip(fx), len(fx)))

    def add_tet(self, v, t, strain, volume, pos, density, patch_dim, fixed=()):
        v = np.ascontiguousarray(v, np.float64)
        t = np.ascontiguousarray(t, np.int32)
        fx = np.ascontiguousarray(fixed, np.int32)
        self._keep += [v, t, fx]
        _check(lib().pdipc_add_tet(self.h, _dp(v), len(v), _ip(t), len(t), strain, volume, pos, density,
                                   int(patch_dim), _ip(fx), len(fx)))

    def add_static(self, v, f, patch_dim):
        v = np.ascontiguousarray(v, np.float64)
        f = np.ascontiguousarray(f, np.int32)
        self._keep += [v, f]
        _check(lib().pdipc_add_static(self.h, _dp(v), len(v), _ip(f), len(f), int(patch_dim)))

    # --- initialization()
    def host_setup(self):
        _check(lib().pdipc_host_setup(self.h))
        c = self.counters()
        self.total, self.sim = c["total_points"], c["sim_points"]

    def initialization(self):
        _check(lib().pdipc_init(self.h))
        c = self.counters()
        self.total, self.sim = c["total_points"], c["sim_points"]

    # --- run(frame) -> advance
    def run(self, nsteps=1):
        _check(lib().pdipc_step(self.h, int(nsteps)))

    step = run

    def set_precision(self, precision):
        _check(lib().pdipc_set_precision(self.h, int(precision)))

    # --- postRun / state
    def x(self):
        out = np.zeros((self.total, 3))
        _check(lib().pdipc_get_x(self.h, _dp(out), self.total))
        return out

    post_run = x

    def set_x(self, x):
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 3)[: self.sim])
        _check(lib().pdipc_set_x(self.h, _dp(x), len(x)))

    def v(self):
        out = np.zeros((self.sim, 3))
        _check(lib().pdipc_get_v(self.h, _dp(out), self.sim))
        return out

    def set_v(self, v):
        v = np.ascontiguousarray(np.asarray(v, np.float64).reshape(-1, 3)[: self.sim])
        _check(lib().pdipc_set_v(self.h, _dp(v), len(v)))

    def counters(self):
        c = Counters()
        _check(lib().pdipc_get_counters(self.h, ctypes.byref(c)))
        return c.as_dict()

    def save_simulator(self, path, as_double=False):
        _check(lib().pdipc_save_frame(self.h, path.encode(), 1 if as_double else 0))

    def save_state(self, path):
        """Full-state checkpoint: x and v in FP64 with the scene's topology fingerprint."""
        _check(lib().pdipc_save_state(self.h, os.fspath(path).encode()))

    def load_state(self, path):
        """Restores a save_state file of the same scene (raises for another scene)."""
        _check(lib().pdipc_load_state(self.h, os.fspath(path).encode()))

    def candidates(self):
        n = ctypes.c_int()
        _check(lib().pdipc_get_candidates(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(5 * max(n.value, 1), np.int32)
        _check(lib().pdipc_get_candidates(self.h, _ip(out), n.value, ctypes.byref(n)))
        return out[: 5 * n.value].reshape(-1, 5)

    def patch_pairs(self):
        n = ctypes.c_int()
        _check(lib().pdipc_get_patch_pairs(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(2 * max(n.value, 1), np.int32)
        _check(lib().pdipc_get_patch_pairs(self.h, _ip(out), n.value, ctypes.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    def sell_layout(self, lanes=8):
        """The SELL layout of M^2 the device upload builds for `lanes` lanes per
        row, from the host precompute (after host_setup; no device needed)."""
        sz = np.zeros(6, np.int64)
        szp = sz.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        _check(lib().pdipc_debug_sell_layout(self.h, lanes, szp, None, None, None, None, None, None))
        ns, ne, uw, c16, nperm, rows = (int(v) for v in sz)
        off = np.zeros(ns + 1, np.int64)
        w = np.zeros(max(ns, 1), np.int32)
        col = np.zeros(max(ne, 1), np.int32)
        flag = np.zeros(max(ne, 1), np.int32)
        col16 = np.zeros(max(ne, 1), np.int16)
        perm = np.zeros(max(nperm, 1), np.int32)
        _check(lib().pdipc_debug_sell_layout(self.h, lanes, szp, off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                             _ip(w), _ip(col), _ip(flag),
                                             col16.ctypes.data_as(ctypes.POINTER(ctypes.c_short)), _ip(perm)))
        return dict(slice_off=off, slice_w=w[:ns], col=col[:ne], row_flag=flag[:ne],
                    col16=col16[:ne] if c16 else None, perm=perm[:nperm] if nperm else None,
                    uniform_w=uw, rows_per_slice=rows)

    def stage_layout(self, lanes=8):
        """The sweeps' LDS staging tables for `lanes` lanes per row (host
        precompute only), or None when staging does not apply."""
        sz = np.zeros(6, np.int64)
        szp = sz.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        _check(lib().pdipc_debug_stage_layout(self.h, lanes, szp, None, None, None, None))
        ok, nt, nu, umax, ne, ts = (int(v) for v in sz)
        if not ok:
            return None
        off = np.zeros(nt + 1, np.int32)
        base = np.zeros(nt, np.int32)
        tcol = np.zeros(max(nu, 1), np.uint16)
        loc = np.zeros(max(ne, 1), np.uint16)
        us = ctypes.POINTER(ctypes.c_ushort)
        _check(lib().pdipc_debug_stage_layout(self.h, lanes, szp, _ip(off), _ip(base), tcol.ctypes.data_as(us),
                                              loc.ctypes.data_as(us)))
        return dict(tile_off=off, tile_base=base, tile_col=tcol[:nu], loc=loc[:ne], stage_max=umax,
                    tile_slices=ts)

    def narrow_stage_layout(self):
        """The sweeps' narrow staging tables (8 lanes per row, one-byte
        indices; host precompute only), or None when they do not apply."""
        sz = np.zeros(6, np.int64)
        szp = sz.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        _check(lib().pdipc_debug_narrow_stage_layout(self.h, szp, None, None, None, None))
        ok, nt, nu, umax, ne, ts = (int(v) for v in sz)
        if not ok:
            return None
        off = np.zeros(nt + 1, np.int32)
        base = np.zeros(nt, np.int32)
        tcol = np.zeros(max(nu, 1), np.uint16)
        loc = np.zeros(max(ne, 1), np.uint8)
        _check(lib().pdipc_debug_narrow_stage_layout(self.h, szp, _ip(off), _ip(base),
                                                     tcol.ctypes.data_as(ctypes.POINTER(ctypes.c_ushort)),
                                                     loc.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        return dict(tile_off=off, tile_base=base, tile_col=tcol[:nu], loc=loc[:ne], stage_max=umax,
                    tile_slices=ts)

    def sweep_staging(self):
        """How the sweeps read their operand: 0 L2 gathers, 1 staged in LDS
        (16-bit indices), 2 narrow staging (one-byte indices)."""
        v = ctypes.c_int()
        _check(lib().pdipc_sweep_staging(self.h, ctypes.byref(v)))
        return v.value

    def model_dat(self, model, which):
        n = ctypes.c_int()
        _check(lib().pdipc_get_model_dat(self.h, model, which, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.int32)
        _check(lib().pdipc_get_model_dat(self.h, model, which, _ip(out), n.value, ctypes.byref(n)))
        return out[: n.value]

    # --- teacher-forced phases
    def debug_broad_phase(self, xold, dx):
        a = np.ascontiguousarray(np.asarray(xold, np.float64).reshape(-1))
        b = np.ascontiguousarray(np.asarray(dx, np.float64).reshape(-1))
        n = ctypes.c_int()
        _check(lib().pdipc_debug_broad_phase(self.h, _dp(a), _dp(b), ctypes.byref(n)))
        return self.candidates()

    def debug_toi(self, x0, x1):
        a = np.ascontiguousarray(np.asarray(x0, np.float64).reshape(-1))
        b = np.ascontiguousarray(np.asarray(x1, np.float64).reshape(-1))
        t = ctypes.c_double()
        _check(lib().pdipc_debug_toi(self.h, _dp(a), _dp(b), ctypes.byref(t)))
        return t.value

    def toi_noalloc(self, x0_dev, x1_dev, toi_dev, stream=0):
        """pdipc_toi_noalloc on device addresses (ints): AoS positions x0, x1
        and one double for the result; enqueued on `stream` (0: the handle's)."""
        _check(lib().pdipc_toi_noalloc(self.h, ctypes.c_void_p(x0_dev), ctypes.c_void_p(x1_dev),
                                       ctypes.c_void_p(toi_dev), ctypes.c_void_p(stream or None)))

    def debug_collision(self, x):
        a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
        d = np.zeros(self.sim)
        r = np.zeros((self.sim, 3))
        n = ctypes.c_int()
        _check(lib().pdipc_debug_collision(self.h, _dp(a), _dp(d), _dp(r), ctypes.byref(n)))
        return d, r, n.value

    def debug_local(self, x, init_rhs):
        a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
        b = np.ascontiguousarray(np.asarray(init_rhs, np.float64).reshape(-1))
        r = np.zeros((self.sim, 3))
        _check(lib().pdipc_debug_local(self.h, _dp(a), _dp(b), _dp(r)))
        return r

    def debug_solve(self, diag, rhs, x0):
        d = np.ascontiguousarray(diag, np.float64)
        r = np.ascontiguousarray(np.asarray(rhs, np.float64).reshape(-1))
        a = np.ascontiguousarray(np.asarray(x0, np.float64).reshape(-1))
        out = np.zeros((self.sim, 3))
        _check(lib().pdipc_debug_solve(self.h, _dp(d), _dp(r), _dp(a), _dp(out)))
        return out

    def kernel_timing(self, enable=True):
        """In-step kernel timing on (direct launches, each timed kernel's own
        start/stop recorded) or off; turning it on again starts from zero."""
        _check(lib().pdipc_kernel_timing(self.h, 1 if enable else 0))

    def kernel_times(self):
        """{class: (summed microseconds, launches)} since timing was enabled."""
        n = len(TIMER_CLASSES)
        us = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        _check(lib().pdipc_get_kernel_times(self.h, _dp(us), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), n))
        return {name: (float(us[i]), int(cnt[i])) for i, name in enumerate(TIMER_CLASSES)}

    def time_kernel(self, name, reps=50):
        """(mean microseconds per launch, algorithmic bytes per launch) of one kernel."""
        us, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().pdipc_time_kernel(self.h, KERNEL_IDS[name], int(reps), ctypes.byref(us), ctypes.byref(b)))
        return us.value, b.value

    # --- per-triangle stiffness and the reverse-mode derivative
    def set_cloth_stiffness(self, model, k):
        k = np.ascontiguousarray(k, np.float64)
        self._keep.append(k)
        _check(lib().pdipc_set_cloth_stiffness(self.h, int(model), _dp(k), len(k)))

    def rollout(self, nsteps, checkpoint_every=10):
        """`nsteps` FP64 steps for a later backward: (x, v) kept every
        `checkpoint_every` steps, or (0) every step's full tape kept in HBM."""
        _check(lib().pdipc_rollout(self.h, int(nsteps), int(checkpoint_every)))

    def backward(self, dx=None, dv=None, n_tris=None):
        """dL/dx, dL/dv of the rolled-out final state -> (dL/dx0, dL/dv0, dL/dk)
        for the initial state and every cloth triangle's stiffness."""
        keep = [np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, 3)[: self.sim]) for a in (dx, dv)
                if a is not None]
        sx = _dp(keep[0]) if dx is not None else None
        sv = _dp(keep[-1]) if dv is not None else None
        dx0 = np.zeros((self.sim, 3))
        dv0 = np.zeros((self.sim, 3))
        nk = int(n_tris) if n_tris is not None else 0
        dk = np.zeros(max(nk, 1))
        _check(lib().pdipc_backward(self.h, sx, sv, _dp(dx0), _dp(dv0), _dp(dk) if nk else None, nk))
        return dx0, dv0, dk[:nk]

    def debug_solve_vjp(self, diag, rhs, x0, xbar_out):
        d = np.ascontiguousarray(diag, np.float64)
        r = np.ascontiguousarray(np.asarray(rhs, np.float64).reshape(-1))
        a = np.ascontiguousarray(np.asarray(x0, np.float64).reshape(-1))
        g = np.ascontiguousarray(np.asarray(xbar_out, np.float64).reshape(-1))
        xb = np.zeros((self.sim, 3))
        bb = np.zeros((self.sim, 3))
        _check(lib().pdipc_debug_solve_vjp(self.h, _dp(d), _dp(r), _dp(a), _dp(g), _dp(xb), _dp(bb)))
        return xb, bb

    def debug_local_vjp(self, x, bbar):
        a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
        g = np.ascontiguousarray(np.asarray(bbar, np.float64).reshape(-1))
        xb = np.zeros((self.sim, 3))
        _check(lib().pdipc_debug_local_vjp(self.h, _dp(a), _dp(g), _dp(xb)))
        return xb

    def debug_tape(self, step, outer, which, l=0, k=0):
        out = np.zeros(self.sim if which == 2 else (4 if which == 5 else 3 * self.sim))
        _check(lib().pdipc_debug_tape(self.h, step, outer, which, l, k, _dp(out)))
        return out if which in (2, 5) else out.reshape(-1, 3)

    def adjoint_stats(self):
        out = np.zeros(4)
        _check(lib().pdipc_adjoint_stats(self.h, _dp(out)))
        return dict(outer_iters=int(out[0]), tape_bytes=float(out[1]), segments=int(out[2]), steps=int(out[3]))

    def count_intersections(self):
        """FP64 edge-triangle crossings at the current positions (0: no two
        triangles intersect)."""
        n = ctypes.c_longlong()
        _check(lib().pdipc_count_intersections(self.h, ctypes.byref(n)))
        return n.value

    def debug_lbvh(self):
        n = ctypes.c_int()
        _check(lib().pdipc_debug_lbvh(self.h, None, None, None, None, ctypes.byref(n)))
        P = n.value
        keys = np.zeros(max(P, 1), np.uint64)
        left = np.zeros(max(P - 1, 1), np.int32)
        right = np.zeros(max(P - 1, 1), np.int32)
        parent = np.zeros(max(2 * P - 1, 1), np.int32)
        _check(lib().pdipc_debug_lbvh(self.h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)), _ip(left),
                                      _ip(right), _ip(parent), ctypes.byref(n)))
        return keys[:P], left[: max(P - 1, 0)], right[: max(P - 1, 0)], parent[: max(2 * P - 1, 0)]
