"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle
(oracle/pdipc_oracle.cpp).  Importable only from tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg; the product never imports it.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of the same source (the mutation tests load
# builds with planted mistakes this way, in a separate process)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

# order of the flat parameter array (enum ParamIndex in pdipc_oracle.cpp)
PARAM_ORDER = ["dt", "gravity", "dhat", "kappa", "lg_iters", "cheb_iters", "max_outer",
               "residual_tol", "omega", "rho", "sigma_lo", "sigma_hi", "ccd_filter", "toi_scale",
               "ccd_iters", "ccd_step", "ccd_gap", "static_mass", "inflation",
               "skip_static_pairs", "touch_policy"]

COUNTER_NAMES = ["outer", "lg", "patch_pairs", "candidates", "active_pairs", "toi_min", "residual",
                 "nnz_r1", "nnz_r2", "n_2paths"]

_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (n <= 0: all processors).
    The results do not depend on the count."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [_D, ctypes.c_int, ctypes.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_add_cloth.argtypes = [vp, _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, _I, ctypes.c_int]
        L.orc_add_tet.argtypes = L.orc_add_cloth.argtypes
        L.orc_add_static.argtypes = [vp, _D, ctypes.c_int, _I, ctypes.c_int, ctypes.c_int]
        for name in ["orc_init"]:
            getattr(L, name).argtypes = [vp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_sizes.argtypes = [vp, _I]
        for name in ["orc_get_x", "orc_set_x", "orc_get_v", "orc_set_v", "orc_counters", "orc_elastic_diag", "orc_mass"]:
            getattr(L, name).argtypes = [vp, _D]
        L.orc_candidates.argtypes = [vp, _I, ctypes.c_int]
        L.orc_patch_pairs.argtypes = [vp, _I, ctypes.c_int]
        L.orc_debug_broad_phase.argtypes = [vp, _D, _D]
        L.orc_debug_toi.argtypes = [vp, _D, _D]
        L.orc_debug_toi.restype = ctypes.c_double
        L.orc_debug_collision.argtypes = [vp, _D, _D, _D]
        L.orc_debug_local.argtypes = [vp, _D, _D, _D]
        L.orc_debug_solve.argtypes = [vp, _D, _D, _D, _D]
        L.orc_model_dat.argtypes = [vp, ctypes.c_int, ctypes.c_int, _I]
        L.orc_sys_matrix.argtypes = [vp, _I, _I, _D]
        L.orc_boundary_faces.argtypes = [vp, _I]
        L.orc_patches.argtypes = [vp, _I]
        L.orc_math_strain_projection.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D]
        L.orc_math_svd.argtypes = [_D, _D, _D, _D]
        L.orc_math_pt_dist2.argtypes = [_D, _D, _D, _D, _D]
        L.orc_math_pt_dist2.restype = ctypes.c_double
        L.orc_math_ee_dist2.argtypes = [_D, _D, _D, _D, _D]
        L.orc_math_ee_dist2.restype = ctypes.c_double
        L.orc_math_ccd.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_math_ccd.restype = ctypes.c_double
        L.orc_math_barrier.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_math_barrier.restype = ctypes.c_double
        L.orc_set_tri_stiffness.argtypes = [vp, ctypes.c_int, _D, ctypes.c_int, _D]
        L.orc_dual_dirs.restype = ctypes.c_int
        L.orc_dual_set.argtypes = [vp, ctypes.c_int, _D, _D]
        L.orc_dual_get.argtypes = [vp, ctypes.c_int, _D, _D]
        _lib = L
    return _lib


def dual_dirs():
    return int(lib().orc_dual_dirs())


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """The CPU restatement of the reference step; `use_float` mimics the
    reference's FP32 build."""

    def __init__(self, scene, use_float=False, dual=False, tri_k_tangents=None):
        """`dual` runs the FP64 step on forward-mode dual numbers carrying
        `dual_dirs()` directional derivatives; `tri_k_tangents` maps a cloth's
        model index to its (dirs, ntri) stiffness seeds."""
        L = lib()
        prm = np.array([float(scene["params"][k]) for k in PARAM_ORDER])
        self.dual = bool(dual)
        self.h = L.orc_create(_dp(prm), len(prm), 2 if dual else (1 if use_float else 0))
        self._keep = []
        order = scene.get("model_order")
        if order is None:  # cloths then tets, unless the scene gives the add order
            order = [("cloth", i) for i in range(len(scene["cloths"]))] + \
                    [("tet", i) for i in range(len(scene["tets"]))]
        for mi, (kind, i) in enumerate(order):
            if kind == "cloth":
                c = scene["cloths"][i]
                v, f, fx = _f64(c["v"]), _i32(c["f"]), _i32(c["fixed"])
                self._keep += [v, f, fx]
                L.orc_add_cloth(self.h, _dp(v), len(v), _ip(f), len(f), c["s0"], c["s1"], c["s2"], c["density"],
                                int(c["patch_dim"]), _ip(fx), len(fx))
                if c.get("tri_k") is not None:
                    k = _f64(c["tri_k"])
                    tg = None
                    if tri_k_tangents is not None and mi in tri_k_tangents:
                        tg = _f64(tri_k_tangents[mi])
                        assert tg.shape == (dual_dirs(), len(f))
                    if L.orc_set_tri_stiffness(self.h, mi, _dp(k), len(k), _dp(tg) if tg is not None else None) != 0:
                        raise ValueError("per-triangle stiffness does not match the cloth")
            else:
                t = scene["tets"][i]
                v, tt, fx = _f64(t["v"]), _i32(t["t"]), _i32(t["fixed"])
                self._keep += [v, tt, fx]
                L.orc_add_tet(self.h, _dp(v), len(v), _ip(tt), len(tt), t["s0"], t["s1"], t["s2"], t["density"],
                              int(t["patch_dim"]), _ip(fx), len(fx))
        for s in scene["statics"]:
            v, f = _f64(s["v"]), _i32(s["f"])
            self._keep += [v, f]
            L.orc_add_static(self.h, _dp(v), len(v), _ip(f), len(f), int(s["patch_dim"]))
        if L.orc_init(self.h) != 0:
            raise RuntimeError("oracle init failed")
        sz = np.zeros(7, np.int32)
        L.orc_sizes(self.h, _ip(sz))
        self.total, self.sim, self.cloth_points, self.tets, self.patches, self.bfaces, self.bedges = map(int, sz)

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    def step(self, n=1):
        lib().orc_step(self.h, n)

    # dual mode: values and tangents of the simulated points, tangents (dirs, sim, 3)
    def dual_set(self, which, val, tangents=None):
        val = _f64(np.asarray(val).reshape(-1)[: 3 * self.sim])
        tg = None
        if tangents is not None:
            tg = _f64(np.asarray(tangents).reshape(dual_dirs(), -1)[:, : 3 * self.sim])
        lib().orc_dual_set(self.h, 0 if which == "x" else 1, _dp(val), _dp(tg) if tg is not None else None)

    def dual_get(self, which):
        val = np.zeros(3 * self.sim)
        tg = np.zeros((dual_dirs(), 3 * self.sim))
        lib().orc_dual_get(self.h, 0 if which == "x" else 1, _dp(val), _dp(tg))
        return val.reshape(-1, 3), tg.reshape(dual_dirs(), -1, 3)

    def x(self):
        out = np.zeros(3 * self.total)
        lib().orc_get_x(self.h, _dp(out))
        return out.reshape(-1, 3)

    def set_x(self, x):
        x = _f64(np.asarray(x).reshape(-1)[: 3 * self.sim])
        lib().orc_set_x(self.h, _dp(x))

    def v(self):
        out = np.zeros(3 * self.sim)
        lib().orc_get_v(self.h, _dp(out))
        return out.reshape(-1, 3)

    def set_v(self, v):
        v = _f64(np.asarray(v).reshape(-1)[: 3 * self.sim])
        lib().orc_set_v(self.h, _dp(v))

    def counters(self):
        out = np.zeros(len(COUNTER_NAMES))
        lib().orc_counters(self.h, _dp(out))
        return dict(zip(COUNTER_NAMES, out.tolist()))

    def candidates(self):
        n = lib().orc_candidates(self.h, None, 0)
        out = np.zeros(5 * max(n, 1), np.int32)
        lib().orc_candidates(self.h, _ip(out), n)
        return out[: 5 * n].reshape(-1, 5)

    def patch_pairs(self):
        n = lib().orc_patch_pairs(self.h, None, 0)
        out = np.zeros(2 * max(n, 1), np.int32)
        lib().orc_patch_pairs(self.h, _ip(out), n)
        return out[: 2 * n].reshape(-1, 2)

    def debug_broad_phase(self, xold, dx):
        a, b = _f64(np.asarray(xold).reshape(-1)), _f64(np.asarray(dx).reshape(-1))
        lib().orc_debug_broad_phase(self.h, _dp(a), _dp(b))
        return self.candidates()

    def debug_toi(self, x0, x1):
        a, b = _f64(np.asarray(x0).reshape(-1)), _f64(np.asarray(x1).reshape(-1))
        return lib().orc_debug_toi(self.h, _dp(a), _dp(b))

    def debug_collision(self, x):
        a = _f64(np.asarray(x).reshape(-1))
        d = np.zeros(self.sim)
        r = np.zeros(3 * self.sim)
        n = lib().orc_debug_collision(self.h, _dp(a), _dp(d), _dp(r))
        return d, r.reshape(-1, 3), n

    def debug_local(self, x, init_rhs):
        a, b = _f64(np.asarray(x).reshape(-1)), _f64(np.asarray(init_rhs).reshape(-1))
        r = np.zeros(3 * self.sim)
        lib().orc_debug_local(self.h, _dp(a), _dp(b), _dp(r))
        return r.reshape(-1, 3)

    def debug_solve(self, diag, rhs, x0):
        d, r, a = _f64(diag), _f64(np.asarray(rhs).reshape(-1)), _f64(np.asarray(x0).reshape(-1))
        out = np.zeros(3 * self.sim)
        lib().orc_debug_solve(self.h, _dp(d), _dp(r), _dp(a), _dp(out))
        return out.reshape(-1, 3)

    def elastic_diag(self):
        out = np.zeros(self.sim)
        lib().orc_elastic_diag(self.h, _dp(out))
        return out

    def mass(self):
        out = np.zeros(self.sim)
        lib().orc_mass(self.h, _dp(out))
        return out

    def sys_matrix(self):
        """sysHess = elastic + M/dt^2 as a dense (sim x sim) array (small scenes only)."""
        nnz = lib().orc_sys_matrix(self.h, None, None, None)
        off = np.zeros(self.sim + 1, np.int32)
        cols = np.zeros(max(nnz, 1), np.int32)
        vals = np.zeros(max(nnz, 1))
        lib().orc_sys_matrix(self.h, _ip(off), _ip(cols), _dp(vals))
        A = np.zeros((self.sim, self.sim))
        for i in range(self.sim):
            A[i, cols[off[i]:off[i + 1]]] = vals[off[i]:off[i + 1]]
        return A

    def boundary_faces(self):
        n = lib().orc_boundary_faces(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_boundary_faces(self.h, _ip(out))
        return out[:n].reshape(-1, 3)

    def patch_lists(self):
        n = lib().orc_patches(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_patches(self.h, _ip(out))
        P, pos, res = int(out[0]), 1, []
        for _ in range(P):
            k = int(out[pos])
            res.append(out[pos + 1: pos + 1 + k].copy())
            pos += 1 + k
        return res

    def model_dat(self, model_index, which):
        """int32 image of the model's boundary.dat (which=0) or patchInfo.dat (which=1)."""
        n = lib().orc_model_dat(self.h, model_index, which, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_model_dat(self.h, model_index, which, _ip(out))
        return out[:n]


# ---- pure-math entry points -------------------------------------------------
def strain_projection(F, lo=0.999, hi=1.001):
    F = _f64(np.asarray(F).reshape(9))
    P = np.zeros(9)
    lib().orc_math_strain_projection(_dp(F), lo, hi, _dp(P))
    return P.reshape(3, 3)


def svd(F):
    F = _f64(np.asarray(F).reshape(9))
    U, s, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().orc_math_svd(_dp(F), _dp(U), _dp(s), _dp(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def pt_dist2(p, a, b, c):
    bary = np.zeros(3)
    d = lib().orc_math_pt_dist2(*[_dp(_f64(q)) for q in (p, a, b, c)], _dp(bary))
    return d, bary


def ee_dist2(a0, a1, b0, b1):
    rs = np.zeros(2)
    d = lib().orc_math_ee_dist2(*[_dp(_f64(q)) for q in (a0, a1, b0, b1)], _dp(rs))
    return d, rs


def ccd(typ, x, dx, touch_policy=1, iters=10, step=0.9, gap=0.1):
    """ACCD time of impact of one record (the reference constants by default)."""
    x, dx = _f64(np.asarray(x).reshape(12)), _f64(np.asarray(dx).reshape(12))
    return lib().orc_math_ccd(typ, _dp(x), _dp(dx), touch_policy, iters, step, gap)


def barrier(d2, dh2, kappa):
    return lib().orc_math_barrier(d2, dh2, kappa)
