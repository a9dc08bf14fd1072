"""ctypes bindings of the CPU oracle (gf_oracle.c) and of the reference library
built from source (oracle/_ref/libnbody_ref.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module; the product never does.
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libnbody_ref.so")

_dbl = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_ci, _cd, _cf = ctypes.c_int, ctypes.c_double, ctypes.c_float


def build():
    """Compiles liboracle.so (and _ref/ where the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_accel_active.restype = ctypes.c_longlong
        L.orc_accel_active.argtypes = [_ci, _dbl, _dbl, _dbl, _cd, _cd, _vp, _ci, _ci]
        L.orc_accel_direct.argtypes = [_ci, _dbl, _dbl, _ci, _vp, _dbl, _cd]
        L.orc_tree_export.restype = _ci
        L.orc_tree_export.argtypes = [_vp] * 13
        L.orc_kernel.restype = _cd
        L.orc_kernel.argtypes = [_cd, _cd]
        L.orc_kernel_grad.restype = _cd
        L.orc_kernel_grad.argtypes = [_cd, _cd]
        L.orc_sph_forces.argtypes = [_ci, _ci, _dbl, _dbl, _dbl, _i64, _dbl, _dbl, _cd, _cd, _cd,
                                     _dbl, _dbl, _dbl, _dbl, _vp]
        L.orc_knn_brute.argtypes = [_ci, _dbl, _ci, _i64, _dbl]
        L.orc_knn_from_candidates.restype = _ci
        L.orc_knn_from_candidates.argtypes = [_ci, _ci, _dbl, _i64, _ci, _i64, _dbl]
        L.orc_num_threads.restype = _ci
        L.orc_render.argtypes = [_ci, _dbl, _u8, _ci, _dbl, _dbl, _cd, _ci, _ci, _f32]
        L.orc_tonemap.argtypes = [_ci, _ci, _ci, _f32, _f32, _ci, _ci, _f32, _cf, _cf, _cf, _u8]
        _lib = L
    return _lib


def ref_lib():
    """The reference engine built from its own source; None if it was not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        L = ctypes.CDLL(REF_LIB)
        L.nbody_accel_active.argtypes = [_ci, _dbl, _dbl, _dbl, _cd, _cd, _u8, _ci]
        L.nbody_accel_direct.argtypes = [_ci, _dbl, _dbl, _dbl, _cd]
        L.nbody_num_threads.restype = _ci
        L.ref_tree_export.restype = _ci
        L.ref_tree_export.argtypes = [_vp] * 9
        L.nbody_render.argtypes = [_ci, _dbl, _u8, _ci, _dbl, _dbl, _cd, _ci, _ci, _f32]
        L.nbody_tonemap.argtypes = [_ci, _ci, _ci, _f32, _f32, _ci, _ci, _f32, _cf, _cf, _cf,
                                    _u8]
        _ref = L
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


# ------------------------------------------------------------------ gravity

def accel(pos, mass, theta, eps, active=None, rebuild=True, mode=1, acc=None):
    """Barnes-Hut accelerations (cpp/nbody.cpp:323-411 restated).  mode 0 is the
    reference FP32 arithmetic, mode 1 the FP64 one.  Returns (acc, interactions)."""
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    if acc is None:
        acc = np.zeros_like(pos)
    act = None if active is None else np.ascontiguousarray(active, np.uint8)
    inter = lib().orc_accel_active(len(mass), pos, mass, acc, float(theta), float(eps),
                                   _p(act), int(rebuild), int(mode))
    return acc, int(inter)


def direct(pos, mass, eps, targets=None):
    """FP64 direct sum (cpp/nbody.cpp:428-441) for all particles or a subset."""
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    if targets is None:
        nt, tg = len(mass), None
    else:
        tg = np.ascontiguousarray(targets, np.int64)
        nt = len(tg)
    acc = np.zeros((nt, 3))
    lib().orc_accel_direct(len(mass), pos, mass, nt, _p(tg), acc, float(eps))
    return acc


def tree_export():
    """Tree of the last oracle build/refit, DFS order (nbody.cpp layout)."""
    L = lib()
    nn = L.orc_tree_export(*([None] * 13))
    n = int(_tree_n())
    out = dict(start=np.empty(nn, np.int32), count=np.empty(nn, np.int32),
               next=np.empty(nn, np.int32), leaf=np.empty(nn, np.int32),
               depth=np.empty(nn, np.int32), com4=np.empty((nn, 4), np.float32),
               open_r=np.empty(nn, np.float32), dcom4=np.empty((nn, 4)),
               dopen_r=np.empty(nn), box4=np.empty((nn, 4)), idx=np.empty(n, np.int32),
               groups=np.empty(nn + 1, np.int32))
    ng = ctypes.c_int(0)
    L.orc_tree_export(*[_p(out[k]) for k in ("start", "count", "next", "leaf", "depth", "com4",
                                               "open_r", "dcom4", "dopen_r", "box4", "idx",
                                               "groups")], ctypes.byref(ng))
    out["groups"] = out["groups"][: ng.value]
    return out


def _tree_n():
    # n of the last build = count of the root node
    nn = lib().orc_tree_export(*([None] * 13))
    c = np.empty(nn, np.int32)
    lib().orc_tree_export(None, _p(c), *([None] * 11))
    return c[0] if nn else 0


def ref_accel(pos, mass, theta, eps, active=None, rebuild=True, acc=None):
    """The reference nbody_accel_active (built from source), or None if absent."""
    L = ref_lib()
    if L is None:
        return None
    pos = np.ascontiguousarray(pos, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    if acc is None:
        acc = np.zeros_like(pos)
    act = np.ones(len(mass), np.uint8) if active is None else np.ascontiguousarray(active, np.uint8)
    L.nbody_accel_active(len(mass), pos, mass, acc, float(theta), float(eps), act, int(rebuild))
    return acc


def ref_tree_export(n):
    L = ref_lib()
    nn = L.ref_tree_export(*([None] * 9))
    out = dict(start=np.empty(nn, np.int32), count=np.empty(nn, np.int32),
               next=np.empty(nn, np.int32), leaf=np.empty(nn, np.int32),
               com4=np.empty((nn, 4), np.float32), open_r=np.empty(nn, np.float32),
               idx=np.empty(n, np.int32), groups=np.empty(nn + 1, np.int32))
    ng = ctypes.c_int(0)
    L.ref_tree_export(*[_p(out[k]) for k in ("start", "count", "next", "leaf", "com4", "open_r",
                                               "idx", "groups")], ctypes.byref(ng))
    out["groups"] = out["groups"][: ng.value]
    return out


# ---------------------------------------------------------------------- SPH

def kernel(r, h):
    """sph.py:31-40 restated (scalar)."""
    return lib().orc_kernel(float(r), float(h))


def kernel_grad(r, h):
    """sph.py:43-54 restated (scalar)."""
    return lib().orc_kernel_grad(float(r), float(h))


def sph_forces(pos, vel, mass, nb, h, u, gamma, alpha, beta):
    """sph.py:57-106 restated in FP64.  Returns rho, acc, dudt, dt_i, divv."""
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    mass = np.ascontiguousarray(mass, np.float64)
    nb = np.ascontiguousarray(nb, np.int64)
    h = np.ascontiguousarray(h, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    n, k = nb.shape
    rho, dudt, dt_i, divv = (np.empty(n) for _ in range(4))
    acc = np.empty((n, 3))
    lib().orc_sph_forces(n, k, pos, vel, mass, nb, h, u, float(gamma), float(alpha), float(beta),
                         rho, acc, dudt, dt_i, _p(divv))
    return rho, acc, dudt, dt_i, divv


def knn(pos, k):
    """Exact kNN by brute force (sph.py:165-166 semantics): (dist, nb), self first."""
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    nb = np.empty((n, k), np.int64)
    d = np.empty((n, k))
    lib().orc_knn_brute(n, pos, k, nb, d)
    return d, nb


def knn_from_candidates(pos, cand, k):
    """sph.py:109-135 restated: returns (ok, nb, dist)."""
    pos = np.ascontiguousarray(pos, np.float64)
    cand = np.ascontiguousarray(cand, np.int64)
    n, m = cand.shape
    nb = np.zeros((n, k), np.int64)
    d = np.zeros((n, k))
    ok = lib().orc_knn_from_candidates(n, m, pos, cand, k, nb, d)
    return bool(ok), nb, d


def smoothing_length(dk, eps):
    """simulation.py:84: h = max(d_k / 2, eps / 2)."""
    return np.maximum(0.5 * dk, 0.5 * eps)


def render(pos, group, ngroups, rot, center, scale, W, H):
    """cpp/nbody.cpp:445-469 restated: (ngroups, H, W) float32 deposit."""
    img = np.empty((ngroups, H, W), np.float32)
    lib().orc_render(len(group), np.ascontiguousarray(pos, np.float64),
                     np.ascontiguousarray(group, np.uint8), int(ngroups),
                     np.ascontiguousarray(rot, np.float64), np.ascontiguousarray(center, np.float64),
                     float(scale), int(W), int(H), img)
    return img


def tonemap(img, glow, colors, inv_ref, K, exposure):
    """cpp/nbody.cpp:474-503 restated: (H, W, 3) uint8."""
    ng, H, W = img.shape
    out = np.empty((H, W, 3), np.uint8)
    lib().orc_tonemap(W, H, ng, np.ascontiguousarray(img, np.float32),
                      np.ascontiguousarray(glow, np.float32), glow.shape[2], glow.shape[1],
                      np.ascontiguousarray(colors, np.float32).ravel(), float(inv_ref), float(K),
                      float(exposure), out)
    return out


def ref_render(pos, group, ngroups, rot, center, scale, W, H):
    """The reference nbody_render (built from source), or None if absent."""
    R = ref_lib()
    if R is None:
        return None
    img = np.empty((ngroups, H, W), np.float32)
    R.nbody_render(len(group), np.ascontiguousarray(pos, np.float64),
                   np.ascontiguousarray(group, np.uint8), int(ngroups),
                   np.ascontiguousarray(rot, np.float64), np.ascontiguousarray(center, np.float64),
                   float(scale), int(W), int(H), img)
    return img


def ref_tonemap(img, glow, colors, inv_ref, K, exposure):
    """The reference nbody_tonemap (built from source), or None if absent."""
    R = ref_lib()
    if R is None:
        return None
    ng, H, W = img.shape
    out = np.empty((H, W, 3), np.uint8)
    R.nbody_tonemap(W, H, ng, np.ascontiguousarray(img, np.float32),
                    np.ascontiguousarray(glow, np.float32), glow.shape[2], glow.shape[1],
                    np.ascontiguousarray(colors, np.float32).ravel(), float(inv_ref), float(K),
                    float(exposure), out)
    return out
