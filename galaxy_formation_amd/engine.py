"""Typed interface to the gravity / rendering engine, with the reference's
API (galaxy_formation/engine.py:39-72): Engine.gravity, gravity_direct,
project and tonemap take and return the same numpy arrays.  The work runs in
libgf.so on the GPU; nothing here computes physics on the host.
"""

import numpy as np

from . import _lib

PRECISIONS = {"fp32": 0, "fp64": 1, "fp32q": 2}


class Engine:
    """Thin, typed interface to libgf.so (engine.py:39-72 of the reference).

    precision: "fp32" (FP32 pair arithmetic on group-relative coordinates,
    compensated accumulation; the default), "fp64", or "fp32q" (as fp32 with
    the accepted nodes' quadrupoles).  Tree, centres of mass and opening tests
    are FP64 in all three.  quad_tolerance (fp32q): a node whose quadrupole
    term is estimated below this share of its monopole is evaluated as a
    monopole (0 = every node with its quadrupole).  quad_relative_tolerance
    (fp32q, 0 = off): where the acc array passed to gravity holds the previous
    step's accelerations, a node whose quadrupole term is bounded below this
    fraction of the group's smallest |acc| is evaluated as a monopole."""

    def __init__(self, precision="fp32", quad_tolerance=None, quad_relative_tolerance=None):
        self.lib = _lib.load()
        self.precision = precision
        if quad_tolerance is not None:
            self.quad_tolerance = quad_tolerance
        if quad_relative_tolerance is not None:
            self.quad_relative_tolerance = quad_relative_tolerance

    @property
    def quad_relative_tolerance(self):
        return float(self.lib.gf_get_quad_relative_tolerance())

    @quad_relative_tolerance.setter
    def quad_relative_tolerance(self, tol):
        _lib.check(self.lib.gf_set_quad_relative_tolerance(float(tol)))

    @property
    def quad_tolerance(self):
        return float(self.lib.gf_get_quad_tolerance())

    @quad_tolerance.setter
    def quad_tolerance(self, tol):
        _lib.check(self.lib.gf_set_quad_tolerance(float(tol)))

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, p):
        if p not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self._precision = p

    def gravity(self, pos, mass, acc, theta, eps, active=None, rebuild=True):
        """Barnes-Hut accelerations (G = 1) written into acc for the active
        particles.  rebuild=False refits the previous tree instead of building
        a new one (engine.py:45-50).  The reference checks dtype and layout
        only; the shapes are checked here too, because the library reads and
        writes 3 n doubles of pos and acc and n of mass and active."""
        n = len(mass)
        if np.shape(pos) != (n, 3) or np.shape(acc) != (n, 3):
            raise ValueError(f"pos and acc must have shape ({n}, 3), got {np.shape(pos)} "
                             f"and {np.shape(acc)}")
        if active is not None and np.size(active) != n:
            raise ValueError(f"active must hold {n} entries, got {np.size(active)}")
        act = np.ones(n, np.uint8) if active is None else np.ascontiguousarray(active, dtype=np.uint8)
        _lib.check(self.lib.gf_set_precision(PRECISIONS[self._precision]))
        _lib.check(self.lib.nbody_accel_active(n, pos, mass, acc, theta, eps, _lib.ptr(act),
                                               int(rebuild)))

    def gravity_direct(self, pos, mass, eps):
        """Exact O(N^2) accelerations (engine.py:52-56)."""
        if np.shape(pos) != (len(mass), 3):
            raise ValueError(f"pos must have shape ({len(mass)}, 3), got {np.shape(pos)}")
        acc = np.zeros_like(pos)
        _lib.check(self.lib.nbody_accel_direct(len(mass), pos, mass, acc, eps))
        return acc

    def gravity_direct_subset(self, pos, mass, eps, targets):
        """Exact FP64 accelerations of the particles `targets` only."""
        tg = np.ascontiguousarray(targets, np.int64)
        acc = np.zeros((len(tg), 3))
        _lib.check(self.lib.gf_accel_direct_subset(len(mass), pos, mass, len(tg), _lib.ptr(tg),
                                                   acc, eps))
        return acc

    def last_interactions(self):
        """Interactions of the last gravity call (sum of list lengths over the
        evaluated particles)."""
        return int(self.lib.gf_last_interactions())

    def project(self, pos, group, n_groups, rot, scale, size):
        """Bilinear deposit of the particles, seen through the rotation rot
        (rows: screen x, screen y, depth) at `scale` pixels per unit, into
        one size x size float32 plane per colour group; the screen is
        centred on the origin (engine.py:58-63)."""
        img = np.empty((n_groups, size, size), np.float32)
        _lib.check(self.lib.nbody_render(len(group), np.ascontiguousarray(pos), group, n_groups,
                                         np.ascontiguousarray(rot), np.zeros(3), scale, size, size,
                                         img))
        return img

    def tonemap(self, img, glow, colors, inv_ref, K, exposure):
        """The (H, W, 3) uint8 frame: each group's deposit plane on a log
        scale plus its glow plane (any coarser gh x gw grid, sampled
        bilinearly), weighted by the group's RGB colour and saturated by
        1 - exp(-exposure x) (engine.py:65-72)."""
        ng, H, W = img.shape
        out = np.empty((H, W, 3), np.uint8)
        _lib.check(self.lib.nbody_tonemap(W, H, ng, img, np.ascontiguousarray(glow), glow.shape[2],
                                          glow.shape[1],
                                          np.ascontiguousarray(colors, dtype=np.float32).ravel(),
                                          inv_ref, K, exposure, out))
        return out

    # -- parity helpers ----------------------------------------------------

    def tree_info(self):
        import ctypes
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.gf_tree_info(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def tree_export(self):
        """The tree of the last gravity call, DFS order (cpp/nbody.cpp layout)."""
        nn, ng, _ = self.tree_info()
        out = dict(start=np.empty(nn, np.int32), count=np.empty(nn, np.int32),
                   next=np.empty(nn, np.int32), leaf=np.empty(nn, np.int32),
                   depth=np.empty(nn, np.int32), com4=np.empty((nn, 4)), open_r=np.empty(nn),
                   box4=np.empty((nn, 4)), groups=np.empty(ng, np.int32))
        _lib.check(self.lib.gf_tree_export(nn, *[_lib.ptr(out[k]) for k in
                                                 ("start", "count", "next", "leaf", "depth",
                                                  "com4", "open_r", "box4")], None,
                                           _lib.ptr(out["groups"])))
        n = int(out["count"][0]) if nn else 0
        idx = np.empty(n, np.int32)
        _lib.check(self.lib.gf_tree_export(nn, None, None, None, None, None, None, None, None,
                                           _lib.ptr(idx), None))
        out["idx"] = idx
        return out
