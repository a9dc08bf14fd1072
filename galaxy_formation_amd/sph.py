"""Smoothed Particle Hydrodynamics of the gas, with the reference's module
interface (galaxy_formation/sph.py): kernel, kernel_grad, sph_forces,
knn_from_candidates, NeighbourFinder and the constants N_NEIGHBOURS,
N_CANDIDATES, ALPHA, BETA.  The arithmetic runs in libgf.so on the GPU; the
host code here only moves arrays and keeps the index bookkeeping of
NeighbourFinder (sph.py:151-170).

Density, forces and time step follow sph.py:57-106:

    rho_i = sum_j m_j W(r_ij, h_i)
    P_i = (gamma - 1) rho_i u_i
    dv_i/dt = -sum_j m_j [P_i/rho_i^2 grad W(h_i) + P_j/rho_j^2 grad W(h_j) + Pi_ij grad W_avg]
    du_i/dt = sum_j m_j [P_i/rho_i^2 grad W(h_i) + Pi_ij/2 grad W_avg] . v_ij

with the cubic-spline kernel of support 2h and Monaghan-Gingold viscosity.
"""

import numpy as np

from . import _lib

N_NEIGHBOURS = 32            # sph.py:26
N_CANDIDATES = 48            # sph.py:27
ALPHA, BETA = 1.0, 2.0       # sph.py:28


def _kernel_eval(r, h, want_w, want_g):
    r_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(r, np.float64), np.broadcast(r, h).shape))
    h_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(h, np.float64), r_arr.shape))
    rf, hf = r_arr.ravel(), h_arr.ravel()
    w = np.empty_like(rf) if want_w else None
    g = np.empty_like(rf) if want_g else None
    _lib.check(_lib.load().gf_sph_kernel(rf.size, rf, hf, _lib.ptr(w), _lib.ptr(g)))
    out = w if want_w else g
    out = out.reshape(r_arr.shape)
    return float(out) if out.ndim == 0 else out


def kernel(r, h):
    """Cubic spline (M4) kernel with support 2h (sph.py:31-40).  Scalars or
    arrays (broadcast)."""
    return _kernel_eval(r, h, True, False)


def kernel_grad(r, h):
    """(1/r) dW/dr: multiply by the separation vector to get grad W
    (sph.py:43-54)."""
    return _kernel_eval(r, h, False, True)


def sph_forces(pos, vel, mass, nb, h, u, gamma, alpha, beta, rho, acc, dudt, dt_i, divv=None):
    """Density, pressure + viscosity acceleration, du/dt and Courant time step
    of every gas particle (sph.py:57-106).  nb holds the neighbour indices
    (the particle itself included).  Outputs rho, acc, dudt, dt_i (and divv,
    the velocity divergence, when given) are written in place."""
    nb = np.ascontiguousarray(nb, np.int64)
    n, k = nb.shape
    outs = {"rho": rho, "acc": acc, "dudt": dudt, "dt_i": dt_i, "divv": divv}
    tmp = {}
    for name, arr in outs.items():
        if arr is None:
            tmp[name] = None
        elif arr.flags.c_contiguous and arr.dtype == np.float64:
            tmp[name] = arr
        else:
            tmp[name] = np.empty(arr.shape, np.float64)
    _lib.check(_lib.load().gf_sph_forces(
        n, k, np.ascontiguousarray(pos, np.float64), np.ascontiguousarray(vel, np.float64),
        np.ascontiguousarray(mass, np.float64), nb, np.ascontiguousarray(h, np.float64),
        np.ascontiguousarray(u, np.float64), float(gamma), float(alpha), float(beta),
        *[_lib.ptr(tmp[k_]) for k_ in ("rho", "acc", "dudt", "dt_i", "divv")]))
    for name, arr in outs.items():
        if arr is not None and tmp[name] is not arr:
            arr[...] = tmp[name]


def knn_from_candidates(pos, cand, k, nb, dist):
    """k nearest neighbours of each particle among its candidates (local
    indices, -1 = none), written to nb and dist (sph.py:109-135).  Returns
    False if some particle does not have enough candidates."""
    cand = np.ascontiguousarray(cand, np.int64)
    n, m = cand.shape
    nb_c = nb if (nb.flags.c_contiguous and nb.dtype == np.int64) else np.ascontiguousarray(nb, np.int64)
    d_c = dist if (dist.flags.c_contiguous and dist.dtype == np.float64) else np.ascontiguousarray(dist, np.float64)
    ok = _lib.check(_lib.load().gf_knn_from_candidates_host(
        n, m, np.ascontiguousarray(pos, np.float64), cand, int(k), nb_c, d_c))
    if nb_c is not nb:
        nb[...] = nb_c
    if d_c is not dist:
        dist[...] = d_c
    return bool(ok)


def knn(p, k, r_init=None):
    """Exact k nearest neighbours of every point of p, itself first:
    (dist (n, k), nb (n, k) int64, passes), rows ascending in distance (ties
    by index) -- the cKDTree query of sph.py:165-166 on the GPU."""
    p = np.ascontiguousarray(p, np.float64)
    n = len(p)
    nb = np.empty((n, k), np.int64)
    d = np.empty((n, k))
    mean = np.zeros(1)
    r0 = None if r_init is None else np.ascontiguousarray(r_init, np.float64)
    passes = _lib.check(_lib.load().gf_knn(n, p, int(k), _lib.ptr(r0), nb, d, _lib.ptr(mean)))
    return d, nb, {"passes": int(passes), "mean_passes": float(mean[0])}


class NeighbourFinder:
    """k nearest gas neighbours of every gas particle (sph.py:138-170).

    A full search (N_CANDIDATES candidates) is done when `fresh` is True.
    Otherwise the k nearest are re-selected among the candidates remembered
    from the last full search (a Verlet list).  Indices are global particle
    indices in the cache, so the list survives gas particles turning into
    stars between two searches."""

    def __init__(self, n_total, k=None, n_candidates=None):
        self.n_total = n_total
        self.k = N_NEIGHBOURS if k is None else k
        self.n_candidates = N_CANDIDATES if n_candidates is None else n_candidates
        self.cache = None
        self.last_passes = None

    def find(self, g, p, fresh):
        """g: global indices of gas particles, p: their positions. Returns
        (dist, nb) with nb in local (0..len(g)-1) indices."""
        k = self.k
        if not fresh and self.cache is not None:
            row = self.cache["row"][g]
            if (row >= 0).all():
                loc = np.full(self.n_total, -1, np.int64)
                loc[g] = np.arange(g.size)
                cand = loc[self.cache["cand"][row]]      # candidates that are still gas
                nb = np.empty((g.size, k), np.int64)
                d = np.empty((g.size, k))
                if knn_from_candidates(p, cand, k, nb, d):
                    return d, nb
        kq = min(self.n_candidates, g.size)
        d, nb, info = knn(p, kq)
        self.last_passes = info
        row = np.full(self.n_total, -1, np.int64)
        row[g] = np.arange(g.size)
        self.cache = {"row": row, "cand": g[nb]}
        return d[:, :k], np.ascontiguousarray(nb[:, :k])
