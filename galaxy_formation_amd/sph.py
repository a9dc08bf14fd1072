"""Smoothed Particle Hydrodynamics of the gas, with the reference's module
interface (galaxy_formation/sph.py): kernel, kernel_grad, sph_forces,
knn_from_candidates, NeighbourFinder and the constants N_NEIGHBOURS,
N_CANDIDATES, ALPHA, BETA.  The arithmetic runs in libgf.so on the GPU,
NeighbourFinder's candidate cache included; the host code here only moves
arrays.

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


class VerletCache:
    """Sizes of the candidate cache NeighbourFinder holds on the GPU (the
    lists stay in device memory between calls; `export` copies them out)."""

    def __init__(self, n_total, rows, candidates):
        self.n_total, self.rows, self.candidates = n_total, rows, candidates

    def export(self):
        """(owner, cand): owner (n_total,) the cache row of each particle, -1
        where it was not gas at the search; cand (rows, candidates) global
        particle indices, nearest first."""
        owner = np.empty(self.n_total, np.int64)
        cand = np.empty((self.rows, self.candidates), np.int64)
        _lib.check(_lib.load().gf_nbcache_export(_lib.ptr(owner), _lib.ptr(cand)))
        return owner, cand


class NeighbourFinder:
    """k nearest gas neighbours of every gas particle, the particle itself
    first (the contract of sph.py:138-170).

    find(g, p, fresh) takes the global indices g of the current gas
    particles and their positions p, and returns (dist, nb), nb in local
    indices 0 .. len(g)-1.  With fresh=True, or when the cache cannot serve
    the call, it searches all gas for each particle's n_candidates nearest
    and keeps them on the GPU as global particle indices
    (gf_nbcache_search).  Otherwise the k nearest are re-selected on the GPU
    among those candidates that are still gas (gf_nbcache_reselect); a
    particle without a cached row, or left with fewer than k candidates,
    sends the call back to a full search.  The cache is the library's (one
    per process, as the library's other state); `cache` describes it."""

    def __init__(self, n_total, k=None, n_candidates=None):
        self.n_total = int(n_total)
        self.k = N_NEIGHBOURS if k is None else int(k)
        self.n_candidates = N_CANDIDATES if n_candidates is None else int(n_candidates)
        if not 1 <= self.k <= self.n_candidates <= 64:
            raise ValueError("need 1 <= k <= n_candidates <= 64")
        self.cache = None
        self.full_searches = 0

    def find(self, g, p, fresh):
        g = np.ascontiguousarray(g, np.int64)
        p = np.ascontiguousarray(p, np.float64)
        n, k = g.size, self.k
        nb = np.empty((n, k), np.int64)
        d = np.empty((n, k))
        L = _lib.load()
        if not fresh and self.cache is not None and n:
            if _lib.check(L.gf_nbcache_reselect(self.n_total, n, g, p, k, nb, d)) == 1:
                return d, nb
        if n == 0:
            return d, nb
        m = min(self.n_candidates, n)
        if m < k:                 # fewer gas particles than k: every one is a neighbour
            nb, d = nb[:, :m].copy(), d[:, :m].copy()
        _lib.check(L.gf_nbcache_search(self.n_total, n, g, p, m, min(k, m), nb, d))
        self.cache = VerletCache(self.n_total, n, m)
        self.full_searches += 1
        return d, nb
