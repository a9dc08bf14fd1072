"""Star formation and supernova feedback with the reference module's
interface (galaxy_formation/star_formation.py): form_stars(sim, dt),
supernova_feedback(sim, new), T_SF_MAX, E_SN_PER_MSUN; plus knn_query, the
k-nearest query of points against a point set that the feedback needs
(the cKDTree query of star_formation.py:59).

Star formation (star_formation.py:35-47): gas with rho > rho_sf and
T < 3e4 K becomes a star particle during dt with probability
p = 1 - exp(-sf_eff dt / t_ff), t_ff = sqrt(3 pi / (32 G rho)); the draws are
the simulation's own rng.uniform(size=m), taken in the same order.

Feedback (star_formation.py:50-65): each new star heats its fb_ngb nearest
gas particles with fb_eff 1e49 erg per Msun formed, shared equally, and
switches their cooling off for fb_myr Myr.

The candidate selection, the neighbour query and the heating run in
libgf.so (gf_sf_candidates_dev, gf_sf_select_dev, gf_knn_query_dev,
gf_feedback_heat_dev).  Neighbours are exact and ascending in (distance,
index); cKDTree leaves the order of equal distances open.
"""

import numpy as np

from . import _lib
from .cooling import _dev, _torch
from .ics import GAS, STAR
from .units import MSUN_G, MYR

T_SF_MAX = 3e4          # star_formation.py:31
E_SN_PER_MSUN = 1e49    # star_formation.py:32


def knn_query(src, q, k):
    """(dist, nb): the k nearest rows of src (n, 3) to each point of q (m, 3),
    nb (m, k) int64 ascending in (distance, index), as
    cKDTree(src).query(q, k) returns them for k >= 1."""
    torch = _torch()
    src = np.ascontiguousarray(src, np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 3)
    ds, dq = _dev(torch, src), _dev(torch, q)
    nb = torch.empty(len(q) * k, dtype=torch.int64, device="cuda")
    dist = torch.empty(len(q) * k, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().gf_knn_query_dev(len(src), ds.data_ptr(), len(q), dq.data_ptr(), int(k),
                                            nb.data_ptr(), dist.data_ptr(), None))
    return dist.cpu().numpy().reshape(len(q), k), nb.cpu().numpy().reshape(len(q), k)


def form_stars(sim, dt):
    """Turns dense, cold gas into stars; returns the global indices of the
    new stars (star_formation.py:35-47).  Uses sim.gas_idx, rho_gas, u,
    u_per_K, rho_sf, cfg.sf_eff, rng, t; sets sim.kind and sim.birth."""
    g = sim.gas_idx
    if len(g) == 0:
        return g[:0]
    torch = _torch()
    L = _lib.load()
    dg, du, drho = _dev(torch, g, np.int64), _dev(torch, sim.u), _dev(torch, sim.rho_gas)
    rows = torch.empty(len(g), dtype=torch.int64, device="cuda")
    p = torch.empty(len(g), dtype=torch.float64, device="cuda")
    m = _lib.check(L.gf_sf_candidates_dev(len(g), dg.data_ptr(), du.numel(), du.data_ptr(),
                                          drho.data_ptr(), float(sim.u_per_K), float(sim.rho_sf),
                                          float(sim.cfg.sf_eff), float(dt), rows.data_ptr(),
                                          p.data_ptr(), None))
    if m == 0:
        return g[:0]
    draws = _dev(torch, sim.rng.uniform(size=m))
    picked = torch.empty(m, dtype=torch.int64, device="cuda")
    c = _lib.check(L.gf_sf_select_dev(m, rows.data_ptr(), p.data_ptr(), draws.data_ptr(),
                                      picked.data_ptr(), None))
    new = g[picked[:c].cpu().numpy()]
    sim.kind[new] = STAR
    sim.birth[new] = sim.t
    return new


def supernova_feedback(sim, new):
    """Heats the gas neighbours of the new star particles and blocks their
    cooling (star_formation.py:50-65).  Uses sim.cfg (fb_eff, fb_ngb, fb_myr),
    kind, pos, mass, u, no_cool_until, t; updates u and no_cool_until."""
    cfg = sim.cfg
    if cfg.fb_eff <= 0 or new.size == 0:
        return
    g = np.flatnonzero(sim.kind == GAS)
    if g.size < 10:
        return
    k = min(cfg.fb_ngb, g.size)
    _, nb = knn_query(sim.pos[g], sim.pos[new], k)
    targets = g[nb.reshape(len(new), k)]
    e_per_g = cfg.fb_eff * E_SN_PER_MSUN / MSUN_G
    until = sim.t + cfg.fb_myr / MYR
    torch = _torch()
    dt_, dm_new = _dev(torch, targets, np.int64), _dev(torch, sim.mass[new])
    dmass, du, dncu = _dev(torch, sim.mass), _dev(torch, sim.u), _dev(torch, sim.no_cool_until)
    _lib.check(_lib.load().gf_feedback_heat_dev(len(new), k, dt_.data_ptr(), dm_new.data_ptr(),
                                                float(e_per_g), float(until), len(sim.mass),
                                                dmass.data_ptr(), du.data_ptr(), dncu.data_ptr(),
                                                None))
    sim.u[:] = du.cpu().numpy()
    sim.no_cool_until[:] = dncu.cpu().numpy()
