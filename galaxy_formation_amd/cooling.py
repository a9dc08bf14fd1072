"""Radiative cooling of a primordial (H + He, no metals) gas with the
reference module's interface (galaxy_formation/cooling.py):
cooling_function(T), cool(u, T, rho, dt, blocked, u_floor), T_FLOOR; and
thermodynamics(sim, dt), the reference's Simulation._thermodynamics
(simulation.py:97-104).

    du/dt = -Lambda(T) n_H^2 / rho,   n_H = X_H rho / m_H,

with Lambda the fits of Cen (1992) for collisional excitation of H and He+
plus bremsstrahlung, zero at or below 1e4 K.  The arithmetic runs in
libgf.so (gf_cooling_function_dev, gf_cool_dev, gf_thermo_dev) in FP64 with
the reference's operation order; numpy arrays go to the GPU and back.
"""

import numpy as np

from . import _lib

T_FLOOR = 1e4            # cooling.py:21


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.GFError("the gas physics runs on the GPU; no HIP device is visible")
    return torch


def _dev(torch, a, dtype=np.float64):
    return torch.from_numpy(np.ascontiguousarray(a, dtype).ravel()).cuda()


def cooling_function(T):
    """Lambda(T) in erg cm^3 / s, to be multiplied by n_H^2 (cooling.py:24-32)."""
    torch = _torch()
    T = np.asarray(T, np.float64)
    d = _dev(torch, T)
    out = torch.empty_like(d)
    _lib.check(_lib.load().gf_cooling_function_dev(d.numel(), d.data_ptr(), out.data_ptr(), None))
    return out.cpu().numpy().reshape(T.shape)


def cool(u, T, rho, dt, blocked, u_floor):
    """New internal energy after a step dt (code units), at least u_floor;
    particles with blocked=True do not cool (cooling.py:35-43)."""
    torch = _torch()
    u = np.asarray(u, np.float64)
    du_, dT, drho = _dev(torch, u), _dev(torch, T), _dev(torch, rho)
    blk = _dev(torch, np.broadcast_to(np.asarray(blocked, bool), u.shape), np.uint8)
    out = torch.empty_like(du_)
    _lib.check(_lib.load().gf_cool_dev(u.size, du_.data_ptr(), dT.data_ptr(), drho.data_ptr(),
                                       float(dt), blk.data_ptr(), float(u_floor), out.data_ptr(),
                                       None))
    return out.cpu().numpy().reshape(u.shape)


def thermodynamics(sim, dt):
    """Compression / shock heating (explicit), then radiative cooling, of the
    gas rows of `sim` (simulation.py:97-104): sim.u[g] = max(sim.u[g] +
    dt sim.dudt, u_floor), then cooled with T = u / u_per_K, except where
    sim.no_cool_until[g] > sim.t.  Uses sim.gas_idx, rho_gas, dudt, u,
    no_cool_until, t, u_per_K, u_floor; updates sim.u in place."""
    if sim.rho_gas.size == 0:
        return
    torch = _torch()
    L = _lib.load()
    g = _dev(torch, sim.gas_idx, np.int64)
    u = _dev(torch, sim.u)
    dudt, rho = _dev(torch, sim.dudt), _dev(torch, sim.rho_gas)
    ncu = _dev(torch, sim.no_cool_until)
    _lib.check(L.gf_thermo_dev(g.numel(), g.data_ptr(), u.numel(), u.data_ptr(), dudt.data_ptr(),
                               rho.data_ptr(), ncu.data_ptr(), float(sim.t), float(dt),
                               float(sim.u_per_K), float(sim.u_floor), None))
    bad = _lib.check(L.gf_gas_errors_dev(None))
    if bad:
        raise _lib.GFError(f"thermodynamics: {bad} gas indices outside the particle arrays")
    sim.u[:] = u.cpu().numpy()
