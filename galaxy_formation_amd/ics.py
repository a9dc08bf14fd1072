"""Synthetic particle sets for the force-evaluation workloads (code units:
length 3.5 kpc, mass 4e10 Msun, G = 1, as galaxy_formation/units.py:17-35).

Particles are ordered gas first, then dark matter, as the reference's
make_cloud does (galaxy_formation/initial_conditions.py:41-45).  Every
generator draws from numpy's PCG64 with a fixed seed, so the same call gives
the same bytes on every machine with this numpy.
"""

from dataclasses import dataclass

import numpy as np

from .units import G_KPC, K_B, KMS, KPC, M_H, MSUN, MU_IONISED, U_CGS  # noqa: F401

GAMMA = 5.0 / 3.0
U_PER_K = K_B / ((GAMMA - 1) * MU_IONISED * M_H) / U_CGS   # simulation.py:40

GAS, STAR, DM = 0, 1, 2         # initial_conditions.py:22


@dataclass
class Particles:
    pos: np.ndarray      # (n, 3) f64
    vel: np.ndarray      # (n, 3) f64
    mass: np.ndarray     # (n,) f64
    u: np.ndarray        # (n,) f64, internal energy (0 for dark matter)
    kind: np.ndarray     # (n,) int8
    eps: float           # Plummer softening, code units

    @property
    def n(self):
        return len(self.mass)

    @property
    def gas(self):
        return np.flatnonzero(self.kind == GAS)


def _isotropic(n, rng):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def make_cloud(n_gas, n_dm, seed=3, r0_kpc=100.0, mass_msun=2.8e11, f_baryon=0.16, spin=0.05,
               t_gas=1e4, eps_kpc=0.1):
    """Rotating cloud: radii r = R sqrt(u) (rho ~ 1/r), isotropic directions,
    solid-body rotation about z with spin lambda (Peebles convention,
    J |E|^(1/2) / (G M^(5/2)) with E the potential energy, as
    initial_conditions.py:50-53 uses it for its uniform cloud).  For rho ~ 1/r inside R,
    I = M R^2 / 3 and E = -2 G M^2 / (3 R), so
    Omega = 3 lambda sqrt(G M) / (sqrt(2/3) R^(3/2))."""
    rng = np.random.default_rng(seed)
    M = mass_msun / MSUN
    R = r0_kpc / KPC
    n = n_gas + n_dm
    pos = np.concatenate([_isotropic(n_gas, rng) * (R * np.sqrt(rng.uniform(0, 1, n_gas)))[:, None],
                          _isotropic(n_dm, rng) * (R * np.sqrt(rng.uniform(0, 1, n_dm)))[:, None]])
    mass = np.concatenate([np.full(n_gas, f_baryon * M / max(n_gas, 1)),
                           np.full(n_dm, (1 - f_baryon) * M / max(n_dm, 1))])
    omega = 3 * spin * np.sqrt(M) / (np.sqrt(2.0 / 3.0) * R ** 1.5)
    vel = np.cross(np.array([0.0, 0.0, omega]), pos)
    kind = np.concatenate([np.full(n_gas, GAS, np.int8), np.full(n_dm, DM, np.int8)])
    u = np.where(kind == GAS, t_gas * U_PER_K, 0.0)
    return Particles(np.ascontiguousarray(pos), np.ascontiguousarray(vel), mass, u, kind,
                     eps_kpc / KPC)


def make_disk(n_gas, n_dm, seed=3, mass_msun=2.8e11, f_baryon=0.16, r_d_kpc=3.0, z_d_kpc=0.3,
              a_kpc=20.0, r_t_kpc=100.0, t_min=1e4, t_max=1e6, eps_kpc=0.1):
    """Collapsed-disk state: exponential gas disk (scale length r_d, exponential
    vertical profile with scale height z_d) inside a Hernquist dark-matter halo
    (scale radius a) truncated at r_t by inverse-CDF sampling
    r = a sqrt(s) / (1 - sqrt(s)), s ~ U(0, M(<r_t)/M).  Gas moves on circular
    orbits v_c = sqrt(M(<r) / r) (spherically enclosed mass of all particles)
    with a 10 % Gaussian dispersion; halo particles get isotropic Gaussian
    velocities of 1D dispersion v_c / sqrt(3).  Gas temperature is
    log-uniform on [t_min, t_max]."""
    rng = np.random.default_rng(seed)
    M = mass_msun / MSUN
    r_d, z_d, a, r_t = r_d_kpc / KPC, z_d_kpc / KPC, a_kpc / KPC, r_t_kpc / KPC
    # disk: cylindrical radius from the Gamma(2) law of an exponential disk
    # (truncated at r_t), height from a two-sided exponential
    R = rng.gamma(2.0, r_d, size=n_gas)
    while np.any(R > r_t):
        bad = R > r_t
        R[bad] = rng.gamma(2.0, r_d, size=int(bad.sum()))
    phi = rng.uniform(0, 2 * np.pi, n_gas)
    z = rng.laplace(0.0, z_d, n_gas)
    disk = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    # halo
    smax = (r_t / (r_t + a)) ** 2
    sq = np.sqrt(rng.uniform(0, smax, n_dm))
    r_h = a * sq / (1 - sq)
    halo = _isotropic(n_dm, rng) * r_h[:, None]
    pos = np.ascontiguousarray(np.concatenate([disk, halo]))
    mass = np.concatenate([np.full(n_gas, f_baryon * M / max(n_gas, 1)),
                           np.full(n_dm, (1 - f_baryon) * M / max(n_dm, 1))])
    # circular speed from the spherically enclosed mass
    r = np.linalg.norm(pos, axis=1)
    order = np.argsort(r, kind="stable")
    menc = np.empty_like(r)
    menc[order] = np.cumsum(mass[order])
    vc = np.sqrt(menc / np.maximum(r, 1e-6))
    vel = np.empty_like(pos)
    ez = np.array([0.0, 0.0, 1.0])
    g = slice(0, n_gas)
    phat = np.cross(ez, disk)
    phat /= np.linalg.norm(phat, axis=1, keepdims=True) + 1e-30
    vel[g] = vc[g, None] * phat + 0.1 * vc[g, None] * rng.normal(size=(n_gas, 3))
    vel[n_gas:] = (vc[n_gas:, None] / np.sqrt(3)) * rng.normal(size=(n_dm, 3))
    kind = np.concatenate([np.full(n_gas, GAS, np.int8), np.full(n_dm, DM, np.int8)])
    T = np.exp(rng.uniform(np.log(t_min), np.log(t_max), n_gas))
    u = np.concatenate([T * U_PER_K, np.zeros(n_dm)])
    return Particles(pos, np.ascontiguousarray(vel), mass, u, kind, eps_kpc / KPC)


def gaussian_blob(n, seed=0):
    """The reference's Barnes-Hut test input (tests/test_physics.py:22-24):
    standard-normal positions from default_rng(0), equal masses 1/n."""
    rng = np.random.default_rng(seed)
    pos = np.ascontiguousarray(rng.normal(size=(n, 3)))
    return pos, np.full(n, 1.0 / n)


def config(name):
    """The workloads by name: cloud10k (5k gas + 5k DM), cloud200k
    (100k + 100k), disk1m (500k + 500k), disk4m (2M + 2M)."""
    if name == "cloud10k":
        return make_cloud(5000, 5000)
    if name == "cloud200k":
        return make_cloud(100000, 100000)
    if name == "disk1m":
        return make_disk(500000, 500000)
    if name == "disk4m":
        return make_disk(2000000, 2000000)
    raise ValueError(f"unknown workload {name!r}")


def sha256(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
