"""CPU restatement (numpy, FP64) of the gas physics around the force path:
radiative cooling (galaxy_formation/cooling.py:24-43), the sub-step
thermodynamics (simulation.py:97-104), star formation
(star_formation.py:35-47) and supernova feedback (star_formation.py:50-65),
with a brute-force k-nearest query in place of scipy's cKDTree.

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module; the product
never does.

Pins (tests/test_gas.py): cooling-function values evaluated by hand at four
temperatures, its floor and high-temperature asymptote, du linear in dt and
in rho, the unit constants against the values the reference's units.py
docstring quotes; the k-nearest query against scipy's cKDTree (an
independent implementation); the np.add.at heating against a hand-summed
case with a particle heated by two stars.
"""

import numpy as np

# galaxy_formation/units.py:17-35, evaluated as that file evaluates them
G_KPC = 4.30091e-6
KPC = 3.5
MSUN = 4e10
KMS = np.sqrt(G_KPC * MSUN / KPC)
MYR = KPC / KMS * 977.79
M_H = 1.6726e-24
K_B = 1.381e-16
MSUN_G = 1.989e33
KPC_CM = 3.086e21
YR_S = 3.156e7
RHO_CGS = MSUN * MSUN_G / (KPC * KPC_CM) ** 3
U_CGS = (KMS * 1e5) ** 2
SEC_PER_MYR = 1e6 * YR_S
X_H = 0.76
MU_IONISED = 0.6

T_FLOOR = 1e4            # cooling.py:21
T_SF_MAX = 3e4           # star_formation.py:31
E_SN_PER_MSUN = 1e49     # star_formation.py:32
GAS, STAR = 0, 1         # initial_conditions.py:22


def cooling_function(T):
    """cooling.py:24-32."""
    T = np.maximum(T, 1.0)
    f = 1 / (1 + np.sqrt(T / 1e5))
    lam = (7.5e-19 * np.exp(-118348 / T) * f
           + 5.54e-17 * T ** -0.397 * np.exp(-473638 / T) * f
           + 1.42e-27 * 1.1 * np.sqrt(T))
    return np.where(T > T_FLOOR, lam, 0.0)


def cool(u, T, rho, dt, blocked, u_floor):
    """cooling.py:35-43."""
    rho_cgs = rho * RHO_CGS
    nH = X_H * rho_cgs / M_H
    dudt_cgs = cooling_function(T) * nH ** 2 / rho_cgs
    du = dudt_cgs * dt * MYR * SEC_PER_MYR / U_CGS
    du[blocked] = 0.0
    return np.maximum(u - du, u_floor)


def thermodynamics(u, g, dudt, rho, no_cool_until, t, dt, u_per_K, u_floor):
    """simulation.py:97-104 on a copy of u (all particles); returns it."""
    u = u.copy()
    u[g] = np.maximum(u[g] + dt * dudt, u_floor)
    u[g] = cool(u[g], u[g] / u_per_K, rho, dt, no_cool_until[g] > t, u_floor)
    return u


def sf_candidates(g, u, rho, u_per_K, rho_sf, sf_eff, dt):
    """star_formation.py:37-43: gas rows of dense, cold gas and p."""
    dense = (rho > rho_sf) & (u[g] / u_per_K < T_SF_MAX)
    rows = np.flatnonzero(dense)
    t_ff = np.sqrt(3 * np.pi / (32 * rho[dense]))
    p = 1 - np.exp(-sf_eff * dt / t_ff)
    return rows, p


def knn_query(src, q, k):
    """Brute force: for each query point the k nearest sources, ascending in
    (distance, index); (nb, dist)."""
    nb = np.empty((len(q), k), np.int64)
    dist = np.empty((len(q), k))
    for i in range(len(q)):
        d2 = ((src - q[i]) ** 2).sum(1)
        order = np.lexsort((np.arange(len(src)), d2))[:k]
        nb[i] = order
        dist[i] = np.sqrt(d2[order])
    return nb, dist


def feedback_heat(u, no_cool_until, mass, targets, new_mass, fb_eff, k, until):
    """star_formation.py:60-65 on copies; returns (u, no_cool_until)."""
    u, ncu = u.copy(), no_cool_until.copy()
    e_per_g = fb_eff * E_SN_PER_MSUN / MSUN_G
    E = e_per_g * new_mass * MSUN * MSUN_G
    du = E[:, None] / k / (mass[targets] * MSUN * MSUN_G) / U_CGS
    np.add.at(u, targets.ravel(), du.ravel())
    ncu[targets.ravel()] = until
    return u, ncu
