"""FP64 numpy oracle of the step loop around the force path.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the device-resident step
loop (galaxy_formation_amd/stepper.py); the product never imports it.

Forces restates the caller of the force path,
GalaxyFormation.accelerations / _hydro (galaxy_formation/simulation.py:61-95):
gravity of the active particles (inactive rows keep their last value, as
nbody_accel_active leaves them, cpp/nbody.cpp:419-421), then, when there are
more than 40 gas particles (simulation.py:80), neighbours, h = max(d_k / 2,
eps / 2) (:84) and sph_forces (:88-90), the sum acc = acc_grav + acc_sph
(:68), dt_ind = min(0.3 sqrt(eps / max(|a|, 1e-30)), dt_courant) (:69-70)
and dt_grav = 0.3 sqrt(eps / max(max |a|, 1e-30)) (:71).  Gravity here is
the FP64 direct sum (gf_oracle.c orc_accel_direct) and the SPH sums the
C restatement of sph.py.

BlockLeapfrog restates the block-timestep kick-drift-kick leapfrog of the
device stepper: the long step / sub-step pattern of simulation.py:114-149
(slow particles kicked by the long step, fast ones sub-stepped, everybody
drifting each sub-step) widened to a power-of-two ladder of `rungs` levels.
It is pinned by mathematics, not by the device code: a particle alone on
its rung under a linear force follows the closed form of the leapfrog map
(tests/test_step_oracle.py).
"""

import numpy as np

from . import oracle as O

GAMMA, ALPHA, BETA = 5.0 / 3.0, 1.0, 2.0


class Forces:
    """simulation.py:61-95 over FP64 direct-sum gravity and the SPH oracle,
    for particles with the gas first (rows 0 .. ng-1)."""

    def __init__(self, n, ng, eps, k):
        self.n, self.ng, self.eps, self.k = n, ng, eps, k
        self.acc_grav = np.zeros((n, 3))
        self.acc_sph = np.zeros((n, 3))
        self.dt_courant = np.full(n, np.inf)
        self.dudt = np.zeros(ng)
        self.h = np.zeros(ng)
        self.rho = np.zeros(ng)

    def __call__(self, pos, vel, mass, u, active=None):
        """Returns (acc, dudt of the gas, dt_ind, dt_grav)."""
        if active is None:
            self.acc_grav[:] = O.direct(pos, mass, self.eps)
        else:
            rows = np.flatnonzero(active)
            if rows.size:
                self.acc_grav[rows] = O.direct(pos, mass, self.eps, targets=rows)
        self._hydro(pos, vel, mass, u)
        acc = self.acc_grav + self.acc_sph
        a_mag = np.sqrt(np.einsum("ij,ij->i", acc, acc))
        dt_ind = np.minimum(0.3 * np.sqrt(self.eps / np.maximum(a_mag, 1e-30)), self.dt_courant)
        dt_grav = 0.3 * np.sqrt(self.eps / max(a_mag.max(), 1e-30))
        return acc, self.dudt, dt_ind, dt_grav

    def _hydro(self, pos, vel, mass, u):
        ng = self.ng
        self.acc_sph[:] = 0.0
        self.dt_courant[:] = np.inf
        self.dudt = np.zeros(ng)
        if ng <= 40:
            return
        p = np.ascontiguousarray(pos[:ng])
        d, nb = O.knn(p, self.k)
        self.h = O.smoothing_length(d[:, -1], self.eps)
        rho, a, dudt, dt_i, _ = O.sph_forces(p, np.ascontiguousarray(vel[:ng]),
                                             np.ascontiguousarray(mass[:ng]), nb, self.h,
                                             np.ascontiguousarray(u[:ng]), GAMMA, ALPHA, BETA)
        self.acc_sph[:ng] = a
        self.dt_courant[:ng] = dt_i
        self.dudt, self.rho = dudt, rho


class BlockLeapfrog:
    """Power-of-two block time steps.  `forces(pos, vel, mass, u, active)`
    returns (acc, dudt of the gas rows, dt_ind, ...); active None = all.

    The smallest step dt_min is the `quantile` quantile of dt_ind (linear
    interpolation), chosen again when every particle is at a boundary; rung
    r steps by dt_min 2^(rungs-1-r), the largest such step not above dt_ind
    (rungs-1 also takes those below dt_min).  A sub-step drifts every
    particle by dt_min; particles at a boundary of their step are active:
    their forces are evaluated, they take the closing half kick of their
    old step, may change rung (to a longer step only where that step has a
    boundary now), and take the opening half kick of their new step.  Gas
    energies are kicked with du/dt alongside velocities and floored at
    u_floor."""

    def __init__(self, pos, vel, mass, u, ng, forces, rungs, u_floor, quantile=0.01):
        self.pos, self.vel = pos.astype(np.float64).copy(), vel.astype(np.float64).copy()
        self.mass, self.u = mass.astype(np.float64).copy(), u.astype(np.float64).copy()
        self.ng, self.forces, self.R = ng, forces, rungs
        self.u_floor, self.q = u_floor, quantile
        self.n = len(self.mass)

    def _rung_of(self, dt):
        r = np.ceil(np.log2(np.maximum(self.dt_max / dt, 1.0)))
        return np.clip(r, 0, self.R - 1).astype(np.int64)

    def _ladder(self, dt_ind):
        self.dt_min = float(np.quantile(dt_ind, self.q))
        self.dt_max = self.dt_min * 2.0 ** (self.R - 1)
        self.rung = self._rung_of(dt_ind)

    def _kick(self, sel, dt_r):
        h = np.where(sel, 0.5 * dt_r, 0.0)
        self.vel += h[:, None] * self.acc
        g = self.ng
        if g:
            ug = self.u[:g]
            self.u[:g] = np.where(sel[:g], np.maximum(ug + h[:g] * self.dudt, self.u_floor), ug)

    def _eval(self, active):
        out = self.forces(self.pos, self.vel, self.mass, self.u, active)
        self.acc, self.dudt, self.dt_ind = out[0], out[1], out[2]

    def start(self):
        self.k = 0
        self._eval(None)
        self._ladder(self.dt_ind)
        self._kick(np.ones(self.n, bool), self.dt_max / 2.0 ** self.rung)

    def substep(self):
        R = self.R
        self.k += 1
        self.pos += self.dt_min * self.vel
        period = 2 ** (R - 1 - self.rung)
        active = (self.k % period) == 0
        self._eval(active)
        self._kick(active, self.dt_max / 2.0 ** self.rung)
        if self.k % 2 ** (R - 1) == 0:
            self._ladder(self.dt_ind)
        else:
            new = self._rung_of(self.dt_ind)
            ok = (self.k % (2 ** (R - 1 - new))) == 0
            new = np.where((new < self.rung) & ~ok, self.rung, new)
            self.rung = np.where(active, new, self.rung)
        self._kick(active, self.dt_max / 2.0 ** self.rung)
        return active
