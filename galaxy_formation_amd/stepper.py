"""Device-resident force evaluation and a leapfrog driver over libgf.so.

ForceEvaluator reproduces the call pattern of the reference's
GalaxyFormation.accelerations / _hydro (galaxy_formation/simulation.py:61-95):
Barnes-Hut gravity of every particle, then the gas: neighbours, h,
density, pressure + viscosity forces, Courant time steps, and
dt_ind = min(0.3 sqrt(eps / |a|), dt_courant) (simulation.py:68-72).  State
stays in device memory (PyTorch tensors); all physics runs in libgf.so.

The neighbour solve can start from three first radii:
  "plain"    the gas tree's cell densities (no history);
  "prev"     d_k of the previous evaluation;
  "learned"  the MLP proposal of learned_h.LearnedH.
The converged neighbours, h and densities do not depend on the choice.

Stepper integrates with a global kick-drift-kick leapfrog whose step is the
10th percentile of the particles' admissible steps (the long step of
simulation.py:117) times a factor; it records what the learned proposal
trains on.  (The reference's two-level sub-stepping, cooling and star
formation are not part of this force path.)
"""

import contextlib

import numpy as np

from . import _lib
from .ics import U_PER_K

GAMMA, ALPHA, BETA = 5.0 / 3.0, 1.0, 2.0


class ForceEvaluator:
    def __init__(self, n, n_gas, eps, theta=0.5, k=64, precision="fp32", device="cuda:0",
                 stream=None, quad_relative_tolerance=None, sph_precision=None,
                 gas_tree="gravity"):
        import torch
        self.torch = torch
        self.L = _lib.load()
        self.dev = torch.device(device)
        self.n, self.ng = n, n_gas
        self.eps, self.theta, self.k = eps, theta, k
        self.prec = {"fp32": 0, "fp64": 1, "fp32q": 2}[precision]
        self.stream = stream
        # fp32q: the relative quadrupole rule, fed by acc_grav, which keeps the
        # previous evaluation's gravity (None leaves the library's setting)
        self.quad_rel = quad_relative_tolerance
        # density and forces: 0 FP64 pair terms, 1 FP32 pair terms with FP64
        # sums (None leaves the library's setting)
        self.sph_prec = sph_precision
        # the gas tree: "gravity" takes the gravity tree's Morton order
        # restricted to the gas (the first n_gas rows) when this evaluation
        # built that tree from the same positions; "own" sorts the gas alone
        if gas_tree not in ("gravity", "own"):
            raise ValueError(f"unknown gas tree {gas_tree!r}")
        self.gas_tree = gas_tree
        self._built_pos = None      # positions of this evaluation's gravity tree build
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.dev)  # noqa: E731
        self.acc_grav = z(n, 3)
        self.acc = z(n, 3)
        g = max(n_gas, 1)
        self.h, self.rho, self.acc_sph = z(g), z(g), z(g, 3)
        self.dudt, self.dt_c, self.divv = z(g), z(g), z(g)
        self.dk, self.cell_rho, self.cell_count, self.r0 = z(g), z(g), z(g), z(g)
        self.dt_ind = z(n)
        self.dt_grav = torch.zeros((), dtype=torch.float64, device=self.dev)
        self.passes = 0

    @property
    def sp(self):
        if self.stream is not None:
            return self.stream.cuda_stream
        return self.torch.cuda.current_stream().cuda_stream or None

    def gravity(self, pos, mass, active=None, rebuild=True, ev=None):
        chk, L = _lib.check, self.L
        if self.prec == 2 and self.quad_rel is not None:
            chk(L.gf_set_quad_relative_tolerance(float(self.quad_rel)))
        if ev is not None:
            ev[0].record()
        self._built_pos = pos.data_ptr() if rebuild else None
        chk(L.gf_tree_build_dev(self.n, pos.data_ptr(), mass.data_ptr(), self.theta, self.sp)
            if rebuild else
            L.gf_gravity_dev(self.n, pos.data_ptr(), mass.data_ptr(), self.acc_grav.data_ptr(),
                             self.theta, self.eps, None if active is None else active.data_ptr(),
                             0, self.prec, self.sp))
        if ev is not None:
            ev[1].record()
        if rebuild:
            chk(L.gf_walk_dev(self.n, None if active is None else active.data_ptr(),
                              self.acc_grav.data_ptr(), self.eps, self.prec, self.sp))
        if ev is not None:
            ev[2].record()

    def hydro(self, pos, vel, mass, u, first="plain", dk_prev=None, divv_prev=None,
              dtc_prev=None, dt=None, model=None, ev=None, margin=0.0):
        """Gas of the first n_gas particles.  Returns the number of passes.
        first="prev" starts from d_k(prev) exp(margin)."""
        chk, L, ng = _lib.check, self.L, self.ng
        if ng <= 40:
            # simulation.py:78-80: no hydro for 40 gas particles or fewer
            for t in (self.h, self.rho, self.acc_sph, self.dudt, self.divv, self.dk):
                t.zero_()
            self.dt_c.fill_(float("inf"))
            self.passes = 0
            return 0
        if self.gas_tree == "gravity" and self._built_pos == pos.data_ptr():
            # the gas tree from the gravity tree built from these positions
            # (the stream waits for that build's sort)
            chk(L.gf_hydro_setup_from_tree_dev(ng, pos.data_ptr(), vel.data_ptr(),
                                               mass.data_ptr(), u.data_ptr(), self.sp))
        else:
            chk(L.gf_hydro_setup_dev(ng, pos.data_ptr(), vel.data_ptr(), mass.data_ptr(),
                                     u.data_ptr(), self.sp))
        if ev is not None:
            ev[3].record()
        r0 = None
        if first == "prev":
            r0 = dk_prev if margin == 0.0 else self.torch.mul(dk_prev, float(np.exp(margin)),
                                                              out=self.r0)
        elif first == "learned":
            self.cell_density()
            x = model.features(self.k, dk_prev, self.cell_rho, self.cell_count, divv_prev,
                               dtc_prev, dt, stream=self.stream)
            r0 = model.propose(x, dk_prev, out=self.r0, stream=self.stream)
        elif first != "plain":
            raise ValueError(f"unknown first-radius policy {first!r}")
        self.passes = chk(L.gf_hydro_knn_dev(self.k, None if r0 is None else r0.data_ptr(),
                                             self.sp))
        if ev is not None:
            ev[4].record()
        if self.sph_prec is not None:
            chk(L.gf_set_sph_precision(int(self.sph_prec)))
        chk(L.gf_hydro_forces_dev(self.eps, GAMMA, ALPHA, BETA, self.h.data_ptr(),
                                  self.rho.data_ptr(), self.acc_sph.data_ptr(),
                                  self.dudt.data_ptr(), self.dt_c.data_ptr(),
                                  self.divv.data_ptr(), self.sp))
        chk(L.gf_hydro_dk_dev(self.dk.data_ptr(), self.sp))
        if ev is not None:
            ev[5].record()
        return self.passes

    def cell_density(self):
        """Density and particle count of the smallest gas-tree cell holding
        >= k particles around each gas particle (current gas tree)."""
        _lib.check(self.L.gf_hydro_cell_density_dev(self.k, self.cell_rho.data_ptr(),
                                                    self.cell_count.data_ptr(), self.sp))
        return self.cell_rho, self.cell_count

    def evaluate(self, pos, vel, mass, u, **hydro_kw):
        """acc = gravity + SPH; dt_ind (simulation.py:61-72)."""
        self.gravity(pos, mass)
        passes = self.hydro(pos[:self.ng], vel[:self.ng], mass[:self.ng], u[:self.ng], **hydro_kw)
        self.combine()
        return passes

    def combine(self):
        """The caller's sum and time steps (simulation.py:68-71): acc =
        acc_grav + acc_sph (gas rows), dt_ind = min(0.3 sqrt(eps / max(|a|,
        1e-30)), dt_courant) with dt_courant only on the gas, and dt_grav =
        0.3 sqrt(eps / max(max |a|, 1e-30)) (a device scalar)."""
        ng = self.ng
        _lib.check(self.L.gf_combine_dev(self.n, ng, self.eps, self.acc_grav.data_ptr(),
                                         self.acc_sph.data_ptr() if ng else None,
                                         self.dt_c.data_ptr() if ng else None,
                                         self.acc.data_ptr(), self.dt_ind.data_ptr(),
                                         self.dt_grav.data_ptr(), self.sp))


class Stepper:
    """Global-step KDK leapfrog over a ForceEvaluator, recording per step
    (t, d_k, div v, Courant dt, cell density, cell count) of the gas for the
    learned proposal."""

    def __init__(self, particles, theta=0.5, k=64, precision="fp32", device="cuda:0"):
        import torch
        self.torch = torch
        dev = torch.device(device)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        p = particles
        self.ng = int((p.kind == 0).sum())
        assert (np.flatnonzero(p.kind == 0) == np.arange(self.ng)).all(), "gas first"
        self.pos, self.vel, self.mass, self.u = T(p.pos), T(p.vel), T(p.mass), T(p.u)
        self.u_floor = 1e4 * U_PER_K        # the cooling floor, T = 1e4 K (cooling.py:21)
        self.F = ForceEvaluator(p.n, self.ng, p.eps, theta, k, precision, device)
        self.t = 0.0
        self.records = []

    def _record(self):
        rho, cnt = self.F.cell_density()
        self.records.append((self.t, self.F.dk.clone(), self.F.divv.clone(), self.F.dt_c.clone(),
                             rho.clone(), cnt.clone()))

    def run(self, steps, dt_factor=1.0, first="plain", model=None):
        torch, F = self.torch, self.F
        F.evaluate(self.pos, self.vel, self.mass, self.u)
        self._record()
        passes = []
        for _ in range(steps):
            dt = float(dt_factor * torch.quantile(F.dt_ind, 0.10).item())
            self.vel += 0.5 * dt * F.acc
            self.pos += dt * self.vel
            if self.ng:
                self.u[:self.ng] = torch.clamp(self.u[:self.ng] + dt * F.dudt, min=self.u_floor)
            dk_prev, divv_prev, dtc_prev = F.dk.clone(), F.divv.clone(), F.dt_c.clone()
            passes.append(F.evaluate(self.pos, self.vel, self.mass, self.u, first=first,
                                     dk_prev=dk_prev, divv_prev=divv_prev, dtc_prev=dtc_prev,
                                     dt=dt, model=model))
            self.vel += 0.5 * dt * F.acc
            self.t += dt
            self._record()
        return passes


def training_pairs(records, model, k, max_gap=4):
    """Samples from a run: for snapshot pairs (s, s + g), g <= max_gap, the
    features (d_k, div v and Courant dt at s, the cell at s + g, dt =
    t_{s+g} - t_s) and the target log d_k(s + g) - log d_k(s).  Returns the
    BF16 features normalised with the model's current statistics (n, 4), the
    raw FP32 features (n, 4) and the targets (n,), all on the device."""
    import torch
    xs, raws, targets = [], [], []
    for s in range(len(records)):
        for g in range(1, max_gap + 1):
            if s + g >= len(records):
                break
            t0, dk0, dv0, dtc0, _, _ = records[s]
            t1, dk1, _, _, rho1, cnt1 = records[s + g]
            x, raw = model.features(k, dk0, rho1, cnt1, dv0, dtc0, float(t1 - t0), raw=True)
            xs.append(x)
            raws.append(raw)
            targets.append((torch.log(dk1) - torch.log(dk0)).float())
    return torch.cat(xs), torch.cat(raws), torch.cat(targets)


class RungStepper:
    """Device-resident hierarchical (block) leapfrog with `rungs` time-step
    levels, the pattern of simulation.py:114-149 (a long step and sub-steps
    for the fast particles) widened to 2^(rungs-1) sub-steps per long step.

    The smallest step dt_min is the 1st percentile of the particles' admissible
    steps dt_ind = min(0.3 sqrt(eps/|a|), dt_Courant) (simulation.py:68-72),
    chosen again at every long-step boundary (as simulation.py:117 chooses
    its long step each cycle); rung r steps by dt_min 2^(rungs-1-r), the
    largest step not above dt_ind (rung rungs-1 also takes the particles
    below dt_min).  Every sub-step
    drifts all particles by dt_min; the particles at a boundary of their step
    are active: gravity is evaluated for them only, on a tree rebuilt every
    sub-step (refit = True refits it between long steps instead, as the
    reference does, rebuild = 0 of nbody_accel_active; a refit widens every
    opening radius by the largest drift since the build, which on the GPU
    costs more walk than a new tree costs build), they take
    the closing and opening half kicks (v, and u of the gas with du/dt), and
    may change rung (to a longer step only where it is also a boundary).  The
    gas hydro (neighbours, h, density, forces) is evaluated for all gas at
    every sub-step; only the active particles use it.  Cooling, star formation
    and feedback are not part of this force path.

    With overlap (the default), the gas pipeline runs on a second stream
    beside the gravity tree build and walk: it waits for the sub-step's drift
    and the forces are summed once both are done.

    Energy and angular momentum are measured where every particle is at a
    boundary (every 2^(rungs-1) sub-steps), on the synchronised velocities
    and energies (the opening half kicks taken back): kinetic + thermal +
    potential, the potential by the FP64 direct sum (gf_potential_dev)."""

    def __init__(self, particles, theta=0.5, k=64, precision="fp32", rungs=6, first="plain",
                 model=None, device="cuda:0", refit=False, quad_relative_tolerance=None,
                 sph_precision=None, overlap=True, gas_tree="gravity"):
        import torch
        self.torch = torch
        dev = torch.device(device)
        self.dev = dev
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        p = particles
        self.ng = int((p.kind == 0).sum())
        assert (np.flatnonzero(p.kind == 0) == np.arange(self.ng)).all(), "gas first"
        self.n = p.n
        self.eps = p.eps
        self.pos, self.vel, self.mass, self.u = T(p.pos), T(p.vel), T(p.mass), T(p.u)
        self.u_floor = 1e4 * U_PER_K
        self.R = rungs
        self.refit = refit
        self.first, self.model = first, model
        self.F = ForceEvaluator(p.n, self.ng, p.eps, theta, k, precision, device,
                                quad_relative_tolerance=quad_relative_tolerance,
                                sph_precision=sph_precision, gas_tree=gas_tree)
        self.L = self.F.L
        self.phi = torch.zeros(p.n, dtype=torch.float64, device=dev)
        self.rung = torch.zeros(p.n, dtype=torch.int64, device=dev)
        self.active = torch.zeros(p.n, dtype=torch.uint8, device=dev)   # this sub-step's
        self._counts = torch.zeros(1024, dtype=torch.int32, device=dev)  # active per sub-step
        self.active_counts = []   # device scalars (views of _counts), read at the end
        self.k = 0
        self.passes = []
        self.L_sync = []
        self.t_hydro = 0.0        # time of the last hydro evaluation
        self.prev = None          # (d_k, div v, dt_Courant) of that evaluation
        # overlap: the gas on a second stream beside the gravity, forked after
        # the drift and joined before the forces are summed
        self.gas_stream = torch.cuda.Stream(device=dev) if overlap and self.ng else None
        self._fork, self._join = torch.cuda.Event(), torch.cuda.Event()

    # -- forces ----------------------------------------------------------------

    def _forces(self, active, rebuild, t):
        """active: uint8 per particle, or None for all."""
        torch, F, gs = self.torch, self.F, self.gas_stream
        if gs is not None:
            self._fork.record()
            gs.wait_event(self._fork)
        F.gravity(self.pos, self.mass, active=active, rebuild=rebuild)
        kw = {}
        if self.first == "learned" and self.prev is not None:
            kw = dict(first="learned", dk_prev=self.prev[0], divv_prev=self.prev[1],
                      dtc_prev=self.prev[2], dt=max(t - self.t_hydro, 1e-30), model=self.model)
        elif self.first == "prev" and self.prev is not None:
            kw = dict(first="prev", dk_prev=self.prev[0])
        else:
            kw = dict(first="plain")
        if self.ng:
            with torch.cuda.stream(gs) if gs is not None else contextlib.nullcontext():
                self.passes.append(F.hydro(self.pos[:self.ng], self.vel[:self.ng],
                                           self.mass[:self.ng], self.u[:self.ng], **kw))
                if self.ng > 40:
                    self.prev = (F.dk.clone(), F.divv.clone(), F.dt_c.clone())
                    self.t_hydro = t
            if gs is not None:
                self._join.record(gs)
                torch.cuda.current_stream().wait_event(self._join)
        F.combine()

    def _ladder(self):
        """dt_min = the 1st percentile of dt_ind (linear interpolation, FP64),
        dt_max = dt_min 2^(rungs-1); the rungs follow in the kick of mode 2."""
        self.dt_min = float(self.torch.quantile(self.F.dt_ind, 0.01).item())
        self.dt_max = self.dt_min * 2 ** (self.R - 1)

    def _kick(self, mode, active):
        """gf_step_kick_dev: mode 0 half kicks of the active particles by
        dt_max 2^-rung / 2 (v, and u of the gas with du/dt); mode 1 first
        moves the active particles to the rung of their dt_ind, to a longer
        step only where this sub-step is also its boundary; mode 2 first moves
        every particle to it (a new ladder).  No host synchronisation."""
        F, ng = self.F, self.ng
        _lib.check(self.L.gf_step_kick_dev(
            self.n, ng, self.k, self.R, self.dt_max, mode, active.data_ptr(), F.acc.data_ptr(),
            F.dudt.data_ptr() if ng else None, F.dt_ind.data_ptr(), self.rung.data_ptr(),
            self.vel.data_ptr(), self.u.data_ptr() if ng else None, self.u_floor, F.sp))

    # -- diagnostics -----------------------------------------------------------

    def energy(self):
        """(kinetic, thermal, potential, |L|) of the synchronised state."""
        torch, _lib = self.torch, _lib_module()
        m = self.mass
        ke = 0.5 * torch.sum(m * torch.sum(self.vel * self.vel, dim=1))
        te = torch.sum(m[:self.ng] * self.u[:self.ng]) if self.ng else torch.zeros((), dtype=m.dtype)
        _lib.check(self.L.gf_potential_dev(self.n, self.pos.data_ptr(), m.data_ptr(),
                                           self.phi.data_ptr(), self.eps, None))
        torch.cuda.synchronize()
        pe = 0.5 * torch.sum(m * self.phi)
        Lv = torch.sum(m[:, None] * torch.linalg.cross(self.pos, self.vel), dim=0)
        return float(ke), float(te), float(pe), float(torch.linalg.vector_norm(Lv))

    @property
    def active_fraction(self):
        return [float(c.item()) / self.n for c in self.active_counts]

    def angular_momentum(self):
        torch = self.torch
        Lv = torch.sum(self.mass[:, None] * torch.linalg.cross(self.pos, self.vel), dim=0)
        return float(torch.linalg.vector_norm(Lv))

    # -- integration -------------------------------------------------------------

    def start(self):
        """Forces of every particle, rungs, opening half kicks."""
        torch = self.torch
        self.t = 0.0
        self.k = 0
        self._forces(None, True, 0.0)
        self._ladder()
        self._kick(2, torch.zeros_like(self.active))           # the rungs only
        self.E0 = self.energy()
        self._kick(0, torch.ones_like(self.active))

    def substep(self):
        torch = self.torch
        R = self.R
        self.k += 1
        self.t += self.dt_min
        if self.k >= len(self._counts):
            grown = torch.zeros(2 * len(self._counts), dtype=torch.int32, device=self.dev)
            grown[:len(self._counts)] = self._counts
            self._counts = grown
        self.active_counts.append(self._counts[self.k])
        # every particle drifts; active: the particles at a boundary of their step
        _lib.check(self.L.gf_step_drift_dev(self.n, self.dt_min, self.k, R, self.pos.data_ptr(),
                                            self.vel.data_ptr(), self.rung.data_ptr(),
                                            self.active.data_ptr(),
                                            self._counts.data_ptr() + 4 * self.k, self.F.sp))
        long_step = self.k % 2 ** (R - 1) == 0
        self._forces(self.active, long_step or not self.refit, self.t)
        self._kick(0, self.active)                             # closing half kicks
        if long_step:
            # every particle is at a boundary: a new ladder of steps
            self.L_sync.append(self.angular_momentum())
            self._ladder()
            self._kick(2, self.active)                         # opening half kicks
        else:
            self._kick(1, self.active)                         # new rungs, opening half kicks

    def synchronised_energy(self):
        """Energy and |L| at a long-step boundary (every particle just took its
        opening half kick with the forces of this sub-step): the kicks are
        taken back on copies, measured, and the state restored."""
        torch = self.torch
        assert self.k % 2 ** (self.R - 1) == 0, "not at a long-step boundary"
        v0, u0 = self.vel.clone(), self.u.clone()
        h = 0.5 * self.dt_max / 2.0 ** self.rung.double()
        self.vel -= h[:, None] * self.F.acc
        if self.ng:
            self.u[:self.ng] = self.u[:self.ng] - h[:self.ng] * self.F.dudt
        try:
            return self.energy()
        finally:
            self.vel.copy_(v0)
            self.u.copy_(u0)


def _lib_module():
    return _lib
