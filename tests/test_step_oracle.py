"""CPU pins of the step-loop oracle (oracle/step_oracle.py) by mathematics:
the leapfrog's closed form on linear forces, its exactness for constant
forces whatever the rungs do, hand-worked rung ladders, and the reference
caller's formulas for dt_ind and dt_grav (simulation.py:68-71)."""

import numpy as np

from oracle import step_oracle as S


def _fixed(acc_fn, dt_fn, ng=0):
    """A force function of positions only, with dt_ind from dt_fn(call)."""
    calls = [0]

    def forces(pos, vel, mass, u, active=None):
        c = calls[0]
        calls[0] += 1
        return acc_fn(pos), np.zeros(ng), dt_fn(c), 0.0
    return forces


def _leapfrog_closed_form(x0, v0, w, tau, n):
    """n kick-drift-kick steps of size tau on x'' = -w^2 x, by the Chebyshev
    identity M^n = (sin n phi M - sin (n-1) phi I) / sin phi of the one-step
    map M (det 1, trace 2 cos phi = 2 - (w tau)^2)."""
    c = 1.0 - 0.5 * (w * tau) ** 2
    M = np.array([[c, tau], [-w * w * tau * (1.0 - 0.25 * (w * tau) ** 2), c]])
    phi = np.arccos(c)
    Mn = (np.sin(n * phi) * M - np.sin((n - 1) * phi) * np.eye(2)) / np.sin(phi)
    return Mn @ np.array([x0, v0])


def test_block_leapfrog_equals_closed_form_per_rung():
    """Independent oscillators on three rungs (steps h, 2h, 4h): after two
    long steps each follows the closed form of its own leapfrog, with the
    opening half kick taken back to synchronise the velocity."""
    h = 0.01
    dt_ind = np.array([h, 2 * h, 4 * h, 3 * h, 8.5 * h])
    step = np.array([h, 2 * h, 4 * h, 2 * h, 4 * h])       # by hand: 2^-ceil(log2(4h / dt))
    w = np.array([3.0, 1.5, 0.7, 2.2, 1.1])
    rng = np.random.default_rng(0)
    x0, v0 = rng.normal(size=(5, 3)), rng.normal(size=(5, 3))
    f = _fixed(lambda p: -(w ** 2)[:, None] * p, lambda c: dt_ind)
    B = S.BlockLeapfrog(x0, v0, np.ones(5), np.zeros(0), 0, f, rungs=3, u_floor=0.0, quantile=0.0)
    B.start()
    assert B.dt_min == h and list(B.rung) == [2, 1, 0, 1, 0]
    for _ in range(8):
        B.substep()
    vs = B.vel - 0.5 * step[:, None] * B.acc
    for i in range(5):
        nsteps = int(round(8 * h / step[i]))
        for d in range(3):
            x, v = _leapfrog_closed_form(x0[i, d], v0[i, d], w[i], step[i], nsteps)
            assert abs(B.pos[i, d] - x) < 1e-13 and abs(vs[i, d] - v) < 1e-13


def test_block_leapfrog_exact_for_constant_force_whatever_the_rungs():
    """A leapfrog is exact for a constant force, so at every long-step
    boundary x = x0 + v0 t + a t^2 / 2 and v = v0 + a t hold for every
    particle, through any sequence of rung changes."""
    rng = np.random.default_rng(1)
    n = 40
    a = rng.normal(size=(n, 3))
    x0, v0 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    h = 1e-2
    sched = rng.choice([h, 2 * h, 3 * h, 4 * h, 8 * h, 16 * h, 0.5 * h], size=(200, n))
    f = _fixed(lambda p: a.copy(), lambda c: sched[c])
    B = S.BlockLeapfrog(x0, v0, np.ones(n), np.zeros(0), 0, f, rungs=4, u_floor=0.0,
                        quantile=0.0)
    B.start()
    changes, t = 0, 0.0
    for k in range(1, 33):
        before = B.rung.copy()
        t += B.dt_min
        B.substep()
        changes += int((B.rung != before).sum())
        if k % 8 == 0:
            vs = B.vel - 0.5 * (B.dt_max / 2.0 ** B.rung)[:, None] * a
            np.testing.assert_allclose(B.pos, x0 + v0 * t + 0.5 * a * t * t, rtol=0, atol=1e-13)
            np.testing.assert_allclose(vs, v0 + a * t, rtol=0, atol=1e-13)
    assert changes > 50


def test_rung_changes_by_hand():
    """Rungs 0-2 (steps 4h, 2h, h).  Particle 0 asks for a longer step at
    sub-step 1: it keeps rung 2 until the long-step boundary at 4.  Particle
    1 (rung 1) asks for a shorter one at sub-step 2, where it is active: it
    moves at once.  Particle 2 (rung 0) asks for a shorter one at sub-step
    1, where it is inactive: nothing changes until it is active at 4."""
    h = 1.0
    base = np.array([h, 2 * h, 4 * h, h])                     # particle 3 fixes dt_min = h

    def dt(c):                                                # call c = sub-step c
        d = base.copy()
        if c >= 1:
            d[0] = 4 * h
            d[2] = h
        if c >= 2:
            d[1] = h
        return d
    f = _fixed(lambda p: np.zeros_like(p), dt)
    B = S.BlockLeapfrog(np.zeros((4, 3)), np.zeros((4, 3)), np.ones(4), np.zeros(0), 0, f,
                        rungs=3, u_floor=0.0, quantile=0.0)
    B.start()
    seq = [list(B.rung)]
    for _ in range(4):
        B.substep()
        seq.append(list(B.rung))
    assert seq == [[2, 1, 0, 2], [2, 1, 0, 2], [2, 2, 0, 2], [2, 2, 0, 2], [0, 2, 2, 2]]


def test_gas_energy_kicks_and_floor():
    """u takes the same half kicks as v with du/dt, floored at u_floor."""
    h = 0.5
    dudt = np.array([2.0, -10.0])

    def forces(pos, vel, mass, u, active=None):
        return np.zeros_like(pos), dudt, np.full(len(pos), h), 0.0
    B = S.BlockLeapfrog(np.zeros((2, 3)), np.zeros((2, 3)), np.ones(2), np.array([1.0, 1.0]), 2,
                        forces, rungs=1, u_floor=0.25, quantile=0.0)
    B.start()                               # opening half kick: u + h / 2 du/dt
    np.testing.assert_allclose(B.u, [1.5, 0.25])
    B.substep()                             # closing + opening: + h du/dt
    np.testing.assert_allclose(B.u, [2.5, 0.25])


def test_forces_glue_is_simulation_py_formulas():
    """acc = acc_grav + acc_sph; dt_ind = min(0.3 sqrt(eps / |a|),
    dt_courant) with dt_courant = inf off the gas; dt_grav from the largest
    |a|; no hydro at 40 gas particles or fewer (simulation.py:68-80)."""
    rng = np.random.default_rng(2)
    n, ng, eps = 300, 120, 0.05
    pos = rng.normal(size=(n, 3))
    vel = 0.1 * rng.normal(size=(n, 3))
    mass = np.full(n, 1.0 / n)
    u = np.full(n, 0.02)
    F = S.Forces(n, ng, eps, k=32)
    acc, dudt, dt_ind, dt_grav = F(pos, vel, mass, u)
    from oracle import oracle as O
    g = O.direct(pos, mass, eps)
    d, nb = O.knn(pos[:ng], 32)
    hh = np.maximum(0.5 * d[:, -1], 0.5 * eps)
    rho, a, du, dt_i, _ = O.sph_forces(pos[:ng], vel[:ng], mass[:ng], nb, hh, u[:ng], 5 / 3, 1, 2)
    exp = g.copy()
    exp[:ng] += a
    np.testing.assert_array_equal(acc, exp)
    amag = np.sqrt((exp ** 2).sum(1))
    dtg = 0.3 * np.sqrt(eps / amag)
    np.testing.assert_allclose(dt_ind[ng:], dtg[ng:], rtol=1e-15, atol=0)
    np.testing.assert_allclose(dt_ind[:ng], np.minimum(dtg[:ng], dt_i), rtol=1e-15, atol=0)
    assert abs(dt_grav - 0.3 * np.sqrt(eps / amag.max())) < 1e-15 * dt_grav
    np.testing.assert_array_equal(dudt, du)
    # 40 gas particles: gravity only
    F40 = S.Forces(n, 40, eps, k=32)
    acc40, dudt40, dt40, _ = F40(pos, vel, mass, u)
    np.testing.assert_array_equal(acc40, g)
    np.testing.assert_allclose(dt40, 0.3 * np.sqrt(eps / np.sqrt((g ** 2).sum(1))), rtol=1e-15,
                               atol=0)
    assert not dudt40.any()


def test_forces_keep_inactive_gravity_rows():
    rng = np.random.default_rng(3)
    n = 50
    pos, mass = rng.normal(size=(n, 3)), np.full(n, 0.02)
    F = S.Forces(n, 0, 0.05, k=8)
    a0 = F(pos, np.zeros_like(pos), mass, np.zeros(n))[0].copy()
    act = np.zeros(n, bool)
    act[::3] = True
    a1 = F(pos + 0.1, np.zeros_like(pos), mass, np.zeros(n), act)[0]
    np.testing.assert_array_equal(a1[~act], a0[~act])
    from oracle import oracle as O
    np.testing.assert_array_equal(a1[act], O.direct(pos + 0.1, mass, 0.05)[act])
