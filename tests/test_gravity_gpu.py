"""GPU parity of the Barnes-Hut gravity path (libgf.so) against the reference
fixtures (tests/golden, written by the reference cpp/nbody.cpp built from
source) and the CPU oracle (oracle/gf_oracle.c)."""

import numpy as np
import pytest

from galaxy_formation_amd import ics
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from galaxy_formation_amd.engine import Engine
    return Engine("fp64")


def relerr(a, b):
    nb = np.linalg.norm(b, axis=1)
    d = np.linalg.norm(a - b, axis=1)
    return np.where(nb > 0, d / np.where(nb > 0, nb, 1), d)


def run(eng, pos, mass, theta, eps, precision, active=None, rebuild=True, acc=None):
    eng.precision = precision
    if acc is None:
        acc = np.zeros_like(pos)
    eng.gravity(pos, mass, acc, theta, eps, active=active, rebuild=rebuild)
    return acc


CASES = {
    "bh_gauss3k_t05": lambda: ics.gaussian_blob(3000, 1),
    "bh_cloud10k_t05": lambda: (ics.config("cloud10k").pos, ics.config("cloud10k").mass),
}


def _massless2k():
    rng = np.random.default_rng(5)
    pos = np.ascontiguousarray(rng.uniform(-1, 1, size=(2000, 3)))
    mass = rng.uniform(0.5, 1.5, 2000) / 2000
    mass[rng.random(2000) < 0.1] = 0.0
    return pos, mass


CASES["bh_massless2k_t05"] = _massless2k


@pytest.mark.parametrize("name", sorted(CASES))
def test_tree_topology_equals_reference(eng, golden, name):
    """Node ranges, leaf flags, next pointers, tree order and groups are the
    reference's (cpp/nbody.cpp:202-256), bit for bit."""
    g = golden(name)
    pos, mass = CASES[name]()
    assert ics.sha256(pos, mass) == str(g["input_sha"])
    run(eng, pos, mass, float(g["theta"]), float(g["eps"]), "fp64")
    t = eng.tree_export()
    for k in ("start", "count", "next", "leaf", "idx", "groups"):
        np.testing.assert_array_equal(t[k], g["tree_" + k], err_msg=k)
    # reference CoM: FP64 sums of FP32-rounded children, stored FP32: equal
    # to a few FP32 ulps at the coordinate scale
    scale = np.abs(pos).max()
    np.testing.assert_allclose(t["com4"][:, :3].astype(np.float32), g["tree_com4"][:, :3],
                               rtol=0, atol=4 * 2.0 ** -23 * scale)
    np.testing.assert_allclose(t["com4"][:, 3].astype(np.float32), g["tree_com4"][:, 3], rtol=4e-7)


@pytest.mark.parametrize("name", ["bh_cloud200k_t05", "bh_disk1m_t05", "bh_gauss20k_t07"])
def test_tree_topology_hash_equals_reference(eng, golden, name):
    g = golden(name)
    if name.startswith("bh_gauss"):
        pos, mass = ics.gaussian_blob(20000, 0)
    else:
        c = ics.config(name.split("_")[1])
        pos, mass = c.pos, c.mass
    assert ics.sha256(pos, mass) == str(g["input_sha"])
    run(eng, pos, mass, float(g["theta"]), float(g["eps"]), "fp64")
    nn, ng, deep = eng.tree_info()
    assert deep == 0
    assert (nn, ng) == (int(g["n_nodes"]), int(g["n_groups"]))
    t = eng.tree_export()
    assert ics.sha256(t["start"], t["count"], t["next"], t["leaf"]) == str(g["topo_sha"])
    assert ics.sha256(t["idx"]) == str(g["idx_sha"])
    assert ics.sha256(t["groups"]) == str(g["groups_sha"])


def test_single_pass_scans_repeat_and_never_give_up(eng):
    """The chain and group scans of a 200k build exceed one block's 65,536
    entries and run as one launch with decoupled look-back, whose state
    carries over from launch to launch (an epoch per launch, never cleared).
    Builds of other sizes in between, then the same build again, give the
    same tree, and no look-back wait was abandoned."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    c = ics.config("cloud200k")
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
    first = eng.tree_export()
    for n in (70000, 150000, 120001):
        run(eng, np.ascontiguousarray(c.pos[:n]), np.ascontiguousarray(c.mass[:n]), 0.5, c.eps,
            "fp64")
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
    again = eng.tree_export()
    for k in ("start", "count", "next", "leaf", "idx", "groups"):
        np.testing.assert_array_equal(again[k], first[k], err_msg=k)
    assert L.gf_lookback_timeouts() == 0


EDGE = 65536 + 4096 * 4


@pytest.mark.parametrize("n", [EDGE, EDGE + 1, EDGE + 3, EDGE + 12, EDGE + 4095])
def test_single_pass_build_steps_at_tile_edges_equal_oracle(eng, n):
    """Node bases, tree order and groups from the single-pass launches
    (tiles of 4096 items, 4 a thread) equal the oracle's for particle counts
    at and just past a tile boundary, where a thread's items and its window
    of LCPs run past the end of the arrays."""
    rng = np.random.default_rng(n)
    pos = np.ascontiguousarray(rng.normal(size=(n, 3)))
    mass = np.full(n, 1.0 / n)
    active = np.zeros(n, np.uint8)              # the build alone: no group is walked
    run(eng, pos, mass, 0.5, 0.01, "fp64", active=active)
    t = eng.tree_export()
    O.accel(pos, mass, 0.5, 0.01, active=active, mode=1)
    o = O.tree_export()
    for k in ("start", "count", "next", "leaf", "idx", "groups"):
        np.testing.assert_array_equal(t[k], o[k], err_msg=k)


def test_tree_fp64_centres_equal_oracle(eng):
    """FP64 centres of mass and opening radii are the oracle's, bit for bit
    (same summation order, no contraction)."""
    c = ics.make_cloud(3000, 3000, seed=4)
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
    t = eng.tree_export()
    O.accel(c.pos, c.mass, 0.5, c.eps, mode=1)
    o = O.tree_export()
    np.testing.assert_array_equal(t["box4"], o["box4"])
    np.testing.assert_array_equal(t["com4"], o["dcom4"])
    np.testing.assert_array_equal(t["open_r"], o["dopen_r"])


@pytest.mark.parametrize("top", [1, 16])
def test_moments_above_subtrees_fallback(eng, top):
    """The build finishes the nodes above its LDS subtrees (41 of them for
    this cloud) from a depth-sorted LDS list, or, past the list's capacity
    (lowered here to `top`), from the tree's arrays at every depth: the same
    centres of mass and opening radii, bit for bit and equal to the
    oracle's, and the same quadrupoles (through the quadrupole walk)."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    c = ics.make_cloud(20000, 20000, seed=7)
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
    ref_t = eng.tree_export()
    ref_q = run(eng, c.pos, c.mass, 0.5, c.eps, "fp32q")
    _lib.check(L.gf_set_test_limits(0, 0, top))
    try:
        run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
        t = eng.tree_export()
        q = run(eng, c.pos, c.mass, 0.5, c.eps, "fp32q")
    finally:
        _lib.check(L.gf_set_test_limits(0, 0, 0))
    nn = len(t["start"])
    size = np.minimum(t["next"], nn) - np.arange(nn)
    assert (size > 256).sum() > top
    np.testing.assert_array_equal(t["com4"], ref_t["com4"])
    np.testing.assert_array_equal(t["open_r"], ref_t["open_r"])
    np.testing.assert_array_equal(q, ref_q)
    O.accel(c.pos, c.mass, 0.5, c.eps, mode=1)
    o = O.tree_export()
    np.testing.assert_array_equal(t["com4"], o["dcom4"])
    np.testing.assert_array_equal(t["open_r"], o["dopen_r"])


@pytest.mark.parametrize("theta", [0.4, 0.5, 0.7])
def test_fp64_equals_oracle(eng, theta):
    c = ics.config("cloud10k")
    acc = run(eng, c.pos, c.mass, theta, c.eps, "fp64")
    inter = eng.last_interactions()
    ref, oi = O.accel(c.pos, c.mass, theta, c.eps, mode=1)
    assert inter == oi
    assert relerr(acc, ref).max() < 1e-12


@pytest.mark.parametrize("theta", [0.4, 0.5, 0.7])
def test_fp32_close_to_oracle_fp64(eng, theta):
    c = ics.config("cloud10k")
    acc = run(eng, c.pos, c.mass, theta, c.eps, "fp32")
    ref, _ = O.accel(c.pos, c.mass, theta, c.eps, mode=1)
    e = relerr(acc, ref)
    assert np.sqrt(np.mean(e ** 2)) < 2e-6 and e.max() < 5e-5


@pytest.mark.parametrize("name", ["bh_cloud10k_t04", "bh_cloud10k_t05", "bh_cloud10k_t07",
                                  "bh_cloud200k_t04", "bh_cloud200k_t05", "bh_cloud200k_t07",
                                  "bh_disk1m_t05"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_accel_close_to_reference(eng, golden, name, precision):
    """Against the reference's own FP32 Barnes-Hut at the same theta: the
    same tree and acceptance rule, so differences are FP32 rounding of the
    reference plus the rare node whose acceptance flips under it."""
    g = golden(name)
    c = ics.config(name.split("_")[1])
    acc = run(eng, c.pos, c.mass, float(g["theta"]), float(g["eps"]), precision)
    e = relerr(acc[g["sub"]], g["acc_sub"])
    assert np.median(e) < 1e-5
    assert np.quantile(e, 0.99) < 1e-4
    assert e.max() < 5e-3


def test_barnes_hut_matches_direct_sum(eng):
    """tests/test_physics.py:20-30 of the reference: median < 1 %."""
    pos, mass = ics.gaussian_blob(20000, 0)
    for p in ("fp32", "fp64"):
        acc = run(eng, pos, mass, 0.7, 0.01, p)
        exact = eng.gravity_direct(pos, mass, eps=0.01)
        assert np.median(relerr(acc, exact)) < 1e-2


def test_direct_equals_oracle(eng):
    c = ics.make_cloud(3000, 3000, seed=9)
    a = eng.gravity_direct(c.pos, c.mass, c.eps)
    b = O.direct(c.pos, c.mass, c.eps)
    assert relerr(a, b).max() < 1e-13
    sub = np.array([5, 17, 4000, 5999])
    np.testing.assert_allclose(eng.gravity_direct_subset(c.pos, c.mass, c.eps, sub), b[sub],
                               rtol=1e-13)


def test_theta_zero_is_direct_sum(eng):
    """theta -> 0 opens every node down to the leaves: the direct sum."""
    c = ics.make_cloud(2000, 2000, seed=2)
    acc = run(eng, c.pos, c.mass, 1e-6, c.eps, "fp64")
    assert relerr(acc, O.direct(c.pos, c.mass, c.eps)).max() < 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp64", "fp32q"])
def test_active_subset(eng, prec):
    """Inactive rows are left bit-unchanged; active rows are the all-active
    result (nbody.cpp:344-348, :393, :404-407).  Only a group's active
    members are evaluated, compacted onto the lanes, so a particle's list is
    split into other slices than when all are active: the FP32 partial sums
    then round differently: within 5e-5 (the FP32 tolerance against the
    oracle), and bit-equal in FP64 up to the order of the slice sums."""
    from galaxy_formation_amd.engine import Engine
    e = eng if prec == "fp32" else Engine(prec)
    c = ics.config("cloud10k")
    full = run(e, c.pos, c.mass, 0.5, c.eps, prec)
    on = np.zeros(c.n, bool)
    for frac, seed in ((0.2, 3), (0.7, 4), (0.02, 5)):
        act = (np.random.default_rng(seed).random(c.n) < frac).astype(np.uint8)
        acc = np.full_like(c.pos, np.nan)
        run(e, c.pos, c.mass, 0.5, c.eps, prec, active=act, acc=acc)
        on = act.astype(bool)
        err = relerr(acc[on], full[on])
        if prec == "fp64":
            assert err.max() < 1e-13
        else:
            assert err.max() < 5e-5 and np.median(err) < 1e-6
        assert np.isnan(acc[~on]).all()
    acc0 = np.full_like(c.pos, 3.5)
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp32", active=np.zeros(c.n, np.uint8), acc=acc0)
    assert (acc0 == 3.5).all()


def test_refit_matches_oracle_and_reference(eng, golden):
    """rebuild = 0 refits the previous tree (nbody.cpp:262-305): same cells,
    new centres of mass, opening radii enlarged by the largest displacement."""
    g = golden("bh_refit_cloud10k_t05")
    c = ics.config("cloud10k")
    assert ics.sha256(c.pos, c.mass) == str(g["input_sha"])
    pos1 = np.ascontiguousarray(c.pos + float(g["drift"]) * c.vel)
    act = g["active"]
    for p in ("fp64", "fp32"):
        run(eng, c.pos, c.mass, 0.5, c.eps, p)
        acc1 = np.full((c.n, 3), 7.25)
        run(eng, pos1, c.mass, 0.5, c.eps, p, active=act, rebuild=False, acc=acc1)
        on = act.astype(bool)
        assert (acc1[~on] == 7.25).all()
        e = relerr(acc1[on], g["acc1"][on])
        assert np.median(e) < 1e-5 and e.max() < 5e-3
        O.accel(c.pos, c.mass, 0.5, c.eps, mode=1)
        ref1 = np.full((c.n, 3), 7.25)
        O.accel(pos1, c.mass, 0.5, c.eps, active=act, rebuild=False, mode=1, acc=ref1)
        e = relerr(acc1[on], ref1[on])
        assert e.max() < (1e-12 if p == "fp64" else 5e-5)
    # refit with zero displacement is the build
    run(eng, c.pos, c.mass, 0.5, c.eps, "fp64")
    t0 = eng.tree_export()
    a0 = run(eng, c.pos, c.mass, 0.5, c.eps, "fp64", rebuild=False)
    t1 = eng.tree_export()
    np.testing.assert_array_equal(t0["com4"], t1["com4"])
    np.testing.assert_array_equal(t0["open_r"], t1["open_r"])
    np.testing.assert_array_equal(a0, run(eng, c.pos, c.mass, 0.5, c.eps, "fp64"))


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9, 65, 700])
def test_small_and_ragged_sizes(eng, n):
    pos, mass = ics.gaussian_blob(n, 7)
    for p in ("fp32", "fp64"):
        acc = run(eng, pos, mass, 0.5, 0.05, p)
        ref, _ = O.accel(pos, mass, 0.5, 0.05, mode=1)
        assert relerr(acc, ref).max() < (1e-12 if p == "fp64" else 5e-5)


def test_every_group_size_equals_oracle(eng):
    """A cloud of n <= 64 particles is one group of n particles, so n = 1..64
    runs the evaluation with every chunk size: one part of 64 / n slices, or
    two parts (32, 16 or 8 particles, then the rest) with their own slices."""
    for n in range(1, 65):
        pos, mass = ics.gaussian_blob(n, 100 + n)
        ref, _ = O.accel(pos, mass, 0.5, 0.05, mode=1)
        ref2, _ = O.accel(pos, mass, 0.5, 0.05, mode=2)
        for p, r, tol in (("fp64", ref, 1e-12), ("fp32", ref, 5e-5), ("fp32q", ref2, 5e-5)):
            acc = run(eng, pos, mass, 0.5, 0.05, p)
            err = relerr(acc, r).max()
            assert err < tol, (n, p, err)


def test_empty_input(eng):
    acc = np.zeros((0, 3))
    eng.gravity(np.zeros((0, 3)), np.zeros(0), acc, 0.5, 0.1)


def test_clustered_duplicates(eng):
    """Many particles at identical positions (deep cells)."""
    rng = np.random.default_rng(12)
    pos = np.ascontiguousarray(np.repeat(rng.normal(size=(300, 3)), 6, axis=0))
    mass = np.full(len(pos), 1.0 / len(pos))
    acc = run(eng, pos, mass, 0.5, 0.05, "fp64")
    ref, _ = O.accel(pos, mass, 0.5, 0.05, mode=1)
    assert relerr(acc, ref).max() < 1e-12


def test_deep_cells_at_the_key_depth(eng):
    """Sites of 12 particles each, coincident or a few 1e-10 apart: the
    21-level keys cannot separate them, so their cells become leaves of
    more than 8 particles (counted as deep cells), where the reference
    divides them down to depth 48 (cpp/nbody.cpp:122).  A node of
    coincident particles has its monopole at their common position, so the
    accelerations equal the oracle's (which follows the reference's depth)
    to FP64 rounding."""
    rng = np.random.default_rng(21)
    sites = rng.normal(size=(400, 3))
    pos = np.repeat(sites, 12, axis=0)
    pos[len(pos) // 2:] += rng.normal(scale=1e-10, size=(len(pos) - len(pos) // 2, 3))
    pos = np.ascontiguousarray(pos)
    mass = rng.uniform(0.5, 1.5, len(pos)) / len(pos)
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6)):
        acc = run(eng, pos, mass, 0.5, 0.05, prec)
        assert eng.tree_info()[2] > 0                       # deep cells present
        ref, _ = O.accel(pos, mass, 0.5, 0.05, mode=1)
        assert relerr(acc, ref).max() < tol, prec


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_groups_over_64_particles_through_the_global_stack(eng, prec):
    """Sites of 100 coincident particles: each site is a depth-capped leaf
    and a group of two 64-particle chunks.  With the LDS stack cut to 8
    entries every group is redone by the global-stack walk: accelerations
    and the interaction count equal the unconstrained walk's, and the
    accelerations equal the oracle's to FP64 rounding (FP64)."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(22)
    pos = np.ascontiguousarray(np.repeat(rng.normal(size=(60, 3)), 100, axis=0))
    mass = rng.uniform(0.5, 1.5, len(pos)) / len(pos)
    ref = run(eng, pos, mass, 0.5, 0.05, prec)
    inter = eng.last_interactions()
    assert eng.tree_info()[2] >= 60
    _lib.check(L.gf_set_test_limits(8, 0, 0))
    try:
        acc = run(eng, pos, mass, 0.5, 0.05, prec)
        assert eng.last_interactions() == inter
    finally:
        _lib.check(L.gf_set_test_limits(0, 0, 0))
    np.testing.assert_array_equal(acc, ref)
    if prec == "fp64":
        oref, _ = O.accel(pos, mass, 0.5, 0.05, mode=1)
        assert relerr(acc, oref).max() < 1e-12


@pytest.mark.parametrize("limit", [1, 64, 300])
def test_global_stack_walk_equals_lds_walk(eng, limit):
    """With the LDS stack cut to `limit` entries, groups overflow and are
    walked again from global-memory stacks: the same nodes in the same order,
    hence bit-identical accelerations and interaction counts."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    c = ics.config("cloud10k")
    ref = {p: (run(eng, c.pos, c.mass, 0.5, c.eps, p), eng.last_interactions())
           for p in ("fp32", "fp64")}
    _lib.check(L.gf_set_test_limits(limit, 0, 0))
    try:
        for p in ("fp32", "fp64"):
            acc = run(eng, c.pos, c.mass, 0.5, c.eps, p)
            np.testing.assert_array_equal(acc, ref[p][0])
            assert eng.last_interactions() == ref[p][1]
    finally:
        _lib.check(L.gf_set_test_limits(0, 0, 0))


@pytest.mark.parametrize("theta", [0.4, 0.5, 0.7])
def test_fp32_quadrupole_close_to_oracle(eng, theta):
    """Mode 2 (FP32 evaluation with node quadrupoles) against the oracle's
    FP64 mode 2: the same interaction set, FP32 rounding apart.  Tolerance 0:
    every accepted node with its quadrupole, as the oracle."""
    c = ics.config("cloud10k")
    tol = eng.quad_tolerance
    eng.quad_tolerance = 0.0
    try:
        acc = run(eng, c.pos, c.mass, theta, c.eps, "fp32q")
    finally:
        eng.quad_tolerance = tol
    inter = eng.last_interactions()
    ref, oi = O.accel(c.pos, c.mass, theta, c.eps, mode=2)
    assert inter == oi
    e = relerr(acc, ref)
    assert np.sqrt(np.mean(e ** 2)) < 2e-6 and e.max() < 5e-5


def test_quadrupoles_reduce_the_error_vs_direct_sum(eng):
    """The node quadrupoles cut the Barnes-Hut truncation error at the same
    opening angle (and the same interaction count)."""
    c = ics.config("cloud10k")
    exact = eng.gravity_direct(c.pos, c.mass, c.eps)
    for theta in (0.5, 0.7):
        e1 = relerr(run(eng, c.pos, c.mass, theta, c.eps, "fp32"), exact)
        e2 = relerr(run(eng, c.pos, c.mass, theta, c.eps, "fp32q"), exact)
        assert np.sqrt(np.mean(e2 ** 2)) < 0.4 * np.sqrt(np.mean(e1 ** 2))


def test_quadrupole_tolerance_trades_little_accuracy():
    """The default tolerance (0.01) evaluates the nodes whose quadrupole term
    is negligible as monopoles: fewer quadrupole entries, the same entry
    count, and an error vs the direct sum within the 1e-3 gate and close to
    the all-quadrupole one."""
    from galaxy_formation_amd import _lib
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp32q")
    L = _lib.load()
    pos, mass = ics.gaussian_blob(20000, 3)
    eps = 0.01
    exact = e.gravity_direct(pos, mass, eps)
    res = {}
    for tol in (0.0, 0.01):
        e.quad_tolerance = tol
        acc = run(e, pos, mass, 0.5, eps, "fp32q")
        nodes = np.zeros(1, np.int64)
        _lib.check(L.gf_node_interactions_dev(_lib.ptr(nodes)))
        res[tol] = (acc, e.last_interactions(), int(nodes[0]))
    e.quad_tolerance = 0.01
    (a0, i0, q0), (a1, i1, q1) = res[0.0], res[0.01]
    assert i0 == i1 and q1 < q0
    r0 = np.sqrt(np.mean(relerr(a0, exact) ** 2))
    r1 = np.sqrt(np.mean(relerr(a1, exact) ** 2))
    assert r1 < 1e-3 and r1 < 2.0 * r0


def test_fused_quadrupole_pass_equals_separate_pass():
    """The first quadrupole walk on a tree computes the quadrupoles in a pass
    of their own; later builds compute them with the centres of mass.  The
    accelerations are bit-identical."""
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp64")
    c = ics.config("cloud10k")
    run(e, c.pos, c.mass, 0.5, c.eps, "fp64")           # a monopole walk: no quadrupoles wanted
    a1 = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q")      # separate pass
    a2 = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q")      # fused into the build
    np.testing.assert_array_equal(a1, a2)
    a3 = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q", rebuild=False)   # refit, fused
    np.testing.assert_array_equal(a1, a3)


def _prev_gravity(eng, c, dt):
    """FP64 gravity of the particles drifted back by dt: the previous step's
    accelerations a simulation holds in its acc array."""
    return run(eng, np.ascontiguousarray(c.pos - dt * c.vel), c.mass, 0.5, c.eps, "fp64")


def test_quad_relative_rule_without_history_is_the_share_rule():
    """With the relative rule on, a group with an active member whose entry
    row is zero or not finite keeps the share rule: entry accelerations of
    zeros, or a NaN in every group, give bit-identical results to the rule
    switched off.  Inactive rows do not count: huge values there change
    nothing either."""
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp32q")
    c = ics.config("cloud10k")
    base = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q")
    try:
        e.quad_relative_tolerance = 1e-4
        np.testing.assert_array_equal(run(e, c.pos, c.mass, 0.5, c.eps, "fp32q"), base)
        nan = np.full_like(c.pos, np.nan)
        np.testing.assert_array_equal(run(e, c.pos, c.mass, 0.5, c.eps, "fp32q", acc=nan), base)
        # half the particles active: the inactive rows hold 1e300, the active
        # ones zeros.  Only active members are evaluated, on other slices than
        # when all are active, so the active rows match within FP32 rounding
        act = (np.arange(c.n) % 2).astype(np.uint8)
        entry = np.where(act[:, None] == 1, np.zeros_like(c.pos), 1e300)
        got = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q", active=act, acc=entry)
        on = act == 1
        assert (got[~on] == 1e300).all()
        err = relerr(got[on], base[on])
        assert err.max() < 2e-5 and np.median(err) < 1e-6
    finally:
        e.quad_relative_tolerance = 0.0


def test_quad_relative_rule_trades_little_accuracy():
    """Fed the previous step's gravity, the relative rule keeps the
    interaction set, evaluates fewer quadrupoles than the share rule, stays
    within the 1e-3 gate and close to the all-quadrupole error, and depends
    little on whether the entry rows are the previous step's or exact."""
    from galaxy_formation_amd import _lib
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp32q")
    L = _lib.load()
    c = ics.config("cloud10k")
    exact = e.gravity_direct(c.pos, c.mass, c.eps)
    prev = _prev_gravity(e, c, 0.5)
    now = run(e, c.pos, c.mass, 0.5, c.eps, "fp64")
    assert np.sqrt(np.mean(relerr(prev, now) ** 2)) > 1e-3      # a real previous step

    def quad(tol, rel, entry):
        e.quad_tolerance, e.quad_relative_tolerance = tol, rel
        acc = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q", acc=entry.copy())
        nodes = np.zeros(1, np.int64)
        _lib.check(L.gf_node_interactions_dev(_lib.ptr(nodes)))
        return acc, e.last_interactions(), int(nodes[0])

    try:
        a0, i0, q0 = quad(0.0, 0.0, prev)         # every accepted node with its quadrupole
        a1, i1, q1 = quad(0.01, 0.0, prev)        # the share rule
        a2, i2, q2 = quad(0.01, 1e-4, prev)       # the relative rule
        a3, i3, q3 = quad(0.01, 1e-4, exact)
    finally:
        e.quad_tolerance, e.quad_relative_tolerance = 0.01, 0.0
    assert i0 == i1 == i2 == i3
    assert q2 < q1 < q0
    r0, r2, r3 = (np.sqrt(np.mean(relerr(a, exact) ** 2)) for a in (a0, a2, a3))
    assert r2 < 1e-3 and r2 < 1.5 * r0
    assert abs(r2 - r3) < 0.1 * r0


def test_quad_relative_tolerance_orders_the_quadrupole_count():
    """A larger relative tolerance demotes more nodes: the quadrupole entries
    fall and the error against the direct sum does not fall."""
    from galaxy_formation_amd import _lib
    from galaxy_formation_amd.engine import Engine
    e = Engine("fp32q")
    L = _lib.load()
    c = ics.config("cloud10k")
    exact = e.gravity_direct(c.pos, c.mass, c.eps)
    prev = _prev_gravity(e, c, 0.5)
    q, r = [], []
    try:
        for rel in (1e-6, 1e-5, 1e-4, 1e-3, 1e-2):
            e.quad_relative_tolerance = rel
            acc = run(e, c.pos, c.mass, 0.5, c.eps, "fp32q", acc=prev.copy())
            nodes = np.zeros(1, np.int64)
            _lib.check(L.gf_node_interactions_dev(_lib.ptr(nodes)))
            q.append(int(nodes[0]))
            r.append(np.sqrt(np.mean(relerr(acc, exact) ** 2)))
    finally:
        e.quad_relative_tolerance = 0.0
    assert all(x > y for x, y in zip(q, q[1:])), q
    assert r[-1] > r[0]


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp32q"])
def test_walk_grid_does_not_change_results(precision):
    """The walk draws groups from a queue and each group's list and sums are
    its own, so the persistent grid (gf_set_walk_blocks) changes nothing:
    bit-identical accelerations and the same interaction count for the
    default grid, a quarter of it and a grid that is not a multiple of the
    CU count."""
    from galaxy_formation_amd import _lib
    from galaxy_formation_amd.engine import Engine
    L = _lib.load()
    e = Engine(precision)
    c = ics.config("cloud10k")
    base = run(e, c.pos, c.mass, 0.5, c.eps, precision)
    i0 = e.last_interactions()
    try:
        for blocks in (1024, 257):
            _lib.check(L.gf_set_walk_blocks(blocks))
            np.testing.assert_array_equal(run(e, c.pos, c.mass, 0.5, c.eps, precision), base)
            assert e.last_interactions() == i0
    finally:
        _lib.check(L.gf_set_walk_blocks(0))


def test_walk_records_packed_ahead_equal_packed_by_the_walk():
    """gf_walk_prepare_dev packs the walk's records right after the build;
    the walk then reuses them.  Accelerations are bit-identical to a walk
    that packs its own records, for every precision, for a second walk on
    the same build, and when the records were packed for another precision
    (the walk packs them again)."""
    import torch

    from galaxy_formation_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    c = ics.config("cloud10k")
    pos = torch.from_numpy(np.ascontiguousarray(c.pos)).to(dev)
    mass = torch.from_numpy(np.ascontiguousarray(c.mass)).to(dev)
    n = len(c.pos)

    def walk(prec, prepare=None, twice=False):
        acc = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        _lib.check(L.gf_tree_build_dev(n, pos.data_ptr(), mass.data_ptr(), 0.5, None))
        if prepare is not None:
            _lib.check(L.gf_walk_prepare_dev(prepare, None))
        _lib.check(L.gf_walk_dev(n, None, acc.data_ptr(), c.eps, prec, None))
        if twice:
            _lib.check(L.gf_walk_dev(n, None, acc.data_ptr(), c.eps, prec, None))
        torch.cuda.synchronize()
        return acc.cpu().numpy()

    for prec in (0, 1, 2):
        base = walk(prec)
        np.testing.assert_array_equal(walk(prec, prepare=prec), base)
        np.testing.assert_array_equal(walk(prec, prepare=prec, twice=True), base)
        other = 2 if prec != 2 else 0
        np.testing.assert_array_equal(walk(prec, prepare=other), base)
