"""GPU parity of the SPH path (libgf.so) against the CPU oracle
(oracle/gf_oracle.c: brute-force kNN, the h rule and sph_forces restated in
FP64) and scipy's cKDTree (the reference's kNN library, sph.py:166)."""

import numpy as np
import pytest

from galaxy_formation_amd import ics, sph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def cloud_gas(n_gas, seed=3):
    c = ics.make_cloud(n_gas, n_gas, seed=seed)
    g = c.gas
    return (np.ascontiguousarray(c.pos[g]), np.ascontiguousarray(c.vel[g]),
            np.ascontiguousarray(c.mass[g]), np.ascontiguousarray(c.u[g]), c.eps)


def disk_gas(n_gas, seed=3):
    c = ics.make_disk(n_gas, n_gas, seed=seed)
    g = c.gas
    return (np.ascontiguousarray(c.pos[g]), np.ascontiguousarray(c.vel[g]),
            np.ascontiguousarray(c.mass[g]), np.ascontiguousarray(c.u[g]), c.eps)


def test_kernel_is_normalised():
    """tests/test_physics.py:33-38 of the reference."""
    h = 0.7
    r = np.linspace(0, 2 * h, 20001)
    w = sph.kernel(r, h)
    assert abs(np.trapezoid(4 * np.pi * r ** 2 * w, r) - 1) < 1e-4
    assert sph.kernel(0.3, 0.7) == O.kernel(0.3, 0.7)


def test_kernel_and_gradient_equal_oracle():
    rng = np.random.default_rng(1)
    h = rng.uniform(0.01, 2.0, 5000)
    r = h * rng.uniform(0, 2.5, 5000)
    r[:10] = 0.0
    w = sph.kernel(r, h)
    g = sph.kernel_grad(r, h)
    np.testing.assert_array_equal(w, [O.kernel(a, b) for a, b in zip(r, h)])
    np.testing.assert_array_equal(g, [O.kernel_grad(a, b) for a, b in zip(r, h)])


@pytest.mark.parametrize("k", [32, 48, 64, 96])
def test_knn_equals_bruteforce(k):
    """Exact kNN: neighbour lists (order included) and distances equal the
    brute-force oracle, bit for bit."""
    pos, *_ = cloud_gas(4000, seed=5)
    d, nb, info = sph.knn(pos, k)
    od, onb = O.knn(pos, k)
    np.testing.assert_array_equal(nb, onb)
    np.testing.assert_array_equal(d, od)
    assert (nb[:, 0] == np.arange(len(pos))).all()
    assert info["passes"] >= 1


def test_knn_matches_ckdtree_on_disk():
    from scipy.spatial import cKDTree
    pos, *_ = disk_gas(30000, seed=8)
    d, nb, _ = sph.knn(pos, 48)
    sd, snb = cKDTree(pos).query(pos, k=48)
    np.testing.assert_allclose(d, sd, rtol=1e-14, atol=0)
    same = (nb == snb).all(axis=1)
    # rows may only differ where two neighbours are at the same distance
    for i in np.flatnonzero(~same):
        assert set(nb[i]) == set(snb[i]) or np.any(np.diff(sd[i]) == 0)


@pytest.mark.parametrize("scale", [0.05, 1.0, 20.0])
def test_knn_independent_of_first_radius(scale):
    """The radius iteration converges to the same exact answer from any
    starting radius; only the number of passes changes."""
    pos, *_ = cloud_gas(3000, seed=6)
    d0, nb0, _ = sph.knn(pos, 32)
    r0 = np.full(len(pos), scale * np.median(d0[:, -1]))
    d, nb, info = sph.knn(pos, 32, r_init=r0)
    np.testing.assert_array_equal(nb, nb0)
    np.testing.assert_array_equal(d, d0)
    if scale < 0.1:
        assert info["passes"] > 1 and info["mean_passes"] > 1


def test_knn_small_inputs():
    rng = np.random.default_rng(2)
    for n, k in ((1, 1), (5, 5), (50, 7), (65, 64)):
        pos = rng.normal(size=(n, 3))
        d, nb, _ = sph.knn(pos, k)
        od, onb = O.knn(pos, k)
        np.testing.assert_array_equal(nb, onb)
        np.testing.assert_array_equal(d, od)


def test_sph_forces_equal_oracle():
    pos, vel, mass, u, eps = cloud_gas(5000, seed=7)
    od, onb = O.knn(pos, 32)
    h = O.smoothing_length(od[:, -1], eps)
    n = len(pos)
    rho, dudt, dt, divv = (np.empty(n) for _ in range(4))
    acc = np.empty((n, 3))
    sph.sph_forces(pos, vel, mass, onb, h, u, GAMMA, sph.ALPHA, sph.BETA, rho, acc, dudt, dt,
                   divv=divv)
    orho, oacc, odudt, odt, odivv = O.sph_forces(pos, vel, mass, onb, h, u, GAMMA, sph.ALPHA,
                                                 sph.BETA)
    np.testing.assert_allclose(rho, orho, rtol=1e-13)
    scale = np.abs(oacc).max()
    np.testing.assert_allclose(acc, oacc, rtol=1e-11, atol=1e-13 * scale)
    np.testing.assert_allclose(dudt, odudt, rtol=1e-11, atol=1e-13 * np.abs(odudt).max())
    np.testing.assert_allclose(dt, odt, rtol=1e-13)
    np.testing.assert_allclose(divv, odivv, rtol=1e-11, atol=1e-13 * np.abs(odivv).max())


def test_knn_from_candidates_equals_oracle():
    pos, *_ = cloud_gas(3000, seed=9)
    od, onb = O.knn(pos, 48)
    cand = onb.copy()
    rng = np.random.default_rng(4)
    perm = rng.permuted(cand, axis=1)          # candidates in any order
    perm[rng.random(perm.shape) < 0.05] = -1   # some became stars
    perm[:, 0] = np.arange(len(pos))
    ok, nb_o, d_o = O.knn_from_candidates(pos, perm, 32)
    nb = np.zeros((len(pos), 32), np.int64)
    d = np.zeros((len(pos), 32))
    ok_g = sph.knn_from_candidates(pos, perm, 32, nb, d)
    assert ok_g == ok
    if ok:
        np.testing.assert_array_equal(nb, nb_o)
        np.testing.assert_array_equal(d, d_o)
    # a row without enough candidates makes the call fail (sph.py:128-130)
    bad = perm.copy()
    bad[10, 5:] = -1
    assert not sph.knn_from_candidates(pos, bad, 32, nb, d)


def test_neighbour_finder_full_and_verlet():
    """NeighbourFinder (sph.py:138-170): the full search returns the exact k
    nearest; a Verlet re-selection after a small drift equals a full search
    when every true neighbour is still a candidate."""
    c = ics.make_cloud(3000, 3000, seed=12)
    g = c.gas
    nf = sph.NeighbourFinder(c.n)
    p = np.ascontiguousarray(c.pos[g])
    d, nb = nf.find(g, p, True)
    od, onb = O.knn(p, 32)
    np.testing.assert_array_equal(nb, onb)
    np.testing.assert_array_equal(d, od)
    p2 = np.ascontiguousarray(p + 1e-3 * c.vel[g])
    d2, nb2 = nf.find(g, p2, False)
    od2, onb2 = O.knn(p2, 32)
    np.testing.assert_array_equal(nb2, onb2)
    np.testing.assert_allclose(d2, od2, rtol=1e-15)


def test_neighbour_finder_after_star_formation():
    """The cache keeps global indices: after some gas particles turn into
    stars (they leave g), the re-selection drops them from the candidate
    lists and equals the oracle's knn_from_candidates over those lists
    (sph.py:155-163), and the exact neighbours where the true ones are still
    candidates, without a new search.  A gas list with a particle the cache
    does not know, or rows left with fewer than k candidates, fall back to a
    full search."""
    c = ics.make_cloud(3000, 3000, seed=31)
    g = c.gas
    nf = sph.NeighbourFinder(c.n)
    p = np.ascontiguousarray(c.pos[g])
    nf.find(g, p, True)
    owner, cand = nf.cache.export()
    assert (owner[g] == np.arange(len(g))).all() and (np.delete(owner, g) == -1).all()
    _, onb48 = O.knn(p, 48)
    np.testing.assert_array_equal(cand, g[onb48])
    # 3 % of the gas becomes stars
    rng = np.random.default_rng(1)
    keep = np.sort(rng.choice(len(g), int(0.97 * len(g)), replace=False))
    g2 = g[keep]
    p2 = np.ascontiguousarray(c.pos[g2] + 1e-4 * c.vel[g2])
    d2, nb2 = nf.find(g2, p2, False)
    assert nf.full_searches == 1
    loc = np.full(c.n, -1, np.int64)
    loc[g2] = np.arange(len(g2))
    ok, onb, od = O.knn_from_candidates(p2, loc[cand[owner[g2]]], 32)
    assert ok
    np.testing.assert_array_equal(nb2, onb)
    np.testing.assert_array_equal(d2, od)
    ed, enb = O.knn(p2, 32)
    served = (np.sort(onb, 1) == np.sort(enb, 1)).all(1)
    assert served.mean() > 0.99
    # a particle that was not gas at the search: full search
    g3 = np.concatenate([g2, [c.n - 1]])
    p3 = np.ascontiguousarray(c.pos[g3])
    d3, nb3 = nf.find(g3, p3, False)
    assert nf.full_searches == 2
    ed3, enb3 = O.knn(p3, 32)
    np.testing.assert_array_equal(nb3, enb3)
    np.testing.assert_array_equal(d3, ed3)
    # most of the gas gone: rows with fewer than 32 candidates left
    g4 = g3[::3]
    p4 = np.ascontiguousarray(c.pos[g4])
    d4, nb4 = nf.find(g4, p4, False)
    assert nf.full_searches == 3
    np.testing.assert_array_equal(nb4, O.knn(p4, 32)[1])


def test_hydro_pipeline_equals_oracle():
    """Device-resident pipeline (gas tree, kNN, h, density, forces) against
    the oracle's brute-force kNN + h rule + sph_forces."""
    import torch

    from galaxy_formation_amd import _lib
    pos, vel, mass, u, eps = disk_gas(20000, seed=4)
    n, k = len(pos), 64
    dev = torch.device("cuda:0")
    t = {name: torch.from_numpy(a).to(dev) for name, a in
         (("pos", pos), ("vel", vel), ("mass", mass), ("u", u))}
    out = {name: torch.zeros(shape, dtype=torch.float64, device=dev) for name, shape in
           (("h", n), ("rho", n), ("acc", (n, 3)), ("dudt", n), ("dt", n), ("divv", n))}
    torch.cuda.synchronize()
    L = _lib.load()
    passes = _lib.check(L.gf_hydro_dev(n, t["pos"].data_ptr(), t["vel"].data_ptr(),
                                       t["mass"].data_ptr(), t["u"].data_ptr(), eps, k, None,
                                       GAMMA, sph.ALPHA, sph.BETA,
                                       *[out[x].data_ptr() for x in ("h", "rho", "acc", "dudt",
                                                                     "dt", "divv")], None))
    _lib.check(L.gf_sync(None))
    assert passes >= 1
    od, onb = O.knn(pos, k)
    h = O.smoothing_length(od[:, -1], eps)
    orho, oacc, odudt, odt, odivv = O.sph_forces(pos, vel, mass, onb, h, u, GAMMA, sph.ALPHA,
                                                 sph.BETA)
    np.testing.assert_array_equal(out["h"].cpu().numpy(), h)
    np.testing.assert_allclose(out["rho"].cpu().numpy(), orho, rtol=1e-13)
    np.testing.assert_allclose(out["acc"].cpu().numpy(), oacc, rtol=1e-10,
                               atol=1e-12 * np.abs(oacc).max())
    np.testing.assert_allclose(out["dt"].cpu().numpy(), odt, rtol=1e-13)
    np.testing.assert_allclose(out["dudt"].cpu().numpy(), odudt, rtol=1e-10,
                               atol=1e-12 * np.abs(odudt).max())
    np.testing.assert_allclose(out["divv"].cpu().numpy(), odivv, rtol=1e-10,
                               atol=1e-12 * np.abs(odivv).max())


def test_hydro_pipeline_fp32_pairs_close_to_oracle():
    """The device pipeline with FP32 pair arithmetic (gf_set_sph_precision(1),
    the fast path's density and forces) against the oracle's FP64 kNN + h +
    sph_forces: h bit-equal (the neighbours do not depend on it), density
    within 1e-6 relative (the gate against the plain solve is 1e-4), forces,
    du/dt and div v within 2e-5 of their largest magnitude, Courant steps
    within 1e-5 relative."""
    import torch

    from galaxy_formation_amd import _lib
    pos, vel, mass, u, eps = disk_gas(20000, seed=4)
    n, k = len(pos), 64
    dev = torch.device("cuda:0")
    t = {name: torch.from_numpy(a).to(dev) for name, a in
         (("pos", pos), ("vel", vel), ("mass", mass), ("u", u))}
    out = {name: torch.zeros(shape, dtype=torch.float64, device=dev) for name, shape in
           (("h", n), ("rho", n), ("acc", (n, 3)), ("dudt", n), ("dt", n), ("divv", n))}
    torch.cuda.synchronize()
    L = _lib.load()
    _lib.check(L.gf_set_sph_precision(1))
    try:
        _lib.check(L.gf_hydro_dev(n, t["pos"].data_ptr(), t["vel"].data_ptr(),
                                  t["mass"].data_ptr(), t["u"].data_ptr(), eps, k, None,
                                  GAMMA, sph.ALPHA, sph.BETA,
                                  *[out[x].data_ptr() for x in ("h", "rho", "acc", "dudt",
                                                                "dt", "divv")], None))
        _lib.check(L.gf_sync(None))
    finally:
        _lib.check(L.gf_set_sph_precision(0))
    assert L.gf_get_sph_precision() == 0
    od, onb = O.knn(pos, k)
    h = O.smoothing_length(od[:, -1], eps)
    orho, oacc, odudt, odt, odivv = O.sph_forces(pos, vel, mass, onb, h, u, GAMMA, sph.ALPHA,
                                                 sph.BETA)
    np.testing.assert_array_equal(out["h"].cpu().numpy(), h)
    np.testing.assert_allclose(out["rho"].cpu().numpy(), orho, rtol=1e-6)
    for name, ref in (("acc", oacc), ("dudt", odudt), ("divv", odivv)):
        got = out[name].cpu().numpy()
        assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max(), name
    np.testing.assert_allclose(out["dt"].cpu().numpy(), odt, rtol=1e-5)


def test_special_cases():
    """Closed-form cases: u = v = 0 gives zero force; a receding pair has no
    viscosity; the interior of a uniform lattice has rho ~ rho_0."""
    n1 = 12
    ax = np.arange(n1, dtype=np.float64)
    pos = np.ascontiguousarray(np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3))
    pos += np.random.default_rng(0).normal(scale=1e-6, size=pos.shape)   # break distance ties
    n = len(pos)
    mass = np.ones(n)
    vel = np.zeros((n, 3))
    u = np.zeros(n)
    d, nb, _ = sph.knn(pos, 32)
    h = np.maximum(0.5 * d[:, -1], 0.01)
    rho, dudt, dt = np.empty(n), np.empty(n), np.empty(n)
    acc = np.empty((n, 3))
    sph.sph_forces(pos, vel, mass, nb, h, u, GAMMA, 1.0, 2.0, rho, acc, dudt, dt)
    assert np.abs(acc).max() == 0 and np.abs(dudt).max() == 0
    inner = np.all((pos > 3) & (pos < n1 - 4), axis=1)
    assert np.allclose(rho[inner], 1.0, rtol=0.1)
    # two particles moving apart: pressure only, no viscous heating term
    pos2 = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    vel2 = np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    nb2 = np.array([[0, 1], [1, 0]])
    h2 = np.array([1.0, 1.0])
    u2 = np.array([1.0, 1.0])
    r2, a2, du2, dt2 = np.empty(2), np.empty((2, 3)), np.empty(2), np.empty(2)
    sph.sph_forces(pos2, vel2, np.ones(2), nb2, h2, u2, GAMMA, 1.0, 2.0, r2, a2, du2, dt2)
    ci = np.sqrt(GAMMA * (GAMMA - 1))
    np.testing.assert_allclose(dt2, 0.3 / (ci + 1.2 * ci), rtol=1e-14)   # mu_max = 0


def _device_solve(pos, k, r_init=None):
    """The device-resident path (fast pass + fallbacks): unordered neighbour
    sets, d_k and the number of passes."""
    import torch

    from galaxy_formation_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    n = len(pos)
    tp = torch.from_numpy(np.ascontiguousarray(pos)).to(dev)
    tm = torch.ones(n, dtype=torch.float64, device=dev)
    r0 = None if r_init is None else torch.from_numpy(np.ascontiguousarray(r_init)).to(dev)
    torch.cuda.synchronize()
    _lib.check(L.gf_hydro_setup_dev(n, tp.data_ptr(), None, tm.data_ptr(), None, None))
    passes = _lib.check(L.gf_hydro_knn_dev(k, None if r0 is None else r0.data_ptr(), None))
    nb = np.empty((n, k), np.int64)
    dk = np.empty(n)
    _lib.check(L.gf_hydro_neighbours(nb, _lib.ptr(dk)))
    return nb, dk, passes


@pytest.mark.parametrize("blocks", [1, 37, 512])
def test_neighbours_independent_of_pass_blocks(blocks):
    """The fast pass is a persistent kernel whose blocks take groups from a
    shared cursor: any number of blocks gives the same exact neighbours."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    pos, *_ = disk_gas(20000, seed=17)
    od, onb = O.knn(pos, 64)
    _lib.check(L.gf_set_knn_blocks(blocks))
    try:
        assert L.gf_get_knn_blocks() == blocks
        nb, dk, _ = _device_solve(pos, 64, 1.05 * od[:, -1])
    finally:
        _lib.check(L.gf_set_knn_blocks(0))
    np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
    np.testing.assert_array_equal(dk, od[:, -1])


@pytest.mark.parametrize("members", [1, 16, 40, 64])
def test_neighbours_independent_of_merged_items(members):
    """The first pass from a caller's radii serves runs of sibling groups
    of at most `members` particles as one work item (1: every group alone):
    the neighbours and d_k are those of brute force whatever the items."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    pos, *_ = disk_gas(20000, seed=19)
    od, onb = O.knn(pos, 64)
    _lib.check(L.gf_set_knn_merge(members))
    try:
        assert L.gf_get_knn_merge() == members
        for first in (1.02, 1.3, 0.8):
            nb, dk, _ = _device_solve(pos, 64, first * od[:, -1])
            np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
            np.testing.assert_array_equal(dk, od[:, -1])
    finally:
        _lib.check(L.gf_set_knn_merge(0))


@pytest.mark.parametrize("k,scale", [(32, None), (64, 1.0), (64, 0.4), (64, 3.0), (96, None)])
def test_device_neighbour_sets_equal_bruteforce(k, scale):
    pos, *_ = disk_gas(20000, seed=11)
    od, onb = O.knn(pos, k)
    r0 = None if scale is None else scale * od[:, -1]
    nb, dk, passes = _device_solve(pos, k, r0)
    np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
    np.testing.assert_array_equal(dk, od[:, -1])
    if scale is not None and scale < 1:
        assert passes > 1


def test_device_neighbours_with_exact_ties():
    """A perfect lattice: many neighbours at exactly equal distances, broken
    by index as the brute-force order (d^2, index) does."""
    ax = np.arange(9, dtype=np.float64)
    pos = np.ascontiguousarray(np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3))
    for k in (32, 64):
        od, onb = O.knn(pos, k)
        nb, dk, _ = _device_solve(pos, k)
        np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
        np.testing.assert_array_equal(dk, od[:, -1])


@pytest.mark.parametrize("first", [None, 1.2])
@pytest.mark.parametrize("k,cap", [(32, 40), (64, 70), (64, 300)])
def test_neighbours_with_small_candidate_lists(k, cap, first):
    """Candidate lists cut to `cap` entries push groups into member mode,
    radius shrinking and the heap pass; the neighbour sets and d_k stay
    those of brute force.  Both shapes of the fast pass: the cell-density
    start (first None) and a caller's first radius (first x d_k)."""
    from galaxy_formation_amd import _lib
    L = _lib.load()
    pos, *_ = disk_gas(6000, seed=13)
    od, onb = O.knn(pos, k)
    _lib.check(L.gf_set_test_limits(0, cap, 0))
    try:
        nb, dk, passes = _device_solve(pos, k, None if first is None else first * od[:, -1])
        st = np.zeros(6, np.int64)
        _lib.check(L.gf_hydro_stats(st, 6))
    finally:
        _lib.check(L.gf_set_test_limits(0, 0, 0))
    np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
    np.testing.assert_array_equal(dk, od[:, -1])
    if cap < 100:
        assert st[1] > 0          # groups handed to the heap pass


@pytest.mark.parametrize("sizes", [(15000, 15000), (20000, 65536 + 4096 * 4 + 3 - 20000)])
@pytest.mark.parametrize("first", [None, 1.05])
def test_gas_tree_from_the_gravity_tree_equals_oracle(first, sizes):
    """The gas tree as the gravity tree's Morton order restricted to the gas
    (gf_hydro_setup_from_tree_dev: the gas is the first rows of the gravity
    tree's particles, no sort of its own): neighbours, d_k and h equal brute
    force, and density and forces the oracle's, from the cell-density start
    and from a caller's radii.  The second size puts the gravity tree's
    particle count just past a 4096-item tile of the compaction."""
    import torch

    from galaxy_formation_amd import _lib
    L = _lib.load()
    c = ics.make_disk(sizes[0], sizes[1], seed=23)
    g = c.gas
    assert (g == np.arange(len(g))).all()
    n, ng, k = len(c.pos), len(g), 64
    dev = torch.device("cuda:0")
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    pos, vel, mass, u = T(c.pos), T(c.vel), T(c.mass), T(c.u)
    gp = np.ascontiguousarray(c.pos[g])
    od, onb = O.knn(gp, k)
    r0 = None if first is None else T(first * od[:, -1])
    torch.cuda.synchronize()
    _lib.check(L.gf_tree_build_dev(n, pos.data_ptr(), mass.data_ptr(), 0.5, None))
    _lib.check(L.gf_hydro_setup_from_tree_dev(ng, pos.data_ptr(), vel.data_ptr(), mass.data_ptr(),
                                              u.data_ptr(), None))
    _lib.check(L.gf_hydro_knn_dev(k, None if r0 is None else r0.data_ptr(), None))
    nb = np.empty((ng, k), np.int64)
    dk = np.empty(ng)
    _lib.check(L.gf_hydro_neighbours(nb, _lib.ptr(dk)))
    np.testing.assert_array_equal(np.sort(nb, axis=1), np.sort(onb, axis=1))
    np.testing.assert_array_equal(dk, od[:, -1])
    out = {name: torch.zeros(shape, dtype=torch.float64, device=dev) for name, shape in
           (("h", ng), ("rho", ng), ("acc", (ng, 3)), ("dudt", ng), ("dt", ng), ("divv", ng))}
    _lib.check(L.gf_set_sph_precision(0))
    _lib.check(L.gf_hydro_forces_dev(c.eps, GAMMA, sph.ALPHA, sph.BETA,
                                     *[out[x].data_ptr() for x in ("h", "rho", "acc", "dudt",
                                                                   "dt", "divv")], None))
    _lib.check(L.gf_sync(None))
    h = O.smoothing_length(od[:, -1], c.eps)
    orho, oacc, _, _, _ = O.sph_forces(gp, np.ascontiguousarray(c.vel[g]),
                                       np.ascontiguousarray(c.mass[g]), onb, h,
                                       np.ascontiguousarray(c.u[g]), GAMMA, sph.ALPHA, sph.BETA)
    np.testing.assert_array_equal(out["h"].cpu().numpy(), h)
    np.testing.assert_allclose(out["rho"].cpu().numpy(), orho, rtol=1e-13)
    np.testing.assert_allclose(out["acc"].cpu().numpy(), oacc, rtol=1e-10,
                               atol=1e-12 * np.abs(oacc).max())


def test_gas_tree_from_the_gravity_tree_refuses_other_positions():
    """The set-up needs the positions of the last gravity build."""
    import torch

    from galaxy_formation_amd import _lib
    L = _lib.load()
    c = ics.make_cloud(2000, 2000, seed=5)
    dev = torch.device("cuda:0")
    pos = torch.from_numpy(np.ascontiguousarray(c.pos)).to(dev)
    other = pos.clone()
    mass = torch.from_numpy(np.ascontiguousarray(c.mass)).to(dev)
    _lib.check(L.gf_tree_build_dev(len(c.pos), pos.data_ptr(), mass.data_ptr(), 0.5, None))
    with pytest.raises(_lib.GFError):
        _lib.check(L.gf_hydro_setup_from_tree_dev(2000, other.data_ptr(), None, mass.data_ptr(),
                                                  None, None))
    with pytest.raises(_lib.GFError):
        _lib.check(L.gf_hydro_setup_from_tree_dev(len(c.pos) + 1, pos.data_ptr(), None,
                                                  mass.data_ptr(), None, None))
