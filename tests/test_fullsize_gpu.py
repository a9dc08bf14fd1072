"""The fast path's accuracy gates at the full 1M-particle disk (500k gas),
through the C ABI on device arrays: gravity against the FP64 direct sum on a
seeded 2,000-particle subset, and the FP32-pair SPH sums against the FP64
ones on the same neighbour table."""

import numpy as np
import pytest

from galaxy_formation_amd import ics

pytestmark = pytest.mark.gpu

GAMMA, ALPHA, BETA = 5.0 / 3.0, 1.0, 2.0


@pytest.fixture(scope="module")
def disk():
    import torch

    from galaxy_formation_amd import _lib
    c = ics.config("disk1m")
    dev = torch.device("cuda:0")
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    return {"c": c, "L": _lib.load(), "lib": _lib, "torch": torch, "pos": T(c.pos),
            "vel": T(c.vel), "mass": T(c.mass), "u": T(c.u), "dev": dev}


def test_quadrupole_relative_rule_meets_the_gate_on_the_1m_disk(disk):
    """FP32 + quadrupoles with the relative rule (1e-4), its entry rows the
    FP64 gravity one step earlier (every particle drifted back by 5e-3):
    RMS relative error under 1e-3 against the direct sum, where the
    monopoles of the reference method are above it; the same interaction
    count as the monopole walk, and a few quadrupole entries per list."""
    torch, L, lib, c = disk["torch"], disk["L"], disk["lib"], disk["c"]
    n, eps = c.n, c.eps
    p = lambda t: t.data_ptr()  # noqa: E731
    z = lambda: torch.zeros((n, 3), dtype=torch.float64, device=disk["dev"])  # noqa: E731
    prev, acc_q, acc_m = z(), z(), z()
    back = (disk["pos"] - 5e-3 * disk["vel"]).contiguous()
    torch.cuda.synchronize()              # the library runs on the null stream
    lib.check(L.gf_tree_build_dev(n, p(back), p(disk["mass"]), 0.5, None))
    lib.check(L.gf_walk_dev(n, None, p(prev), eps, 1, None))
    lib.check(L.gf_tree_build_dev(n, p(disk["pos"]), p(disk["mass"]), 0.5, None))
    lib.check(L.gf_walk_dev(n, None, p(acc_m), eps, 0, None))
    lib.check(L.gf_sync(None))
    inter = np.zeros(1, np.int64)
    lib.check(L.gf_interactions_dev(lib.ptr(inter)))
    i_mono = int(inter[0])
    acc_q.copy_(prev)
    torch.cuda.synchronize()
    lib.check(L.gf_set_quad_relative_tolerance(1e-4))
    try:
        lib.check(L.gf_walk_dev(n, None, p(acc_q), eps, 2, None))
    finally:
        lib.check(L.gf_set_quad_relative_tolerance(0.0))
    lib.check(L.gf_sync(None))
    lib.check(L.gf_interactions_dev(lib.ptr(inter)))
    nodes = np.zeros(1, np.int64)
    lib.check(L.gf_node_interactions_dev(lib.ptr(nodes)))
    assert int(inter[0]) == i_mono
    assert 0 < int(nodes[0]) < 0.2 * i_mono
    sub = np.sort(np.random.default_rng(0).choice(n, 2000, replace=False))
    subt = torch.from_numpy(sub).to(disk["dev"])
    exact = torch.zeros((len(sub), 3), dtype=torch.float64, device=disk["dev"])
    torch.cuda.synchronize()
    lib.check(L.gf_direct_dev(n, p(disk["pos"]), p(disk["mass"]), len(sub), p(subt), p(exact),
                              eps, None))
    lib.check(L.gf_sync(None))
    ex = exact.cpu().numpy()

    def rms(a):
        e = np.linalg.norm(a.cpu().numpy()[sub] - ex, axis=1) / np.linalg.norm(ex, axis=1)
        return float(np.sqrt(np.mean(e ** 2)))
    r_q, r_m = rms(acc_q), rms(acc_m)
    assert r_q < 1e-3 < r_m, (r_q, r_m)


def test_fp32_pair_sph_matches_fp64_on_the_1m_disk_gas(disk):
    """The 500k gas of the 1M disk: the FP32-pair density and forces against
    the FP64 ones on the same solve (neighbours and h bit-equal by
    construction), density within 1e-6 relative (the gate is 1e-4), forces
    within 1e-4 of their largest magnitude."""
    torch, L, lib, c = disk["torch"], disk["L"], disk["lib"], disk["c"]
    ng = len(c.gas)
    assert (c.gas == np.arange(ng)).all()
    p = lambda t: t.data_ptr()  # noqa: E731
    mk = lambda *s: torch.zeros(s, dtype=torch.float64, device=disk["dev"])  # noqa: E731
    torch.cuda.synchronize()
    lib.check(L.gf_hydro_setup_dev(ng, p(disk["pos"]), p(disk["vel"]), p(disk["mass"]),
                                   p(disk["u"]), None))
    lib.check(L.gf_hydro_knn_dev(64, None, None))
    out = {}
    try:
        for prec in (0, 1):
            o = [mk(ng), mk(ng), mk(ng, 3), mk(ng), mk(ng), mk(ng)]
            torch.cuda.synchronize()
            lib.check(L.gf_set_sph_precision(prec))
            lib.check(L.gf_hydro_forces_dev(c.eps, GAMMA, ALPHA, BETA, *[p(x) for x in o], None))
            lib.check(L.gf_sync(None))
            out[prec] = [x.cpu().numpy() for x in o]
    finally:
        lib.check(L.gf_set_sph_precision(0))
    h64, rho64, a64 = out[0][0], out[0][1], out[0][2]
    h32, rho32, a32 = out[1][0], out[1][1], out[1][2]
    np.testing.assert_array_equal(h32, h64)
    assert np.max(np.abs(rho32 - rho64) / rho64) < 1e-6
    assert np.abs(a32 - a64).max() < 1e-4 * np.abs(a64).max()


def test_learned_radius_on_the_1m_disk_gas(disk):
    """The learned first radius on the 500k gas of the 1M disk (the workload
    of the learned-h configuration), with the committed network
    (data/learned_h_k64.npz).  The previous state is the gas one step
    earlier: drifted back by D, the 10th percentile of the admissible time
    steps (the long step of simulation.py:117).  From the cell-density start
    (the plain iteration), from d_k(prev) exp(fixed margin) and from the
    learned proposal: h and the neighbour sets are bit-equal to the plain
    solve and rho within 1e-4 of it (equal to rounding); the learned start
    takes fewer passes than the plain iteration, and overshoots the
    converged d_k by less than the fixed margin does."""
    import os

    from galaxy_formation_amd.learned_h import LearnedH
    from galaxy_formation_amd.stepper import ForceEvaluator
    torch, lib, c = disk["torch"], disk["lib"], disk["c"]
    L = disk["L"]
    ng = len(c.gas)
    model = LearnedH().load(os.path.join(os.path.dirname(ics.__file__), "data",
                                         "learned_h_k64.npz"))
    F = ForceEvaluator(c.n, ng, c.eps, theta=0.5, k=64, precision="fp64", sph_precision=0)
    pos, vel, mass, u = disk["pos"], disk["vel"], disk["mass"], disk["u"]
    F.evaluate(pos, vel, mass, u)
    D = float(torch.quantile(F.dt_ind, 0.10).item())
    back = (pos[:ng] - D * vel[:ng]).contiguous()
    F.hydro(back, vel[:ng], mass[:ng], u[:ng])
    dk_prev, divv_prev, dtc_prev = F.dk.clone(), F.divv.clone(), F.dt_c.clone()
    out = {}
    for name, first, margin in (("plain", "plain", 0.0),
                                ("fixed_margin", "prev", model.margin_fixed),
                                ("learned", "learned", 0.0)):
        print(f"disk1m gas neighbours: {name}", flush=True)
        F.hydro(pos[:ng], vel[:ng], mass[:ng], u[:ng], first=first, dk_prev=dk_prev,
                divv_prev=divv_prev, dtc_prev=dtc_prev, dt=D, model=model, margin=margin)
        torch.cuda.synchronize()
        nb = np.empty((ng, 64), np.int64)
        lib.check(L.gf_hydro_neighbours(nb, None))
        tot = np.zeros(1, np.int64)
        lib.check(L.gf_hydro_passes(None, lib.ptr(tot)))
        over = None if first == "plain" else float(torch.log(F.r0 / F.dk).mean().item())
        out[name] = dict(h=F.h.cpu().numpy(), rho=F.rho.cpu().numpy(), nb=np.sort(nb, axis=1),
                         mean_passes=tot[0] / ng, launches=F.passes, overshoot=over)
        print({k: v for k, v in out[name].items() if k not in ("h", "rho", "nb")}, flush=True)
    for name in ("fixed_margin", "learned"):
        np.testing.assert_array_equal(out[name]["h"], out["plain"]["h"])
        np.testing.assert_array_equal(out[name]["nb"], out["plain"]["nb"])
        assert np.max(np.abs(out[name]["rho"] - out["plain"]["rho"]) / out["plain"]["rho"]) < 1e-4
    assert out["learned"]["mean_passes"] < out["plain"]["mean_passes"]
    assert out["learned"]["overshoot"] < out["fixed_margin"]["overshoot"]
