"""The learned first-radius network (libgf.so gf_mlp_* kernels) against the
FP64 numpy oracle (oracle/mlp_oracle.py), which is itself pinned by central
finite differences; and what the trained network gains over the plain
radius iteration, with the converged neighbours unchanged."""

import numpy as np
import pytest

from oracle import mlp_oracle as M


def _sample(n=1000, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=M.NPARAMS)
    x = rng.normal(size=(n, 4))
    t = np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2] - 0.2 * x[:, 3]
    return p, x, t


# --------------------------------------------------------------- CPU: oracle

def test_oracle_gradient_matches_finite_differences():
    p, x, t = _sample(200)
    g = M.grad(p, x, t)
    idx = np.random.default_rng(1).choice(M.NPARAMS, 80, replace=False)
    idx = np.concatenate([idx, [0, 127, 128, 160, 1183, 1215, 1216, 1248]])
    fd = M.fd_grad(p, x, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-5, 1e30], np.float32)
    bits, back = M.bf16_round(vals)
    assert bits[0] == 0x3F80
    assert back[1] == 1.0                 # halfway between 1.0 and 1.0078125: to even
    assert back[2] == 1.0078125           # above halfway: up
    assert back[3] == -2.5
    rel = np.abs(back[4:] - vals[4:].astype(np.float64)) / np.abs(vals[4:])
    assert (rel < 2 ** -8).all()


def test_oracle_pinball_gradient_matches_finite_differences():
    """The training loss (tau = 0.999): gradient against central differences,
    on samples whose error stays clear of the kink at y = t."""
    p, x, t = _sample(400, seed=2)
    y, _ = M.forward(p, x)
    keep = np.abs(y - t) > 1e-3
    x, t = x[keep], t[keep]
    assert keep.sum() > 300
    for tau in (0.999, 0.5, 0.1):
        g = M.grad(p, x, t, tau)
        idx = np.random.default_rng(5).choice(M.NPARAMS, 60, replace=False)
        idx = np.concatenate([idx, [0, 127, 128, 160, 1183, 1215, 1216, 1248]])
        fd = M.fd_grad(p, x, t, idx, tau=tau)
        np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-10)


def test_oracle_pinball_loss_closed_forms():
    """Hand values of the pinball loss, and its defining property: for a
    constant prediction b the derivative of the mean loss is
    (1 - tau) P(t <= b) - tau P(t > b), which changes sign at the
    tau-quantile of the targets."""
    y = np.array([3.0, 1.0, 2.0])
    t = np.array([1.0, 3.0, 2.0])
    np.testing.assert_allclose(M.sample_loss(y, t, 0.9), [0.2, 1.8, 0.0], rtol=1e-15)
    np.testing.assert_allclose(M.dloss_dy(y, t, 0.9), [0.1, -0.9, 0.1], rtol=1e-15)
    np.testing.assert_allclose(M.sample_loss(y, t, 0.0), [2.0, 2.0, 0.0], rtol=1e-15)
    # a network whose output is b3 alone (every other parameter zero)
    targets = np.arange(1000, dtype=np.float64)             # quantile tau at ~1000 tau
    x = np.zeros((1000, 4))
    for tau in (0.9, 0.999):
        q = np.quantile(targets, tau)
        for b, sign in ((q - 5.0, -1.0), (q + 5.0, 1.0)):
            p = np.zeros(M.NPARAMS)
            p[-1] = b
            assert np.sign(M.grad(p, x, targets, tau)[-1]) == sign


def test_oracle_adam_three_steps_by_hand():
    """Three Adam steps (b1 0.9, b2 0.999, lr 0.1, eps 0) with gradients 1, -1,
    2, worked by hand: the bias corrections 1 - 0.9^s and 1 - 0.999^s of steps
    2 and 3 are 0.19, 0.271 and 0.001999, 0.002997001."""
    p, m, v = np.zeros(1), np.zeros(1), np.zeros(1)
    out = []
    for s, g in enumerate((1.0, -1.0, 2.0), start=1):
        p, m, v = M.adam(p, m, v, np.array([g]), s, 0.1, eps=0.0)
        out.append((p[0], m[0], v[0]))
    # step 1: m 0.1, v 0.001, corrected to 1 and 1
    np.testing.assert_allclose(out[0], (-0.1, 0.1, 0.001), rtol=1e-14)
    # step 2: m 0.09 - 0.1, v 0.000999 + 0.001, corrected to -1/19 and 1
    p2 = -0.1 + 0.1 / 19.0
    np.testing.assert_allclose(out[1], (p2, -0.01, 0.001999), rtol=1e-12)
    # step 3: m -0.009 + 0.2, v 0.001997001 + 0.004
    p3 = p2 - 0.1 * (0.191 / 0.271) / np.sqrt(0.005997001 / 0.002997001)
    np.testing.assert_allclose(out[2], (p3, 0.191, 0.005997001), rtol=1e-12)


def test_oracle_train_is_repeated_adam_on_the_gradient():
    """train() applies adam() to grad() at the current parameters each step,
    and reports the loss before each step."""
    p, x, t = _sample(100, seed=8)
    pt, _, _, sums = M.train(p, x, t, 3, 1e-2, tau=0.999)
    q, m, v = p.copy(), np.zeros(M.NPARAMS), np.zeros(M.NPARAMS)
    for s in range(1, 4):
        y, _ = M.forward(q, x)
        assert abs(sums[s - 1] - np.sum(M.sample_loss(y, t, 0.999))) < 1e-12
        q, m, v = M.adam(q, m, v, M.grad(q, x, t, 0.999), s, 1e-2)
    np.testing.assert_array_equal(pt, q)


def test_oracle_adam_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = M.adam(p, np.zeros(2), np.zeros(2), g, 1, lr=0.1)
    # the first Adam step moves every parameter by lr against its gradient's sign
    np.testing.assert_allclose(p1, p - 0.1 * np.sign(g), rtol=1e-7)


# -------------------------------------------------------------- GPU: kernels

@pytest.mark.gpu
def test_kernel_gradient_equals_oracle():
    from galaxy_formation_amd import _lib
    p, x, t = _sample(1000, seed=3)
    bits, xq = M.bf16_round(x.astype(np.float32))
    p32 = p.astype(np.float32)
    g = np.empty(M.NPARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    _lib.check(_lib.load().gf_mlp_grad_host(len(t), np.ascontiguousarray(bits.reshape(-1, 4)),
                                            t.astype(np.float32), p32, 0.0, g, _lib.ptr(loss)))
    go = M.grad(p32.astype(np.float64), xq.reshape(-1, 4), t.astype(np.float32).astype(np.float64))
    scale = np.abs(go).max()
    np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-5 * scale)
    y, _ = M.forward(p32.astype(np.float64), xq.reshape(-1, 4))
    sse = np.sum((y - t.astype(np.float32)) ** 2)
    assert abs(loss[0] - sse) < 1e-4 * sse


@pytest.mark.gpu
def test_kernel_pinball_gradient_equals_oracle():
    """The gradient training uses (pinball loss, tau = 0.999, the quantile of
    train_learned_h.py) and its summed loss."""
    from galaxy_formation_amd import _lib
    p, x, t = _sample(1000, seed=6)
    bits, xq = M.bf16_round(x.astype(np.float32))
    p32 = p.astype(np.float32)
    t32 = t.astype(np.float32)
    for tau in (0.999, 0.9):
        g = np.empty(M.NPARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        _lib.check(_lib.load().gf_mlp_grad_host(len(t), np.ascontiguousarray(bits.reshape(-1, 4)),
                                                t32, p32, tau, g, _lib.ptr(loss)))
        xo = xq.reshape(-1, 4)
        go = M.grad(p32.astype(np.float64), xo, t32.astype(np.float64), tau)
        np.testing.assert_allclose(g, go, rtol=2e-3, atol=2e-5 * np.abs(go).max())
        y, _ = M.forward(p32.astype(np.float64), xo)
        lo = np.sum(M.sample_loss(y, t32.astype(np.float64), tau))
        assert abs(loss[0] - lo) < 1e-4 * lo


@pytest.mark.gpu
def test_kernel_five_adam_steps_equal_oracle():
    """Five full-batch Adam steps on the pinball loss (tau = 0.999) by the
    kernels, against the oracle's five steps from the same parameters, each
    oracle step taking its own FP64 gradient."""
    import torch

    from galaxy_formation_amd.learned_h import LearnedH
    p, x, t = _sample(2000, seed=9)
    bits, xq = M.bf16_round(x.astype(np.float32))
    t32 = t.astype(np.float32)
    model = LearnedH()
    model.params = torch.from_numpy(p.astype(np.float32)).cuda()
    xd = torch.from_numpy(bits.reshape(-1, 4).view(np.int16)).cuda()
    td = torch.from_numpy(t32).cuda()
    lr = 1e-2
    model.step_count = 0
    model.train(xd, td, steps=5, batch=len(t), lr=lr, lr_end=lr, tau=0.999)
    p5 = model.params.cpu().numpy().astype(np.float64)
    po, mo, vo, _ = M.train(p.astype(np.float32).astype(np.float64), xq.reshape(-1, 4),
                            t32.astype(np.float64), 5, lr, tau=0.999)
    np.testing.assert_allclose(model.m.cpu().numpy(), mo, rtol=1e-3, atol=1e-3 * np.abs(mo).max())
    # every parameter moved by about 5 lr; the kernels' FP32 sums agree to
    # a small fraction of one step (a gradient perturbed by 1e-6 of its
    # largest entry moves the oracle's parameters by 0.012 lr; a wrong bias
    # correction moves them by about lr per step)
    assert np.abs(p5 - po).max() < 0.05 * lr
    assert np.abs(p5 - p).max() > 3 * lr


@pytest.mark.gpu
def test_kernel_forward_and_adam_equal_oracle():
    import torch

    from galaxy_formation_amd.learned_h import LearnedH
    p, x, t = _sample(3000, seed=4)
    bits, xq = M.bf16_round(x.astype(np.float32))
    model = LearnedH()
    model.params = torch.from_numpy(p.astype(np.float32)).cuda()
    xd = torch.from_numpy(bits.reshape(-1, 4).view(np.int16)).cuda()
    y = model.predict(xd).cpu().numpy()
    yo, _ = M.forward(p.astype(np.float32).astype(np.float64), xq.reshape(-1, 4))
    np.testing.assert_allclose(y, yo, rtol=1e-5, atol=1e-5)
    # one Adam step on the full batch
    td = torch.from_numpy(t.astype(np.float32)).cuda()
    g, _ = model.grad(xd, td)
    model.step_count = 0
    model.train(xd, td, steps=1, batch=len(t), lr=1e-2, lr_end=1e-2)
    p1 = model.params.cpu().numpy()
    po, _, _ = M.adam(p.astype(np.float32).astype(np.float64), np.zeros(M.NPARAMS),
                      np.zeros(M.NPARAMS), g.cpu().numpy().astype(np.float64), 1, 1e-2)
    np.testing.assert_allclose(p1, po, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_training_reduces_loss():
    import torch

    from galaxy_formation_amd.learned_h import LearnedH
    p, x, t = _sample(20000, seed=5)
    bits, _ = M.bf16_round(x.astype(np.float32))
    model = LearnedH(seed=1)
    xd = torch.from_numpy(bits.reshape(-1, 4).view(np.int16)).cuda()
    td = torch.from_numpy(t.astype(np.float32)).cuda()
    before = float(((model.predict(xd) - td) ** 2).mean())
    model.train(xd, td, steps=600, batch=2048, lr=3e-3, lr_end=3e-4)
    after = float(((model.predict(xd) - td) ** 2).mean())
    assert after < 0.1 * before


@pytest.mark.gpu
def test_learned_radius_gains_and_keeps_neighbours():
    """Train on snapshots of a cloud run (5,000 + 5,000 particles), then from
    one later state solve the neighbours four ways: plain iteration, d_k of
    the previous step, d_k of the previous step times the fixed margin that
    covers the same quantile of the training drifts (the method without the
    network), learned proposal.  Neighbours, h and densities are identical.
    The learned start needs fewer passes than the plain iteration and bare
    d_k(prev), and at the fixed margin's coverage it starts closer to the
    answer: its first radii overshoot d_k by less."""
    import torch

    from galaxy_formation_amd import ics
    from galaxy_formation_amd.learned_h import LearnedH
    from galaxy_formation_amd.stepper import Stepper, training_pairs
    model = LearnedH(seed=2)
    S = Stepper(ics.make_cloud(5000, 5000, seed=21), k=64)
    S.run(24, dt_factor=2.0)
    raw = training_pairs(S.records, model, 64, 3)[1]
    model.fit_norm(raw)
    x, _, t = training_pairs(S.records, model, 64, 3)
    model.train(x, t, steps=800, batch=4096, lr=3e-3, lr_end=2e-4, tau=0.97)
    model.calibrate_margin(x, t, 0.97)
    margin_fixed = float(torch.quantile(t.double(), 0.97).item())
    # one more step from the run's end, four ways
    F = S.F
    dt = float(2.0 * torch.quantile(F.dt_ind, 0.10).item())
    S.vel += 0.5 * dt * F.acc
    S.pos += dt * S.vel
    dk_prev, divv_prev, dtc_prev = F.dk.clone(), F.divv.clone(), F.dt_c.clone()
    out = {}
    from galaxy_formation_amd import _lib
    for name, first, margin in (("plain", "plain", 0.0), ("prev", "prev", 0.0),
                                ("fixed_margin", "prev", margin_fixed),
                                ("learned", "learned", 0.0)):
        F.evaluate(S.pos, S.vel, S.mass, S.u, first=first, dk_prev=dk_prev, divv_prev=divv_prev,
                   dtc_prev=dtc_prev, dt=dt, model=model, margin=margin)
        tot = np.zeros(1, np.int64)
        _lib.check(_lib.load().gf_hydro_passes(None, _lib.ptr(tot)))
        # first radius over the converged d_k (log), for the proposals
        over = float(torch.log(F.r0 / F.dk).mean().item()) if first != "plain" and (
            first == "learned" or margin) else None
        out[name] = dict(h=F.h.cpu().numpy(), rho=F.rho.cpu().numpy(),
                         mean_passes=tot[0] / S.ng, passes=F.passes, overshoot=over)
    print({k: (v["mean_passes"], v["overshoot"]) for k, v in out.items()})
    for name in ("prev", "fixed_margin", "learned"):
        np.testing.assert_array_equal(out[name]["h"], out["plain"]["h"])
        np.testing.assert_allclose(out[name]["rho"], out["plain"]["rho"], rtol=1e-13)
    assert out["learned"]["mean_passes"] < out["plain"]["mean_passes"]
    assert out["learned"]["mean_passes"] < out["prev"]["mean_passes"]
    assert out["learned"]["mean_passes"] < 1.1
    # the gain over the method without the network: the same coverage
    # (passes within 0.05) from first radii that overshoot less
    assert out["learned"]["mean_passes"] <= out["fixed_margin"]["mean_passes"] + 0.05
    assert out["learned"]["overshoot"] < out["fixed_margin"]["overshoot"]
