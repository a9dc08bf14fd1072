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
    one later state solve the neighbours three ways: plain iteration, d_k of
    the previous step, learned proposal.  Neighbours, h and densities are
    identical; the learned start needs the fewest passes."""
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
    # one more step from the run's end, three ways
    F = S.F
    dt = float(2.0 * torch.quantile(F.dt_ind, 0.10).item())
    S.vel += 0.5 * dt * F.acc
    S.pos += dt * S.vel
    dk_prev, divv_prev, dtc_prev = F.dk.clone(), F.divv.clone(), F.dt_c.clone()
    out = {}
    for first in ("plain", "prev", "learned"):
        F.evaluate(S.pos, S.vel, S.mass, S.u, first=first, dk_prev=dk_prev, divv_prev=divv_prev,
                   dtc_prev=dtc_prev, dt=dt, model=model)
        tot = np.zeros(1, np.int64)
        from galaxy_formation_amd import _lib
        _lib.check(_lib.load().gf_hydro_passes(None, _lib.ptr(tot)))
        out[first] = dict(h=F.h.cpu().numpy(), rho=F.rho.cpu().numpy(),
                          mean_passes=tot[0] / S.ng, passes=F.passes)
    for first in ("prev", "learned"):
        np.testing.assert_array_equal(out[first]["h"], out["plain"]["h"])
        np.testing.assert_allclose(out[first]["rho"], out["plain"]["rho"], rtol=1e-13)
    assert out["learned"]["mean_passes"] < out["plain"]["mean_passes"]
    assert out["learned"]["mean_passes"] < out["prev"]["mean_passes"]
    assert out["learned"]["mean_passes"] < 1.1
