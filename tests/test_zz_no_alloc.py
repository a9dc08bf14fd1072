"""gf_mlp_train_step_noalloc_dev (the Adam step whose partial gradients sit in
the library's static device array) against gf_mlp_train_step_dev (the same
kernels with a buffer the library allocates): bit-identical parameters,
moments and losses over two calls in a row, the second with fewer samples,
so that a stale partial left by the first call would show."""

import numpy as np
import pytest

from oracle import mlp_oracle as M

MAX_SAMPLES = 65536


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 4))
    t = np.sin(x[:, 0]) + 0.3 * x[:, 1] * x[:, 2] - 0.2 * x[:, 3]
    bits, _ = M.bf16_round(x.astype(np.float32))
    return bits.reshape(-1, 4).view(np.int16), t.astype(np.float32)


@pytest.mark.gpu
def test_noalloc_train_step_equals_allocating_step():
    import torch

    from galaxy_formation_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(7)
    p0 = torch.from_numpy(rng.normal(scale=0.3, size=M.NPARAMS).astype(np.float32)).to(dev)
    state = {name: [p0.clone(), torch.zeros_like(p0), torch.zeros_like(p0)]
             for name in ("alloc", "noalloc")}
    fn = {"alloc": L.gf_mlp_train_step_dev, "noalloc": L.gf_mlp_train_step_noalloc_dev}
    # (samples, step, tau): the largest batch the scratch holds, then a
    # smaller one on the pinball branch
    for n, step, tau in ((MAX_SAMPLES, 1, 0.0), (3001, 2, 0.999)):
        xb, tb = _batch(n, seed=n)
        x = torch.from_numpy(xb).to(dev)
        t = torch.from_numpy(tb).to(dev)
        loss = {}
        for name in ("alloc", "noalloc"):
            p, m, v = state[name]
            out = torch.full((1,), float("nan"), dtype=torch.float32, device=dev)
            print(f"call {name} n={n} step={step} tau={tau}", flush=True)
            _lib.check(fn[name](n, x.data_ptr(), t.data_ptr(), p.data_ptr(), m.data_ptr(),
                                v.data_ptr(), step, 2e-3, 0.9, 0.999, 1e-8, tau, out.data_ptr(),
                                None))
            torch.cuda.synchronize()
            loss[name] = out.cpu().numpy()
        for a, b in zip(state["alloc"], state["noalloc"]):
            assert torch.equal(a, b)
        assert np.isfinite(loss["alloc"]).all()
        assert np.array_equal(loss["alloc"], loss["noalloc"])
        assert not torch.equal(state["alloc"][0], p0)


@pytest.mark.gpu
def test_noalloc_train_step_refuses_more_samples_than_its_scratch():
    from galaxy_formation_amd import _lib
    L = _lib.load()
    rc = L.gf_mlp_train_step_noalloc_dev(MAX_SAMPLES + 1, None, None, None, None, None, 1, 1e-3,
                                         0.9, 0.999, 1e-8, 0.0, None, None)
    assert rc < 0
    assert b"exceed" in L.gf_last_error()
