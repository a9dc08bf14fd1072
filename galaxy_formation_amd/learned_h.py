"""Learned first radius of the neighbour solve.

A 4 -> 32 -> 32 -> 1 SiLU MLP (FP32 weights, BF16 inputs) maps the features
of a gas particle, built from the previous step's d_k (distance to the k-th
neighbour; h = max(d_k, eps) / 2), the density of the smallest gas-tree
cell holding >= k particles, div v and dt, and made free of length and time
scales so that one network serves every resolution:

    x0 = log(d_k(prev) / r_cell),  r_cell = (3 k / (4 pi n_cell))^(1/3)
    x1 = log(cell count / k)
    x2 = div v(prev) * dt
    x3 = log(dt / dt_Courant(prev))

to a correction y of log d_k: the proposal is r = d_k(prev) exp(y + margin).
The plain radius iteration (gf_hydro_knn_dev) decides the result: a proposal only
sets where the iteration starts, so neighbour sets, d_k and densities are the
ones of the plain solve whatever the network says.

Forward, backward and Adam run in libgf.so's own kernels (gf_mlp_*); this
module holds the parameters and the training loop.  PyTorch provides device
memory and minibatch index shuffling.
"""

import numpy as np

from . import _lib

NPARAMS = 1249
H, NIN = 32, 4


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (layout W1, b1, W2, b2, W3, b3)."""
    rng = np.random.default_rng(seed)

    def glorot(fan_out, fan_in):
        a = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-a, a, size=(fan_out, fan_in))
    p = np.concatenate([glorot(H, NIN).ravel(), np.zeros(H), glorot(H, H).ravel(), np.zeros(H),
                        glorot(1, H).ravel() * 0.1, [0.0]])
    assert p.size == NPARAMS
    return p.astype(np.float32)


class LearnedH:
    def __init__(self, device="cuda:0", seed=0):
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.L = _lib.load()
        assert self.L.gf_mlp_params() == NPARAMS
        self.params = torch.from_numpy(init_params(seed)).to(self.dev)
        self.m = torch.zeros(NPARAMS, dtype=torch.float32, device=self.dev)
        self.v = torch.zeros_like(self.m)
        self.norm = torch.tensor([0, 0, 0, 0, 1, 1, 1, 1], dtype=torch.float32, device=self.dev)
        self.margin = 0.0
        self.margin_fixed = 0.0   # margin that d_k(prev) alone needs for the same coverage
        self.step_count = 0

    # -- features ------------------------------------------------------------

    def features(self, k, dk_prev, cell_rho, cell_count, divv_prev, dtc_prev, dt, raw=False,
                 stream=None):
        """BF16 features (n, 4) as int16 bit patterns; optionally the raw FP32
        features too."""
        torch = self.torch
        n = dk_prev.numel()
        x = torch.empty((n, NIN), dtype=torch.int16, device=self.dev)
        r = torch.empty((n, NIN), dtype=torch.float32, device=self.dev) if raw else None
        _lib.check(self.L.gf_mlp_features_dev(
            n, int(k), dk_prev.data_ptr(), cell_rho.data_ptr(), cell_count.data_ptr(),
            divv_prev.data_ptr(), dtc_prev.data_ptr(), float(dt), self.norm.data_ptr(),
            x.data_ptr(), None if r is None else r.data_ptr(), _stream(stream)))
        return (x, r) if raw else x

    def fit_norm(self, raw):
        """Feature means and standard deviations from raw (n, 4) features."""
        mu = raw.double().mean(0)
        sd = raw.double().std(0).clamp_min(1e-6)
        self.norm = self.torch.cat([mu, sd]).float().contiguous()

    # -- network ---------------------------------------------------------------

    def predict(self, x, stream=None):
        out = self.torch.empty(x.shape[0], dtype=self.torch.float32, device=self.dev)
        _lib.check(self.L.gf_mlp_forward_dev(x.shape[0], x.data_ptr(), self.params.data_ptr(),
                                             out.data_ptr(), _stream(stream)))
        return out

    def propose(self, x, dk_prev, margin=None, out=None, stream=None):
        """First search radius r = d_k(prev) exp(y + margin) (FP64, device)."""
        if out is None:
            out = self.torch.empty(x.shape[0], dtype=self.torch.float64, device=self.dev)
        m = self.margin if margin is None else margin
        _lib.check(self.L.gf_mlp_propose_dev(x.shape[0], x.data_ptr(), self.params.data_ptr(),
                                             dk_prev.data_ptr(), float(m), out.data_ptr(),
                                             _stream(stream)))
        return out

    def grad(self, x, target, tau=0.0, stream=None):
        g = self.torch.empty(NPARAMS, dtype=self.torch.float32, device=self.dev)
        loss = self.torch.empty(1, dtype=self.torch.float32, device=self.dev)
        _lib.check(self.L.gf_mlp_grad_dev(x.shape[0], x.data_ptr(), target.data_ptr(),
                                          self.params.data_ptr(), float(tau), g.data_ptr(),
                                          loss.data_ptr(), _stream(stream)))
        return g, loss

    def train(self, x, target, steps=2000, batch=8192, lr=2e-3, lr_end=1e-4, seed=0, log_every=0,
              tau=0.0):
        """Adam on the mean squared error (tau <= 0) or on the pinball loss
        of the tau-quantile, minibatches drawn without replacement each
        epoch, learning rate decayed geometrically."""
        torch = self.torch
        n = x.shape[0]
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        perm = torch.randperm(n, device=self.dev, generator=gen)
        pos = 0
        hist = []
        for s in range(steps):
            if pos + batch > n:
                perm = torch.randperm(n, device=self.dev, generator=gen)
                pos = 0
            idx = perm[pos:pos + batch]
            pos += batch
            xb = x.index_select(0, idx).contiguous()
            tb = target.index_select(0, idx).contiguous()
            self.step_count += 1
            lr_s = lr * (lr_end / lr) ** (s / max(1, steps - 1))
            _lib.check(self.L.gf_mlp_train_step_dev(xb.shape[0], xb.data_ptr(), tb.data_ptr(),
                                                    self.params.data_ptr(), self.m.data_ptr(),
                                                    self.v.data_ptr(), self.step_count, lr_s, 0.9,
                                                    0.999, 1e-8, float(tau), loss.data_ptr(),
                                                    _stream(None)))
            if log_every and (s % log_every == 0 or s == steps - 1):
                hist.append((s, float(loss.item()) / xb.shape[0]))
        return hist

    def calibrate_margin(self, x, target, quantile=0.97):
        """Margin such that a fraction `quantile` of the samples' true log d_k
        lies below log d_k(prev) + y + margin (their first pass resolves).
        After quantile training the margin is only a small correction."""
        res = (target - self.predict(x)).double()
        self.margin = float(self.torch.quantile(res[:min(len(res), 1 << 24)], quantile).item())
        return self.margin

    # -- persistence -----------------------------------------------------------

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy(),
                 margin=np.float64(self.margin), margin_fixed=np.float64(self.margin_fixed))

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            self.params = self.torch.from_numpy(f["params"].astype(np.float32)).to(self.dev)
            self.norm = self.torch.from_numpy(f["norm"].astype(np.float32)).to(self.dev)
            self.margin = float(f["margin"])
            self.margin_fixed = float(f["margin_fixed"]) if "margin_fixed" in f.files else 0.0
        return self


def _stream(stream):
    """Handle of `stream`, or of PyTorch's current stream (0 = the default)."""
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream or None
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
