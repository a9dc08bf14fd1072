"""FP64 numpy oracle of the learned smoothing-length MLP (4 -> 32 -> 32 -> 1,
SiLU).  TEST INFRASTRUCTURE ONLY: used by tests/ to check the GPU forward,
backward and Adam kernels; the product never imports it.

Parameter layout (1249 values): W1[32][4], b1[32], W2[32][32], b2[32],
W3[32], b3.  The model has no counterpart in the reference.  Its gradients
(squared error and pinball loss) are pinned by central finite differences,
and Adam by a three-step sequence worked by hand (tests/test_mlp.py).
"""

import numpy as np

H, NIN = 32, 4
NPARAMS = H * NIN + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * NIN].reshape(H, NIN); o += H * NIN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def silu(z):
    return z / (1.0 + np.exp(-z))


def forward(p, x):
    """x (n, 4) -> y (n,), with the intermediates."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = x @ W1.T + b1
    a1 = silu(z1)
    z2 = a1 @ W2.T + b2
    a2 = silu(z2)
    y = a2 @ W3 + b3
    return y, (z1, a1, z2, a2)


def sample_loss(y, t, tau=0.0):
    """Per-sample loss: tau <= 0 the squared error 0.5 (y - t)^2; 0 < tau < 1
    the pinball loss of the tau-quantile, max(tau (t - y), (1 - tau) (y - t))."""
    e = y - t
    if tau > 0.0:
        return np.where(e >= 0.0, (1.0 - tau) * e, -tau * e)
    return 0.5 * e * e


def loss(p, x, t, tau=0.0):
    """Mean of sample_loss over the n samples."""
    y, _ = forward(p, x)
    return np.sum(sample_loss(y, t, tau)) / len(t)


def dloss_dy(y, t, tau=0.0):
    """d sample_loss / dy; at y == t the pinball loss takes the right-hand
    derivative 1 - tau (the subgradient the kernels use)."""
    e = y - t
    if tau > 0.0:
        return np.where(e >= 0.0, 1.0 - tau, -tau)
    return e


def grad(p, x, t, tau=0.0):
    """Analytic gradient of loss()."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (z1, a1, z2, a2) = forward(p, x)
    n = len(t)
    dy = dloss_dy(y, t, tau) / n
    s2 = 1 / (1 + np.exp(-z2))
    dz2 = (dy[:, None] * W3[None, :]) * s2 * (1 + z2 * (1 - s2))
    s1 = 1 / (1 + np.exp(-z1))
    dz1 = (dz2 @ W2) * s1 * (1 + z1 * (1 - s1))
    g = np.concatenate([(dz1.T @ x).ravel(), dz1.sum(0), (dz2.T @ a1).ravel(), dz2.sum(0),
                        a2.T @ dy, [dy.sum()]])
    return g


def fd_grad(p, x, t, idx, h=1e-6, tau=0.0):
    """Central finite differences of loss() for the parameters idx."""
    out = []
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        out.append((loss(pp, x, t, tau) - loss(pm, x, t, tau)) / (2 * h))
    return np.array(out)


def bf16_round(x):
    """Round float32 values to BF16 (nearest even); returns (uint16 bits,
    the rounded values as float64)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    back = (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return r, back


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def train(p, x, t, steps, lr, tau=0.0, b1=0.9, b2=0.999, eps=1e-8):
    """`steps` full-batch Adam steps on loss(); returns the parameters and
    the summed loss before each step (squared errors for tau <= 0, as the
    kernels report it, else pinball losses)."""
    p = np.asarray(p, np.float64).copy()
    m, v = np.zeros_like(p), np.zeros_like(p)
    sums = []
    for s in range(1, steps + 1):
        y, _ = forward(p, x)
        e = y - t
        sums.append(float(np.sum(e * e) if tau <= 0.0 else np.sum(sample_loss(y, t, tau))))
        p, m, v = adam(p, m, v, grad(p, x, t, tau), s, lr, b1, b2, eps)
    return p, m, v, sums
