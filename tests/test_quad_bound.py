"""The bound behind the walk's relative quadrupole rule (no GPU needed): the
quadrupole term of an accepted node's field, a_q = 5/2 (d.Q.d) d / r^7 -
Q.d / r^5 for its traceless quadrupole Q, has |a_q| <= 2.858 |Q|_F / r^4.
It follows from |a_q| <= 3.5 |Q|_2 / r^4 and |Q|_2 <= sqrt(2/3) |Q|_F for a
traceless symmetric Q; the walk stores (2.858 |Q|_F)^(1/4) per node."""

import numpy as np


def quad_term(Q, d):
    r2 = np.sum(d * d, axis=-1)
    Qd = np.einsum("...ij,...j->...i", Q, d)
    dQd = np.sum(d * Qd, axis=-1)
    return (2.5 * dQd / r2 ** 3.5)[..., None] * d - Qd / r2[..., None] ** 2.5


def random_traceless(rng, n):
    A = rng.normal(size=(n, 3, 3))
    S = 0.5 * (A + A.transpose(0, 2, 1))
    return S - np.trace(S, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3


def test_norm_inequality_is_tight_for_an_axisymmetric_tensor():
    Q = np.diag([1.0, -0.5, -0.5])
    assert np.isclose(np.linalg.norm(Q, 2), np.sqrt(2.0 / 3.0) * np.linalg.norm(Q))
    assert np.isclose(3.5 * np.sqrt(2.0 / 3.0), 2.858, rtol=1e-3)
    assert 3.5 * np.sqrt(2.0 / 3.0) <= 2.858 * (1 + 1e-4)


def test_quadrupole_term_is_below_the_bound():
    rng = np.random.default_rng(5)
    Q = random_traceless(rng, 4000)
    d = rng.normal(size=(4000, 3)) * rng.uniform(0.1, 10.0, size=(4000, 1))
    r = np.linalg.norm(d, axis=1)
    aq = np.linalg.norm(quad_term(Q, d), axis=1)
    bound = 2.858 * np.linalg.norm(Q, axis=(1, 2)) / r ** 4
    ratio = aq / bound
    assert ratio.max() <= 1.0
    # the bound is within a small factor of the worst case, so a constant
    # taken too large (a weaker rule) shows here
    assert ratio.max() > 0.3
    # along an eigenvector of Q = diag(1, -1/2, -1/2) the term is 1.5 / r^4
    Qa = np.diag([1.0, -0.5, -0.5])[None]
    x = np.array([[2.0, 0.0, 0.0]])
    assert np.isclose(np.linalg.norm(quad_term(Qa, x)), 1.5 / 2.0 ** 4)
