"""Pins of the network restatement oracle/learned_ref.py (CPU): its float64 gradient against
central finite differences, the bf16 rounding against the bit patterns, and the trainable mask."""
import numpy as np

from oracle import learned_ref as R


def _case(seed, shape=(4, 5, 6)):
    rng = np.random.default_rng(seed)
    m = R.mask()
    p = rng.normal(0, 0.3, 4 * R.LP) * m
    feat = np.zeros(shape + (R.CH,))
    feat[..., :8] = rng.normal(0, 1, shape + (8,))
    lab = np.stack([(rng.random(shape) < 0.3).astype(float), rng.uniform(0, 2, shape)], axis=-1)
    return p, feat, lab, m


def test_gradient_matches_finite_differences():
    p, feat, lab, m = _case(0)
    g, _ = R.gradient(feat, lab, p, pos_weight=3.0, lam=0.5, exact=True)
    assert np.all(g[m == 0] == 0)

    def loss(q):
        _, z = R.forward(feat, q, exact=True)
        (bce, size), _ = R.loss_grad(z, lab, 3.0, 0.5)
        return bce + 0.5 * size

    rng = np.random.default_rng(1)
    idx = rng.choice(np.nonzero(m)[0], 60, replace=False)
    idx = np.concatenate([idx, [l * R.LP + 27 * 256 + c for l in range(4) for c in (0, 1)]])  # biases
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_size_term_weight_is_lambda():
    """The size term enters the loss with weight lambda (the device's dm = 2 lambda y (m - t) / n)."""
    p, feat, lab, _ = _case(2)
    g1, (_, s1) = R.gradient(feat, lab, p, pos_weight=1.0, lam=0.0, exact=True)
    g2, (_, s2) = R.gradient(feat, lab, p, pos_weight=1.0, lam=2.0, exact=True)
    assert s1 == s2 and s1 > 0
    assert np.abs(g1 - g2).max() > 0


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-30, 0.0], np.float32)
    r = R.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie to even
    assert r[2] == 1.0 + 2 ** -6            # tie to even, upward
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-30) <= 3.0e-30 * 2 ** -8
    bits = (np.array([1.0], np.float32).view(np.uint32) >> 16).astype(np.uint16)
    assert R.bf16_bits(bits)[0] == 1.0
