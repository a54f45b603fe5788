"""Pins the CPU oracle of the learned smoother's network (oracle/oracle_mlp.c):
gradient by central finite differences, forward by special weight sets whose
output is fixed by the normal CDF (scipy), an independent numpy forward over
the documented parameter layout, Adam's first steps in closed form, and the
bfloat16 round-to-nearest-even bit patterns."""
import numpy as np
import pytest
from scipy.special import ndtr

NI, NH, NO = 20, 64, 10
OW1, OB1 = 0, NH * NI
OW2, OB2 = OB1 + NH, OB1 + NH + NH * NH
OW3, OB3 = OB2 + NH, OB2 + NH + NO * NH
NP = OB3 + NO


def random_params(rng, scale=0.3):
    return rng.normal(scale=scale, size=NP)


def test_layout(orc):
    assert orc.mlp_n_params() == NP == 6154


def test_gradient_central_differences(orc):
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x = rng.uniform(-1, 1, (7, NI))
    t = rng.normal(size=(7, NO))
    loss, g = orc.mlp_loss_grad(p, x, t)
    assert loss == pytest.approx(orc.mlp_loss(p, x, t), rel=1e-14)
    # every bias and a random sample of weights of each layer
    idx = list(range(OB1, OB1 + NH)) + list(range(OB2, OB2 + NH)) + list(range(OB3, OB3 + NO))
    for lo, hi in ((OW1, OB1), (OW2, OB2), (OW3, OB3)):
        idx += list(rng.choice(np.arange(lo, hi), 40, replace=False))
    h = 1e-6
    fd = []
    for e in idx:
        pp, pm = p.copy(), p.copy()
        pp[e] += h
        pm[e] -= h
        fd.append((orc.mlp_loss(pp, x, t) - orc.mlp_loss(pm, x, t)) / (2 * h))
    fd = np.array(fd)
    assert np.max(np.abs(fd - g[idx])) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_gradient_planted_transpose_fails(orc):
    """A backward pass with W2 transposed (a plausible mistake) is caught by
    the finite-difference check: here the mistake is planted in the oracle's
    output by swapping the W2 gradient's indices."""
    rng = np.random.default_rng(2)
    p = random_params(rng)
    x = rng.uniform(-1, 1, (5, NI))
    t = rng.normal(size=(5, NO))
    _, g = orc.mlp_loss_grad(p, x, t)
    gw2 = g[OW2:OB2].reshape(NH, NH).T.ravel()
    e = OW2 + 3 * NH + 17
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[e] += h
    pm[e] -= h
    fd = (orc.mlp_loss(pp, x, t) - orc.mlp_loss(pm, x, t)) / (2 * h)
    tol = 1e-7 * max(1.0, abs(fd))
    assert abs(fd - g[e]) < tol
    assert abs(fd - gw2[e - OW2]) > 100 * tol


def test_forward_closed_form_gelu(orc):
    """W2 = 0: y = W3 gelu(b2) + b3 for every input, gelu(a) = a Phi(a)."""
    rng = np.random.default_rng(3)
    p = random_params(rng)
    p[OW2:OB2] = 0.0
    b2 = np.array([-2.0, -1.0, -0.5, 0.0, 0.5, 1.0, 2.0, 3.0] * 8)
    p[OB2:OW3] = b2
    x = rng.uniform(-1, 1, (3, NI))
    y = orc.mlp_forward(p, x)
    h2 = b2 * ndtr(b2)
    want = p[OW3:OB3].reshape(NO, NH) @ h2 + p[OB3:]
    for r in range(3):
        np.testing.assert_allclose(y[r], want, rtol=1e-14, atol=1e-14)
    assert 1.0 * ndtr(1.0) == pytest.approx(0.8413447460685429, rel=1e-15)


def test_forward_matches_numpy_layout(orc):
    """Independent numpy forward over the documented layout W[out][in]."""
    from scipy.special import erf
    rng = np.random.default_rng(4)
    p = random_params(rng)
    x = rng.uniform(-1, 1, (9, NI))
    W1, b1 = p[OW1:OB1].reshape(NH, NI), p[OB1:OW2]
    W2, b2 = p[OW2:OB2].reshape(NH, NH), p[OB2:OW3]
    W3, b3 = p[OW3:OB3].reshape(NO, NH), p[OB3:]
    g = lambda a: 0.5 * a * (1 + erf(a / np.sqrt(2)))
    want = g(g(x @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    np.testing.assert_allclose(orc.mlp_forward(p, x), want, rtol=1e-12, atol=1e-13)


def test_bf16_forward_close(orc):
    rng = np.random.default_rng(5)
    p = random_params(rng)
    x = rng.uniform(-1, 1, (50, NI))
    y, yb = orc.mlp_forward(p, x), orc.mlp_forward(p, x, bf16=True)
    assert np.max(np.abs(y - yb)) < 3e-2 * np.max(np.abs(y))
    assert np.max(np.abs(y - yb)) > 0


def test_adam_closed_form(orc):
    """Step 1 and 2 with a constant gradient g: m_hat = g, v_hat = g^2, so
    p moves by lr g / (|g| + eps) each step (Kingma & Ba, Algorithm 1)."""
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.array([1.0, 1.0, 1.0, 1.0])
    m, v = np.zeros(4), np.zeros(4)
    lr, eps = 1e-2, 1e-8
    orc.adam(p, g, m, v, lr, 1)
    step = lr * g / (np.abs(g) + eps)
    np.testing.assert_allclose(p, 1.0 - step, rtol=1e-14)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-14)
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=1e-14)
    orc.adam(p, g, m, v, lr, 2)
    np.testing.assert_allclose(p, 1.0 - 2 * step, rtol=1e-13)


@pytest.mark.parametrize("f,bits", [
    (1.0, 0x3F80), (-2.0, 0xC000), (0.0, 0x0000),
    (1.0 + 2.0 ** -8, 0x3F80),        # tie -> even (down)
    (1.0 + 3 * 2.0 ** -8, 0x3F82),    # tie -> even (up)
    (1.0 + 2.0 ** -8 + 2.0 ** -20, 0x3F81),  # above the tie -> up
    (3.140625, 0x4049), (float("inf"), 0x7F80), (-float("inf"), 0xFF80),
    (3.3895313892515355e38, 0x7F7F),  # largest bf16
])
def test_bf16_rounding(orc, f, bits):
    assert orc.bf16_bits(f) == bits


def test_bf16_nan_stays_nan(orc):
    b = orc.bf16_bits(float("nan"))
    assert (b & 0x7F80) == 0x7F80 and (b & 0x7F) != 0
