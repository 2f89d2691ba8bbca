"""Pins of the MLP gradient oracle: analytic FP64 gradient against central
finite differences of the loss, and the BF16 rounding helper against
closed-form cases."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_closed_form():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 3.0e-39], np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    np.testing.assert_array_equal(got[:4], [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6])
    assert got[4] == -2.5


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d = 20
    L = M.layout(d)
    p = np.zeros(L["total"], np.float64)
    for name, scale in (("w1", 0.3), ("b1", 0.1), ("w2", 0.1), ("b2", 0.1), ("w3", 0.2), ("b3", 0.1)):
        o, s = L[name]
        p[o:o + s] = rng.standard_normal(s) * scale
    # padding entries (rows of W1 past d, columns of W3/b3 past 12) are zero
    o, s = L["w1"]
    w1 = p[o:o + s].reshape(L["kinp"], M.H)
    w1[d:] = 0
    o, s = L["w3"]
    p[o:o + s].reshape(M.H, M.OUTP)[:, M.CODE:] = 0
    o, s = L["b3"]
    p[o + M.CODE:o + s] = 0
    X = rng.standard_normal((50, d)).astype(np.float32)
    B = 8
    rows = rng.integers(0, 50, 3 * B)
    loss, g = M.loss_grad(p, X, rows, B, d, use_bf16=False)
    idx = []
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        o, s = L[name]
        idx += list(rng.choice(np.arange(o, o + s), 6, replace=False))
    eps = 1e-6
    for i in idx:
        if g[i] == 0.0 and abs(p[i]) == 0.0:
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad(pp, X, rows, B, d, False)[0] - M.loss_grad(pm, X, rows, B, d, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_bits_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    d = 12
    L = M.layout(d)
    p = np.zeros(L["total"], np.float64)
    for name, scale in (("w1", 0.3), ("b1", 0.1), ("w2", 0.1), ("b2", 0.1), ("w3", 0.2), ("b3", 0.1)):
        o, s = L[name]
        p[o:o + s] = rng.standard_normal(s) * scale
    o, s = L["w1"]
    p[o:o + s].reshape(L["kinp"], M.H)[d:] = 0
    o, s = L["w3"]
    p[o:o + s].reshape(M.H, M.OUTP)[:, M.CODE:] = 0
    o, s = L["b3"]
    p[o + M.CODE:o + s] = 0
    X = rng.standard_normal((40, d)).astype(np.float32)
    codes = rng.integers(0, 4096, 40)
    B = 6
    rows = rng.integers(0, 40, 3 * B)
    loss, g = M.loss_grad_bits(p, X, rows, B, codes, d, use_bf16=False)
    # closed form at the loss level: t = tanh(z) recomputed independently
    t = np.tanh(M.forward(p, X[rows], d, use_bf16=False)[4])
    sgn = np.where((codes[rows][:, None] >> np.arange(12)) & 1, 1.0, -1.0)
    ref = ((t - sgn) ** 2).mean() + ((t[:B] - t[B:2 * B]) ** 2).mean()
    assert abs(loss - ref) < 1e-12
    idx = []
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        o, s = L[name]
        idx += list(rng.choice(np.arange(o, o + s), 6, replace=False))
    eps = 1e-6
    for i in idx:
        if g[i] == 0.0 and abs(p[i]) == 0.0:
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_grad_bits(pp, X, rows, B, codes, d, False)[0] -
              M.loss_grad_bits(pm, X, rows, B, codes, d, False)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_gelu_kernel_constants_match_erf():
    """The kernels' GELU evaluates the normal tail with the Abramowitz &
    Stegun 7.1.26 erfc form; its constants, read from the HIP source and
    evaluated in FP64, must give GELU and its derivative within the
    published 1.5e-7 bound of the exact-erf forms the oracle uses."""
    import math
    import pathlib
    import re
    src = (pathlib.Path(__file__).resolve().parents[1] / "gpu_hdbscan_amd" / "csrc" / "mlp.hip").read_text()
    body = src[src.index("NormTail norm_tail(float x) {"):]
    body = body[:body.index("return r;")]
    c = [float(v) for v in re.findall(r"(-?\d+\.\d+)f", body)]
    p, one, a5, a4, a3, a2, a1, ex = c
    assert one == 1.0
    x = np.linspace(-12, 12, 200001)
    t = 1.0 / (1.0 + p * np.abs(x))
    poly = (((a5 * t + a4) * t + a3) * t + a2) * t + a1
    e = np.exp2(x * x * ex)
    q = t * poly * e
    gelu = np.where(x >= 0, x - x * q, x * q)
    cdf = np.where(x >= 0, 1 - q, q)
    grad = cdf + x * 0.39894228040143268 * e
    assert np.max(np.abs(gelu - M.gelu(x)) / np.maximum(np.abs(x), 1)) < 1.5e-7
    assert np.max(np.abs(grad - M.gelu_grad(x)) / np.maximum(np.abs(x), 1)) < 1.5e-7
    assert abs(ex + 0.5 / math.log(2)) < 1e-12
