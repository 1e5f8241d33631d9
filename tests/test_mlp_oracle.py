"""Pins of the FP64 radiance-network oracle (the checker of the GPU training
kernels): its gradient against central finite differences of its own loss,
and its positional encoding and layer wiring against closed forms."""
import numpy as np

from oracle import hvx_oracle as O

W1, B1, W2, B2, W3, B3, W4, B4 = 0, 8192, 8256, 12352, 12416, 16512, 16576, 16768


def _inputs(rng, n, lod_max):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([d, rng.uniform(0, lod_max, (n, 1))], axis=1).astype(np.float32)


def test_parameter_layout():
    assert B4 + 3 == O.MLP_PARAMS


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    P = (rng.normal(size=O.MLP_PARAMS) * 0.2).astype(np.float32)
    x = _inputs(rng, 6, 2.0)
    t = np.concatenate([rng.normal(size=(6, 3)), np.zeros((6, 1))], axis=1).astype(np.float32)
    g, _ = O.mlp_gradient(P, x, t, 2.0)
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in
                          [(W1, B1), (B1, W2), (W2, B2), (B2, W3), (W3, B3), (B3, W4), (W4, B4), (B4, B4 + 3)]])
    for i in idx:
        h = np.float32(1e-4)
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp = O.mlp_gradient(Pp, x, t, 2.0)[1]
        lm = O.mlp_gradient(Pm, x, t, 2.0)[1]
        fd = (lp - lm) / (float(Pp[i]) - float(Pm[i]))
        assert abs(fd - g[i]) <= 1e-6 + 1e-3 * abs(g[i]), (i, fd, g[i])


def test_output_bias_only():
    P = np.zeros(O.MLP_PARAMS, np.float32)
    P[B4:B4 + 3] = [1.0, -2.0, 0.5]
    y = O.mlp_forward(P, _inputs(np.random.default_rng(1), 5, 1.0), 1.0)
    assert np.array_equal(y, np.tile([1.0, -2.0, 0.5], (5, 1)))


def test_encoding_feature_wiring():
    """A network that routes one encoding feature to output 0 through unit
    weights gives 1 + sin(pi 2^(k/2) x_j) (or cos), x_3 = 2 lod/lod_max - 1."""
    rng = np.random.default_rng(2)
    x = _inputs(rng, 16, 3.0)
    x64 = x.astype(np.float64)
    xin = np.concatenate([x64[:, :3], 2.0 * x64[:, 3:] / 3.0 - 1.0], axis=1)
    for j, k, s in [(0, 0, 0), (1, 5, 1), (2, 15, 0), (3, 7, 1), (3, 0, 0)]:
        P = np.zeros(O.MLP_PARAMS, np.float32)
        P[W1 + j * 32 + 2 * k + s] = 1.0   # W1[0][feature]
        P[B1] = 1.0                        # h1_0 = 1 + feature >= 0
        P[W2] = 1.0
        P[W3] = 1.0
        P[W4] = 1.0
        y = O.mlp_forward(P, x, 3.0)[:, 0]
        f = np.pi * 2.0 ** (k / 2.0)
        ref = 1.0 + (np.cos if s else np.sin)(f * xin[:, j])
        assert np.allclose(y, ref, rtol=0, atol=1e-9), (j, k, s)


def test_fp32_encoding_against_sin_cos():
    """The kernels' FP32 encoding against sin/cos of f_k x in float64, f_k =
    pi 2^(k/2): within 3e-7 absolute for even k (exact argument; k = 2, 3
    mod 4 by one doubling) and 4e-5 for odd k (the rounded argument, up to 90.5
    turns at k = 15) over [-1, 1], exact at x = 0 and at the even-k zeros."""
    x = np.concatenate([np.linspace(-1, 1, 40001), [-1.0, 1.0, 0.5, -0.25, 1e-30, -1e-30]]).astype(np.float32)
    sc = O.mlp_sincos(x)
    k = np.arange(16)
    f = np.pi * 2.0 ** (k // 2) * np.where(k % 2 == 1, np.sqrt(2.0), 1.0)
    t = x.astype(np.float64)[:, None] * f[None, :]
    err = np.maximum(np.abs(sc[:, :, 0] - np.sin(t)), np.abs(sc[:, :, 1] - np.cos(t))).max(axis=0)
    assert err[0::2].max() <= 3e-7 and err[1::2].max() <= 4e-5, err
    # |sin|, |cos| <= 1 and sin^2 + cos^2 = 1 to FP32 rounding
    assert np.abs(sc).max() <= 1.0
    assert np.abs((sc.astype(np.float64) ** 2).sum(axis=2) - 1.0).max() <= 1e-6
    z = O.mlp_sincos(np.zeros(1, np.float32))[0]
    assert np.all(z[:, 0] == 0.0) and np.all(z[:, 1] == 1.0)
    # x = +-1 at even k >= 2: f_k x is a whole number of turns
    for xv in (1.0, -1.0):
        e = O.mlp_sincos(np.array([xv], np.float32))[0]
        for kk in range(2, 16, 2):
            assert abs(e[kk, 0]) == 0.0 and e[kk, 1] == 1.0
        assert abs(e[0, 0]) == 0.0 and e[0, 1] == -1.0  # f_0 = pi
