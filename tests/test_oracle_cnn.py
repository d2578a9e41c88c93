"""Pins the numpy network oracle (oracle/cnn_oracle.py) before it judges the
GPU network: bf16 rounding on known values, convolution against a direct
loop and identity kernels, the backward operators as exact adjoints of the
forward ones, and the whole network's parameter gradients against central
finite differences of its own FP64 forward pass."""
import numpy as np
import pytest

from oracle import cnn_oracle as co


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 0.0,
                     1.0 + 2.0 ** -7 + 2.0 ** -8])
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0, -2.5, 0.0, 1.0 + 2.0 ** -6])
    np.testing.assert_array_equal(co.bf16(vals), want)


def _direct_conv(X, Wt, b, stride):
    N, H, W, Ci = X.shape
    Co = Wt.shape[0]
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    Y = np.zeros((N, Ho, Wo, Co))
    for n in range(N):
        for oy in range(Ho):
            for ox in range(Wo):
                for ky in range(3):
                    for kx in range(3):
                        iy, ix = oy * stride + ky - 1, ox * stride + kx - 1
                        if 0 <= iy < H and 0 <= ix < W:
                            Y[n, oy, ox] += Wt[:, ky, kx, :] @ X[n, iy, ix]
    return Y + b


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_direct_loop(stride):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 5, 7, 3))
    Wt = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(co.conv_fwd(X, Wt, b, stride, 0), _direct_conv(X, Wt, b, stride),
                               rtol=1e-12, atol=1e-12)


def test_identity_and_shift_kernels():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((1, 6, 5, 4))
    Wt = np.zeros((4, 3, 3, 4))
    Wt[:, 1, 1, :] = np.eye(4)
    np.testing.assert_array_equal(co.conv_fwd(X, Wt, None, 1, 0), X)
    np.testing.assert_array_equal(co.conv_fwd(X, Wt, None, 2, 0), X[:, ::2, ::2])
    # tap (0, 0) reads the pixel up-left: a shift with a zero first row/column
    Ws = np.zeros((4, 3, 3, 4))
    Ws[:, 0, 0, :] = np.eye(4)
    Y = co.conv_fwd(X, Ws, None, 1, 0)
    np.testing.assert_array_equal(Y[:, 1:, 1:], X[:, :-1, :-1])
    assert not Y[:, 0].any() and not Y[:, :, 0].any()


@pytest.mark.parametrize("stride,H,W", [(1, 6, 7), (2, 7, 6), (2, 8, 8)])
def test_backward_operators_are_adjoints(stride, H, W):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2, H, W, 3))
    Wt = rng.standard_normal((5, 3, 3, 3))
    dZ = rng.standard_normal(co.conv_fwd(X, Wt, None, stride, 0).shape)
    lhs = np.sum(co.conv_fwd(X, Wt, None, stride, 0) * dZ)
    np.testing.assert_allclose(np.sum(X * co.conv_bwd_data(dZ, Wt, H, W, stride)), lhs, rtol=1e-11)
    np.testing.assert_allclose(np.sum(Wt * co.conv_bwd_weight(X, dZ, stride)), lhs, rtol=1e-11)


def test_upcat_backward_is_adjoint():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((2, 3, 4, 5))
    S = rng.standard_normal((2, 5, 7, 2))   # odd sizes: the upsampled A is cropped
    dO = rng.standard_normal((2, 5, 7, 7))
    lhs = np.sum(co.upcat(A, S) * dO)
    dA, dS = co.upcat_bwd(dO, A.shape, 5)
    np.testing.assert_allclose(np.sum(A * dA) + np.sum(S * dS), lhs, rtol=1e-12)


def _tiny_params(rng, scale=1.0):
    ps = []
    for ci, cout, _, _ in co.LAYERS:
        ps.append(rng.standard_normal(cout * 9 * ci) * scale * np.sqrt(2.0 / (9 * ci)))
    for _, cout, _, _ in co.LAYERS:
        ps.append(rng.standard_normal(cout) * 0.1)
    return np.concatenate(ps)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    params = _tiny_params(rng)
    x = np.zeros((1, 6, 5, 64))
    x[..., :33] = rng.uniform(0, 1, (1, 6, 5, 33))
    lo, hi = 0.05, 1.0
    # targets away from every prediction, so the L1 term is smooth
    tgt = np.full((1, 6, 5), 0.3)
    tgt[0, 0, 0] = 0.0  # no target
    act = co.forward(params, x, round_bf16=False)
    u, _ = co.head_outputs(act[14], lo, hi)
    assert np.all(np.abs(u - 0.3) > 0.02)
    e0 = u - tgt   # the likelihood term's error is held at its base value

    def loss(p):
        a = co.forward(p, x, round_bf16=False)
        _, l1, nll = co.head_loss(a[14], tgt, lo, hi, 0.5, e_fixed=e0)
        return l1 + 0.5 * nll

    g, _, _ = co.backward(params, act, tgt, lo, hi, 0.5, round_bf16=False)
    offs = np.cumsum([0] + [c * 9 * i for i, c, _, _ in co.LAYERS])
    nw = offs[-1]
    boffs = nw + np.cumsum([0] + [c for _, c, _, _ in co.LAYERS])
    picks = [offs[k] + rng.integers(0, offs[k + 1] - offs[k]) for k in (0, 2, 6, 7, 8, 12, 14)]
    picks += [boffs[k] + rng.integers(0, boffs[k + 1] - boffs[k]) for k in (1, 9, 14)]
    for k in picks:
        e = 1e-5
        p1, p2 = params.copy(), params.copy()
        p1[k] += e
        p2[k] -= e
        fd = (loss(p1) - loss(p2)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])
