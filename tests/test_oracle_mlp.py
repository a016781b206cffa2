"""Pins of the surrogate network's CPU restatement (oracle/mlp.py).

The FP64 gradient is checked against central finite differences of the FP64
loss, and the bfloat16 rounding against torch's own float32 -> bfloat16
conversion and hand values."""
import numpy as np
import pytest

from oracle import mlp as OM


def rand_net(rng, dims):
    return [(rng.normal(0, np.sqrt(2.0 / k), (n, k)), rng.normal(0, 0.1, n))
            for k, n in zip(dims[:-1], dims[1:])]


def test_bf16_hand_values():
    x = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, -2.5, 0.0, 3.0e38, np.inf, 1.0 + 2.0**-7])
    y = OM.bf16(x)
    assert y[0] == 1.0
    assert y[1] == 1.0                      # halfway: ties to the even mantissa
    assert y[2] == 1.0 + 2.0**-6            # halfway: ties to the even mantissa
    assert y[3] == -2.5
    assert y[4] == 0.0
    assert np.isinf(y[6])
    assert y[7] == 1.0 + 2.0**-7            # representable


def test_bf16_matches_torch_conversion():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OM.bf16(x), ref)


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    dims = [5, 8, 8, 3]
    net = rand_net(rng, dims)
    X = rng.normal(size=(7, 5))
    T = rng.normal(size=(7, 3))
    loss, grads = OM.loss_and_grads(net, X, T)
    assert abs(loss - OM.loss_only(net, X, T)) <= 1e-15 * max(1.0, loss)
    h = 1e-6
    worst = 0.0
    for li, (W, b) in enumerate(net):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + h
                lp = OM.loss_only(net, X, T)
                arr[idx] = old - h
                lm = OM.loss_only(net, X, T)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                worst = max(worst, abs(fd - g[idx]) / (abs(fd) + 1e-6))
    assert worst <= 1e-6


def test_gradient_sign_errors_are_caught():
    """A planted mistake (SiLU derivative without its z-term) must fail the
    finite-difference comparison used above."""
    rng = np.random.default_rng(12)
    net = rand_net(rng, [4, 6, 2])
    X = rng.normal(size=(5, 4))
    T = rng.normal(size=(5, 2))
    loss, grads = OM.loss_and_grads(net, X, T)
    # planted: dz = da * s (missing the s*z*(1-s) term)
    Y, cache = OM.forward(net, X)
    dz = 2.0 * (Y - T) / Y.size
    a0, z0 = cache[0]
    s = 1.0 / (1.0 + np.exp(-z0))
    bad_dW0 = ((dz @ net[1][0]) * s).T @ a0
    h = 1e-6
    W = net[0][0]
    fd = np.zeros_like(W)
    for i in range(W.shape[0]):
        for j in range(W.shape[1]):
            o = W[i, j]
            W[i, j] = o + h
            lp = OM.loss_only(net, X, T)
            W[i, j] = o - h
            lm = OM.loss_only(net, X, T)
            W[i, j] = o
            fd[i, j] = (lp - lm) / (2 * h)
    assert np.max(np.abs(fd - grads[0][0])) <= 1e-8
    assert np.max(np.abs(fd - bad_dW0)) > 1e-4


def test_flatten_layout():
    rng = np.random.default_rng(1)
    net = rand_net(rng, [3, 4, 2])
    f = OM.flatten(net)
    assert f.size == 3 * 4 + 4 + 4 * 2 + 2
    assert np.array_equal(f[:12], net[0][0].ravel()) and np.array_equal(f[12:16], net[0][1])
