"""BF16 MFMA network kernels (include/winds_mlp.h) against oracle/mlp.py.

Tolerances: against the oracle with the device's bfloat16 roundings
emulated, the forward output agrees to 1e-3 and each layer's gradient to
1e-2 normwise (what is left is FP32 accumulation order and the rare bfloat16
rounding it flips); against the plain FP64 network the gradients agree to
5e-2 normwise (bfloat16 weights and activations)."""
import numpy as np
import pytest

from conftest import flush_print
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def make(n_in=40, hidden=64, layers=2, n_out=3, max_batch=256, seed=5):
    from winds.mlp import MLP
    return MLP(n_in, hidden, layers, n_out, max_batch, seed)


def data(B, n_in, n_out, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, n_in)).astype(np.float32)
    T = (0.3 * rng.normal(size=(B, n_out))).astype(np.float32)
    return X, T


def normwise(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("B", [200, 64, 1])
def test_forward_matches_oracle(need_gpu, B):
    import torch
    net = make()
    flat = net.get_params()
    params = [(W.astype(np.float64), b.astype(np.float64)) for W, b in net.unflatten(flat)]
    X, _ = data(B, 40, 3, 1)
    Y = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    Yo, _ = OM.forward(params, X, bf=True)
    e = normwise(Y, Yo)
    flush_print(f"[mlp forward B={B}] normwise {e:.2e}")
    assert e <= 1e-3


def test_gradients_match_oracle(need_gpu):
    import torch
    net = make()
    # move the output layer off its small initial scale so every layer's
    # gradient is well conditioned
    flat = net.get_params()
    rng = np.random.default_rng(2)
    pairs = net.unflatten(flat)
    pairs[-1][0][:] = rng.normal(0, 0.2, pairs[-1][0].shape)
    pairs[-1][1][:] = rng.normal(0, 0.1, pairs[-1][1].shape)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)
    params = [(W.astype(np.float64), b.astype(np.float64)) for W, b in net.unflatten(flat)]
    X, T = data(200, 40, 3, 3)
    g, loss = net.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lo, go = OM.loss_and_grads(params, X, T, bf=True)
    l64, g64 = OM.loss_and_grads(params, X, T, bf=False)
    assert abs(loss - lo) <= 1e-3 * lo
    gd = net.unflatten(g)
    for i, ((dW, db), (oW, ob), (fW, fb)) in enumerate(zip(gd, go, g64)):
        eW, eb = normwise(dW, oW), normwise(db, ob)
        fW_e = normwise(dW, fW)
        flush_print(f"[mlp grad layer {i}] W {eW:.2e} b {eb:.2e} vs fp64 {fW_e:.2e}")
        assert eW <= 1e-2 and eb <= 1e-2
        assert fW_e <= 5e-2


def test_first_adam_step_moves_by_lr_times_sign(need_gpu):
    import torch
    net = make(seed=9)
    p0 = net.get_params()
    X, T = data(128, 40, 3, 4)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    g, _ = net.gradients(Xd, Td)
    lr = 1e-3
    net.train_step(Xd, Td, lr)
    torch.cuda.synchronize()
    p1 = net.get_params()
    big = np.abs(g) > 1e-4 * np.max(np.abs(g))
    step = (p1 - p0)[big]
    # Adam's first step is lr * g / (|g| + eps') = lr * sign(g) up to eps
    assert np.allclose(step, -lr * np.sign(g[big]), rtol=1e-2, atol=1e-6)


def test_training_fits_a_smooth_field(need_gpu):
    import torch
    from winds.mlp import MLP
    net = MLP(4, 128, 3, 3, 1024, seed=1)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(8192, 4)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) * X[:, 1], np.cos(X[:, 2] + X[:, 3]), X[:, 0] * X[:, 3]],
                 axis=1).astype(np.float32)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    first = None
    for it in range(600):
        i = (it * 1024) % 8192
        net.train_step(Xd[i:i + 1024], Td[i:i + 1024], 2e-3, loss)
        if it == 0:
            first = float(loss.item())
    Y = torch.cat([net.forward(Xd[i:i + 1024]) for i in range(0, 8192, 1024)]).cpu().numpy()
    mse = float(np.mean((Y - T) ** 2))
    flush_print(f"[mlp train] first loss {first:.3e} final mse {mse:.3e}")
    assert mse <= 0.02 * first


def test_calls_follow_the_default_stream(need_gpu):
    """Inputs written by torch on the default stream just before the call are
    what the network reads, and its output is complete when torch reads it."""
    import torch
    net = make(max_batch=4096)
    params = [(W.astype(np.float64), b.astype(np.float64)) for W, b in net.unflatten(net.get_params())]
    torch.manual_seed(0)
    worst = 0.0
    for k in range(8):
        X = torch.randn(4096, 40, device="cuda") * (1.0 + k)
        Y = net.forward(X)
        Yo, _ = OM.forward(params, X.cpu().numpy(), bf=True)
        worst = max(worst, normwise(Y.cpu().numpy(), Yo))
    assert worst <= 1e-3


def test_batch_limits(need_gpu):
    import torch
    import winds
    net = make(max_batch=100)            # rounded up to 128 rows
    X = torch.zeros((129, 40), dtype=torch.float32, device="cuda")
    with pytest.raises(winds.WindsError):
        net.forward(X)
    Y = net.forward(X[:0])
    assert Y.shape == (0, 3)


def test_graph_replayed_steps_match_eager_steps(need_gpu):
    """Repeated train_step calls on the same batch buffers replay a recorded
    graph of the step from the second call on; a twin network fed the same
    values from alternating buffers runs every step eagerly.  Between steps
    the program does eager work of its own (a forward pass, torch kernels,
    synchronisations); after every step both networks hold the same
    parameters bit for bit, the losses agree (to the rounding of their
    atomic sum), and the learning rate and Adam
    scalars of each replay are the current ones (they differ every step)."""
    import torch
    a = make(hidden=256, layers=4, n_in=49, max_batch=1024)
    b = make(hidden=256, layers=4, n_in=49, max_batch=1024)
    assert np.array_equal(a.get_params(), b.get_params())
    torch.manual_seed(3)
    X = torch.randn(1000, 49, device="cuda")
    T = torch.randn(1000, 3, device="cuda") * 0.3
    X2, T2 = X.clone(), T.clone()
    la = torch.zeros(1, device="cuda")
    lb = torch.zeros(1, device="cuda")
    for it in range(6):
        lr = 1e-3 * (1.0 + 0.5 * it)
        a.train_step(X, T, lr, la)                       # same buffers: graph from step 2
        xb, tb = (X, T) if it % 2 else (X2, T2)
        b.train_step(xb, tb, lr, lb)                     # alternating buffers: eager
        _ = a.forward(X[:100])                           # eager work between replays
        X.add_(0.01)
        X2.add_(0.01)
        torch.cuda.synchronize()
        assert np.array_equal(a.get_params(), b.get_params()), it
        # the loss is an atomic sum over workgroups: equal to rounding
        assert abs(float(la.item()) - float(lb.item())) <= 1e-6 * float(lb.item()), it


@pytest.mark.parametrize("shape", [(49, 256, 4, 3), (200, 128, 2, 5), (7, 64, 1, 32),
                                   (256, 96, 3, 1)])
def test_fused_forward_equals_the_per_layer_forward(need_gpu, shape):
    """winds_mlp_forward in one kernel per 64 rows (activations in LDS) and
    as one GEMM per layer: the same MFMA sequence and epilogues, so bitwise
    the same outputs, for ragged batches and widths that are not multiples
    of the 128-k slices."""
    import torch
    from winds.mlp import MLP
    n_in, hidden, layers, n_out = shape
    net = MLP(n_in, hidden, layers, n_out, 10240, seed=11)
    rng = np.random.default_rng(5)
    flat = net.get_params()
    net.set_params((flat + 0.05 * rng.standard_normal(flat.shape)).astype(np.float32))
    for B in (1, 63, 64, 65, 1000, 6786, 10240):
        X = torch.from_numpy(rng.standard_normal((B, n_in)).astype(np.float32)).cuda()
        net.set_fused(True)
        Yf = net.forward(X).cpu().numpy()
        net.set_fused(False)
        Yl = net.forward(X).cpu().numpy()
        assert np.array_equal(Yf.view(np.uint32), Yl.view(np.uint32)), (shape, B)
    net.set_fused(True)


def test_fused_forward_refuses_wide_layers(need_gpu):
    from winds.mlp import MLP
    net = MLP(16, 512, 1, 2, 128, seed=1)
    with pytest.raises(Exception):
        net.set_fused(True)
    net.set_fused(False)


def test_fused_training_forward_equals_the_per_layer_forward(need_gpu):
    """Training steps whose forward pass runs in one kernel per 32 rows and
    steps whose forward runs one GEMM per layer: bitwise the same weights
    after every step, and the same losses."""
    import torch
    from winds.mlp import MLP
    rng = np.random.default_rng(9)
    a = MLP(49, 256, 4, 3, 4096, seed=3)
    b = MLP(49, 256, 4, 3, 4096, seed=3)
    b.set_fused(False)
    for B in (4096, 1000, 33):
        X = torch.from_numpy(rng.standard_normal((B, 49)).astype(np.float32)).cuda()
        T = torch.from_numpy(0.3 * rng.standard_normal((B, 3)).astype(np.float32)).cuda()
        la, lb = torch.zeros(1, device="cuda"), torch.zeros(1, device="cuda")
        for _ in range(3):
            a.train_step(X, T, 1e-3, la)
            b.train_step(X, T, 1e-3, lb)
        torch.cuda.synchronize()
        assert np.array_equal(a.get_params(), b.get_params()), B
        # the loss is a sum of workgroup partials added atomically
        assert abs(float(la.item()) - float(lb.item())) <= 1e-6 * abs(float(lb.item())), B
