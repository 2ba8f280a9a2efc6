"""The surrogate's CPU oracle, pinned against independent implementations:
torch's bf16 conversion, tanh-approximate GELU, autograd and Adam, and
central finite differences. No GPU."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 20000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 3.4e38, -3.4e38, 1e-40, np.inf, -np.inf],
                                 np.float32)]).astype(np.float32)
    # values exactly halfway between two bf16 numbers (ties to even)
    ties = (np.arange(1, 2000, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    ours = M.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_gelu_matches_torch_tanh_gelu_and_its_derivative():
    z = np.linspace(-8, 8, 4001)
    t = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    g = torch.nn.functional.gelu(t, approximate="tanh")
    g.sum().backward()
    np.testing.assert_allclose(M.gelu(z), g.detach().numpy(), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(M.gelu_grad(z), t.grad.numpy(), rtol=1e-10, atol=1e-13)
    # limits and the special point
    assert M.gelu(0.0) == 0.0 and M.gelu_grad(0.0) == 0.5


def _torch_loss(lay, params, X, T):
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    h = torch.zeros((X.shape[0], lay.ins[0]), dtype=torch.float64)
    h[:, :lay.d_in] = torch.tensor(X, dtype=torch.float64)
    for li, ((ow, ob), i, o) in enumerate(zip(lay.offsets, lay.ins, lay.outs)):
        W = p[ow:ow + i * o].reshape(i, o)
        b = p[ob:ob + o]
        z = h @ W + b
        h = torch.nn.functional.gelu(z, approximate="tanh") if li + 1 < len(lay.offsets) else z
    loss = ((h[:, :lay.d_out] - torch.tensor(T, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy()


@pytest.mark.parametrize("shape", [(5, 7, 2, 3), (16, 40, 3, 8)])
def test_analytic_gradients_match_autograd_and_finite_differences(shape):
    lay = M.Layout(*shape)
    rng = np.random.default_rng(7)
    params = lay.random_params(rng)
    X = rng.standard_normal((13, lay.d_in))
    T = rng.standard_normal((13, lay.d_out))
    Y, cache = M.forward(lay, params, X, emulate=False)
    g = M.backward(lay, cache, Y, T, emulate=False)
    loss_t, g_t = _torch_loss(lay, params, X, T)
    assert abs(M.loss_of(Y, T) - loss_t) < 1e-12
    np.testing.assert_allclose(g, g_t, rtol=1e-9, atol=1e-12)
    # padding carries no gradient
    assert np.all(g[lay.mask() == 0] == 0)
    # central differences on a sample of real parameters
    real = np.flatnonzero(lay.mask())
    idx = rng.choice(real, size=min(60, real.size), replace=False)
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        f1 = M.loss_of(M.forward(lay, p1, X, emulate=False)[0], T)
        f2 = M.loss_of(M.forward(lay, p2, X, emulate=False)[0], T)
        fd = (f1 - f2) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_emulated_pass_stays_close_to_float64():
    """bf16 rounding moves the outputs and gradients by about 2^-8
    relative: the emulation is a perturbation of the exact network."""
    lay = M.Layout(16, 64, 3, 8)
    rng = np.random.default_rng(3)
    params = lay.random_params(rng)
    X = rng.standard_normal((64, 16))
    T = rng.standard_normal((64, 8))
    Y, c = M.forward(lay, params, X, emulate=False)
    Ye, ce = M.forward(lay, params, X, emulate=True)
    assert np.abs(Y - Ye).max() < 0.05 * np.abs(Y).max()
    g = M.backward(lay, c, Y, T, emulate=False)
    ge = M.backward(lay, ce, Ye, T, emulate=True)
    assert np.linalg.norm(g - ge) < 0.05 * np.linalg.norm(g)


def test_adam_matches_torch():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, 1):
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        p, m, v = M.adam(p, g, m, v, t, 1e-2)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)
    mask = (np.arange(50) % 3 != 0).astype(float)
    p2, _, _ = M.adam(p0, grads[0], np.zeros(50), np.zeros(50), 1, 1e-2, mask=mask)
    assert np.array_equal(p2[mask == 0], p0[mask == 0])
