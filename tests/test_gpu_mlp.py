"""The surrogate's HIP kernels against the CPU oracle (oracle/mlp_oracle.py):
the MFMA GEMM in all four operand layouts, the forward pass, the gradients of
every parameter, one Adam step, and training on a smooth map."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    return M.bf16_round(x).view(np.uint32).__rshift__(16).astype(np.uint16)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("M_,N_,K_", [(64, 64, 32), (104, 72, 40), (256, 200, 296), (8, 16, 8), (128, 8, 520),
                                      (384, 256, 512), (520, 136, 64), (128, 128, 24)])
def test_gemm_all_operand_layouts(gpu, ta, tb, M_, N_, K_):
    from axicor_amd.surrogate import gemm
    rng = np.random.default_rng(M_ * 7 + N_ * 3 + K_ + 2 * ta + tb)
    a = M.bf16_round(rng.standard_normal((M_, K_)).astype(np.float32)).astype(np.float64)
    b = M.bf16_round(rng.standard_normal((K_, N_)).astype(np.float32)).astype(np.float64)
    A_st = a.T if ta else a        # stored layouts
    B_st = b.T if tb else b
    c = gemm(bf16_bits(A_st.astype(np.float32)), bf16_bits(B_st.astype(np.float32)), ta, tb)
    ref = a @ b
    bound = np.abs(a) @ np.abs(b)
    # fp32 accumulation of exact bf16 products: |error| <= K * 2^-24 * sum|a||b|
    err = np.abs(c - ref)
    assert np.all(err <= 2 * K_ * 2.0 ** -24 * bound + 1e-30), float((err / (bound + 1e-30)).max())


def test_gemm_ragged_rows(gpu):
    """M not a multiple of 8 or of the 64- and 128-row tiles (row-major A
    only), with N below and above 128 (the two tile sizes)."""
    from axicor_amd.surrogate import gemm
    rng = np.random.default_rng(11)
    for M_, N_ in ((1, 24), (37, 24), (65, 24), (130, 24), (129, 256), (300, 136), (1000, 512)):
        a = M.bf16_round(rng.standard_normal((M_, 48)).astype(np.float32)).astype(np.float64)
        b = M.bf16_round(rng.standard_normal((N_, 48)).astype(np.float32)).astype(np.float64)
        c = gemm(bf16_bits(a.astype(np.float32)), bf16_bits(b.astype(np.float32)), False, True)
        np.testing.assert_allclose(c, a @ b.T, rtol=0, atol=1e-4 * (np.abs(a) @ np.abs(b.T)).max())


@pytest.mark.parametrize("M_,N_,K_", [(8300, 520, 200), (32768, 128, 512), (16385, 264, 24)])
def test_gemm_big_tiles(gpu, M_, N_, K_):
    """Products with at least 256 tiles of 128x128 and both operands
    contiguous in K take the double-buffered 128x128 kernel: ragged M and N,
    K a multiple of 8 but not of the 32-deep step, and K below one step."""
    from axicor_amd.surrogate import gemm
    rng = np.random.default_rng(M_ + N_ + K_)
    a = M.bf16_round(rng.standard_normal((M_, K_)).astype(np.float32)).astype(np.float64)
    b = M.bf16_round(rng.standard_normal((N_, K_)).astype(np.float32)).astype(np.float64)
    c = gemm(bf16_bits(a.astype(np.float32)), bf16_bits(b.astype(np.float32)), False, True)
    ref = a @ b.T
    bound = np.abs(a) @ np.abs(b.T)
    err = np.abs(c - ref)
    assert np.all(err <= 2 * K_ * 2.0 ** -24 * bound + 1e-30), float((err / (bound + 1e-30)).max())


def _net(seed=1, max_batch=512):
    from axicor_amd.surrogate import Mlp
    net = Mlp(16, 512, 3, 8, max_batch=max_batch, seed=seed)
    lay = M.Layout(16, 512, 3, 8)
    assert net.n_params == lay.n_params and net.ins == lay.ins and net.outs == lay.outs
    return net, lay


def test_initial_parameters_respect_the_padding(gpu):
    net, lay = _net()
    p = net.get_params().astype(np.float64)
    mask = lay.mask()
    assert np.all(p[mask == 0] == 0)
    for (W, b), ri in zip(lay.split(p), lay.real_in):
        lim = np.sqrt(6.0 / ri)
        assert np.abs(W).max() <= lim and np.abs(W).max() > 0.9 * lim
        assert np.all(b == 0)
    net.destroy()


def test_forward_matches_oracle(gpu):
    net, lay = _net()
    rng = np.random.default_rng(2)
    params = lay.random_params(rng).astype(np.float32)
    net.set_params(params)
    X = rng.standard_normal((300, 16)).astype(np.float32)   # 300: ragged last tile
    Y = net.predict(X)
    Yo, _ = M.forward(lay, params.astype(np.float64), X, emulate=True)
    Yx, _ = M.forward(lay, params.astype(np.float64), X, emulate=False)
    rel = np.linalg.norm(Y - Yo) / np.linalg.norm(Yo)
    rel_exact = np.linalg.norm(Y - Yx) / np.linalg.norm(Yx)
    print(f"forward: rel error vs bf16-emulating oracle {rel:.2e}, vs float64 network {rel_exact:.2e}")
    assert rel < 5e-3
    assert rel_exact < 3e-2
    # deterministic, and chunking over max_batch gives the same rows
    assert np.array_equal(net.predict(X), Y)
    X2 = rng.standard_normal((1200, 16)).astype(np.float32)
    Y2 = net.predict(X2)
    assert np.array_equal(Y2[:300], net.predict(X2[:300]))
    assert np.array_equal(Y2[1024:], net.predict(X2[1024:]))
    net.destroy()


def test_fused_forward_equals_layer_by_layer(gpu):
    """The fused kernel (128 rows per workgroup, activations in LDS through
    all layers) runs the same MFMA steps in the same order and the same
    epilogues as the layer-by-layer pass: the outputs are equal bit for bit,
    for ragged row counts and for more rows than the layer-by-layer buffers
    hold (chunked there, one launch fused)."""
    import time
    from axicor_amd._ffi import day_lib
    from axicor_amd.engine import DeviceBuffer
    net, lay = _net(seed=7, max_batch=4096)
    rng = np.random.default_rng(8)
    net.set_params(lay.random_params(rng).astype(np.float32))
    assert net.set_fused(True)
    for n in (1, 127, 128, 129, 300, 4096, 10001):
        X = rng.standard_normal((n, 16)).astype(np.float32)
        net.set_fused(True)
        Yf = net.predict(X)
        net.set_fused(False)
        Yl = net.predict(X)
        assert np.array_equal(Yf, Yl), f"n={n}: max |diff| {np.abs(Yf - Yl).max()}"
    # Throughput on a million candidates, both ways.
    n = 1 << 20
    dx = DeviceBuffer.from_array(rng.standard_normal((n, 16)).astype(np.float32))
    dy = DeviceBuffer(n * 8 * 4)
    for fused in (True, False):
        net.set_fused(fused)
        net.forward_device(dx.ptr, n, dy.ptr)
        day_lib().gpu_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            net.forward_device(dx.ptr, n, dy.ptr)
        day_lib().gpu_device_synchronize()
        print(f"forward of {n} rows, {'fused' if fused else 'layer by layer'}: "
              f"{(time.perf_counter() - t0) / 3 * 1e3:.2f} ms")
    dx.free()
    dy.free()
    net.set_fused(True)
    net.destroy()


def test_forward_with_big_tiles_matches_oracle(gpu):
    """A batch of 40,000 rows: every layer's product has at least 313 tiles
    of 128x128 and takes the double-buffered kernel."""
    net, lay = _net(seed=3, max_batch=40000)
    rng = np.random.default_rng(4)
    params = lay.random_params(rng).astype(np.float32)
    net.set_params(params)
    X = rng.standard_normal((40000, 16)).astype(np.float32)
    Y = net.predict(X)
    Yo, _ = M.forward(lay, params.astype(np.float64), X, emulate=True)
    rel = np.linalg.norm(Y - Yo) / np.linalg.norm(Yo)
    print(f"forward (big tiles): rel error vs bf16-emulating oracle {rel:.2e}")
    assert rel < 5e-3
    net.destroy()


def test_gradients_match_oracle(gpu):
    net, lay = _net()
    rng = np.random.default_rng(3)
    params = lay.random_params(rng).astype(np.float32)
    net.set_params(params)
    X = rng.standard_normal((257, 16)).astype(np.float32)
    T = rng.standard_normal((257, 8)).astype(np.float32)
    g, loss = net.gradients(X, T)
    Yo, cache = M.forward(lay, params.astype(np.float64), X, emulate=True)
    go = M.backward(lay, cache, Yo, T, emulate=True)
    assert abs(loss - M.loss_of(Yo, T)) < 1e-3 * M.loss_of(Yo, T)
    assert np.all(g[lay.mask() == 0] == 0)
    for li, ((gW, gb), (oW, ob)) in enumerate(zip(lay.split(g.astype(np.float64)), lay.split(go))):
        rw = np.linalg.norm(gW - oW) / np.linalg.norm(oW)
        rb = np.linalg.norm(gb - ob) / np.linalg.norm(ob)
        print(f"layer {li}: rel error dW {rw:.2e} db {rb:.2e}")
        assert rw < 2e-2 and rb < 2e-2, (li, rw, rb)
    net.destroy()


def test_adam_step_matches_oracle(gpu):
    net, lay = _net()
    rng = np.random.default_rng(4)
    params = lay.random_params(rng).astype(np.float32)
    net.set_params(params)
    X = rng.standard_normal((200, 16)).astype(np.float32)
    T = rng.standard_normal((200, 8)).astype(np.float32)
    g, _ = net.gradients(X, T)
    lr = 1e-3
    net.train_step(X, T, lr)
    got = net.get_params().astype(np.float64)
    mask = lay.mask()
    want, _, _ = M.adam(params.astype(np.float64), g.astype(np.float64), np.zeros(lay.n_params),
                        np.zeros(lay.n_params), 1, lr, mask=mask)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6 * lr)
    assert np.all(got[mask == 0] == 0)
    net.destroy()


def test_training_fits_a_smooth_map(gpu):
    net, lay = _net(seed=9, max_batch=512)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((16, 8)) / 4
    Bm = rng.standard_normal((16, 8)) / 4

    def f(X):
        return (np.sin(X @ A) + 0.5 * (X @ Bm) ** 2).astype(np.float32)

    Xv = rng.uniform(-1, 1, (1024, 16)).astype(np.float32)
    base = float(np.mean((f(Xv) - f(Xv).mean(axis=0)) ** 2))
    losses = []
    for step in range(400):
        X = rng.uniform(-1, 1, (512, 16)).astype(np.float32)
        losses.append(net.train_step(X, f(X), 1e-3))
    val = float(np.mean((net.predict(Xv) - f(Xv)) ** 2))
    print(f"training: first loss {losses[0]:.4f} last {losses[-1]:.5f}; validation {val:.5f} vs variance {base:.4f}")
    assert losses[-1] < losses[0] / 10
    assert val < base / 20
    net.destroy()
