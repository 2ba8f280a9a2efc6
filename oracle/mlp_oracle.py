"""CPU oracle of the surrogate MLP (axicor_amd/csrc/surrogate.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker. The product
never imports it.

The surrogate has no reference counterpart. The reference's parameter search
is Optuna TPE (examples/bayesian_search.py:63-187), so this restates the
network from its own definition:

* layers d_in -> hidden x depth -> d_out, GELU (tanh form) after each hidden
  layer, linear output;
* loss = mean over the batch and the d_out real outputs of (y - t)^2;
* Adam with bias correction.

Parameters use the library's padded layout: per layer W[in_pad][out_pad]
row-major, then b[out_pad]. d_in and hidden are padded to multiples of 32,
d_out to a multiple of 16.

With ``emulate=True`` the forward and backward passes round to bf16 where the
GPU kernels store bf16: the input, the weights the GEMMs read, the hidden
activations and the pre-activation gradients. Everything else is float64.
With ``emulate=False`` everything is float64, for finite-difference checks.

Pinned in tests/test_mlp_oracle.py by:
* bf16 rounding against torch's float32 -> bfloat16 conversion;
* GELU and its derivative against torch's tanh-approximate GELU and autograd;
* analytic gradients against central finite differences and torch autograd;
* Adam against torch.optim.Adam.
"""
from __future__ import annotations

import numpy as np


def pad(x: int, m: int) -> int:
    return (x + m - 1) // m * m


def bf16_round(x) -> np.ndarray:
    """float32 -> bfloat16 -> float32, round to nearest even (NaN kept NaN)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


C_GELU = np.sqrt(2.0 / np.pi)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    du = C_GELU * (1.0 + 3.0 * 0.044715 * z * z)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * du


class Layout:
    """Padded layer sizes and the offsets of every W and b in the flat array."""

    def __init__(self, d_in: int, hidden: int, depth: int, d_out: int):
        self.d_in, self.hidden, self.depth, self.d_out = d_in, hidden, depth, d_out
        self.ins = [pad(d_in, 32)] + [pad(hidden, 32)] * depth
        self.outs = [pad(hidden, 32)] * depth + [pad(d_out, 16)]
        self.real_in = [d_in] + [hidden] * depth
        self.real_out = [hidden] * depth + [d_out]
        self.offsets = []
        off = 0
        for i, o in zip(self.ins, self.outs):
            self.offsets.append((off, off + i * o))
            off += i * o + o
        self.n_params = off

    def split(self, flat: np.ndarray):
        """[(W[in][out], b[out]), ...] as views of flat."""
        out = []
        for (ow, ob), i, o in zip(self.offsets, self.ins, self.outs):
            out.append((flat[ow:ow + i * o].reshape(i, o), flat[ob:ob + o]))
        return out

    def mask(self) -> np.ndarray:
        """1 on real parameters, 0 on padding."""
        m = np.zeros(self.n_params, np.float64)
        for (W, b), ri, ro in zip(self.split(m), self.real_in, self.real_out):
            W[:ri, :ro] = 1
            b[:ro] = 1
        return m

    def random_params(self, rng: np.random.Generator, scale: float = 1.0) -> np.ndarray:
        p = np.zeros(self.n_params, np.float64)
        for (W, b), ri, ro in zip(self.split(p), self.real_in, self.real_out):
            lim = np.sqrt(6.0 / ri) * scale
            W[:ri, :ro] = rng.uniform(-lim, lim, size=(ri, ro))
            b[:ro] = rng.uniform(-0.1, 0.1, size=ro)
        return p


def forward(lay: Layout, params: np.ndarray, X: np.ndarray, emulate: bool = True):
    """Returns (Y[n][d_out] float64, cache)."""
    rnd = bf16_round if emulate else (lambda a: np.asarray(a, np.float64))
    n = X.shape[0]
    H = np.zeros((n, lay.ins[0]), np.float64)
    H[:, :lay.d_in] = rnd(X)
    Hs, Zs, W16s = [H], [], []
    layers = lay.split(np.asarray(params, np.float64))
    for li, (W, b) in enumerate(layers):
        W16 = np.asarray(rnd(W), np.float64)
        W16s.append(W16)
        Z = Hs[-1] @ W16 + b
        Zs.append(Z)
        if li + 1 < len(layers):
            Hs.append(np.asarray(rnd(gelu(Z)), np.float64))
    Y = Zs[-1][:, :lay.d_out]
    return Y, (Hs, Zs, W16s)


def loss_of(Y: np.ndarray, T: np.ndarray) -> float:
    return float(np.mean((Y - T) ** 2))


def backward(lay: Layout, cache, Y: np.ndarray, T: np.ndarray, emulate: bool = True) -> np.ndarray:
    """Gradient of loss_of(Y, T) wrt the flat padded parameters."""
    rnd = bf16_round if emulate else (lambda a: np.asarray(a, np.float64))
    Hs, Zs, W16s = cache
    n = Y.shape[0]
    grads = np.zeros(lay.n_params, np.float64)
    gl = lay.split(grads)
    dZ = np.zeros((n, lay.outs[-1]), np.float64)
    dZ[:, :lay.d_out] = np.asarray(rnd(2.0 * (Y - T) / (n * lay.d_out)), np.float64)
    for li in range(len(W16s) - 1, -1, -1):
        gW, gb = gl[li]
        gW[:] = Hs[li].T @ dZ
        gb[:] = dZ.sum(axis=0)
        if li > 0:
            dZ = np.asarray(rnd((dZ @ W16s[li].T) * gelu_grad(Zs[li - 1])), np.float64)
    return grads


def adam(params, grads, m, v, t: int, lr: float, mask=None, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t counts from 1); returns (params, m, v).
    Entries with mask 0 are left as they are."""
    m2 = b1 * m + (1 - b1) * grads
    v2 = b2 * v + (1 - b2) * grads * grads
    p2 = params - lr * (m2 / (1 - b1 ** t)) / (np.sqrt(v2 / (1 - b2 ** t)) + eps)
    if mask is not None:
        keep = mask == 0
        m2 = np.where(keep, m, m2)
        v2 = np.where(keep, v, v2)
        p2 = np.where(keep, params, p2)
    return p2, m2, v2
