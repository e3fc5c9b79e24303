"""The surrogate network's CPU oracle (oracle/mlp_oracle.cpp) pinned against
independent implementations: PyTorch's float64 autograd for the loss and its
gradient, PyTorch's bfloat16 conversion for the rounding, torch.optim.Adam
for the optimiser step, and central finite differences of the oracle's own
FP64 forward pass.  The reference has no learned component, so these are
the checks that fix the oracle (no reference golden vectors exist for it).
"""
import numpy as np
import pytest
import torch

import oracle as O

P = O.MLP_PARAMS
SIZES = [(256, 5), (256, 256), (256, 256), (256, 256), (9, 256)]


def _random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for n, k in SIZES:
        parts.append(rng.uniform(-1, 1, n * k) * scale * np.sqrt(6.0 / k))
        parts.append(rng.uniform(-0.1, 0.1, n))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == P
    return p


def _random_rows(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    t = np.zeros((n, 8), np.float32)
    t[:, 0] = rng.integers(0, 3, n)
    t[: n // 8, 0] = -1                                  # rays without a target
    t[:, 1] = rng.uniform(0.2, 3.0, n)
    t[:, 2] = rng.uniform(-1.0, 0.4, n)
    th, ph = rng.uniform(0, np.pi, n), rng.uniform(-np.pi, np.pi, n)
    t[:, 3:7] = np.stack([np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)], 1)
    t[:, 7] = rng.integers(0, 2, n)
    return x, t


def _torch_net(p, bf16=False):
    """Unpacked layers as float64 tensors (optionally rounded to bfloat16)."""
    out, o = [], 0
    for n, k in SIZES:
        W = torch.tensor(p[o:o + n * k].reshape(n, k), dtype=torch.float64)
        o += n * k
        b = torch.tensor(p[o:o + n], dtype=torch.float64)
        o += n
        if bf16:
            W = W.to(torch.bfloat16).to(torch.float64)
        out += [W, b]
    return out


def _torch_forward(layers, x, bf16=False):
    if bf16:
        xf = torch.tensor(x, dtype=torch.float32)
        hi = xf.to(torch.bfloat16).to(torch.float32)
        lo = (xf - hi).to(torch.bfloat16).to(torch.float32)
        h = hi.double() + lo.double()
    else:
        h = torch.tensor(x, dtype=torch.float64)
    for i in range(4):
        W, b = layers[2 * i], layers[2 * i + 1]
        h = torch.nn.functional.silu(h @ W.T + b)
        if bf16 and i < 3:                  # the last hidden layer enters as an exact hi + lo pair
            h = h.float().to(torch.bfloat16).double()
    W, b = layers[8], layers[9]
    return h @ W.T + b


def _torch_loss(y, t, w):
    t = torch.tensor(t, dtype=torch.float64)
    cls = t[:, 0].long()
    valid = cls >= 0
    ce = torch.nn.functional.cross_entropy(y[valid, :3], cls[valid], reduction="sum")
    disk = (cls == 2).double()
    esc = (t[:, 7] > 0.5).double() * valid.double()
    reg = (w[0] * disk * (y[:, 3] - t[:, 1]) ** 2 + w[1] * disk * (y[:, 4] - t[:, 2]) ** 2
           + w[2] * esc * ((y[:, 5:9] - t[:, 3:7]) ** 2).sum(1))
    return (ce + reg.sum()) / y.shape[0]


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.normal(0, 10, 2000), rng.uniform(-1e-3, 1e-3, 500),
                           [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 65504.0, 1e-40, 0.0]])
    vals = vals.astype(np.float32)
    ref = torch.tensor(vals).to(torch.bfloat16).float().numpy()
    got = np.array([O.mlp_bf16(float(v)) for v in vals], np.float32)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_matches_torch(bf16):
    p = _random_params(5)
    x, _ = _random_rows(64)
    y = O.mlp_forward(p, x, emulate=bf16)
    ref = _torch_forward(_torch_net(p, bf16), x, bf16).numpy()
    assert np.allclose(y, ref, rtol=1e-12 if not bf16 else 1e-6, atol=1e-12 if not bf16 else 1e-6)


def test_fp64_gradient_matches_torch_autograd():
    p = _random_params(7)
    x, t = _random_rows(48, seed=8)
    w = (3.0, 2.0, 0.5)
    loss, g = O.mlp_gradients(p, x, t, w, emulate=False)
    layers = [a.clone().requires_grad_(True) for a in _torch_net(p)]
    L = _torch_loss(_torch_forward(layers, x), t, w)
    L.backward()
    ref = torch.cat([a.grad.reshape(-1) for a in layers]).numpy()
    assert loss == pytest.approx(L.item(), rel=1e-12)
    assert np.allclose(g, ref, rtol=1e-9, atol=1e-13)


def test_fp64_gradient_finite_differences():
    """Central differences of the oracle's own FP64 loss on sampled
    parameters of every layer (weights and biases)."""
    p = _random_params(11).astype(np.float64)
    x, t = _random_rows(16, seed=12)
    w = (1.0, 1.0, 1.0)
    _, g = O.mlp_gradients(p.astype(np.float32), x, t, w, emulate=False)
    rng = np.random.default_rng(13)
    bounds = np.cumsum([0] + [n * k + n for n, k in SIZES])
    idx = np.concatenate([rng.integers(bounds[i], bounds[i + 1], 12) for i in range(5)])
    base = p.astype(np.float32)
    for i in idx:
        # parameters are float32 in the oracle's interface: step by a power of two
        h = 2.0 ** -10 * max(1.0, abs(float(base[i])))
        pp, pm = base.copy(), base.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = O.mlp_gradients(pp, x, t, w, emulate=False)
        lm, _ = O.mlp_gradients(pm, x, t, w, emulate=False)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        assert fd == pytest.approx(g[i], rel=2e-3, abs=1e-7), (i, fd, g[i])


def test_emulated_gradient_close_to_fp64():
    """The BF16 rounding points move the gradient by BF16-level amounts only."""
    p = _random_params(17)
    x, t = _random_rows(64, seed=18)
    _, g64 = O.mlp_gradients(p, x, t, emulate=False)
    _, gbf = O.mlp_gradients(p, x, t, emulate=True)
    rel = np.linalg.norm(gbf - g64) / np.linalg.norm(g64)
    assert rel < 0.05, rel


def test_adam_matches_torch():
    rng = np.random.default_rng(21)
    w0 = rng.normal(size=1000).astype(np.float32)
    grads = [rng.normal(size=1000).astype(np.float32) for _ in range(3)]
    w, m, v = w0.copy(), np.zeros_like(w0), np.zeros_like(w0)
    for k, g in enumerate(grads, 1):
        O.mlp_adam(w, g, m, v, 1e-2, k)
    tw = torch.tensor(w0.astype(np.float64), requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        opt.zero_grad()
        tw.grad = torch.tensor(g.astype(np.float64))
        opt.step()
    assert np.allclose(w, tw.detach().numpy(), rtol=1e-5, atol=1e-6)
