"""Pins of the scheduler network's CPU oracle (oracle/mlp.py): its gradient equals torch.autograd's
in float64 and central finite differences; its forward equals torch's layers; its Adam step equals
torch.optim.Adam's."""

from __future__ import annotations

import numpy as np
import pytest

import mlp

torch = pytest.importorskip("torch")


def batch(seed, count=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.PARAMS)
    x = rng.normal(0, 2, (count, 12))
    t = rng.normal(0, 1, (count, 3))
    mask = (rng.random((count, 3)) < 0.8).astype(np.float64)
    mask[0] = 1
    mean = rng.normal(0, 1, 12)
    scale = rng.uniform(0.5, 2, 12)
    return p, x, t, mask, mean, scale


def torch_loss(p, x, t, mask, mean, scale):
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    w, i = {}, 0
    for name, shape in mlp.SIZES:
        k = int(np.prod(shape))
        w[name] = pt[i:i + k].reshape(shape)
        i += k
    xn = (torch.tensor(x) - torch.tensor(mean)) * torch.tensor(scale)
    h1 = torch.relu(torch.nn.functional.linear(xn, w["W1"], w["b1"]))
    h2 = torch.relu(torch.nn.functional.linear(h1, w["W2"], w["b2"]))
    y = torch.nn.functional.linear(h2, w["W3"], w["b3"])
    m = torch.tensor(mask)
    value = torch.sum(m * (y - torch.tensor(t)) ** 2) / torch.sum(m)
    value.backward()
    return float(value.detach()), pt.grad.numpy(), y.detach().numpy()


def test_parameter_count():
    assert mlp.PARAMS == 1571


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forward_and_gradient_match_torch_autograd(seed):
    p, x, t, mask, mean, scale = batch(seed)
    value, grad = mlp.loss_grad(p, x, t, mask, mean, scale)
    tv, tg, ty = torch_loss(p, x, t, mask, mean, scale)
    y, _ = mlp.forward(p, x, mean, scale)
    assert np.allclose(y, ty, rtol=1e-12, atol=1e-12)
    assert abs(value - tv) <= 1e-12 * max(1.0, abs(tv))
    assert np.allclose(grad, tg, rtol=1e-10, atol=1e-12)


def test_gradient_matches_central_differences():
    p, x, t, mask, mean, scale = batch(7, count=12)
    _, grad = mlp.loss_grad(p, x, t, mask, mean, scale)
    rng = np.random.default_rng(0)
    for i in rng.choice(mlp.PARAMS, size=60, replace=False):
        h = 1e-6
        up, dn = p.copy(), p.copy()
        up[i] += h
        dn[i] -= h
        fd = (mlp.loss(up, x, t, mask, mean, scale) - mlp.loss(dn, x, t, mask, mean, scale)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), i


def test_a_planted_sign_error_is_caught():
    """The pins above are sensitive: flipping one term of the backward pass breaks them."""
    p, x, t, mask, mean, scale = batch(4)
    _, grad = mlp.loss_grad(p, x, t, mask, mean, scale)
    bad = grad.copy()
    w3 = 32 * 12 + 32 + 32 * 32 + 32
    bad[w3:w3 + 96] *= -1.0
    _, tg, _ = torch_loss(p, x, t, mask, mean, scale)
    assert not np.allclose(bad, tg, rtol=1e-6, atol=1e-9)


def test_adam_matches_torch_optim():
    p, x, t, mask, mean, scale = batch(5)
    ours, _ = mlp.adam(p, x, t, mask, steps=7, lr=1e-2, mean=mean, scale=scale)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(7):
        opt.zero_grad()
        _, g = mlp.loss_grad(pt.detach().numpy(), x, t, mask, mean, scale)
        pt.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(ours, pt.detach().numpy(), rtol=1e-9, atol=1e-12)
