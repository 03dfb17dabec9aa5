"""Pins the CPU restatement of the U-net (oracle/unet.py): torch's conv3d /
conv_transpose3d semantics, central finite differences and torch autograd. CPU only."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet as U


def torch_forward(p, x):
    d = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in U.unpack(p.astype(np.float64)).items()}
    W = lambda name: d[name + '.W'].reshape(d[name + '.W'].shape[0], d[name + '.W'].shape[1], 3, 3, 3)
    xt = torch.tensor(x, dtype=torch.float64)[None]
    c = lambda a, nm, s: F.conv3d(a, W(nm), d[nm + '.b'], stride=s, padding=1)
    t = lambda a, nm: F.conv_transpose3d(a, W(nm), d[nm + '.b'], stride=2, padding=1)
    a0 = torch.relu(c(xt, 'L0', 1))
    a1 = torch.relu(c(a0, 'D1', 2))
    a2 = torch.relu(c(a1, 'L1', 1))
    a3 = torch.relu(c(a2, 'D2', 2))
    a4 = torch.relu(c(a3, 'L2', 1))
    a5 = torch.relu(t(a4, 'U1') + a2)
    a6 = torch.relu(t(a5, 'U0') + a0)
    return c(a6, 'OUT', 1)[0], d


def problem(n=9, seed=0):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed).astype(np.float64) + 0.05 * rng.standard_normal(U.NPARAM)
    x = rng.standard_normal((4, n, n, n))
    gy = rng.standard_normal((4, n, n, n))
    return p, x, gy


def test_sizes():
    assert U.coarse(65) == 33 and U.coarse(33) == 17
    assert U.NPARAM == sum(co * ci * 27 + co for _, _, ci, co, _ in U.LAYERS)


@pytest.mark.parametrize('n', [5, 9])
def test_forward_matches_torch(n):
    p, x, _ = problem(n)
    y = U.forward(p, x)
    yt, _ = torch_forward(p, x)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)


def test_backward_matches_torch_autograd():
    p, x, gy = problem(9, 1)
    _, g = U.backward(p, x, gy)
    yt, d = torch_forward(p, x)
    (yt * torch.tensor(gy)).sum().backward()
    gt = np.concatenate([np.concatenate([d[nm + '.W'].grad.numpy().ravel(), d[nm + '.b'].grad.numpy().ravel()])
                         for nm, *_ in U.LAYERS])
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-12)


def test_backward_matches_central_differences():
    p, x, gy = problem(5, 2)
    _, g = U.backward(p, x, gy)
    rng = np.random.default_rng(3)
    h = 1e-6
    for j in rng.integers(0, U.NPARAM, 25):
        e = np.zeros_like(p)
        e[j] = h
        fd = (np.sum(gy * U.forward(p + e, x)) - np.sum(gy * U.forward(p - e, x))) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(g[j])), (j, fd, g[j])


def test_bf16_mode_close():
    p, x, gy = problem(9, 4)
    y = U.forward(p, x)
    yb = U.forward(p, x, bf16=True)
    rel = np.linalg.norm(yb - y) / np.linalg.norm(y)
    assert 1e-5 < rel < 5e-2
