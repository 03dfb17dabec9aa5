"""CPU oracle of the per-DOF MLP initial-guess network -- TEST INFRASTRUCTURE.

Only the test-suite (and the smoke/bench checker legs) may import this module.

Network (the learned initial guess of a transient solve; no reference counterpart):
    x in R^4 -> 128 -> 128 -> 128 -> 1, ReLU between layers,
    y = W4 relu(W3 relu(W2 relu(W1 x + b1) + b2) + b3) + b4
Loss: mean squared error 0.5/B * sum_s (y_s - t_s)^2.
Parameter layout (float32, contiguous): W1[128,4], b1[128], W2[128,128], b2[128],
W3[128,128], b3[128], W4[1,128], b4[1].
Features of a DOF with states u0 (newest) .. u3 and segment scales s, U:
((u0-u1)/s, (u1-u2)/s, (u2-u3)/s, u0/U); target (u_next - u0)/s; guess u0 + s*y.
This component has no counterpart in the reference; parity of the GPU kernels
is pinned by this restatement, which is itself pinned by central finite
differences and an independent torch-autograd gradient (tests/test_oracle_mlp.py).

With ``bf16=True`` weights and the stored activations are rounded to bfloat16
(round to nearest even) exactly where the GPU kernels round them; all sums are
accumulated in float64 here (the kernels accumulate in float32).
"""
from __future__ import annotations

import numpy as np

H = 128
IN = 4
SHAPES = [('W1', (H, IN)), ('b1', (H,)), ('W2', (H, H)), ('b2', (H,)), ('W3', (H, H)), ('b3', (H,)),
          ('W4', (1, H)), ('b4', (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, s in SHAPES:
        k = int(np.prod(s))
        out[name] = p[o:o + k].reshape(s)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, s in SHAPES:
        if name.startswith('W'):
            fan_in = s[1]
            d[name] = rng.standard_normal(s) * np.sqrt(2.0 / fan_in)
        else:
            d[name] = np.zeros(s)
    d['W4'] *= 0.1
    return pack(d).astype(np.float32)


def forward(p, X, bf16=False, keep=False):
    """X: (B, 4). Returns y (B,) and, if keep, the pre-activations/activations per layer."""
    if bf16:
        q = {k: bf16_round(v) for k, v in unpack(np.asarray(p, dtype=np.float32)).items()}
    else:
        q = unpack(np.asarray(p, dtype=np.float64))
    q = {k: v.astype(np.float64) for k, v in q.items()}
    a = np.asarray(X, dtype=np.float64)
    acts = [a]
    zs = []
    for L in (1, 2, 3):
        z = a @ q[f'W{L}'].T + q[f'b{L}']
        a = np.maximum(z, 0.0)
        if bf16:
            a = bf16_round(a).astype(np.float64)
        zs.append(z)
        acts.append(a)
    y = (a @ q['W4'].T + q['b4'])[:, 0]
    if keep:
        return y, zs, acts, q
    return y


def loss_and_grad(p, X, T, bf16=False):
    """Loss 0.5/B sum (y - t)^2 and its gradient (packed like the parameters).

    bf16=True mirrors the GPU kernels: dL/dy is a float32, and each
    back-propagated dz_L = (dL/da_L) o [a_L > 0] is rounded to bfloat16 (it is an
    MFMA operand) before it feeds the weight gradients and the next layer."""
    y, zs, acts, q = forward(p, X, bf16=bf16, keep=True)
    B = X.shape[0]
    r = y - np.asarray(T, dtype=np.float64)
    loss = 0.5 * np.sum(r * r) / B
    d = r / B                                              # dL/dy
    if bf16:
        d = d.astype(np.float32).astype(np.float64)
    g = {'W4': (d @ acts[3])[None, :], 'b4': np.array([d.sum()])}
    delta = d[:, None] * q['W4'][0][None, :]               # dL/da3
    for L in (3, 2, 1):
        delta = delta * (acts[L] > 0)                      # dL/dz_L
        if bf16:
            delta = bf16_round(delta).astype(np.float64)
        g[f'W{L}'] = delta.T @ acts[L - 1]
        g[f'b{L}'] = delta.sum(0)
        if L > 1:
            delta = delta @ q[f'W{L}']
    return loss, pack(g)


def adam(p, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015) with bias correction; returns updated (p, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(u0, u1, u2, u3, s, U):
    """Per-DOF features of the last four states (u0 newest): scaled increments and the scaled state."""
    return np.stack([(u0 - u1) / s, (u1 - u2) / s, (u2 - u3) / s, u0 / U], 1)


def scales(snap, seg_off):
    """Per segment: s = rms of the increments u_j - u_{j-1}, U = rms of u_j (j >= 1); 1 where 0."""
    snap = np.asarray(snap, dtype=np.float64)
    s, U = [], []
    for k in range(len(seg_off) - 1):
        blk = snap[:, seg_off[k]:seg_off[k + 1]]
        si = np.sqrt(np.mean((blk[1:] - blk[:-1]) ** 2))
        sv = np.sqrt(np.mean(blk[1:] ** 2))
        s.append(si if si > 1e-300 else 1.0)
        U.append(sv if sv > 1e-300 else 1.0)
    return np.array(s), np.array(U)


def guess(p, hist, seg_off, s, U, bf16=True, base=0):
    """u0 + base (u0 - u1) + s_k * f(features) per DOF; hist = [u0, u1, u2, u3]."""
    n = hist[0].size
    sd = np.empty(n)
    Ud = np.empty(n)
    for k in range(len(seg_off) - 1):
        sd[seg_off[k]:seg_off[k + 1]] = np.float32(s[k])
        Ud[seg_off[k]:seg_off[k + 1]] = np.float32(U[k])
    X = features(*[np.asarray(h, dtype=np.float64) for h in hist], sd, Ud).astype(np.float32)
    return hist[0] + base * (hist[0] - hist[1]) + sd * forward(p, X, bf16=bf16)


def target(u_next, u0, u1, s, base=0):
    """Training target of a sample: (u_next - u0 - base (u0 - u1)) / s."""
    return (u_next - u0 - base * (u0 - u1)) / s
