"""CPU oracle of the 3-level U-net preconditioner correction -- TEST INFRASTRUCTURE.

Only the test-suite (and the smoke/bench checker legs) may import this module.
No counterpart in the reference; the GPU kernels are pinned by this
restatement, itself pinned by central finite differences and torch autograd
(tests/test_oracle_unet.py).

Data live on the vertex grid of a structured (n-1)^3 cube mesh, channel-major:
x[c, k, j, i]. Levels: n0 = n, n1 = (n0-1)/2+1, n2 = (n1-1)/2+1 (vertex-centred
coarsening: coarse vertex i is fine vertex 2i). Layers (3x3x3 kernels, padding 1):

    a0 = relu(conv (x;  4 -> C0))          level 0
    a1 = relu(conv2(a0; C0 -> C1))         stride 2 -> level 1
    a2 = relu(conv (a1; C1 -> C1))
    a3 = relu(conv2(a2; C1 -> C2))         stride 2 -> level 2
    a4 = relu(conv (a3; C2 -> C2))
    a5 = relu(tconv2(a4; C2 -> C1) + a2)   transposed stride 2 -> level 1, skip
    a6 = relu(tconv2(a5; C1 -> C0) + a0)   -> level 0, skip
    y  =      conv (a6; C0 -> 4)           linear

conv stride s:   out[o, z] = b[o] + sum_{i,d} W[o, i, d] in[i, s z + d - 1]
tconv stride 2:  out[o, z] = b[o] + sum_{i,d} W[i, o, d] in[i, y],  2 y + d - 1 = z
(d in {0,1,2}^3 is the kernel offset; out-of-range inputs are 0). Weights are
stored [out][in][27] for conv and [in][out][27] for tconv (PyTorch's layouts).

With bf16=True the input, the weights and biases and every activation a_k are
rounded to bfloat16 where the GPU rounds them; sums are float64 here.
"""
from __future__ import annotations

import numpy as np

from .mlp import bf16_round

CIN = 4
C0, C1, C2 = 8, 16, 32
# (name, kind, cin, cout, stride); kind: 'c' conv, 't' transposed conv
LAYERS = [('L0', 'c', CIN, C0, 1), ('D1', 'c', C0, C1, 2), ('L1', 'c', C1, C1, 1), ('D2', 'c', C1, C2, 2),
          ('L2', 'c', C2, C2, 1), ('U1', 't', C2, C1, 2), ('U0', 't', C1, C0, 2), ('OUT', 'c', C0, CIN, 1)]


def layer_sizes():
    out = []
    for name, kind, ci, co, s in LAYERS:
        out.append((name, co * ci * 27, co))
    return out


NPARAM = sum(w + b for _, w, b in layer_sizes())


def unpack(p):
    d, o = {}, 0
    for (name, kind, ci, co, s) in LAYERS:
        nw = co * ci * 27
        shape = (co, ci, 27) if kind == 'c' else (ci, co, 27)
        d[name + '.W'] = p[o:o + nw].reshape(shape)
        o += nw
        d[name + '.b'] = p[o:o + co]
        o += co
    return d


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for (name, kind, ci, co, s) in LAYERS:
        fan_in = ci * 27 if kind == 'c' else ci * 27 / 8.0
        w = rng.standard_normal(co * ci * 27) * np.sqrt(2.0 / fan_in)
        if name == 'OUT':
            w *= 0.1
        parts += [w, np.zeros(co)]
    return np.concatenate(parts).astype(np.float32)


def coarse(n):
    return (n - 1) // 2 + 1


_OFF = [(dz, dy, dx) for dz in range(3) for dy in range(3) for dx in range(3)]


def _shift_slices(n_in, n_out, s, d):
    """For out index z in [0, n_out): in index s*z + d - 1. Returns (out slice, in slice) of the valid range."""
    lo = 0
    while s * lo + d - 1 < 0:
        lo += 1
    hi = n_out
    while hi > lo and s * (hi - 1) + d - 1 > n_in - 1:
        hi -= 1
    return slice(lo, hi), slice(s * lo + d - 1, s * (hi - 1) + d - 1 + 1, s)


def conv(x, W, b, s):
    """x: (ci, n, n, n); W: (co, ci, 27); stride s; returns (co, m, m, m)."""
    ci, n = x.shape[0], x.shape[1]
    m = (n - 1) // s + 1
    co = W.shape[0]
    out = np.broadcast_to(b[:, None, None, None], (co, m, m, m)).astype(np.float64).copy()
    for k, (dz, dy, dx) in enumerate(_OFF):
        oz, iz = _shift_slices(n, m, s, dz)
        oy, iy = _shift_slices(n, m, s, dy)
        ox, ix = _shift_slices(n, m, s, dx)
        out[:, oz, oy, ox] += np.einsum('oi,izyx->ozyx', W[:, :, k], x[:, iz, iy, ix])
    return out


def conv_adjoint_input(g, W, s, n_in):
    """Adjoint of conv w.r.t. its input: g (co, m, m, m) -> (ci, n_in^3)."""
    co, m = g.shape[0], g.shape[1]
    ci = W.shape[1]
    out = np.zeros((ci, n_in, n_in, n_in))
    for k, (dz, dy, dx) in enumerate(_OFF):
        oz, iz = _shift_slices(n_in, m, s, dz)
        oy, iy = _shift_slices(n_in, m, s, dy)
        ox, ix = _shift_slices(n_in, m, s, dx)
        out[:, iz, iy, ix] += np.einsum('oi,ozyx->izyx', W[:, :, k], g[:, oz, oy, ox])
    return out


def conv_grad_weight(g, x, s):
    """dL/dW for out = conv(x, W, b, s) given g = dL/dout: (co, ci, 27)."""
    ci, n = x.shape[0], x.shape[1]
    co, m = g.shape[0], g.shape[1]
    gw = np.zeros((co, ci, 27))
    for k, (dz, dy, dx) in enumerate(_OFF):
        oz, iz = _shift_slices(n, m, s, dz)
        oy, iy = _shift_slices(n, m, s, dy)
        ox, ix = _shift_slices(n, m, s, dx)
        gw[:, :, k] = np.einsum('ozyx,izyx->oi', g[:, oz, oy, ox], x[:, iz, iy, ix])
    return gw


def tconv(x, W, b, n_out):
    """Transposed stride-2 conv, the adjoint of the stride-2 conv: x (ci, m^3) at the coarse
    level, W (ci, co, 27) -> (co, n_out^3)."""
    return conv_adjoint_input(x, W, 2, n_out) + b[:, None, None, None]


def tconv_adjoint_input(g, W):
    """Adjoint of tconv w.r.t. its input: the stride-2 conv with the same weights [ci][co]."""
    return conv(g, W, np.zeros(W.shape[0]), 2)


def _tconv_gw(g, x):
    # out[o, 2y + d - 1] += W[i, o, d] x[i, y]  =>  gW[i, o, d] = sum_y x[i, y] g[o, 2y + d - 1]
    ci, m = x.shape[0], x.shape[1]
    co, n = g.shape[0], g.shape[1]
    gw = np.zeros((ci, co, 27))
    for k, (dz, dy, dx) in enumerate(_OFF):
        oz, iz = _shift_slices(n, m, 2, dz)
        oy, iy = _shift_slices(n, m, 2, dy)
        ox, ix = _shift_slices(n, m, 2, dx)
        gw[:, :, k] = np.einsum('izyx,ozyx->io', x[:, oz, oy, ox], g[:, iz, iy, ix])
    return gw


def forward(p, x, bf16=False, keep=False):
    """x: (4, n, n, n). Returns y (4, n, n, n) [and the activations]."""
    if bf16:
        q = {k: bf16_round(v).astype(np.float64) for k, v in unpack(np.asarray(p, dtype=np.float32)).items()}
        x = bf16_round(np.asarray(x, dtype=np.float32)).astype(np.float64)
    else:
        q = unpack(np.asarray(p, dtype=np.float64))
        x = np.asarray(x, dtype=np.float64)
    n0 = x.shape[1]
    n1, n2 = coarse(n0), coarse(coarse(n0))
    R = (lambda a: bf16_round(a.astype(np.float32)).astype(np.float64)) if bf16 else (lambda a: a)
    relu = lambda a: np.maximum(a, 0.0)
    a0 = R(relu(conv(x, q['L0.W'], q['L0.b'], 1)))
    a1 = R(relu(conv(a0, q['D1.W'], q['D1.b'], 2)))
    a2 = R(relu(conv(a1, q['L1.W'], q['L1.b'], 1)))
    a3 = R(relu(conv(a2, q['D2.W'], q['D2.b'], 2)))
    a4 = R(relu(conv(a3, q['L2.W'], q['L2.b'], 1)))
    a5 = R(relu(tconv(a4, q['U1.W'], q['U1.b'], n1) + a2))
    a6 = R(relu(tconv(a5, q['U0.W'], q['U0.b'], n0) + a0))
    y = conv(a6, q['OUT.W'], q['OUT.b'], 1)
    if keep:
        return y, dict(x=x, a0=a0, a1=a1, a2=a2, a3=a3, a4=a4, a5=a5, a6=a6), q
    return y


def backward(p, x, gy, bf16=False, acts=None):
    """Gradient of sum(gy * y) w.r.t. the parameters (packed like p).

    acts: optional activations {x, a0..a6} to back-propagate through instead of this
    module's own forward pass (e.g. the ones a GPU forward pass stored), so that a
    comparison of the backward pass is not blurred by bf16 rounding flips upstream."""
    y, a, q = forward(p, x, bf16=bf16, keep=True)
    if acts is not None:
        a = {k: np.asarray(v, dtype=np.float64) for k, v in acts.items()}
    g = {}
    gy = np.asarray(gy, dtype=np.float64)
    g['OUT.W'] = conv_grad_weight(gy, a['a6'], 1)
    g['OUT.b'] = gy.sum((1, 2, 3))
    g6 = conv_adjoint_input(gy, q['OUT.W'], 1, a['a6'].shape[1]) * (a['a6'] > 0)
    g['U0.W'] = _tconv_gw(g6, a['a5'])
    g['U0.b'] = g6.sum((1, 2, 3))
    g5 = tconv_adjoint_input(g6, q['U0.W']) * (a['a5'] > 0)
    g['U1.W'] = _tconv_gw(g5, a['a4'])
    g['U1.b'] = g5.sum((1, 2, 3))
    g4 = tconv_adjoint_input(g5, q['U1.W']) * (a['a4'] > 0)
    g['L2.W'] = conv_grad_weight(g4, a['a3'], 1)
    g['L2.b'] = g4.sum((1, 2, 3))
    g3 = conv_adjoint_input(g4, q['L2.W'], 1, a['a3'].shape[1]) * (a['a3'] > 0)
    g['D2.W'] = conv_grad_weight(g3, a['a2'], 2)
    g['D2.b'] = g3.sum((1, 2, 3))
    g2 = (conv_adjoint_input(g3, q['D2.W'], 2, a['a2'].shape[1]) + g5) * (a['a2'] > 0)
    g['L1.W'] = conv_grad_weight(g2, a['a1'], 1)
    g['L1.b'] = g2.sum((1, 2, 3))
    g1 = conv_adjoint_input(g2, q['L1.W'], 1, a['a1'].shape[1]) * (a['a1'] > 0)
    g['D1.W'] = conv_grad_weight(g1, a['a0'], 2)
    g['D1.b'] = g1.sum((1, 2, 3))
    g0 = (conv_adjoint_input(g1, q['D1.W'], 2, a['a0'].shape[1]) + g6) * (a['a0'] > 0)
    g['L0.W'] = conv_grad_weight(g0, a['x'], 1)
    g['L0.b'] = g0.sum((1, 2, 3))
    parts = []
    for (name, kind, ci, co, s) in LAYERS:
        parts += [g[name + '.W'].ravel(), g[name + '.b'].ravel()]
    return y, np.concatenate(parts)
