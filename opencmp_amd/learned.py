"""Learned initial guess for transient solves (ocmp_mlp_* in include/ocmp.h).

A per-DOF MLP reads the last four states of each DOF and predicts the next
one; the prediction becomes the initial guess of the next linear solve. The
reference starts every solve from the previous solution held in the
GridFunction (opencmp/models/ins.py:323-359,
opencmp/solvers/time_integration_schemes.py:26-80); :func:`choose_guess`
keeps that plain guess whenever the learned one does not lower the FP64
residual, so the learned path never starts a solve from a worse point.

Training runs on the GPU through the library's own forward, backward and Adam
kernels on snapshots the plain solver produced.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from .core import Context, _ptr

NPARAM = L.MLP_NPARAM
HID = L.MLP_HIDDEN
SHAPES = [('W1', (HID, 4)), ('b1', (HID,)), ('W2', (HID, HID)), ('b2', (HID,)), ('W3', (HID, HID)),
          ('b3', (HID,)), ('W4', (1, HID)), ('b4', (1,))]


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights, zero biases; the output layer is scaled down so the
    untrained network predicts increments near 0."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith('W'):
            w = rng.standard_normal(shape) * math.sqrt(2.0 / shape[1])
            if name == 'W4':
                w *= 0.1
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(shape).ravel())
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAM
    return p


class MLPGuess:
    """Device-resident network, its Adam state and the per-segment feature scales."""

    def __init__(self, ctx: Optional[Context] = None, params: Optional[np.ndarray] = None, seed: int = 0,
                 max_batch: int = 65536, base: int = 0):
        """base 1: the network predicts the correction to the linear extrapolation 2 u0 - u1
        (training targets and the guess follow, ocmp_mlp_set_base); base 0: to u0."""
        self.ctx = ctx or Context.default()
        self.lib = self.ctx.lib
        self.max_batch = int(max_batch)
        p = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f'expected {NPARAM} parameters, got {p.size}')
        h = C.c_void_p()
        L.check(self.lib.ocmp_mlp_create(self.ctx.h, p.ctypes.data_as(C.c_void_p), self.max_batch, C.byref(h)),
                'mlp_create')
        self.h = h
        self.losses: List[float] = []
        L.check(self.lib.ocmp_mlp_set_base(self.h, int(base)), 'mlp_set_base')

    @property
    def base(self) -> int:
        b = C.c_int()
        L.check(self.lib.ocmp_mlp_get_base(self.h, C.byref(b)), 'mlp_get_base')
        return b.value

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                self.lib.ocmp_mlp_destroy(h)
            except Exception:
                pass

    # ------------------------------------------------------------ parameters and scales
    @property
    def params(self) -> np.ndarray:
        out = np.empty(NPARAM, dtype=np.float32)
        L.check(self.lib.ocmp_mlp_get_params(self.h, out.ctypes.data_as(C.c_void_p)), 'mlp_get_params')
        return out

    @params.setter
    def params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f'expected {NPARAM} parameters, got {p.size}')
        L.check(self.lib.ocmp_mlp_set_params(self.h, p.ctypes.data_as(C.c_void_p)), 'mlp_set_params')

    def set_scales(self, seg_off: Sequence[int], s: Sequence[float], U: Sequence[float]) -> None:
        k = len(seg_off) - 1
        off = (C.c_int64 * (k + 1))(*[int(v) for v in seg_off])
        L.check(self.lib.ocmp_mlp_set_scales(self.h, k, off, (C.c_double * k)(*s), (C.c_double * k)(*U)),
                'mlp_set_scales')

    def scales(self) -> Tuple[List[int], List[float], List[float]]:
        k = C.c_int()
        off = (C.c_int64 * (L.MLP_MAXSEG + 1))()
        s = (C.c_double * L.MLP_MAXSEG)()
        U = (C.c_double * L.MLP_MAXSEG)()
        L.check(self.lib.ocmp_mlp_get_scales(self.h, C.byref(k), off, s, U), 'mlp_get_scales')
        n = k.value
        return list(off[:n + 1]), list(s[:n]), list(U[:n])

    def fit_scales(self, snap: torch.Tensor, seg_off: Sequence[int]) -> None:
        """snap: (nsnap, n) float64 on the device, oldest first."""
        snap = self._snap(snap)
        k = len(seg_off) - 1
        off = (C.c_int64 * (k + 1))(*[int(v) for v in seg_off])
        L.check(self.lib.ocmp_mlp_fit_scales(self.h, C.c_void_p(_ptr(snap)), snap.shape[0], snap.stride(0), k, off),
                'mlp_fit_scales')

    # ------------------------------------------------------------ kernels
    def forward(self, X: torch.Tensor) -> torch.Tensor:
        X = self._x(X)
        Y = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        L.check(self.lib.ocmp_mlp_forward(self.h, X.shape[0], C.c_void_p(_ptr(X)), C.c_void_p(_ptr(Y))),
                'mlp_forward')
        return Y

    def grad(self, X: torch.Tensor, T: torch.Tensor) -> Tuple[torch.Tensor, float]:
        X, T = self._x(X), self._t(T, X.shape[0])
        g = torch.empty(NPARAM, dtype=torch.float32, device=X.device)
        loss = C.c_double()
        L.check(self.lib.ocmp_mlp_grad(self.h, X.shape[0], C.c_void_p(_ptr(X)), C.c_void_p(_ptr(T)),
                                       C.c_void_p(_ptr(g)), C.byref(loss)), 'mlp_grad')
        return g, loss.value

    def train_step(self, X: torch.Tensor, T: torch.Tensor, lr: float, want_loss: bool = False) -> Optional[float]:
        X, T = self._x(X), self._t(T, X.shape[0])
        loss = C.c_double()
        L.check(self.lib.ocmp_mlp_train_step(self.h, X.shape[0], C.c_void_p(_ptr(X)), C.c_void_p(_ptr(T)),
                                             C.c_float(lr), C.byref(loss) if want_loss else None),
                'mlp_train_step')
        return loss.value if want_loss else None

    def sample(self, snap: torch.Tensor, n: int, seed: int, batch: int,
               X: Optional[torch.Tensor] = None, T: Optional[torch.Tensor] = None):
        snap = self._snap(snap)
        if X is None:
            X = torch.empty((batch, 4), dtype=torch.float32, device=snap.device)
        if T is None:
            T = torch.empty(batch, dtype=torch.float32, device=snap.device)
        L.check(self.lib.ocmp_mlp_sample(self.h, C.c_void_p(_ptr(snap)), snap.shape[0], snap.stride(0), int(n),
                                         C.c_uint64(seed & (2 ** 64 - 1)), int(batch), C.c_void_p(_ptr(X)),
                                         C.c_void_p(_ptr(T))), 'mlp_sample')
        return X, T

    def guess(self, hist: Sequence[torch.Tensor], out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """hist = [u^n, u^{n-1}, u^{n-2}, u^{n-3}] (device f64); returns u^n + s*f(features)."""
        if len(hist) != 4:
            raise ValueError('the guess needs the last four states')
        n = hist[0].numel()
        for h in hist:
            if h.dtype != torch.float64 or not h.is_cuda or h.numel() != n or not h.is_contiguous():
                raise ValueError('history vectors must be contiguous float64 device tensors of one length')
        if out is None:
            out = torch.empty_like(hist[0])
        L.check(self.lib.ocmp_mlp_guess(self.h, n, *[C.c_void_p(_ptr(h)) for h in hist], C.c_void_p(_ptr(out))),
                'mlp_guess')
        return out

    # ------------------------------------------------------------ training loop
    def train(self, snap: torch.Tensor, seg_off: Sequence[int], steps: int = 2000, batch: int = 65536,
              lr: float = 2e-3, lr_min: float = 1e-4, seed: int = 0, log_every: int = 100) -> List[float]:
        """Fit scales on the snapshots, then `steps` Adam steps on fresh random batches
        with a cosine learning-rate decay; returns the sampled loss history."""
        snap = self._snap(snap)
        if batch > self.max_batch:
            raise ValueError(f'batch {batch} exceeds max_batch {self.max_batch}')
        self.fit_scales(snap, seg_off)
        n = int(seg_off[-1])
        X = torch.empty((batch, 4), dtype=torch.float32, device=snap.device)
        T = torch.empty(batch, dtype=torch.float32, device=snap.device)
        out = []
        for it in range(steps):
            self.sample(snap, n, seed * 1000003 + it, batch, X, T)
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(1, steps)))
            want = (it % log_every == 0) or it == steps - 1
            loss = self.train_step(X, T, a, want_loss=want)
            if want:
                out.append(loss)
        self.losses.extend(out)
        return out

    # ------------------------------------------------------------ helpers
    @staticmethod
    def _snap(snap: torch.Tensor) -> torch.Tensor:
        if snap.dim() != 2 or snap.dtype != torch.float64 or not snap.is_cuda or snap.stride(1) != 1:
            raise ValueError('snapshots must be a (nsnap, n) float64 device tensor with unit row stride')
        return snap

    @staticmethod
    def _x(X: torch.Tensor) -> torch.Tensor:
        if X.dim() != 2 or X.shape[1] != 4 or X.dtype != torch.float32 or not X.is_cuda:
            raise ValueError('X must be an (n, 4) float32 device tensor')
        return X.contiguous()

    @staticmethod
    def _t(T: torch.Tensor, n: int) -> torch.Tensor:
        if T.numel() != n or T.dtype != torch.float32 or not T.is_cuda:
            raise ValueError('T must be an (n,) float32 device tensor')
        return T.contiguous()


def residual_norms(a, b: torch.Tensor, x: torch.Tensor) -> Tuple[float, float]:
    """(||b - A x||, ||b||) over the free dofs, FP64 (ocmp_residual_norms)."""
    out = (C.c_double * 2)()
    L.check(a.ctx.lib.ocmp_residual_norms(a.h, C.c_void_p(_ptr(b)), C.c_void_p(_ptr(x)), out), 'residual_norms')
    return out[0], out[1]


def choose_guess(a, b: torch.Tensor, plain: torch.Tensor, learned: torch.Tensor) -> Tuple[str, float, float]:
    """Fallback rule: keep the learned guess only if its FP64 residual is lower
    than the plain guess's. Copies the winner into `plain` (the solve's start
    vector) and returns (choice, plain residual, learned residual)."""
    rp, _ = residual_norms(a, b, plain)
    rl, _ = residual_norms(a, b, learned)
    if math.isfinite(rl) and rl < rp:
        plain.copy_(learned)
        return 'learned', rp, rl
    return 'plain', rp, rl


def choose_best(a, b: torch.Tensor, plain: torch.Tensor, candidates) -> Tuple[str, float, float]:
    """Fallback rule over several candidates: keep the one with the lowest FP64 residual
    if it beats the plain guess. Copies the winner into `plain` and returns (choice,
    plain residual, residual of the best candidate). The residuals are enqueued back to back
    into device memory (ocmp_residual_norms_noalloc) and read with one synchronisation."""
    xs = [plain] + [v for _, v in candidates]
    out = torch.empty(2 * len(xs), dtype=torch.float64, device=b.device)
    for k, x in enumerate(xs):
        L.check(a.ctx.lib.ocmp_residual_norms_noalloc(a.h, C.c_void_p(_ptr(b)), C.c_void_p(_ptr(x)),
                                                      C.c_void_p(out.data_ptr() + 16 * k)), 'residual_norms_noalloc')
    res = out.cpu().tolist()
    rp = res[0]
    best, rb, vb = 'plain', math.inf, None
    for k, (name, v) in enumerate(candidates):
        r = res[2 * (k + 1)]
        if math.isfinite(r) and r < rb:
            best, rb, vb = name, r, v
    if vb is not None and rb < rp:
        plain.copy_(vb)
        return best, rp, rb
    return 'plain', rp, rb


def choose_best_async(a, b: torch.Tensor, plain: torch.Tensor, candidates):
    """choose_best without a synchronisation: the FP64 residuals, the choice and the copy of the
    winner into `plain` are enqueued on the device, and the returned function reads the outcome
    (choice, plain residual, residual of the best candidate) once that work has run. The rule is
    choose_best's: the first candidate with the lowest finite residual, kept only if that residual
    is below the plain guess's; a NaN plain residual keeps the plain guess."""
    xs = [plain] + [v for _, v in candidates]
    out = torch.empty(2 * len(xs), dtype=torch.float64, device=b.device)
    for k, x in enumerate(xs):
        L.check(a.ctx.lib.ocmp_residual_norms_noalloc(a.h, C.c_void_p(_ptr(b)), C.c_void_p(_ptr(x)),
                                                      C.c_void_p(out.data_ptr() + 16 * k)), 'residual_norms_noalloc')
    r = out[0::2]
    key = torch.where(torch.isfinite(r), r, torch.full_like(r, math.inf))
    key[0] = torch.where(torch.isnan(r[0]), torch.full_like(r[0], -math.inf), key[0])
    idx = torch.argmin(key)  # the first minimum: plain on a tie, the earlier candidate among equals
    for k, (_, v) in enumerate(candidates, 1):
        torch.where(idx == k, v, plain, out=plain)
    names = ['plain'] + [name for name, _ in candidates]

    def outcome() -> Tuple[str, float, float]:
        res = out.cpu().tolist()
        rb = min((res[2 * k] for k in range(1, len(xs)) if math.isfinite(res[2 * k])), default=math.inf)
        return names[int(idx.item())], res[0], rb
    return outcome


class GuessPolicy:
    """The last four accepted states of a transient run (device ring), optional
    snapshot recording for training, and the initial-guess candidates of the next
    step: the trained network's prediction and, optionally, the linear extrapolation."""

    def __init__(self, ctx: Context, n: int, net: Optional[MLPGuess] = None, max_snapshots: int = 0,
                 n_learned: Optional[int] = None, extrapolate: bool = False):
        """n_learned: the network (and the extrapolation) cover dofs [0, n_learned) (e.g. the
        velocity blocks of a saddle-point model); the rest keep the plain guess, the newest
        state. extrapolate: also offer 2 u^n - u^{n-1} on those dofs."""
        self.ctx = ctx
        self.n = int(n)
        self.n_learned = self.n if n_learned is None else int(n_learned)
        self.net = net
        self.extrapolate = bool(extrapolate)
        self.ring = [ctx.zeros(self.n) for _ in range(4)]
        self.count = 0
        self.snap = (torch.empty((max_snapshots, self.n), dtype=torch.float64, device=ctx.torch_device)
                     if max_snapshots else None)
        self.nsnap = 0
        self.out = ctx.empty(self.n)
        self.out_x = ctx.empty(self.n) if self.extrapolate else None

    def push(self, u: torch.Tensor) -> None:
        self.ring[self.count % 4].copy_(u)
        self.count += 1
        if self.snap is not None and self.nsnap < self.snap.shape[0]:
            self.snap[self.nsnap].copy_(u)
            self.nsnap += 1

    def history(self) -> List[torch.Tensor]:
        """[u^n, u^{n-1}, u^{n-2}, u^{n-3}]"""
        return [self.ring[(self.count - 1 - k) % 4] for k in range(4)]

    def ready(self) -> bool:
        return (self.net is not None and self.count >= 4) or (self.extrapolate and self.count >= 2)

    def propose(self) -> torch.Tensor:
        """The network's guess."""
        k = self.n_learned
        h = self.history()
        self.net.guess([v[:k] for v in h], out=self.out[:k])
        if k < self.n:
            self.out[k:].copy_(h[0][k:])
        return self.out

    def candidates(self) -> List[Tuple[str, torch.Tensor]]:
        """(name, vector) of every guess the policy can offer now."""
        out = []
        if self.extrapolate and self.count >= 2:
            k = self.n_learned
            h = self.history()
            torch.add(h[0][:k], h[0][:k] - h[1][:k], out=self.out_x[:k])
            if k < self.n:
                self.out_x[k:].copy_(h[0][k:])
            out.append(('extrapolated', self.out_x))
        if self.net is not None and self.count >= 4:
            out.append(('learned', self.propose()))
        return out

    def snapshots(self) -> torch.Tensor:
        return self.snap[:self.nsnap]


class UNetPrecond:
    """Learned U-net correction of a point-Jacobi / Schur-diagonal preconditioner
    (ocmp_unet_* in include/ocmp.h), used by the 'FGMRes' solver once attached.

    The network lives on the vertex grid of an N^3 x 6-tet structured mesh; its
    four channels are the vertex dofs of four scalar blocks."""

    NPARAM = L.UNET_NPARAM

    def __init__(self, N: int, ctx: Optional[Context] = None, params: Optional[np.ndarray] = None, seed: int = 0):
        self.ctx = ctx or Context.default()
        self.lib = self.ctx.lib
        self.N = int(N)
        p = self.init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.NPARAM:
            raise ValueError(f'expected {self.NPARAM} parameters, got {p.size}')
        h = C.c_void_p()
        L.check(self.lib.ocmp_unet_create(self.ctx.h, self.N, p.ctypes.data_as(C.c_void_p), C.byref(h)),
                'unet_create')
        self.h = h
        self.losses: List[float] = []

    LAYERS = [(4, 8, 'c'), (8, 16, 'c'), (16, 16, 'c'), (16, 32, 'c'), (32, 32, 'c'), (32, 16, 't'), (16, 8, 't'),
              (8, 4, 'c')]

    @classmethod
    def init_params(cls, seed: int = 0) -> np.ndarray:
        """He-normal 3x3x3 kernels (fan-in / 8 for the stride-2 transposed ones), zero
        biases, the output layer scaled down so the untrained correction is small."""
        rng = np.random.default_rng(seed)
        parts = []
        for k, (ci, co, kind) in enumerate(cls.LAYERS):
            fan = ci * 27 if kind == 'c' else ci * 27 / 8.0
            w = rng.standard_normal(ci * co * 27) * math.sqrt(2.0 / fan)
            if k == len(cls.LAYERS) - 1:
                w *= 0.1
            parts += [w, np.zeros(co)]
        return np.concatenate(parts).astype(np.float32)

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                self.lib.ocmp_unet_destroy(h)
            except Exception:
                pass

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.NPARAM, dtype=np.float32)
        L.check(self.lib.ocmp_unet_get_params(self.h, out.ctypes.data_as(C.c_void_p)), 'unet_get_params')
        return out

    @params.setter
    def params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        L.check(self.lib.ocmp_unet_set_params(self.h, p.ctypes.data_as(C.c_void_p)), 'unet_set_params')

    def set_channels(self, dof_off: Sequence[int], in_scale: Sequence[float], out_scale: Sequence[float]) -> None:
        L.check(self.lib.ocmp_unet_set_channels(self.h, (C.c_int64 * 4)(*[int(v) for v in dof_off]),
                                                (C.c_double * 4)(*in_scale), (C.c_double * 4)(*out_scale)),
                'unet_set_channels')

    def channels(self):
        off = (C.c_int64 * 4)()
        a = (C.c_double * 4)()
        b = (C.c_double * 4)()
        L.check(self.lib.ocmp_unet_get_channels(self.h, off, a, b), 'unet_get_channels')
        return list(off), list(a), list(b)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x: (4, (N+1)^3) float32 device grid -> y of the same shape."""
        n3 = (self.N + 1) ** 3
        if x.shape != (4, n3) or x.dtype != torch.float32 or not x.is_cuda:
            raise ValueError(f'x must be a (4, {n3}) float32 device tensor')
        x = x.contiguous()
        y = torch.empty_like(x)
        L.check(self.lib.ocmp_unet_forward(self.h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y))), 'unet_forward')
        return y

    def grad(self, x: torch.Tensor, gy: torch.Tensor) -> torch.Tensor:
        x, gy = x.contiguous(), gy.contiguous()
        g = torch.empty(self.NPARAM, dtype=torch.float32, device=x.device)
        L.check(self.lib.ocmp_unet_grad(self.h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(gy)), C.c_void_p(_ptr(g))),
                'unet_grad')
        return g

    def activation(self, layer: int) -> torch.Tensor:
        """Activation of the last forward pass (-1: input, 0..6: a0..a6, 7: output), float32."""
        cnt = C.c_int64()
        L.check(self.lib.ocmp_unet_activation(self.h, int(layer), None, C.byref(cnt)), 'unet_activation')
        out = torch.empty(cnt.value, dtype=torch.float32, device=self.ctx.torch_device)
        L.check(self.lib.ocmp_unet_activation(self.h, int(layer), C.c_void_p(_ptr(out)), C.byref(cnt)),
                'unet_activation')
        return out

    def fit_channels(self, pre, samples: torch.Tensor, dof_off: Sequence[int]) -> None:
        L.check(self.lib.ocmp_unet_fit_channels(self.h, pre.h, C.c_void_p(_ptr(samples)), samples.shape[0],
                                                samples.stride(0), (C.c_int64 * 4)(*[int(v) for v in dof_off])),
                'unet_fit_channels')

    def train(self, a, pre, samples: torch.Tensor, dof_off: Sequence[int], steps: int = 200, batch: int = 4,
              lr: float = 1e-3, seed: int = 0) -> List[float]:
        """Fit the channel scales, then `steps` Adam steps of the self-supervised loss on
        random batches of the sample rows; returns the loss per step."""
        if samples.dim() != 2 or samples.dtype != torch.float64 or samples.stride(1) != 1:
            raise ValueError('samples must be a (ns, n) float64 device tensor')
        self.fit_channels(pre, samples, dof_off)
        rng = np.random.default_rng(seed)
        ns = samples.shape[0]
        out = []
        for _ in range(steps):
            idx = (C.c_int32 * batch)(*[int(v) for v in rng.integers(0, ns, batch)])
            loss = C.c_double()
            L.check(self.lib.ocmp_unet_train_step(self.h, a.h, pre.h, C.c_void_p(_ptr(samples)), ns,
                                                  samples.stride(0), idx, batch, C.c_float(lr), C.byref(loss)),
                    'unet_train_step')
            out.append(loss.value)
        self.losses.extend(out)
        return out

    def attach(self, pre) -> None:
        L.check(self.lib.ocmp_precond_attach_unet(pre.h, self.h), 'precond_attach_unet')
        pre.unet = self

    @staticmethod
    def detach(pre) -> None:
        L.check(pre.a.ctx.lib.ocmp_precond_attach_unet(pre.h, None), 'precond_attach_unet')
        pre.unet = None


def residual_samples(a, pre, b: torch.Tensor, n_samples: int, every: int, x0: Optional[torch.Tensor] = None,
                     solver: str = 'GMRes', restart: int = 50, precision: str = 'fp64') -> torch.Tensor:
    """Normalised residuals r_k / ||r_k|| of a plain solve from x0 (default 0; pass the
    Dirichlet lift for a problem driven by boundary data), taken every `every`
    iterations: the vectors a preconditioner meets during that solve (training data)."""
    from .core import solve as _solve
    n = b.numel()
    out = torch.zeros((n_samples, n), dtype=torch.float64, device=b.device)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    r = torch.empty_like(b)
    rn = C.c_double()
    for k in range(n_samples):
        if k > 0:
            _solve(a, b, x, pre, solver, tol=1e-300, maxit=every, restart=restart, precision=precision)
        L.check(a.ctx.lib.ocmp_residual(a.h, C.c_void_p(_ptr(b)), C.c_void_p(_ptr(x)), C.c_void_p(_ptr(r)),
                                        C.byref(rn)), 'residual')
        L.check(a.ctx.lib.ocmp_axpby(a.ctx.h, n, 1.0 / rn.value if rn.value > 0 else 1.0, C.c_void_p(_ptr(r)), 0.0,
                                     C.c_void_p(_ptr(out[k]))), 'axpby')
    return out
