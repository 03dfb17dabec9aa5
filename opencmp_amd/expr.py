"""Config-file arithmetic expressions for the GPU path.

OpenCMP writes boundary data, initial conditions, sources and reference
solutions as arithmetic strings of x, y, z and t (``-4*(y-0.5)*(y+0.5) *
ramp(t)``, ``[sin(y)*exp(-2*t), 0.0]``), parsed by pyparsing into NGSolve
CoefficientFunctions (opencmp/config_functions/parse_arithmetic.py:37-268).
Here the same language is parsed into a small syntax tree that is

* compiled to the postfix programs of ``OCMP_FN_EXPR`` (include/ocmp.h), which
  the HIP kernels evaluate at their own quadrature and interpolation points, and
* evaluated with numpy on the host, for values that do not depend on space
  (model parameters, time factors).

Grammar and precedence follow parse_arithmetic.py:86-118: ``+ -`` < ``* /`` <
``^`` (right associative); a leading sign belongs to the atom, so ``-x^2`` is
``(-x)^2`` as in the reference. ``pi`` and ``e`` are case-insensitive constants;
``[a, b]`` is a vector. Functions: sin cos tan exp sqrt abs trunc round sgn and
the reference's tanh, sig, H (H_s) and ramp (ramp_cos) from
opencmp/helpers/math.py:23-131.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib as L

# opcodes, as declared in include/ocmp.h
OP_END, OP_CONST, OP_X, OP_Y, OP_Z, OP_T = 0, 1, 2, 3, 4, 5
OP_NEG, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW = 10, 11, 12, 13, 14, 15
OP_SIN, OP_COS, OP_TAN, OP_EXP, OP_SQRT, OP_TANH, OP_SIG, OP_ABS, OP_SGN, OP_TRUNC, OP_ROUND = range(20, 31)
OP_HS, OP_HS2, OP_RAMP, OP_RAMP4 = 40, 41, 50, 53
MAX_DEPTH = 8

_UNARY = {'sin': OP_SIN, 'cos': OP_COS, 'tan': OP_TAN, 'exp': OP_EXP, 'sqrt': OP_SQRT, 'tanh': OP_TANH,
          'sig': OP_SIG, 'abs': OP_ABS, 'sgn': OP_SGN, 'trunc': OP_TRUNC, 'round': OP_ROUND}
_BINARY = {'+': OP_ADD, '-': OP_SUB, '*': OP_MUL, '/': OP_DIV, '^': OP_POW}
_CONSTANTS = {'pi': math.pi, 'e': math.e}


class ExprError(ValueError):
    pass


# ------------------------------------------------------------------ syntax tree

@dataclass(frozen=True)
class Num:
    v: float


@dataclass(frozen=True)
class Var:
    name: str      # x | y | z | t


@dataclass(frozen=True)
class Un:
    op: str        # '-' or a function name
    a: 'Node'


@dataclass(frozen=True)
class Bin:
    op: str
    a: 'Node'
    b: 'Node'


@dataclass(frozen=True)
class Call:
    name: str      # H, ramp (several arities)
    args: Tuple['Node', ...]


Node = Union[Num, Var, Un, Bin, Call]

_TOKEN = re.compile(r'\s*(?:(?P<num>\d+(?:\.\d*)?(?:[eE][+-]?\d+)?)|(?P<name>[A-Za-z_][A-Za-z_]*)|(?P<op>[-+*/^(),\[\]]))')


def _tokenize(s: str) -> List[Tuple[str, str]]:
    out, pos = [], 0
    s = s.strip()
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            raise ExprError(f'cannot parse {s!r} at {s[pos:]!r}')
        pos = m.end()
        kind = m.lastgroup
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, text: str):
        self.text = text
        self.toks = _tokenize(text)
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self, val=None):
        k, v = self.peek()
        if k is None or (val is not None and v != val):
            raise ExprError(f'expected {val!r} in {self.text!r}')
        self.i += 1
        return v

    def parse(self):
        node = self.expr_or_vec()
        if self.i != len(self.toks):
            raise ExprError(f'unexpected {self.toks[self.i][1]!r} in {self.text!r}')
        return node

    def expr_or_vec(self):
        return self.expr()

    def bin(self, op, a, b):
        if isinstance(a, tuple) or isinstance(b, tuple):
            raise ExprError(f'arithmetic on a vector in {self.text!r}')
        return Bin(op, a, b)

    def expr(self):
        node = self.term()
        while self.peek()[1] in ('+', '-'):
            op = self.take()
            node = self.bin(op, node, self.term())
        return node

    def term(self):
        node = self.factor()
        while self.peek()[1] in ('*', '/'):
            op = self.take()
            node = self.bin(op, node, self.factor())
        return node

    def factor(self):
        node = self.atom()
        if self.peek()[1] == '^':
            self.take()
            node = self.bin('^', node, self.factor())     # right associative
        return node

    def atom(self):
        signs = []
        while self.peek()[1] in ('+', '-'):
            signs.append(self.take())
        k, v = self.peek()
        if v == '(':
            self.take()
            node = self.expr()
            self.take(')')
        elif v == '[':
            self.take()
            comps = [self.expr()]
            while self.peek()[1] == ',':
                self.take()
                comps.append(self.expr())
            self.take(']')
            node = ('vec', tuple(comps))
        elif k == 'num':
            self.take()
            node = Num(float(v))
        elif k == 'name':
            self.take()
            if self.peek()[1] == '(':
                self.take()
                args = [self.expr()]
                while self.peek()[1] == ',':
                    self.take()
                    args.append(self.expr())
                self.take(')')
                node = _call(v, args, self.text)
            elif v in ('x', 'y', 'z', 't'):
                node = Var(v)
            elif v.lower() in _CONSTANTS:
                node = Num(_CONSTANTS[v.lower()])
            else:
                raise ExprError(f'unknown name {v!r} in {self.text!r} (model variables and IMPORT are not '
                                'supported on the GPU path)')
        else:
            raise ExprError(f'unexpected {v!r} in {self.text!r}')
        # a leading '-' negates the atom (parse_arithmetic.py:54-63: one 'unary -' per sign)
        for sgn in reversed(signs):
            if sgn == '-':
                if isinstance(node, tuple):
                    raise ExprError('a vector cannot be negated')
                node = Un('-', node)
        return node


def _call(name: str, args: List[Node], text: str) -> Node:
    if name in _UNARY:
        if len(args) != 1:
            raise ExprError(f'{name} takes one argument in {text!r}')
        return Un(name, args[0])
    if name == 'H':
        if len(args) not in (1, 2):
            raise ExprError(f'H takes (t) or (t, w) in {text!r}')
        return Call('H', tuple(args))
    if name == 'ramp':
        if not 1 <= len(args) <= 4:
            raise ExprError(f'ramp takes 1 to 4 arguments in {text!r}')
        defaults = [Num(0.0), Num(1.0), Num(1.0)]     # p1, p2, tr of ramp_cos
        return Call('ramp', tuple(list(args) + defaults[len(args) - 1:]) if len(args) > 1 else tuple(args))
    raise ExprError(f'unknown function {name!r} in {text!r}')


def parse(text: str):
    """A scalar expression (Node) or a vector (tuple of Nodes) from a config string."""
    node = _Parser(str(text)).parse()
    if isinstance(node, tuple):
        return node[1]
    return node


# ------------------------------------------------------------------ evaluation and compilation

def variables(node: Node) -> set:
    if isinstance(node, Num):
        return set()
    if isinstance(node, Var):
        return {node.name}
    if isinstance(node, Un):
        return variables(node.a)
    if isinstance(node, Bin):
        return variables(node.a) | variables(node.b)
    return set().union(*[variables(a) for a in node.args])


def _sig(v):
    return 1.0 / (1.0 + np.exp(-v))


def _tanh(v):
    v = np.where(v - 350.0 > 0.0, 350.0, v)
    e2 = np.exp(2.0 * v)
    return (e2 - 1.0) / (e2 + 1.0)


def _ramp(t, p1, p2, tr):
    return np.where(tr - t > 0.0, 0.5 * (p1 - p2) * np.cos(t * np.pi / tr) + 0.5 * (p1 + p2), p2)


def evaluate(node: Node, x=0.0, y=0.0, z=0.0, t=0.0):
    """numpy evaluation (arrays broadcast)."""
    with np.errstate(all='ignore'):
        return _ev(node, dict(x=x, y=y, z=z, t=t))


def _ev(n, env):
    if isinstance(n, Num):
        return n.v
    if isinstance(n, Var):
        return env[n.name]
    if isinstance(n, Un):
        a = _ev(n.a, env)
        f = {'-': np.negative, 'sin': np.sin, 'cos': np.cos, 'tan': np.tan, 'exp': np.exp, 'sqrt': np.sqrt,
             'tanh': _tanh, 'sig': _sig, 'abs': np.abs, 'trunc': np.trunc, 'round': np.rint,
             'sgn': lambda v: np.where(v < -1e-12, -1.0, np.where(v > 1e-12, 1.0, 0.0))}[n.op]
        return f(a)
    if isinstance(n, Bin):
        a, b = _ev(n.a, env), _ev(n.b, env)
        return {'+': np.add, '-': np.subtract, '*': np.multiply, '/': np.divide, '^': np.power}[n.op](a, b)
    args = [_ev(a, env) for a in n.args]
    if n.name == 'H':
        w = args[1] if len(args) > 1 else 0.1
        return _sig(2.0 * 9.21025 / w * args[0] - 9.21025)
    if len(args) == 1:
        return _ramp(args[0], 0.0, 1.0, 1.0)
    return _ramp(*args)


def compile_program(node: Node) -> List[float]:
    """Postfix program of OCMP_FN_EXPR; raises if it needs more than 8 stack slots."""
    code: List[float] = []
    depth = [0, 0]   # current, max

    def push(k=1):
        depth[0] += k
        depth[1] = max(depth[1], depth[0])

    def emit(n):
        if isinstance(n, Num):
            code.extend([OP_CONST, n.v])
            push()
        elif isinstance(n, Var):
            code.append({'x': OP_X, 'y': OP_Y, 'z': OP_Z, 't': OP_T}[n.name])
            push()
        elif isinstance(n, Un):
            emit(n.a)
            code.append(OP_NEG if n.op == '-' else _UNARY[n.op])
        elif isinstance(n, Bin):
            emit(n.a)
            emit(n.b)
            code.append(_BINARY[n.op])
            depth[0] -= 1
        else:
            for a in n.args:
                emit(a)
            k = len(n.args)
            if n.name == 'H':
                code.append(OP_HS if k == 1 else OP_HS2)
            else:
                code.append(OP_RAMP if k == 1 else OP_RAMP4)
            depth[0] -= k - 1

    emit(node)
    code.append(OP_END)
    if depth[1] > MAX_DEPTH:
        raise ExprError(f'expression needs {depth[1]} stack slots, more than {MAX_DEPTH}')
    return [float(c) for c in code]


# ------------------------------------------------------------------ time parameter and bound expressions

class Parameter:
    """A mutable scalar (ngsolve.Parameter as used for the time, opencmp/solvers/base_solver.py)."""

    def __init__(self, value: float = 0.0):
        self.value = float(value)

    def Set(self, v: float):
        self.value = float(v)

    def Get(self) -> float:
        return self.value


class Expr:
    """A parsed scalar expression; ``bind`` makes it an ``Fn`` of the GPU terms."""

    def __init__(self, text_or_node, text: Optional[str] = None):
        if isinstance(text_or_node, str):
            self.text = text_or_node
            node = parse(text_or_node)
            if isinstance(node, tuple):
                raise ExprError(f'{text_or_node!r} is a vector; use vector()')
            self.node = node
        else:
            self.node = text_or_node
            self.text = text or repr(text_or_node)
        self.vars = variables(self.node)

    def is_constant(self) -> bool:
        return not self.vars

    def depends_on_space(self) -> bool:
        return bool(self.vars & {'x', 'y', 'z'})

    def value(self, t: float = 0.0) -> float:
        if self.depends_on_space():
            raise ExprError(f'{self.text!r} depends on space')
        return float(evaluate(self.node, t=t))

    def __call__(self, x=0.0, y=0.0, z=0.0, t=0.0):
        return evaluate(self.node, x, y, z, t)

    def bind(self, ctx, t_param: Optional[Parameter] = None, scale: float = 1.0, shift: float = 0.0) -> 'BoundExpr':
        return BoundExpr(self, ctx, t_param, scale, shift)

    def __repr__(self):
        return f'Expr({self.text!r})'


class BoundExpr:
    """An expression registered on a context: ``fn_args()`` gives (FN_EXPR, (id, t, scale, shift))
    with the time parameter's current value, so a form assembled at a later time sees that time."""

    def __init__(self, e: Expr, ctx, t_param: Optional[Parameter], scale: float, shift: float):
        self.expr = e
        self.ctx = ctx
        self.t_param = t_param
        self.scale = float(scale)
        self.shift = float(shift)
        self.id = program_id(ctx, compile_program(e.node))

    def fn_args(self) -> Tuple[int, Tuple[float, ...]]:
        t = self.t_param.value if self.t_param is not None else 0.0
        return (L.FN_EXPR, (float(self.id), float(t), self.scale, self.shift))

    def scaled(self, scale: float = 1.0, shift: float = 0.0) -> 'BoundExpr':
        b = BoundExpr.__new__(BoundExpr)
        b.expr, b.ctx, b.t_param, b.id = self.expr, self.ctx, self.t_param, self.id
        b.scale, b.shift = self.scale * scale, self.shift * scale + shift
        return b


def program_id(ctx, code: Sequence[float]) -> int:
    """Register a program once per context (identical programs share an id)."""
    import ctypes
    cache: Dict[tuple, int] = ctx.__dict__.setdefault('_expr_ids', {})
    key = tuple(code)
    if key not in cache:
        arr = (ctypes.c_double * len(code))(*code)
        out = ctypes.c_int32()
        L.check(ctx.lib.ocmp_expr_create(ctx.h, len(code), arr, ctypes.byref(out)), 'expr_create')
        cache[key] = out.value
    return cache[key]


def vector(text: str) -> List[Expr]:
    """The components of a vector expression ``[a, b, ...]`` (a scalar gives one component)."""
    node = parse(text)
    if isinstance(node, tuple):
        return [Expr(n, text) for n in node]
    return [Expr(node, text)]


def resolve_fn(fn) -> Tuple[int, Sequence[float]]:
    """(FN_*, params) of a closed form, or of a bound expression at its time parameter's value."""
    if hasattr(fn, 'fn_args'):
        return fn.fn_args()
    return fn
