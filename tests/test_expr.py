"""Config-expression parsing and compilation (host side, no GPU).

Expected values are worked out by hand from the reference grammar and helper
functions (opencmp/config_functions/parse_arithmetic.py:37-268,
opencmp/helpers/math.py:23-131), including the two places where the config
language differs from Python: a leading minus binds tighter than ``^`` and
``^`` groups to the right.
"""
import math

import numpy as np
import pytest

from opencmp_amd import expr as E


def run_program(code, x=0.0, y=0.0, z=0.0, t=0.0):
    """A plain postfix interpreter of the device program format (include/ocmp.h)."""
    st, pc = [], 0
    while True:
        op = int(code[pc]); pc += 1
        if op == E.OP_END:
            assert len(st) == 1
            return st[0]
        if op == E.OP_CONST:
            st.append(code[pc]); pc += 1
        elif op in (E.OP_X, E.OP_Y, E.OP_Z, E.OP_T):
            st.append({E.OP_X: x, E.OP_Y: y, E.OP_Z: z, E.OP_T: t}[op])
        elif E.OP_ADD <= op <= E.OP_POW:
            b, a = st.pop(), st.pop()
            if op == E.OP_ADD:
                st.append(a + b)
            elif op == E.OP_SUB:
                st.append(a - b)
            elif op == E.OP_MUL:
                st.append(a * b)
            elif op == E.OP_DIV:
                st.append(a / b)
            else:
                st.append(a ** b)
        elif op == E.OP_NEG:
            st.append(-st.pop())
        elif op == E.OP_HS2:
            w, tt = st.pop(), st.pop()
            st.append(1 / (1 + math.exp(-(2 * 9.21025 / w * tt - 9.21025))))
        elif op == E.OP_RAMP4:
            tr, p2, p1, tt = st.pop(), st.pop(), st.pop(), st.pop()
            st.append(0.5 * (p1 - p2) * math.cos(tt * math.pi / tr) + 0.5 * (p1 + p2) if tr - tt > 0 else p2)
        else:
            a = st.pop()
            f = {E.OP_SIN: math.sin, E.OP_COS: math.cos, E.OP_TAN: math.tan, E.OP_EXP: math.exp,
                 E.OP_SQRT: math.sqrt, E.OP_ABS: abs, E.OP_TRUNC: math.trunc, E.OP_ROUND: round,
                 E.OP_SIG: lambda v: 1 / (1 + math.exp(-v)),
                 E.OP_TANH: lambda v: (math.exp(2 * min(v, 350)) - 1) / (math.exp(2 * min(v, 350)) + 1),
                 E.OP_SGN: lambda v: -1.0 if v < -1e-12 else (1.0 if v > 1e-12 else 0.0),
                 E.OP_HS: lambda v: 1 / (1 + math.exp(-(2 * 9.21025 / 0.1 * v - 9.21025))),
                 E.OP_RAMP: lambda v: 0.5 * (0 - 1) * math.cos(v * math.pi) + 0.5 if 1 - v > 0 else 1.0}[op]
            st.append(float(f(a)))


CASES = [
    # (text, point (x, y, z, t), value worked out by hand)
    ('-4*(y-0.5)*(y+0.5) * ramp(t)', (0.0, 0.25, 0, 0.5), -4 * (-0.25) * 0.75 * 0.5),
    ('-4*(y-0.5)*(y+0.5) * ramp(t)', (0.0, 0.25, 0, 3.0), 0.75),
    ('5.0/(2.0*0.001)*(0.2*y - y^2)', (0.3, 0.1, 0, 0), 2500.0 * 0.01),
    ('5.0*(1.0 - x) + 5.0', (0.25, 0, 0, 0), 8.75),
    ('-0.25*(cos(2*x) + cos(2*y))*exp(-4*t)', (0.0, 0.0, 0, 0.0), -0.5),
    ('-x^2', (3.0, 0, 0, 0), 9.0),            # (-x)^2: the sign belongs to the atom
    ('0-x^2', (3.0, 0, 0, 0), -9.0),
    ('2^3^2', (0, 0, 0, 0), 512.0),           # right associative
    ('2*-3', (0, 0, 0, 0), -6.0),
    ('--2', (0, 0, 0, 0), 2.0),
    ('PI + e - pi - E', (0, 0, 0, 0), 0.0),
    ('1.5e-3*1e3', (0, 0, 0, 0), 1.5),
    # H_s goes from 1 - 0.9999 at t = 0 to 0.9999 at t = w (opencmp/helpers/math.py:89-113); the
    # reference's shift 9.21025 gives those values to about 2e-4 relative
    ('H(0)', (0, 0, 0, 0), 1e-4),
    ('H(0.1)', (0, 0, 0, 0), 0.9999),
    ('H(t, 1.0)', (0, 0, 0, 1.0), 0.9999),
    ('H(0.05)', (0, 0, 0, 0), 0.5),
    ('ramp(t, 2, 6, 4)', (0, 0, 0, 2.0), 4.0),  # halfway up a cosine ramp from 2 to 6 over 4
    ('ramp(t, 2, 6, 4)', (0, 0, 0, 9.0), 6.0),
    ('ramp(t, 3)', (0, 0, 0, 0.0), 3.0),
    ('sig(0) + tanh(0) + tanh(1000)', (0, 0, 0, 0), 1.5),
    ('sgn(-2) + sgn(1e-13) + sgn(5)', (0, 0, 0, 0), 0.0),
    ('trunc(-2.7) + round(2.5) + round(3.5) + abs(-1)', (0, 0, 0, 0), -2 + 2 + 4 + 1),
    ('sqrt(x*x + y*y + z*z)', (2.0, 3.0, 6.0, 0), 7.0),
    ('tan(pi/4)', (0, 0, 0, 0), 1.0),
]


@pytest.mark.parametrize('text,pt,val', CASES)
def test_evaluate_and_program_agree_with_hand_values(text, pt, val):
    e = E.Expr(text)
    x, y, z, t = pt
    rel = 3e-4 if text.startswith('H(') else 1e-12
    assert float(e(x, y, z, t)) == pytest.approx(val, rel=rel, abs=1e-12)
    assert run_program(E.compile_program(e.node), x, y, z, t) == pytest.approx(float(e(x, y, z, t)), rel=1e-14,
                                                                              abs=1e-15)


def test_vectors_and_components():
    comps = E.vector('[-cos(x)*sin(y)*exp(-2*t), sin(x)*cos(y)*exp(-2*t)]')
    assert len(comps) == 2
    x, y, t = 0.3, 1.1, 0.05
    assert comps[0](x, y, 0, t) == pytest.approx(-math.cos(x) * math.sin(y) * math.exp(-2 * t))
    assert comps[1](x, y, 0, t) == pytest.approx(math.sin(x) * math.cos(y) * math.exp(-2 * t))
    assert len(E.vector('x + 1')) == 1
    # numpy arrays broadcast
    xs = np.linspace(0, 1, 5)
    np.testing.assert_allclose(E.Expr('x^2 + t')(xs, 0, 0, 2.0), xs ** 2 + 2.0)


def test_constant_and_time_only_values():
    assert E.Expr('0.001').is_constant() and E.Expr('0.001').value() == 0.001
    e = E.Expr('exp(-2*t)')
    assert not e.depends_on_space() and e.value(t=0.5) == pytest.approx(math.exp(-1.0))
    with pytest.raises(E.ExprError):
        E.Expr('x + 1').value()


@pytest.mark.parametrize('bad', ['x +', 'foo(1)', 'q + 1', '[1, 2] * 3', '-[1, 2]', '(x', 'sin(1, 2)',
                                 'IMPORT(f)', 'x $ y'])
def test_rejects_what_the_gpu_path_cannot_evaluate(bad):
    with pytest.raises(E.ExprError):
        v = E.parse(bad)
        if isinstance(v, tuple):
            raise E.ExprError('vector')


def test_stack_depth_is_checked():
    # a left-leaning sum needs 2 slots, a right-leaning nest of products one slot per level
    E.compile_program(E.parse('+'.join(['x'] * 40)))
    deep = 'x*(' * 7 + 'x' + ')' * 7
    assert run_program(E.compile_program(E.parse(deep)), 2.0) == 2.0 ** 8
    with pytest.raises(E.ExprError):
        E.compile_program(E.parse('x*(' * 8 + 'x' + ')' * 8))
