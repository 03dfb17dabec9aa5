"""Host-side Ritz values and Newton-basis shifts of the s-step GMRES (no device calls): the
Hessenberg eigenvalues against numpy's LAPACK eigenvalues, the Leja order on a worked example,
and the three-term basis step against the complex product it stands for."""
import ctypes as C

import numpy as np
import pytest

from opencmp_amd import _lib as L


def eigvals(H):
    n = H.shape[0]
    Hc = np.asfortranarray(H, dtype=np.float64)
    wr = np.zeros(n)
    wi = np.zeros(n)
    dp = C.POINTER(C.c_double)
    L.check(L.load().ocmp_hessenberg_eigvals(n, Hc.ctypes.data_as(dp), n, wr.ctypes.data_as(dp),
                                               wi.ctypes.data_as(dp)), 'eigvals')
    return wr + 1j * wi


def shifts(z, s):
    z = np.asarray(z, dtype=np.complex128)
    wr = np.ascontiguousarray(z.real)
    wi = np.ascontiguousarray(z.imag)
    d = np.zeros(s)
    c = np.zeros(s)
    dp = C.POINTER(C.c_double)
    L.check(L.load().ocmp_newton_shifts(len(z), wr.ctypes.data_as(dp), wi.ctypes.data_as(dp), s,
                                          d.ctypes.data_as(dp), c.ctypes.data_as(dp)), 'shifts')
    return d, c


def match(a, b):
    """Largest distance from each of a to its nearest unused b (a permutation match)."""
    b = list(b)
    worst = 0.0
    for x in a:
        k = int(np.argmin([abs(x - y) for y in b]))
        worst = max(worst, abs(x - b[k]))
        b.pop(k)
    return worst


@pytest.mark.parametrize('n', [1, 2, 3, 5, 10, 30, 50])
def test_random_hessenberg_matches_lapack(n):
    rng = np.random.default_rng(n)
    H = np.triu(rng.standard_normal((n, n)), -1)
    z = eigvals(H)
    ref = np.linalg.eigvals(H)
    assert match(z, ref) <= 1e-9 * max(1.0, np.abs(ref).max())
    # conjugate pairs come out as pairs
    assert match(z, np.conj(z)) <= 1e-9 * max(1.0, np.abs(ref).max())


def test_companion_matrix_roots():
    """The companion matrix (upper Hessenberg) of a polynomial with roots on a circle and on the
    real axis has exactly those roots as eigenvalues."""
    roots = np.concatenate([np.exp(2j * np.pi * np.arange(1, 6) / 11) * 1.5, [0.5, -0.7, 2.0]])
    roots = np.concatenate([roots, np.conj(roots[:5])])
    coef = np.real(np.poly(roots))
    n = len(roots)
    H = np.zeros((n, n))
    H[0, :] = -coef[1:] / coef[0]
    H[np.arange(1, n), np.arange(n - 1)] = 1.0
    assert match(eigvals(H), roots) <= 1e-8


def test_reducible_hessenberg():
    """Zero subdiagonal entries split the matrix; each diagonal block keeps its eigenvalues."""
    A = np.array([[2.0, 1.0], [-5.0, 0.0]])   # eigenvalues 1 +- 2i
    B = np.array([[3.0]])
    Cb = np.array([[0.0, -4.0], [1.0, 0.0]])  # +- 2i
    H = np.zeros((5, 5))
    H[:2, :2] = A
    H[2:3, 2:3] = B
    H[3:, 3:] = Cb
    H[0, 2:] = [7.0, -1.0, 2.0]
    ref = np.concatenate([np.linalg.eigvals(A), [3.0], np.linalg.eigvals(Cb)])
    assert match(eigvals(H), ref) <= 1e-12


def test_arnoldi_hessenberg_carries_the_spectrum():
    """A full Arnoldi run on a nonsymmetric matrix gives H = Q^T M Q: same eigenvalues as M."""
    rng = np.random.default_rng(7)
    n = 24
    M = rng.standard_normal((n, n)) / np.sqrt(n) + np.eye(n)
    Q = np.zeros((n, n + 1))
    H = np.zeros((n + 1, n))
    q = rng.standard_normal(n)
    Q[:, 0] = q / np.linalg.norm(q)
    for j in range(n):
        w = M @ Q[:, j]
        for _ in range(2):
            h = Q[:, :j + 1].T @ w
            w = w - Q[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        if j + 1 < n:
            Q[:, j + 1] = w / H[j + 1, j]
    assert match(eigvals(H[:n, :n]), np.linalg.eigvals(M)) <= 1e-9


def test_leja_order_worked_example():
    # moduli 3, 2, sqrt(5): 3 first; then -2 (distance 5 against 2*sqrt(2) for 1 +- 2i); then the
    # pair, as (Re, 0) followed by (Re, Im^2)
    d, c = shifts([1 + 2j, 3, -2, 1 - 2j], 4)
    assert np.allclose(d, [3, -2, 1, 1]) and np.allclose(c, [0, 0, 0, 4])
    # a pair that does not fit enters by its real part
    d, c = shifts([1 + 2j, 3, -2, 1 - 2j], 3)
    assert np.allclose(d, [3, -2, 1]) and np.allclose(c, [0, 0, 0])
    # largest modulus complex first: pair first
    d, c = shifts([0.5, 2 + 3j, 2 - 3j], 3)
    assert np.allclose(d, [2, 2, 0.5]) and np.allclose(c, [0, 9, 0])
    # fewer Ritz values than shifts: the last shift repeats
    d, c = shifts([2.0], 3)
    assert np.allclose(d, [2, 2, 2]) and np.allclose(c, 0)


def test_leja_points_maximise_distance_products():
    rng = np.random.default_rng(3)
    z = rng.standard_normal(40)   # real Ritz values
    d, c = shifts(z + 0j, 5)
    assert np.all(c == 0)
    assert abs(d[0]) == np.abs(z).max()
    for t in range(1, 5):
        score = lambda x: np.sum(np.log(np.abs(x - d[:t])))
        rest = [x for x in z if not np.any(np.isclose(x, d[:t]))]
        assert score(d[t]) >= max(score(x) for x in rest) - 1e-12


def test_three_term_step_is_the_complex_product():
    """v_{t+1} = (M - d_t) v_t + c_t v_{t-1} with the shifts of a conjugate pair gives
    (M - z)(M - conj z) v, so the block spans the Newton polynomial basis."""
    rng = np.random.default_rng(11)
    n = 8
    M = rng.standard_normal((n, n))
    z = np.linalg.eigvals(M)
    d, c = shifts(z, 5)
    v = rng.standard_normal(n)
    vs = [v]
    for t in range(5):
        w = M @ vs[-1] - d[t] * vs[-1]
        if c[t] != 0.0:
            w = w + c[t] * vs[-2]
        vs.append(w)
    # rebuild the (complex) shifts the real steps stand for
    zs = []
    t = 0
    while t < 5:
        if t + 1 < 5 and c[t + 1] != 0.0:
            b = np.sqrt(c[t + 1])
            zs += [d[t] + 1j * b, d[t] - 1j * b]
            t += 2
        else:
            zs.append(d[t])
            t += 1
    ref = v.astype(np.complex128)
    for zz in zs:
        ref = M @ ref - zz * ref
    assert np.allclose(vs[-1], ref.real, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    assert np.abs(ref.imag).max() <= 1e-10 * np.abs(ref).max()
    # and every shift is (the real part of) an eigenvalue of M
    for t in range(5):
        assert np.min(np.abs(z.real - d[t])) <= 1e-12


def test_bad_arguments_raise():
    dp = C.POINTER(C.c_double)
    x = np.zeros(4)
    rc = L.load().ocmp_hessenberg_eigvals(2, x.ctypes.data_as(dp), 1, x.ctypes.data_as(dp), x.ctypes.data_as(dp))
    assert rc == L.OCMP_ERR_ARG
