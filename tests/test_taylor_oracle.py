"""CPU pins of the Taylor-map generator oracle: its term bookkeeping and
sigmoid derivatives against sympy's symbolic differentiation, its zero-spiral
maps against the separable closed form, the direction of the spiral shift,
and a planted sign slip in the chain rule."""
import math

import numpy as np
import pytest
from numpy.polynomial import Polynomial

from oracle import taylor_gen as T

sp = pytest.importorskip("sympy")


def _sym_terms_value(terms, f, r, th):
    tot = 0
    for (a, b, p), c in terms.items():
        tot += c * r ** (-p) * sp.diff(f, r, a, th, b) if (a or b) else c * r ** (-p) * f
    return tot


def test_laplacian_terms_match_sympy():
    r, th = sp.symbols("r theta", positive=True)
    f = r ** sp.Rational(38, 5) * sp.cos(3 * th) + sp.exp(r / 3) * sp.sin(th) ** 2
    lap = f
    for n in range(1, 4):
        lap = sp.diff(lap, r, 2) + sp.diff(lap, r) / r + sp.diff(lap, th, 2) / r ** 2
        got = _sym_terms_value(T.laplacian_terms(n), f, r, th)
        pt = {r: sp.Rational(43, 5), th: sp.Rational(3, 10)}
        assert abs(float((got - lap).subs(pt).evalf(30))) < 1e-9 * max(1.0, abs(float(lap.subs(pt).evalf(30))))


def test_chain_terms_match_sympy():
    r = sp.symbols("r", positive=True)
    u = sp.symbols("u")
    G = sp.Function("G")
    s = sp.Rational(1, 7) * r ** 3 - sp.Rational(2, 3) * r ** 2 + sp.Rational(1, 5) * r + 1
    for order in range(1, 6):
        want = sp.diff(G(u).subs(u, -s), r, order)
        got = 0
        for (n, m), c in T.chain_terms(order).items():
            pr = 1
            for j, e in enumerate(m):
                pr *= sp.diff(s, r, j) ** e if j else s ** e
            got += int(c) * (sp.diff(G(u), u, n) if n else G(u)).subs(u, -s) * pr
        diff = sp.simplify(sp.expand(want.doit() - got.doit()))
        assert diff == 0, (order, diff)


def test_sigmoid_and_shape_derivatives_match_sympy():
    x = sp.symbols("x")
    x0, x1, w0, h = 0.2, 0.45, 0.06, -0.728
    k = 16 / w0
    shape = h / (1 + sp.exp(-k * (x - x0))) / (1 + sp.exp(k * (x - x1)))
    pts = np.array([0.15, 0.2, 0.23, 0.33, 0.44, 0.5])
    for order in range(0, 7):
        fx = sp.lambdify(x, sp.diff(shape, x, order), "mpmath")
        want = np.array([float(fx(p)) for p in pts])
        got = T.shape_derivative(pts, x0, x1, w0, h, order)
        scale = np.max(np.abs(want)) + 1e-300
        assert np.max(np.abs(got - want)) / scale < 1e-9, order        # sigma-polynomial rounding grows with k^order


def _poles():
    return [(np.deg2rad(7.5 / 30 * 30), np.deg2rad(16.875), np.deg2rad(9.375 * 0.2), 1.0),
            (np.deg2rad(19.6875), np.deg2rad(22.5), np.deg2rad(2.8125 * 0.35), -0.728)]


def test_zero_spiral_maps_match_separable_closed_form():
    """For s = 0, Lap^n (a r^k G(theta)) = a sum_i g_ni r^(k-2n) G^(2i) with
    g from (k - 2j)^2 acting on r powers: checked against a direct
    recursion on (r power, theta derivative) monomials."""
    a, k = 0.35, 7.6
    r_axis = np.linspace(8.0, 9.3, 7)
    theta = np.linspace(0, np.pi / 6, 181)
    zero = Polynomial([0.0])
    bz, br, bf = T.generate_maps(3, r_axis, theta, _poles(), zero, (a, k))
    # closed form: Lap (r^m G^(j)) = m^2 r^(m-2) G^(j) + r^(m-2) G^(j+2)
    for n in range(4):
        poly = {(k, 0): 1.0}
        for _ in range(n):
            nxt = {}
            for (m, j), c in poly.items():
                nxt[(m - 2, j)] = nxt.get((m - 2, j), 0.0) + c * m * m
                nxt[(m - 2, j + 2)] = nxt.get((m - 2, j + 2), 0.0) + c
            poly = nxt
        R = r_axis[:, None]
        lap = sum(c * a * R ** m * T.flutter_derivative(theta, _poles(), j)[None, :] for (m, j), c in poly.items())
        dlr = sum(c * a * m * R ** (m - 1) * T.flutter_derivative(theta, _poles(), j)[None, :]
                  for (m, j), c in poly.items())
        dlt = sum(c * a * R ** m * T.flutter_derivative(theta, _poles(), j + 1)[None, :] for (m, j), c in poly.items())
        c2n, c2n1 = (-1) ** n / math.factorial(2 * n), (-1) ** n / math.factorial(2 * n + 1)
        for got, want in ((bz[n], c2n * lap), (br[n], c2n1 * dlr), (bf[n], c2n1 * dlt / R)):
            assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), n


def test_exact_shift_maps_match_sympy_on_a_spiral():
    """use_roll=False evaluates G at theta - s(r) exactly: compared with
    sympy's derivatives of A(r) G(theta - s(r)).  At each point G is replaced
    by its degree-5 Taylor polynomial about u0 = theta - s(r) (its
    derivatives up to order 5 at u0 are exact), which keeps the symbolic
    expressions small."""
    r, th, u = sp.symbols("r theta u")
    a, k = 0.35, 7.6
    sc = Polynomial([0.1, -0.03, 0.004])             # s(r) in rad
    s_sym = sum(sp.Float(c, 30) * r ** i for i, c in enumerate(sc.coef))
    x0, x1, w0, h = _poles()[1]
    kk = 16 / w0
    G = h / (1 + sp.exp(-kk * (u - x0))) / (1 + sp.exp(kk * (u - x1)))
    gd = [sp.lambdify(u, sp.diff(G, u, q), "mpmath") for q in range(6)]
    r_axis = np.array([8.3, 8.9])
    theta = np.linspace(0.3, 0.45, 3)
    bz, br, bf = T.generate_maps(2, r_axis, theta, [_poles()[1]], sc, (a, k), use_roll=False)
    for ir, rv in enumerate(r_axis):
        for jt, tv in ((0, theta[0]), (2, theta[2])) if ir == 0 else ((1, theta[1]),):
            u0 = float(tv - sc(rv))
            w = th - s_sym - sp.Float(u0, 30)
            P = sum(sp.Float(float(gd[q](u0)), 30) * w ** q / math.factorial(q) for q in range(6))
            lap = a * r ** sp.Float(k, 30) * P
            for n in range(3):
                c2n, c2n1 = (-1) ** n / math.factorial(2 * n), (-1) ** n / math.factorial(2 * n + 1)
                exprs = (c2n * lap, c2n1 * sp.diff(lap, r), c2n1 * sp.diff(lap, th) / r)
                for got, e in zip((bz[n], br[n], bf[n]), exprs):
                    want = float(e.subs({r: sp.Float(rv, 30), th: sp.Float(tv, 30)}).evalf(25))
                    scale = np.max(np.abs(got))
                    assert abs(got[ir, jt] - want) <= 1e-9 * scale, (n, ir, jt, got[ir, jt], want)
                lap = sp.diff(lap, r, 2) + sp.diff(lap, r) / r + sp.diff(lap, th, 2) / r ** 2


def test_positive_spiral_moves_the_pattern_to_larger_azimuth():
    theta = np.linspace(0, np.pi / 6, 601)
    step = theta[1] - theta[0]
    r_axis = np.array([8.0, 9.0])
    s = Polynomial([-8.0 * 20 * step, 20 * step])     # s(8) = 0, s(9) = 20 steps
    bz, _, _ = T.generate_maps(0, r_axis, theta, _poles(), s, (0.35, 7.6))
    shift = int(np.argmax(bz[0, 1])) - int(np.argmax(bz[0, 0]))
    assert shift in (19, 20)                          # int() of 20 steps in floating point


def test_planted_chain_rule_sign_slip_is_caught():
    """With du/dr = +s' instead of -s' the odd r-derivatives flip sign: the
    sympy comparison of the order-1 chain terms fails."""
    good = T.chain_terms(1)
    bad = {(n, m): -c if n == 1 else c for (n, m), c in good.items()}
    assert good != bad and good[(1, (0, 1))] == -1.0
