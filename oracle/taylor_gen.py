"""Oracle (test infrastructure only): CPU restatement of the reference's
Taylor-coefficient map generator, used by tests to check the GPU generator.
Nothing in the product imports this module.

Follows src/FFAG_HighOrder.py:
* the mid-plane field is B(r, theta) = A(r) G(theta - s(r));
* ``laplacian_terms`` -- the polar Laplacian applied n times as a sum of
  c r^-p d^a/dr^a d^b/dtheta^b (apply_r_derivative, apply_theta_derivative,
  apply_laplacian_combined, merge_similar_terms; :524-654);
* ``chain_terms`` -- d^i/dr^i G(theta - s(r)) as a sum of
  c G^(n) prod_j (s^(j))^m_j (calculate_first_derivative, merge_terms,
  calculate_higher_order_derivative; :11-91);
* ``sigmoid_derivative`` -- derivatives of the logistic as polynomials in
  sigma (sigmoid_nth_derivative; :212-243); ``shape_derivative`` -- the
  Leibniz product of the two edges of one pole (:319-349), summed over poles
  (generate_shapes_with_derivatives; :372-390);
* ``roll`` -- the azimuthal shift by int(s / dtheta) whole grid steps,
  periodic over the grid (shift_function; :393-397);
* ``power_derivative`` -- d^n/dr^n of a r^k or of a power series
  (recursive_nth_derivative; :448-486);
* ``generate_maps`` -- Bz_n = (-1)^n/(2n)! Lap^n B,
  Br_n = (-1)^n/(2n+1)! d/dr Lap^n B, Bfi_n = (-1)^n/(2n+1)! / r d/dtheta Lap^n B,
  evaluated row by row in r like Generate_Bmaps_MultiProcess /
  compute_taylor_coefficients_group (:913-998, no mean-flutter division).

Pins (tests/test_taylor_oracle.py): ``laplacian_terms`` and ``chain_terms``
against sympy's symbolic derivatives; ``sigmoid_derivative`` and
``shape_derivative`` against sympy; the zero-spiral maps against the
separable closed form; the roll direction (a positive spiral moves the field
pattern to larger azimuth); a planted sign slip in the chain rule.
"""
import math
from collections import defaultdict

import numpy as np


def _merge(terms):
    out = defaultdict(float)
    for key, c in terms:
        out[key] += c
    return {k: c for k, c in out.items() if c != 0}


def d_dr(terms):
    """d/dr of sum c r^-p d^a_r d^b_theta B: (a+1, b, p) with c, and
    (a, b, p+1) with -c p when p != 0."""
    new = []
    for (a, b, p), c in terms.items():
        new.append(((a + 1, b, p), c))
        if p != 0:
            new.append(((a, b, p + 1), -c * p))
    return _merge(new)


def d_dtheta(terms):
    return _merge([((a, b + 1, p), c) for (a, b, p), c in terms.items()])


def laplacian(terms):
    """d2/dr2 + (1/r) d/dr + (1/r^2) d2/dtheta2."""
    rr = d_dr(d_dr(terms))
    r1 = {(a, b, p + 1): c for (a, b, p), c in d_dr(terms).items()}
    tt = {(a, b, p + 2): c for (a, b, p), c in d_dtheta(d_dtheta(terms)).items()}
    return _merge(list(rr.items()) + list(r1.items()) + list(tt.items()))


def laplacian_terms(n):
    t = {(0, 0, 0): 1.0}
    for _ in range(n):
        t = laplacian(t)
    return t


def chain_terms(order):
    """d^order/dr^order G(theta - s(r)) = sum c G^(n) prod_j (s^(j))^m_j:
    {(n, (m_0, m_1, ...)): c}, m_0 being the power of s itself (always 0)."""
    cur = {(0, (0,)): 1.0}
    for _ in range(order):
        new = []
        for (n, m), c in cur.items():
            m = list(m)
            g = m.copy()
            if len(g) > 1:
                g[1] += 1
            else:
                g.append(1)
            new.append(((n + 1, tuple(g)), -c))            # G^(n+1)(u) du/dr, du/dr = -s'
            for i in range(len(m)):
                if m[i] > 0:
                    h = m.copy()
                    h[i] -= 1
                    if i + 1 < len(h):
                        h[i + 1] += 1
                    else:
                        h.append(1)
                    new.append(((n, tuple(h)), c * m[i]))
        cur = _merge(new)
    return cur


def sigmoid(x, x0, k):
    return 1.0 / (1.0 + np.exp(-k * (x - x0)))


def sigmoid_derivative(x, x0, k, order):
    """order-th derivative of 1/(1 + exp(-k (x - x0))) as P(sigma) with
    P_1 = k sigma - k sigma^2 and P_{i+1} = P_i'(sigma) P_1(sigma)."""
    s = sigmoid(x, x0, k)
    if order == 0:
        return s
    p1 = np.array([-k, k, 0.0])                          # highest power first
    p = p1
    for _ in range(order - 1):
        n = len(p) - 1
        dp = np.array([p[i] * (n - i) for i in range(n)])
        p = np.convolve(dp, p1)
    val = np.zeros_like(x)
    for c in p:
        val = val * s + c
    return val


def shape_derivative(x, x0, x1, w0, height, order):
    """order-th derivative of height sigma(x; x0, 16/w0) sigma(x; x1, -16/w0)."""
    k = 16.0 / w0
    if order == 0:
        return height * sigmoid(x, x0, k) * sigmoid(x, x1, -k)
    tot = np.zeros_like(x)
    for i in range(order + 1):
        tot += math.comb(order, i) * sigmoid_derivative(x, x0, k, i) * sigmoid_derivative(x, x1, -k, order - i)
    return height * tot


def flutter_derivative(theta, poles, order):
    """Sum over poles (x0, x1, w0, height) of shape_derivative."""
    g = np.zeros_like(theta)
    for x0, x1, w0, h in poles:
        g += shape_derivative(theta, x0, x1, w0, h, order)
    return g


def roll(theta, y, shift):
    """y moved by int(shift / dtheta) whole steps, periodically over the grid."""
    return np.roll(y, int(shift / (theta[1] - theta[0])))


def power_derivative(a, k, n, r):
    """d^n/dr^n of a r^k (a scalar), or of sum_i c_i (r - k)^i when ``a`` is
    the coefficient list, highest power first."""
    if np.isscalar(a):
        kv = np.append(np.arange(k, k - n, -1), 1)
        return a * np.prod(kv) * r ** (k - n)
    coefs = list(a)[::-1]
    out = np.zeros_like(np.asarray(r, dtype=np.float64))
    for i, c in enumerate(coefs):
        if i >= n:
            out = out + c * math.factorial(i) / math.factorial(i - n) * (r - k) ** (i - n)
    return out


def _mixed(a_r, b_th, r, theta, poles, s_poly, coef_k, use_roll):
    """d^a/dr^a d^b/dtheta^b [A(r) G(theta - s(r))] on the theta row at r."""
    tot = np.zeros_like(theta)
    s_val = s_poly(r)
    for i in range(a_r + 1):
        gi = np.zeros_like(theta)
        for (n, m), c in chain_terms(i).items():
            if use_roll:
                gq = roll(theta, flutter_derivative(theta, poles, n + b_th), s_val)
            else:
                gq = flutter_derivative(theta - s_val, poles, n + b_th)
            pr = 1.0
            for j, e in enumerate(m):
                pr *= s_poly.deriv(m=j)(r) ** e
            gi += gq * c * pr
        tot += math.comb(a_r, i) * gi * power_derivative(coef_k[0], coef_k[1], a_r - i, r)
    return tot


def _apply(terms, r, theta, poles, s_poly, coef_k, use_roll):
    out = np.zeros_like(theta)
    for (a, b, p), c in terms.items():
        out += _mixed(a, b, r, theta, poles, s_poly, coef_k, use_roll) * (c / r ** p)
    return out


def generate_maps(max_order, r_axis, theta, poles, s_poly, coef_k, use_roll=True):
    """(Bz, Br, Bfi) stacks of shape (max_order+1, len(r_axis), len(theta)).
    ``poles``: rows (x0, x1, w0, height) with height = sign |polarity|;
    ``s_poly``: a numpy Polynomial for s(r); ``coef_k``: (a, k)."""
    nr, nt = len(r_axis), len(theta)
    bz = np.zeros((max_order + 1, nr, nt))
    br = np.zeros_like(bz)
    bf = np.zeros_like(bz)
    lap = {(0, 0, 0): 1.0}
    for n in range(max_order + 1):
        c2n = (-1) ** n / math.factorial(2 * n)
        c2n1 = (-1) ** n / math.factorial(2 * n + 1)
        lr, lt = d_dr(lap), d_dtheta(lap)
        for ir, r in enumerate(r_axis):
            bz[n, ir] = c2n * _apply(lap, r, theta, poles, s_poly, coef_k, use_roll)
            br[n, ir] = c2n1 * _apply(lr, r, theta, poles, s_poly, coef_k, use_roll)
            bf[n, ir] = c2n1 / r * _apply(lt, r, theta, poles, s_poly, coef_k, use_roll)
        lap = laplacian(lap)
    return bz, br, bf
