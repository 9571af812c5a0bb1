"""CPU restatement of the field surrogate -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` use this module,
as the checker of the GPU surrogate kernels; the product package never
imports it.

The surrogate has no counterpart in the reference: it stands in for the
Taylor-map gather and z-extrapolation ``Bmap_add_all_orders``
(src/FFAG_MathTools.py:1139-1188), i.e. for
``dBz = sum_n Bz_n z^(2n)``, ``dBfi = sum_n Bfi_n z^(2n+1)``,
``dBr = sum_n Br_n z^(2n+1)`` (n >= 1).

Network (same definition as the HIP kernels):

* inputs ``a = (r - r_c) / r_h``, ``u = (theta mod sector) / sector``,
  ``c = z / z_t``;
* 32 Fourier features ``rev_f = m_f u + wa_f a + wc_f c^2`` (revolutions),
  feature ``2f = sin(2 pi rev_f)``, ``2f+1 = cos(2 pi rev_f)``; the integer
  ``m_f`` make the features periodic over the sector;
* four hidden layers of 128 SiLU units, then 3 linear outputs ``o``;
* field ``dBz = S_z c^2 g(o_0)``, ``dBfi = S_fi c^3 g(o_1)``,
  ``dBr = S_r c^3 g(o_2)`` with ``g(o) = o`` or ``delta sinh(o)``, so the z
  parity of the Taylor series holds exactly.

``forward``/``backward`` can round to bfloat16 exactly where the GPU does
(weights, features, hidden activations and the backward signals that feed the
matrix cores) so that GPU-vs-oracle differences reduce to FP32 summation
order.  In that mode they also follow the GPU's scaled pre-activations:
the packed W_0 and hidden biases carry KAPPA = -log2(e) and W_4 carries
1 / KAPPA, so each hidden layer computes y' = KAPPA y and the activation
a' = y' / (1 + 2^y') = KAPPA silu(y) (BF16-rounded), and the weight gradients
of layers 1..4, taken from the scaled activations, are divided by KAPPA.
With ``bf16=False`` everything is float64 and unscaled: that mode is what
the finite-difference gradient check pins, and the BF16 mode must agree with
it to BF16 rounding.  ``adam_step`` is Algorithm 1 of Kingma & Ba (2015).
"""
from __future__ import annotations

import numpy as np

NFREQ, NFEAT, HID, NOUT = 32, 64, 128, 3
LAYERS = [(NFEAT, HID), (HID, HID), (HID, HID), (HID, HID), (HID, NOUT)]


def param_offsets():
    """Offsets of (W_l, b_l) in the flat FP32 parameter vector; W_l is
    stored [in][out] row-major."""
    offs, o = [], 0
    for fi, fo in LAYERS:
        offs.append((o, o + fi * fo))
        o += fi * fo + fo
    return offs, o


OFFSETS, NPARAM = param_offsets()


def unflatten(p):
    return [(p[w:w + fi * fo].reshape(fi, fo), p[b:b + fo]) for (w, b), (fi, fo) in zip(OFFSETS, LAYERS)]


def init_params(seed=0):
    """He-uniform style init (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for (w, b), (fi, fo) in zip(OFFSETS, LAYERS):
        lim = np.sqrt(6.0 / fi) * (0.5 if fo == NOUT else 1.0)
        p[w:w + fi * fo] = rng.uniform(-lim, lim, fi * fo).astype(np.float32)
    return p


def bf16(x):
    """Round to the nearest bfloat16 (ties to even) through float32."""
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(a, u, c2, m, wa, wc):
    rev = np.outer(u, m) + np.outer(a, wa) + np.outer(c2, wc)
    rev = rev - np.floor(rev)
    f = np.empty((a.size, NFEAT))
    f[:, 0::2] = np.sin(2 * np.pi * rev)
    f[:, 1::2] = np.cos(2 * np.pi * rev)
    return f


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


KAPPA = np.float32(-1.4426950408889634)        # -log2(e), as the GPU's FP32 constant
INV_KAPPA = np.float32(-0.6931471805599453)    # -ln(2)


def _scales(li, nl):
    """(weight factor, bias factor, input scale) of layer li in the BF16 mode."""
    if li == 0:
        return KAPPA, KAPPA, 1.0
    if li == nl - 1:
        return INV_KAPPA, None, KAPPA
    return None, KAPPA, KAPPA


def forward(p, a, u, c2, m, wa, wc, bf=True):
    """Returns (o, cache); o is (n, 3).  In the BF16 mode the cache holds
    the scaled pre-activations KAPPA y and activations BF16(KAPPA silu)."""
    L = unflatten(np.asarray(p, np.float64))
    if not bf:
        x = features(a, u, c2, m, wa, wc)
        xs, zs = [x], []
        for li, (W, b) in enumerate(L):
            z = x @ W + b
            if li == len(L) - 1:
                return z, (xs, zs, L)
            zs.append(z)
            x = silu(z)
            xs.append(x)
    x = bf16(features(a, u, c2, m, wa, wc))
    xs, zs = [x], []
    for li, (W, b) in enumerate(L):
        fw, fb, _ = _scales(li, len(L))
        W32, b32 = W.astype(np.float32), b.astype(np.float32)
        Wp = bf16(W32 * fw if fw is not None else W32)
        bp = (b32 * fb if fb is not None else b32).astype(np.float64)
        z = x @ Wp + bp
        if li == len(L) - 1:
            return z, (xs, zs, L)
        zs.append(z)
        x = bf16(z / (1.0 + np.exp2(z)))
        xs.append(x)


def backward(cache, dy, bf=True):
    """Gradient of sum(dy * o) with respect to the flat parameters."""
    xs, zs, L = cache
    g = np.zeros(NPARAM)
    if not bf:
        d = np.asarray(dy, np.float64)
        for li in range(len(L) - 1, -1, -1):
            (w, b), (fi, fo) = OFFSETS[li], LAYERS[li]
            g[w:w + fi * fo] = (xs[li].T @ d).ravel()
            g[b:b + fo] = d.sum(0)
            if li == 0:
                break
            d = (d @ L[li][0].T) * dsilu(zs[li - 1])
        return g
    d = bf16(dy)
    for li in range(len(L) - 1, -1, -1):
        (w, b), (fi, fo) = OFFSETS[li], LAYERS[li]
        _, _, xin = _scales(li, len(L))
        g[w:w + fi * fo] = (xs[li].T @ d).ravel() / float(xin)
        g[b:b + fo] = d.sum(0)
        if li == 0:
            break
        dx = d @ bf16(L[li][0]).T
        d = bf16(dx * dsilu(zs[li - 1] * float(INV_KAPPA)))
    return g


def loss_and_dy(o, t):
    """Mean squared error over all outputs and its gradient d loss / d o."""
    e = o - t
    return float(np.mean(e * e)), 2.0 * e / e.size


def adam_step(p, g, mom, vel, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update in place (float64 math on float32 storage)."""
    mom[:] = beta1 * mom + (1 - beta1) * g
    vel[:] = beta2 * vel + (1 - beta2) * g * g
    mh = mom / (1 - beta1 ** step)
    vh = vel / (1 - beta2 ** step)
    p[:] = p - lr * mh / (np.sqrt(vh) + eps)


def field(o, c, scale, delta=0.0):
    """Network outputs to the Taylor-correction field (dBz, dBfi, dBr); with
    delta > 0 the outputs are first mapped through delta * sinh."""
    g = delta * np.sinh(o) if delta > 0 else o
    c2 = c * c
    return np.stack([scale[0] * c2 * g[:, 0], scale[1] * c2 * c * g[:, 1], scale[2] * c2 * c * g[:, 2]], 1)
