"""Pins the CPU restatement of the field surrogate (oracle/mlp.py) before it
is used to check the GPU kernels: backward against central finite
differences, Adam's first step against its closed form, BF16 rounding
against torch's conversion, and the feature/field symmetries the network is
built to have."""
import numpy as np
import pytest

from oracle import mlp as O
from pyffa_amd import FFAG_Surrogate as S
from pyffa_amd import _lib


def small_inputs(n=16, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, n)
    u = rng.uniform(0, 1, n)
    c2 = rng.uniform(0, 1, n)
    return a, u, c2


def test_parameter_layout_matches_library():
    assert O.NPARAM == _lib.MLP_NPARAM
    p = S.init_params(0)
    assert p.size == O.NPARAM
    # W_l stored [in][out]: the bias blocks sit right after each weight block
    L = O.unflatten(p.astype(np.float64))
    assert [W.shape for W, _ in L] == [(64, 128), (128, 128), (128, 128), (128, 128), (128, 3)]
    assert all(np.all(b == 0) for _, b in L)


def test_backward_matches_finite_differences():
    m, wa, wc = S.default_frequencies()
    a, u, c2 = small_inputs()
    p = S.init_params(5).astype(np.float64)
    rng = np.random.default_rng(9)
    p += rng.normal(0, 0.05, p.size)          # non-zero biases too
    t = rng.normal(0, 1, (a.size, 3))

    def loss(q):
        o, _ = O.forward(q, a, u, c2, m, wa, wc, bf=False)
        return O.loss_and_dy(o, t)[0]

    o, cache = O.forward(p, a, u, c2, m, wa, wc, bf=False)
    _, dy = O.loss_and_dy(o, t)
    g = O.backward(cache, dy, bf=False)
    idx = np.concatenate([rng.choice(np.arange(w, w + fi * fo), 6, replace=False) for (w, b), (fi, fo) in
                          zip(O.OFFSETS, O.LAYERS)] + [np.arange(b, b + 3) for (w, b), _ in zip(O.OFFSETS, O.LAYERS)])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_planted_backward_bug_is_caught():
    """A transposed weight in the backward pass fails the same check."""
    m, wa, wc = S.default_frequencies()
    a, u, c2 = small_inputs(8)
    p = S.init_params(5).astype(np.float64)
    t = np.ones((a.size, 3))
    o, cache = O.forward(p, a, u, c2, m, wa, wc, bf=False)
    _, dy = O.loss_and_dy(o, t)
    xs, zs, L = cache
    good = O.backward(cache, dy, bf=False)
    bad_L = [(W.copy(), b) for W, b in L]
    bad_L[2] = (bad_L[2][0].T.copy(), bad_L[2][1])       # W_2 transposed
    bad = O.backward((xs, zs, bad_L), dy, bf=False)
    w1 = O.OFFSETS[1][0]
    assert np.max(np.abs(bad[w1:w1 + 128 * 128] - good[w1:w1 + 128 * 128])) > 1e-6


def test_adam_first_step_closed_form():
    """After one step from zero moments, m_hat / sqrt(v_hat) = sign(g)."""
    rng = np.random.default_rng(1)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    p0 = p.copy()
    mom, vel = np.zeros(50), np.zeros(50)
    O.adam_step(p, g, mom, vel, 1, lr=1e-3, eps=0.0)
    assert np.allclose(p, p0 - 1e-3 * np.sign(g), rtol=0, atol=1e-15)


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(size=5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0, 3.0e38])])
    x32 = x.astype(np.float32)
    ref = torch.tensor(x32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(O.bf16(x32), ref)


def test_features_periodic_over_the_sector():
    m, wa, wc = S.default_frequencies()
    assert np.all(m == np.round(m))
    a, u, c2 = small_inputs(32)
    f1 = O.features(a, u, c2, m, wa, wc)
    f2 = O.features(a, u + 1.0, c2, m, wa, wc)
    assert np.allclose(f1, f2, atol=1e-9)
    # sin^2 + cos^2 = 1 for every frequency
    assert np.allclose(f1[:, 0::2] ** 2 + f1[:, 1::2] ** 2, 1.0)


def test_field_transform_parity_and_midplane():
    rng = np.random.default_rng(4)
    o = rng.normal(size=(10, 3))
    c = rng.uniform(-1, 1, 10)
    s = (0.2, 0.1, 0.001)
    f, fm = O.field(o, c, s), O.field(o, -c, s)
    assert np.allclose(f[:, 0], fm[:, 0])            # Bz even in z
    assert np.allclose(f[:, 1:], -fm[:, 1:])         # Bfi, Br odd in z
    assert np.all(O.field(o, np.zeros(10), s) == 0)  # no correction on the mid-plane


def test_bf16_forward_close_to_fp64_forward():
    """The rounding model only perturbs the outputs at the BF16 level."""
    m, wa, wc = S.default_frequencies()
    a, u, c2 = small_inputs(256)
    p = S.init_params(0)
    o1, _ = O.forward(p, a, u, c2, m, wa, wc, bf=True)
    o2, _ = O.forward(p, a, u, c2, m, wa, wc, bf=False)
    rel = np.sqrt(np.mean((o1 - o2) ** 2) / np.mean(o2 ** 2))
    assert 1e-4 < rel < 3e-2


def test_scaled_bf16_mode_agrees_with_the_unscaled_float64_mode(monkeypatch):
    """The BF16 mode's scaled pre-activations (KAPPA = -log2 e folded into
    the packed W_0, hidden biases and W_4) must give the float64 network's
    outputs and gradients to BF16 rounding, non-zero biases included; a
    gradient that forgets to undo the activations' KAPPA fails."""
    m, wa, wc = S.default_frequencies()
    a, u, c2 = small_inputs(512, seed=7)
    rng = np.random.default_rng(13)
    p = S.init_params(3).astype(np.float64) + rng.normal(0, 0.05, O.NPARAM)
    t = rng.normal(0, 1, (a.size, 3))

    def grads(bf):
        o, cache = O.forward(p, a, u, c2, m, wa, wc, bf=bf)
        _, dy = O.loss_and_dy(o, t)
        return o, O.backward(cache, dy, bf=bf)

    o64, g64 = grads(False)
    o16, g16 = grads(True)
    assert np.sqrt(np.mean((o16 - o64) ** 2) / np.mean(o64 ** 2)) < 3e-2
    for (w, b), (fi, fo) in zip(O.OFFSETS, O.LAYERS):
        for sl in (slice(w, w + fi * fo), slice(b, b + fo)):
            assert np.linalg.norm(g16[sl] - g64[sl]) / np.linalg.norm(g64[sl]) < 5e-2
    monkeypatch.setattr(O, "_scales", lambda li, nl: (O.KAPPA if li == 0 else (O.INV_KAPPA if li == nl - 1 else None),
                                                      None if li == nl - 1 else O.KAPPA, 1.0))
    _, bad = grads(True)
    w1, _ = O.OFFSETS[1]
    sl = slice(w1, w1 + 128 * 128)
    assert np.linalg.norm(bad[sl] - g64[sl]) / np.linalg.norm(g64[sl]) > 0.3
