"""GPU tests of the azimuth ordering of the FP32 tier: the permutation is the
stable sort of the bin keys (against numpy's stable argsort), every state
array moves with its particle, and tracking does not depend on the order."""
import ctypes

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def keys_of(fi, flags, nbins):
    v = fi * (1.0 / (2.0 * np.pi))
    u = v - np.floor(v)
    k = np.minimum((u * nbins).astype(np.int64), nbins - 1)
    live = ((flags & 1) != 0) & ((flags & 2) != 0) & np.isfinite(fi)
    return np.where(live, k, nbins)


@pytest.mark.parametrize("n,count,nbins", [(10000, 10000, 65535), (70001, 50000, 1000), (5000, 4097, 7)])
def test_sort_is_the_stable_key_order(torch_cuda, n, count, nbins):
    torch = torch_cuda
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    print(f"[sort] n={n} count={count} nbins={nbins}", flush=True)
    rng = np.random.default_rng(n)
    M = cases.bunch("stable", n=n, seed=1)
    M[:, 4] = rng.uniform(-20.0, 40.0, n)            # several turns, both signs
    B = FFAG_Bunch(M, precision="fp32")
    fl = np.full(n, 3, np.uint8)
    fl[rng.random(n) < 0.1] = 1                      # injected, lost
    fl[count:] = 0                                   # not injected
    B.flags.copy_(torch.tensor(fl, device="cuda"))
    B._uninject_head = count
    before = {k: getattr(B, k).cpu().numpy().copy() for k in B._STATE}
    B.sort_by_azimuth(nbins=nbins)
    order = np.argsort(keys_of(before["fi"][:count], fl[:count], nbins), kind="stable")
    pid = B.pid.cpu().numpy()
    assert np.array_equal(pid[:count], order)
    assert np.array_equal(pid[count:], np.arange(count, n))
    full = np.concatenate([order, np.arange(count, n)])
    for k in B._STATE:
        a = getattr(B, k).cpu().numpy()
        want = before[k][:, full] if a.ndim == 2 else before[k][full]
        assert np.array_equal(a, want), k
    # the host mirror is in the input order
    assert np.array_equal(B.ordered("r").cpu().numpy(), before["r"])


def _track(gp, M, steps, sort_every, sc=None):
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    grid = sc["grid"] if sc else (16, 32, 32, 4.0)
    B = FFAG_Bunch(M, marcosize=sc["marcosize"] if sc else 1.0, sc_grid_size=grid, BunchType="Boris",
                   precision="fp32")
    if sc:
        cases.set_bunch_seo(B)
    lp = dict(stop_condition=dict(max_stepsN=steps, max_turn=1e9, max_time=1.0), time_step=0.05e-9,
              step_dumps=None, enable_SC=bool(sc), SC_step=20, precision="fp32", verbose=False,
              sort_every=sort_every, max_block=100)   # same launch boundaries in both runs
    FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
    return B.LocalBunch.copy()


def test_push_is_independent_of_the_order(torch_cuda, tmp_path):
    """Without SC each particle's push is independent of the others: sorting
    every 500 steps gives bit-identical states after 1200 steps."""
    print("[sort] fp32 push 1200 steps, sorted vs unsorted", flush=True)
    _, gp = cases.machines("stable", max_order=2, rf=dict(v0=100e3), aperture=cases.APERTURE, tmpdir=str(tmp_path))
    M = cases.bunch("stable", n=5000, seed=4)
    a = _track(gp, M, 1200, 500)
    b = _track(gp, M, 1200, 0)
    assert np.array_equal(a, b)


def test_sc_tracking_with_sort_matches_unsorted(torch_cuda, tmp_path):
    """With SC the moment sums change order, so states agree to rounding."""
    print("[sort] fp32 with SC 1200 steps, sorted vs unsorted", flush=True)
    _, gp = cases.machines("stable", max_order=2, rf=dict(v0=100e3), aperture=cases.APERTURE, tmpdir=str(tmp_path))
    n = 4000
    M = cases.bunch("stable", n=n, seed=9, fi_spread=1.0)
    sc = dict(grid=(16, 32, 32, 4.0), marcosize=1e11 / n)
    a = _track(gp, M, 1200, 200, sc)
    b = _track(gp, M, 1200, 0, sc)
    alive = (a[:, 8] != 0) & (b[:, 8] != 0)
    assert np.mean((a[:, 8] != 0) != (b[:, 8] != 0)) < 0.002
    for c in (0, 2, 4):
        d = np.abs(a[alive, c] - b[alive, c])
        print(f"[sort] col {c}: max |diff| {d.max():.3e}, median {np.median(d):.3e}", flush=True)
        assert np.median(d) < 1e-9
