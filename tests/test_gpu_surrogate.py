"""GPU tests of the field surrogate: matrix-core forward pass and gradients
against the CPU oracle (oracle/mlp.py, with the same BF16 rounding points),
Adam against its restatement, bit-reproducible training, the field it learns
against the FP64 map method, and the FP32 push driven by it."""
import ctypes

import numpy as np
import pytest

from oracle import mlp as O
from tests import cases

pytestmark = pytest.mark.gpu

ZT = 0.01


@pytest.fixture(scope="module")
def gp4():
    _, gp = cases.machines("stable", max_order=4)
    return gp


def make(torch, gp, seed=0):
    from pyffa_amd.FFAG_Surrogate import FFAG_FieldSurrogate
    return FFAG_FieldSurrogate(gp, zt=ZT, seed=seed)


def oracle_inputs(sur, pts):
    """Network inputs in FP32 exactly as the kernels form them."""
    d = sur.desc()
    r, fi, z = (pts[k].astype(np.float32) for k in range(3))
    a = (r - np.float32(d.r_c)) * np.float32(d.inv_rh)
    v = fi * np.float32(d.inv_sector)
    u = v - np.floor(v)
    c = z * np.float32(1.0 / d.zt)
    return a.astype(np.float64), u.astype(np.float64), (c * c).astype(np.float64)


def random_points(n, seed, sur):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(sur.r_min, sur.r_max, n), rng.uniform(-np.pi, np.pi, n),
                     rng.uniform(-ZT, ZT, n)]).astype(np.float32)


def perturbed_params(seed):
    from pyffa_amd.FFAG_Surrogate import init_params
    rng = np.random.default_rng(seed)
    p = init_params(seed)
    p += rng.normal(0, 0.02, p.size).astype(np.float32)
    return p


def test_forward_matches_oracle(torch_cuda, gp4):
    torch = torch_cuda
    print("[surrogate] forward vs oracle, 4000 points", flush=True)
    sur = make(torch, gp4)
    p = perturbed_params(11)
    sur.params.copy_(torch.tensor(p, device="cuda"))
    sur._pack()
    pts = random_points(4000, 1, sur)
    out = sur.eval(torch.tensor(pts, device="cuda"), mode=1).cpu().numpy().T
    m, wa, wc = sur.freq
    ref, _ = O.forward(p, *oracle_inputs(sur, pts), m, wa, wc, bf=True)
    scale = np.sqrt(np.mean(ref ** 2, axis=0))
    err = np.abs(out - ref) / scale
    print("[surrogate] forward rel err rms", np.sqrt(np.mean(err ** 2)), "max", err.max(), flush=True)
    assert np.sqrt(np.mean(err ** 2)) < 5e-3
    assert err.max() < 5e-2
    # mode 0 applies the field transform
    fld = sur.eval(torch.tensor(pts, device="cuda"), mode=0).cpu().numpy().T
    c = pts[2].astype(np.float64) / ZT
    assert np.allclose(fld, O.field(out.astype(np.float64), c, sur.scale, sur.delta), rtol=1e-4, atol=1e-12)


def test_gradients_match_oracle(torch_cuda, gp4):
    torch = torch_cuda
    print("[surrogate] gradient vs oracle, batch 3000", flush=True)
    sur = make(torch, gp4)
    p = perturbed_params(12)
    sur.params.copy_(torch.tensor(p, device="cuda"))
    sur._pack()
    n = 3000                                   # not a multiple of the 512-point chunk
    pts = random_points(n, 2, sur)
    tgt = np.random.default_rng(3).normal(0, 1, (3, n)).astype(np.float32)
    g, loss = sur.grad(torch.tensor(pts, device="cuda"), torch.tensor(tgt, device="cuda"))
    g = g.cpu().numpy().astype(np.float64)
    m, wa, wc = sur.freq
    o, cache = O.forward(p, *oracle_inputs(sur, pts), m, wa, wc, bf=True)
    l_ref, dy = O.loss_and_dy(o, tgt.T.astype(np.float64))
    g_ref = O.backward(cache, dy, bf=True)
    assert abs(float(loss.item()) / l_ref - 1) < 1e-3
    for li, ((w, b), (fi, fo)) in enumerate(zip(O.OFFSETS, O.LAYERS)):
        for name, sl in (("W", slice(w, w + fi * fo)), ("b", slice(b, b + fo))):
            rel = np.linalg.norm(g[sl] - g_ref[sl]) / np.linalg.norm(g_ref[sl])
            print(f"[surrogate] layer {li} {name} grad rel err {rel:.2e}", flush=True)
            assert rel < 2e-2, (li, name, rel)
    # a batch that is a whole number of 512-point chunks (no padding)
    g2, _ = sur.grad(torch.tensor(pts[:, :2048].copy(), device="cuda"), torch.tensor(tgt[:, :2048].copy(),
                                                                                    device="cuda"))
    o2, cache2 = O.forward(p, *oracle_inputs(sur, pts[:, :2048]), m, wa, wc, bf=True)
    _, dy2 = O.loss_and_dy(o2, tgt[:, :2048].T.astype(np.float64))
    g2_ref = O.backward(cache2, dy2, bf=True)
    assert np.linalg.norm(g2.cpu().numpy() - g2_ref) / np.linalg.norm(g2_ref) < 2e-2


def test_adam_matches_oracle(torch_cuda, gp4):
    torch = torch_cuda
    sur = make(torch, gp4)
    p = sur.params.cpu().numpy().astype(np.float64)
    mom, vel = np.zeros_like(p), np.zeros_like(p)
    rng = np.random.default_rng(4)
    for step in (1, 2, 3):
        g = rng.normal(0, 1e-2, p.size)
        sur.adam(torch.tensor(g, dtype=torch.float32, device="cuda"), lr=1e-3)
        O.adam_step(p, g.astype(np.float32).astype(np.float64), mom, vel, step, lr=1e-3)
        gp_ = sur.params.cpu().numpy()
        assert np.max(np.abs(gp_ - p)) < 1e-6, step
    assert sur.step == 3


def test_training_is_bit_reproducible(torch_cuda, gp4):
    torch = torch_cuda
    print("[surrogate] reproducibility: 2 x 40 iterations", flush=True)
    runs = []
    for _ in range(2):
        sur = make(torch, gp4, seed=3)
        sur.build_dataset(200_000, seed=5)
        sur.train(iters=40, batch=4096, lr=2e-3, seed=6, log_every=0)
        runs.append(sur.params.cpu().numpy())
    assert np.array_equal(runs[0], runs[1])


@pytest.fixture(scope="module")
def trained(torch_cuda, gp4):
    sur = make(torch_cuda, gp4, seed=0)
    print("[surrogate] dataset 1M + 2500 iterations", flush=True)
    sur.build_dataset(1_000_000, seed=1)
    hist = sur.train(iters=2500, batch=8192, lr=2e-3, seed=2, log_every=500)
    print("[surrogate] loss history", hist, flush=True)
    return sur


def test_trained_network_beats_the_method_without_maps(trained):
    """What the network gains: with the order-0 analytic field alone (no
    Taylor maps resident) the field error is that of the whole Taylor part;
    order 0 + the network must be at least 20x closer to the order-4 map
    method, with ~116 KB of weights instead of ~375 MB of FP32 maps."""
    e = trained.field_error(n=500_000, seed=9)
    print("[surrogate] field error", e, flush=True)
    assert e["rel_rms"] < e["order0_rel_rms"] / 20.0
    assert e["rel_rms"] < 1e-3
    h = trained.loss_history
    assert h[-1][1] < h[0][1] / 100.0


def test_push_with_surrogate_tracks_map_push(torch_cuda, trained, tmp_path):
    """FP32 tier, 2000 Boris steps of 4096 particles (with the aperture), three
    ways: the order-4 map gather (the method), order 0 + the network, and order
    0 alone (the method without maps).  The network's trajectories must stay
    at least 10x closer to the method's than order 0 alone."""
    torch = torch_cuda
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    M = cases.bunch("stable", n=4096, seed=0)
    _, gp_map = cases.machines("stable", max_order=4, aperture=cases.APERTURE)
    _, gp_o0 = cases.machines("stable", max_order=0, aperture=cases.APERTURE)
    runs = {}
    for name, gp, sur in (("map", gp_map, None), ("net", gp_map, trained), ("order0", gp_o0, None)):
        print(f"[surrogate] push 2000 steps: {name}", flush=True)
        B = FFAG_Bunch(M, marcosize=1.0, sc_grid_size=(16, 32, 32, 4.0), BunchType="Boris", precision="fp32")
        lp = dict(stop_condition=dict(max_stepsN=2000, max_turn=1e9, max_time=1.0), time_step=0.05e-9,
                  step_dumps=None, enable_SC=False, SC_step=20, precision="fp32", verbose=False)
        if sur is not None:
            lp["surrogate"] = sur
        FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
        runs[name] = B.LocalBunch.copy()
    ref = runs["map"]
    dev = {}
    for name in ("net", "order0"):
        a = runs[name]
        both = (a[:, 8] != 0) & (ref[:, 8] != 0)
        lost_diff = np.mean((a[:, 8] != 0) != (ref[:, 8] != 0))
        d = {c: np.quantile(np.abs(a[both, c] - ref[both, c]), 0.99) for c in (0, 2, 4)}
        dev[name] = d
        print(f"[surrogate] {name}: alive {both.sum()}, lost-flag differences {lost_diff:.4f}, "
              f"p99 |dr| {d[0]:.2e} m, |dz| {d[2]:.2e} m, |dphi| {d[4]:.2e} rad", flush=True)
        if name == "net":
            assert lost_diff < 0.005
    for c in (0, 2):
        assert dev["net"][c] < dev["order0"][c] / 10.0, (c, dev)


def test_surrogate_drives_8m_protons_with_space_charge(torch_cuda, trained, tmp_path):
    """The surrogate workload at its full size: 8M protons (1e13), SC on the
    (512, 256, 256) grid every 20 steps, RF and aperture, the FP32 tier with
    the network in place of the order-4 map gather for |z| <= zt and the maps
    beyond; 100 steps (5 macro steps).  The run is bit-reproducible, the
    deposit of the final bunch holds exactly 2^(2 bits) units per survivor
    and the counts add up, and the bunch moments stay within 1e-3 of the
    FP32 order-4 map tier's on the same input."""
    torch = torch_cuda
    from pyffa_amd import workloads as W
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch, fixed_point_bits
    from pyffa_amd.FFAG_track import FFAG_RungeKutta
    n, grid, steps = 8_000_000, (512, 256, 256, 4.0), 100
    print(f"[case] surrogate 8M with SC, {steps} steps", flush=True)
    gp = W.global_parameters("stable", max_order=4, rf=dict(v0=100e3), aperture=W.APERTURE, tmpdir=str(tmp_path))
    M = W.bunch("stable", n=n, seed=0)
    runs = {}
    for name, sur in (("net", trained), ("net2", trained), ("map", None)):
        B = FFAG_Bunch(M, marcosize=1e13 / n, sc_grid_size=grid, BunchType="Boris", precision="fp32")
        W.set_bunch_seo(B)
        lp = dict(stop_condition=dict(max_stepsN=steps, max_turn=1e9, max_time=1.0), time_step=0.05e-9,
                  step_dumps=None, enable_SC=True, SC_type=1, SC_step=20, precision="fp32", verbose=False)
        if sur is not None:
            lp["surrogate"] = sur
        FFAG_RungeKutta().rk4_solve_dt_bunch3_withSC_boris(None, 0.0, B, gp, lp)
        runs[name] = {k: B.ordered(k) for k in ("r", "vr", "z", "vz", "fi", "fidot", "flags")}
        if name == "net":
            alive = int(((B.flags & 2) != 0).sum().item())
            B.sc_charge_grid = B.sc_cell_counts = True
            B.sc_solve_device(first=False, s_last=0.95)
            assert B.sc_result()["count"] == alive
            assert int(B.sc_view("grid").sum().item()) == alive << (2 * fixed_point_bits(n))
            assert int(B.sc_view("slice_count").sum().item()) == alive
            assert int(B.sc_view("cell_count").sum().item()) == alive
        del B
    for k in runs["net"]:
        assert torch.equal(runs["net"][k], runs["net2"][k]), k
    a, b = runs["net"], runs["map"]
    both = ((a["flags"] & 2) != 0) & ((b["flags"] & 2) != 0)
    assert both.sum().item() > 0.99 * n

    def mom(S):
        r, z = S["r"][both], S["z"][both]
        v = torch.sqrt(S["vr"][both] ** 2 + S["vz"][both] ** 2 + (r * S["fidot"][both]) ** 2)
        xp, zp = S["vr"][both] / v, S["vz"][both] / v

        def emit(u, up):
            u, up = u - u.mean(), up - up.mean()
            return float(torch.sqrt((u * u).mean() * (up * up).mean() - (u * up).mean() ** 2))
        return float(r.mean()), float(z.mean()), emit(r, xp), emit(z, zp)
    ma, mb = mom(a), mom(b)
    print("net vs map moments", ma, mb, flush=True)
    assert abs(ma[0] - mb[0]) < 1e-6 and abs(ma[1] - mb[1]) < 1e-6
    assert abs(ma[2] / mb[2] - 1) < 1e-3 and abs(ma[3] / mb[3] - 1) < 1e-3
