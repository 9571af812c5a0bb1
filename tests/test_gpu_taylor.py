"""GPU Taylor-map generator (ffa_taylor_maps) against the oracle's
restatement of the reference generator, on spiral and zero-spiral lattices,
both output layouts and both output types, plus the main_Bmap folder flow."""
import json

import numpy as np
import pytest
from numpy.polynomial import Polynomial

from oracle import taylor_gen as T

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

POLES = [(np.deg2rad(7.5), np.deg2rad(16.875), np.deg2rad(9.375 * 0.2), (1.0, 1)),
         (np.deg2rad(19.6875), np.deg2rad(22.5), np.deg2rad(2.8125 * 0.35), (0.728, -1))]


def _oracle_poles():
    return [(x0, x1, w0, sg * m) for x0, x1, w0, (m, sg) in POLES]


def _setup(spiral_deg, nr=13, nphi=241, rmin=8.0, rmax=9.3):
    from pyffa_amd import FFAG_HighOrder as H
    r = np.linspace(rmin, rmax, nr)
    fi = np.linspace(0, np.pi / 6, nphi)
    s_fit = Polynomial.fit(r, H.calculate_s(r, 0.0, spiral_deg), 5)
    return r, fi, s_fit


@pytest.mark.parametrize("spiral_deg", [35.0, 0.0, -50.0])
@pytest.mark.parametrize("max_order", [0, 1, 3])
def test_maps_match_oracle(spiral_deg, max_order):
    from pyffa_amd import FFAG_HighOrder as H
    print(f"taylor maps spiral {spiral_deg} order {max_order}", flush=True)
    r, fi, s_fit = _setup(spiral_deg)
    gen = H.TaylorMapGenerator(max_order, r, fi, POLES, s_fit, (0.35, 7.6))
    st = gen.run("stacks", torch.float64)
    torch.cuda.synchronize()
    st = st.cpu().numpy()
    bz, br, bf = T.generate_maps(max_order, r, fi, _oracle_poles(), s_fit, (0.35, 7.6))
    for comp, ref in ((H.BZ, bz), (H.BR, br), (H.BFI, bf)):
        for n in range(max_order + 1):
            scale = np.max(np.abs(ref[n]))
            assert scale > 0
            err = np.max(np.abs(st[comp, n] - ref[n])) / scale
            assert err < 1e-12, (comp, n, err)


def test_layouts_and_types_agree_bitwise():
    from pyffa_amd import FFAG_HighOrder as H
    r, fi, s_fit = _setup(35.0, nr=37, nphi=301)
    gen = H.TaylorMapGenerator(3, r, fi, POLES, s_fit, (0.35, 7.6))
    st = gen.run("stacks", torch.float64)
    tr = gen.run("tracking", torch.float64)
    st32 = gen.run("stacks", torch.float32)
    tr32 = gen.run("tracking", torch.float32)
    torch.cuda.synchronize()
    # tracking[j][i][3 (n-1) + (Bz, Bfi, Br)] == stacks[(Bz, Br, Bfi)][n][i][j]
    want = torch.stack([st[H.BZ, 1:], st[H.BFI, 1:], st[H.BR, 1:]], dim=1)       # (n, 3, nr, nphi)
    want = want.permute(3, 2, 0, 1).reshape(fi.size, r.size, -1)
    assert torch.equal(tr, want)
    assert torch.equal(st32, st.to(torch.float32))
    assert torch.equal(tr32, tr.to(torch.float32))


def test_zero_spiral_matches_closed_form_of_the_lattice_module():
    from pyffa_amd import FFAG_HighOrder as H
    from pyffa_amd import lattice as Lt
    from pyffa_amd.FFAG_Field import FFAG_Bfield_analytical
    lat = Lt.ScalingFFAG(rstep=0.01, theta_step_deg=0.05, B0=1.2 / 8.5 ** 7.6)
    bz, br, bf = Lt.taylor_maps(lat, 3)
    cfg = lat.to_config()["Bmap"]
    r, fi = H.grid_from_config(cfg)
    conv = H.BmapParams_to_ConvexParams(cfg)
    gen = H.TaylorMapGenerator(3, r, fi, conv, Polynomial([0.0]), tuple(cfg["coefficients_tupel"]))
    st = gen.run("stacks", torch.float64)
    torch.cuda.synchronize()
    st = st.cpu().numpy()
    assert np.allclose(r, lat.r_axis()) and np.allclose(fi, lat.fi_axis())
    assert FFAG_Bfield_analytical.from_arrays(lat.to_config(), 0, (bz, br, bf)).r_size == r.size
    # the closed form evaluates G^(q) its own way: at q = 7 and edge slopes
    # k ~ 900 the two evaluations differ by rounding of order 1e-11
    for comp, ref in ((H.BZ, bz), (H.BR, br), (H.BFI, bf)):
        for n in range(4):
            err = np.max(np.abs(st[comp, n] - ref[n])) / np.max(np.abs(ref[n]))
            assert err < 1e-10, (comp, n, err)


def test_main_bmap_folder_loads_in_the_tracker(tmp_path):
    from pyffa_amd import FFAG_HighOrder as H
    from pyffa_amd import lattice as Lt
    from pyffa_amd.FFAG_Field import FFAG_Bfield_analytical
    lat = Lt.ScalingFFAG(rstep=0.01, theta_step_deg=0.05, B0=1.2 / 8.5 ** 7.6)
    cfg = lat.to_config()
    cfg["Bmap"]["SpiralAngle_deg"] = 20.0
    cfg["config"] = {"folder": "maps", "expand_order": 2}
    path = tmp_path / "config_Bmap.json"
    path.write_text(json.dumps(cfg))
    out = H.main_Bmap(str(path))
    B = FFAG_Bfield_analytical(str(tmp_path / "maps" / "config_Bmap.json"), 2)
    assert B.Bz_stack.shape == (3, B.r_size, B.fi_size)
    saved = json.loads((tmp_path / "maps" / "config_Bmap.json").read_text())
    assert len(saved["Bmap"]["Bmean"]) == B.r_size and out.endswith("maps")
    np.testing.assert_allclose(saved["Bmap"]["Bmean"], B.Bz_stack[0].mean(axis=1))
    # the same maps as the oracle
    r, fi = H.grid_from_config(cfg["Bmap"])
    s_fit = Polynomial.fit(r, H.calculate_s(r, 0.0, 20.0), 5)
    poles = [(x0, x1, w0, sg * m) for x0, x1, w0, (m, sg) in H.BmapParams_to_ConvexParams(cfg["Bmap"])]
    rows = [0, r.size // 2, r.size - 1]
    bz, br, bf = T.generate_maps(2, r[rows], fi, poles, s_fit, tuple(cfg["Bmap"]["coefficients_tupel"]))
    for st, ref in ((B.Bz_stack, bz), (B.Br_stack, br), (B.Bfi_stack, bf)):
        for n in range(3):
            assert np.max(np.abs(st[n][rows] - ref[n])) <= 1e-12 * np.max(np.abs(ref[n]))


def test_invalid_arguments_are_rejected():
    import ctypes
    from pyffa_amd import _lib
    ctx = _lib.Context(0, torch.cuda.current_stream().cuda_stream)
    g = _lib.FfaTaylorGen(n_orders=9, n_q=18, n_poles=1, layout=0, out_f32=0, nr=4, n_phi=4)
    assert ctx.lib.ffa_taylor_maps(ctx.h, ctypes.byref(g)) == -1
    g = _lib.FfaTaylorGen(n_orders=2, n_q=5, n_poles=1, layout=0, out_f32=0, nr=4, n_phi=4)
    assert ctx.lib.ffa_taylor_maps(ctx.h, ctypes.byref(g)) == -1
    g = _lib.FfaTaylorGen(n_orders=1, n_q=2, n_poles=1, layout=1, out_f32=0, nr=4, n_phi=4)
    assert ctx.lib.ffa_taylor_maps(ctx.h, ctypes.byref(g)) == -1      # tracking layout needs an order >= 1
    g = _lib.FfaTaylorGen(n_orders=2, n_q=4, n_poles=1, layout=0, out_f32=0, nr=4, n_phi=4)
    assert ctx.lib.ffa_taylor_maps(ctx.h, ctypes.byref(g)) == -1      # NULL arrays
    assert b"NULL" in ctx.lib.ffa_last_error(ctx.h)
