"""The stage functions of pyffa_amd.FFAG_SC_new (the reference's
src/FFAG_SC_new.py names, on the GPU) against the oracle's restatement of
each stage on the same inputs, and Build_2_5D_SliceGrid_sbs's 18-tuple
(src/FFAG_Bunch.py:1352) against the oracle's solve, also after the FP32
tier has reordered the particles."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from tests import cases

pytestmark = pytest.mark.gpu
L = O.lib()
p, d, i64, ci = O._p, O._d, O._i64, O._i


def _bunch(n, seed, lost_every=97):
    M = cases.bunch("stable", n=n, seed=seed, fi_spread=0.6)
    M[:, 7] = 1.0
    M[:, 8] = 1.0
    M[::lost_every, 8] = 0.0
    return M


@pytest.mark.parametrize("rk", [False, True])
def test_local_coords_and_divide_arc(torch_cuda, rk):
    from pyffa_amd import FFAG_SC_new as SC
    print(f"[case] stages local coords rk={rk}", flush=True)
    n = 30000
    M = _bunch(n, 11)
    if rk:
        M = cases.bunch("stable", n=n, seed=11, solver="rk4", fi_spread=0.6)
        M[:, 7] = M[:, 8] = 1.0
    surv = M[:, 8] != 0
    bufs = [np.zeros(n) for _ in range(6)] + [np.zeros(n, np.int64)]
    fn = SC.GetLocalCoordinates_2p5_rk if rk else SC.GetLocalCoordinates_2p5
    r, fi, z, g, mg, fm, idx, j, sg = fn(M, surv, *bufs, O.C)
    ref = np.zeros(6 * n)
    sgr = ctypes.c_double()
    cnt = L.ora_sc_local_coords(p(M), i64(n), p(surv.astype(np.uint8)), d(O.C), ci(1 if rk else 0), p(ref),
                                ctypes.byref(sgr))
    assert j == cnt == surv.sum()
    for k, a in enumerate((r, fi, z, fm)):
        np.testing.assert_array_equal(a, ref[(k if k < 3 else 4) * n: (k if k < 3 else 4) * n + j], err_msg=str(k))
    np.testing.assert_allclose(g, ref[3 * n: 3 * n + j], rtol=4e-16)
    np.testing.assert_array_equal(idx, ref[5 * n: 5 * n + j].astype(np.int64))
    assert np.all(mg == 1.0)
    assert sg == pytest.approx(sgr.value, rel=1e-13)
    # covering arc of the survivors' azimuths: min / max are exact operations
    arc = np.zeros(3)
    L.ora_divide_arc2(p(np.ascontiguousarray(fm)), i64(j), p(arc))
    got = SC.divide_arc2(fm)
    assert got == tuple(arc)
    # the merge of several ranks' arcs is host arithmetic
    s, e = SC.merge_multiple_arcs([(got[0], got[1]), (got[0] + 0.1, got[1] + 0.1)])
    assert 0.0 <= s < 2 * np.pi and 0.0 <= e < 2 * np.pi


def test_slices_deposit_and_fields(torch_cuda):
    from pyffa_amd import FFAG_SC_new as SC
    print("[case] stages slices / deposit / poisson / field / gather", flush=True)
    n, ns, nz, nx = 40000, 40, 48, 96
    M = _bunch(n, 12)
    seo = cases.oracle_seo()
    ref = O.sc_solve(M.copy(), M[:, 8] != 0, seo, ns, nx, nz, 4.0, 1e13 / n)
    cnt = ref["count"]
    r = M[M[:, 8] != 0, 0]
    z = ref["z"]
    fi = M[M[:, 8] != 0, 4]
    fm = fi - 2 * np.pi * np.floor(fi / (2 * np.pi))           # GetLocalCoordinates_2p5's reduction
    r0 = ref["r0"]
    # slicing and moments
    x = np.zeros(cnt)
    r2d, z2d, x2d = (np.zeros((ns, cnt)) for _ in range(3))
    sl = np.zeros(cnt, np.int32)
    out = SC.ComputeSliceBoundsAndCoordsSBS_merged(r, z, fm, r0, x, ref["slice_step"], ref["fi_s"], ns, r2d, z2d, x2d,
                                                   sl)
    xg, _, _, _, counts, slg, sx, sz, sx2, sz2 = out
    np.testing.assert_array_equal(xg, ref["x"])
    np.testing.assert_array_equal(slg, ref["slice_idx"])
    np.testing.assert_array_equal(counts, ref["counts"])
    np.testing.assert_allclose([sx, sz, sx2, sz2], ref["sums"], rtol=1e-11, atol=1e-17)
    for s in range(ns):
        members = np.flatnonzero(ref["slice_idx"] == s)          # particle order within the slice
        np.testing.assert_array_equal(r2d[s, :members.size], r[members])
        np.testing.assert_array_equal(z2d[s, :members.size], z[members])
        np.testing.assert_array_equal(x2d[s, :members.size], ref["x"][members])
    # deposit: the oracle's fixed-point grid scaled, bit for bit; close to the
    # reference's float32 accumulation
    bl = ref["bunch_length"]
    q = 1e13 / n * O.Q
    rho = SC.DistributeChargeSBS_Njit(z2d, x2d, counts, q, bl, ref["xmin"], ref["xmax"], ref["zmin"], ref["zmax"], nx,
                                      nz, ns)
    want = (ref["grid_fixed"].astype(np.float64) * (ref["qscale"] / float(1 << (2 * ref["bits"])))).astype(np.float32)
    np.testing.assert_array_equal(rho, want)
    ref32 = np.zeros((ns, nz, nx), np.float32)
    order = np.argsort(ref["slice_idx"], kind="stable").astype(np.int64)
    L.ora_deposit_f32(p(ref["x"]), p(z), p(order), i64(cnt), p(ref["slice_idx"]), d(q), d(bl), d(ref["xmin"]),
                      d(ref["xmax"]), d(ref["zmin"]), d(ref["zmax"]), ci(nx), ci(nz), ci(ns), p(ref32))
    assert np.abs(rho - ref32).max() <= 1e-5 * np.abs(ref32).max()
    # degenerate box: zeros
    zero = SC.DistributeChargeSBS_Njit(z2d, x2d, counts, q, bl, 1.0, 1.0, 0.0, 1.0, nx, nz, ns)
    assert not zero.any()
    # Green's function: the reference's float32 arithmetic, bit for bit
    G = SC.make_fft_kernel_open_numba(nx, nz, ref["dx"], ref["dz"])
    Gr = np.zeros((nz, nx), np.float32)
    L.ora_green(ci(nx), ci(nz), d(ref["dx"]), d(ref["dz"]), p(Gr))
    np.testing.assert_array_equal(G, Gr)
    # Poisson: complex64 LDS FFT against the FP64 FFT of the same density
    phi = SC.poisson_solve_slices(rho, ref["dx"], ref["dz"])
    phi_ref = np.real(np.fft.ifft2(np.fft.fft2(rho.astype(np.float64), axes=(1, 2)) * Gr.astype(np.float64)[None],
                                   axes=(1, 2)))
    e = np.abs(phi - phi_ref).max() / np.abs(phi_ref).max()
    print("poisson", e, flush=True)
    assert e < 2e-6
    # field of each slice from the GPU potential
    Ex = np.zeros((ns, nz, nx), np.float32)
    Ez = np.zeros_like(Ex)
    for k in range(ns):
        SC.compute_field_from_potential_numba_SBS(phi[k], k, ref["dx"], ref["dz"], bl, Ex, Ez, ref["mean_gamma_m1"])
    Exr = np.zeros((ns, nz, nx))
    Ezr = np.zeros_like(Exr)
    ph64 = np.ascontiguousarray(phi.astype(np.float64))
    for k in range(ns):
        L.ora_field_from_potential(p(ph64[k]), ci(nz), ci(nx), d(ref["dx"]), d(ref["dz"]), d(bl),
                                   d(ref["mean_gamma_m1"]), p(Exr[k]), p(Ezr[k]))
    for a, b in ((Ex, Exr), (Ez, Ezr)):
        assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max()
    Ex0, Ez0 = Ex.copy(), Ez.copy()
    SC.compute_field_from_potential_numba_SBS(phi[0], 0, 0.0, ref["dz"], bl, Ex0, Ez0, ref["mean_gamma_m1"])
    assert not Ex0[0].any() and not Ez0[0].any() and np.array_equal(Ex0[1:], Ex[1:])
    # longitudinal field
    sgrid = ref["s_grid"]
    qf = counts.astype(np.float64) * q
    Ef = SC.total_field_1d_numba(sgrid, qf)
    Efr = np.zeros(ns)
    L.ora_total_field_1d(p(np.ascontiguousarray(sgrid)), p(qf), ci(ns), p(Efr))
    np.testing.assert_allclose(Ef, Efr, rtol=1e-13, atol=1e-300)
    # gather into the particle matrix
    Mg = M.copy()
    Mr = M.copy()
    ep, ezp, efp = np.zeros(cnt), np.zeros(cnt), np.zeros(cnt)
    idx = ref["idx"].astype(np.int64)
    SC.interpolate_fields_to_particles_SBS_numba(Ex, Ez, Ef, ref["x"], z, ref["pr0"], ref["slice_idx"], idx,
                                                 ref["xmin"], ref["zmin"], ref["dx"], ref["dz"], bl, ep, ezp, efp, Mg)
    L.ora_sc_gather(p(Ex.astype(np.float64)), p(Ez.astype(np.float64)), ci(ns), ci(nz), ci(nx), p(ref["x"]), p(z),
                    p(ref["pr0"]), p(ref["slice_idx"]), p(ref["idx"]), i64(cnt), d(ref["xmin"]), d(ref["zmin"]),
                    d(ref["dx"]), d(ref["dz"]), d(bl), p(Mr))
    np.testing.assert_allclose(Mg[:, 10:13], Mr[:, 10:13], rtol=1e-12, atol=1e-12 * np.abs(Mr[:, 10:13]).max())
    np.testing.assert_allclose(np.stack([ep, ezp, efp], 1), Mr[idx, 10:13], rtol=1e-12,
                               atol=1e-12 * np.abs(Mr[:, 10:13]).max())


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_build_2_5d_tuple_matches_the_oracle(torch_cuda, precision):
    """All 18 elements of Build_2_5D_SliceGrid_sbs (src/FFAG_Bunch.py:1352):
    charge cube, potential, field cubes, box, bunch length, longitudinal field,
    s grid, slice charges, and the survivors' r, unwrapped fi, x, r0, z in the
    input order; for the FP32 tier after an azimuth sort has reordered the
    particles on the device."""
    from pyffa_amd.FFAG_Bunch import FFAG_Bunch
    print(f"[case] Build_2_5D tuple {precision}", flush=True)
    n, ns, nz, nx = 20000, 32, 64, 64
    M = _bunch(n, 13)
    M[:, 4] += 2 * np.pi * 3          # unwrapped azimuths several turns on
    marcosize = 1e13 / n
    B = FFAG_Bunch(M, marcosize=marcosize, sc_grid_size=(ns, nx, nz, 4.0), BunchType="Boris", precision=precision)
    cases.set_bunch_seo(B)
    if precision == "fp32":
        B._uninject_head = n
        B.sort_by_azimuth()
        assert not np.array_equal(B.pid.cpu().numpy(), np.arange(n))
    out = B.Build_2_5D_SliceGrid_sbs(None, None)
    assert len(out) == 18
    res = B.sc_result()
    alive = M[:, 8] != 0
    ref = O.sc_solve(M.copy(), alive, cases.oracle_seo(), ns, nx, nz, 4.0, marcosize,
                     bounds_override=(res["xmin"], res["xmax"], res["zmin"], res["zmax"]))
    (rho, phi, ex, ez, xmin, xmax, zmin, zmax, bl, ef, s_grid, rho_f, cr, cfi, cx, cr0, cz, fi_grid) = out
    np.testing.assert_allclose(rho, ref["rho"].astype(np.float32), rtol=1e-6, atol=1e-6 * np.abs(ref["rho"]).max())
    for a, b in ((phi, ref["phi"]), (ex, ref["Ex"]), (ez, ref["Ez"])):
        assert a.dtype == np.float32 and a.shape == (ns, nz, nx)
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max()
    assert (xmin, xmax, zmin, zmax) == (ref["xmin"], ref["xmax"], ref["zmin"], ref["zmax"])
    assert bl == pytest.approx(ref["bunch_length"], rel=1e-12)
    np.testing.assert_allclose(ef, ref["Ef"], rtol=1e-10, atol=1e-12 * np.abs(ref["Ef"]).max())
    np.testing.assert_allclose(s_grid, ref["s_grid"], rtol=1e-12)
    np.testing.assert_array_equal(rho_f, ref["counts"] * (marcosize * O.Q))     # count_slices * q_part
    np.testing.assert_array_equal(cr, M[alive, 0])
    np.testing.assert_array_equal(cfi, M[alive, 4])                  # unwrapped, as the reference returns it
    np.testing.assert_allclose(cx, ref["x"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(cr0, ref["r0"], rtol=1e-12)
    np.testing.assert_array_equal(cz, M[alive, 2])
    np.testing.assert_allclose(fi_grid, np.linspace(ref["fi_s"], ref["fi_e"], ns), rtol=1e-15)
