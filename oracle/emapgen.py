"""Oracle (test infrastructure only): CPU restatement of the reference's
spiral RF-gap map construction (examples/main_Emap.py), used by tests to
check the GPU kernels behind ``ffa_spiral_gap_maps``.  Nothing in the
product imports this module.

* ``nearest``  spiral_distance_direction_rphi_with_flag (:267-387): Newton
               on r for the nearest point of phi = s_ref + a ln(r / r_ref),
               r clamped to [r_min, r_max], the end flags, the distance and
               the unit normal;
* ``maps``     the Gaussian profile along the normal, its (Er, Ephi)
               projection (:390-395, 906-925) and Ez/z from div E = 0 with
               deriv_r / deriv_phi (:428-445, 930-939).

Pinned by tests/test_emap_oracle.py: nodes on the line have distance 0 and
the peak field along the normal; nodes moved along the normal by d have
distance d; the profile integrates to U across the line; a flipped normal
side (a planted slip) fails the orientation check.
"""
import numpy as np


def nearest(R, PHI, r_ref, s_ref, a, r_min, r_max, n_iter=10, side=1.0):
    """(dist, ux, uy, flag) for node arrays R, PHI (continuous azimuth)."""
    one_plus_a2 = 1.0 + a * a
    x_p, y_p = R * np.cos(PHI), R * np.sin(PHI)
    r = np.clip(R, r_min, r_max)
    r = np.where(r <= 0.0, r_min, r)
    active = np.ones(R.shape, bool)
    for _ in range(n_iter):
        s = s_ref + a * np.log(r / r_ref)
        cs, sn = np.cos(s), np.sin(s)
        x, y = r * cs, r * sn
        x1, y1 = cs - a * sn, sn + a * cs
        invr = 1.0 / r
        x2, y2 = -(a * invr) * (sn + a * cs), (a * invr) * (cs - a * sn)
        dx, dy = x - x_p, y - y_p
        g = dx * x1 + dy * y1
        gp = one_plus_a2 + dx * x2 + dy * y2
        stop = active & (np.abs(gp) < 1e-14)
        active &= ~stop
        r_new = np.clip(r - g / np.where(gp == 0, 1.0, gp), r_min, r_max)
        conv = active & (np.abs(r_new - r) < 1e-12 * (1.0 + r))
        r = np.where(active, r_new, r)
        active &= ~conv
        if not active.any():
            break
    eps_end = 1e-12 * (1.0 + r_max)
    flag = np.where(np.abs(r - r_min) < eps_end, 1, np.where(np.abs(r - r_max) < eps_end, 2, 0))
    s = s_ref + a * np.log(r / r_ref)
    cs, sn = np.cos(s), np.sin(s)
    vx, vy = x_p - r * cs, y_p - r * sn
    d = np.sqrt(vx * vx + vy * vy)
    inv_t = 1.0 / np.sqrt(one_plus_a2)
    tx, ty = (cs - a * sn) * inv_t, (sn + a * cs) * inv_t
    return d, -ty * side, tx * side, flag


def maps(r_axis, phi_axis, r_ref, s_ref, a, r_min, r_max, U, delta, n_iter=10, side=1.0):
    """(Er, Ephi, Ez/z, dist, flag) on the [nr][nphi] grid."""
    R, PHI = np.meshgrid(r_axis, phi_axis, indexing="ij")
    d, ux, uy, flag = nearest(R, PHI, r_ref, s_ref, a, r_min, r_max, n_iter, side)
    eh = (U / (delta * np.sqrt(2.0 * np.pi))) * np.exp(-0.5 * (d / delta) ** 2)
    c, s = np.cos(PHI), np.sin(PHI)
    er = eh * (ux * c + uy * s)
    ef = eh * (-ux * s + uy * c)
    er[flag != 0] = 0.0
    ef[flag != 0] = 0.0
    dr, dphi = r_axis[1] - r_axis[0], phi_axis[1] - phi_axis[0]
    der = np.empty_like(er)
    der[1:-1] = (er[2:] - er[:-2]) / (2.0 * dr)
    der[0] = (er[1] - er[0]) / dr
    der[-1] = (er[-1] - er[-2]) / dr
    dph = np.empty_like(ef)
    dph[:, 1:-1] = (ef[:, 2:] - ef[:, :-2]) / (2.0 * dphi)
    dph[:, 0] = (ef[:, 1] - ef[:, 0]) / dphi
    dph[:, -1] = (ef[:, -1] - ef[:, -2]) / dphi
    ez = -(er / R + der + (1.0 / R) * dph)
    return er, ef, ez, d, flag
