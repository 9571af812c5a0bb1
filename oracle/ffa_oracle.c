/*
 * ffa_oracle.c -- CPU FP64 restatement of PyFFA's tracking hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline; the product (pyffa_amd/libffa_hip.so) never links or calls it.
 *
 * Every function restates one reference Numba kernel, operating on the same
 * data layout the reference uses: the particle matrix is the (N, 16) row-major
 * float64 ``LocalBunch`` (columns per src/FFAG_Bunch.py:23-62), field-map
 * stacks are (n_max+1, NR, NPHI) C-contiguous float64 (src/FFAG_Field.py:
 * 691-706).  Arithmetic is written in the reference's operation order and the
 * file is compiled with -ffp-contract=off so no fused multiply-adds appear,
 * matching Numba's default (non-fastmath) code generation.  Loops that the
 * reference runs under ``prange`` carry an OpenMP pragma (the CPU baseline);
 * loops the reference runs serially stay serial.
 *
 * Deviation: the reference compiles Bmap_add_all_orders with fastmath=True
 * (src/FFAG_MathTools.py:1139), which lets LLVM reassociate the sums; this
 * restatement keeps the written order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NCOL 16
#define TWO_PI (2.0 * 3.141592653589793)
#define PI_D 3.141592653589793

/* Clamp helper used before float->int casts so that non-finite or huge
 * inputs (lost particles) give a defined index instead of C undefined
 * behaviour; finite in-range values are unaffected. */
static inline double clamp_for_cast(double v) {
    if (!(v > -1e9)) return -1e9;
    if (v > 1e9) return 1e9;
    return v;
}

/* fast_mod_parallel, src/FFAG_MathTools.py:1553-1557 (column view of M). */
void ora_fast_mod(const double *M, int64_t n, double b, double *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double a = M[i * NCOL + 4];
        out[i] = a - b * floor(a / b);
    }
}

/* GetBfields_Njit, src/FFAG_Field.py:1333-1426.  Order-0 analytic field
 * with z-linear Br, Bfi; skips particles that are not injected and alive
 * (their outputs keep stale values, as in the reference). */
void ora_get_bfields(const double *M, int64_t n, const double *fi_vect,
                     double *Bz, double *Br, double *Bfi,
                     const double *s_poly, int n_s, const double *sp_poly, int n_sp,
                     double B0, double k_value, int nsectors,
                     const double *convex, int m_seg) {
    const double sector_period = 2.0 * PI_D / (double)nsectors;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double *row = M + i * NCOL;
        if (row[7] != 1.0 || row[8] != 1.0) continue;
        double ri = row[0], fii = fi_vect[i], zi = row[2];
        double val = 0.0;
        for (int a = 0; a < n_s; ++a) val = val * ri + s_poly[a];
        double spiral = val;
        val = 0.0;
        for (int a = 0; a < n_sp; ++a) val = val * ri + sp_poly[a];
        double spiral_p = val;
        double B0r, B0r_p;
        if (k_value == 0.0) { B0r = B0; B0r_p = 0.0; }
        else { B0r = B0 * pow(ri, k_value); B0r_p = B0 * k_value * pow(ri, k_value - 1.0); }
        double fs = fii - spiral;
        fs -= sector_period * floor(fs / sector_period);
        double F = 0.0, Fp = 0.0;
        for (int j = 0; j < m_seg; ++j) {
            const double *cp = convex + 5 * j;
            double x0 = cp[0], x1 = cp[1], W0 = cp[2], y01 = cp[3], y02 = cp[4];
            double yc = y02 * fabs(y01);
            double ks = 16.0 / W0;
            double s1 = 1.0 / (1.0 + exp(-ks * (fs - x0)));
            double s2 = 1.0 / (1.0 + exp(ks * (fs - x1)));
            F += yc * s1 * s2;
            double s1p = ks * s1 * (1.0 - s1);
            double s2p = -ks * s2 * (1.0 - s2);
            Fp += yc * (s1p * s2 + s1 * s2p);
        }
        Bz[i] = B0r * F;
        Bfi[i] = B0r * Fp / ri * zi;
        Br[i] = (B0r_p * F - B0r * Fp * spiral_p) * zi;
    }
}

/* Bmap_add_all_orders, src/FFAG_MathTools.py:1139-1188.  Bilinear gather of
 * Taylor orders 1..max_order; index by int() truncation; all particles. */
void ora_bmap_add_all_orders(double X0, double dX, int64_t NX, double Y0, double dY, int64_t NY,
                             const double *Bz_stack, const double *Bfi_stack, const double *Br_stack,
                             const double *M, int64_t n, const double *fi_vect, int max_order,
                             double *out_Bz, double *out_Bfi, double *out_Br) {
    const int64_t plane = NX * NY;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double x = M[k * NCOL + 0], y = fi_vect[k];
        int64_t i = (int64_t)clamp_for_cast((x - X0) / dX);
        int64_t j = (int64_t)clamp_for_cast((y - Y0) / dY);
        if (i < 0) i = 0; else if (i >= NX - 1) i = NX - 2;
        if (j < 0) j = 0; else if (j >= NY - 1) j = NY - 2;
        double tx = (x - (X0 + (double)i * dX)) / dX;
        double ty = (y - (Y0 + (double)j * dY)) / dY;
        double w00 = (1.0 - tx) * (1.0 - ty), w01 = (1.0 - tx) * ty;
        double w10 = tx * (1.0 - ty), w11 = tx * ty;
        int64_t i1 = i + 1, j1 = j + 1;
        double z = M[k * NCOL + 2];
        double z2 = z * z, z_even = z2, z_odd = z_even * z;
        for (int nn = 1; nn <= max_order; ++nn) {
            const double *a = Bz_stack + nn * plane, *b = Bfi_stack + nn * plane, *c = Br_stack + nn * plane;
            double bz = w00 * a[i * NY + j] + w01 * a[i * NY + j1] + w10 * a[i1 * NY + j] + w11 * a[i1 * NY + j1];
            double bfi = w00 * b[i * NY + j] + w01 * b[i * NY + j1] + w10 * b[i1 * NY + j] + w11 * b[i1 * NY + j1];
            double br = w00 * c[i * NY + j] + w01 * c[i * NY + j1] + w10 * c[i1 * NY + j] + w11 * c[i1 * NY + j1];
            out_Bz[k] += bz * z_even;
            out_Bfi[k] += bfi * z_odd;
            out_Br[k] += br * z_odd;
            if (nn != max_order) { z_even *= z2; z_odd *= z2; }
        }
    }
}

/* bharmonics_kernel_1d, src/FFAG_Field.py:1541-1587 (serial in the
 * reference).  Uses the raw (unwrapped) azimuth; Br, Bfi terms are zero. */
void ora_bharmonics(const double *M, int64_t n, int nh, const int64_t *m_arr, const double *amp,
                    const double *phase, const double *rmin, const double *rmax,
                    double *Bz, double *Br, double *Bfi) {
    for (int64_t i = 0; i < n; ++i) {
        double r = M[i * NCOL + 0], phi = M[i * NCOL + 4];
        if (r <= 0.0) continue;
        double bz = 0.0, br = 0.0, bfi = 0.0;
        for (int k = 0; k < nh; ++k)
            if (rmin[k] <= r && r <= rmax[k]) bz += amp[k] * cos((double)m_arr[k] * phi + phase[k]);
        Bz[i] += bz; Br[i] += br; Bfi[i] += bfi;
    }
}

/* emap_kernel, src/FFAG_Utils.py:984-1036, one call per gap (the caller
 * loops gaps in order, src/FFAG_Field.py:284-306, so a later gap zeroes the
 * field of particles outside it). */
void ora_emap_gap(const double *M, int64_t n, double acc_Etot, double theta_n,
                  const double *poly_x, const double *poly_y, int nv,
                  double bx0, double bx1, double by0, double by1,
                  double *Er, double *Efi, double *Ez, int32_t *gap_idx, int gap_index) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double r = M[k * NCOL + 0], fi = M[k * NCOL + 4];
        double x = r * cos(fi), y = r * sin(fi);
        if (x < bx0 || x > bx1 || y < by0 || y > by1) { Er[k] = 0.0; Efi[k] = 0.0; Ez[k] = 0.0; continue; }
        int inside = 1;
        double prev = 0.0;
        for (int j = 0; j < nv; ++j) {
            double x0 = poly_x[j], y0 = poly_y[j];
            double x1 = poly_x[(j + 1) % nv], y1 = poly_y[(j + 1) % nv];
            double cross = (x1 - x0) * (y - y0) - (y1 - y0) * (x - x0);
            if (j == 0) prev = cross;
            else if (cross * prev < 0.0) { inside = 0; break; }
        }
        if (inside) {
            double dth = theta_n - fi;
            Er[k] = acc_Etot * cos(dth);
            Efi[k] = acc_Etot * sin(dth);
            Ez[k] = 0.0;
            gap_idx[k] = gap_index;
        } else {
            Er[k] = 0.0; Efi[k] = 0.0; Ez[k] = 0.0; gap_idx[k] = 0;
        }
    }
}

/* FFAG_EField_spiral.Interpolation2D_EMap, src/FFAG_Field.py:464-498:
 * interpolate_spiral_emap_unique_gap_kernel, shift_to_nearest_window and
 * Bilinear_interp_2D_vect_uniform_3fields, src/FFAG_MathTools.py:939-1135.
 * maps: (nr, nphi) row-major Er, Ez/z, Ephi.  A point outside every gap
 * window gets zero fields and gap index 0 (the reference's -1 selects an
 * rf_shift entry that multiplies a zero field). */
void ora_emap_spiral(const double *M, int64_t n, const double *gap_az, int n_gap, double rf_amp, double r0,
                     double dr, int64_t nr, double phi0, double dphi, int64_t nphi, const double *Er_map,
                     const double *Ez_map, const double *Ephi_map, double *Er, double *Ez, double *Efi,
                     int32_t *gap_idx) {
    const double twopi = 2.0 * 3.141592653589793;
    const double phi_min = phi0, phi_max = phi0 + dphi * (double)(nphi - 1);
    const double center = 0.5 * (phi_min + phi_max);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        const double r = M[k * NCOL + 0], z = M[k * NCOL + 2], fi = M[k * NCOL + 4];
        Er[k] = 0.0; Ez[k] = 0.0; Efi[k] = 0.0; gap_idx[k] = 0;
        int hit = -1;
        double loc = 0.0;
        for (int g = 0; g < n_gap; ++g) {
            const double ph = fi - gap_az[g];
            const double kk = nearbyint((center - ph) / twopi);      /* np.round: half to even */
            const double pl = ph + kk * twopi;
            if (pl >= phi_min && pl <= phi_max) { hit = g; loc = pl; break; }
        }
        if (hit < 0) continue;
        int64_t i = (int64_t)floor(clamp_for_cast((r - r0) / dr));
        int64_t j = (int64_t)floor(clamp_for_cast((loc - phi0) / dphi));
        if (i < 0) i = 0; else if (i >= nr - 1) i = nr - 2;
        if (j < 0) j = 0; else if (j >= nphi - 1) j = nphi - 2;
        const double xx0 = r0 + (double)i * dr, yy0 = phi0 + (double)j * dphi;
        const double tx = (r - xx0) / dr, ty = (loc - yy0) / dphi;
        const double w00 = (1.0 - tx) * (1.0 - ty), w01 = (1.0 - tx) * ty, w10 = tx * (1.0 - ty), w11 = tx * ty;
        const int64_t a = i * nphi + j, b = (i + 1) * nphi + j;
        const double er = w00 * Er_map[a] + w01 * Er_map[a + 1] + w10 * Er_map[b] + w11 * Er_map[b + 1];
        const double ez = w00 * Ez_map[a] + w01 * Ez_map[a + 1] + w10 * Ez_map[b] + w11 * Ez_map[b + 1];
        const double ef = w00 * Ephi_map[a] + w01 * Ephi_map[a + 1] + w10 * Ephi_map[b] + w11 * Ephi_map[b + 1];
        Er[k] = rf_amp * er;
        Ez[k] = rf_amp * ez * z;
        Efi[k] = rf_amp * ef;
        gap_idx[k] = hit;
    }
}

/* cylindrical_to_cartesian, src/FFAG_MathTools.py:1969-2014.  cart holds
 * 12 planes of n: x y z vx vy vz Ex Ey Ez Bx By Bz. */
void ora_cyl2cart(const double *M, int64_t n, const double *E_r, const double *E_z, const double *E_fi,
                  const double *Br, const double *Bz0, const double *Bfi,
                  const int32_t *gap_idx, const double *rf_shift, int apply_rf, double *cart) {
    double *X = cart, *Y = cart + n, *Z = cart + 2 * n, *VX = cart + 3 * n, *VY = cart + 4 * n, *VZ = cart + 5 * n;
    double *EX = cart + 6 * n, *EY = cart + 7 * n, *EZ = cart + 8 * n, *BX = cart + 9 * n, *BY = cart + 10 * n, *BZ = cart + 11 * n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double *row = M + i * NCOL;
        double r = row[0], rdot = row[1], z = row[2], zdot = row[3], fi = row[4], fidot = row[5];
        double sinFi = sin(fi), cosFi = cos(fi);
        double sin_RF = apply_rf ? sin(row[9] + rf_shift[gap_idx[i]]) : 0.0;
        double ez = E_z[i] * sin_RF, er = E_r[i] * sin_RF, efi = E_fi[i] * sin_RF;
        ez += row[11]; er += row[10]; efi += row[12];
        X[i] = r * cosFi; Y[i] = r * sinFi; Z[i] = z;
        VX[i] = rdot * cosFi - r * fidot * sinFi;
        VY[i] = rdot * sinFi + r * fidot * cosFi;
        VZ[i] = zdot;
        EX[i] = (er * cosFi - efi * sinFi);
        EY[i] = (er * sinFi + efi * cosFi);
        EZ[i] = ez;
        BX[i] = -(Br[i] * cosFi - Bfi[i] * sinFi);
        BY[i] = -(Br[i] * sinFi + Bfi[i] * cosFi);
        BZ[i] = -Bz0[i];
    }
}

/* boris_push_cartesian_njit, src/FFAG_MathTools.py:1877-1966.  Writes the
 * pushed coordinates into out (6 planes); clears the survive column of M on
 * an aperture loss but still pushes the particle this step. */
void ora_boris(double *M, int64_t n, const double *cart, double dt, double q, double m, double c,
               int aperture_enable, const double *aperture, double *out) {
    const double *X = cart, *Y = cart + n, *Z = cart + 2 * n, *VX = cart + 3 * n, *VY = cart + 4 * n, *VZ = cart + 5 * n;
    const double *EX = cart + 6 * n, *EY = cart + 7 * n, *EZ = cart + 8 * n, *BX = cart + 9 * n, *BY = cart + 10 * n, *BZ = cart + 11 * n;
    double *XO = out, *YO = out + n, *ZO = out + 2 * n, *VXO = out + 3 * n, *VYO = out + 4 * n, *VZO = out + 5 * n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *row = M + i * NCOL;
        if (row[7] != 1.0 || row[8] != 1.0) {
            XO[i] = X[i]; YO[i] = Y[i]; ZO[i] = Z[i]; VXO[i] = VX[i]; VYO[i] = VY[i]; VZO[i] = VZ[i];
            continue;
        }
        if (aperture_enable) {
            double r = row[0], z = Z[i];
            if (r > aperture[1] || r < aperture[0] || z > aperture[3] || z < aperture[2]) row[8] = 0.0;
        }
        double vx = VX[i], vy = VY[i], vz = VZ[i];
        double v2 = vx * vx + vy * vy + vz * vz;
        double gamma = 1.0 / sqrt(1.0 - v2 / (c * c));
        double px = gamma * m * vx, py = gamma * m * vy, pz = gamma * m * vz;
        px += 0.5 * dt * q * EX[i]; py += 0.5 * dt * q * EY[i]; pz += 0.5 * dt * q * EZ[i];
        double p2 = px * px + py * py + pz * pz;
        double gm = sqrt(1.0 + p2 / (m * m * c * c));
        double cb = 0.5 * dt * q / (gm * m);
        double tx = cb * BX[i], ty = cb * BY[i], tz = cb * BZ[i];
        double pxp = px + (py * tz - pz * ty);
        double pyp = py + (pz * tx - px * tz);
        double pzp = pz + (px * ty - py * tx);
        double t2 = tx * tx + ty * ty + tz * tz;
        double s = 2.0 / (1.0 + t2);
        double sx = s * tx, sy = s * ty, sz = s * tz;
        px += (pyp * sz - pzp * sy);
        py += (pzp * sx - pxp * sz);
        pz += (pxp * sy - pyp * sx);
        px += 0.5 * dt * q * EX[i]; py += 0.5 * dt * q * EY[i]; pz += 0.5 * dt * q * EZ[i];
        double p2n = px * px + py * py + pz * pz;
        double gn = sqrt(1.0 + p2n / (m * m * c * c));
        double vxn = px / (gn * m), vyn = py / (gn * m), vzn = pz / (gn * m);
        XO[i] = X[i] + vxn * dt; YO[i] = Y[i] + vyn * dt; ZO[i] = Z[i] + vzn * dt;
        VXO[i] = vxn; VYO[i] = vyn; VZO[i] = vzn;
    }
}

/* convert_cartesian_to_cylindrical, src/FFAG_MathTools.py:2017-2060.  Runs
 * for every particle; unwraps the azimuth with atan2 differences. */
void ora_cart2cyl(double *M, int64_t n, const double *outp, const double *cart, double dRF) {
    const double *XN = outp, *YN = outp + n, *ZN = outp + 2 * n, *VXN = outp + 3 * n, *VYN = outp + 4 * n, *VZN = outp + 5 * n;
    const double *XO = cart, *YO = cart + n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *row = M + i * NCOL;
        double xn = XN[i], yn = YN[i];
        double r_new = sqrt(xn * xn + yn * yn);
        double fb = atan2(YO[i], XO[i]);
        double fa = atan2(yn, xn);
        double d = fa - fb;
        if (d > PI_D) d -= 2.0 * PI_D;
        else if (d < -PI_D) d += 2.0 * PI_D;
        double fi_new = row[4] + d;
        double vr = VXN[i] * cos(fi_new) + VYN[i] * sin(fi_new);
        double vfi = -VXN[i] * sin(fi_new) + VYN[i] * cos(fi_new);
        row[0] = r_new; row[1] = vr; row[2] = ZN[i]; row[3] = VZN[i]; row[4] = fi_new;
        row[5] = r_new > 1e-12 ? vfi / r_new : 0.0;
        row[9] += dRF;
    }
}

/* RK4_equations_of_motion_njit, src/FFAG_MathTools.py:2205-2266, on a
 * stage matrix S (N, 16); writes dxdt columns 0-5 and 9. */
void ora_rk4_rhs(const double *S, int64_t n, const double *E_r, const double *E_z, const double *E_fi,
                 const double *B_r, const double *B_z, const double *B_fi,
                 double q, double c, double E0_J, double *dxdt, double dRF_rate,
                 const double *rf_shift, const int32_t *gap_idx, int apply_rf) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const double *row = S + i * NCOL;
        double *d = dxdt + i * NCOL;
        if (row[7] != 1.0 || row[8] != 1.0) { d[9] = dRF_rate; continue; }
        double r = row[0], rdot = row[1], zdot = row[3], Et = row[5];
        double ig = E0_J / Et;
        double inv_gamma2 = ig * ig;
        double beta2 = 1.0 - inv_gamma2;
        double v2 = beta2 * (c * c);
        double vperp2 = v2 - rdot * rdot - zdot * zdot;
        double fidot = sqrt(vperp2) / r;
        double sin_RF = apply_rf ? sin(row[9] + rf_shift[gap_idx[i]]) : 0.0;
        double ez = E_z[i] * sin_RF, er = E_r[i] * sin_RF, efi = E_fi[i] * sin_RF;
        ez += row[11]; er += row[10]; efi += row[12];
        double Edv = rdot * er + (r * fidot) * efi + zdot * ez;
        d[0] = rdot;
        d[1] = (r * fidot * fidot + q * (c * c) / Et * (er - r * fidot * B_z[i] + zdot * B_fi[i])
                - q * rdot / Et * Edv);
        d[2] = zdot;
        d[3] = (q * (c * c) / Et * (ez + r * fidot * B_r[i] - rdot * B_fi[i]) - q * zdot / Et * Edv);
        d[4] = fidot;
        d[5] = q * Edv;
        d[9] = dRF_rate;
    }
}

/* RK4_Post_Step_njit, src/FFAG_MathTools.py:2269-2299. */
void ora_rk4_post(double *M, int64_t n, const double *dstep) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double *row = M + i * NCOL;
        const double *d = dstep + i * NCOL;
        if (row[7] != 1.0 || row[8] != 1.0) { row[9] += d[9]; continue; }
        for (int c = 0; c < 6; ++c) row[c] += d[c];
        row[9] += d[9];
    }
}

/* ---------------------------------------------------------------------- */
/* 2.5D space charge (src/FFAG_SC_new.py)                                   */
/* ---------------------------------------------------------------------- */

/* GetLocalCoordinates_2p5 (_rk when rk != 0), src/FFAG_SC_new.py:74-148.
 * out: 6 planes of n (r, fi, z, gamma, fiMod, idx-as-double); returns count,
 * *sum_gamma the sequential sum. */
int64_t ora_sc_local_coords(const double *M, int64_t n, const uint8_t *step_survive, double c,
                            int rk, double *out, double *sum_gamma) {
    double *R = out, *FI = out + n, *Z = out + 2 * n, *G = out + 3 * n, *FM = out + 4 * n, *IX = out + 5 * n;
    const double E0J = 938.2723e6 * 1.60217662e-19;
    int64_t j = 0;
    double sg = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!step_survive[i]) continue;
        const double *row = M + i * NCOL;
        double r = row[0], drdt = row[1], z = row[2], dzdt = row[3], fi = row[4];
        double gamma;
        if (rk) {
            gamma = row[5] / E0J;
        } else {
            double rd = r * row[5];
            double v2 = drdt * drdt + dzdt * dzdt + rd * rd;
            gamma = 1.0 / sqrt(1.0 - v2 / (c * c));
        }
        R[j] = r; Z[j] = z; FI[j] = fi; G[j] = gamma;
        FM[j] = fi - TWO_PI * floor(fi / TWO_PI);
        IX[j] = (double)i;
        sg += gamma;
        ++j;
    }
    *sum_gamma = sg;
    return j;
}

/* divide_arc2, src/FFAG_SC_new.py:534-566. res = (arc_start, arc_end, span). */
void ora_divide_arc2(const double *fi, int64_t n, double *res) {
    const double shifts[4] = {0.0, PI_D, PI_D * 0.5, PI_D * 1.5};
    double mn[4], mx[4];
    for (int k = 0; k < 4; ++k) { mn[k] = 1e30; mx[k] = -1e30; }
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) {
            double a = fi[i] + shifts[k], b = 2.0 * PI_D;
            double p = a - b * floor(a / b);
            if (p < mn[k]) mn[k] = p;
            if (p > mx[k]) mx[k] = p;
        }
    int best = 0;
    double bs = mx[0] - mn[0];
    for (int k = 1; k < 4; ++k) if (mx[k] - mn[k] < bs) { bs = mx[k] - mn[k]; best = k; }
    res[0] = mn[best] - shifts[best];
    res[1] = mx[best] - shifts[best];
    res[2] = bs;
}

/* Bilinear_interp_2D_vect_uniform, src/FFAG_MathTools.py:827-936 (serial). */
void ora_bilinear_uniform(double X0, double dX, int64_t NX, double Y0, double dY, int64_t NY,
                          const double *Zt, const double *Xi, const double *Yi, int64_t n, double *values) {
    for (int64_t k = 0; k < n; ++k) {
        double x = Xi[k], y = Yi[k];
        int64_t i = (int64_t)floor(clamp_for_cast((x - X0) / dX));
        int64_t j = (int64_t)floor(clamp_for_cast((y - Y0) / dY));
        if (i < 0) i = 0; else if (i >= NX - 1) i = NX - 2;
        if (j < 0) j = 0; else if (j >= NY - 1) j = NY - 2;
        double xx0 = X0 + (double)i * dX, yy0 = Y0 + (double)j * dY;
        double tx = (x - xx0) / dX, ty = (y - yy0) / dY;
        double z00 = Zt[i * NY + j], z01 = Zt[i * NY + j + 1], z10 = Zt[(i + 1) * NY + j], z11 = Zt[(i + 1) * NY + j + 1];
        values[k] = (1 - tx) * (1 - ty) * z00 + (1 - tx) * ty * z01 + tx * (1 - ty) * z10 + tx * ty * z11;
    }
}

/* ComputeSliceBoundsAndCoordsSBS_merged, src/FFAG_SC_new.py:449-531, without
 * the (ns x N) per-slice copies: x and the slice index are returned per
 * particle and the per-slice lists are recovered by a stable order.
 * sums = (sum_x, sum_z, sum_x2, sum_z2), sequential as in the reference. */
void ora_sc_slices(const double *r, const double *z, const double *fi_mod, const double *r0, int64_t n,
                   double slice_step, double fi_s, int ns, double *x_out, int32_t *slice_idx,
                   int32_t *counts, double *sums) {
    double sx = 0.0, sz = 0.0, sx2 = 0.0, sz2 = 0.0;
    for (int s = 0; s < ns; ++s) counts[s] = 0;
    for (int64_t i = 0; i < n; ++i) {
        double x = r[i] - r0[i];
        double zz = z[i];
        x_out[i] = x;
        sx += x; sz += zz; sx2 += x * x; sz2 += zz * zz;
        double fi = fi_mod[i];
        if (fi < fi_s) fi += TWO_PI;
        int64_t s = (int64_t)clamp_for_cast((fi - fi_s) / slice_step);
        if (s < 0) s = 0; else if (s >= ns) s = ns - 1;
        slice_idx[i] = (int32_t)s;
        counts[s] += 1;
    }
    sums[0] = sx; sums[1] = sz; sums[2] = sx2; sums[3] = sz2;
}

/* Shared CIC cell/weight computation of DistributeChargeSBS_Njit,
 * src/FFAG_SC_new.py:244-312. */
static inline void cic_cell(double x, double z, double xmin, double zmin, double inv_dx, double inv_dz,
                            int nx, int nz, int *ix, int *iz, double *fx, double *fz) {
    double rx = (x - xmin) * inv_dx, rz = (z - zmin) * inv_dz;
    int64_t jx = (int64_t)floor(clamp_for_cast(rx)), jz = (int64_t)floor(clamp_for_cast(rz));
    double gx = rx - (double)jx, gz = rz - (double)jz;
    if (jx < 0) { jx = 0; gx = 0.0; } else if (jx >= nx - 1) { jx = nx - 2; gx = 1.0; }
    if (jz < 0) { jz = 0; gz = 0.0; } else if (jz >= nz - 1) { jz = nz - 2; gz = 1.0; }
    *ix = (int)jx; *iz = (int)jz; *fx = gx; *fz = gz;
}

/* DistributeChargeSBS_Njit as the reference runs it: float32 grid
 * accumulated slice by slice, particles in their original order within a
 * slice.  order[] lists particle indices grouped by slice (stable). */
void ora_deposit_f32(const double *x, const double *z, const int64_t *order, int64_t n,
                     const int32_t *slice_idx, double q_macro, double bunch_len,
                     double xmin, double xmax, double zmin, double zmax, int nx, int nz, int ns, float *rho) {
    memset(rho, 0, sizeof(float) * (size_t)ns * nz * nx);
    double dx = (xmax - xmin) / nx, dz = (zmax - zmin) / nz;
    if (dx <= 0.0 || dz <= 0.0 || bunch_len <= 0.0) return;
    double inv_dx = 1.0 / dx, inv_dz = 1.0 / dz;
    double q = (q_macro / bunch_len * ns) * (1.0 / (dx * dz));
    for (int64_t t = 0; t < n; ++t) {
        int64_t i = order[t];
        int ix, iz; double fx, fz;
        cic_cell(x[i], z[i], xmin, zmin, inv_dx, inv_dz, nx, nz, &ix, &iz, &fx, &fz);
        float *g = rho + (size_t)slice_idx[i] * nz * nx;
        double w00 = (1.0 - fx) * (1.0 - fz) * q, w10 = (fx) * (1.0 - fz) * q;
        double w01 = (1.0 - fx) * (fz) * q, w11 = (fx) * (fz) * q;
        g[iz * nx + ix] = (float)((double)g[iz * nx + ix] + w00);
        g[iz * nx + ix + 1] = (float)((double)g[iz * nx + ix + 1] + w10);
        g[(iz + 1) * nx + ix] = (float)((double)g[(iz + 1) * nx + ix] + w01);
        g[(iz + 1) * nx + ix + 1] = (float)((double)g[(iz + 1) * nx + ix + 1] + w11);
    }
}

/* Fixed-point restatement of the same CIC deposit (the MI355X path's
 * reproducible form): each particle's axis fractions are rounded to
 * ``bits`` fractional bits, a = rint(fx 2^bits), b = rint(fz 2^bits), and the
 * four corners receive the exact integer products (2^bits-a)(2^bits-b),
 * a(2^bits-b), (2^bits-a)b, ab -- 2^(2 bits) units per particle.  Also counts
 * particles per CIC base cell.  Integer sums are order-independent, so this
 * grid is bit-exact against any summation order. */
void ora_deposit_fixed(const double *x, const double *z, int64_t n, const int32_t *slice_idx,
                       double xmin, double xmax, double zmin, double zmax, int nx, int nz, int ns,
                       int bits, int64_t *grid, int32_t *cell_count) {
    memset(grid, 0, sizeof(int64_t) * (size_t)ns * nz * nx);
    memset(cell_count, 0, sizeof(int32_t) * (size_t)ns * nz * nx);
    double dx = (xmax - xmin) / nx, dz = (zmax - zmin) / nz;
    if (dx <= 0.0 || dz <= 0.0) return;
    double inv_dx = 1.0 / dx, inv_dz = 1.0 / dz;
    const int64_t one = (int64_t)1 << bits;
    const double scale = (double)one;
    for (int64_t i = 0; i < n; ++i) {
        int ix, iz; double fx, fz;
        cic_cell(x[i], z[i], xmin, zmin, inv_dx, inv_dz, nx, nz, &ix, &iz, &fx, &fz);
        int64_t a = (int64_t)rint(fx * scale), b = (int64_t)rint(fz * scale);
        int64_t *g = grid + (size_t)slice_idx[i] * nz * nx;
        g[iz * nx + ix] += (one - a) * (one - b);
        g[iz * nx + ix + 1] += a * (one - b);
        g[(iz + 1) * nx + ix] += (one - a) * b;
        g[(iz + 1) * nx + ix + 1] += a * b;
        cell_count[(size_t)slice_idx[i] * nz * nx + iz * nx + ix] += 1;
    }
}

/* make_fft_kernel_open_numba, src/FFAG_SC_new.py:731-760: periodic Green's
 * function 1/(eps0 k^2) with k=0 set to 0, in the reference's float32
 * arithmetic. */
void ora_green(int nx, int nz, double dx, double dz, float *kernel) {
    const double eps0 = 8.854187817e-12;
    float dkx = (float)(2.0 * PI_D / (nx * dx));
    float dkz = (float)(2.0 * PI_D / (nz * dz));
    float *kx = (float *)malloc(sizeof(float) * nx), *kz = (float *)malloc(sizeof(float) * nz);
    for (int i = 0; i < nx; ++i) { int idx = i <= nx / 2 ? i : i - nx; kx[i] = (float)((double)dkx * idx); }
    for (int j = 0; j < nz; ++j) { int idz = j <= nz / 2 ? j : j - nz; kz[j] = (float)((double)dkz * idz); }
    for (int j = 0; j < nz; ++j) {
        float kz2 = kz[j] * kz[j];
        for (int i = 0; i < nx; ++i) {
            float k2 = kx[i] * kx[i] + kz2;
            kernel[j * nx + i] = k2 == 0.0f ? 0.0f : (float)(1.0 / (eps0 * (double)k2));
        }
    }
    free(kx); free(kz);
}

/* compute_field_from_potential_numba_SBS, src/FFAG_SC_new.py:672-728, for one
 * slice; phi (nz, nx) float64 here, Ex/Ez float64 outputs. */
void ora_field_from_potential(const double *phi, int nz, int nx, double dr, double dz, double bunch_len,
                              double mean_gamma_m1, double *Ex, double *Ez) {
    double gamma = mean_gamma_m1 + 1.0;
    double ig2 = 1.0 / (gamma * gamma);
    if (dr == 0.0 || dz == 0.0 || bunch_len == 0.0) {
        for (int t = 0; t < nz * nx; ++t) { Ex[t] = 0.0; Ez[t] = 0.0; }
        return;
    }
    double i2r = 0.5 / dr, i2z = 0.5 / dz;
#define P(j, i) phi[(j) * nx + (i)]
    for (int j = 1; j < nz - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            Ex[j * nx + i] = -(P(j, i + 1) - P(j, i - 1)) * i2r * ig2;
            Ez[j * nx + i] = -(P(j + 1, i) - P(j - 1, i)) * i2z * ig2;
        }
    for (int j = 0; j < nz; ++j) {
        int jp = j + 1 < nz - 1 ? j + 1 : nz - 1, jm = j - 1 > 0 ? j - 1 : 0;
        Ex[j * nx + 0] = -(P(j, 1) - P(j, 0)) / dr * ig2;
        Ez[j * nx + 0] = -(P(jp, 0) - P(jm, 0)) / dz * ig2;
        Ex[j * nx + nx - 1] = -(P(j, nx - 1) - P(j, nx - 2)) / dr * ig2;
        Ez[j * nx + nx - 1] = -(P(jp, nx - 1) - P(jm, nx - 1)) / dz * ig2;
    }
    for (int i = 1; i < nx - 1; ++i) {
        Ex[i] = -(P(0, i + 1) - P(0, i - 1)) * i2r * ig2;
        Ez[i] = -(P(1, i) - P(0, i)) / dz * ig2;
        Ex[(nz - 1) * nx + i] = -(P(nz - 1, i + 1) - P(nz - 1, i - 1)) * i2r * ig2;
        Ez[(nz - 1) * nx + i] = -(P(nz - 1, i) - P(nz - 2, i)) / dz * ig2;
    }
#undef P
}

/* total_field_1d_numba, src/FFAG_SC_new.py:763-782. */
void ora_total_field_1d(const double *s, const double *q, int n, double *E) {
    const double eps0 = 8.854187817e-12;
    double K = 1.0 / (4.0 * PI_D * eps0);
    for (int i = 0; i < n; ++i) E[i] = 0.0;
    if (n < 2 || s[1] - s[0] <= 0.0) return;
    for (int i = 0; i < n; ++i) {
        double Ei = 0.0;
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            double r = s[i] - s[j];
            double sg = r > 0 ? 1.0 : (r < 0 ? -1.0 : 0.0);
            Ei += q[j] * sg / (r * r);
        }
        E[i] = K * Ei;
    }
}

/* interpolate_fields_to_particles_SBS_numba, src/FFAG_SC_new.py:870-972.
 * Ex/Ez cubes (ns, nz, nx) float64; writes M columns 10-12 of the survivor
 * rows listed in idx. */
void ora_sc_gather(const double *Exc, const double *Ezc, int ns, int nz, int nx,
                   const double *x, const double *z, const double *pr0, const int32_t *slice_idx,
                   const double *idx, int64_t n, double x0, double z0, double dx, double dz,
                   double bunch_len, double *M) {
    int bad = (dx <= 0.0) || (dz <= 0.0) || (bunch_len <= 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        double *row = M + (int64_t)idx[k] * NCOL;
        if (bad) { row[10] = 0.0; row[11] = 0.0; row[12] = 0.0; continue; }
        int s = slice_idx[k];
        if (s < 0) s = 0; else if (s >= ns) s = ns - 1;
        const double *Ex = Exc + (size_t)s * nz * nx, *Ez = Ezc + (size_t)s * nz * nx;
        double t = (x[k] - x0) / dx, u = (z[k] - z0) / dz;
        int64_t i0 = (int64_t)floor(clamp_for_cast(t)), j0 = (int64_t)floor(clamp_for_cast(u));
        if (i0 < 0) i0 = 0; else if (i0 > nx - 2) i0 = nx - 2;
        if (j0 < 0) j0 = 0; else if (j0 > nz - 2) j0 = nz - 2;
        double al = t - (double)i0, be = u - (double)j0;
        double w00 = (1.0 - al) * (1.0 - be), w10 = al * (1.0 - be), w01 = (1.0 - al) * be, w11 = al * be;
        double er = w00 * Ex[j0 * nx + i0] + w10 * Ex[j0 * nx + i0 + 1] + w01 * Ex[(j0 + 1) * nx + i0] + w11 * Ex[(j0 + 1) * nx + i0 + 1];
        double ez = w00 * Ez[j0 * nx + i0] + w10 * Ez[j0 * nx + i0 + 1] + w01 * Ez[(j0 + 1) * nx + i0] + w11 * Ez[(j0 + 1) * nx + i0 + 1];
        double p = pr0[k];
        row[10] = er * (1.0 / sqrt(1.0 + p * p));
        row[11] = ez;
        row[12] = er * (-p / sqrt(1.0 + p * p));
    }
}

/* ---------------------------------------------------------------------- */
/* Multi-step driver: the Boris / RK4 loops of src/FFAG_track.py:297-579 and  */
/* :65-294 without space charge (injection, field evaluation, push, stop      */
/* checks skipped), so long parity runs and the CPU baseline do not pay a     */
/* Python round trip per step.                                                */
/* ---------------------------------------------------------------------- */
typedef struct {
    double r_min, r_step, fi_min, fi_step;
    int64_t r_size, fi_size;
    const double *Bz_stack, *Br_stack, *Bfi_stack;
    int32_t max_order;
    const double *s_poly; int32_t n_s;
    const double *sp_poly; int32_t n_sp;
    double B0, k_value;
    int32_t nsectors;
    const double *convex; int32_t m_seg;
    int32_t n_harm;
    const int64_t *harm_m; const double *harm_amp, *harm_phase, *harm_rmin, *harm_rmax;
    int32_t emap_enable, n_gap;
    double gap_width;
    const double *gap_theta;   /* n_gap */
    const double *gap_px, *gap_py;   /* n_gap x 4 */
    const double *gap_box;     /* n_gap x 4 */
    const double *rf_shift;
    const double *freq_curve; int64_t n_freq;   /* (n_freq, >=3) row-major, ncols below */
    int32_t freq_cols;
    int32_t aperture_enable;
    const double *aperture;
    double q, c, m0, E0;
    /* spiral gaps (gap_model 1): gap_theta holds the gap azimuths */
    int32_t gap_model;
    int64_t emap_nr, emap_nphi;
    double emap_r0, emap_dr, emap_phi0, emap_dphi;
    const double *emap_er, *emap_ez, *emap_ephi;
} ora_machine;

static void freq_at(const ora_machine *m, double t, double *f, double *v) {
    if (!m->freq_curve || m->n_freq < 2) { *f = 0.0; *v = 0.0; return; }
    const double *fc = m->freq_curve;
    int nc = m->freq_cols;
    double X0 = fc[0], dX = fc[nc] - fc[0];
    int64_t n = m->n_freq;
    int64_t i = (int64_t)floor(clamp_for_cast((t - X0) / dX));
    if (i < 0) i = 0; else if (i >= n - 1) i = n - 2;
    double x0 = X0 + (double)i * dX, x1 = x0 + dX;
    *f = fc[i * nc + 1] + (t - x0) / (x1 - x0) * (fc[(i + 1) * nc + 1] - fc[i * nc + 1]);
    *v = fc[i * nc + 2] + (t - x0) / (x1 - x0) * (fc[(i + 1) * nc + 2] - fc[i * nc + 2]);
}

typedef struct {
    double *fi_vect, *Bz, *Br, *Bfi, *Er, *Ez, *Efi, *cart, *out, *dxdt, *stage, *acc;
    int32_t *gap;
} ora_scratch;

static int scratch_alloc(ora_scratch *s, int64_t n) {
    memset(s, 0, sizeof *s);
    s->fi_vect = calloc(n, 8); s->Bz = calloc(n, 8); s->Br = calloc(n, 8); s->Bfi = calloc(n, 8);
    s->Er = calloc(n, 8); s->Ez = calloc(n, 8); s->Efi = calloc(n, 8);
    s->cart = calloc(12 * n, 8); s->out = calloc(6 * n, 8);
    s->dxdt = calloc(NCOL * n, 8); s->stage = calloc(NCOL * n, 8); s->acc = calloc(NCOL * n, 8);
    s->gap = calloc(n, 4);
    return s->fi_vect && s->Bz && s->Br && s->Bfi && s->Er && s->Ez && s->Efi && s->cart && s->out && s->dxdt &&
           s->stage && s->acc && s->gap;
}

static void scratch_free(ora_scratch *s) {
    free(s->fi_vect); free(s->Bz); free(s->Br); free(s->Bfi); free(s->Er); free(s->Ez); free(s->Efi);
    free(s->cart); free(s->out); free(s->dxdt); free(s->stage); free(s->acc); free(s->gap);
}

static void field_eval_c(const ora_machine *m, const double *M, int64_t n, ora_scratch *s, double t, double *f,
                         double *v) {
    ora_fast_mod(M, n, 3.141592653589793 * 2 / (double)m->nsectors, s->fi_vect);
    ora_get_bfields(M, n, s->fi_vect, s->Bz, s->Br, s->Bfi, m->s_poly, m->n_s, m->sp_poly, m->n_sp, m->B0,
                    m->k_value, m->nsectors, m->convex, m->m_seg);
    if (m->max_order >= 1)
        ora_bmap_add_all_orders(m->r_min, m->r_step, m->r_size, m->fi_min, m->fi_step, m->fi_size, m->Bz_stack,
                                m->Bfi_stack, m->Br_stack, M, n, s->fi_vect, m->max_order, s->Bz, s->Bfi, s->Br);
    if (m->n_harm > 0)
        ora_bharmonics(M, n, m->n_harm, m->harm_m, m->harm_amp, m->harm_phase, m->harm_rmin, m->harm_rmax, s->Bz,
                       s->Br, s->Bfi);
    freq_at(m, t, f, v);
    if (!m->emap_enable) {
        memset(s->Er, 0, n * 8); memset(s->Efi, 0, n * 8); memset(s->Ez, 0, n * 8);
    }
    if (m->gap_model == 1) {
        if (m->emap_enable)
            ora_emap_spiral(M, n, m->gap_theta, m->n_gap, *v * (1.0 / m->gap_width), m->emap_r0, m->emap_dr,
                            m->emap_nr, m->emap_phi0, m->emap_dphi, m->emap_nphi, m->emap_er, m->emap_ez,
                            m->emap_ephi, s->Er, s->Ez, s->Efi, s->gap);
        else
            for (int64_t k = 0; k < n; ++k) s->gap[k] = 0;
        return;
    }
    for (int g = 0; g < m->n_gap; ++g)
        ora_emap_gap(M, n, (1.0 / m->gap_width) * *v, m->gap_theta[g] + PI_D / 2.0, m->gap_px + 4 * g,
                     m->gap_py + 4 * g, 4, m->gap_box[4 * g], m->gap_box[4 * g + 1], m->gap_box[4 * g + 2],
                     m->gap_box[4 * g + 3], s->Er, s->Efi, s->Ez, s->gap, g);
}

/* Injection (src/FFAG_Bunch.py:265-303): inj_t is the static sorted
 * injection-time column [ns]; *head is the first not-yet-injected index. */
static void inject_c(double *M, int64_t n, const double *inj_t, int64_t *head, double t, double dt) {
    double hi = t * 1e9, lo = (t - dt) * 1e9;
    int64_t a = *head, b;
    while (a < n && inj_t[a] < lo) ++a;     /* searchsorted side=left */
    b = a;
    while (b < n && inj_t[b] <= hi) ++b;    /* searchsorted side=right */
    if (b > a) {
        for (int64_t i = a; i < b; ++i) { M[i * NCOL + 7] = 1.0; M[i * NCOL + 8] = 1.0; }
        *head = b;
    }
}

/* Runs nsteps Boris (rk4 == 0) or RK4 (rk4 != 0) steps from t0; returns the
 * number of steps run (or -1 on allocation failure). */
int64_t ora_run(const ora_machine *m, double *M, int64_t n, const double *inj_t, int64_t *head, double t0,
                double dt, int64_t nsteps, int rk4) {
    ora_scratch s;
    if (!scratch_alloc(&s, n)) { scratch_free(&s); return -1; }
    double t = t0;
    for (int64_t k = 0; k < nsteps; ++k) {
        inject_c(M, n, inj_t, head, t, dt);
        double f, v;
        if (!rk4) {
            field_eval_c(m, M, n, &s, t, &f, &v);
            double dRF = f * PI_D * 2 * dt;
            ora_cyl2cart(M, n, s.Er, s.Ez, s.Efi, s.Br, s.Bz, s.Bfi, s.gap, m->rf_shift, m->emap_enable, s.cart);
            ora_boris(M, n, s.cart, dt, m->q, m->m0, m->c, m->aperture_enable, m->aperture, s.out);
            ora_cart2cyl(M, n, s.out, s.cart, dRF);
        } else {
            /* rk4_step_malloc: k_i = h * rhs(stage_i), stages r + 0.5 k1, r + 0.5 k2, r + k3 */
            const double tt[4] = {t, t + 0.5 * dt, t + 0.5 * dt, t + dt};
            const double *src = M;
            for (int st = 0; st < 4; ++st) {
                field_eval_c(m, src, n, &s, tt[st], &f, &v);
                ora_rk4_rhs(src, n, s.Er, s.Ez, s.Efi, s.Br, s.Bz, s.Bfi, m->q, m->c, m->E0, s.dxdt, f * PI_D * 2,
                            m->rf_shift, s.gap, m->emap_enable);
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; ++i) {
                    for (int c = 0; c < NCOL; ++c) {
                        double kk = dt * s.dxdt[i * NCOL + c];
                        double *acc = s.acc + i * NCOL + c;
                        /* (k1 + 2 k2 + 2 k3 + k4) / 6, summed left to right */
                        if (st == 0) *acc = kk;
                        else if (st == 3) *acc = *acc + kk;
                        else *acc = *acc + 2 * kk;
                        double base = M[i * NCOL + c];
                        if (st == 0 || st == 1) s.stage[i * NCOL + c] = base + 0.5 * kk;
                        else if (st == 2) s.stage[i * NCOL + c] = base + kk;
                    }
                }
                src = s.stage;
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n * NCOL; ++i) s.acc[i] = s.acc[i] / 6;
            ora_rk4_post(M, n, s.acc);
        }
        t = t + dt;
    }
    scratch_free(&s);
    return nsteps;
}

#include <omp.h>
/* Thread count of the OpenMP loops above (CPU-baseline timing and tests). */
void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ora_get_threads(void) { return omp_get_max_threads(); }
