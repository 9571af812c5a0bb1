// FP64 push kernels of the tracking hot path (gfx950).
//
// This translation unit is compiled with -ffp-contract=off: every expression
// keeps the reference's operation order without fused multiply-adds, so the
// FP64 path differs from the CPU restatement only by the last-ulp behaviour
// of the transcendental functions.
//
// Layout: one thread per particle, structure-of-arrays FP64 state (each
// array read and written once per step, fully coalesced), one flag byte per
// particle, Taylor maps interleaved [r][phi][order][Bz,Bfi,Br] so the two
// corner pairs of a bilinear gather are two contiguous runs of
// 2*max_order*3 doubles.
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;

struct F64Fields {
    double Bz, Br, Bfi;
};

// Order-0 analytic field, GetBfields_Njit (src/FFAG_Field.py:1333-1426).
__device__ __forceinline__ void order0_f64(const ffa_machine &m, double ri, double fii, double zi,
                                           double &Bz, double &Br, double &Bfi) {
    const double sector_period = 2.0 * FFA_PI / (double)m.nsectors;
    double val = 0.0;
    for (int a = 0; a < m.n_spoly; ++a) val = val * ri + m.s_poly[a];
    const double spiral = val;
    val = 0.0;
    for (int a = 0; a < m.n_sppoly; ++a) val = val * ri + m.sp_poly[a];
    const double spiral_p = val;
    double B0r, B0r_p;
    if (m.k_value == 0.0) {
        B0r = m.B0;
        B0r_p = 0.0;
    } else {
        B0r = m.B0 * pow(ri, m.k_value);
        B0r_p = m.B0 * m.k_value * pow(ri, m.k_value - 1.0);
    }
    double fs = fii - spiral;
    fs -= sector_period * floor(fs / sector_period);
    double F = 0.0, Fp = 0.0;
    for (int j = 0; j < m.n_seg; ++j) {
        const double x0 = m.convex[j][0], x1 = m.convex[j][1], W0 = m.convex[j][2];
        const double yc = m.convex[j][4] * fabs(m.convex[j][3]);
        const double ks = 16.0 / W0;
        const double s1 = 1.0 / (1.0 + exp(-ks * (fs - x0)));
        const double s2 = 1.0 / (1.0 + exp(ks * (fs - x1)));
        F += yc * s1 * s2;
        const double s1p = ks * s1 * (1.0 - s1);
        const double s2p = -ks * s2 * (1.0 - s2);
        Fp += yc * (s1p * s2 + s1 * s2p);
    }
    Bz = B0r * F;
    Bfi = B0r * Fp / ri * zi;
    Br = (B0r_p * F - B0r * Fp * spiral_p) * zi;
}

// Bmap_add_all_orders (src/FFAG_MathTools.py:1139-1188) on the interleaved
// azimuth-major map [phi][r][order][Bz, Bfi, Br],
// with the z powers of order 1 passed in: (z2, z^2, z^3) for the field,
// (z2, zt^2, zt^3) for the coefficient form of ffa_bfield_eval mode 2.
__device__ __forceinline__ void taylor_add_pow_f64(const ffa_machine &m, double x, double y, double z2, double ze0,
                                                   double zo0, double &oBz, double &oBfi, double &oBr) {
    const int NX = m.r_size, NY = m.fi_size, no = m.max_order;
    double fx = (x - m.r_min) / m.r_step;
    double fy = (y - m.fi_min) / m.fi_step;
    // int() truncation; non-finite / huge inputs clamp to the map edge
    fx = fx > -1e9 ? (fx < 1e9 ? fx : 1e9) : -1e9;
    fy = fy > -1e9 ? (fy < 1e9 ? fy : 1e9) : -1e9;
    int64_t i = (int64_t)fx, j = (int64_t)fy;
    if (i < 0) i = 0; else if (i >= NX - 1) i = NX - 2;
    if (j < 0) j = 0; else if (j >= NY - 1) j = NY - 2;
    const double tx = (x - (m.r_min + (double)i * m.r_step)) / m.r_step;
    const double ty = (y - (m.fi_min + (double)j * m.fi_step)) / m.fi_step;
    const double w00 = (1.0 - tx) * (1.0 - ty), w01 = (1.0 - tx) * ty;
    const double w10 = tx * (1.0 - ty), w11 = tx * ty;
    const double *T = reinterpret_cast<const double *>(m.taylor);
    const int64_t stride_j = 3 * (int64_t)no;
    // azimuth-major map: (i, j), (i+1, j) contiguous, then (i, j+1), (i+1, j+1)
    const double *a0 = T + (j * NX + i) * stride_j;
    const double *a1 = T + ((j + 1) * NX + i) * stride_j;
    double ze = ze0, zo = zo0;
    for (int n = 0; n < no; ++n) {
        const double *p00 = a0 + 3 * n, *p10 = a0 + stride_j + 3 * n;
        const double *p01 = a1 + 3 * n, *p11 = a1 + stride_j + 3 * n;
        const double bz = w00 * p00[0] + w01 * p01[0] + w10 * p10[0] + w11 * p11[0];
        const double bfi = w00 * p00[1] + w01 * p01[1] + w10 * p10[1] + w11 * p11[1];
        const double br = w00 * p00[2] + w01 * p01[2] + w10 * p10[2] + w11 * p11[2];
        oBz += bz * ze;
        oBfi += bfi * zo;
        oBr += br * zo;
        if (n != no - 1) { ze *= z2; zo *= z2; }
    }
}

__device__ __forceinline__ void taylor_add_f64(const ffa_machine &m, double x, double y, double z,
                                               double &oBz, double &oBfi, double &oBr) {
    const double z2 = z * z;
    taylor_add_pow_f64(m, x, y, z2, z2, z2 * z, oBz, oBfi, oBr);
}

// bharmonics_kernel_1d (src/FFAG_Field.py:1541-1587): raw azimuth.
__device__ __forceinline__ void harmonics_f64(const ffa_machine &m, double r, double phi, double &Bz) {
    if (r <= 0.0) return;
    double bz = 0.0;
    for (int k = 0; k < m.n_harm; ++k)
        if (m.harm_rmin[k] <= r && r <= m.harm_rmax[k]) bz += m.harm_amp[k] * cos(m.harm_m[k] * phi + m.harm_phase[k]);
    Bz += bz;
}

// emap_kernel over the gaps in order (src/FFAG_Utils.py:984-1036 called per
// gap from src/FFAG_Field.py:296-306): a later gap zeroes the field of a
// particle outside it, so only the last gap can act.
__device__ __forceinline__ void emap_f64(const ffa_machine &m, double x, double y, double fi, double v0,
                                         double &Er, double &Efi, double &Ez, int &gap) {
    const double acc = m.acc_etot * v0;
    for (int g = 0; g < m.n_gap; ++g) {
        if (x < m.gap_box[g][0] || x > m.gap_box[g][1] || y < m.gap_box[g][2] || y > m.gap_box[g][3]) {
            Er = 0.0; Efi = 0.0; Ez = 0.0;
            continue;
        }
        bool inside = true;
        double prev = 0.0;
        for (int j = 0; j < 4; ++j) {
            const double x0 = m.gap_px[g][j], y0 = m.gap_py[g][j];
            const double x1 = m.gap_px[g][(j + 1) & 3], y1 = m.gap_py[g][(j + 1) & 3];
            const double cross = (x1 - x0) * (y - y0) - (y1 - y0) * (x - x0);
            if (j == 0) prev = cross;
            else if (cross * prev < 0.0) { inside = false; break; }
        }
        if (inside) {
            const double dth = (m.gap_theta[g] + FFA_PI / 2.0) - fi;
            Er = acc * cos(dth);
            Efi = acc * sin(dth);
            Ez = 0.0;
            gap = g;
        } else {
            Er = 0.0; Efi = 0.0; Ez = 0.0;
            gap = 0;
        }
    }
}

// FFAG_EField_spiral.Interpolation2D_EMap (src/FFAG_Field.py:464-498):
// interpolate_spiral_emap_unique_gap_kernel with shift_to_nearest_window and
// Bilinear_interp_2D_vect_uniform_3fields (src/FFAG_MathTools.py:939-1135).
// The first gap whose window holds the raw azimuth (moved by the nearest
// whole number of turns) acts; its index selects rf_shift.  Indices are
// clamped to the map, so the weights extrapolate outside it, as there.
__device__ __forceinline__ void emap_spiral_f64(const ffa_machine &m, double r, double fi, double z, double v0,
                                                double &Er, double &Efi, double &Ez, int &gap) {
    Er = 0.0; Efi = 0.0; Ez = 0.0;
    gap = 0;                     // the reference's -1 (no gap): the field is zero either way
    const double twopi = 2.0 * FFA_PI;
    const double phi_min = m.emap_phi0, phi_max = m.emap_phi0 + m.emap_dphi * (double)(m.emap_nphi - 1);
    const double center = 0.5 * (phi_min + phi_max);
    int hit = -1;
    double loc = 0.0;
    for (int g = 0; g < m.n_gap; ++g) {
        const double ph = fi - m.gap_theta[g];
        const double k = rint((center - ph) / twopi);
        const double pl = ph + k * twopi;
        if (pl >= phi_min && pl <= phi_max) { hit = g; loc = pl; break; }
    }
    if (hit < 0) return;
    const int NX = m.emap_nr, NY = m.emap_nphi;
    double fx = (r - m.emap_r0) / m.emap_dr, fy = (loc - m.emap_phi0) / m.emap_dphi;
    fx = fx > -1e9 ? (fx < 1e9 ? fx : 1e9) : -1e9;
    fy = fy > -1e9 ? (fy < 1e9 ? fy : 1e9) : -1e9;
    int64_t i = (int64_t)floor(fx), j = (int64_t)floor(fy);
    if (i < 0) i = 0; else if (i >= NX - 1) i = NX - 2;
    if (j < 0) j = 0; else if (j >= NY - 1) j = NY - 2;
    const double xx0 = m.emap_r0 + (double)i * m.emap_dr, yy0 = m.emap_phi0 + (double)j * m.emap_dphi;
    const double tx = (r - xx0) / m.emap_dr, ty = (loc - yy0) / m.emap_dphi;
    const double w00 = (1.0 - tx) * (1.0 - ty), w01 = (1.0 - tx) * ty;
    const double w10 = tx * (1.0 - ty), w11 = tx * ty;
    const double *E = reinterpret_cast<const double *>(m.emap);
    const double *e00 = E + (i * NY + j) * 3, *e01 = e00 + 3;
    const double *e10 = E + ((i + 1) * NY + j) * 3, *e11 = e10 + 3;
    const double er = w00 * e00[0] + w01 * e01[0] + w10 * e10[0] + w11 * e11[0];
    const double ezc = w00 * e00[1] + w01 * e01[1] + w10 * e10[1] + w11 * e11[1];
    const double eph = w00 * e00[2] + w01 * e01[2] + w10 * e10[2] + w11 * e11[2];
    const double rf_amp = v0 * m.acc_etot;
    Er = rf_amp * er;
    Ez = rf_amp * ezc * z;
    Efi = rf_amp * eph;
    gap = hit;
}

// Field evaluation of one particle (fast_mod + order 0 + Taylor + harmonics
// + gap field), for injected-and-alive particles.
__device__ __forceinline__ void fields_f64(const ffa_machine &m, double r, double fi, double z, double sf,
                                           double cf, double v0, double &Bz, double &Br, double &Bfi,
                                           double &Er, double &Efi, double &Ez, int &gap) {
    const double P = FFA_PI * 2.0 / (double)m.nsectors;
    const double fim = fi - P * floor(fi / P);
    order0_f64(m, r, fim, z, Bz, Br, Bfi);
    if (m.max_order >= 1) taylor_add_f64(m, r, fim, z, Bz, Bfi, Br);
    if (m.n_harm > 0) harmonics_f64(m, r, fi, Bz);
    Er = 0.0; Efi = 0.0; Ez = 0.0;
    if (m.rf_enable) {
        if (m.gap_model == 1) emap_spiral_f64(m, r, fi, z, v0, Er, Efi, Ez, gap);
        else emap_f64(m, r * cf, r * sf, fi, v0, Er, Efi, Ez, gap);
    }
}

__global__ __launch_bounds__(BLOCK) void k_inject(uint8_t *flags, int64_t b, int64_t e) {
    const int64_t i = b + blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i < e) flags[i] = (uint8_t)(flags[i] | FFA_F_INJ | FFA_F_ALIVE);
}

// One FP64 Boris step of one particle: A2 extrapolation + A6..A15, on the
// cylindrical state in place (every particle, live or not, goes through the
// reference's cylindrical -> Cartesian -> cylindrical round trip).  Shared
// by the one-step kernel and the multi-step kernel, so both give the same
// bits (this unit keeps the operation order: no contraction).
// e_pre / e_prepre: the particle's SC endpoint fields (r, z, fi).
__device__ __forceinline__ void boris_step_f64(const ffa_machine &m, const ffa_step &s, double &r, double &vr,
                                               double &z, double &vz, double &fi, double &fidot, double &rf,
                                               uint8_t &fl, const double e_pre[3], const double e_prepre[3]) {
    double esr = e_pre[0], esz = e_pre[1], esf = e_pre[2];
    if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) {
        esr = esr + s.sc_s * (esr - e_prepre[0]);
        esz = esz + s.sc_s * (esz - e_prepre[1]);
        esf = esf + s.sc_s * (esf - e_prepre[2]);
    }
    const bool live = (fl & FFA_F_INJ) && (fl & FFA_F_ALIVE);
    const double sinFi = sin(fi), cosFi = cos(fi);
    double Bz = 0.0, Br = 0.0, Bfi = 0.0, Er = 0.0, Efi = 0.0, Ez = 0.0;
    int gap = 0;
    if (live) fields_f64(m, r, fi, z, sinFi, cosFi, s.v0, Bz, Br, Bfi, Er, Efi, Ez, gap);
    // cylindrical_to_cartesian (src/FFAG_MathTools.py:1969-2014)
    const double sin_RF = m.rf_enable ? sin(rf + m.rf_shift[gap]) : 0.0;
    double ezt = Ez * sin_RF, ert = Er * sin_RF, eft = Efi * sin_RF;
    ezt += esz; ert += esr; eft += esf;
    const double x = r * cosFi, y = r * sinFi;
    const double vx = vr * cosFi - r * fidot * sinFi;
    const double vy = vr * sinFi + r * fidot * cosFi;
    double xn = x, yn = y, zn = z, vxn = vx, vyn = vy, vzn = vz;
    if (live) {
        // boris_push_cartesian_njit (src/FFAG_MathTools.py:1877-1966)
        if (m.aperture_enable &&
            (r > m.aperture[1] || r < m.aperture[0] || z > m.aperture[3] || z < m.aperture[2]))
            fl = (uint8_t)(fl & ~FFA_F_ALIVE);
        const double Ex = ert * cosFi - eft * sinFi, Ey = ert * sinFi + eft * cosFi, Ezc = ezt;
        const double Bx = -(Br * cosFi - Bfi * sinFi), By = -(Br * sinFi + Bfi * cosFi), Bzc = -Bz;
        const double q = m.q, mm = m.m0, c = m.c, dt = s.dt;
        const double v2 = vx * vx + vy * vy + vz * vz;
        const double gamma = 1.0 / sqrt(1.0 - v2 / (c * c));
        double px = gamma * mm * vx, py = gamma * mm * vy, pz = gamma * mm * vz;
        px += 0.5 * dt * q * Ex; py += 0.5 * dt * q * Ey; pz += 0.5 * dt * q * Ezc;
        const double p2 = px * px + py * py + pz * pz;
        const double gm = sqrt(1.0 + p2 / (mm * mm * c * c));
        const double cb = 0.5 * dt * q / (gm * mm);
        const double tx = cb * Bx, ty = cb * By, tz = cb * Bzc;
        const double pxp = px + (py * tz - pz * ty);
        const double pyp = py + (pz * tx - px * tz);
        const double pzp = pz + (px * ty - py * tx);
        const double t2 = tx * tx + ty * ty + tz * tz;
        const double sc = 2.0 / (1.0 + t2);
        const double sx = sc * tx, sy = sc * ty, sz = sc * tz;
        px += (pyp * sz - pzp * sy);
        py += (pzp * sx - pxp * sz);
        pz += (pxp * sy - pyp * sx);
        px += 0.5 * dt * q * Ex; py += 0.5 * dt * q * Ey; pz += 0.5 * dt * q * Ezc;
        const double p2n = px * px + py * py + pz * pz;
        const double gn = sqrt(1.0 + p2n / (mm * mm * c * c));
        vxn = px / (gn * mm); vyn = py / (gn * mm); vzn = pz / (gn * mm);
        xn = x + vxn * dt; yn = y + vyn * dt; zn = z + vzn * dt;
    }
    // convert_cartesian_to_cylindrical (src/FFAG_MathTools.py:2017-2060)
    const double r_new = sqrt(xn * xn + yn * yn);
    const double fb = atan2(y, x), fa = atan2(yn, xn);
    double d = fa - fb;
    if (d > FFA_PI) d -= 2.0 * FFA_PI;
    else if (d < -FFA_PI) d += 2.0 * FFA_PI;
    const double fi_new = fi + d;
    double sn, cn;
    sincos(fi_new, &sn, &cn);
    const double vr_n = vxn * cn + vyn * sn;
    const double vfi = -vxn * sn + vyn * cn;
    r = r_new;
    vr = vr_n;
    z = zn;
    vz = vzn;
    fi = fi_new;
    fidot = r_new > 1e-12 ? vfi / r_new : 0.0;
    rf = rf + s.dRF;
}

// One FP64 Boris step of every particle.
__global__ __launch_bounds__(BLOCK) void k_boris_f64(ffa_particles p, const ffa_machine *__restrict__ mp,
                                                    ffa_step s) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    double r = p.r[i], vr = p.vr[i], z = p.z[i], vz = p.vz[i], fi = p.fi[i], fidot = p.fidot[i];
    double rf = p.rf_phase[i];
    const uint8_t fl0 = p.flags[i];
    uint8_t fl = fl0;
    const double *ep = reinterpret_cast<const double *>(p.esc_pre);
    const double *epp = reinterpret_cast<const double *>(p.esc_prepre);
    const double e0[3] = {ep[i], ep[n + i], ep[2 * n + i]};
    double e1[3] = {0.0, 0.0, 0.0};
    if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) { e1[0] = epp[i]; e1[1] = epp[n + i]; e1[2] = epp[2 * n + i]; }
    boris_step_f64(*mp, s, r, vr, z, vz, fi, fidot, rf, fl, e0, e1);
    p.r[i] = r;
    p.vr[i] = vr;
    p.z[i] = z;
    p.vz[i] = vz;
    p.fi[i] = fi;
    p.fidot[i] = fidot;
    p.rf_phase[i] = rf;
    if (fl != fl0) p.flags[i] = fl;
}

// nsteps FP64 Boris steps of every particle in one launch, the state held in
// registers between steps: the same steps as nsteps launches of k_boris_f64
// (with the injection ranges inj[2 k], inj[2 k + 1] of step k applied first,
// as ffa_inject does), bit for bit, without the state's round trip through
// HBM at every step.  steps / inj: the context's device step table.
__global__ __launch_bounds__(BLOCK) void k_boris_f64_multi(ffa_particles p, const ffa_machine *__restrict__ mp,
                                                          const ffa_step *__restrict__ steps,
                                                          const int64_t *__restrict__ inj, int nsteps) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    double r = p.r[i], vr = p.vr[i], z = p.z[i], vz = p.vz[i], fi = p.fi[i], fidot = p.fidot[i];
    double rf = p.rf_phase[i];
    const uint8_t fl0 = p.flags[i];
    uint8_t fl = fl0;
    const double *ep = reinterpret_cast<const double *>(p.esc_pre);
    const double *epp = reinterpret_cast<const double *>(p.esc_prepre);
    const double e0[3] = {ep[i], ep[n + i], ep[2 * n + i]};
    double e1[3] = {0.0, 0.0, 0.0};
    // the SC flags do not change inside a block of steps
    if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) { e1[0] = epp[i]; e1[1] = epp[n + i]; e1[2] = epp[2 * n + i]; }
    const ffa_machine &m = *mp;
    for (int k = 0; k < nsteps; ++k) {
        if (i >= inj[2 * k] && i < inj[2 * k + 1]) fl = (uint8_t)(fl | FFA_F_INJ | FFA_F_ALIVE);
        boris_step_f64(m, steps[k], r, vr, z, vz, fi, fidot, rf, fl, e0, e1);
    }
    p.r[i] = r;
    p.vr[i] = vr;
    p.z[i] = z;
    p.vz[i] = vz;
    p.fi[i] = fi;
    p.fidot[i] = fidot;
    p.rf_phase[i] = rf;
    if (fl != fl0) p.flags[i] = fl;
}

// RK4_equations_of_motion_njit (src/FFAG_MathTools.py:2205-2266) for one live
// particle state Y = (r, vr, z, vz, fi, Et, rf).
__device__ __forceinline__ void rk4_rhs_f64(const ffa_machine &m, const double *Y, double esr, double esz,
                                            double esf, double rate, double v0, double *d) {
    const double r = Y[0], rdot = Y[1], z = Y[2], zdot = Y[3], fi = Y[4], Et = Y[5], rf = Y[6];
    double sf, cf;
    sincos(fi, &sf, &cf);
    double Bz = 0.0, Br = 0.0, Bfi = 0.0, Er = 0.0, Efi = 0.0, Ez = 0.0;
    int gap = 0;
    fields_f64(m, r, fi, z, sf, cf, v0, Bz, Br, Bfi, Er, Efi, Ez, gap);
    const double q = m.q, c = m.c;
    const double ig = m.E0 / Et;
    const double inv_gamma2 = ig * ig;
    const double beta2 = 1.0 - inv_gamma2;
    const double v2 = beta2 * (c * c);
    const double vperp2 = v2 - rdot * rdot - zdot * zdot;
    const double fidot = sqrt(vperp2) / r;
    const double sin_RF = m.rf_enable ? sin(rf + m.rf_shift[gap]) : 0.0;
    double ez = Ez * sin_RF, er = Er * sin_RF, efi = Efi * sin_RF;
    ez += esz; er += esr; efi += esf;
    const double Edv = rdot * er + (r * fidot) * efi + zdot * ez;
    d[0] = rdot;
    d[1] = (r * fidot * fidot + q * (c * c) / Et * (er - r * fidot * Bz + zdot * Bfi) - q * rdot / Et * Edv);
    d[2] = zdot;
    d[3] = (q * (c * c) / Et * (ez + r * fidot * Br - rdot * Bfi) - q * zdot / Et * Edv);
    d[4] = fidot;
    d[5] = q * Edv;
    d[6] = rate;
}

__global__ __launch_bounds__(BLOCK) void k_rk4_f64(ffa_particles p, const ffa_machine *__restrict__ mp,
                                                  ffa_step s) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const ffa_machine &m = *mp;
    const int64_t n = p.n;
    const uint8_t fl = p.flags[i];
    const double h = s.dt;
    const bool live = (fl & FFA_F_INJ) && (fl & FFA_F_ALIVE);
    double Y[7] = {p.r[i], p.vr[i], p.z[i], p.vz[i], p.fi[i], p.fidot[i], p.rf_phase[i]};
    if (!live) {
        // RK4_Post_Step_njit: only the RF phase advances
        const double k1 = h * s.rk_rate[0], k2 = h * s.rk_rate[1], k3 = h * s.rk_rate[1], k4 = h * s.rk_rate[2];
        p.rf_phase[i] = Y[6] + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0;
        return;
    }
    const double *ep = reinterpret_cast<const double *>(p.esc_pre);
    const double *epp = reinterpret_cast<const double *>(p.esc_prepre);
    double esr = ep[i], esz = ep[n + i], esf = ep[2 * n + i];
    if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) {
        esr = esr + s.sc_s * (esr - epp[i]);
        esz = esz + s.sc_s * (esz - epp[n + i]);
        esf = esf + s.sc_s * (esf - epp[2 * n + i]);
    }
    double d[7], k1[7], k2[7], k3[7], k4[7], S[7];
    rk4_rhs_f64(m, Y, esr, esz, esf, s.rk_rate[0], s.rk_v0[0], d);
    for (int c = 0; c < 7; ++c) { k1[c] = h * d[c]; S[c] = Y[c] + 0.5 * k1[c]; }
    rk4_rhs_f64(m, S, esr, esz, esf, s.rk_rate[1], s.rk_v0[1], d);
    for (int c = 0; c < 7; ++c) { k2[c] = h * d[c]; S[c] = Y[c] + 0.5 * k2[c]; }
    rk4_rhs_f64(m, S, esr, esz, esf, s.rk_rate[1], s.rk_v0[1], d);
    for (int c = 0; c < 7; ++c) { k3[c] = h * d[c]; S[c] = Y[c] + k3[c]; }
    rk4_rhs_f64(m, S, esr, esz, esf, s.rk_rate[2], s.rk_v0[2], d);
    for (int c = 0; c < 7; ++c) { k4[c] = h * d[c]; Y[c] = Y[c] + (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]) / 6.0; }
    p.r[i] = Y[0]; p.vr[i] = Y[1]; p.z[i] = Y[2]; p.vz[i] = Y[3];
    p.fi[i] = Y[4]; p.fidot[i] = Y[5]; p.rf_phase[i] = Y[6];
}

// Stop-condition reduction: fixed block partition, fixed-order combine.
constexpr int RED_BLOCKS = 512;

__global__ __launch_bounds__(BLOCK) void k_phi_partials(const double *__restrict__ fi, int64_t n,
                                                       double *__restrict__ part) {
    __shared__ double ss[BLOCK], sc[BLOCK];
    double acc = 0.0, cnt = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * BLOCK) {
        const double v = fi[i];
        if (v == v) { acc += v; cnt += 1.0; }
    }
    ss[threadIdx.x] = acc;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { ss[threadIdx.x] += ss[threadIdx.x + w]; sc[threadIdx.x] += sc[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = ss[0]; part[2 * blockIdx.x + 1] = sc[0]; }
}

__global__ __launch_bounds__(BLOCK) void k_phi_final(const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double ss[BLOCK], sc[BLOCK];
    double a = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < RED_BLOCKS; b += BLOCK) { a += part[2 * b]; c += part[2 * b + 1]; }
    ss[threadIdx.x] = a;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { ss[threadIdx.x] += ss[threadIdx.x + w]; sc[threadIdx.x] += sc[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = ss[0]; out[1] = sc[0]; }
}

// ffa_bfield_eval: static field at points (order 0 + Taylor + harmonics, the
// Taylor part, or its coefficient form).
__global__ __launch_bounds__(BLOCK) void k_bfield_eval(const ffa_machine *__restrict__ mp, const double *__restrict__ r,
                                                       const double *__restrict__ fi, const double *__restrict__ z,
                                                       int64_t n, double *__restrict__ out, int mode, double zt) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= n) return;
    const ffa_machine &m = *mp;
    const double P = FFA_PI * 2.0 / (double)m.nsectors;
    const double ri = r[i], fii = fi[i], zi = z[i];
    const double fim = fii - P * floor(fii / P);
    double Bz = 0.0, Br = 0.0, Bfi = 0.0;
    if (mode == 0) {
        order0_f64(m, ri, fim, zi, Bz, Br, Bfi);
        if (m.max_order >= 1) taylor_add_f64(m, ri, fim, zi, Bz, Bfi, Br);
        if (m.n_harm > 0) harmonics_f64(m, ri, fii, Bz);
    } else if (m.max_order >= 1) {
        const double z2 = zi * zi;
        if (mode == 1) taylor_add_pow_f64(m, ri, fim, z2, z2, z2 * zi, Bz, Bfi, Br);
        else taylor_add_pow_f64(m, ri, fim, z2, zt * zt, zt * zt * zt, Bz, Bfi, Br);
    }
    out[i] = Bz;
    out[n + i] = Bfi;
    out[2 * n + i] = Br;
}

// FunctionForSEO (src/FFAG_track.py:1597-1638): azimuth-stepped equations of
// a particle of fixed momentum in the order-0 field plus harmonics.  The
// azimuth reductions are the reference's Python / numpy float modulo (fmod,
// exact) rather than the tracker's floor form.
__device__ __forceinline__ void seo_rhs(const ffa_machine &m, double qp, double t, const double *x, double *d) {
    const double two_pi = 2.0 * FFA_PI;
    const double P = two_pi / (double)m.nsectors;
    double fi = fmod(t, two_pi);
    if (fi < 0.0) fi += two_pi;
    double fim = fmod(fi, P);
    if (fim < 0.0) fim += P;
    const double r = x[0], pr = x[1], z = x[2], pz = x[3];
    double Bz, Br, Bfi;
    order0_f64(m, r, fim, z, Bz, Br, Bfi);
    if (m.n_harm > 0) harmonics_f64(m, r, fi, Bz);
    const double pf = sqrt(1.0 - pr * pr - pz * pz);
    d[0] = r * pr / pf;
    d[1] = pf - qp * (r * Bz - r * pz / pf * Bfi);
    d[2] = r * pz / pf;
    d[3] = qp * (r * Br - r * pr / pf * Bfi);
}

// rk4_step (src/FFAG_track.py:33-40) repeated on the linspace grid of
// rk4_solve; one particle per thread, the trajectory optionally recorded
// component-major so each store is coalesced.
__global__ __launch_bounds__(BLOCK) void k_orbit_rk4(const ffa_machine *__restrict__ mp, ffa_orbit o) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= o.n) return;
    const ffa_machine &m = *mp;
    const int64_t n = o.n;
    double x[4] = {o.state[i], o.state[n + i], o.state[2 * n + i], o.state[3 * n + i]};
    const double qp = o.q_over_p[i], h = o.h;
    if (o.traj)
        for (int c = 0; c < 4; ++c) o.traj[c * n + i] = x[c];
    for (int s = 0; s < o.nsteps; ++s) {
        const double t = (double)s * o.t_step + o.t0;
        double k1[4], k2[4], k3[4], k4[4], y[4], d[4];
        seo_rhs(m, qp, t, x, d);
        for (int c = 0; c < 4; ++c) { k1[c] = h * d[c]; y[c] = x[c] + 0.5 * k1[c]; }
        seo_rhs(m, qp, t + 0.5 * h, y, d);
        for (int c = 0; c < 4; ++c) { k2[c] = h * d[c]; y[c] = x[c] + 0.5 * k2[c]; }
        seo_rhs(m, qp, t + 0.5 * h, y, d);
        for (int c = 0; c < 4; ++c) { k3[c] = h * d[c]; y[c] = x[c] + k3[c]; }
        seo_rhs(m, qp, t + h, y, d);
        for (int c = 0; c < 4; ++c) {
            k4[c] = h * d[c];
            x[c] = x[c] + (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]) / 6.0;
        }
        if (o.traj)
            for (int c = 0; c < 4; ++c) o.traj[((size_t)(s + 1) * 4 + c) * n + i] = x[c];
    }
    for (int c = 0; c < 4; ++c) o.state[c * n + i] = x[c];
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

hipError_t launch_orbit_rk4(ffa_ctx_s *c, const ffa_orbit *o) {
    if (o->n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_orbit_rk4, dim3(grid_for(o->n)), dim3(BLOCK), 0, c->stream, ctx_machine_dev(c), *o);
    return hipGetLastError();
}

hipError_t launch_inject(ffa_ctx_s *c, const ffa_particles *p, int64_t b, int64_t e) {
    if (e <= b) return hipSuccess;
    hipLaunchKernelGGL(k_inject, dim3(grid_for(e - b)), dim3(BLOCK), 0, c->stream, p->flags, b, e);
    return hipGetLastError();
}

hipError_t launch_boris_f64(ffa_ctx_s *c, const ffa_particles *p, const ffa_step *s) {
    if (p->n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_boris_f64, dim3(grid_for(p->n)), dim3(BLOCK), 0, c->stream, *p, ctx_machine_dev(c), *s);
    return hipGetLastError();
}

hipError_t launch_boris_f64_multi(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps) {
    if (p->n == 0 || nsteps == 0) return hipSuccess;
    const ffa_step *steps = reinterpret_cast<const ffa_step *>(c->dev + FFA_CTX_STEPS_OFF);
    const int64_t *inj = reinterpret_cast<const int64_t *>(c->dev + FFA_CTX_INJ_OFF);
    hipLaunchKernelGGL(k_boris_f64_multi, dim3(grid_for(p->n)), dim3(BLOCK), 0, c->stream, *p, ctx_machine_dev(c),
                       steps, inj, nsteps);
    return hipGetLastError();
}

hipError_t launch_rk4_f64(ffa_ctx_s *c, const ffa_particles *p, const ffa_step *s) {
    if (p->n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_rk4_f64, dim3(grid_for(p->n)), dim3(BLOCK), 0, c->stream, *p, ctx_machine_dev(c), *s);
    return hipGetLastError();
}

hipError_t launch_reduce_phi(ffa_ctx_s *c, const ffa_particles *p, double *out_dev) {
    double *part = reinterpret_cast<double *>(c->dev + FFA_CTX_RED_OFF);
    hipLaunchKernelGGL(k_phi_partials, dim3(RED_BLOCKS), dim3(BLOCK), 0, c->stream, p->fi, p->n, part);
    hipLaunchKernelGGL(k_phi_final, dim3(1), dim3(BLOCK), 0, c->stream, part, out_dev);
    return hipGetLastError();
}

hipError_t launch_bfield_eval(ffa_ctx_s *c, const double *r, const double *fi, const double *z, int64_t n,
                              double *out, int mode, double zt) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bfield_eval, dim3(grid_for(n)), dim3(BLOCK), 0, c->stream, ctx_machine_dev(c), r, fi, z, n,
                       out, mode, zt);
    return hipGetLastError();
}
