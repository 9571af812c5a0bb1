// FP32-field tier of the Boris push (gfx950): nsteps steps per launch with
// the particle state held in registers.
//
// Precision split: position, velocity, RF phase and the time base stay FP64
// (multi-thousand-turn stability); the field evaluation (order-0 analytic
// field, Taylor-map gather, harmonics, RF gap field, SC endpoint
// extrapolation) is FP32 with hardware transcendentals; the Boris rotation and
// the position update are FP64.
//
// The state is converted to Cartesian once per launch and back once at the
// end, instead of once per step as the reference does
// (src/FFAG_MathTools.py:1969-2060): cos(phi), sin(phi) of the field rotation
// are x/r, y/r, and the unwrapped azimuth is rebuilt from atan2 plus a count
// of crossings of the atan2 branch cut, which is the reference's
// |d phi| > pi unwrapping rule applied step by step.
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ void fields_f32(const ffa_machine &m, float r, float th, float z, float xf, float yf,
                                           float v0, float rf_red, float &Bz, float &Br, float &Bfi, float &Er,
                                           float &Efi, float &Ez, float &sin_rf) {
    // fast_mod on the branch-cut azimuth (equal to the unwrapped one mod the sector)
    const float P = (float)(FFA_TWO_PI / (double)m.nsectors);
    float fim = th - P * floorf(th / P);
    float spiral = 0.0f, spiral_p = 0.0f;
    for (int a = 0; a < m.n_spoly; ++a) spiral = spiral * r + (float)m.s_poly[a];
    for (int a = 0; a < m.n_sppoly; ++a) spiral_p = spiral_p * r + (float)m.sp_poly[a];
    const float kv = (float)m.k_value;
    float B0r, B0r_p;
    if (kv == 0.0f) {
        B0r = (float)m.B0;
        B0r_p = 0.0f;
    } else {
        const float lr = __logf(r);
        B0r = (float)m.B0 * __expf(kv * lr);
        B0r_p = B0r * kv / r;
    }
    float fs = fim - spiral;
    fs -= P * floorf(fs / P);
    float F = 0.0f, Fp = 0.0f;
    for (int j = 0; j < m.n_seg; ++j) {
        const float x0 = (float)m.convex[j][0], x1 = (float)m.convex[j][1];
        const float ks = (float)(16.0 / m.convex[j][2]);
        const float yc = (float)(m.convex[j][4] * fabs(m.convex[j][3]));
        const float s1 = 1.0f / (1.0f + __expf(-ks * (fs - x0)));
        const float s2 = 1.0f / (1.0f + __expf(ks * (fs - x1)));
        F += yc * s1 * s2;
        Fp += yc * ks * s1 * s2 * ((1.0f - s1) - (1.0f - s2));
    }
    Bz = B0r * F;
    Bfi = B0r * Fp / r * z;
    Br = (B0r_p * F - B0r * Fp * spiral_p) * z;
    if (m.max_order >= 1) {
        const int NX = m.r_size, NY = m.fi_size, no = m.max_order;
        float fx = (r - (float)m.r_min) / (float)m.r_step;
        float fy = (fim - (float)m.fi_min) / (float)m.fi_step;
        fx = fminf(fmaxf(fx, -1.0f), (float)NX);
        fy = fminf(fmaxf(fy, -1.0f), (float)NY);
        int i = (int)fx, j = (int)fy;
        i = i < 0 ? 0 : (i >= NX - 1 ? NX - 2 : i);
        j = j < 0 ? 0 : (j >= NY - 1 ? NY - 2 : j);
        const float tx = fx - (float)i, ty = fy - (float)j;
        const float w00 = (1.0f - tx) * (1.0f - ty), w01 = (1.0f - tx) * ty, w10 = tx * (1.0f - ty), w11 = tx * ty;
        const float *T = reinterpret_cast<const float *>(m.taylor);
        const int sj = 3 * no;
        const float *a0 = T + ((int64_t)i * NY + j) * sj;
        const float *a1 = T + ((int64_t)(i + 1) * NY + j) * sj;
        const float z2 = z * z;
        float ze = z2, zo = z2 * z;
        for (int n = 0; n < no; ++n) {
            const float *p00 = a0 + 3 * n, *p01 = a0 + sj + 3 * n, *p10 = a1 + 3 * n, *p11 = a1 + sj + 3 * n;
            Bz += (w00 * p00[0] + w01 * p01[0] + w10 * p10[0] + w11 * p11[0]) * ze;
            Bfi += (w00 * p00[1] + w01 * p01[1] + w10 * p10[1] + w11 * p11[1]) * zo;
            Br += (w00 * p00[2] + w01 * p01[2] + w10 * p10[2] + w11 * p11[2]) * zo;
            ze *= z2;
            zo *= z2;
        }
    }
    if (m.n_harm > 0 && r > 0.0f) {
        for (int k = 0; k < m.n_harm; ++k)
            if ((float)m.harm_rmin[k] <= r && r <= (float)m.harm_rmax[k])
                Bz += (float)m.harm_amp[k] * __cosf((float)m.harm_m[k] * th + (float)m.harm_phase[k]);
    }
    Er = 0.0f; Efi = 0.0f; Ez = 0.0f;
    sin_rf = 0.0f;
    if (m.rf_enable) {
        int gap = 0;
        const float acc = (float)(m.acc_etot) * v0;
        for (int g = 0; g < m.n_gap; ++g) {
            if (xf < (float)m.gap_box[g][0] || xf > (float)m.gap_box[g][1] || yf < (float)m.gap_box[g][2] ||
                yf > (float)m.gap_box[g][3]) {
                Er = 0.0f; Efi = 0.0f;
                continue;
            }
            bool inside = true;
            float prev = 0.0f;
            for (int j = 0; j < 4; ++j) {
                const float x0 = (float)m.gap_px[g][j], y0 = (float)m.gap_py[g][j];
                const float x1 = (float)m.gap_px[g][(j + 1) & 3], y1 = (float)m.gap_py[g][(j + 1) & 3];
                const float cross = (x1 - x0) * (yf - y0) - (y1 - y0) * (xf - x0);
                if (j == 0) prev = cross;
                else if (cross * prev < 0.0f) { inside = false; break; }
            }
            if (inside) {
                float sd, cd;
                __sincosf((float)(m.gap_theta[g] + FFA_PI / 2.0) - th, &sd, &cd);
                Er = acc * cd;
                Efi = acc * sd;
                gap = g;
            } else {
                Er = 0.0f; Efi = 0.0f;
                gap = 0;
            }
        }
        sin_rf = __sinf(rf_red + (float)m.rf_shift[gap]);
    }
}

template <typename ET>
__global__ __launch_bounds__(BLOCK) void k_boris_fused(ffa_particles p, const ffa_machine *__restrict__ mp,
                                                      const ffa_step *__restrict__ steps,
                                                      const int64_t *__restrict__ inj, int nsteps) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const ffa_machine &m = *mp;
    const int64_t n = p.n;
    uint8_t fl = p.flags[i];
    double rf = p.rf_phase[i];
    const ET *ep = reinterpret_cast<const ET *>(p.esc_pre);
    const ET *epp = reinterpret_cast<const ET *>(p.esc_prepre);
    const float e0r = (float)ep[i], e0z = (float)ep[n + i], e0f = (float)ep[2 * n + i];
    const float d_r = e0r - (float)epp[i], d_z = e0z - (float)epp[n + i], d_f = e0f - (float)epp[2 * n + i];
    const double r0 = p.r[i], fi0 = p.fi[i];
    double s0, c0;
    sincos(fi0, &s0, &c0);
    double x = r0 * c0, y = r0 * s0, z = p.z[i];
    const double vr0 = p.vr[i], rfd = r0 * p.fidot[i];
    double vx = vr0 * c0 - rfd * s0, vy = vr0 * s0 + rfd * c0, vz = p.vz[i];
    const double th0 = atan2(y, x);
    int wind = 0;
    bool moved = false;
    const double q = m.q, mm = m.m0, cl = m.c;
    const double mc2 = mm * mm * cl * cl;
    for (int k = 0; k < nsteps; ++k) {
        const ffa_step &s = steps[k];
        if (i >= inj[2 * k] && i < inj[2 * k + 1]) fl = (uint8_t)(fl | FFA_F_INJ | FFA_F_ALIVE);
        const bool live = (fl & FFA_F_INJ) && (fl & FFA_F_ALIVE);
        if (!live) { rf += s.dRF; continue; }
        moved = true;
        float esr = e0r, esz = e0z, esf = e0f;
        if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) {
            const float sc = (float)s.sc_s;
            esr = e0r + sc * d_r; esz = e0z + sc * d_z; esf = e0f + sc * d_f;
        }
        const float xf = (float)x, yf = (float)y, zf = (float)z;
        const float rr = sqrtf(xf * xf + yf * yf);
        const float th = atan2f(yf, xf);
        const double rfr = rf - FFA_TWO_PI * rint(rf * (1.0 / FFA_TWO_PI));
        float Bz, Br, Bfi, Er, Efi, Ez, srf;
        fields_f32(m, rr, th, zf, xf, yf, (float)s.v0, (float)rfr, Bz, Br, Bfi, Er, Efi, Ez, srf);
        const float cf = xf / rr, sf = yf / rr;
        const float ert = Er * srf + esr, eft = Efi * srf + esf, ezt = Ez * srf + esz;
        const double Ex = (double)(ert * cf - eft * sf), Ey = (double)(ert * sf + eft * cf), Ezc = (double)ezt;
        const double Bx = -(double)(Br * cf - Bfi * sf), By = -(double)(Br * sf + Bfi * cf), Bzc = -(double)Bz;
        if (m.aperture_enable) {
            const double rd = (double)rr;
            if (rd > m.aperture[1] || rd < m.aperture[0] || z > m.aperture[3] || z < m.aperture[2])
                fl = (uint8_t)(fl & ~FFA_F_ALIVE);
        }
        // relativistic Boris, FP64 (src/FFAG_MathTools.py:1918-1966)
        const double dt = s.dt;
        const double hq = 0.5 * dt * q;
        const double v2 = vx * vx + vy * vy + vz * vz;
        const double gmm = mm / sqrt(1.0 - v2 / (cl * cl));
        double px = gmm * vx + hq * Ex, py = gmm * vy + hq * Ey, pz = gmm * vz + hq * Ezc;
        const double gm = sqrt(1.0 + (px * px + py * py + pz * pz) / mc2);
        const double cb = hq / (gm * mm);
        const double tx = cb * Bx, ty = cb * By, tz = cb * Bzc;
        const double pxp = px + (py * tz - pz * ty), pyp = py + (pz * tx - px * tz), pzp = pz + (px * ty - py * tx);
        const double sc2 = 2.0 / (1.0 + tx * tx + ty * ty + tz * tz);
        px += sc2 * (pyp * tz - pzp * ty) + hq * Ex;
        py += sc2 * (pzp * tx - pxp * tz) + hq * Ey;
        pz += sc2 * (pxp * ty - pyp * tx) + hq * Ezc;
        const double ign = 1.0 / (mm * sqrt(1.0 + (px * px + py * py + pz * pz) / mc2));
        vx = px * ign; vy = py * ign; vz = pz * ign;
        const double xo = x, yo = y;
        x += vx * dt; y += vy * dt; z += vz * dt;
        // atan2 branch cut (negative x axis) crossings
        if (xo + x < 0.0) {
            if (yo >= 0.0 && y < 0.0) ++wind;
            else if (yo < 0.0 && y >= 0.0) --wind;
        }
        rf += s.dRF;
    }
    p.rf_phase[i] = rf;
    if (fl != p.flags[i]) p.flags[i] = fl;
    if (!moved) return;
    const double rn = sqrt(x * x + y * y);
    const double fin = fi0 + (atan2(y, x) - th0 + FFA_TWO_PI * (double)wind);
    const double cn = x / rn, sn = y / rn;
    p.r[i] = rn;
    p.vr[i] = vx * cn + vy * sn;
    p.z[i] = z;
    p.vz[i] = vz;
    p.fi[i] = fin;
    p.fidot[i] = (-vx * sn + vy * cn) / rn;
}

}  // namespace

hipError_t launch_boris_fused_f32(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps) {
    if (p->n == 0 || nsteps == 0) return hipSuccess;
    const ffa_step *steps = reinterpret_cast<const ffa_step *>(c->dev + FFA_CTX_STEPS_OFF);
    const int64_t *inj = reinterpret_cast<const int64_t *>(c->dev + FFA_CTX_INJ_OFF);
    const unsigned g = (unsigned)((p->n + BLOCK - 1) / BLOCK);
    if (p->esc_f32)
        hipLaunchKernelGGL(k_boris_fused<float>, dim3(g), dim3(BLOCK), 0, c->stream, *p, ctx_machine_dev(c), steps,
                           inj, nsteps);
    else
        hipLaunchKernelGGL(k_boris_fused<double>, dim3(g), dim3(BLOCK), 0, c->stream, *p, ctx_machine_dev(c), steps,
                           inj, nsteps);
    return hipGetLastError();
}
