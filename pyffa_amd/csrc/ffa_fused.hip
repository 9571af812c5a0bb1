// FP32-field tier of the Boris push (gfx950): nsteps steps per launch with
// the particle state held in registers.
//
// Precision split: position, velocity, RF phase and the time base stay FP64
// (multi-thousand-turn stability); the field evaluation (order-0 analytic
// field, Taylor-map gather, harmonics, RF gap field, SC endpoint
// extrapolation) is FP32 with hardware transcendentals and reciprocals; the
// Boris rotation and the position update are FP64.
//
// The state is converted to Cartesian once per launch and back once at the
// end, instead of once per step as the reference does
// (src/FFAG_MathTools.py:1969-2060): cos(phi), sin(phi) of the field rotation
// are x/r, y/r.  The azimuth is carried unwrapped in FP64, advanced by the
// exact turning angle of each step's chord (the reference's |d phi| < pi
// unwrapping of atan2 differences, src/FFAG_MathTools.py:2035-2060), and the
// stored azimuth at the end is the FP64 atan2 of the final position in the
// turn the carried azimuth is in.  Between steps the
// kernel carries u = gamma v, so the Boris step needs two reciprocal square
// roots and one division (the reference recomputes gamma from v each step;
// the two agree to FP64 rounding).
#include "ffa_mlp.h"
#include "ffa_sc_gather.h"

namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware base-2 logarithm and exponential (about 1 ulp; the FP32 tier's
// field error budget is far above that)
__device__ __forceinline__ float hw_log2f(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float hw_exp2f(float x) { return __builtin_amdgcn_exp2f(x); }

typedef const __attribute__((address_space(1))) float gfloat;

// Spiral RF gaps in FP32 (the FP64 path follows the reference's operation
// order): first gap whose window, moved by the nearest whole number of
// turns, holds the azimuth; bilinear fields on (r, local azimuth), weights
// extrapolating outside the map as the reference's do.
__device__ __forceinline__ void spiral_gap_f32(const ffa_fconst *__restrict__ k, float r, float th, float z, float acc,
                                               float &Er, float &Efi, float &Ez, int &gap) {
    int hit = -1;
    float loc = 0.0f;
    for (int g = 0; g < k->n_gap; ++g) {
        const float ph = th - k->gap_az[g];
        const float kk = rintf((k->emap_center - ph) * 0.15915494309189535f);
        const float pl = ph + kk * 6.283185307179586f;
        if (pl >= k->emap_phi0 && pl <= k->emap_phi_max) { hit = g; loc = pl; break; }
    }
    if (hit < 0) return;
    const int NX = k->emap_nr, NY = k->emap_nphi;
    const float fx = (r - k->emap_r0) * k->emap_inv_dr, fy = (loc - k->emap_phi0) * k->emap_inv_dphi;
    int i = (int)floorf(fminf(fmaxf(fx, -1.0f), (float)NX));
    int j = (int)floorf(fminf(fmaxf(fy, -1.0f), (float)NY));
    i = i < 0 ? 0 : (i >= NX - 1 ? NX - 2 : i);
    j = j < 0 ? 0 : (j >= NY - 1 ? NY - 2 : j);
    const float tx = fx - (float)i, ty = fy - (float)j;
    const float w00 = (1.0f - tx) * (1.0f - ty), w01 = (1.0f - tx) * ty, w10 = tx * (1.0f - ty), w11 = tx * ty;
    const double *e00 = k->emap + ((int64_t)i * NY + j) * 3, *e01 = e00 + 3;
    const double *e10 = k->emap + ((int64_t)(i + 1) * NY + j) * 3, *e11 = e10 + 3;
    const float er = w00 * (float)e00[0] + w01 * (float)e01[0] + w10 * (float)e10[0] + w11 * (float)e11[0];
    const float ezc = w00 * (float)e00[1] + w01 * (float)e01[1] + w10 * (float)e10[1] + w11 * (float)e11[1];
    const float eph = w00 * (float)e00[2] + w01 * (float)e01[2] + w10 * (float)e10[2] + w11 * (float)e11[2];
    Er = acc * er;
    Ez = acc * ezc * z;
    Efi = acc * eph;
    gap = hit;
}

// Radius and inverse radius of an FP32 position: one reciprocal square root.
__device__ __forceinline__ void radius_f32(float xf, float yf, float &r2, float &rr, float &ir) {
    r2 = xf * xf + yf * yf;
    ir = __builtin_amdgcn_rsqf(r2);
    rr = r2 * ir;
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) f4v gfloat4;

// Orders 1..NO of the Taylor gather written out for one map order (the
// kernel instance for order-NO maps; order 2 is written out in fields_f32).
// Each of the cell's two runs (6 NO floats: corner (i, .) then corner
// (i + 1, .), orders 1..NO, (Bz, Bphi, Br) each) is read as one block, which
// the compiler issues as 16-byte loads instead of 4 NO 12-byte loads; the
// four corners' sums per order are completed before the order's z power
// multiplies them, as in the other paths.
template <int NO>
__device__ __forceinline__ void taylor_runs(gfloat *a0, gfloat *a1, float w00, float w01, float w10, float w11,
                                            float z2, float z, float &Bz, float &Bfi, float &Br) {
    constexpr int L = 6 * NO, H = 3 * NO;
    float c0[L], c1[L];
#pragma unroll
    for (int q = 0; q < L; ++q) {
        c0[q] = a0[q];
        c1[q] = a1[q];
    }
    float ze = z2, zo = z2 * z;
#pragma unroll
    for (int n = 0; n < NO; ++n) {
        const float bz = w00 * c0[3 * n] + w01 * c1[3 * n] + w10 * c0[H + 3 * n] + w11 * c1[H + 3 * n];
        const float bf = w00 * c0[3 * n + 1] + w01 * c1[3 * n + 1] + w10 * c0[H + 3 * n + 1] + w11 * c1[H + 3 * n + 1];
        const float br = w00 * c0[3 * n + 2] + w01 * c1[3 * n + 2] + w10 * c0[H + 3 * n + 2] + w11 * c1[H + 3 * n + 2];
        Bz += bz * ze;
        Bfi += bf * zo;
        Br += br * zo;
        ze *= z2;
        zo *= z2;
    }
}

// Field at (r, th, z): th is the azimuth within a turn of the branch cut
// (harmonics, gap direction, spiral-gap window), fim the same azimuth reduced
// to the sector [0, P) in FP64 before rounding (the map index and the
// order-0 field: fast_mod of the reference, src/FFAG_MathTools.py:1553-1557).
// locd is fim before rounding.  The order-0 flutter F and its slope F' come
// from the sector table (per-interval cubic Hermite on exact values and
// slopes: relative error < 1e-7 of max |F| and < 1e-6 of max |F'| at 8192
// intervals per sector), indexed in FP64 from locd when there is no spiral,
// so that the steep edges see the FP64 azimuth rather than its FP32
// rounding.
//
// ORD > 0: the kernel instance for order-ORD maps (its own register budget),
// with the gather written out for that order.
template <int ORD = 0>
__device__ __forceinline__ void fields_f32(const ffa_fconst *__restrict__ k, float r, float inv_r, float th, float fim,
                                           double locd, float z, float xf, float yf, float v0, double rf, float &Bz,
                                           float &Br, float &Bfi, float &Er, float &Efi, float &Ez, float &sin_rf,
                                           bool taylor = true) {
    float spiral_p = 0.0f, t;
    int it;
    if (k->n_sp > 0) {
        float spiral = 0.0f;
        for (int a = 0; a < k->n_sp; ++a) spiral = spiral * r + k->sp[a];
        float fs = fim - spiral;
        fs -= k->P * floorf(fs * k->invP);
        const float u = fmaxf(fs * k->flut_scale, 0.0f);
        it = min((int)u, FFA_FLUT_N - 1);
        t = u - (float)it;
    } else {
        const double u = fmax(locd * k->flut_scale_d, 0.0);
        it = min((int)u, FFA_FLUT_N - 1);
        t = (float)(u - (double)it);
    }
    const f4v fc = *((gfloat4 *)(k->flut) + it);  // global address space: global_load, not flat
    for (int a = 0; a < k->n_spp; ++a) spiral_p = spiral_p * r + k->spp[a];
    // B0 r^k (k = 0: exp2(0) = 1, B0 exactly)
    const float B0r = k->B0 * hw_exp2f(k->kv * hw_log2f(r));
    const float B0r_p = B0r * k->kv * inv_r;
    const float F = fmaf(t, fmaf(t, fmaf(t, fc.w, fc.z), fc.y), fc.x);
    const float Fp = fmaf(t, fmaf(t, 3.0f * fc.w, 2.0f * fc.z), fc.y) * k->flut_scale;
    Bz = B0r * F;
    Bfi = B0r * Fp * inv_r * z;
    Br = (B0r_p * F - B0r * Fp * spiral_p) * z;
    if (k->no >= 1 && taylor) {
        // cell (int truncation, clamped to [0, N - 2]) and weights of the
        // position clamped to one cell beyond the map
        const float fx = __builtin_amdgcn_fmed3f((r - k->r_min) * k->inv_rstep, -1.0f, k->NXf);
        const float fy = __builtin_amdgcn_fmed3f((fim - k->fi_min) * k->inv_fistep, -1.0f, k->NYf);
        const int i = (int)__builtin_amdgcn_fmed3f(fx, 0.0f, k->NXm2f);
        const int j = (int)__builtin_amdgcn_fmed3f(fy, 0.0f, k->NYm2f);
        const float tx = fx - (float)i, ty = fy - (float)j;
        const float w00 = (1.0f - tx) * (1.0f - ty), w01 = (1.0f - tx) * ty, w10 = tx * (1.0f - ty), w11 = tx * ty;
        const int sj = 3 * k->no;
        const float z2 = z * z;
        // azimuth-major map: (i, j), (i+1, j) and (i, j+1), (i+1, j+1) are two
        // contiguous runs of 2 sj floats; the lanes of a wave (one azimuth
        // cell after the ordering) read neighbouring runs.  32-bit offsets
        // (the host caps the map at 2^32 elements).
        gfloat *T = (gfloat *)(k->T);          // global address space: global_load, not flat
        const unsigned cell = __umul24((unsigned)j, (unsigned)k->NX) + (unsigned)i;
        float ze = z2, zo = z2 * z;
        if constexpr (ORD > 0) {
            gfloat *p0 = T + cell * (unsigned)(3 * ORD), *p1 = p0 + (unsigned)(k->NX * 3 * ORD);
            taylor_runs<ORD>(p0, p1, w00, w01, w10, w11, z2, z, Bz, Bfi, Br);
        } else if (k->no == 2) {
            gfloat *a0 = T + cell * 6u, *a1 = a0 + 6u * (unsigned)k->NX;
            // 12 floats per run: order 1 (Bz, Bfi, Br), order 2, for i and i + 1
            float c0[12], c1[12];
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                c0[q] = a0[q];
                c1[q] = a1[q];
            }
#pragma unroll
            for (int n = 0; n < 2; ++n) {
                const float bz = w00 * c0[3 * n] + w01 * c1[3 * n] + w10 * c0[6 + 3 * n] + w11 * c1[6 + 3 * n];
                const float bf = w00 * c0[3 * n + 1] + w01 * c1[3 * n + 1] + w10 * c0[7 + 3 * n] + w11 * c1[7 + 3 * n];
                const float br = w00 * c0[3 * n + 2] + w01 * c1[3 * n + 2] + w10 * c0[8 + 3 * n] + w11 * c1[8 + 3 * n];
                Bz += bz * ze;
                Bfi += bf * zo;
                Br += br * zo;
                ze *= z2;
                zo *= z2;
            }
        } else {
            gfloat *p0 = T + cell * (unsigned)sj, *p1 = p0 + (unsigned)(k->NX * sj);
            for (int n = 0; n < k->no; ++n) {
                gfloat *p00 = p0 + 3 * n, *p10 = p0 + sj + 3 * n, *p01 = p1 + 3 * n, *p11 = p1 + sj + 3 * n;
                Bz += (w00 * p00[0] + w01 * p01[0] + w10 * p10[0] + w11 * p11[0]) * ze;
                Bfi += (w00 * p00[1] + w01 * p01[1] + w10 * p10[1] + w11 * p11[1]) * zo;
                Br += (w00 * p00[2] + w01 * p01[2] + w10 * p10[2] + w11 * p11[2]) * zo;
                ze *= z2;
                zo *= z2;
            }
        }
    }
    if (k->n_harm > 0 && r > 0.0f) {
        for (int q = 0; q < k->n_harm; ++q)
            if (k->harm[q][3] <= r && r <= k->harm[q][4])
                Bz += k->harm[q][1] * __cosf(k->harm[q][0] * th + k->harm[q][2]);
    }
    Er = 0.0f; Efi = 0.0f; Ez = 0.0f;
    sin_rf = 0.0f;
    if (k->rf) {
        int gap = 0;
        const float acc = k->acc * v0;
        if (k->gap_model == 1) spiral_gap_f32(k, r, th, z, acc, Er, Efi, Ez, gap);
        else if (k->n_gap > 0) {
            // The reference evaluates the gaps in order and each one sets or
            // zeroes the field (src/FFAG_Field.py:296-306), so the last gap
            // alone decides it: a particle outside the last gap's box or
            // polygon gets E = 0 whatever the earlier gaps gave, and the gap
            // index only matters where E != 0.  Evaluate that gap only.
            const int g = k->n_gap - 1;
            if (!(xf < k->gap_box[g][0] || xf > k->gap_box[g][1] || yf < k->gap_box[g][2] ||
                  yf > k->gap_box[g][3])) {
                bool inside = true;
                float prev = 0.0f;
                for (int q = 0; q < 4; ++q) {
                    const float x0 = k->gap_px[g][q], y0 = k->gap_py[g][q];
                    const float x1 = k->gap_px[g][(q + 1) & 3], y1 = k->gap_py[g][(q + 1) & 3];
                    const float cross = (x1 - x0) * (yf - y0) - (y1 - y0) * (xf - x0);
                    if (q == 0) prev = cross;
                    else if (cross * prev < 0.0f) { inside = false; break; }
                }
                if (inside) {
                    float sd, cd;
                    __sincosf(k->gap_thn[g] - th, &sd, &cd);
                    Er = acc * cd;
                    Efi = acc * sd;
                    gap = g;
                }
            }
        }
        // the RF factor only matters inside a gap (E = 0 elsewhere)
        if (Er != 0.0f || Efi != 0.0f || Ez != 0.0f) {
            const double rfr = rf - FFA_TWO_PI * rint(rf * (1.0 / FFA_TWO_PI));
            sin_rf = __sinf((float)rfr + k->rf_shift[gap]);
        }
    }
}

// SUR: the field surrogate (weights staged in LDS, one 512-thread workgroup
// per CU) replaces the Taylor-map gather for particles with |z| <= zt; every
// lane runs the step loop to the end so the wave-wide matrix-core work sees
// all 64 lanes.
constexpr int SUR_BLOCK = 512;
// Waves per SIMD of the order-6 instance: its written-out gather holds two
// 144-byte runs; at 4 waves (128 registers) a few values spill outside the
// step loop, at 3 (no spill) it ran 8 % slower on the 32M spiral cycle.
#ifndef ORD_WAVES
#define ORD_WAVES 4
#endif

// Without the surrogate the kernel is held to 96 VGPRs (5 waves per SIMD):
// the extra wave hides more of the gathers' and constants' latency than the
// few spilled registers cost.
//
// GATHER (1: from the (Ex, Ez) cubes, 2: differentiating phi): the launch
// first applies the gather and endpoint roll of the SC solve whose results
// g describes (ffa_sc_solve_field), particle by particle with the solve's
// own arithmetic, then pushes from the new endpoints; p's endpoint pointers
// are the ones the solve saw (the caller swaps them after this launch).
template <typename ET, bool SUR, int GATHER = 0, int ORD = 0>
__global__ __launch_bounds__(SUR ? SUR_BLOCK : BLOCK)
__attribute__((amdgpu_waves_per_eu(SUR ? 1 : (ORD > 0 ? ORD_WAVES : 5), SUR ? 8 : (ORD > 0 ? ORD_WAVES : 5)))) void k_boris_fused(ffa_particles p, const ffa_machine *__restrict__ mp,
                                                      const ffa_fconst *__restrict__ kc,
                                                      const ffa_step *__restrict__ steps,
                                                      const int64_t *__restrict__ inj, int nsteps,
                                                      const char *__restrict__ sur_w,
                                                      const mlp::fdesc *__restrict__ md, sc_gather_src g) {
    const int64_t i = xcd_block() * (int64_t)(SUR ? SUR_BLOCK : BLOCK) + threadIdx.x;
    __shared__ __attribute__((aligned(16))) char lds[SUR ? mlp::FWD_BYTES : 16];
    if constexpr (SUR) {
        for (int k = threadIdx.x; k < mlp::FWD_BYTES / 16; k += SUR_BLOCK)
            reinterpret_cast<uint4 *>(lds)[k] = reinterpret_cast<const uint4 *>(sur_w)[k];
        __syncthreads();
    } else {
        if (i >= p.n) return;
    }
    const bool inb = i < p.n;
    const int64_t ii = inb ? i : 0;    // lanes past the end load particle 0 and never store
    const ffa_machine &m = *mp;
    const int64_t n = p.n;
    uint8_t fl = inb ? p.flags[ii] : (uint8_t)0;
    double rf = p.rf_phase[ii];
    // the state, loaded before the gather's dependent loads so that their
    // round trips overlap
    const double r0 = p.r[ii], fi0 = p.fi[ii], z0 = p.z[ii], vr0 = p.vr[ii], fd0 = p.fidot[ii], vz0 = p.vz[ii];
    float e0r, e0z, e0f, d_r, d_z, d_f;
    if constexpr (GATHER != 0) {
        // the solve's gather and roll for this particle (ffa_sc_gather.h):
        // the new pre-step field goes into the prepre memory, the pre memory
        // (on the first solve also overwritten) holds the new prepre field
        ET *gp = reinterpret_cast<ET *>(p.esc_pre);
        ET *gpp = reinterpret_cast<ET *>(p.esc_prepre);
        const ffa_sc_result *res = g.res;
        const int nx = g.sp.nx, nz = g.sp.nz;
        const scg::cellw cw = scg::cell_weights(g.pslice[ii], g.px[ii], z0, g.sp.ns, nx, nz, res->xmin, res->zmin,
                                                1.0 / res->dx, 1.0 / res->dz);
        const scg::field_scale fsc =
            GATHER == 2 ? scg::make_field_scale(res->dx, res->dz, res->mean_gamma_m1 + 1.0) : scg::field_scale{};
        float2 a00, a10, a01, a11;
        scg::corners<GATHER == 2>(cw, g.exz, g.phi, nx, nz, fsc, a00, a10, a01, a11);
        ET a, b, c;
        fl = scg::roll<ET>(ii, n, fl, res->bad_grid == 0.0, cw, a00, a10, a01, a11, g.ppr0[ii], g.sp, gp, gpp, a, b, c);
        e0r = (float)a; e0z = (float)b; e0f = (float)c;
        const bool sc_both = (fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE);
        d_r = sc_both ? e0r - (float)gp[ii] : 0.0f;
        d_z = sc_both ? e0z - (float)gp[n + ii] : 0.0f;
        d_f = sc_both ? e0f - (float)gp[2 * n + ii] : 0.0f;
    } else {
        const ET *ep = reinterpret_cast<const ET *>(p.esc_pre);
        const ET *epp = reinterpret_cast<const ET *>(p.esc_prepre);
        e0r = (float)ep[ii]; e0z = (float)ep[n + ii]; e0f = (float)ep[2 * n + ii];
        // the endpoint slope, zero unless both endpoints exist (the SC flags do
        // not change inside a launch): E = E_pre + s (E_pre - E_prepre)
        const bool sc_both = (fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE);
        d_r = sc_both ? e0r - (float)epp[ii] : 0.0f;
        d_z = sc_both ? e0z - (float)epp[n + ii] : 0.0f;
        d_f = sc_both ? e0f - (float)epp[2 * n + ii] : 0.0f;
    }
    double s0, c0;
    sincos(fi0, &s0, &c0);
    double x = r0 * c0, y = r0 * s0, z = z0;
    const double rfd = r0 * fd0;
    const double vx0 = vr0 * c0 - rfd * s0, vy0 = vr0 * s0 + rfd * c0;
    const double q = m.q, mm = m.m0, cl = m.c;
    const double inv_c2 = 1.0 / (cl * cl);
    // u = gamma v
    const double g0 = 1.0 / sqrt(1.0 - (vx0 * vx0 + vy0 * vy0 + vz0 * vz0) * inv_c2);
    double ux = g0 * vx0, uy = g0 * vy0, uz = g0 * vz0;
    double vx = vx0, vy = vy0, vz = vz0;
    // azimuth of the position, carried unwrapped in FP64 from the launch's
    // exact atan2 and advanced per step by the exact angle of the step's
    // chord; the sector-local azimuth of the field is reduced from it in FP64
    const double th0 = atan2(y, x);
    double thd = th0;
    const double Pd = FFA_TWO_PI / (double)m.nsectors, invPd = (double)m.nsectors / FFA_TWO_PI;
    bool moved = false;
    const double q_over_m = q / mm;
    const bool aperture = m.aperture_enable != 0;
    const double ap0 = m.aperture[0], ap1 = m.aperture[1], ap2 = m.aperture[2], ap3 = m.aperture[3];
    for (int kstep = 0; kstep < nsteps; ++kstep) {
        const ffa_step &s = steps[kstep];
        const int64_t ilo = inj[2 * kstep], ihi = inj[2 * kstep + 1];
        // ilo <= i < ihi as one unsigned compare (the host checks ilo <= ihi)
        if ((uint64_t)(i - ilo) < (uint64_t)(ihi - ilo)) fl = (uint8_t)(fl | FFA_F_INJ | FFA_F_ALIVE);
        const bool live = (fl & FFA_F_INJ) && (fl & FFA_F_ALIVE);
        bool use_sur = false;
        float sBz = 0.0f, sBfi = 0.0f, sBr = 0.0f;
        const float xf = (float)x, yf = (float)y, zf = (float)z;
        float r2f, rr, ir;
        radius_f32(xf, yf, r2f, rr, ir);
        const float th = (float)thd;
        const double locd = thd - Pd * floor(thd * invPd);
        const float fim = (float)locd;
        if constexpr (SUR) {
            float a, u, c;
            mlp::inputs(md, rr, fim, zf, a, u, c);
            use_sur = live && fabsf(c) <= 1.0f;
            if (__ballot(use_sur) != 0) {          // wave-uniform: all 64 lanes evaluate
                float o[3];
                mlp::forward_wave(reinterpret_cast<const __bf16 *>(lds),
                                  reinterpret_cast<const float *>(lds + mlp::FWD_BF16 * 2), md, a, u, c * c, o);
                // lanes outside |z| <= zt keep the map gather (fallback) and no network term
                if (use_sur) mlp::field_from_outputs(md, c, o, sBz, sBfi, sBr);
            }
        }
        if (!live) { rf += s.dRF; continue; }
        moved = true;
        const float sc = (float)s.sc_s;
        const float esr = fmaf(sc, d_r, e0r), esz = fmaf(sc, d_z, e0z), esf = fmaf(sc, d_f, e0f);
        float Bz, Br, Bfi, Er, Efi, Ez, srf;
        fields_f32<ORD>(kc, rr, ir, th, fim, locd, zf, xf, yf, (float)s.v0, rf, Bz, Br, Bfi, Er, Efi, Ez, srf,
                   SUR ? (!use_sur && md->fallback != 0) : true);
        if constexpr (SUR) {
            Bz += sBz;
            Bfi += sBfi;
            Br += sBr;
        }
        const float cf = xf * ir, sf = yf * ir;
        const float ert = Er * srf + esr, eft = Efi * srf + esf, ezt = Ez * srf + esz;
        const double Ex = (double)(ert * cf - eft * sf), Ey = (double)(ert * sf + eft * cf), Ezc = (double)ezt;
        const double Bx = -(double)(Br * cf - Bfi * sf), By = -(double)(Br * sf + Bfi * cf), Bzc = -(double)Bz;
        if (aperture) {
            const double rd = (double)rr;
            if (rd > ap1 || rd < ap0 || z > ap3 || z < ap2) fl = (uint8_t)(fl & ~FFA_F_ALIVE);
        }
        // relativistic Boris on u = gamma v (src/FFAG_MathTools.py:1918-1966)
        const double dt = s.dt;
        const double eps = 0.5 * dt * q_over_m;
        double mx = ux + eps * Ex, my = uy + eps * Ey, mz = uz + eps * Ezc;
        // 1 / gamma at the half step only scales the rotation vector of the
        // FP32 field: single precision suffices (the rotation conserves |u|
        // for any t); the final 1 / gamma, which sets the velocity, is FP64
        const double ig = (double)__builtin_amdgcn_rsqf((float)(1.0 + (mx * mx + my * my + mz * mz) * inv_c2));
        const double tx = eps * ig * Bx, ty = eps * ig * By, tz = eps * ig * Bzc;
        const double pxp = mx + (my * tz - mz * ty), pyp = my + (mz * tx - mx * tz), pzp = mz + (mx * ty - my * tx);
        // 2 / (1 + t^2): |t| ~ omega dt / 2 is small, so the series to t^6
        // (relative error <= t^8 <= 1e-16) replaces the FP64 division
        const double t2 = tx * tx + ty * ty + tz * tz;
        double sc2 = 2.0 * (1.0 - t2 * (1.0 - t2 * (1.0 - t2)));
        if (t2 >= 1e-4) sc2 = 2.0 / (1.0 + t2);
        ux = mx + sc2 * (pyp * tz - pzp * ty) + eps * Ex;
        uy = my + sc2 * (pzp * tx - pxp * tz) + eps * Ey;
        uz = mz + sc2 * (pxp * ty - pyp * tx) + eps * Ezc;
        // 1 / gamma = (1 + u^2 / c^2)^(-1/2): hardware estimate and one
        // third-order correction (the argument is >= 1 and finite)
        const double g2 = 1.0 + (ux * ux + uy * uy + uz * uz) * inv_c2;
        const double y0 = __builtin_amdgcn_rsq(g2);
        const double eg = fma(-g2 * y0, y0, 1.0);
        const double ign = fma(y0 * eg, fma(eg, 0.375, 0.5), y0);
        vx = ux * ign; vy = uy * ign; vz = uz * ign;
        x += vx * dt; y += vy * dt; z += vz * dt;
        // azimuth advance: the chord from r to r + v dt turns by dth with
        // tan dth = (r x v) dt / (r^2 + (r . v) dt); |dth| ~ 1e-3, so FP32 and
        // atan t = t - t^3 / 3 leave ~1e-10 rad per step
        {
            const float vxf = (float)vx, vyf = (float)vy, dtf = (float)dt;
            const float cr = (xf * vyf - yf * vxf) * dtf;
            const float den = fmaf(xf * vxf + yf * vyf, dtf, r2f);
            const float tq = cr * rcpf(den);
            thd += (double)(tq - tq * tq * tq * (1.0f / 3.0f));
        }
        rf += s.dRF;
    }
    if (!inb) return;
    p.rf_phase[i] = rf;
    if (fl != p.flags[i]) p.flags[i] = fl;
    if (!moved) return;
    const double rn = sqrt(x * x + y * y);
    const double irn = 1.0 / rn;
    // unwrapped azimuth: the exact end angle, in the turn the carried one is in
    const double a_end = atan2(y, x);
    const double fin = fi0 + (a_end - th0 + FFA_TWO_PI * rint((thd - a_end) * (1.0 / FFA_TWO_PI)));
    const double cn = x * irn, sn = y * irn;
    p.r[i] = rn;
    p.vr[i] = vx * cn + vy * sn;
    p.z[i] = z;
    p.vz[i] = vz;
    p.fi[i] = fin;
    p.fidot[i] = (-vx * sn + vy * cn) * irn;
}

// ffa_bfield_eval mode 3: the static field exactly as the FP32 tier's push
// evaluates it at a position, widened to double: the FP64 Cartesian
// position rounded to FP32 for the radius, the FP64 azimuth reduced to the
// sector before rounding, fields_f32 on the FP32 map.
template <int ORD>
__global__ __launch_bounds__(BLOCK) void k_bfield_eval_f32(const ffa_fconst *__restrict__ kc, int nsectors,
                                                           const double *__restrict__ r, const double *__restrict__ fi,
                                                           const double *__restrict__ z, int64_t n,
                                                           double *__restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= n) return;
    double s, c;
    sincos(fi[i], &s, &c);
    const double x = r[i] * c, y = r[i] * s;
    const float xf = (float)x, yf = (float)y, zf = (float)z[i];
    float r2f, rr, ir;
    radius_f32(xf, yf, r2f, rr, ir);
    const double thd = atan2(y, x);
    const double Pd = FFA_TWO_PI / (double)nsectors;
    const double locd = thd - Pd * floor(thd * ((double)nsectors / FFA_TWO_PI));
    const float th = (float)thd, fim = (float)locd;
    float Bz, Br, Bfi, Er, Efi, Ez, srf;
    fields_f32<ORD>(kc, rr, ir, th, fim, locd, zf, xf, yf, 0.0f, 0.0, Bz, Br, Bfi, Er, Efi, Ez, srf);
    out[i] = (double)Bz;
    out[n + i] = (double)Bfi;
    out[2 * n + i] = (double)Br;
}

}  // namespace

hipError_t launch_bfield_eval_f32(ffa_ctx_s *c, const double *r, const double *fi, const double *z, int64_t n,
                                  double *out) {
    if (n == 0) return hipSuccess;
    // the push's instance: written-out order-6 gather for order-6 maps
    if (c->host_machine.max_order == 6)
        hipLaunchKernelGGL(k_bfield_eval_f32<6>, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                           c->stream, ctx_fconst_dev(c), c->host_machine.nsectors, r, fi, z, n, out);
    else
        hipLaunchKernelGGL(k_bfield_eval_f32<0>, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0,
                           c->stream, ctx_fconst_dev(c), c->host_machine.nsectors, r, fi, z, n, out);
    return hipGetLastError();
}

// The push without the surrogate: the endpoint type of the particles, the
// gather mode GATHER, and the kernel instance for order-6 maps when the
// machine has them (its own register budget; every other order runs the
// general instance).
template <int GATHER>
static void launch_push(ffa_ctx_s *c, const ffa_particles *p, unsigned blocks, const ffa_step *steps,
                        const int64_t *inj, int32_t nsteps, const mlp::fdesc *md, const sc_gather_src &g) {
    const bool ord6 = c->host_machine.max_order == 6;
    const ffa_machine *m = ctx_machine_dev(c);
    const ffa_fconst *k = ctx_fconst_dev(c);
    if (p->esc_f32) {
        if (ord6)
            hipLaunchKernelGGL((k_boris_fused<float, false, GATHER, 6>), dim3(blocks), dim3(BLOCK), 0, c->stream, *p,
                               m, k, steps, inj, nsteps, nullptr, md, g);
        else
            hipLaunchKernelGGL((k_boris_fused<float, false, GATHER, 0>), dim3(blocks), dim3(BLOCK), 0, c->stream, *p,
                               m, k, steps, inj, nsteps, nullptr, md, g);
    } else {
        if (ord6)
            hipLaunchKernelGGL((k_boris_fused<double, false, GATHER, 6>), dim3(blocks), dim3(BLOCK), 0, c->stream, *p,
                               m, k, steps, inj, nsteps, nullptr, md, g);
        else
            hipLaunchKernelGGL((k_boris_fused<double, false, GATHER, 0>), dim3(blocks), dim3(BLOCK), 0, c->stream, *p,
                               m, k, steps, inj, nsteps, nullptr, md, g);
    }
}

hipError_t launch_boris_fused_f32(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps) {
    if (p->n == 0 || nsteps == 0) return hipSuccess;
    const sc_gather_src gs{};
    const ffa_step *steps = reinterpret_cast<const ffa_step *>(c->dev + FFA_CTX_STEPS_OFF);
    const int64_t *inj = reinterpret_cast<const int64_t *>(c->dev + FFA_CTX_INJ_OFF);
    const mlp::fdesc *md = reinterpret_cast<const mlp::fdesc *>(c->dev + FFA_CTX_MLP_OFF);
    const char *w = static_cast<const char *>(c->mlp_packed);
    if (c->mlp_packed != nullptr) {
        const unsigned g = (unsigned)((p->n + SUR_BLOCK - 1) / SUR_BLOCK);
        if (p->esc_f32)
            hipLaunchKernelGGL((k_boris_fused<float, true>), dim3(g), dim3(SUR_BLOCK), 0, c->stream, *p,
                               ctx_machine_dev(c), ctx_fconst_dev(c), steps, inj, nsteps, w, md, gs);
        else
            hipLaunchKernelGGL((k_boris_fused<double, true>), dim3(g), dim3(SUR_BLOCK), 0, c->stream, *p,
                               ctx_machine_dev(c), ctx_fconst_dev(c), steps, inj, nsteps, w, md, gs);
        return hipGetLastError();
    }
    const unsigned g = (unsigned)((p->n + BLOCK - 1) / BLOCK);
    launch_push<0>(c, p, g, steps, inj, nsteps, md, gs);
    return hipGetLastError();
}

hipError_t launch_boris_fused_f32_gather(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps, const sc_gather_src &g,
                                         bool from_phi) {
    if (p->n == 0) return hipSuccess;
    const ffa_step *steps = reinterpret_cast<const ffa_step *>(c->dev + FFA_CTX_STEPS_OFF);
    const int64_t *inj = reinterpret_cast<const int64_t *>(c->dev + FFA_CTX_INJ_OFF);
    const mlp::fdesc *md = reinterpret_cast<const mlp::fdesc *>(c->dev + FFA_CTX_MLP_OFF);
    const unsigned blocks = (unsigned)((p->n + BLOCK - 1) / BLOCK);
    if (from_phi) launch_push<2>(c, p, blocks, steps, inj, nsteps, md, g);
    else launch_push<1>(c, p, blocks, steps, inj, nsteps, md, g);
    return hipGetLastError();
}
