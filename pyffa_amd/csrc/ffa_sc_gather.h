// Per-particle space-charge gather and endpoint roll, shared by the SC
// solve's gather kernel (ffa_sc.hip) and the FP32 push that applies a
// solve's gather in its own prologue (ffa_fused.hip).
//
// interpolate_fields_to_particles_SBS_numba (src/FFAG_SC_new.py:870-972):
// bilinear (Ex, Ez) in the particle's slice at (x, z) with floor and clamp,
// Esc_r = Ex / sqrt(1 + pr0^2), Esc_z = Ez, Esc_phi = -Ex pr0 / sqrt(1 + pr0^2),
// fused with the Pre/PrePre endpoint roll of src/FFAG_track.py:356-382.
//
// The FP64 arithmetic follows the reference's operation order with every
// product rounded on its own: each product that feeds a sum passes through
// an empty asm statement, so no compiler, whatever its contraction setting,
// can fuse it into a multiply-add.  The two translation units that include
// this file (one built without contraction, one with) therefore produce the
// same bits.
#pragma once
#include "ffa_internal.h"

namespace scg {

// A rounded FP64 value the optimiser cannot see through (no instruction).
__device__ __forceinline__ double rounded(double v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Scale factors of the stencils: central differences 1 / (2 d gamma^2),
// one-sided ones 1 / (d gamma^2) (the reference's x-edge and z-edge rows).
struct field_scale {
    float cx1, cx2, cz1, cz2;
};

__device__ __forceinline__ field_scale make_field_scale(double dr, double dz, double gamma) {
    const double ig2 = 1.0 / (gamma * gamma);
    return {(float)(ig2 / dr), (float)(0.5 / dr * ig2), (float)(ig2 / dz), (float)(0.5 / dz * ig2)};
}

// (Ex, Ez) of cell (j, i) from phi at the cell (pc), above and below it in z
// (pu, pd: rows clamped to the grid) and left and right of it in x (lf, rt:
// columns clamped to the row).  Shared by the field kernels and the gather
// that differentiates phi itself, so both give the same float32 values.  At
// the x edges the x difference is one-sided and the z difference is the
// clamped-row one scaled by 1 / (dz gamma^2) (the reference's edge rows);
// on the z edges (not at an x edge) the z difference is one-sided.
__device__ __forceinline__ float2 cell_field(float pc, float pu, float pd, float lf, float rt, int i, int j, int nx,
                                             int nz, const field_scale &f) {
    const bool xe = i == 0 || i == nx - 1;
    const bool z0 = !xe && j == 0, z1 = !xe && j == nz - 1;
    const float dx = i == 0 ? rt - pc : (i == nx - 1 ? pc - lf : rt - lf);
    const float dz = z0 ? pu - pc : (z1 ? pc - pd : pu - pd);
    return make_float2(-dx * (xe ? f.cx1 : f.cx2), -dz * ((xe || z0 || z1) ? f.cz1 : f.cz2));
}

// (Ex, Ez) at the gather's four cells (i0 .. i0 + 1, j0 .. j0 + 1, with
// 0 <= i0 <= nx - 2 and 0 <= j0 <= nz - 2) straight from phi: the 4 x 4
// stencil block minus its corners, rows and columns clamped to the grid as
// the field kernels clamp them.
__device__ __forceinline__ void cells_field_from_phi(const float *P, int i0, int j0, int nx, int nz,
                                                     const field_scale &f, float2 &a00, float2 &a10, float2 &a01,
                                                     float2 &a11) {
    int64_t row[4];
    int col[4];
    for (int k = 0; k < 4; ++k) {
        const int j = j0 - 1 + k, i = i0 - 1 + k;
        row[k] = (int64_t)(j < 0 ? 0 : (j > nz - 1 ? nz - 1 : j)) * nx;
        col[k] = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
    }
    const float v01 = P[row[0] + col[1]], v02 = P[row[0] + col[2]];
    const float v10 = P[row[1] + col[0]], v11 = P[row[1] + col[1]], v12 = P[row[1] + col[2]], v13 = P[row[1] + col[3]];
    const float v20 = P[row[2] + col[0]], v21 = P[row[2] + col[1]], v22 = P[row[2] + col[2]], v23 = P[row[2] + col[3]];
    const float v31 = P[row[3] + col[1]], v32 = P[row[3] + col[2]];
    a00 = cell_field(v11, v21, v01, v10, v12, i0, j0, nx, nz, f);
    a10 = cell_field(v12, v22, v02, v11, v13, i0 + 1, j0, nx, nz, f);
    a01 = cell_field(v21, v31, v11, v20, v22, i0, j0 + 1, nx, nz, f);
    a11 = cell_field(v22, v32, v12, v21, v23, i0 + 1, j0 + 1, nx, nz, f);
}

// Cell and bilinear weights of a particle: slice s clamped to [0, ns), the
// cell by floor of the position clamped to +-1e9 and then to
// [0, n - 2], weights from the unclamped offsets (they extrapolate outside
// the grid as the reference's do).
struct cellw {
    int64_t cb;            // (s, j0, i0) cell offset in the (ns, nz, nx) cube
    int ci, cj;
    double w00, w10, w01, w11;
};

__device__ __forceinline__ cellw cell_weights(int s, double xv, double zv, int ns, int nx, int nz, double xmin,
                                              double zmin, double idx, double idz) {
    cellw c;
    if (s < 0) s = 0; else if (s >= ns) s = ns - 1;
    const double t = (xv - xmin) * idx, v = (zv - zmin) * idz;
    const double ct = t > -1e9 ? (t < 1e9 ? t : 1e9) : -1e9, cv = v > -1e9 ? (v < 1e9 ? v : 1e9) : -1e9;
    int64_t i0 = (int64_t)floor(ct), j0 = (int64_t)floor(cv);
    if (i0 < 0) i0 = 0; else if (i0 > nx - 2) i0 = nx - 2;
    if (j0 < 0) j0 = 0; else if (j0 > nz - 2) j0 = nz - 2;
    const double al = t - (double)i0, be = v - (double)j0;
    c.w00 = (1.0 - al) * (1.0 - be);
    c.w10 = al * (1.0 - be);
    c.w01 = (1.0 - al) * be;
    c.w11 = al * be;
    c.ci = (int)i0;
    c.cj = (int)j0;
    c.cb = (int64_t)s * nz * nx + j0 * nx + i0;
    return c;
}

// The four (Ex, Ez) pairs of the cell: from the interleaved cubes, or
// differentiated from phi (FROM_PHI; same float32 values).
template <bool FROM_PHI>
__device__ __forceinline__ void corners(const cellw &c, const float2 *exz, const float *phi, int nx, int nz,
                                        const field_scale &fsc, float2 &a00, float2 &a10, float2 &a01, float2 &a11) {
    if constexpr (FROM_PHI) {
        const int64_t sb = c.cb - ((int64_t)c.cj * nx + c.ci);
        cells_field_from_phi(phi + sb, c.ci, c.cj, nx, nz, fsc, a00, a10, a01, a11);
    } else {
        a00 = exz[c.cb];
        a10 = exz[c.cb + 1];
        a01 = exz[c.cb + nx];
        a11 = exz[c.cb + nx + 1];
    }
}

// The solve's endpoint update of particle i (flags fl at the solve): the new
// pre-step field (a, b, c) = (Esc_r, Esc_z, Esc_phi) in the endpoint type,
// stored into the memory epp points to (the caller then swaps its pre and
// prepre pointers); on the first solve also into ep.  A live particle gets
// the gathered field; a dead one keeps its old pre-step field; on a
// degenerate grid a live particle with both endpoints keeps the field the
// step before the solve extrapolated (the reference leaves LocalBunch
// columns 10-12 as that step wrote them).  Returns the new flags.
template <typename ET>
__device__ __forceinline__ uint8_t roll(int64_t i, int64_t n, uint8_t fl, bool good, const cellw &c, float2 a00,
                                        float2 a10, float2 a01, float2 a11, double pr, const ffa_sc_params &sp,
                                        ET *ep, ET *epp, ET &a, ET &b, ET &cc) {
    const bool alive = fl & FFA_F_ALIVE;
    // the old pre-step field is read only where it survives the solve
    double o_r = 0.0, o_z = 0.0, o_f = 0.0;
    if (!(alive && good)) { o_r = (double)ep[i]; o_z = (double)ep[n + i]; o_f = (double)ep[2 * n + i]; }
    double nr = o_r, nz_ = o_z, nf = o_f;
    if (alive) {
        if (good) {
            const double er = rounded(c.w00 * a00.x) + rounded(c.w10 * a10.x) + rounded(c.w01 * a01.x) +
                              rounded(c.w11 * a11.x);
            const double ez = rounded(c.w00 * a00.y) + rounded(c.w10 * a10.y) + rounded(c.w01 * a01.y) +
                              rounded(c.w11 * a11.y);
            const double ip = 1.0 / sqrt(1.0 + rounded(pr * pr));
            nr = er * ip;
            nz_ = ez;
            nf = er * (-pr * ip);
        } else if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) {
            nr = o_r + rounded(sp.s_last * (o_r - (double)epp[i]));
            nz_ = o_z + rounded(sp.s_last * (o_z - (double)epp[n + i]));
            nf = o_f + rounded(sp.s_last * (o_f - (double)epp[2 * n + i]));
        }
    }
    a = (ET)nr;
    b = (ET)nz_;
    cc = (ET)nf;
    epp[i] = a; epp[n + i] = b; epp[2 * n + i] = cc;
    if (sp.first) {
        ep[i] = a; ep[n + i] = b; ep[2 * n + i] = cc;
        return (uint8_t)((fl & ~(FFA_F_SCPRE | FFA_F_SCPREPRE)) | (alive ? FFA_F_SCPRE : 0));
    }
    return (uint8_t)((fl & ~(FFA_F_SCPRE | FFA_F_SCPREPRE)) | ((fl & FFA_F_SCPRE) ? FFA_F_SCPREPRE : 0) |
                     (alive ? FFA_F_SCPRE : 0));
}

}  // namespace scg
