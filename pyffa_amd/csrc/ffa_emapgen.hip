// Spiral RF-gap field maps (gfx950).
//
// The gap's centre line is the log spiral phi(r) = s_ref + a ln(r / r_ref),
// a = tan(alpha), r in [r_min, r_max] (examples/main_Emap.py:64-69).  At each
// map node (r, phi) the field is a Gaussian of the distance d to the nearest
// point of that line, U / (delta sqrt(2 pi)) exp(-d^2 / (2 delta^2)), along
// the line's normal (projected to (Er, Ephi)); nodes whose nearest point is a
// line end get no field (main_Emap.py:865-931).  The nearest point comes from
// the reference's Newton iteration on r with clamping to the line's range
// (spiral_distance_direction_rphi_with_flag, :267-387); one node per thread.
// Ez/z follows from div E = 0 with the reference's one-sided / central
// differences (deriv_r, deriv_phi, :428-445).
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;

__global__ __launch_bounds__(BLOCK) void k_spiral_gap_field(ffa_spiral_gap g) {
    const int64_t idx = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    const int64_t total = (int64_t)g.nr * g.nphi;
    if (idx >= total) return;
    const int ir = (int)(idx / g.nphi), jp = (int)(idx - (int64_t)ir * g.nphi);
    const double r_p = g.r_axis[ir], phi_p = g.phi_axis[jp];
    const double a = g.a, one_plus_a2 = 1.0 + a * a;
    const double x_p = r_p * cos(phi_p), y_p = r_p * sin(phi_p);
    double r = r_p;
    if (r < g.r_min) r = g.r_min;
    if (r > g.r_max) r = g.r_max;
    if (r <= 0.0) r = g.r_min;
    for (int it = 0; it < g.n_iter; ++it) {
        const double s = g.s_ref + a * log(r / g.r_ref);
        const double cs = cos(s), sn = sin(s);
        const double x = r * cs, y = r * sn;
        const double x1 = cs - a * sn, y1 = sn + a * cs;
        const double invr = 1.0 / r;
        const double x2 = -(a * invr) * (sn + a * cs), y2 = (a * invr) * (cs - a * sn);
        const double dx = x - x_p, dy = y - y_p;
        const double gv = dx * x1 + dy * y1;
        const double gp = one_plus_a2 + dx * x2 + dy * y2;
        if (fabs(gp) < 1e-14) break;
        double r_new = r - gv / gp;
        if (r_new < g.r_min) r_new = g.r_min;
        if (r_new > g.r_max) r_new = g.r_max;
        if (fabs(r_new - r) < 1e-12 * (1.0 + r)) {
            r = r_new;
            break;
        }
        r = r_new;
    }
    const double eps_end = 1e-12 * (1.0 + g.r_max);
    const int flag = fabs(r - g.r_min) < eps_end ? 1 : (fabs(r - g.r_max) < eps_end ? 2 : 0);
    const double s = g.s_ref + a * log(r / g.r_ref);
    const double cs = cos(s), sn = sin(s);
    const double vx = x_p - r * cs, vy = y_p - r * sn;
    const double d = sqrt(vx * vx + vy * vy);
    const double inv_t = 1.0 / sqrt(one_plus_a2);
    const double tx = (cs - a * sn) * inv_t, ty = (sn + a * cs) * inv_t;
    const double nx = -ty * g.side, ny = tx * g.side;
    double er = 0.0, ef = 0.0;
    if (flag == 0) {
        const double eh = (g.U / (g.delta * sqrt(2.0 * FFA_PI))) * exp(-0.5 * (d / g.delta) * (d / g.delta));
        const double c = cos(phi_p), sp = sin(phi_p);
        er = eh * (nx * c + ny * sp);
        ef = eh * (-nx * sp + ny * c);
    }
    g.er[idx] = er;
    g.ephi[idx] = ef;
    g.dist[idx] = d;
    g.flag[idx] = flag;
}

// Ez/z = -(Er / r + dEr/dr + (1/r) dEphi/dphi); the derivatives are central
// inside and one-sided on the first and last row / column.
__global__ __launch_bounds__(BLOCK) void k_spiral_gap_ez(ffa_spiral_gap g) {
    const int64_t idx = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    const int64_t total = (int64_t)g.nr * g.nphi;
    if (idx >= total) return;
    const int nr = g.nr, np = g.nphi;
    const int ir = (int)(idx / np), jp = (int)(idx - (int64_t)ir * np);
    const double *Er = g.er, *Ef = g.ephi;
    double der, def;
    if (ir == 0) der = (Er[idx + np] - Er[idx]) / g.dr;
    else if (ir == nr - 1) der = (Er[idx] - Er[idx - np]) / g.dr;
    else der = (Er[idx + np] - Er[idx - np]) / (2.0 * g.dr);
    if (jp == 0) def = (Ef[idx + 1] - Ef[idx]) / g.dphi;
    else if (jp == np - 1) def = (Ef[idx] - Ef[idx - 1]) / g.dphi;
    else def = (Ef[idx + 1] - Ef[idx - 1]) / (2.0 * g.dphi);
    const double R = g.r_axis[ir];
    g.ez_over_z[idx] = -(Er[idx] / R + der + (1.0 / R) * def);
}

}  // namespace

hipError_t launch_spiral_gap(ffa_ctx_s *c, const ffa_spiral_gap *g) {
    const int64_t total = (int64_t)g->nr * g->nphi;
    const unsigned nb = (unsigned)((total + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_spiral_gap_field, dim3(nb), dim3(BLOCK), 0, c->stream, *g);
    hipLaunchKernelGGL(k_spiral_gap_ez, dim3(nb), dim3(BLOCK), 0, c->stream, *g);
    return hipGetLastError();
}
