// 2.5D space-charge solve of the tracking hot path (gfx950).
//
// Pipeline of one solve (every kernel on the context stream, no host sync):
//   k_sc_stage1        per survivor: gamma, phi mod 2pi; per-block partial
//                      sums of gamma and the four covering-arc min/max
//   k_sc_final1        1 block: covering arc, slice step, mean gamma,
//                      perimeter, bunch length (device-side scalars); zeroes
//                      the per-slice counters
//   k_sc_locate_moments per survivor: SEO r0/pr0, x = r - r0, slice index,
//                      per-block partial moments, slice histogram
//   k_sc_final2        1 block: grid box, dx, dz, charge scale, slice offsets
//   k_sc_bin           per survivor: CIC cell + fixed-point weights packed in
//                      one 64-bit record, written slice-sorted (two-level
//                      reservation: one global atomic per block and slice)
//   k_sc_rowsort       per slice: records ordered by CIC row, row offsets;
//                      the blocks also fill the per-solve FFT twiddles and
//                      Green's function table
//   k_sc_deposit_*     one workgroup per (slice, z band) up to 2^23 cells,
//                      per (slice pair, z band) above: LDS int64 accumulation
//                      of the records (integer sums: bit-identical for any
//                      order, no device-scope atomics), plain stores of the
//                      band: the int64 grid, or above 2^23 cells the scaled
//                      float32 charge packed as re/im into the FFT buffer
//   k_fft_rows_fwd*    row FFTs of the packed slice pairs (reading and
//                      scaling the int64 grid on small grids)
//   k_fft_cols_green   column FFT, x periodic Green's function, inverse
//   k_fft_rows_inv     inverse row FFTs, 1/(nx nz), unpack to phi
//   k_sc_field         -grad(phi)/gamma^2 with the reference's edge stencils,
//                      stored as (Ex, Ez) cubes (only when sp.field_cubes)
//   k_sc_gather_roll   bilinear gather + pr0 projection, endpoint roll; without
//                      the cubes it applies the same stencils to phi at the
//                      particle's four cells
// Particle-side arithmetic is FP64 in the reference's operation order (this
// unit is compiled with -ffp-contract=off); the grid is float32 like the
// reference's (complex64 FFTs).  Reductions use a fixed block partition and
// a fixed combine order, so a solve is bit-reproducible run to run.
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;
constexpr int NB = FFA_SC_BLOCKS;
constexpr int UNR = 4;       // particles per pass of the reduction loops
constexpr double EPS0 = 8.854187817e-12;

__device__ __forceinline__ double gamma_of(const ffa_particles &p, int64_t i, int rk, double c, double E0J) {
    if (rk) return p.fidot[i] / E0J;
    const double r = p.r[i], rd = r * p.fidot[i], vr = p.vr[i], vz = p.vz[i];
    const double v2 = vr * vr + vz * vz + rd * rd;
    return 1.0 / sqrt(1.0 - v2 / (c * c));
}

__device__ __forceinline__ double mod2pi(double fi) { return fi - FFA_TWO_PI * floor(fi / FFA_TWO_PI); }

__device__ __forceinline__ double combine(int op /*0 sum,1 min,2 max*/, double a, double b) {
    return op == 0 ? a + b : (op == 1 ? (b < a ? b : a) : (b > a ? b : a));
}

// Block-wide sum / min / max of NV values: an xor-shuffle tree inside each
// wave (both partners of a pair compute the same commutative combine, so
// every lane ends with the wave's result), then the BLOCK / 64 wave results
// in wave order.  Fixed pairing and order: bit-reproducible run to run.
// Every thread returns the block's values; contains one __syncthreads.
template <int NV>
__device__ __forceinline__ void block_reduce(double (*sh)[BLOCK / 64], double *v, const int *op) {
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = combine(op[k], v[k], __shfl_xor(v[k], off));
    const int wave = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NV; ++k) sh[k][wave] = v[k];
    __syncthreads();
    for (int k = 0; k < NV; ++k) {
        double a = sh[k][0];
        for (int w = 1; w < BLOCK / 64; ++w) a = combine(op[k], a, sh[k][w]);
        v[k] = a;
    }
}

// Wave-aggregated LDS counter update: lanes holding the same key (the
// common case after the azimuth ordering, when a wave's particles share a
// slice) are served by one atomic from their first lane, which returns the
// group's base; each lane's value is base + its rank among the group's
// lower lanes.  Up to PEEL distinct keys per wave are grouped; the rest fall
// back to one atomic per lane.  Returns the counter's value before this
// lane's increment.
constexpr int PEEL = 4;

__device__ __forceinline__ int32_t wave_counter_add(int32_t *ctr, int key, bool valid) {
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    bool todo = valid;
    int32_t out = 0;
    for (int k = 0; k < PEEL; ++k) {
        const uint64_t act = __ballot(todo);
        if (act == 0ull) break;                                  // wave-uniform
        const int leader = __builtin_ctzll(act);
        const int k0 = __shfl(key, leader);
        const bool mine = todo && key == k0;
        const uint64_t grp = __ballot(mine);
        int32_t base = 0;
        if (lane == leader) base = atomicAdd(&ctr[k0], (int32_t)__popcll(grp));
        base = __shfl(base, leader);
        if (mine) out = base + (int32_t)__popcll(grp & below);
        todo = todo && !mine;
    }
    if (todo) out = atomicAdd(&ctr[key], 1);
    return out;
}

// GetLocalCoordinates_2p5(_rk) + divide_arc2's scan (src/FFAG_SC_new.py:74-148, 534-566).
__global__ __launch_bounds__(BLOCK) void k_sc_stage1(ffa_particles p, int rk, double c, double E0J,
                                                    double *part) {
    __shared__ double sh[10][BLOCK / 64];
    const double shifts[4] = {0.0, FFA_PI, FFA_PI * 0.5, FFA_PI * 1.5};
    double v[10] = {0.0, 0.0, 1e30, 1e30, 1e30, 1e30, -1e30, -1e30, -1e30, -1e30};
    // UNR particles per pass with their loads issued together; a thread
    // still visits its particles i, i + S, i + 2S, ... in that order
    const int64_t S = (int64_t)NB * BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i0 < p.n; i0 += UNR * S) {
        // flags, then every state load, then the arithmetic: no branch between
        // the particles' loads, so their round trips overlap
        bool ok[UNR];
        int64_t jj[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int64_t i = i0 + u * S;
            const uint8_t f = p.flags[i < p.n ? i : 0];
            ok[u] = (i < p.n) & ((f & FFA_F_ALIVE) != 0);
            jj[u] = ok[u] ? i : 0;
        }
        double r[UNR], fd[UNR], vr[UNR], vz[UNR], fi[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            r[u] = p.r[jj[u]];
            fd[u] = p.fidot[jj[u]];
            vr[u] = p.vr[jj[u]];
            vz[u] = p.vz[jj[u]];
            fi[u] = p.fi[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if (!ok[u]) continue;
            double g;                                  // gamma_of on the loaded state
            if (rk) {
                g = fd[u] / E0J;
            } else {
                const double rd = r[u] * fd[u];
                const double v2 = vr[u] * vr[u] + vz[u] * vz[u] + rd * rd;
                g = 1.0 / sqrt(1.0 - v2 / (c * c));
            }
            v[0] += g;
            v[1] += 1.0;
            const double fm = mod2pi(fi[u]);
            for (int k = 0; k < 4; ++k) {
                const double a = fm + shifts[k], b = 2.0 * FFA_PI;
                const double q = a - b * floor(a / b);
                v[2 + k] = q < v[2 + k] ? q : v[2 + k];
                v[6 + k] = q > v[6 + k] ? q : v[6 + k];
            }
        }
    }
    const int op[10] = {0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
    block_reduce<10>(sh, v, op);
    if (threadIdx.x == 0)
        for (int k = 0; k < 10; ++k) part[16 * blockIdx.x + k] = v[k];
}

__device__ double py_mod(double a, double b) {  // Python float % (sign of divisor)
    double r = fmod(a, b);
    if (r != 0.0 && ((r < 0.0) != (b < 0.0))) r += b;
    return r;
}

__device__ double np_interp(double x, const double *xp, const double *fp, int n) {
    if (n == 1) return fp[0];
    if (x < xp[0]) return fp[0];
    if (x > xp[n - 1]) return fp[n - 1];
    if (x == xp[n - 1]) return fp[n - 1];
    int lo = 0, hi = n - 1;            // largest j with xp[j] <= x
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= x) lo = mid; else hi = mid;
    }
    const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
    double r = slope * (x - xp[lo]) + fp[lo];
    if (r != r) r = slope * (x - xp[lo + 1]) + fp[lo + 1];
    return r;
}

// Covering arc (divide_arc2 + merge_multiple_arcs for one rank,
// src/FFAG_SC_new.py:534-606; src/FFAG_Bunch.py:1203-1215), mean gamma and
// bunch length (src/FFAG_Bunch.py:1217-1294).
constexpr int EK_LDS = 1024;

__global__ __launch_bounds__(BLOCK) void k_sc_final1(const double *part, ffa_sc_params sp, ffa_sc_result *res,
                                                    int32_t *slc) {
    __shared__ double sh[10][BLOCK / 64];
    __shared__ double ek_ax[EK_LDS], ek_per[EK_LDS];
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK) slc[k] = 0;
    // the perimeter table is searched by one thread: stage it in LDS so the
    // binary search does not wait on a global load per probe
    const bool stage = sp.n_ek <= EK_LDS;
    if (stage)
        for (int k = threadIdx.x; k < sp.n_ek; k += BLOCK) {
            ek_ax[k] = sp.seo_ek_axis[k];
            ek_per[k] = sp.seo_perimeter[k];
        }
    double v[10] = {0.0, 0.0, 1e30, 1e30, 1e30, 1e30, -1e30, -1e30, -1e30, -1e30};
    const int op[10] = {0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
    for (int b = threadIdx.x; b < NB; b += BLOCK) {
        const double *q = part + 16 * b;
        v[0] += q[0];
        v[1] += q[1];
        for (int k = 2; k < 6; ++k) v[k] = q[k] < v[k] ? q[k] : v[k];
        for (int k = 6; k < 10; ++k) v[k] = q[k] > v[k] ? q[k] : v[k];
    }
    block_reduce<10>(sh, v, op);
    if (threadIdx.x != 0) return;
    const double shifts[4] = {0.0, FFA_PI, FFA_PI * 0.5, FFA_PI * 1.5};
    int best = 0;
    double bs = v[6] - v[2];
    for (int k = 1; k < 4; ++k)
        if (v[6 + k] - v[2 + k] < bs) { bs = v[6 + k] - v[2 + k]; best = k; }
    double s = v[2 + best] - shifts[best], e = v[6 + best] - shifts[best];
    if (e < s) e += 2.0 * FFA_PI;
    double fs = py_mod(s, 2.0 * FFA_PI), fe = py_mod(e, 2.0 * FFA_PI);
    if (fe <= fs) fe += 2.0 * FFA_PI;
    const double cnt = v[1];
    const double mgm1 = cnt > 0.0 ? v[0] / cnt - 1.0 : 0.0;
    const double perim = stage ? np_interp(mgm1 * sp.E0_MeV, ek_ax, ek_per, sp.n_ek)
                               : np_interp(mgm1 * sp.E0_MeV, sp.seo_ek_axis, sp.seo_perimeter, sp.n_ek);
    res->count = cnt;
    res->sum_gamma = v[0];
    for (int k = 0; k < 4; ++k) { res->arc_min[k] = v[2 + k]; res->arc_max[k] = v[6 + k]; }
    res->fi_s = fs;
    res->fi_e = fe;
    res->slice_step = (fe - fs) / (double)(sp.ns - 1);
    res->mean_gamma_m1 = mgm1;
    res->perimeter = perim;
    res->bunch_length = perim * (fe - fs) / FFA_PI / 2.0;
}

// Bilinear_interp_2D_vect_uniform (src/FFAG_MathTools.py:827-936).
__device__ __forceinline__ double seo_bilinear(const double *Z, int NX, int NY, double X0, double dX, double Y0,
                                               double dY, double x, double y) {
    double fx = (x - X0) / dX, fy = (y - Y0) / dY;
    fx = fx > -1e9 ? (fx < 1e9 ? fx : 1e9) : -1e9;
    fy = fy > -1e9 ? (fy < 1e9 ? fy : 1e9) : -1e9;
    int64_t i = (int64_t)floor(fx), j = (int64_t)floor(fy);
    if (i < 0) i = 0; else if (i >= NX - 1) i = NX - 2;
    if (j < 0) j = 0; else if (j >= NY - 1) j = NY - 2;
    const double xx0 = X0 + (double)i * dX, yy0 = Y0 + (double)j * dY;
    const double tx = (x - xx0) / dX, ty = (y - yy0) / dY;
    const double z00 = Z[i * NY + j], z01 = Z[i * NY + j + 1], z10 = Z[(i + 1) * NY + j], z11 = Z[(i + 1) * NY + j + 1];
    return (1 - tx) * (1 - ty) * z00 + (1 - tx) * ty * z01 + tx * (1 - ty) * z10 + tx * ty * z11;
}

// SEO lookup + ComputeSliceBoundsAndCoordsSBS_merged (src/FFAG_SC_new.py:449-531).
constexpr int MAX_NS_LDS = 4096;

// Particles per pass of k_sc_locate_moments: two keeps it under 128 VGPRs,
// so all NB blocks are resident at once (the visiting order i0, i0 + S, ...
// and hence the sums are the same for any value).
constexpr int LUNR = 2;

// Per survivor: x = r - r0(Ek, phi), pr0 and the slice index (written for
// the bin and gather kernels), the moments (sum x, z, x^2, z^2) and the
// slice histogram: per-block counts in LDS, one global atomic per non-empty
// (block, slice) pair (ns <= FFA_SC_MAX_NS).  Fixed block partition and
// per-thread order, so the sums are reproducible run to run; LUNR particles
// per pass have their loads (state, then the SEO table corners) in flight
// together.
__global__ __launch_bounds__(BLOCK) void k_sc_locate_moments(ffa_particles p, ffa_sc_params sp, double c,
                                                            double E0J, const ffa_sc_result *res, double *px,
                                                            double *ppr0, int32_t *pslice, int32_t *slc,
                                                            int32_t *blk_off, double *part) {
    __shared__ double sh[4][BLOCK / 64];
    __shared__ int32_t hist[MAX_NS_LDS];
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK) hist[k] = 0;
    __syncthreads();
    const double fs = res->fi_s, step = res->slice_step;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t S = (int64_t)NB * BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i0 < p.n; i0 += LUNR * S) {
        // phases over the LUNR particles (flags; state; SEO corners; results),
        // all loads ahead of the stores so the particles' memory round trips
        // overlap
        bool ok[LUNR];
        int64_t jj[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            const int64_t i = i0 + u * S;
            const uint8_t f = p.flags[i < p.n ? i : 0];
            ok[u] = (i < p.n) & ((f & FFA_F_ALIVE) != 0);
            jj[u] = ok[u] ? i : 0;
        }
        double ek[LUNR], fm[LUNR], x[LUNR], z[LUNR], fd[LUNR], vr[LUNR], vz[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            x[u] = p.r[jj[u]];
            fd[u] = p.fidot[jj[u]];
            vr[u] = p.vr[jj[u]];
            vz[u] = p.vz[jj[u]];
            fm[u] = p.fi[jj[u]];
            z[u] = p.z[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            // gamma_of's arithmetic on the loaded state
            double g;
            if (sp.rk) {
                g = fd[u] / E0J;
            } else {
                const double rd = x[u] * fd[u];
                const double v2 = vr[u] * vr[u] + vz[u] * vz[u] + rd * rd;
                g = 1.0 / sqrt(1.0 - v2 / (c * c));
            }
            ek[u] = (g - 1) * sp.E0_MeV;
            fm[u] = mod2pi(fm[u]);
        }
        double pr0[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            x[u] -= seo_bilinear(sp.seo_r, sp.n_ek, sp.n_fi, sp.ek0, sp.dek, sp.fi0, sp.dfi, ek[u], fm[u]);
            pr0[u] = seo_bilinear(sp.seo_pr, sp.n_ek, sp.n_fi, sp.ek0, sp.dek, sp.fi0, sp.dfi, ek[u], fm[u]);
        }
        int s[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            double f = fm[u];
            if (f < fs) f += FFA_TWO_PI;
            double fsl = (f - fs) / step;
            fsl = fsl > -1e9 ? (fsl < 1e9 ? fsl : 1e9) : -1e9;
            int64_t sl = (int64_t)fsl;
            if (sl < 0) sl = 0; else if (sl >= sp.ns) sl = sp.ns - 1;
            s[u] = ok[u] ? (int)sl : -1;
            const int64_t i = i0 + u * S;
            if (i < p.n) {
                pslice[i] = s[u];
                if (ok[u]) { px[i] = x[u]; ppr0[i] = pr0[u]; }
            }
        }
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            if (s[u] >= 0) { v[0] += x[u]; v[1] += z[u]; v[2] += x[u] * x[u]; v[3] += z[u] * z[u]; }
            wave_counter_add(hist, s[u] >= 0 ? s[u] : 0, s[u] >= 0);
        }
    }
    const int op[4] = {0, 0, 0, 0};
    block_reduce<4>(sh, v, op);       // contains __syncthreads: hist is complete
    // the slice counter's value before this block's add is where the block's
    // run of records starts inside the slice (entries of slices the block
    // does not hold stay unwritten and are never read)
    int32_t *bo = blk_off + (size_t)blockIdx.x * sp.ns;
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK)
        if (hist[k]) bo[k] = atomicAdd(&slc[k], hist[k]);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[16 * blockIdx.x + 10 + k] = v[k];
}

// Grid box from the moments (src/FFAG_Bunch.py:1259-1322).
__global__ __launch_bounds__(BLOCK) void k_sc_final2(const double *part, ffa_sc_params sp, ffa_sc_result *res,
                                                    const int32_t *slc, int32_t *slice_start) {
    __shared__ double sh[4][BLOCK / 64];
    __shared__ int32_t tot[BLOCK];
    double v[4] = {0.0, 0.0, 0.0, 0.0};       // loads issued before the scan's
    for (int b = threadIdx.x; b < NB; b += BLOCK)
        for (int k = 0; k < 4; ++k) v[k] += part[16 * b + 10 + k];
    {   // exclusive scan of the slice counts (ns <= 4096): each thread owns a
        // contiguous run of ceil(ns / BLOCK) slices
        const int per = (sp.ns + BLOCK - 1) / BLOCK;
        const int k0 = threadIdx.x * per, k1 = min(k0 + per, sp.ns);
        int32_t own = 0;
        for (int k = k0; k < k1; ++k) own += slc[k];
        tot[threadIdx.x] = own;
        __syncthreads();
        for (int w = 1; w < BLOCK; w <<= 1) {      // inclusive Hillis-Steele scan of the run totals
            const int32_t add = threadIdx.x >= w ? tot[threadIdx.x - w] : 0;
            __syncthreads();
            tot[threadIdx.x] += add;
            __syncthreads();
        }
        int32_t acc = tot[threadIdx.x] - own;
        for (int k = k0; k < k1; ++k) {
            slice_start[k] = acc;
            acc += slc[k];
        }
        if (threadIdx.x == BLOCK - 1) slice_start[sp.ns] = tot[BLOCK - 1];
    }
    const int op[4] = {0, 0, 0, 0};
    block_reduce<4>(sh, v, op);
    if (threadIdx.x != 0) return;
    const double cnt = res->count;
    double mx, mz, sx, sz;
    if (cnt > 1.0) {
        mx = v[0] / cnt;
        mz = v[1] / cnt;
        const double vx = v[2] / cnt - mx * mx, vz = v[3] / cnt - mz * mz;
        sx = sqrt(vx > 0.0 ? vx : 0.0);
        sz = sqrt(vz > 0.0 ? vz : 0.0);
    } else {
        mx = mz = sx = sz = 0.0;
    }
    res->sum_x = v[0]; res->sum_z = v[1]; res->sum_x2 = v[2]; res->sum_z2 = v[3];
    const double xmin = mx - sp.sigma_ratio * sx, xmax = mx + sp.sigma_ratio * sx;
    const double zmin = mz - sp.sigma_ratio * sz, zmax = mz + sp.sigma_ratio * sz;
    const double dx = (xmax - xmin) / sp.nx, dz = (zmax - zmin) / sp.nz;
    const double bl = res->bunch_length;
    res->xmin = xmin; res->xmax = xmax; res->zmin = zmin; res->zmax = zmax;
    res->dx = dx; res->dz = dz;
    const bool bad = !(dx > 0.0) || !(dz > 0.0) || !(bl > 0.0) || cnt <= 0.0;
    res->bad_grid = bad ? 1.0 : 0.0;
    res->qscale = bad ? 0.0 : (sp.q_macro / bl * sp.ns) * (1.0 / (dx * dz));
}

// CIC cell and fixed-point weights of one particle (DistributeChargeSBS_Njit
// cell/weight rule, src/FFAG_SC_new.py:244-312): a = rint(fx 2^bits),
// b = rint(fz 2^bits), packed as ix | iz << 10 | a << 20 | b << 41.
__device__ __forceinline__ uint64_t cic_record(double x, double z, const ffa_sc_result *res, int nx, int nz, int bits) {
    const double inv_dx = 1.0 / res->dx, inv_dz = 1.0 / res->dz;
    const double rx = (x - res->xmin) * inv_dx, rz = (z - res->zmin) * inv_dz;
    const double crx = rx > -1e9 ? (rx < 1e9 ? rx : 1e9) : -1e9;
    const double crz = rz > -1e9 ? (rz < 1e9 ? rz : 1e9) : -1e9;
    int64_t ix = (int64_t)floor(crx), iz = (int64_t)floor(crz);
    double fx = rx - (double)ix, fz = rz - (double)iz;
    if (ix < 0) { ix = 0; fx = 0.0; } else if (ix >= nx - 1) { ix = nx - 2; fx = 1.0; }
    if (iz < 0) { iz = 0; fz = 0.0; } else if (iz >= nz - 1) { iz = nz - 2; fz = 1.0; }
    const double scale = (double)((int64_t)1 << bits);
    const uint64_t a = (uint64_t)rint(fx * scale), b = (uint64_t)rint(fz * scale);
    return (uint64_t)ix | ((uint64_t)iz << 10) | (a << 20) | (b << 41);
}

// Bin the survivors into slice-sorted records.  Same block partition and
// per-thread order as k_sc_locate_moments, which counted each block's
// particles per slice and reserved the block's run inside each slice: the
// LDS counters start at the runs' global offsets and each record takes the
// next place of its slice's run.
__global__ __launch_bounds__(BLOCK) void k_sc_bin(ffa_particles p, ffa_sc_params sp, const ffa_sc_result *res,
                                                 const double *px, const int32_t *pslice, const int32_t *slice_start,
                                                 const int32_t *blk_off, uint64_t *rec) {
    __shared__ int32_t place[MAX_NS_LDS];
    if (res->bad_grid != 0.0) return;
    const int32_t *bo = blk_off + (size_t)blockIdx.x * sp.ns;
    // entries of slices the block holds no particle of are stale; unsigned
    // sums keep those (never used) entries well defined
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK) place[k] = (int32_t)((uint32_t)slice_start[k] + (uint32_t)bo[k]);
    __syncthreads();
    const int64_t S = (int64_t)NB * BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i0 < p.n; i0 += UNR * S) {
        int s[UNR];
        double x[UNR], z[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int64_t i = i0 + u * S;
            s[u] = i < p.n ? pslice[i] : -1;
            const int64_t j = s[u] >= 0 ? i : 0;
            x[u] = s[u] >= 0 ? px[j] : 0.0;
            z[u] = s[u] >= 0 ? p.z[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int pos = wave_counter_add(place, s[u] >= 0 ? s[u] : 0, s[u] >= 0);
            if (s[u] >= 0) rec[pos] = cic_record(x[u], z[u], res, sp.nx, sp.nz, sp.bits);
        }
    }
}

// Per-solve tables (tables region): twiddles w^k for k < N of the x and z
// transforms, forward and inverse, at [0..4) * 1024, then the periodic
// Green's function (make_fft_kernel_open_numba, src/FFAG_SC_new.py:731-760)
// in the reference's float32 arithmetic, transposed to [ci][j] so a column
// group reads one contiguous run.
constexpr int TW_STRIDE = 1024;

__device__ __forceinline__ void sc_table_entry(int64_t t, float2 *tw, float *green, const ffa_sc_params &sp,
                                               const ffa_sc_result *res) {
    const int nx = sp.nx, nz = sp.nz;
    if (t < 4 * TW_STRIDE) {
        const int which = (int)(t / TW_STRIDE), k = (int)(t - which * TW_STRIDE);
        const int N = which < 2 ? nx : nz;
        if (k < N) {
            double s, c;
            sincospi(-2.0 * (double)k / (double)N, &s, &c);
            tw[t] = make_float2((float)c, (float)((which & 1) ? -s : s));
        }
        return;
    }
    const int64_t g = t - 4 * TW_STRIDE;
    if (g >= (int64_t)nx * nz) return;
    const int ci = (int)(g / nz), j = (int)(g - (int64_t)ci * nz);
    const float dkx = (float)(2.0 * FFA_PI / (nx * res->dx));
    const float dkz = (float)(2.0 * FFA_PI / (nz * res->dz));
    const int idx = ci <= nx / 2 ? ci : ci - nx;
    const int idz = j <= nz / 2 ? j : j - nz;
    const float kx = (float)((double)dkx * idx), kz = (float)((double)dkz * idz);
    const float kz2 = kz * kz;
    const float k2 = kx * kx + kz2;
    green[g] = k2 == 0.0f ? 0.0f : (float)(1.0 / (EPS0 * (double)k2));
}

// One workgroup per slice: order the slice's records by CIC row iz (LDS
// counting sort; the order inside a row is immaterial, the deposit sums
// integers) and record where each row starts, so a band of rows is one
// contiguous range of records.
constexpr int ROW_BLOCK = 1024;

__global__ __launch_bounds__(ROW_BLOCK) void k_sc_rowsort(ffa_sc_params sp, const ffa_sc_result *res,
                                                         const int32_t *slice_start, const uint64_t *rec,
                                                         uint64_t *rec2, int32_t *row_start, float2 *tw,
                                                         float *green) {
    __shared__ int32_t cnt[1024];
    __shared__ int32_t tot[ROW_BLOCK];
    {   // this block's share of the per-solve tables (read by the FFT kernels)
        const int64_t n_tab = 4 * TW_STRIDE + (int64_t)sp.nx * sp.nz;
        for (int64_t t = blockIdx.x * (int64_t)ROW_BLOCK + threadIdx.x; t < n_tab; t += (int64_t)gridDim.x * ROW_BLOCK)
            sc_table_entry(t, tw, green, sp, res);
    }
    const int nz = sp.nz, s = blockIdx.x;
    int32_t *rs = row_start + (size_t)s * (nz + 1);
    const int r0 = slice_start[s], r1 = slice_start[s + 1];
    const bool ok = res->bad_grid == 0.0;
    for (int k = threadIdx.x; k < nz; k += ROW_BLOCK) cnt[k] = 0;
    __syncthreads();
    if (ok)
        for (int t = r0 + threadIdx.x; t < r1; t += ROW_BLOCK) atomicAdd(&cnt[(int)((rec[t] >> 10) & 1023u)], 1);
    __syncthreads();
    // exclusive scan of the nz <= 1024 row counts
    const int own = threadIdx.x < nz ? cnt[threadIdx.x] : 0;
    tot[threadIdx.x] = own;
    __syncthreads();
    for (int w = 1; w < ROW_BLOCK; w <<= 1) {
        const int add = threadIdx.x >= w ? tot[threadIdx.x - w] : 0;
        __syncthreads();
        tot[threadIdx.x] += add;
        __syncthreads();
    }
    const int excl = tot[threadIdx.x] - own;
    __syncthreads();
    if (threadIdx.x < nz) {
        cnt[threadIdx.x] = r0 + excl;
        rs[threadIdx.x] = r0 + excl;
    }
    if (threadIdx.x == 0) rs[nz] = ok ? r1 : r0;
    __syncthreads();
    if (ok)
        for (int t = r0 + threadIdx.x; t < r1; t += ROW_BLOCK) {
            const uint64_t w = rec[t];
            rec2[atomicAdd(&cnt[(int)((w >> 10) & 1023u)], 1)] = w;
        }
}

// Grids of up to 2^23 cells: one workgroup per (slice, band of bz rows):
// zero an LDS tile, add the corners of the records of rows z0 - 1 .. z1 - 1
// that fall in the band (a contiguous range after k_sc_rowsort), store the
// band of the int64 grid (k_fft_rows_fwd_grid scales it).  Dynamic LDS:
// bz * nx int64 sums, then (when sp.cell_counts) bz * nx int32 base-cell
// counts.
__global__ __launch_bounds__(BLOCK) void k_sc_deposit_slices(ffa_sc_params sp, const ffa_sc_result *res,
                                                           const int32_t *row_start, const uint64_t *rec,
                                                           unsigned long long *grid, int32_t *cell, int bz) {
    extern __shared__ unsigned long long lds[];
    const int nx = sp.nx, nz = sp.nz;
    const int nbands = nz / bz;
    const int s = blockIdx.x / nbands, band = blockIdx.x - s * nbands;
    const int z0 = band * bz, z1 = z0 + bz;
    const int ncell = bz * nx;
    const bool counts = sp.cell_counts != 0;
    unsigned long long *acc = lds;
    int32_t *cnt = reinterpret_cast<int32_t *>(lds + ncell);
    for (int k = threadIdx.x; k < ncell; k += BLOCK) acc[k] = 0ull;
    if (counts)
        for (int k = threadIdx.x; k < ncell; k += BLOCK) cnt[k] = 0;
    __syncthreads();
    if (res->bad_grid == 0.0) {
        const uint64_t one = (uint64_t)1 << sp.bits;
        const int32_t *rs = row_start + (size_t)s * (nz + 1);
        const int r0 = rs[z0 > 0 ? z0 - 1 : 0], r1 = rs[z1];
        for (int t = r0 + threadIdx.x; t < r1; t += BLOCK) {
            const uint64_t w = rec[t];
            const int ix = (int)(w & 1023u), iz = (int)((w >> 10) & 1023u);
            const uint64_t a = (w >> 20) & 0x1FFFFFu, b = (w >> 41) & 0x1FFFFFu;
            if (iz >= z0 && iz < z1) {
                const int o = (iz - z0) * nx + ix;
                atomicAdd(&acc[o], (one - a) * (one - b));
                atomicAdd(&acc[o + 1], a * (one - b));
                if (counts) atomicAdd(&cnt[o], 1);
            }
            if (iz + 1 >= z0 && iz + 1 < z1) {
                const int o = (iz + 1 - z0) * nx + ix;
                atomicAdd(&acc[o], (one - a) * b);
                atomicAdd(&acc[o + 1], a * b);
            }
        }
    }
    __syncthreads();
    const size_t base = (size_t)s * nz * nx + (size_t)z0 * nx;
    for (int k = threadIdx.x; k < ncell; k += BLOCK) grid[base + k] = acc[k];
    if (counts)
        for (int k = threadIdx.x; k < ncell; k += BLOCK) cell[base + k] = cnt[k];
}

// Grids above 2^23 cells: one workgroup per (slice pair, band of bz rows):
// zero an LDS tile per
// slice, add the corners of the records of rows z0 - 1 .. z1 - 1 that fall
// in the band (a contiguous range of each slice after k_sc_rowsort), then
// store the band as the packed FFT input: charge x scale in float32, the pair
// as (re, im), in the rounding the row FFT applied when it read the integer
// grid.  The int64 grid (sp.charge_grid) and the per-cell counts
// (sp.cell_counts) are stored only on request.  Dynamic LDS: 2 bz nx int64
// sums, then (with counts) 2 bz nx int32 base-cell counts.
__global__ __launch_bounds__(BLOCK) void k_sc_deposit_pairs(ffa_sc_params sp, const ffa_sc_result *res,
                                                           const int32_t *row_start, const uint64_t *rec,
                                                           unsigned long long *grid, int32_t *cell, float2 *cbuf,
                                                           int bz) {
    extern __shared__ unsigned long long lds[];
    const int nx = sp.nx, nz = sp.nz;
    const int nbands = nz / bz;
    const int pl = blockIdx.x / nbands, band = blockIdx.x - pl * nbands;
    const int z0 = band * bz, z1 = z0 + bz;
    const int ncell = bz * nx;
    const bool counts = sp.cell_counts != 0;
    const int npair = 2 * pl + 1 < sp.ns ? 2 : 1;          // an odd ns leaves the last plane one slice
    unsigned long long *acc = lds;
    int32_t *cnt = reinterpret_cast<int32_t *>(lds + 2 * ncell);
    for (int k = threadIdx.x; k < 2 * ncell; k += BLOCK) acc[k] = 0ull;
    if (counts)
        for (int k = threadIdx.x; k < 2 * ncell; k += BLOCK) cnt[k] = 0;
    __syncthreads();
    if (res->bad_grid == 0.0) {
        const uint64_t one = (uint64_t)1 << sp.bits;
        for (int t = 0; t < npair; ++t) {
            const int s = 2 * pl + t;
            unsigned long long *a_t = acc + t * ncell;
            int32_t *c_t = cnt + t * ncell;
            const int32_t *rs = row_start + (size_t)s * (nz + 1);
            const int r0 = rs[z0 > 0 ? z0 - 1 : 0], r1 = rs[z1];
            for (int q = r0 + threadIdx.x; q < r1; q += BLOCK) {
                const uint64_t w = rec[q];
                const int ix = (int)(w & 1023u), iz = (int)((w >> 10) & 1023u);
                const uint64_t a = (w >> 20) & 0x1FFFFFu, b = (w >> 41) & 0x1FFFFFu;
                if (iz >= z0 && iz < z1) {
                    const int o = (iz - z0) * nx + ix;
                    atomicAdd(&a_t[o], (one - a) * (one - b));
                    atomicAdd(&a_t[o + 1], a * (one - b));
                    if (counts) atomicAdd(&c_t[o], 1);
                }
                if (iz + 1 >= z0 && iz + 1 < z1) {
                    const int o = (iz + 1 - z0) * nx + ix;
                    atomicAdd(&a_t[o], (one - a) * b);
                    atomicAdd(&a_t[o + 1], a * b);
                }
            }
        }
    }
    __syncthreads();
    const double us = res->bad_grid != 0.0 ? 0.0 : res->qscale / (double)((int64_t)1 << (2 * sp.bits));
    float2 *cb = cbuf + (size_t)pl * nz * nx + (size_t)z0 * nx;
    for (int k = threadIdx.x; k < ncell; k += BLOCK) {
        const float re = (float)((double)(int64_t)acc[k] * us);
        const float im = npair == 2 ? (float)((double)(int64_t)acc[ncell + k] * us) : 0.0f;
        cb[k] = make_float2(re, im);
    }
    for (int t = 0; t < npair; ++t) {
        const size_t base = (size_t)(2 * pl + t) * nz * nx + (size_t)z0 * nx;
        if (sp.charge_grid)
            for (int k = threadIdx.x; k < ncell; k += BLOCK) grid[base + k] = acc[t * ncell + k];
        if (counts)
            for (int k = threadIdx.x; k < ncell; k += BLOCK) cell[base + k] = cnt[t * ncell + k];
    }
}

// ---------------------------------------------------------------------------
// LDS Stockham FFT (radix 2, autosort).  A block transforms `lines` lines of
// length N (N a power of two, N * lines == 1024) held in LDS.
// ---------------------------------------------------------------------------
constexpr int FFT_PTS = 1024;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ void load_twiddles(float2 *tw, const float2 *tab, int N) {
    for (int k = threadIdx.x; k < N; k += BLOCK) tw[k] = tab[k];
}

// In-LDS Stockham FFT of pts / N lines of N points: radix-4 stages
//   a_j = x[q + s (p + j m)], j < 4, m = N / (4 s)
//   y[q + s (4 p + k)] = w^(k p s) sum_j a_j (-+i)^(j k)
// then one radix-2 stage when log2 N is odd; returns the buffer holding the
// result.  Lines are ls >= N elements apart (padding against bank conflicts).
__device__ float2 *lds_fft4(float2 *x, float2 *y, const float2 *tw, int N, int logN, bool inv, int pts = FFT_PTS,
                            int ls = 0) {
    if (ls == 0) ls = N;
    const int qn = N >> 2;
    int s = 1, logs = 0;
    for (; 4 * s <= N; s <<= 2, logs += 2) {
        const int m = N / (4 * s);
        for (int t = threadIdx.x; t < (pts >> 2); t += BLOCK) {
            const int line = t / qn, b = t - line * qn;
            const int pp = b >> logs, q = b & (s - 1);
            const float2 *xl = x + line * ls;
            float2 *yl = y + line * ls;
            const float2 a0 = xl[q + s * pp], a1 = xl[q + s * (pp + m)];
            const float2 a2 = xl[q + s * (pp + 2 * m)], a3 = xl[q + s * (pp + 3 * m)];
            const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y), t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
            const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
            const float2 d = make_float2(a1.x - a3.x, a1.y - a3.y);
            const float2 t3 = inv ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);     // (+-i) d
            const int e = pp * s;
            float2 *o = yl + q + s * (4 * pp);
            o[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
            o[s] = cmul(make_float2(t1.x + t3.x, t1.y + t3.y), tw[e]);
            o[2 * s] = cmul(make_float2(t0.x - t2.x, t0.y - t2.y), tw[2 * e]);
            o[3 * s] = cmul(make_float2(t1.x - t3.x, t1.y - t3.y), tw[3 * e]);
        }
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
    }
    if (s < N) {                        // radix 2, m = N / (2 s)
        const int m = N / (2 * s), half = N >> 1;
        for (int t = threadIdx.x; t < (pts >> 1); t += BLOCK) {
            const int line = t / half, b = t - line * half;
            const int pp = b >> logs, q = b & (s - 1);
            const float2 *xl = x + line * ls;
            float2 *yl = y + line * ls;
            const float2 a = xl[q + s * pp], bb = xl[q + s * (pp + m)];
            yl[q + s * (2 * pp)] = make_float2(a.x + bb.x, a.y + bb.y);
            yl[q + s * (2 * pp + 1)] = cmul(make_float2(a.x - bb.x, a.y - bb.y), tw[pp * s]);
        }
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
    }
    (void)logN;
    return x;
}

// rows, grids of up to 2^23 cells: int64 grid -> float, two slices packed
// as re/im, forward FFT along x; grid -> cbuf.
__global__ __launch_bounds__(BLOCK) void k_fft_rows_fwd_grid(const int64_t *grid, float2 *cbuf, ffa_sc_params sp,
                                                       const ffa_sc_result *res, const float2 *twt, int lognx,
                                                       int lognz) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx, nz = sp.nz, ns = sp.ns;
    const int lines = FFT_PTS / nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;          // global row = plane*nz + j
    const double us = res->bad_grid != 0.0 ? 0.0 : res->qscale / (double)((int64_t)1 << (2 * sp.bits));
    load_twiddles(tw, twt, nx);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) {
        const int l = e >> lognx, i = e & (nx - 1);
        const int64_t row = row0 + l;
        const int64_t plane = row >> lognz, j = row & (nz - 1);
        const int64_t sa = 2 * plane, sb = sa + 1;
        const float re = (float)((double)grid[(sa * nz + j) * nx + i] * us);
        const float im = sb < ns ? (float)((double)grid[(sb * nz + j) * nx + i] * us) : 0.0f;
        b0[e] = make_float2(re, im);
    }
    __syncthreads();
    const float2 *o = lds_fft4(b0, b1, tw, nx, lognx, false);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) cbuf[row0 * nx + e] = o[e];
}

// rows, larger grids: forward FFT along x of the packed slice pairs, in
// place in cbuf.
__global__ __launch_bounds__(BLOCK) void k_fft_rows_fwd(float2 *cbuf, ffa_sc_params sp, const float2 *twt,
                                                       int lognx) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx;
    const int lines = FFT_PTS / nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;          // global row = plane*nz + j
    load_twiddles(tw, twt, nx);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) b0[e] = cbuf[row0 * nx + e];
    __syncthreads();
    const float2 *o = lds_fft4(b0, b1, tw, nx, lognx, false);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) cbuf[row0 * nx + e] = o[e];
}

// columns: forward FFT along z, multiply by the periodic Green's function
// (make_fft_kernel_open_numba, src/FFAG_SC_new.py:731-760), inverse FFT.
// Columns per workgroup: COL_PTS / nz (at most nx).  Wider tiles (4096
// points) measured slower at nz = 256: fewer workgroups, lower occupancy.
constexpr int COL_PTS = 1024;

__host__ __device__ inline int col_width(int nx, int nz) { return COL_PTS / nz < nx ? COL_PTS / nz : nx; }

__global__ __launch_bounds__(BLOCK) void k_fft_cols_green(float2 *cbuf, ffa_sc_params sp, const float2 *twt,
                                                         const float *green, int lognz) {
    // dynamic LDS: b0, b1 (COL_PTS + 32 each, padded columns), tw[nz], twi[nz]
    extern __shared__ float2 lds2[];
    float2 *b0 = lds2, *b1 = lds2 + (COL_PTS + 32), *tw = lds2 + 2 * (COL_PTS + 32), *twi = tw + sp.nz;
    const int nx = sp.nx, nz = sp.nz;
    const int W = col_width(nx, nz);                  // columns per block
    const int PTS = W * nz;
    const int LS = nz + 1;                            // padded column stride in LDS
    const int cpp = nx / W;                           // column groups per plane
    // neighbouring column groups share cache lines: keep them on one XCD
    const unsigned blk = xcd_block();
    const int64_t plane = blk / cpp;
    const int c0 = (blk - plane * cpp) * W;
    float2 *base = cbuf + plane * (int64_t)nz * nx;
    load_twiddles(tw, twt + 2 * TW_STRIDE, nz);
    load_twiddles(twi, twt + 3 * TW_STRIDE, nz);
    for (int e = threadIdx.x; e < PTS; e += BLOCK) {
        const int j = e / W, c = e - j * W;
        b0[c * LS + j] = base[(int64_t)j * nx + c0 + c];
    }
    __syncthreads();
    float2 *o = lds_fft4(b0, b1, tw, nz, lognz, false, PTS, LS);
    float2 *other = (o == b0) ? b1 : b0;
    // Green's function of the block's columns: one contiguous run of the
    // transposed table
    const float *gcol = green + (int64_t)c0 * nz;
    for (int e = threadIdx.x; e < PTS; e += BLOCK) {
        const int c = e / nz, j = e - c * nz;
        const float g = gcol[e];
        float2 &v = o[c * LS + j];
        v = make_float2(v.x * g, v.y * g);
    }
    __syncthreads();
    float2 *r = lds_fft4(o, other, twi, nz, lognz, true, PTS, LS);
    for (int e = threadIdx.x; e < PTS; e += BLOCK) {
        const int j = e / W, c = e - j * W;
        base[(int64_t)j * nx + c0 + c] = r[c * LS + j];
    }
}

// rows: inverse FFT along x, normalise by 1/(nx nz), unpack the slice pair.
__global__ __launch_bounds__(BLOCK) void k_fft_rows_inv(const float2 *cbuf, float *phi, ffa_sc_params sp,
                                                       const float2 *twt, int lognx, int lognz) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx, nz = sp.nz, ns = sp.ns;
    const int lines = FFT_PTS / nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;
    load_twiddles(tw, twt + TW_STRIDE, nx);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) b0[e] = cbuf[row0 * nx + e];
    __syncthreads();
    const float2 *o = lds_fft4(b0, b1, tw, nx, lognx, true);
    const float norm = 1.0f / (float)(nx * nz);
    for (int e = threadIdx.x; e < FFT_PTS; e += BLOCK) {
        const int l = e >> lognx, i = e & (nx - 1);
        const int64_t row = row0 + l;
        const int64_t plane = row >> lognz, j = row & (nz - 1);
        const int64_t sa = 2 * plane, sb = sa + 1;
        phi[(sa * nz + j) * nx + i] = o[e].x * norm;
        if (sb < ns) phi[(sb * nz + j) * nx + i] = o[e].y * norm;
    }
}

// compute_field_from_potential_numba_SBS (src/FFAG_SC_new.py:672-728):
// float32 differences, float64 scaling, float32 store.  A thread computes
// FIELD_ROWS consecutive rows of one column (its loads are independent, so
// their latencies overlap; a wave covers 64 consecutive columns) and stores
// (Ex, Ez) interleaved, so the gather reads both with one access per corner.
constexpr int FIELD_ROWS = 4;
constexpr size_t FROM_PHI_MAX_CELLS = (size_t)1 << 23;   // 32 MB of float32 potential

// (Ex, Ez) of cell (j, i) from phi at the cell (pc), above and below it in z
// (pu, pd: rows clamped to the grid) and left and right of it in x (lf, rt:
// columns clamped to the row).  Shared by k_sc_field and the gather that
// differentiates phi itself, so both give the same float32 values.
__device__ __forceinline__ float2 cell_field(float pc, float pu, float pd, float lf, float rt, int i, int j, int nx,
                                             int nz, double dr, double dz, double i2r, double i2z, double ig2) {
    double ex, ez;
    if (i == 0 || i == nx - 1) {              // one-sided in x; z neighbours clamped
        ex = i == 0 ? -(double)(rt - pc) / dr * ig2 : -(double)(pc - lf) / dr * ig2;
        ez = -(double)(pu - pd) / dz * ig2;
    } else if (j == 0) {
        ex = -(double)(rt - lf) * i2r * ig2;
        ez = -(double)(pu - pc) / dz * ig2;
    } else if (j == nz - 1) {
        ex = -(double)(rt - lf) * i2r * ig2;
        ez = -(double)(pc - pd) / dz * ig2;
    } else {
        ex = -(double)(rt - lf) * i2r * ig2;
        ez = -(double)(pu - pd) * i2z * ig2;
    }
    return make_float2((float)ex, (float)ez);
}

__global__ __launch_bounds__(BLOCK) void k_sc_field(const float *phi, float2 *exz, ffa_sc_params sp,
                                                   const ffa_sc_result *res, int lognx) {
    const int nx = sp.nx, nz = sp.nz;
    const int nb = nz / FIELD_ROWS;
    const int64_t total = (int64_t)sp.ns * nb * nx;
    // the stencil rows j0 - 1 and j0 + FIELD_ROWS belong to neighbouring
    // bands: keep neighbours on one XCD so each row is fetched into one L2
    const int64_t t = xcd_block() * (int64_t)BLOCK + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t & (nx - 1));
    const int64_t r = t >> lognx;
    const int band = (int)(r % nb);
    const int64_t k = r / nb;
    const int j0 = band * FIELD_ROWS;
    const int64_t base = k * (int64_t)nz * nx;
    const double dr = res->dx, dz = res->dz, bl = res->bunch_length;
    if (res->bad_grid != 0.0 || dr == 0.0 || dz == 0.0 || bl == 0.0) {
        for (int c = 0; c < FIELD_ROWS; ++c) exz[base + (int64_t)(j0 + c) * nx + i] = make_float2(0.0f, 0.0f);
        return;
    }
    const double gamma = res->mean_gamma_m1 + 1.0;
    const double ig2 = 1.0 / (gamma * gamma);
    const double i2r = 0.5 / dr, i2z = 0.5 / dz;
    const float *P = phi + base;
    // col[c] = phi(clamp(j0 - 1 + c), i); lf/rt = phi(j, i -+ 1) clamped to the row
    float col[FIELD_ROWS + 2], lf[FIELD_ROWS], rt[FIELD_ROWS];
    const int il = i > 0 ? i - 1 : 0, ir = i < nx - 1 ? i + 1 : nx - 1;
    for (int c = 0; c < FIELD_ROWS + 2; ++c) {
        int j = j0 - 1 + c;
        j = j < 0 ? 0 : (j > nz - 1 ? nz - 1 : j);
        col[c] = P[(int64_t)j * nx + i];
    }
    for (int c = 0; c < FIELD_ROWS; ++c) {
        lf[c] = P[(int64_t)(j0 + c) * nx + il];
        rt[c] = P[(int64_t)(j0 + c) * nx + ir];
    }
    for (int c = 0; c < FIELD_ROWS; ++c) {
        const int j = j0 + c;
        exz[base + (int64_t)j * nx + i] =
            cell_field(col[c + 1], col[c + 2], col[c], lf[c], rt[c], i, j, nx, nz, dr, dz, i2r, i2z, ig2);
    }
}

// (Ex, Ez) at the gather's four cells (i0 .. i0 + 1, j0 .. j0 + 1, with
// 0 <= i0 <= nx - 2 and 0 <= j0 <= nz - 2) straight from phi: the 4 x 4
// stencil block minus its corners, rows and columns clamped to the grid as
// k_sc_field clamps them.
__device__ __forceinline__ void cells_field_from_phi(const float *P, int i0, int j0, int nx, int nz, double dr,
                                                     double dz, double i2r, double i2z, double ig2, float2 &a00,
                                                     float2 &a10, float2 &a01, float2 &a11) {
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
    a00 = cell_field(v11, v21, v01, v10, v12, i0, j0, nx, nz, dr, dz, i2r, i2z, ig2);
    a10 = cell_field(v12, v22, v02, v11, v13, i0 + 1, j0, nx, nz, dr, dz, i2r, i2z, ig2);
    a01 = cell_field(v21, v31, v11, v20, v22, i0, j0 + 1, nx, nz, dr, dz, i2r, i2z, ig2);
    a11 = cell_field(v22, v32, v12, v21, v23, i0 + 1, j0 + 1, nx, nz, dr, dz, i2r, i2z, ig2);
}

// interpolate_fields_to_particles_SBS_numba (src/FFAG_SC_new.py:870-972)
// fused with the endpoint roll of src/FFAG_track.py:356-382.
template <typename ET, bool FROM_PHI>
__global__ __launch_bounds__(BLOCK) void k_sc_gather_roll(ffa_particles p, ffa_sc_params sp, const ffa_sc_result *res,
                                                         const float2 *exz, const float *phi, const double *px,
                                                         const double *ppr0, const int32_t *pslice) {
    const int64_t i = xcd_block() * (int64_t)BLOCK + threadIdx.x;
    if (i >= p.n) return;
    const int64_t n = p.n;
    ET *ep = reinterpret_cast<ET *>(p.esc_pre);
    ET *epp = reinterpret_cast<ET *>(p.esc_prepre);
    const uint8_t fl = p.flags[i];
    const bool alive = fl & FFA_F_ALIVE;
    const bool good = res->bad_grid == 0.0;
    // the old pre-step field is read only where it survives the solve
    double o_r = 0.0, o_z = 0.0, o_f = 0.0;
    if (!(alive && good)) { o_r = (double)ep[i]; o_z = (double)ep[n + i]; o_f = (double)ep[2 * n + i]; }
    double nr = o_r, nz_ = o_z, nf = o_f;
    if (alive) {
        if (good) {
            const int nx = sp.nx, nz = sp.nz;
            int s = pslice[i];
            if (s < 0) s = 0; else if (s >= sp.ns) s = sp.ns - 1;
            const double t = (px[i] - res->xmin) * (1.0 / res->dx), u = (p.z[i] - res->zmin) * (1.0 / res->dz);
            const double pr = ppr0[i];
            const double ct = t > -1e9 ? (t < 1e9 ? t : 1e9) : -1e9, cu = u > -1e9 ? (u < 1e9 ? u : 1e9) : -1e9;
            int64_t i0 = (int64_t)floor(ct), j0 = (int64_t)floor(cu);
            if (i0 < 0) i0 = 0; else if (i0 > nx - 2) i0 = nx - 2;
            if (j0 < 0) j0 = 0; else if (j0 > nz - 2) j0 = nz - 2;
            const double al = t - (double)i0, be = u - (double)j0;
            const double w00 = (1.0 - al) * (1.0 - be), w10 = al * (1.0 - be), w01 = (1.0 - al) * be, w11 = al * be;
            float2 a00, a10, a01, a11;
            if constexpr (FROM_PHI) {
                const double dr = res->dx, dz = res->dz, gm = res->mean_gamma_m1 + 1.0;
                cells_field_from_phi(phi + (size_t)s * nz * nx, (int)i0, (int)j0, nx, nz, dr, dz, 0.5 / dr,
                                     0.5 / dz, 1.0 / (gm * gm), a00, a10, a01, a11);
            } else {
                const float2 *E = exz + (size_t)s * nz * nx;
                const int64_t b = j0 * nx + i0;
                a00 = E[b]; a10 = E[b + 1]; a01 = E[b + nx]; a11 = E[b + nx + 1];
            }
            const double er = w00 * a00.x + w10 * a10.x + w01 * a01.x + w11 * a11.x;
            const double ez = w00 * a00.y + w10 * a10.y + w01 * a01.y + w11 * a11.y;
            const double ip = 1.0 / sqrt(1.0 + pr * pr);
            nr = er * ip;
            nz_ = ez;
            nf = er * (-pr * ip);
        } else if ((fl & FFA_F_SCPRE) && (fl & FFA_F_SCPREPRE)) {
            // degenerate grid: the reference leaves LocalBunch columns 10-12 as
            // the previous step wrote them (the extrapolated field)
            nr = o_r + sp.s_last * (o_r - (double)epp[i]);
            nz_ = o_z + sp.s_last * (o_z - (double)epp[n + i]);
            nf = o_f + sp.s_last * (o_f - (double)epp[2 * n + i]);
        }
    }
    // The two endpoint buffers trade roles at every solve: the new pre-step
    // field goes into the old prepre memory (the caller then swaps its
    // pointers) and the old pre memory, untouched, is the new prepre field.
    // A live particle writes only its new field; a dead one keeps its old.
    const ET a = (ET)nr, b = (ET)nz_, c = (ET)nf;
    epp[i] = a; epp[n + i] = b; epp[2 * n + i] = c;
    uint8_t nfl;
    if (sp.first) {
        ep[i] = a; ep[n + i] = b; ep[2 * n + i] = c;
        nfl = (uint8_t)((fl & ~(FFA_F_SCPRE | FFA_F_SCPREPRE)) | (alive ? FFA_F_SCPRE : 0));
    } else {
        nfl = (uint8_t)((fl & ~(FFA_F_SCPRE | FFA_F_SCPREPRE)) | ((fl & FFA_F_SCPRE) ? FFA_F_SCPREPRE : 0) |
                        (alive ? FFA_F_SCPRE : 0));
    }
    if (nfl != fl) p.flags[i] = nfl;
}

inline unsigned gridn(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

inline int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

}  // namespace

int ffa_sc_layout_compute(int64_t n, int32_t ns, int32_t nz, int32_t nx, ffa_sc_layout *L) {
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t C = (size_t)ns * nz * nx;
    const size_t P = (size_t)((ns + 1) / 2) * nz * nx;
    size_t o = 0;
    L->grid = o; o = al(o + 8 * C);
    L->cell_count = o; o = al(o + 4 * C);
    L->cbuf = o; o = al(o + 8 * P);
    L->phi = o; o = al(o + 4 * C);
    L->exz = o; o = al(o + 8 * C);
    L->px = o; o = al(o + 8 * (size_t)n);
    L->ppr0 = o; o = al(o + 8 * (size_t)n);
    L->pslice = o; o = al(o + 4 * (size_t)n);
    L->slice_count = o; o = al(o + 4 * (size_t)ns);
    L->partials = o; o = al(o + sizeof(double) * 16 * NB);
    L->result = o; o = al(o + sizeof(ffa_sc_result));
    L->records = o; o = al(o + 8 * (size_t)n);
    L->slice_start = o; o = al(o + 4 * ((size_t)ns + 1));
    L->block_offset = o; o = al(o + 4 * (size_t)NB * ns);
    L->records_by_row = o; o = al(o + 8 * (size_t)n);
    L->row_start = o; o = al(o + 4 * (size_t)ns * ((size_t)nz + 1));
    L->tables = o; o = al(o + 8 * 4 * (size_t)TW_STRIDE + 4 * (size_t)nz * nx);
    L->total = o;
    return 0;
}

hipError_t launch_sc_solve(ffa_ctx_s *c, const ffa_particles *p, const ffa_sc_params *sp, void *ws,
                           const ffa_sc_layout *L) {
    char *w = reinterpret_cast<char *>(ws);
    int64_t *grid = reinterpret_cast<int64_t *>(w + L->grid);
    int32_t *cell = reinterpret_cast<int32_t *>(w + L->cell_count);
    float2 *cbuf = reinterpret_cast<float2 *>(w + L->cbuf);
    float *phi = reinterpret_cast<float *>(w + L->phi);
    float2 *exz = reinterpret_cast<float2 *>(w + L->exz);
    double *px = reinterpret_cast<double *>(w + L->px);
    double *ppr0 = reinterpret_cast<double *>(w + L->ppr0);
    int32_t *pslice = reinterpret_cast<int32_t *>(w + L->pslice);
    int32_t *slc = reinterpret_cast<int32_t *>(w + L->slice_count);
    double *part = reinterpret_cast<double *>(w + L->partials);
    ffa_sc_result *res = reinterpret_cast<ffa_sc_result *>(w + L->result);
    const double E0J = 938.2723e6 * 1.60217662e-19;   // src/FFAG_SC_new.py:115
    const size_t C = (size_t)sp->ns * sp->nz * sp->nx;
    hipStream_t st = c->stream;
    uint64_t *rec = reinterpret_cast<uint64_t *>(w + L->records);
    int32_t *sstart = reinterpret_cast<int32_t *>(w + L->slice_start);
    int32_t *blk_off = reinterpret_cast<int32_t *>(w + L->block_offset);
    hipLaunchKernelGGL(k_sc_stage1, dim3(NB), dim3(BLOCK), 0, st, *p, sp->rk, sp->c, E0J, part);
    hipLaunchKernelGGL(k_sc_final1, dim3(1), dim3(BLOCK), 0, st, part, *sp, res, slc);
    hipLaunchKernelGGL(k_sc_locate_moments, dim3(NB), dim3(BLOCK), 0, st, *p, *sp, sp->c, E0J, res, px, ppr0, pslice,
                       slc, blk_off, part);
    hipLaunchKernelGGL(k_sc_final2, dim3(1), dim3(BLOCK), 0, st, part, *sp, res, slc, sstart);
    hipLaunchKernelGGL(k_sc_bin, dim3(NB), dim3(BLOCK), 0, st, *p, *sp, res, px, pslice, sstart, blk_off, rec);
    // Small grids deposit slice by slice into the int64 grid, which the row
    // FFT scales on read; above 2^23 cells the deposit writes slice pairs
    // straight into the packed FFT buffer (half the grid traffic).  Both give
    // the same float32 FFT input.  Measured: the pair path is 3.9 % faster
    // overall at (512, 256, 256) and 1.5 % slower at (256, 128, 128).
    const bool pairs = C > FROM_PHI_MAX_CELLS;
    const int tile_cells = pairs ? 2048 : 4096;                            // <= 48 KB of LDS per tile
    const int bz = (tile_cells / sp->nx) < sp->nz ? (tile_cells / sp->nx) : sp->nz;
    const size_t tile_lds = (size_t)(pairs ? 2 : 1) * bz * sp->nx * (sp->cell_counts ? 12 : 8);
    uint64_t *rec2 = reinterpret_cast<uint64_t *>(w + L->records_by_row);
    int32_t *rstart = reinterpret_cast<int32_t *>(w + L->row_start);
    float2 *twt = reinterpret_cast<float2 *>(w + L->tables);
    float *green = reinterpret_cast<float *>(w + L->tables + 8 * 4 * (size_t)TW_STRIDE);
    hipLaunchKernelGGL(k_sc_rowsort, dim3((unsigned)sp->ns), dim3(ROW_BLOCK), 0, st, *sp, res, sstart, rec, rec2,
                       rstart, twt, green);
    if (pairs)
        hipLaunchKernelGGL(k_sc_deposit_pairs, dim3((unsigned)(((sp->ns + 1) / 2) * (sp->nz / bz))), dim3(BLOCK),
                           tile_lds, st, *sp, res, rstart, rec2, reinterpret_cast<unsigned long long *>(grid), cell, cbuf,
                           bz);
    else
        hipLaunchKernelGGL(k_sc_deposit_slices, dim3((unsigned)(sp->ns * (sp->nz / bz))), dim3(BLOCK), tile_lds, st,
                           *sp, res, rstart, rec2, reinterpret_cast<unsigned long long *>(grid), cell, bz);
    const int lognx = ilog2(sp->nx), lognz = ilog2(sp->nz);
    const int64_t planes = (sp->ns + 1) / 2;
    const int64_t rows = planes * sp->nz;
    const size_t rows_lds = sizeof(float2) * (2 * FFT_PTS + sp->nx);
    const size_t cols_lds = sizeof(float2) * (2 * (COL_PTS + 32) + 2 * sp->nz);
    if (pairs)
        hipLaunchKernelGGL(k_fft_rows_fwd, dim3((unsigned)(rows * sp->nx / FFT_PTS)), dim3(BLOCK), rows_lds, st, cbuf,
                           *sp, twt, lognx);
    else
        hipLaunchKernelGGL(k_fft_rows_fwd_grid, dim3((unsigned)(rows * sp->nx / FFT_PTS)), dim3(BLOCK), rows_lds, st,
                           grid, cbuf, *sp, res, twt, lognx, lognz);
    hipLaunchKernelGGL(k_fft_cols_green, dim3((unsigned)(planes * sp->nx / col_width(sp->nx, sp->nz))), dim3(BLOCK),
                       cols_lds, st, cbuf, *sp, twt, green, lognz);
    hipLaunchKernelGGL(k_fft_rows_inv, dim3((unsigned)(rows * sp->nx / FFT_PTS)), dim3(BLOCK), rows_lds, st, cbuf, phi,
                       *sp, twt, lognx, lognz);
    // The gather from phi reads 12 potential values per particle instead of 4
    // field pairs: cheaper while the potential stays cache-resident (measured
    // faster at 4.2M cells, slower at 33.5M), so larger grids store the cubes
    const bool cubes = sp->field_cubes != 0 || C > FROM_PHI_MAX_CELLS;
    if (cubes)
        hipLaunchKernelGGL(k_sc_field, dim3(gridn((int64_t)(C / FIELD_ROWS))), dim3(BLOCK), 0, st, phi, exz, *sp, res,
                           lognx);
    if (p->n > 0) {
        const dim3 g(gridn(p->n));
        if (p->esc_f32) {
            if (cubes)
                hipLaunchKernelGGL((k_sc_gather_roll<float, false>), g, dim3(BLOCK), 0, st, *p, *sp, res, exz, phi, px,
                                   ppr0, pslice);
            else
                hipLaunchKernelGGL((k_sc_gather_roll<float, true>), g, dim3(BLOCK), 0, st, *p, *sp, res, exz, phi, px,
                                   ppr0, pslice);
        } else {
            if (cubes)
                hipLaunchKernelGGL((k_sc_gather_roll<double, false>), g, dim3(BLOCK), 0, st, *p, *sp, res, exz, phi, px,
                                   ppr0, pslice);
            else
                hipLaunchKernelGGL((k_sc_gather_roll<double, true>), g, dim3(BLOCK), 0, st, *p, *sp, res, exz, phi, px,
                                   ppr0, pslice);
        }
    }
    return hipGetLastError();
}
