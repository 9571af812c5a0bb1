// 2.5D space-charge solve of the tracking hot path (gfx950).
//
// Pipeline of one solve (every kernel on the context stream, no host sync):
//   k_sc_front         per survivor: gamma, phi mod 2pi, SEO r0/pr0 and
//                      x = r - r0; per-block partial sums of gamma, the four
//                      covering-arc min/max and the moments
//   k_sc_final1        1 block: covering arc, slice step, mean gamma,
//                      perimeter, bunch length (device-side scalars); zeroes
//                      the per-slice counters
//   k_sc_slices        per survivor: slice index; slice histogram
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
//   k_fft_rows_inv     inverse row FFTs, 1/(nx nz), unpack to phi (grids
//                      whose gather differentiates phi itself)
//   k_fft_rows_inv_field  inverse row FFTs fused with -grad(phi)/gamma^2 and
//                      the reference's edge stencils, stored as (Ex, Ez)
//                      cubes (above 2^23 cells, or when sp.field_cubes)
//   k_sc_gather_roll   bilinear gather + pr0 projection, endpoint roll; without
//                      the cubes it applies the same stencils to phi at the
//                      particle's four cells
// Particle-side arithmetic is FP64 in the reference's operation order (this
// unit is compiled with -ffp-contract=off); the grid is float32 like the
// reference's (complex64 FFTs).  Reductions use a fixed block partition and
// a fixed combine order, so a solve is bit-reproducible run to run.
#include "ffa_internal.h"
#include "ffa_sc_gather.h"

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

// Particles per pass of k_sc_front: two keep it under 128 VGPRs, so all NB
// blocks are resident at once (the visiting order i0, i0 + S, ... and hence
// the sums are the same for any value).
constexpr int LUNR = 2;

// The per-particle front end of a solve in one pass over the state:
// GetLocalCoordinates_2p5(_rk) and divide_arc2's scan
// (src/FFAG_SC_new.py:74-148, 534-566: sum of gamma, survivor count, the
// covering-arc minima / maxima of the azimuth under four shifts) and the SEO
// lookup of ComputeSliceBoundsAndCoordsSBS_merged (x = r - r0(Ek, phi) and
// pr0 written for the later kernels, the moments sum x, z, x^2, z^2).  None
// of it needs the arc, which only the slice index does (k_sc_slices, after
// k_sc_final1).  Fixed block partition and per-thread order: the sums are
// reproducible run to run.  part[16 b + 0 .. 13] per block.
__global__ __launch_bounds__(BLOCK) void k_sc_front(ffa_particles p, ffa_sc_params sp, double c, double E0J,
                                                   double *px, double *ppr0, double *part) {
    __shared__ double sh[14][BLOCK / 64];
    const double shifts[4] = {0.0, FFA_PI, FFA_PI * 0.5, FFA_PI * 1.5};
    double v[14] = {0.0, 0.0, 1e30, 1e30, 1e30, 1e30, -1e30, -1e30, -1e30, -1e30, 0.0, 0.0, 0.0, 0.0};
    const int64_t S = (int64_t)NB * BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i0 < p.n; i0 += LUNR * S) {
        // flags, then every state load, then the arithmetic: no branch between
        // the particles' loads, so their round trips overlap
        bool ok[LUNR];
        int64_t jj[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            const int64_t i = i0 + u * S;
            const uint8_t f = p.flags[i < p.n ? i : 0];
            ok[u] = (i < p.n) & ((f & FFA_F_ALIVE) != 0);
            jj[u] = ok[u] ? i : 0;
        }
        double fm[LUNR], x[LUNR], z[LUNR], fd[LUNR], vr[LUNR], vz[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            x[u] = p.r[jj[u]];
            fd[u] = p.fidot[jj[u]];
            vr[u] = p.vr[jj[u]];
            vz[u] = p.vz[jj[u]];
            fm[u] = p.fi[jj[u]];
            z[u] = p.z[jj[u]];
        }
        double ek[LUNR];
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
            if (ok[u]) {
                v[0] += g;
                v[1] += 1.0;
                for (int k = 0; k < 4; ++k) {
                    const double a = fm[u] + shifts[k], b = 2.0 * FFA_PI;
                    const double q = a - b * floor(a / b);
                    v[2 + k] = q < v[2 + k] ? q : v[2 + k];
                    v[6 + k] = q > v[6 + k] ? q : v[6 + k];
                }
            }
        }
        double pr0[LUNR];
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            x[u] -= seo_bilinear(sp.seo_r, sp.n_ek, sp.n_fi, sp.ek0, sp.dek, sp.fi0, sp.dfi, ek[u], fm[u]);
            pr0[u] = seo_bilinear(sp.seo_pr, sp.n_ek, sp.n_fi, sp.ek0, sp.dek, sp.fi0, sp.dfi, ek[u], fm[u]);
        }
#pragma unroll
        for (int u = 0; u < LUNR; ++u) {
            if (!ok[u]) continue;
            const int64_t i = i0 + u * S;
            px[i] = x[u];
            ppr0[i] = pr0[u];
            v[10] += x[u];
            v[11] += z[u];
            v[12] += x[u] * x[u];
            v[13] += z[u] * z[u];
        }
    }
    const int op[14] = {0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 0, 0, 0, 0};
    block_reduce<14>(sh, v, op);
    if (threadIdx.x == 0)
        for (int k = 0; k < 14; ++k) part[16 * blockIdx.x + k] = v[k];
}

// Slice index of every survivor from the covering arc (k_sc_final1), and the
// slice histogram: per-block counts in LDS, one global atomic per non-empty
// (block, slice) pair (ns <= FFA_SC_MAX_NS).  Same block partition as
// k_sc_bin, which places each block's records from these offsets.
__global__ __launch_bounds__(BLOCK) void k_sc_slices(ffa_particles p, ffa_sc_params sp, const ffa_sc_result *res,
                                                    int32_t *pslice, int32_t *slc, int32_t *blk_off) {
    __shared__ int32_t hist[MAX_NS_LDS];
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK) hist[k] = 0;
    __syncthreads();
    const double fs = res->fi_s, step = res->slice_step;
    const int64_t S = (int64_t)NB * BLOCK;
    for (int64_t i0 = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i0 < p.n; i0 += UNR * S) {
        bool ok[UNR];
        double fi[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            const int64_t i = i0 + u * S;
            const uint8_t f = p.flags[i < p.n ? i : 0];
            ok[u] = (i < p.n) & ((f & FFA_F_ALIVE) != 0);
            fi[u] = p.fi[ok[u] ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            double f = mod2pi(fi[u]);
            if (f < fs) f += FFA_TWO_PI;
            double fsl = (f - fs) / step;
            fsl = fsl > -1e9 ? (fsl < 1e9 ? fsl : 1e9) : -1e9;
            int64_t sl = (int64_t)fsl;
            if (sl < 0) sl = 0; else if (sl >= sp.ns) sl = sp.ns - 1;
            const int s = ok[u] ? (int)sl : -1;
            const int64_t i = i0 + u * S;
            if (i < p.n) pslice[i] = s;
            wave_counter_add(hist, s >= 0 ? s : 0, s >= 0);
        }
    }
    __syncthreads();
    // the slice counter's value before this block's add is where the block's
    // run of records starts inside the slice (entries of slices the block
    // does not hold stay unwritten and are never read)
    int32_t *bo = blk_off + (size_t)blockIdx.x * sp.ns;
    for (int k = threadIdx.x; k < sp.ns; k += BLOCK)
        if (hist[k]) bo[k] = atomicAdd(&slc[k], hist[k]);
}

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
// per-thread order as k_sc_slices, which counted each block's
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
    const int nbands = (nz + bz - 1) / bz;                  // the last band may be shorter
    const int s = blockIdx.x / nbands, band = blockIdx.x - s * nbands;
    const int z0 = band * bz, z1 = min(z0 + bz, nz);
    const int ncell = (z1 - z0) * nx;
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
    const int nbands = (nz + bz - 1) / bz;                  // the last band may be shorter
    const int pl = blockIdx.x / nbands, band = blockIdx.x - pl * nbands;
    const int z0 = band * bz, z1 = min(z0 + bz, nz);
    const int ncell = (z1 - z0) * nx;
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
// LDS Stockham FFT (mixed radix 4, 2, 3, 5; autosort).  A block transforms
// the lines of one tile, each of length N = 2^a 3^b 5^c <= 1024, held in LDS.
// Power-of-two sizes take radix-4 stages and at most one radix-2 stage with
// shift/mask index arithmetic (P2 = true); other sizes add radix-3 and
// radix-5 stages and divide.
// ---------------------------------------------------------------------------
constexpr int FFT_PTS = 1024;
constexpr int FFT_MAX_STAGES = 12;

struct fft_plan {
    int n, nst;
    int radix[FFT_MAX_STAGES];
};

// Radix sequence of n (4s first, then one 2, then 3s, then 5s); nst = 0 when
// n has a prime factor above 5.
inline fft_plan make_fft_plan(int n) {
    fft_plan p{};
    p.n = n;
    int m = n, k = 0;
    while (m % 4 == 0 && k < FFT_MAX_STAGES) { p.radix[k++] = 4; m /= 4; }
    if (m % 2 == 0 && k < FFT_MAX_STAGES) { p.radix[k++] = 2; m /= 2; }
    while (m % 3 == 0 && k < FFT_MAX_STAGES) { p.radix[k++] = 3; m /= 3; }
    while (m % 5 == 0 && k < FFT_MAX_STAGES) { p.radix[k++] = 5; m /= 5; }
    p.nst = m == 1 ? k : 0;
    return p;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// -i v (forward) or +i v (inverse)
__device__ __forceinline__ float2 rot90(float2 v, bool inv) { return inv ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x); }

__device__ __forceinline__ void load_twiddles(float2 *tw, const float2 *tab, int N) {
    for (int k = threadIdx.x; k < N; k += BLOCK) tw[k] = tab[k];
}

// q = a / d, r = a - q d (d a power of two with log2 ld when P2)
template <bool P2>
__device__ __forceinline__ int idiv(int a, int d, int ld) { return P2 ? (a >> ld) : a / d; }

// In-LDS Stockham FFT of pts / N lines of N points.  Stage with radix R and
// stride s (the product of the earlier radices), m = N / (R s), p < m, q < s:
//   a_j = x[q + s (p + j m)], j < R
//   y[q + s (R p + k)] = w^(k p s) sum_j a_j omega_R^(j k),  w = e^(-+2 pi i / N)
// Every index stays below N.  Lines are ls >= N elements apart (padding
// against bank conflicts).  Returns the buffer holding the result.
// Power-of-two sizes: radix-4 stages and at most one radix-2 stage, every
// index by shifts and masks (integer multiplies and divides are the
// expensive VALU instructions here).
__device__ float2 *lds_fft_p2(float2 *x, float2 *y, const float2 *tw, int N, bool inv, int pts, int ls) {
    const int logN = __builtin_ctz((unsigned)N);
    const int logq = logN - 2;                           // log2 of the butterflies per line
    int logs = 0;
    for (; logs + 2 <= logN; logs += 2) {
        const int s = 1 << logs, logm = logN - 2 - logs;
        const int sm = 1 << (logN - 2);                  // s m
        for (int t = threadIdx.x; t < (pts >> 2); t += BLOCK) {
            const int line = t >> logq, b = t & ((1 << logq) - 1);
            const int pp = b >> logs, q = b & (s - 1);
            const float2 *xl = x + line * ls + q + (pp << logs);
            float2 *yl = y + line * ls + q + (pp << (logs + 2));
            const float2 a0 = xl[0], a1 = xl[sm], a2 = xl[2 * sm], a3 = xl[3 * sm];
            const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
            const float2 t3 = rot90(csub(a1, a3), inv);
            const int e = pp << logs;
            yl[0] = cadd(t0, t2);
            yl[s] = cmul(cadd(t1, t3), tw[e]);
            yl[2 * s] = cmul(csub(t0, t2), tw[2 * e]);
            yl[3 * s] = cmul(csub(t1, t3), tw[3 * e]);
        }
        (void)logm;
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
    }
    if (logs < logN) {                                   // radix 2, s = N / 2
        const int s = 1 << logs, logh = logN - 1;
        for (int t = threadIdx.x; t < (pts >> 1); t += BLOCK) {
            const int line = t >> logh, q = t & (s - 1);  // one butterfly group: pp = 0
            const float2 *xl = x + line * ls + q;
            float2 *yl = y + line * ls + q;
            const float2 a0 = xl[0], a1 = xl[s];
            yl[0] = cadd(a0, a1);
            yl[s] = cmul(csub(a0, a1), tw[0]);
        }
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
    }
    return x;
}

template <bool P2>
__device__ float2 *lds_fft(float2 *x, float2 *y, const float2 *tw, const fft_plan &pl, bool inv, int pts = FFT_PTS,
                           int ls = 0) {
    const int N = pl.n;
    if (ls == 0) ls = N;
    if constexpr (P2) return lds_fft_p2(x, y, tw, N, inv, pts, ls);
    const int lines = pts / N;
    int s = 1, logs = 0;
    for (int st = 0; st < pl.nst; ++st) {
        const int R = pl.radix[st];
        const int m = N / (R * s);
        const int per = N / R;                           // butterflies per line
        const int logper = P2 ? __builtin_ctz((unsigned)per) : 0;
        for (int t = threadIdx.x; t < lines * per; t += BLOCK) {
            const int line = idiv<P2>(t, per, logper), b = t - line * per;
            const int pp = idiv<P2>(b, s, logs), q = b - pp * s;
            const float2 *xl = x + line * ls + q + s * pp;
            float2 *yl = y + line * ls + q + s * (R * pp);
            const int e = pp * s;
            const int sm = s * m;
            if (R == 4) {
                const float2 a0 = xl[0], a1 = xl[sm], a2 = xl[2 * sm], a3 = xl[3 * sm];
                const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
                const float2 t3 = rot90(csub(a1, a3), inv);
                yl[0] = cadd(t0, t2);
                yl[s] = cmul(cadd(t1, t3), tw[e]);
                yl[2 * s] = cmul(csub(t0, t2), tw[2 * e]);
                yl[3 * s] = cmul(csub(t1, t3), tw[3 * e]);
            } else if (R == 2) {
                const float2 a0 = xl[0], a1 = xl[sm];
                yl[0] = cadd(a0, a1);
                yl[s] = cmul(csub(a0, a1), tw[e]);
            } else if (R == 3) {
                // omega = e^(-+2 pi i / 3) = -1/2 -+ i sqrt(3)/2
                const float2 a0 = xl[0], a1 = xl[sm], a2 = xl[2 * sm];
                const float2 sum = cadd(a1, a2);
                const float2 d = rot90(cscale(csub(a1, a2), 0.8660254037844386f), inv);
                const float2 c = csub(a0, cscale(sum, 0.5f));
                yl[0] = cadd(a0, sum);
                yl[s] = cmul(cadd(c, d), tw[e]);
                yl[2 * s] = cmul(csub(c, d), tw[2 * e]);
            } else {
                // radix 5 with c1 = cos(2 pi/5), c2 = cos(4 pi/5), s1 = sin(2 pi/5), s2 = sin(4 pi/5)
                const float c1 = 0.30901699437494745f, c2 = -0.8090169943749475f;
                const float s1 = 0.9510565162951535f, s2 = 0.5877852522924731f;
                const float2 a0 = xl[0], a1 = xl[sm], a2 = xl[2 * sm], a3 = xl[3 * sm], a4 = xl[4 * sm];
                const float2 t1 = cadd(a1, a4), t2 = cadd(a2, a3), t3 = csub(a1, a4), t4 = csub(a2, a3);
                const float2 m1 = cadd(a0, cadd(cscale(t1, c1), cscale(t2, c2)));
                const float2 m2 = cadd(a0, cadd(cscale(t1, c2), cscale(t2, c1)));
                const float2 n1 = rot90(cadd(cscale(t3, s1), cscale(t4, s2)), inv);
                const float2 n2 = rot90(csub(cscale(t3, s2), cscale(t4, s1)), inv);
                yl[0] = cadd(a0, cadd(t1, t2));
                yl[s] = cmul(cadd(m1, n1), tw[e]);
                yl[2 * s] = cmul(cadd(m2, n2), tw[2 * e]);
                yl[3 * s] = cmul(csub(m2, n2), tw[3 * e]);
                yl[4 * s] = cmul(csub(m1, n1), tw[4 * e]);
            }
        }
        __syncthreads();
        float2 *tmp = x; x = y; y = tmp;
        s *= R;
        if (P2) logs += R == 4 ? 2 : 1;
    }
    return x;
}

// Rows of the packed slice pairs: a tile holds `lines` = FFT_PTS / nx whole
// rows (global row = plane * nz + j); rows past the last of the last tile
// are zero-filled and not stored.
__host__ __device__ inline int row_lines(int nx) { return FFT_PTS / nx; }

// rows, grids of up to 2^23 cells: int64 grid -> float, two slices packed
// as re/im, forward FFT along x; grid -> cbuf.
template <bool P2>
__global__ __launch_bounds__(BLOCK) void k_fft_rows_fwd_grid(const int64_t *grid, float2 *cbuf, ffa_sc_params sp,
                                                             const ffa_sc_result *res, const float2 *twt, fft_plan pl,
                                                             int lognx, int lognz, int64_t rows) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx, nz = sp.nz, ns = sp.ns;
    const int lines = row_lines(nx), pts = lines * nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;          // global row = plane*nz + j
    const double us = res->bad_grid != 0.0 ? 0.0 : res->qscale / (double)((int64_t)1 << (2 * sp.bits));
    load_twiddles(tw, twt, nx);
    for (int e = threadIdx.x; e < pts; e += BLOCK) {
        const int l = idiv<P2>(e, nx, lognx), i = e - l * nx;
        const int64_t row = row0 + l;
        float2 v = make_float2(0.0f, 0.0f);
        if (row < rows) {
            const int64_t plane = P2 ? (row >> lognz) : row / nz, j = row - plane * nz;
            const int64_t sa = 2 * plane, sb = sa + 1;
            v.x = (float)((double)grid[(sa * nz + j) * nx + i] * us);
            v.y = sb < ns ? (float)((double)grid[(sb * nz + j) * nx + i] * us) : 0.0f;
        }
        b0[e] = v;
    }
    __syncthreads();
    const float2 *o = lds_fft<P2>(b0, b1, tw, pl, false, pts);
    for (int e = threadIdx.x; e < pts; e += BLOCK)
        if (row0 * nx + e < rows * nx) cbuf[row0 * nx + e] = o[e];
}

// rows, larger grids: forward FFT along x of the packed slice pairs, in
// place in cbuf.
template <bool P2>
__global__ __launch_bounds__(BLOCK) void k_fft_rows_fwd(float2 *cbuf, ffa_sc_params sp, const float2 *twt, fft_plan pl,
                                                        int64_t rows) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx;
    const int lines = row_lines(nx), pts = lines * nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;
    const int64_t lim = rows * nx - row0 * nx;
    load_twiddles(tw, twt, nx);
    for (int e = threadIdx.x; e < pts; e += BLOCK) b0[e] = e < lim ? cbuf[row0 * nx + e] : make_float2(0.0f, 0.0f);
    __syncthreads();
    const float2 *o = lds_fft<P2>(b0, b1, tw, pl, false, pts);
    for (int e = threadIdx.x; e < pts; e += BLOCK)
        if (e < lim) cbuf[row0 * nx + e] = o[e];
}

// columns: forward FFT along z, multiply by the periodic Green's function
// (make_fft_kernel_open_numba, src/FFAG_SC_new.py:731-760), inverse FFT.
// Columns per workgroup: COL_PTS / nz (at most nx); the last group of a
// plane may hold fewer.  Wider tiles (4096 points) measured slower at
// nz = 256: fewer workgroups, lower occupancy.
constexpr int COL_PTS = 1024;

__host__ __device__ inline int col_width(int nx, int nz) { return COL_PTS / nz < nx ? COL_PTS / nz : nx; }
__host__ __device__ inline int col_groups(int nx, int nz) { return (nx + col_width(nx, nz) - 1) / col_width(nx, nz); }

template <bool P2>
__global__ __launch_bounds__(BLOCK) void k_fft_cols_green(float2 *cbuf, ffa_sc_params sp, const float2 *twt,
                                                          const float *green, fft_plan pl) {
    // dynamic LDS: b0, b1 (COL_PTS + 32 each, padded columns), tw[nz], twi[nz]
    extern __shared__ float2 lds2[];
    float2 *b0 = lds2, *b1 = lds2 + (COL_PTS + 32), *tw = lds2 + 2 * (COL_PTS + 32), *twi = tw + sp.nz;
    const int nx = sp.nx, nz = sp.nz;
    const int W = col_width(nx, nz);                  // columns per block
    const int LS = nz + 1;                            // padded column stride in LDS
    const int cpp = col_groups(nx, nz);               // column groups per plane
    // neighbouring column groups share cache lines: keep them on one XCD
    const unsigned blk = xcd_block();
    const int64_t plane = blk / cpp;
    const int c0 = (int)(blk - plane * cpp) * W;
    const int Wb = min(W, nx - c0);                   // columns of this block
    const int PTS = W * nz;
    float2 *base = cbuf + plane * (int64_t)nz * nx;
    load_twiddles(tw, twt + 2 * TW_STRIDE, nz);
    load_twiddles(twi, twt + 3 * TW_STRIDE, nz);
    const int logW = P2 ? __builtin_ctz((unsigned)W) : 0, lognz = P2 ? __builtin_ctz((unsigned)nz) : 0;
    for (int e = threadIdx.x; e < PTS; e += BLOCK) {
        const int j = idiv<P2>(e, W, logW), c = e - j * W;
        b0[c * LS + j] = c < Wb ? base[(int64_t)j * nx + c0 + c] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    float2 *o = lds_fft<P2>(b0, b1, tw, pl, false, PTS, LS);
    float2 *other = (o == b0) ? b1 : b0;
    // Green's function of the block's columns: one contiguous run of the
    // transposed table
    const float *gcol = green + (int64_t)c0 * nz;
    for (int e = threadIdx.x; e < Wb * nz; e += BLOCK) {
        const int c = idiv<P2>(e, nz, lognz), j = e - c * nz;
        const float g = gcol[e];
        float2 &v = o[c * LS + j];
        v = make_float2(v.x * g, v.y * g);
    }
    __syncthreads();
    float2 *r = lds_fft<P2>(o, other, twi, pl, true, PTS, LS);
    for (int e = threadIdx.x; e < PTS; e += BLOCK) {
        const int j = idiv<P2>(e, W, logW), c = e - j * W;
        if (c < Wb) base[(int64_t)j * nx + c0 + c] = r[c * LS + j];
    }
}

// rows: inverse FFT along x, normalise by 1/(nx nz), unpack the slice pair.
template <bool P2>
__global__ __launch_bounds__(BLOCK) void k_fft_rows_inv(const float2 *cbuf, float *phi, ffa_sc_params sp,
                                                        const float2 *twt, fft_plan pl, int lognx, int lognz,
                                                        int64_t rows) {
    extern __shared__ float2 lds2[];           // b0[FFT_PTS], b1[FFT_PTS], tw[nx]
    float2 *b0 = lds2, *b1 = lds2 + FFT_PTS, *tw = lds2 + 2 * FFT_PTS;
    const int nx = sp.nx, nz = sp.nz, ns = sp.ns;
    const int lines = row_lines(nx), pts = lines * nx;
    const int64_t row0 = (int64_t)blockIdx.x * lines;
    const int64_t lim = rows * nx - row0 * nx;
    load_twiddles(tw, twt + TW_STRIDE, nx);
    for (int e = threadIdx.x; e < pts; e += BLOCK) b0[e] = e < lim ? cbuf[row0 * nx + e] : make_float2(0.0f, 0.0f);
    __syncthreads();
    const float2 *o = lds_fft<P2>(b0, b1, tw, pl, true, pts);
    const float norm = 1.0f / (float)(nx * nz);
    for (int e = threadIdx.x; e < pts; e += BLOCK) {
        if (e >= lim) continue;
        const int l = idiv<P2>(e, nx, lognx), i = e - l * nx;
        const int64_t row = row0 + l;
        const int64_t plane = P2 ? (row >> lognz) : row / nz, j = row - plane * nz;
        const int64_t sa = 2 * plane, sb = sa + 1;
        phi[(sa * nz + j) * nx + i] = o[e].x * norm;
        if (sb < ns) phi[(sb * nz + j) * nx + i] = o[e].y * norm;
    }
}

// compute_field_from_potential_numba_SBS (src/FFAG_SC_new.py:672-728):
// float32 differences scaled by float32 factors (the reference scales in
// float64 before its float32 store; the two agree to an ulp of the store),
// float32 store.  A thread computes FIELD_ROWS consecutive rows of one column
// (its loads are independent, so their latencies overlap; a wave covers 64
// consecutive columns) and stores (Ex, Ez) interleaved, so the gather reads
// both with one access per corner.
constexpr int FIELD_ROWS = 4;
constexpr size_t FROM_PHI_MAX_CELLS = (size_t)1 << 23;   // 32 MB of float32 potential

using scg::cell_field;
using scg::field_scale;
using scg::make_field_scale;

// PLANAR: (Ex, Ez) stored as two (ns, nz, nx) cubes ex, ez (the reference's
// Ex_cube / Ez_cube) instead of the interleaved exz cube.
template <bool P2, bool PLANAR>
__global__ __launch_bounds__(BLOCK) void k_sc_field(const float *phi, float2 *exz, float *ex, float *ez,
                                                   ffa_sc_params sp, const ffa_sc_result *res, int lognx) {
    const int nx = sp.nx, nz = sp.nz;
    const int nb = (nz + FIELD_ROWS - 1) / FIELD_ROWS;
    const int64_t total = (int64_t)sp.ns * nb * nx;
    // the stencil rows j0 - 1 and j0 + FIELD_ROWS belong to neighbouring
    // bands: keep neighbours on one XCD so each row is fetched into one L2
    const int64_t t = xcd_block() * (int64_t)BLOCK + threadIdx.x;
    if (t >= total) return;
    const int64_t r = P2 ? (t >> lognx) : t / nx;
    const int i = (int)(t - r * nx);
    const int band = (int)(r % nb);
    const int64_t k = r / nb;
    const int j0 = band * FIELD_ROWS;
    const int rows = min(FIELD_ROWS, nz - j0);        // the last band may be shorter
    const int64_t base = k * (int64_t)nz * nx;
    auto store = [&](int j, float2 v) {
        const int64_t o = base + (int64_t)j * nx + i;
        if constexpr (PLANAR) {
            ex[o] = v.x;
            ez[o] = v.y;
        } else {
            exz[o] = v;
        }
    };
    const double dr = res->dx, dz = res->dz, bl = res->bunch_length;
    if (res->bad_grid != 0.0 || dr == 0.0 || dz == 0.0 || bl == 0.0) {
        for (int c = 0; c < rows; ++c) store(j0 + c, make_float2(0.0f, 0.0f));
        return;
    }
    const field_scale fsc = make_field_scale(dr, dz, res->mean_gamma_m1 + 1.0);
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
        const int j = min(j0 + c, nz - 1);              // rows past the grid are loaded but not stored
        lf[c] = P[(int64_t)j * nx + il];
        rt[c] = P[(int64_t)j * nx + ir];
    }
    for (int c = 0; c < rows; ++c) {
        const int j = j0 + c;
        store(j, cell_field(col[c + 1], col[c + 2], col[c], lf[c], rt[c], i, j, nx, nz, fsc));
    }
}

// Inverse row FFTs fused with compute_field_from_potential_numba_SBS (grids
// whose field cubes are stored): a workgroup transforms the rows of one
// slice pair's band [j0, j0 + R) plus the rows j0 - 1 and j0 + R (clamped to
// the grid, as the stencils clamp them), normalises them to phi in LDS and
// writes the (Ex, Ez) cube of the band's rows with k_sc_field's stencils
// (cell_field), so phi need not make a round trip through HBM.  phi itself
// is stored when `phi` is not null.  Same arithmetic as k_fft_rows_inv
// followed by k_sc_field: bit-identical cubes.
__host__ __device__ inline int field_band_lines(int nx) { return 2048 / nx > 4 ? 2048 / nx : 4; }

template <bool P2>
__global__ __launch_bounds__(BLOCK) void k_fft_rows_inv_field(const float2 *cbuf, float *phi, float2 *exz,
                                                              ffa_sc_params sp, const ffa_sc_result *res,
                                                              const float2 *twt, fft_plan pl, int lognx) {
    extern __shared__ float2 lds2[];           // b0, b1 [lines * nx], tw[nx]
    const int nx = sp.nx, nz = sp.nz, ns = sp.ns;
    const int L = field_band_lines(nx), R = L - 2;
    const int nbands = (nz + R - 1) / R;
    const int plane = (int)(blockIdx.x / nbands), band = (int)(blockIdx.x - (unsigned)plane * nbands);
    const int j0 = band * R, rows = min(R, nz - j0);
    const int pts = (rows + 2) * nx;
    float2 *b0 = lds2, *b1 = lds2 + L * nx, *tw = lds2 + 2 * L * nx;
    load_twiddles(tw, twt + TW_STRIDE, nx);
    const float2 *src = cbuf + (int64_t)plane * nz * nx;
    for (int e = threadIdx.x; e < pts; e += BLOCK) {
        const int l = idiv<P2>(e, nx, lognx), i = e - l * nx;
        int j = j0 - 1 + l;
        j = j < 0 ? 0 : (j > nz - 1 ? nz - 1 : j);
        b0[e] = src[(int64_t)j * nx + i];
    }
    __syncthreads();
    float2 *o = lds_fft<P2>(b0, b1, tw, pl, true, pts);
    const float norm = 1.0f / (float)(nx * nz);
    for (int e = threadIdx.x; e < pts; e += BLOCK) o[e] = make_float2(o[e].x * norm, o[e].y * norm);
    __syncthreads();
    const int64_t sa = 2 * (int64_t)plane, sb = sa + 1;
    const bool two = sb < ns;
    if (phi != nullptr)
        for (int e = threadIdx.x; e < rows * nx; e += BLOCK) {
            const int l = idiv<P2>(e, nx, lognx), i = e - l * nx;
            const float2 v = o[(l + 1) * nx + i];
            phi[(sa * nz + j0 + l) * nx + i] = v.x;
            if (two) phi[(sb * nz + j0 + l) * nx + i] = v.y;
        }
    const double dr = res->dx, dz = res->dz, bl = res->bunch_length;
    const bool zero = res->bad_grid != 0.0 || dr == 0.0 || dz == 0.0 || bl == 0.0;
    const field_scale fsc = zero ? field_scale{} : make_field_scale(dr, dz, res->mean_gamma_m1 + 1.0);
    for (int e = threadIdx.x; e < rows * nx; e += BLOCK) {
        const int l = idiv<P2>(e, nx, lognx), i = e - l * nx, j = j0 + l;
        const float2 *c = o + (l + 1) * nx;
        const float2 pc = c[i], pu = c[nx + i], pd = c[i - nx];
        const float2 lf = c[i > 0 ? i - 1 : 0], rt = c[i < nx - 1 ? i + 1 : nx - 1];
        float2 fa = make_float2(0.0f, 0.0f), fb = fa;
        if (!zero) {
            fa = cell_field(pc.x, pu.x, pd.x, lf.x, rt.x, i, j, nx, nz, fsc);
            fb = cell_field(pc.y, pu.y, pd.y, lf.y, rt.y, i, j, nx, nz, fsc);
        }
        exz[(sa * nz + j) * nx + i] = fa;
        if (two) exz[(sb * nz + j) * nx + i] = fb;
    }
}

// interpolate_fields_to_particles_SBS_numba (src/FFAG_SC_new.py:870-972)
// fused with the endpoint roll of src/FFAG_track.py:356-382.
//
// GATHER_U particles per thread (i and i + BLOCK within the block's run of
// GATHER_U * BLOCK): every particle's stream loads are issued together, then
// every particle's four field pairs, so two dependent round trips are in
// flight per particle instead of one particle's chain at a time (the field
// pairs are scattered reads of a cube that stays in the last-level cache).
constexpr int GATHER_U = 2;

template <typename ET, bool FROM_PHI>
__global__ __launch_bounds__(BLOCK) void k_sc_gather_roll(ffa_particles p, ffa_sc_params sp, const ffa_sc_result *res,
                                                         const float2 *exz, const float *phi, const double *px,
                                                         const double *ppr0, const int32_t *pslice) {
    const int64_t base = xcd_block() * (int64_t)(GATHER_U * BLOCK) + threadIdx.x;
    const int64_t n = p.n;
    if (base >= n) return;
    ET *ep = reinterpret_cast<ET *>(p.esc_pre);
    ET *epp = reinterpret_cast<ET *>(p.esc_prepre);
    const bool good = res->bad_grid == 0.0;
    const int nx = sp.nx, nz = sp.nz;
    const double xmin = res->xmin, zmin = res->zmin, idx = 1.0 / res->dx, idz = 1.0 / res->dz;
    // stream loads of every particle (indices past the end read particle base)
    uint8_t fl[GATHER_U];
    int sl[GATHER_U];
    double xv[GATHER_U], zv[GATHER_U], pr[GATHER_U];
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) {
        const int64_t i = base + u * BLOCK;
        const int64_t j = i < n ? i : base;
        fl[u] = p.flags[j];
        sl[u] = pslice[j];
        xv[u] = px[j];
        zv[u] = p.z[j];
        pr[u] = ppr0[j];
    }
    // cells and weights, then every particle's field pairs
    scg::cellw cw[GATHER_U];
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) cw[u] = scg::cell_weights(sl[u], xv[u], zv[u], sp.ns, nx, nz, xmin, zmin, idx, idz);
    const field_scale fsc = FROM_PHI ? make_field_scale(res->dx, res->dz, res->mean_gamma_m1 + 1.0) : field_scale{};
    float2 a00[GATHER_U], a10[GATHER_U], a01[GATHER_U], a11[GATHER_U];
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) scg::corners<FROM_PHI>(cw[u], exz, phi, nx, nz, fsc, a00[u], a10[u], a01[u], a11[u]);
#pragma unroll
    for (int u = 0; u < GATHER_U; ++u) {
        const int64_t i = base + u * BLOCK;
        if (i >= n) break;
        ET a, b, c;
        const uint8_t nfl = scg::roll<ET>(i, n, fl[u], good, cw[u], a00[u], a10[u], a01[u], a11[u], pr[u], sp, ep, epp,
                                          a, b, c);
        if (nfl != fl[u]) p.flags[i] = nfl;
    }
}

inline unsigned gridn(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

inline int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Periodic Poisson solve of every slice pair: forward row FFTs (reading and
// scaling the int64 grid when `grid` is given, else the packed charge already
// in cbuf), column FFT x Green's function x inverse column FFT, inverse row
// FFTs unpacked to phi.  twt / green: the per-solve tables.
void launch_fft_poisson(hipStream_t st, const ffa_sc_params &sp, const ffa_sc_result *res, const int64_t *grid,
                        float2 *cbuf, float *phi, const float2 *twt, const float *green, bool inv_rows = true) {
    const bool p2 = is_pow2(sp.nx) && is_pow2(sp.nz);
    const fft_plan px = make_fft_plan(sp.nx), pz = make_fft_plan(sp.nz);
    const int lognx = ilog2(sp.nx), lognz = ilog2(sp.nz);
    const int64_t planes = (sp.ns + 1) / 2;
    const int64_t rows = planes * sp.nz;
    const int lines = row_lines(sp.nx);
    const unsigned rblocks = (unsigned)((rows + lines - 1) / lines);
    const unsigned cblocks = (unsigned)(planes * col_groups(sp.nx, sp.nz));
    const size_t rows_lds = sizeof(float2) * (2 * FFT_PTS + sp.nx);
    const size_t cols_lds = sizeof(float2) * (2 * (COL_PTS + 32) + 2 * sp.nz);
    if (p2) {
        if (grid)
            hipLaunchKernelGGL(k_fft_rows_fwd_grid<true>, dim3(rblocks), dim3(BLOCK), rows_lds, st, grid, cbuf, sp, res,
                               twt, px, lognx, lognz, rows);
        else
            hipLaunchKernelGGL(k_fft_rows_fwd<true>, dim3(rblocks), dim3(BLOCK), rows_lds, st, cbuf, sp, twt, px, rows);
        hipLaunchKernelGGL(k_fft_cols_green<true>, dim3(cblocks), dim3(BLOCK), cols_lds, st, cbuf, sp, twt, green, pz);
        if (inv_rows)
            hipLaunchKernelGGL(k_fft_rows_inv<true>, dim3(rblocks), dim3(BLOCK), rows_lds, st, cbuf, phi, sp, twt, px,
                               lognx, lognz, rows);
    } else {
        if (grid)
            hipLaunchKernelGGL(k_fft_rows_fwd_grid<false>, dim3(rblocks), dim3(BLOCK), rows_lds, st, grid, cbuf, sp,
                               res, twt, px, lognx, lognz, rows);
        else
            hipLaunchKernelGGL(k_fft_rows_fwd<false>, dim3(rblocks), dim3(BLOCK), rows_lds, st, cbuf, sp, twt, px,
                               rows);
        hipLaunchKernelGGL(k_fft_cols_green<false>, dim3(cblocks), dim3(BLOCK), cols_lds, st, cbuf, sp, twt, green, pz);
        if (inv_rows)
            hipLaunchKernelGGL(k_fft_rows_inv<false>, dim3(rblocks), dim3(BLOCK), rows_lds, st, cbuf, phi, sp, twt, px,
                               lognx, lognz, rows);
    }
}

// -grad(phi) / gamma^2 of every cell: interleaved exz, or the planar ex / ez
// cubes when exz is null.
void launch_field(hipStream_t st, const ffa_sc_params &sp, const ffa_sc_result *res, const float *phi, float2 *exz,
                  float *ex, float *ez) {
    const int64_t threads = (int64_t)sp.ns * ((sp.nz + FIELD_ROWS - 1) / FIELD_ROWS) * sp.nx;
    const dim3 g(gridn(threads));
    const int lognx = ilog2(sp.nx);
    if (is_pow2(sp.nx)) {
        if (exz) hipLaunchKernelGGL((k_sc_field<true, false>), g, dim3(BLOCK), 0, st, phi, exz, ex, ez, sp, res, lognx);
        else hipLaunchKernelGGL((k_sc_field<true, true>), g, dim3(BLOCK), 0, st, phi, exz, ex, ez, sp, res, lognx);
    } else {
        if (exz) hipLaunchKernelGGL((k_sc_field<false, false>), g, dim3(BLOCK), 0, st, phi, exz, ex, ez, sp, res, lognx);
        else hipLaunchKernelGGL((k_sc_field<false, true>), g, dim3(BLOCK), 0, st, phi, exz, ex, ez, sp, res, lognx);
    }
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

bool sc_gather_inputs(const ffa_sc_params *sp, void *ws, const ffa_sc_layout *L, sc_gather_src *g) {
    char *w = reinterpret_cast<char *>(ws);
    const size_t C = (size_t)sp->ns * sp->nz * sp->nx;
    g->sp = *sp;
    g->res = reinterpret_cast<const ffa_sc_result *>(w + L->result);
    g->exz = reinterpret_cast<const float2 *>(w + L->exz);
    g->phi = reinterpret_cast<const float *>(w + L->phi);
    g->px = reinterpret_cast<const double *>(w + L->px);
    g->ppr0 = reinterpret_cast<const double *>(w + L->ppr0);
    g->pslice = reinterpret_cast<const int32_t *>(w + L->pslice);
    return !(sp->field_cubes != 0 || C > FROM_PHI_MAX_CELLS);
}

// gather = false stops after the field (cubes or phi): the particles'
// endpoint fields and flags are left for launch_boris_fused_f32_gather.
hipError_t launch_sc_solve(ffa_ctx_s *c, const ffa_particles *p, const ffa_sc_params *sp, void *ws,
                           const ffa_sc_layout *L, bool gather) {
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
    hipLaunchKernelGGL(k_sc_front, dim3(NB), dim3(BLOCK), 0, st, *p, *sp, sp->c, E0J, px, ppr0, part);
    hipLaunchKernelGGL(k_sc_final1, dim3(1), dim3(BLOCK), 0, st, part, *sp, res, slc);
    hipLaunchKernelGGL(k_sc_slices, dim3(NB), dim3(BLOCK), 0, st, *p, *sp, res, pslice, slc, blk_off);
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
    const int nbands = (sp->nz + bz - 1) / bz;
    if (pairs)
        hipLaunchKernelGGL(k_sc_deposit_pairs, dim3((unsigned)(((sp->ns + 1) / 2) * nbands)), dim3(BLOCK), tile_lds,
                           st, *sp, res, rstart, rec2, reinterpret_cast<unsigned long long *>(grid), cell, cbuf, bz);
    else
        hipLaunchKernelGGL(k_sc_deposit_slices, dim3((unsigned)(sp->ns * nbands)), dim3(BLOCK), tile_lds, st, *sp, res,
                           rstart, rec2, reinterpret_cast<unsigned long long *>(grid), cell, bz);
    // The gather from phi reads 12 potential values per particle instead of 4
    // field pairs: cheaper while the potential stays cache-resident (measured
    // faster at 4.2M cells, slower at 33.5M), so larger grids store the cubes.
    // With cubes the inverse row FFTs and the field stencils run as one
    // kernel; phi is stored only when the caller asked for the cubes (a
    // checking mode that keeps every intermediate).
    const bool cubes = sp->field_cubes != 0 || C > FROM_PHI_MAX_CELLS;
    launch_fft_poisson(st, *sp, res, pairs ? nullptr : grid, cbuf, phi, twt, green, !cubes);
    if (cubes) {
        const fft_plan plx = make_fft_plan(sp->nx);
        const int L = field_band_lines(sp->nx), R = L - 2;
        const dim3 g((unsigned)(((sp->ns + 1) / 2) * ((sp->nz + R - 1) / R)));
        const size_t lds_b = sizeof(float2) * ((size_t)2 * L * sp->nx + sp->nx);
        float *phi_out = sp->field_cubes != 0 ? phi : nullptr;
        if (is_pow2(sp->nx))
            hipLaunchKernelGGL(k_fft_rows_inv_field<true>, g, dim3(BLOCK), lds_b, st, cbuf, phi_out, exz, *sp, res, twt,
                               plx, ilog2(sp->nx));
        else
            hipLaunchKernelGGL(k_fft_rows_inv_field<false>, g, dim3(BLOCK), lds_b, st, cbuf, phi_out, exz, *sp, res,
                               twt, plx, 0);
    }
    if (gather && p->n > 0) {
        const dim3 g((unsigned)((p->n + GATHER_U * BLOCK - 1) / (GATHER_U * BLOCK)));
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

// ===========================================================================
// Stage-level entry points: the module functions of FFAG_SC_new, each on its
// own (survivor compaction, covering arc, slicing and moments, deposit,
// Green's function, Poisson solve, field, longitudinal field, gather), for
// callers that drive the stages themselves.  Same arithmetic as the fused
// solve above where the reference's stage and the fused pipeline agree; the
// deposit is the fixed-point one (integer sums, bit-reproducible).
// ===========================================================================
namespace {

// GetLocalCoordinates_2p5(_rk) (src/FFAG_SC_new.py:74-148): key 0 for the
// survivors, 1 for the others; a stable sort by key keeps the survivors in
// row order first.
__global__ __launch_bounds__(BLOCK) void k_stage_survive_keys(const uint8_t *survive, int64_t n, uint16_t *keys) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i < n) keys[i] = survive[i] ? 0 : 1;
}

__device__ __forceinline__ double row_gamma(const double *row, int rk, double c, double E0J) {
    if (rk) return row[5] / E0J;
    const double r = row[0], drdt = row[1], dzdt = row[3], rd = r * row[5];
    const double v2 = drdt * drdt + dzdt * dzdt + rd * rd;
    return 1.0 / sqrt(1.0 - v2 / (c * c));
}

// out planes (n each): r, fi, z, gamma, fi mod 2 pi, source row index
__global__ __launch_bounds__(BLOCK) void k_stage_local_coords(const double *bunch, const uint8_t *survive,
                                                             const uint32_t *order, int64_t n, double c, int rk,
                                                             double E0J, double *out) {
    const int64_t j = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (j >= n) return;
    const int64_t i = order[j];
    if (!survive[i]) return;                      // survivors occupy positions [0, count)
    const double *row = bunch + i * 16;
    const double fi = row[4];
    out[j] = row[0];
    out[n + j] = fi;
    out[2 * n + j] = row[2];
    out[3 * n + j] = row_gamma(row, rk, c, E0J);
    out[4 * n + j] = fi - FFA_TWO_PI * floor(fi / FFA_TWO_PI);
    out[5 * n + j] = (double)i;
}

// per-block (count, sum gamma) over the survivors, fixed partition
__global__ __launch_bounds__(BLOCK) void k_stage_gamma_partials(const double *bunch, const uint8_t *survive, int64_t n,
                                                               double c, int rk, double E0J, double *part) {
    __shared__ double sh[2][BLOCK / 64];
    double v[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)NB * BLOCK)
        if (survive[i]) {
            v[0] += 1.0;
            v[1] += row_gamma(bunch + i * 16, rk, c, E0J);
        }
    const int op[2] = {0, 0};
    block_reduce<2>(sh, v, op);
    if (threadIdx.x == 0) { part[16 * blockIdx.x] = v[0]; part[16 * blockIdx.x + 1] = v[1]; }
}

// combine operation of each partial column: 0 sum, 1 min, 2 max
struct reduce_ops {
    int op[8];
};

// generic final: the NV partial columns of the fixed block partition
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_stage_reduce_final(const double *part, reduce_ops ops, double *out) {
    __shared__ double sh[NV][BLOCK / 64];
    int op[NV];
    double v[NV];
    for (int k = 0; k < NV; ++k) {
        op[k] = ops.op[k];
        v[k] = op[k] == 0 ? 0.0 : (op[k] == 1 ? 1e30 : -1e30);
    }
    for (int b = threadIdx.x; b < NB; b += BLOCK)
        for (int k = 0; k < NV; ++k) v[k] = combine(op[k], v[k], part[16 * b + k]);
    block_reduce<NV>(sh, v, op);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) out[k] = v[k];
}

// divide_arc2 (src/FFAG_SC_new.py:534-566): per-block min / max of the
// azimuths shifted by 0, pi, pi/2, 3 pi/2, reduced mod 2 pi
__global__ __launch_bounds__(BLOCK) void k_stage_arc_partials(const double *fi, int64_t n, double *part) {
    __shared__ double sh[8][BLOCK / 64];
    const double shifts[4] = {0.0, FFA_PI, FFA_PI * 0.5, FFA_PI * 1.5};
    double v[8] = {1e30, 1e30, 1e30, 1e30, -1e30, -1e30, -1e30, -1e30};
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)NB * BLOCK)
        for (int k = 0; k < 4; ++k) {
            const double a = fi[i] + shifts[k], b = 2.0 * FFA_PI;
            const double q = a - b * floor(a / b);
            v[k] = q < v[k] ? q : v[k];
            v[4 + k] = q > v[4 + k] ? q : v[4 + k];
        }
    const int op[8] = {1, 1, 1, 1, 2, 2, 2, 2};
    block_reduce<8>(sh, v, op);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) part[16 * blockIdx.x + k] = v[k];
}

__global__ void k_stage_arc_final(const double *mm, double *out) {
    const double shifts[4] = {0.0, FFA_PI, FFA_PI * 0.5, FFA_PI * 1.5};
    int best = 0;
    double bs = mm[4] - mm[0];
    for (int k = 1; k < 4; ++k)
        if (mm[4 + k] - mm[k] < bs) { bs = mm[4 + k] - mm[k]; best = k; }
    out[0] = mm[best] - shifts[best];
    out[1] = mm[4 + best] - shifts[best];
    out[2] = bs;
}

__global__ __launch_bounds__(BLOCK) void k_stage_zero_i32(int32_t *a, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ __launch_bounds__(BLOCK) void k_stage_zero_u64(unsigned long long *a, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK) a[i] = 0ull;
}

// ComputeSliceBoundsAndCoordsSBS_merged (src/FFAG_SC_new.py:449-531):
// x = r - r0, the slice of each particle, the per-slice counts, per-block
// moment partials; key = slice for the stable placement that follows
__global__ __launch_bounds__(BLOCK) void k_stage_slice_keys(const double *r, const double *z, const double *fim,
                                                           const double *r0, int64_t n, double step, double fs, int ns,
                                                           double *x_out, int32_t *sl_out, uint16_t *keys,
                                                           int32_t *counts, double *part) {
    __shared__ double sh[4][BLOCK / 64];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < n; i += (int64_t)NB * BLOCK) {
        const double x = r[i] - r0[i], zz = z[i];
        x_out[i] = x;
        v[0] += x; v[1] += zz; v[2] += x * x; v[3] += zz * zz;
        double f = fim[i];
        if (f < fs) f += FFA_TWO_PI;
        double q = (f - fs) / step;
        q = q > -1e9 ? (q < 1e9 ? q : 1e9) : -1e9;
        int64_t s = (int64_t)q;
        if (s < 0) s = 0; else if (s >= ns) s = ns - 1;
        sl_out[i] = (int32_t)s;
        keys[i] = (uint16_t)s;
        atomicAdd(&counts[s], 1);
    }
    const int op[4] = {0, 0, 0, 0};
    block_reduce<4>(sh, v, op);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[16 * blockIdx.x + k] = v[k];
}

// exclusive scan of ns <= 4096 counts in one block
__global__ __launch_bounds__(BLOCK) void k_stage_scan_counts(const int32_t *counts, int ns, int32_t *start) {
    __shared__ int32_t tot[BLOCK];
    const int per = (ns + BLOCK - 1) / BLOCK;
    const int k0 = threadIdx.x * per, k1 = min(k0 + per, ns);
    int32_t own = 0;
    for (int k = k0; k < k1; ++k) own += counts[k];
    tot[threadIdx.x] = own;
    __syncthreads();
    for (int w = 1; w < BLOCK; w <<= 1) {
        const int32_t add = threadIdx.x >= w ? tot[threadIdx.x - w] : 0;
        __syncthreads();
        tot[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t acc = tot[threadIdx.x] - own;
    for (int k = k0; k < k1; ++k) { start[k] = acc; acc += counts[k]; }
}

// the per-slice lists in particle order: sorted position j holds particle
// order[j] of slice key[j] at place j - start[key[j]] (places at or beyond
// the row stride are not written)
__global__ __launch_bounds__(BLOCK) void k_stage_slice_place(const uint32_t *order, const uint16_t *key,
                                                            const int32_t *start, const double *r, const double *z,
                                                            const double *x, int64_t n, int64_t stride, double *r2d,
                                                            double *z2d, double *x2d) {
    const int64_t j = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (j >= n) return;
    const int64_t i = order[j], s = key[j];
    const int64_t pos = j - start[s];
    if (pos >= stride) return;
    const int64_t o = s * stride + pos;
    if (r2d) r2d[o] = r[i];
    z2d[o] = z[i];
    x2d[o] = x[i];
}

// DistributeChargeSBS_Njit (src/FFAG_SC_new.py:244-312) in fixed point: the
// CIC cell and clamped fractions of the reference, rounded to `bits`
// fractional bits, added as integers (one workgroup per slice)
__global__ __launch_bounds__(BLOCK) void k_stage_deposit(const double *z2d, const double *x2d, int64_t stride,
                                                        const int32_t *counts, int nz, int nx, double xmin, double zmin,
                                                        double inv_dx, double inv_dz, int bits,
                                                        unsigned long long *grid) {
    const int s = blockIdx.x;
    const int64_t cnt = min((int64_t)counts[s], stride);
    const uint64_t one = (uint64_t)1 << bits;
    const double scale = (double)one;
    unsigned long long *g = grid + (size_t)s * nz * nx;
    for (int64_t k = threadIdx.x; k < cnt; k += BLOCK) {
        const double rx = (x2d[s * stride + k] - xmin) * inv_dx, rz = (z2d[s * stride + k] - zmin) * inv_dz;
        const double crx = rx > -1e9 ? (rx < 1e9 ? rx : 1e9) : -1e9;
        const double crz = rz > -1e9 ? (rz < 1e9 ? rz : 1e9) : -1e9;
        int64_t ix = (int64_t)floor(crx), iz = (int64_t)floor(crz);
        double fx = rx - (double)ix, fz = rz - (double)iz;
        if (ix < 0) { ix = 0; fx = 0.0; } else if (ix >= nx - 1) { ix = nx - 2; fx = 1.0; }
        if (iz < 0) { iz = 0; fz = 0.0; } else if (iz >= nz - 1) { iz = nz - 2; fz = 1.0; }
        const uint64_t a = (uint64_t)rint(fx * scale), b = (uint64_t)rint(fz * scale);
        const size_t o = (size_t)iz * nx + ix;
        atomicAdd(&g[o], (unsigned long long)((one - a) * (one - b)));
        atomicAdd(&g[o + 1], (unsigned long long)(a * (one - b)));
        atomicAdd(&g[o + nx], (unsigned long long)((one - a) * b));
        atomicAdd(&g[o + nx + 1], (unsigned long long)(a * b));
    }
}

__global__ __launch_bounds__(BLOCK) void k_stage_grid_to_rho(const int64_t *grid, int64_t C, double us, float *rho) {
    for (int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x; i < C; i += (int64_t)gridDim.x * BLOCK)
        rho[i] = (float)((double)grid[i] * us);
}

// the scalars the FFT tables and the field kernel read
__global__ void k_stage_set_result(ffa_sc_result *res, double dx, double dz, double bl, double mgm1, double bad) {
    res->dx = dx;
    res->dz = dz;
    res->bunch_length = bl;
    res->mean_gamma_m1 = mgm1;
    res->bad_grid = bad;
}

__global__ __launch_bounds__(BLOCK) void k_stage_tables(ffa_sc_params sp, const ffa_sc_result *res, float2 *tw,
                                                       float *green) {
    const int64_t n_tab = 4 * TW_STRIDE + (int64_t)sp.nx * sp.nz;
    for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < n_tab; t += (int64_t)gridDim.x * BLOCK)
        sc_table_entry(t, tw, green, sp, res);
}

// make_fft_kernel_open_numba's (nz, nx) layout from the transposed table
__global__ __launch_bounds__(BLOCK) void k_stage_green_out(const float *green_t, int nz, int nx, float *out) {
    const int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (t >= (int64_t)nz * nx) return;
    const int j = (int)(t / nx), i = (int)(t - (int64_t)j * nx);
    out[t] = green_t[(int64_t)i * nz + j];
}

// real slices -> packed pairs (re: slice 2p, im: slice 2p + 1)
__global__ __launch_bounds__(BLOCK) void k_stage_pack(const float *rho, int ns, int64_t plane_cells, float2 *cbuf) {
    const int64_t planes = (ns + 1) / 2;
    for (int64_t t = blockIdx.x * (int64_t)BLOCK + threadIdx.x; t < planes * plane_cells;
         t += (int64_t)gridDim.x * BLOCK) {
        const int64_t p = t / plane_cells, c = t - p * plane_cells;
        const float re = rho[(2 * p) * plane_cells + c];
        const float im = 2 * p + 1 < ns ? rho[(2 * p + 1) * plane_cells + c] : 0.0f;
        cbuf[t] = make_float2(re, im);
    }
}

// total_field_1d_numba (src/FFAG_SC_new.py:763-782)
__global__ __launch_bounds__(BLOCK) void k_stage_total_field_1d(const double *s, const double *q, int n, double *E) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const double K = 1.0 / (4.0 * FFA_PI * EPS0);
    if (n < 2 || s[1] - s[0] <= 0.0) { E[i] = 0.0; return; }
    const double si = s[i];
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const double r = si - s[j];
        const double sg = r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : 0.0);
        acc += q[j] * sg / (r * r);
    }
    E[i] = K * acc;
}

// interpolate_fields_to_particles_SBS_numba (src/FFAG_SC_new.py:870-972)
__global__ __launch_bounds__(BLOCK) void k_stage_interp(const float *ex, const float *ez, int ns, int nz, int nx,
                                                       const double *x, const double *z, const double *pr0,
                                                       const int32_t *slice, const int64_t *rows, int64_t n, double x0,
                                                       double z0, double dx, double dz, double bl, double *ex_p,
                                                       double *ez_p, double *ef_p, double *bunch, int64_t n_rows) {
    const int64_t k = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (k >= n) return;
    const int64_t row = rows ? rows[k] : -1;
    const bool has_row = bunch && row >= 0 && row < n_rows;
    double vr = 0.0, vz = 0.0, vf = 0.0;
    if (!((dx <= 0.0) || (dz <= 0.0) || (bl <= 0.0))) {
        int s = slice[k];
        s = s < 0 ? 0 : (s >= ns ? ns - 1 : s);
        const float *EX = ex + (size_t)s * nz * nx, *EZ = ez + (size_t)s * nz * nx;
        const double t = (x[k] - x0) / dx, u = (z[k] - z0) / dz;
        const double ct = t > -1e9 ? (t < 1e9 ? t : 1e9) : -1e9, cu = u > -1e9 ? (u < 1e9 ? u : 1e9) : -1e9;
        int64_t i0 = (int64_t)floor(ct), j0 = (int64_t)floor(cu);
        if (i0 < 0) i0 = 0; else if (i0 > nx - 2) i0 = nx - 2;
        if (j0 < 0) j0 = 0; else if (j0 > nz - 2) j0 = nz - 2;
        const double al = t - (double)i0, be = u - (double)j0;
        const double w00 = (1.0 - al) * (1.0 - be), w10 = al * (1.0 - be), w01 = (1.0 - al) * be, w11 = al * be;
        const int64_t b = j0 * nx + i0;
        const double er = w00 * EX[b] + w10 * EX[b + 1] + w01 * EX[b + nx] + w11 * EX[b + nx + 1];
        const double ezv = w00 * EZ[b] + w10 * EZ[b + 1] + w01 * EZ[b + nx] + w11 * EZ[b + nx + 1];
        const double p = pr0[k];
        vr = er * (1 / sqrt(1 + p * p));
        vz = ezv;
        vf = er * (-p / sqrt(1 + p * p));
    }
    ex_p[k] = vr;
    ez_p[k] = vz;
    ef_p[k] = vf;
    if (has_row) {
        bunch[row * 16 + 10] = vr;
        bunch[row * 16 + 11] = vz;
        bunch[row * 16 + 12] = vf;
    }
}

inline unsigned grid_cap(int64_t n) {
    const int64_t g = (n + BLOCK - 1) / BLOCK;
    return (unsigned)(g < 65536 ? (g > 0 ? g : 1) : 65536);
}

constexpr size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

// Workspace of the stage entry points (regions used by different stages may
// overlap in time only within one call): sort buffers for n keys, the int64
// grid, the packed FFT buffer, the FFT tables, partial sums and a result.
size_t sc_stage_workspace_bytes(int64_t n, int32_t ns, int32_t nz, int32_t nx) {
    const size_t C = (size_t)ns * nz * nx, P = (size_t)((ns + 1) / 2) * nz * nx;
    return al256(sort_workspace_bytes(n)) + al256(8 * C) + al256(8 * P) +
           al256(8 * 4 * (size_t)TW_STRIDE + 4 * (size_t)nz * nx) + al256(sizeof(double) * 16 * NB) + 2 * 256 +
           al256(4 * (size_t)ns) * 2;
}

namespace {
struct stage_ws {
    char *sort, *grid, *cbuf, *tables, *part, *res, *misc, *ints;
};
stage_ws stage_layout(void *ws, int64_t n, int32_t ns, int32_t nz, int32_t nx) {
    const size_t C = (size_t)ns * nz * nx, P = (size_t)((ns + 1) / 2) * nz * nx;
    char *w = static_cast<char *>(ws);
    stage_ws L;
    L.sort = w; w += al256(sort_workspace_bytes(n));
    L.grid = w; w += al256(8 * C);
    L.cbuf = w; w += al256(8 * P);
    L.tables = w; w += al256(8 * 4 * (size_t)TW_STRIDE + 4 * (size_t)nz * nx);
    L.part = w; w += al256(sizeof(double) * 16 * NB);
    L.res = w; w += 256;
    L.misc = w; w += 256;
    L.ints = w;
    return L;
}
}  // namespace

hipError_t launch_stage_local_coords(ffa_ctx_s *c, const double *bunch, const uint8_t *survive, int64_t n,
                                     double speed_c, int rk, double *out, double *count_sum, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, n, 1, 1, 1);   // as ffa_sc_local_coords checks it
    const double E0J = 938.2723e6 * 1.60217662e-19;   // src/FFAG_SC_new.py:115
    double *part = reinterpret_cast<double *>(L.part);
    hipLaunchKernelGGL(k_stage_gamma_partials, dim3(NB), dim3(BLOCK), 0, st, bunch, survive, n, speed_c, rk, E0J, part);
    hipLaunchKernelGGL(k_stage_reduce_final<2>, dim3(1), dim3(BLOCK), 0, st, part, reduce_ops{{0, 0}}, count_sum);
    hipError_t e;
    if (n > 0) {
        hipLaunchKernelGGL(k_stage_survive_keys, dim3(gridn(n)), dim3(BLOCK), 0, st, survive, n, sort_key_buffer(L.sort));
        const uint32_t *order;
        const uint16_t *keys;
        e = launch_sort_keys16(c, n, 1, L.sort, &order, &keys);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_stage_local_coords, dim3(gridn(n)), dim3(BLOCK), 0, st, bunch, survive, order, n, speed_c,
                           rk, E0J, out);
    }
    return hipGetLastError();
}

hipError_t launch_stage_divide_arc(ffa_ctx_s *c, const double *fi, int64_t n, double *out, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, 0, 1, 1, 1);   // as ffa_sc_divide_arc checks it
    double *part = reinterpret_cast<double *>(L.part);
    double *mm = reinterpret_cast<double *>(L.res);
    hipLaunchKernelGGL(k_stage_arc_partials, dim3(NB), dim3(BLOCK), 0, st, fi, n, part);
    hipLaunchKernelGGL(k_stage_reduce_final<8>, dim3(1), dim3(BLOCK), 0, st, part, reduce_ops{{1, 1, 1, 1, 2, 2, 2, 2}},
                       mm);
    hipLaunchKernelGGL(k_stage_arc_final, dim3(1), dim3(1), 0, st, mm, out);
    return hipGetLastError();
}

hipError_t launch_stage_slices(ffa_ctx_s *c, const double *r, const double *z, const double *fim, const double *r0,
                               int64_t n, double step, double fs, int ns, double *x, double *r2d, double *z2d,
                               double *x2d, int64_t stride, int32_t *counts, int32_t *sl, double *sums, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, n, ns, 1, 1);
    double *part = reinterpret_cast<double *>(L.part);
    int32_t *start = reinterpret_cast<int32_t *>(L.ints);
    hipError_t e;
    hipLaunchKernelGGL(k_stage_zero_i32, dim3(gridn(ns)), dim3(BLOCK), 0, st, counts, (int64_t)ns);
    hipLaunchKernelGGL(k_stage_slice_keys, dim3(NB), dim3(BLOCK), 0, st, r, z, fim, r0, n, step, fs, ns, x, sl,
                       sort_key_buffer(L.sort), counts, part);
    hipLaunchKernelGGL(k_stage_reduce_final<4>, dim3(1), dim3(BLOCK), 0, st, part, reduce_ops{{0, 0, 0, 0}}, sums);
    hipLaunchKernelGGL(k_stage_scan_counts, dim3(1), dim3(BLOCK), 0, st, counts, ns, start);
    if (n > 0) {
        const uint32_t *order;
        const uint16_t *keys;
        e = launch_sort_keys16(c, n, ns - 1, L.sort, &order, &keys);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_stage_slice_place, dim3(gridn(n)), dim3(BLOCK), 0, st, order, keys, start, r, z, x, n,
                           stride, r2d, z2d, x2d);
    }
    return hipGetLastError();
}

hipError_t launch_stage_deposit(ffa_ctx_s *c, const double *z2d, const double *x2d, int64_t stride,
                                const int32_t *counts, int ns, int nz, int nx, double q_macro, double bl, double xmin,
                                double xmax, double zmin, double zmax, int bits, float *rho, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, 0, ns, nz, nx);
    unsigned long long *grid = reinterpret_cast<unsigned long long *>(L.grid);
    const int64_t C = (int64_t)ns * nz * nx;
    const double dx = (xmax - xmin) / nx, dz = (zmax - zmin) / nz;
    const bool ok = dx > 0.0 && dz > 0.0 && bl > 0.0;
    hipLaunchKernelGGL(k_stage_zero_u64, dim3(grid_cap(C)), dim3(BLOCK), 0, st, grid, C);
    if (ok)
        hipLaunchKernelGGL(k_stage_deposit, dim3((unsigned)ns), dim3(BLOCK), 0, st, z2d, x2d, stride, counts, nz, nx,
                           xmin, zmin, 1.0 / dx, 1.0 / dz, bits, grid);
    const double q = ok ? (q_macro / bl * ns) * (1.0 / (dx * dz)) : 0.0;
    const double us = q / (double)((int64_t)1 << (2 * bits));
    hipLaunchKernelGGL(k_stage_grid_to_rho, dim3(grid_cap(C)), dim3(BLOCK), 0, st,
                       reinterpret_cast<const int64_t *>(grid), C, us, rho);
    return hipGetLastError();
}

hipError_t launch_stage_green(ffa_ctx_s *c, int nx, int nz, double dx, double dz, float *out, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, 0, 2, nz, nx);
    ffa_sc_params sp{};
    sp.ns = 2; sp.nx = nx; sp.nz = nz;
    ffa_sc_result *res = reinterpret_cast<ffa_sc_result *>(L.res);
    float2 *tw = reinterpret_cast<float2 *>(L.tables);
    float *green = reinterpret_cast<float *>(L.tables + 8 * 4 * (size_t)TW_STRIDE);
    hipLaunchKernelGGL(k_stage_set_result, dim3(1), dim3(1), 0, st, res, dx, dz, 1.0, 0.0, 0.0);
    hipLaunchKernelGGL(k_stage_tables, dim3(grid_cap(4 * TW_STRIDE + (int64_t)nx * nz)), dim3(BLOCK), 0, st, sp, res,
                       tw, green);
    hipLaunchKernelGGL(k_stage_green_out, dim3(gridn((int64_t)nx * nz)), dim3(BLOCK), 0, st, green, nz, nx, out);
    return hipGetLastError();
}

hipError_t launch_stage_poisson(ffa_ctx_s *c, const float *rho, int ns, int nz, int nx, double dx, double dz, float *phi,
                                void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, 0, ns, nz, nx);
    ffa_sc_params sp{};
    sp.ns = ns; sp.nx = nx; sp.nz = nz;
    ffa_sc_result *res = reinterpret_cast<ffa_sc_result *>(L.res);
    float2 *tw = reinterpret_cast<float2 *>(L.tables);
    float *green = reinterpret_cast<float *>(L.tables + 8 * 4 * (size_t)TW_STRIDE);
    float2 *cbuf = reinterpret_cast<float2 *>(L.cbuf);
    hipLaunchKernelGGL(k_stage_set_result, dim3(1), dim3(1), 0, st, res, dx, dz, 1.0, 0.0, 0.0);
    hipLaunchKernelGGL(k_stage_tables, dim3(grid_cap(4 * TW_STRIDE + (int64_t)nx * nz)), dim3(BLOCK), 0, st, sp, res,
                       tw, green);
    const int64_t planes = (ns + 1) / 2, pc = (int64_t)nz * nx;
    hipLaunchKernelGGL(k_stage_pack, dim3(grid_cap(planes * pc)), dim3(BLOCK), 0, st, rho, ns, pc, cbuf);
    launch_fft_poisson(st, sp, res, nullptr, cbuf, phi, tw, green);
    return hipGetLastError();
}

hipError_t launch_stage_field(ffa_ctx_s *c, const float *phi, int ns, int nz, int nx, double dr, double dz, double bl,
                              double mgm1, float *ex, float *ez, void *ws) {
    hipStream_t st = c->stream;
    const stage_ws L = stage_layout(ws, 0, 1, 1, 1);   // as ffa_sc_field checks it: only the result block
    ffa_sc_params sp{};
    sp.ns = ns; sp.nx = nx; sp.nz = nz;
    ffa_sc_result *res = reinterpret_cast<ffa_sc_result *>(L.res);
    hipLaunchKernelGGL(k_stage_set_result, dim3(1), dim3(1), 0, st, res, dr, dz, bl, mgm1, 0.0);
    launch_field(st, sp, res, phi, nullptr, ex, ez);
    return hipGetLastError();
}

hipError_t launch_stage_total_field_1d(ffa_ctx_s *c, const double *s, const double *q, int n, double *E) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_total_field_1d, dim3(gridn(n)), dim3(BLOCK), 0, c->stream, s, q, n, E);
    return hipGetLastError();
}

hipError_t launch_stage_interp(ffa_ctx_s *c, const float *ex, const float *ez, int ns, int nz, int nx, const double *x,
                               const double *z, const double *pr0, const int32_t *slice, const int64_t *rows, int64_t n,
                               double x0, double z0, double dx, double dz, double bl, double *ex_p, double *ez_p,
                               double *ef_p, double *bunch, int64_t n_rows) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stage_interp, dim3(gridn(n)), dim3(BLOCK), 0, c->stream, ex, ez, ns, nz, nx, x, z, pr0, slice,
                       rows, n, x0, z0, dx, dz, bl, ex_p, ez_p, ef_p, bunch, n_rows);
    return hipGetLastError();
}
