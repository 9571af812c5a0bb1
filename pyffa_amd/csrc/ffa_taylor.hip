// Taylor-coefficient map generator (gfx950).
//
// For the mid-plane field B(r, theta) = A(r) G(theta - s(r)) the maps of
// src/FFAG_HighOrder.py:936-998 are
//   Bz_n  = (-1)^n/(2n)!      Lap^n B
//   Br_n  = (-1)^n/(2n+1)!    d/dr Lap^n B
//   Bfi_n = (-1)^n/(2n+1)! / r d/dtheta Lap^n B
// and every one of them, at one radius, is a linear combination of the
// flutter derivatives G^(q) evaluated at theta - s(r) (the chain rule moves
// all r dependence into scalar factors of A, s and their derivatives).  The
// host reduces the operator algebra to those scalar weights per radius; the
// GPU does the per-point work in two kernels:
//   k_flutter_table  G^(q)(theta_j), q < n_q, from the sigmoid-product poles
//                    (n_q x n_phi values: a few hundred KB that stay in L2);
//   k_taylor_fill    map = sum_q w_q G^(q)(theta_(j - shift)), one grid point
//                    per thread, every order and component at once.
// The fill writes 8 B (or 4 B) per map value and reads only L2-resident
// tables: it is bound by HBM write bandwidth.
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;
constexpr int MAX_Q = 2 * FFA_TAYLOR_MAX_ORDERS;

// P(sigma) by Horner over the left-padded coefficient row (leading zeros keep
// the value exactly 0 until the first coefficient, as without the padding).
__device__ __forceinline__ double horner(const double *__restrict__ c, int len, double s) {
    double v = 0.0;
    for (int i = 0; i < len; ++i) v = v * s + c[i];
    return v;
}

// One azimuth per thread, NQ derivative orders held in registers; 64-thread
// blocks spread the few thousand azimuths over many CUs.
constexpr int FLUTTER_BLOCK = 64;

template <int NQ>
__global__ __launch_bounds__(FLUTTER_BLOCK) void k_flutter_table(const double *__restrict__ theta, int64_t n_phi,
                                                                const double *__restrict__ poles, int n_poles,
                                                                const double *__restrict__ sig_poly,
                                                                double *__restrict__ table) {
    const int64_t j = blockIdx.x * (int64_t)FLUTTER_BLOCK + threadIdx.x;
    if (j >= n_phi) return;
    const double x = theta[j];
    constexpr int len = NQ + 1;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int p = 0; p < n_poles; ++p) {
        const double x0 = poles[4 * p], x1 = poles[4 * p + 1], k = poles[4 * p + 2], h = poles[4 * p + 3];
        const double s1 = 1.0 / (1.0 + exp(-k * (x - x0)));
        const double s2 = 1.0 / (1.0 + exp(k * (x - x1)));
        const double *rise = sig_poly + (size_t)(2 * p) * NQ * len;
        const double *fall = rise + (size_t)NQ * len;
        double d1[NQ], d2[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            d1[i] = horner(rise + (size_t)i * len, len, s1);
            d2[i] = horner(fall + (size_t)i * len, len, s2);
        }
        acc[0] += h * s1 * s2;
#pragma unroll
        for (int q = 1; q < NQ; ++q) {
            double t = 0.0, binom = 1.0;                  // C(q, i), exact in double for q < 2^53
#pragma unroll
            for (int i = 0; i <= q; ++i) {
                t += binom * d1[i] * d2[q - i];
                binom = binom * (double)(q - i) / (double)(i + 1);
            }
            acc[q] += h * t;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) table[(size_t)q * n_phi + j] = acc[q];
}

template <typename T>
__device__ __forceinline__ void put(void *out, size_t idx, double v) {
    static_cast<T *>(out)[idx] = (T)v;
}

// Row shift reduced to [0, n_phi) in 32-bit arithmetic.
__device__ __forceinline__ int wrap_shift(int32_t sh, int n_phi) {
    int s = sh % n_phi;
    return s < 0 ? s + n_phi : s;
}

// Stacks layout: block row = one radius, threads along theta, so the
// weights are uniform over the block and every store is coalesced.
template <typename T, int NO>
__global__ __launch_bounds__(BLOCK) void k_taylor_fill_stacks(const double *__restrict__ table, int n_phi, int nr,
                                                              const double *__restrict__ weights,
                                                              const int32_t *__restrict__ shift,
                                                              void *__restrict__ out) {
    constexpr int n_orders = NO, n_q = 2 * NO;
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    const int ir = blockIdx.y;
    if (j >= n_phi) return;
    int jj = j - wrap_shift(shift[ir], n_phi);
    if (jj < 0) jj += n_phi;
    double g[MAX_Q];
    for (int q = 0; q < n_q; ++q) g[q] = table[(size_t)q * n_phi + jj];
    const double *w = weights + (size_t)ir * 3 * n_orders * n_q;
    const size_t plane = (size_t)nr * n_phi;
    size_t o = (size_t)ir * n_phi + j;
    for (int c = 0; c < 3; ++c)
        for (int n = 0; n < n_orders; ++n, o += plane) {
            const double *wc = w + (size_t)(c * n_orders + n) * n_q;
            const int qn = 2 * n + 2;
            double acc = wc[0] * g[0];
            for (int q = 1; q < qn; ++q) acc = fma(wc[q], g[q], acc);
            put<T>(out, o, acc);
        }
}

// Tracking layout [theta][r][order 1.., (Bz, Bfi, Br)]: one block = a tile
// of 64 azimuths x TR_RADII<T, NO> consecutive radii.  Each wave takes one
// radius at a time with its lanes on 64 consecutive azimuths, so the
// radius's weights are wave-uniform (scalar loads, once per 64 values) and
// the flutter-table reads are coalesced; the values go to an LDS tile
// [azimuth][radius][slot] (rows padded by one element against bank
// conflicts), which the block stores as 64 contiguous runs of
// TR_RADII * 3 (n_orders - 1) values.
constexpr int TR_AZ = 64;

template <typename T, int NO>
struct tr_tile {
    static constexpr int m = 3 * (NO - 1);
    // radii per wave: the tile stays within 48 KB of LDS
    static constexpr int rpw_fit = (48 * 1024) / (TR_AZ * 4 * m * (int)sizeof(T) + TR_AZ * (int)sizeof(T));
    static constexpr int rpw = rpw_fit < 1 ? 1 : (rpw_fit > 8 ? 8 : rpw_fit);
    static constexpr int radii = 4 * rpw;                    // 4 waves per block
    static constexpr int row = radii * m + 1;                // padded LDS row (one azimuth)
    static constexpr size_t lds_bytes = (size_t)TR_AZ * row * sizeof(T);
};

template <typename T, int NO>
__global__ __launch_bounds__(BLOCK) void k_taylor_fill_tracking(const double *__restrict__ table, int n_phi, int nr,
                                                                const double *__restrict__ weights,
                                                                const int32_t *__restrict__ shift,
                                                                void *__restrict__ out) {
    using tile = tr_tile<T, NO>;
    constexpr int n_orders = NO, n_q = 2 * NO, m = tile::m, RT = tile::radii, ROW = tile::row;
    extern __shared__ unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const int tiles_r = (nr + RT - 1) / RT;
    const int tj = (int)(blockIdx.x / (unsigned)tiles_r);
    const int ir0 = (int)(blockIdx.x - (unsigned)tj * tiles_r) * RT;
    const int j0 = tj * TR_AZ;
    const int lane = (int)(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int j = j0 + lane;
    constexpr int slot[3] = {0, 2, 1};                 // stacks component -> (Bz, Bfi, Br) slot
    for (int k = 0; k < tile::rpw; ++k) {
        const int rl = wave + 4 * k;                   // radius within the tile (wave-uniform)
        const int ir = ir0 + rl;
        if (ir >= nr) break;
        int jj = (j < n_phi ? j : 0) - wrap_shift(shift[ir], n_phi);
        if (jj < 0) jj += n_phi;
        double g[n_q];
#pragma unroll
        for (int q = 0; q < n_q; ++q) g[q] = table[(size_t)q * n_phi + jj];
        const double *w = weights + (size_t)ir * 3 * n_orders * n_q;
        T *dst = lds + (size_t)lane * ROW + rl * m;
#pragma unroll
        for (int n = 1; n < n_orders; ++n)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double *wc = w + (size_t)(c * n_orders + n) * n_q;
                double acc = wc[0] * g[0];
#pragma unroll
                for (int q = 1; q < 2 * n + 2; ++q) acc = fma(wc[q], g[q], acc);
                dst[(n - 1) * 3 + slot[c]] = (T)acc;
            }
    }
    __syncthreads();
    const int rows_r = min(RT, nr - ir0), rows_a = min(TR_AZ, n_phi - j0);
    const int run = rows_r * m;                        // contiguous values per azimuth
    T *o = static_cast<T *>(out);
    for (int e = threadIdx.x; e < rows_a * run; e += BLOCK) {
        const int a = e / run, kk = e - a * run;
        o[((size_t)(j0 + a) * nr + ir0) * m + kk] = lds[(size_t)a * ROW + kk];
    }
}

template <typename T, int NO>
void launch_fill(ffa_ctx_s *c, const ffa_taylor_gen *g) {
    const int n_phi = (int)g->n_phi, nr = (int)g->nr;
    hipLaunchKernelGGL((k_flutter_table<2 * NO>), dim3((unsigned)((n_phi + FLUTTER_BLOCK - 1) / FLUTTER_BLOCK)),
                       dim3(FLUTTER_BLOCK), 0, c->stream, g->theta, g->n_phi, g->poles, g->n_poles, g->sig_poly,
                       g->flutter);
    if (g->layout == 0) {
        const dim3 grid((unsigned)((n_phi + BLOCK - 1) / BLOCK), (unsigned)nr);
        hipLaunchKernelGGL((k_taylor_fill_stacks<T, NO>), grid, dim3(BLOCK), 0, c->stream, g->flutter, n_phi, nr,
                           g->weights, g->shift, g->out);
    } else if constexpr (NO >= 2) {
        using tile = tr_tile<T, NO>;
        const unsigned tiles = (unsigned)((nr + tile::radii - 1) / tile::radii) *
                               (unsigned)((n_phi + TR_AZ - 1) / TR_AZ);
        hipLaunchKernelGGL((k_taylor_fill_tracking<T, NO>), dim3(tiles), dim3(BLOCK), tile::lds_bytes, c->stream,
                           g->flutter, n_phi, nr, g->weights, g->shift, g->out);
    }
}

template <typename T>
void launch_fill_t(ffa_ctx_s *c, const ffa_taylor_gen *g) {
    switch (g->n_orders) {
        case 1: launch_fill<T, 1>(c, g); break;
        case 2: launch_fill<T, 2>(c, g); break;
        case 3: launch_fill<T, 3>(c, g); break;
        case 4: launch_fill<T, 4>(c, g); break;
        case 5: launch_fill<T, 5>(c, g); break;
        case 6: launch_fill<T, 6>(c, g); break;
        case 7: launch_fill<T, 7>(c, g); break;
        default: launch_fill<T, 8>(c, g); break;
    }
}

}  // namespace

static_assert(FFA_TAYLOR_MAX_ORDERS == 8, "the fill dispatch covers 1..8 orders");

hipError_t launch_taylor_maps(ffa_ctx_s *c, const ffa_taylor_gen *g) {
    if (g->out_f32) launch_fill_t<float>(c, g);
    else launch_fill_t<double>(c, g);
    return hipGetLastError();
}
