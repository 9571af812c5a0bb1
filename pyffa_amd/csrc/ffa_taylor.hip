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

__global__ __launch_bounds__(BLOCK) void k_flutter_table(const double *__restrict__ theta, int64_t n_phi,
                                                        const double *__restrict__ poles, int n_poles,
                                                        const double *__restrict__ sig_poly, int n_q,
                                                        double *__restrict__ table) {
    const int64_t j = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (j >= n_phi) return;
    const double x = theta[j];
    const int len = n_q + 1;
    double acc[MAX_Q];
    for (int q = 0; q < n_q; ++q) acc[q] = 0.0;
    for (int p = 0; p < n_poles; ++p) {
        const double x0 = poles[4 * p], x1 = poles[4 * p + 1], k = poles[4 * p + 2], h = poles[4 * p + 3];
        const double s1 = 1.0 / (1.0 + exp(-k * (x - x0)));
        const double s2 = 1.0 / (1.0 + exp(k * (x - x1)));
        const double *rise = sig_poly + (size_t)(2 * p) * n_q * len;
        const double *fall = rise + (size_t)n_q * len;
        double d1[MAX_Q], d2[MAX_Q];
        for (int i = 0; i < n_q; ++i) {
            d1[i] = horner(rise + (size_t)i * len, len, s1);
            d2[i] = horner(fall + (size_t)i * len, len, s2);
        }
        acc[0] += h * s1 * s2;
        for (int q = 1; q < n_q; ++q) {
            double t = 0.0, binom = 1.0;                  // C(q, i), exact in double for q < 2^53
            for (int i = 0; i <= q; ++i) {
                t += binom * d1[i] * d2[q - i];
                binom = binom * (double)(q - i) / (double)(i + 1);
            }
            acc[q] += h * t;
        }
    }
    for (int q = 0; q < n_q; ++q) table[(size_t)q * n_phi + j] = acc[q];
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

// Tracking layout [theta][r][order 1.., (Bz, Bfi, Br)]: one block = one
// azimuth and BLOCK consecutive radii, i.e. one contiguous run of
// BLOCK * 3 (n_orders - 1) values.  Each thread computes its radius's values
// into LDS; the block then stores the run with consecutive lanes on
// consecutive elements.
template <typename T, int NO>
__global__ __launch_bounds__(BLOCK) void k_taylor_fill_tracking(const double *__restrict__ table, int n_phi, int nr,
                                                                const double *__restrict__ weights,
                                                                const int32_t *__restrict__ shift,
                                                                void *__restrict__ out) {
    constexpr int n_orders = NO, n_q = 2 * NO;
    extern __shared__ unsigned char lds_raw[];
    T *lds = reinterpret_cast<T *>(lds_raw);
    const unsigned bpr = (unsigned)(nr + BLOCK - 1) / BLOCK;        // blocks per azimuth (uniform)
    const int j = (int)(blockIdx.x / bpr);
    const int ir0 = (int)(blockIdx.x - (unsigned)j * bpr) * BLOCK;
    const int ir = ir0 + (int)threadIdx.x;
    constexpr int m = 3 * (n_orders - 1);
    const int rows = min(BLOCK, nr - ir0);
    if (ir < nr) {
        int jj = j - wrap_shift(shift[ir], n_phi);
        if (jj < 0) jj += n_phi;
        double g[MAX_Q];
        for (int q = 0; q < n_q; ++q) g[q] = table[(size_t)q * n_phi + jj];
        const double *w = weights + (size_t)ir * 3 * n_orders * n_q;
        constexpr int slot[3] = {0, 2, 1};             // stacks component -> (Bz, Bfi, Br) slot
        for (int n = 1; n < n_orders; ++n)
            for (int c = 0; c < 3; ++c) {
                const double *wc = w + (size_t)(c * n_orders + n) * n_q;
                const int qn = 2 * n + 2;
                double acc = wc[0] * g[0];
                for (int q = 1; q < qn; ++q) acc = fma(wc[q], g[q], acc);
                lds[threadIdx.x * m + (n - 1) * 3 + slot[c]] = (T)acc;
            }
    }
    __syncthreads();
    T *dst = static_cast<T *>(out) + ((size_t)j * nr + ir0) * m;
    for (int k = threadIdx.x; k < rows * m; k += BLOCK) dst[k] = lds[k];
}

template <typename T, int NO>
void launch_fill(ffa_ctx_s *c, const ffa_taylor_gen *g) {
    const int n_phi = (int)g->n_phi, nr = (int)g->nr;
    if (g->layout == 0) {
        const dim3 grid((unsigned)((n_phi + BLOCK - 1) / BLOCK), (unsigned)nr);
        hipLaunchKernelGGL((k_taylor_fill_stacks<T, NO>), grid, dim3(BLOCK), 0, c->stream, g->flutter, n_phi, nr,
                           g->weights, g->shift, g->out);
    } else if constexpr (NO >= 2) {
        const unsigned bpr = (unsigned)((nr + BLOCK - 1) / BLOCK);
        hipLaunchKernelGGL((k_taylor_fill_tracking<T, NO>), dim3(bpr * (unsigned)n_phi), dim3(BLOCK),
                           BLOCK * 3 * (NO - 1) * sizeof(T), c->stream, g->flutter, n_phi, nr, g->weights, g->shift,
                           g->out);
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
    const unsigned tb = (unsigned)((g->n_phi + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_flutter_table, dim3(tb), dim3(BLOCK), 0, c->stream, g->theta, g->n_phi, g->poles,
                       g->n_poles, g->sig_poly, g->n_q, g->flutter);
    if (g->out_f32) launch_fill_t<float>(c, g);
    else launch_fill_t<double>(c, g);
    return hipGetLastError();
}
