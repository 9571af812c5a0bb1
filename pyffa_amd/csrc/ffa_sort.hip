// Periodic particle ordering for the FP32 tier (gfx950): a stable sort of the
// injected particles by azimuth bin, then one gather pass over the state.
//
// Particles that sit next to each other in memory then sit next to each
// other in the machine: the waves of the push read neighbouring Taylor-map
// cells (the map is periodic over a sector, so an azimuth order over 2 pi is
// also an order over the sector) and the space-charge gather reads
// neighbouring slices.  The reference keeps particles in their input order;
// the order is internal here (the host mirror is rebuilt in the input order
// from the carried particle ids).
//
// Sort: key = floor((fi mod 2 pi) * nbins / 2 pi) for live particles, nbins
// for the others (they go last), nbins <= 65535.  Two LSD passes over 8-bit
// digits; each pass is a per-tile digit histogram, one exclusive scan over
// (digit, tile), and a scatter whose in-tile ranks come from wave ballots in
// element order, so the result is stable and does not depend on timing.
#include "ffa_internal.h"

namespace {

constexpr int SBLOCK = 256;
constexpr int PER_THREAD = 16;
constexpr int TILE = SBLOCK * PER_THREAD;     // elements per tile (4096)
constexpr int WAVES = SBLOCK / 64;
constexpr int PER_WAVE = TILE / WAVES;        // 1024 = 16 chunks of 64

__global__ __launch_bounds__(SBLOCK) void k_sort_keys(const double *__restrict__ fi, const uint8_t *__restrict__ flags,
                                                     int64_t count, int nbins, uint16_t *__restrict__ keys,
                                                     uint32_t *__restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)SBLOCK + threadIdx.x;
    if (i >= count) return;
    const uint8_t f = flags[i];
    int key = nbins;
    if ((f & FFA_F_INJ) && (f & FFA_F_ALIVE)) {
        const double v = fi[i] * (1.0 / FFA_TWO_PI);
        const double u = v - floor(v);                   // [0, 1); NaN fails both tests below
        if (u >= 0.0 && u < 1.0) {
            key = (int)(u * (double)nbins);
            if (key >= nbins) key = nbins - 1;
        }
    }
    keys[i] = (uint16_t)key;
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(SBLOCK) void k_sort_hist(const uint16_t *__restrict__ keys, int64_t count, int shift,
                                                     int ntiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * TILE;
    for (int k = 0; k < PER_THREAD; ++k) {
        const int64_t e = base + k * SBLOCK + threadIdx.x;
        if (e < count) atomicAdd(&h[(keys[e] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];      // digit-major
}

// Exclusive scan of the 256 * ntiles counters (digit-major) in three
// coalesced passes: segment sums, a scan of the segment sums, and a scan of
// each segment from its carry.  A segment is one tile's worth (4096).
constexpr int SCAN_SEG = TILE;

// exclusive scan across the block; *total receives the block's sum
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 1; w < SBLOCK; w <<= 1) {
        const uint32_t a = (int)threadIdx.x >= w ? sh[threadIdx.x - w] : 0u;
        __syncthreads();
        sh[threadIdx.x] += a;
        __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    *total = sh[SBLOCK - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SBLOCK) void k_scan_sums(const uint32_t *__restrict__ hist, int64_t total,
                                                     uint32_t *__restrict__ seg_sum) {
    __shared__ uint32_t sh[SBLOCK];
    const int64_t base = (int64_t)blockIdx.x * SCAN_SEG;
    uint32_t v = 0;
    for (int k = 0; k < PER_THREAD; ++k) {
        const int64_t e = base + k * SBLOCK + threadIdx.x;
        if (e < total) v += hist[e];
    }
    uint32_t tot;
    block_excl_scan(v, sh, &tot);
    if (threadIdx.x == 0) seg_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SBLOCK) void k_scan_top(uint32_t *__restrict__ seg_sum, int nseg) {
    __shared__ uint32_t sh[SBLOCK];
    uint32_t carry = 0;
    for (int b = 0; b < nseg; b += SBLOCK) {            // uniform trip count
        const int k = b + (int)threadIdx.x;
        const uint32_t v = k < nseg ? seg_sum[k] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, sh, &tot);
        if (k < nseg) seg_sum[k] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(SBLOCK) void k_scan_apply(uint32_t *__restrict__ hist, int64_t total,
                                                      const uint32_t *__restrict__ seg_off) {
    __shared__ uint32_t sh[SBLOCK];
    // thread t owns the PER_THREAD consecutive counters at t * PER_THREAD
    const int64_t base = (int64_t)blockIdx.x * SCAN_SEG + (int64_t)threadIdx.x * PER_THREAD;
    uint32_t v[PER_THREAD];
    uint32_t s = 0;
    for (int k = 0; k < PER_THREAD; ++k) {
        v[k] = base + k < total ? hist[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t acc = seg_off[blockIdx.x] + block_excl_scan(s, sh, &tot);
    for (int k = 0; k < PER_THREAD; ++k) {
        if (base + k < total) hist[base + k] = acc;
        acc += v[k];
    }
}

__global__ __launch_bounds__(SBLOCK) void k_sort_scatter(const uint16_t *__restrict__ kin,
                                                        const uint32_t *__restrict__ iin, int64_t count, int shift,
                                                        int ntiles, const uint32_t *__restrict__ offs,
                                                        uint16_t *__restrict__ kout, uint32_t *__restrict__ iout) {
    __shared__ uint32_t cnt[WAVES][256];
    for (int k = threadIdx.x; k < WAVES * 256; k += SBLOCK) (&cnt[0][0])[k] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * TILE + (int64_t)w * PER_WAVE;
    uint16_t key[PER_THREAD];
    uint32_t id[PER_THREAD], rank[PER_THREAD];
    // chunks in element order; ranks count the earlier equal digits of this wave
    for (int ch = 0; ch < PER_THREAD; ++ch) {
        const int64_t e = base + ch * 64 + lane;
        const bool valid = e < count;
        key[ch] = valid ? kin[e] : (uint16_t)0;
        id[ch] = valid ? iin[e] : 0u;
        const int d = (key[ch] >> shift) & 255;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const uint32_t before = (uint32_t)__popcll(peers & below);
        const uint32_t base_cnt = cnt[w][d];
        rank[ch] = base_cnt + before;
        // the last lane of each group of equal digits advances the counter
        if (valid && (peers >> lane) == 1ull) cnt[w][d] = base_cnt + before + 1;
    }
    __syncthreads();
    // per-digit offsets of the waves inside the tile
    if (threadIdx.x < 256) {
        uint32_t acc = 0;
        for (int q = 0; q < WAVES; ++q) {
            const uint32_t v = cnt[q][threadIdx.x];
            cnt[q][threadIdx.x] = acc;
            acc += v;
        }
    }
    __syncthreads();
    for (int ch = 0; ch < PER_THREAD; ++ch) {
        const int64_t e = base + ch * 64 + lane;
        if (e >= count) continue;
        const int d = (key[ch] >> shift) & 255;
        const uint32_t pos = offs[d * ntiles + blockIdx.x] + cnt[w][d] + rank[ch];
        kout[pos] = key[ch];
        iout[pos] = id[ch];
    }
}

template <typename ET>
__global__ __launch_bounds__(SBLOCK) void k_gather_particles(ffa_particles s, ffa_particles d,
                                                            const uint32_t *__restrict__ order, int64_t count,
                                                            const int64_t *__restrict__ pid_s,
                                                            int64_t *__restrict__ pid_d) {
    const int64_t k = blockIdx.x * (int64_t)SBLOCK + threadIdx.x;
    if (k >= s.n) return;
    const int64_t j = k < count ? (int64_t)order[k] : k;
    const int64_t n = s.n;
    // every field is loaded before any is stored (the compiler cannot tell
    // the source arrays from the destination ones, so interleaved loads and
    // stores would wait one by one); measured the same at 8M, where the
    // scattered reads, not the per-thread chain, bound the kernel
    const double r = s.r[j], vr = s.vr[j], z = s.z[j], vz = s.vz[j], fi = s.fi[j], fd = s.fidot[j];
    const double rf = s.rf_phase[j];
    const uint8_t fl = s.flags[j];
    const ET *ep = reinterpret_cast<const ET *>(s.esc_pre), *epp = reinterpret_cast<const ET *>(s.esc_prepre);
    ET e[3], f[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e[c] = ep[c * n + j];
        f[c] = epp[c * n + j];
    }
    const int64_t id = pid_s[j];
    d.r[k] = r;
    d.vr[k] = vr;
    d.z[k] = z;
    d.vz[k] = vz;
    d.fi[k] = fi;
    d.fidot[k] = fd;
    d.rf_phase[k] = rf;
    d.flags[k] = fl;
    ET *dp = reinterpret_cast<ET *>(d.esc_pre), *dpp = reinterpret_cast<ET *>(d.esc_prepre);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dp[c * n + k] = e[c];
        dpp[c * n + k] = f[c];
    }
    pid_d[k] = id;
}

}  // namespace

int64_t sort_tiles(int64_t count) { return (count + TILE - 1) / TILE; }

size_t sort_workspace_bytes(int64_t count) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const int64_t nt = sort_tiles(count);
    const int64_t nseg = (nt * 256 + SCAN_SEG - 1) / SCAN_SEG;
    return 2 * al((size_t)count * 2) + 2 * al((size_t)count * 4) + al((size_t)nt * 256 * 4) + al((size_t)nseg * 4);
}

hipError_t launch_sort_order(ffa_ctx_s *c, const ffa_particles *p, int64_t count, int nbins, void *ws,
                             const uint32_t **order_out) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char *w = static_cast<char *>(ws);
    const int64_t nt = sort_tiles(count);
    uint16_t *k0 = reinterpret_cast<uint16_t *>(w);
    uint16_t *k1 = reinterpret_cast<uint16_t *>(w + al((size_t)count * 2));
    uint32_t *i0 = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2));
    uint32_t *i1 = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2) + al((size_t)count * 4));
    uint32_t *hist = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2) + 2 * al((size_t)count * 4));
    uint32_t *seg = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + al((size_t)nt * 256 * 4));
    const int64_t total = nt * 256, nseg = (total + SCAN_SEG - 1) / SCAN_SEG;
    *order_out = i0;
    if (count == 0) return hipSuccess;
    const unsigned g = (unsigned)((count + SBLOCK - 1) / SBLOCK);
    hipLaunchKernelGGL(k_sort_keys, dim3(g), dim3(SBLOCK), 0, c->stream, p->fi, p->flags, count, nbins, k0, i0);
    for (int pass = 0; pass < 2; ++pass) {
        const int shift = 8 * pass;
        uint16_t *ki = pass == 0 ? k0 : k1, *ko = pass == 0 ? k1 : k0;
        uint32_t *ii = pass == 0 ? i0 : i1, *io = pass == 0 ? i1 : i0;
        hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)nt), dim3(SBLOCK), 0, c->stream, ki, count, shift, (int)nt,
                           hist);
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nseg), dim3(SBLOCK), 0, c->stream, hist, total, seg);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SBLOCK), 0, c->stream, seg, (int)nseg);
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nseg), dim3(SBLOCK), 0, c->stream, hist, total, seg);
        hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)nt), dim3(SBLOCK), 0, c->stream, ki, ii, count, shift,
                           (int)nt, hist, ko, io);
    }
    return hipGetLastError();
}

uint16_t *sort_key_buffer(void *ws) { return static_cast<uint16_t *>(ws); }

namespace {
__global__ __launch_bounds__(SBLOCK) void k_sort_iota(int64_t count, uint32_t *__restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)SBLOCK + threadIdx.x;
    if (i < count) idx[i] = (uint32_t)i;
}
}  // namespace

// Stable sort of caller-written 16-bit keys (in sort_key_buffer(ws), count of
// them): one 8-bit pass when every key is below 256 (max_key), else two.
// *order_out: the input index of each sorted position; *keys_out: the
// sorted keys.
hipError_t launch_sort_keys16(ffa_ctx_s *c, int64_t count, int max_key, void *ws, const uint32_t **order_out,
                              const uint16_t **keys_out) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char *w = static_cast<char *>(ws);
    const int64_t nt = sort_tiles(count);
    uint16_t *k0 = reinterpret_cast<uint16_t *>(w);
    uint16_t *k1 = reinterpret_cast<uint16_t *>(w + al((size_t)count * 2));
    uint32_t *i0 = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2));
    uint32_t *i1 = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2) + al((size_t)count * 4));
    uint32_t *hist = reinterpret_cast<uint32_t *>(w + 2 * al((size_t)count * 2) + 2 * al((size_t)count * 4));
    uint32_t *seg = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(hist) + al((size_t)nt * 256 * 4));
    const int64_t total = nt * 256, nseg = (total + SCAN_SEG - 1) / SCAN_SEG;
    const int passes = max_key < 256 ? 1 : 2;
    *order_out = passes == 1 ? i1 : i0;
    *keys_out = passes == 1 ? k1 : k0;
    if (count == 0) return hipSuccess;
    const unsigned g = (unsigned)((count + SBLOCK - 1) / SBLOCK);
    hipLaunchKernelGGL(k_sort_iota, dim3(g), dim3(SBLOCK), 0, c->stream, count, i0);
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = 8 * pass;
        uint16_t *ki = pass == 0 ? k0 : k1, *ko = pass == 0 ? k1 : k0;
        uint32_t *ii = pass == 0 ? i0 : i1, *io = pass == 0 ? i1 : i0;
        hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)nt), dim3(SBLOCK), 0, c->stream, ki, count, shift, (int)nt,
                           hist);
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nseg), dim3(SBLOCK), 0, c->stream, hist, total, seg);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SBLOCK), 0, c->stream, seg, (int)nseg);
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nseg), dim3(SBLOCK), 0, c->stream, hist, total, seg);
        hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)nt), dim3(SBLOCK), 0, c->stream, ki, ii, count, shift,
                           (int)nt, hist, ko, io);
    }
    return hipGetLastError();
}

hipError_t launch_gather_particles(ffa_ctx_s *c, const ffa_particles *src, const ffa_particles *dst,
                                   const uint32_t *order, int64_t count, const int64_t *pid_src, int64_t *pid_dst) {
    if (src->n == 0) return hipSuccess;
    const unsigned g = (unsigned)((src->n + SBLOCK - 1) / SBLOCK);
    if (src->esc_f32)
        hipLaunchKernelGGL(k_gather_particles<float>, dim3(g), dim3(SBLOCK), 0, c->stream, *src, *dst, order, count,
                           pid_src, pid_dst);
    else
        hipLaunchKernelGGL(k_gather_particles<double>, dim3(g), dim3(SBLOCK), 0, c->stream, *src, *dst, order, count,
                           pid_src, pid_dst);
    return hipGetLastError();
}
