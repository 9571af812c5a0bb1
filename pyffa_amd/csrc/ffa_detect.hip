// Azimuth-crossing detection for the position detectors (gfx950).
//
// detect_azimuth_crossing (src/FFAG_MathTools.py:1633-1695): a particle
// alive before the step crosses the target azimuth when, with both azimuths
// reduced to [0, 2 pi) by x - 2 pi floor(x / 2 pi) (fast_mod_parallel,
// src/FFAG_MathTools.py:1553-1557),
//   post >= pre : pre <= target <= post
//   post <  pre : target >= pre  or  target <= post     (the step wrapped)
// The crossing particles' positions are compacted into idx_out with one
// atomic counter (their order is arbitrary; the caller orders them), so only
// the crossing rows ever leave the device.
#include "ffa_internal.h"

namespace {

constexpr int BLOCK = 256;

__global__ __launch_bounds__(BLOCK) void k_detect_crossing(const double *__restrict__ fi_pre,
                                                          const uint8_t *__restrict__ flags_pre,
                                                          const double *__restrict__ fi_post, int64_t n,
                                                          double target, int64_t *__restrict__ idx_out,
                                                          unsigned long long *__restrict__ count) {
    const int64_t i = blockIdx.x * (int64_t)BLOCK + threadIdx.x;
    if (i >= n) return;
    if (!(flags_pre[i] & FFA_F_ALIVE)) return;
    const double P = FFA_TWO_PI;
    const double a = fi_pre[i], b = fi_post[i];
    const double pre = a - P * floor(a / P), post = b - P * floor(b / P);
    bool hit;
    if (post >= pre) hit = pre <= target && target <= post;
    else hit = (target >= pre && target < P) || (target >= 0.0 && target <= post);
    if (hit) idx_out[atomicAdd(count, 1ull)] = i;
}

__global__ void k_zero_count(unsigned long long *count) {
    if (threadIdx.x == 0) count[0] = 0ull;      // a vector store
}

}  // namespace

hipError_t launch_detect_crossing(ffa_ctx_s *c, const double *fi_pre, const uint8_t *flags_pre, const double *fi_post,
                                  int64_t n, double target, int64_t *idx_out, int64_t *count) {
    // the counter is cleared by a kernel, not a memset, so a captured graph replays it correctly
    hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, c->stream, reinterpret_cast<unsigned long long *>(count));
    if (n == 0) return hipGetLastError();
    hipLaunchKernelGGL(k_detect_crossing, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, c->stream,
                       fi_pre, flags_pre, fi_post, n, target, idx_out, reinterpret_cast<unsigned long long *>(count));
    return hipGetLastError();
}
