// C-ABI entry points of libffa_hip.so (declared in include/ffa_hip.h).
// Argument validation happens here, on the host, before any launch; C++
// exceptions never cross the boundary.
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>
#include "ffa_internal.h"

// Workspace of ffa_sc_solve_noalloc: one array per GPU, allocated when the
// code object loads and held until the process ends.  Host code never uses
// this name as an address: each device's address comes from
// hipGetSymbolAddress, taken once per device in ffa_create.
static __device__ __align__(256) char scratch[FFA_SC_NOALLOC_BYTES];

static void flutter_table(const ffa_machine &m, std::vector<float> &t);

int ffa_sc_layout_compute(int64_t n, int32_t ns, int32_t nz, int32_t nx, ffa_sc_layout *L);

namespace {

thread_local std::string g_err;

int fail(ffa_ctx_s *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    g_err = buf;
    return code;
}

int hip_check(ffa_ctx_s *c, hipError_t e, const char *what) {
    if (e == hipSuccess) return FFA_OK;
    return fail(c, FFA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// transverse SC grid sizes the LDS FFT handles: 32..1024 points whose only
// prime factors are 2, 3 and 5
bool fft_size_ok(int v) {
    if (v < 32 || v > 1024) return false;
    for (int f : {2, 3, 5})
        while (v % f == 0) v /= f;
    return v == 1;
}

int check_particles(ffa_ctx_s *c, const ffa_particles *p) {
    if (!p) return fail(c, FFA_ERR_INVALID, "particles is NULL");
    if (p->n < 0) return fail(c, FFA_ERR_INVALID, "negative particle count");
    if (p->n == 0) return FFA_OK;
    if (!p->r || !p->vr || !p->z || !p->vz || !p->fi || !p->fidot || !p->rf_phase || !p->flags || !p->esc_pre ||
        !p->esc_prepre)
        return fail(c, FFA_ERR_INVALID, "a particle array pointer is NULL");
    return FFA_OK;
}

ffa_ctx_s *as_ctx(void *v) { return reinterpret_cast<ffa_ctx_s *>(v); }

// Device address of `scratch` on each GPU (looked up once per device, with
// that device current).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[MAX_DEVICES] = {};

hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

int ffa_create(int device, void *stream, void **ctx_out) {
    if (!ctx_out) return fail(nullptr, FFA_ERR_INVALID, "ctx_out is NULL");
    *ctx_out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return fail(nullptr, FFA_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(nullptr, FFA_ERR_INVALID, "device %d out of range (%d)", device, ndev);
    ffa_ctx_s *c = new (std::nothrow) ffa_ctx_s();
    if (!c) return fail(nullptr, FFA_ERR_HIP, "out of host memory");
    c->device = device;
    c->stream = reinterpret_cast<hipStream_t>(stream);
    if (device >= MAX_DEVICES) {
        delete c;
        return fail(nullptr, FFA_ERR_INVALID, "device %d beyond the %d this library handles", device, MAX_DEVICES);
    }
    e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&c->dev), FFA_CTX_DEV_BYTES);
    if (e == hipSuccess) e = scratch_address(device, &c->scratch);
    if (e != hipSuccess) {
        int rc = fail(nullptr, FFA_ERR_HIP, "context set-up: %s", hipGetErrorString(e));
        if (c->dev) (void)hipFree(c->dev);
        delete c;
        return rc;
    }
    c->dev_bytes = FFA_CTX_DEV_BYTES;
    *ctx_out = c;
    return FFA_OK;
}

int ffa_destroy(void *ctx) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return FFA_OK;
    (void)hipSetDevice(c->device);
    if (c->dev) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(c->dev);
    }
    delete c;
    return FFA_OK;
}

const char *ffa_last_error(void *ctx) {
    ffa_ctx_s *c = as_ctx(ctx);
    return c ? c->err.c_str() : g_err.c_str();
}

int ffa_set_stream(void *ctx, void *stream) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(nullptr, FFA_ERR_INVALID, "ctx is NULL");
    hipStream_t ns = reinterpret_cast<hipStream_t>(stream);
    if (ns != c->stream) {
        // work already queued on the old stream may still read the context's
        // device block (machine, step tables): the new stream waits for it
        hipEvent_t ev;
        hipError_t e = hipSetDevice(c->device);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return hip_check(c, e, "ffa_set_stream");
        e = hipEventRecord(ev, c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ns, ev, 0);
        const hipError_t e2 = hipEventDestroy(ev);
        if (e == hipSuccess) e = e2;
        if (e != hipSuccess) return hip_check(c, e, "ffa_set_stream");
    }
    c->stream = ns;
    return FFA_OK;
}

int ffa_set_machine(void *ctx, const ffa_machine *m) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !m) return fail(c, FFA_ERR_INVALID, "ctx or machine is NULL");
    if (m->nsectors <= 0) return fail(c, FFA_ERR_INVALID, "nsectors must be positive");
    if (m->n_seg < 0 || m->n_seg > FFA_MAX_SEG) return fail(c, FFA_ERR_INVALID, "n_seg out of range");
    if (m->n_spoly < 0 || m->n_spoly > FFA_MAX_POLY || m->n_sppoly < 0 || m->n_sppoly > FFA_MAX_POLY)
        return fail(c, FFA_ERR_INVALID, "spiral polynomial degree out of range");
    if (m->n_harm < 0 || m->n_harm > FFA_MAX_HARM) return fail(c, FFA_ERR_INVALID, "n_harm out of range");
    if (m->n_gap < 0 || m->n_gap > FFA_MAX_GAP) return fail(c, FFA_ERR_INVALID, "n_gap out of range");
    if (m->gap_model != 0 && m->gap_model != 1) return fail(c, FFA_ERR_INVALID, "gap_model must be 0 or 1");
    if (m->gap_model == 1 && m->rf_enable &&
        (!m->emap || m->emap_nr < 2 || m->emap_nphi < 2 || !(m->emap_dr > 0.0) || !(m->emap_dphi > 0.0)))
        return fail(c, FFA_ERR_INVALID, "spiral gaps need a map of >= 2 x 2 nodes with positive steps");
    if (m->max_order < 0) return fail(c, FFA_ERR_INVALID, "max_order negative");
    if (m->taylor_f32 < 0 || m->taylor_f32 > 1) return fail(c, FFA_ERR_INVALID, "taylor_f32 must be 0 or 1");
    if (m->max_order > 0) {
        if (!m->taylor) return fail(c, FFA_ERR_INVALID, "max_order > 0 but no Taylor map");
        if (m->taylor_f32 && (double)m->r_size * m->fi_size * 3.0 * m->max_order >= 4294967296.0)
            return fail(c, FFA_ERR_INVALID, "FP32 Taylor map exceeds 2^32 elements (32-bit offsets)");
        if (m->r_size < 2 || m->fi_size < 2) return fail(c, FFA_ERR_INVALID, "Taylor map needs >= 2 nodes per axis");
        if (!(m->r_step > 0.0) || !(m->fi_step > 0.0)) return fail(c, FFA_ERR_INVALID, "map steps must be positive");
    }
    if (m->rf_enable && m->n_gap > 0 && !(m->acc_etot == m->acc_etot))
        return fail(c, FFA_ERR_INVALID, "acc_etot is NaN");
    static_assert(sizeof(ffa_machine) <= FFA_CTX_FCONST_OFF, "machine block overflows");
    static_assert(FFA_CTX_FCONST_OFF + sizeof(ffa_fconst) <= FFA_CTX_STEPS_OFF, "fconst block overflows");
    ffa_fconst fc;
    ffa_fconst_from_machine(*m, fc);
    std::vector<float> flut;
    flutter_table(*m, flut);
    fc.flut = reinterpret_cast<const float *>(c->dev + FFA_CTX_FLUT_OFF);
    fc.flut_scale_d = (double)FFA_FLUT_N * (double)m->nsectors / FFA_TWO_PI;
    fc.flut_scale = (float)fc.flut_scale_d;
    hipError_t e = hipSetDevice(c->device);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dev + FFA_CTX_MACHINE_OFF, m, sizeof(ffa_machine), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dev + FFA_CTX_FCONST_OFF, &fc, sizeof(ffa_fconst), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dev + FFA_CTX_FLUT_OFF, flut.data(), flut.size() * sizeof(float), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_check(c, e, "ffa_set_machine");
    c->host_machine = *m;
    c->have_machine = true;
    return FFA_OK;
}

int ffa_inject(void *ctx, const ffa_particles *p, int64_t begin, int64_t end) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(nullptr, FFA_ERR_INVALID, "ctx is NULL");
    if (int rc = check_particles(c, p)) return rc;
    if (begin < 0 || end > p->n || begin > end)
        return fail(c, FFA_ERR_INVALID, "inject range [%lld, %lld) outside [0, %lld)", (long long)begin,
                    (long long)end, (long long)p->n);
    return hip_check(c, launch_inject(c, p, begin, end), "ffa_inject");
}

int ffa_push_boris_f64(void *ctx, const ffa_particles *p, const ffa_step *s) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !s) return fail(c, FFA_ERR_INVALID, "ctx or step is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP64 push needs a double Taylor map");
    if (int rc = check_particles(c, p)) return rc;
    if (p->esc_f32) return fail(c, FFA_ERR_INVALID, "the FP64 push needs double SC endpoints");
    return hip_check(c, launch_boris_f64(c, p, s), "ffa_push_boris_f64");
}

int ffa_push_rk4_f64(void *ctx, const ffa_particles *p, const ffa_step *s) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !s) return fail(c, FFA_ERR_INVALID, "ctx or step is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP64 push needs a double Taylor map");
    if (int rc = check_particles(c, p)) return rc;
    if (p->esc_f32) return fail(c, FFA_ERR_INVALID, "the FP64 push needs double SC endpoints");
    return hip_check(c, launch_rk4_f64(c, p, s), "ffa_push_rk4_f64");
}

int ffa_push_f64_multi(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                       int32_t nsteps, int32_t rk4) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !steps || !inj) return fail(c, FFA_ERR_INVALID, "ctx, steps or inj is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP64 push needs a double Taylor map");
    if (int rc = check_particles(c, p)) return rc;
    if (p->esc_f32) return fail(c, FFA_ERR_INVALID, "the FP64 push needs double SC endpoints");
    if (nsteps < 0) return fail(c, FFA_ERR_INVALID, "nsteps negative");
    for (int k = 0; k < nsteps; ++k)
        if (inj[2 * k] < 0 || inj[2 * k + 1] > p->n || inj[2 * k] > inj[2 * k + 1])
            return fail(c, FFA_ERR_INVALID, "inject range of step %d invalid", k);
    for (int k = 0; k < nsteps; ++k) {
        hipError_t e = launch_inject(c, p, inj[2 * k], inj[2 * k + 1]);
        if (e == hipSuccess) e = rk4 ? launch_rk4_f64(c, p, &steps[k]) : launch_boris_f64(c, p, &steps[k]);
        if (e != hipSuccess) return hip_check(c, e, "ffa_push_f64_multi");
    }
    return FFA_OK;
}

int ffa_push_boris_f64_fused(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                             int32_t nsteps) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !steps || !inj) return fail(c, FFA_ERR_INVALID, "ctx, steps or inj is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP64 push needs a double Taylor map");
    if (int rc = check_particles(c, p)) return rc;
    if (p->esc_f32) return fail(c, FFA_ERR_INVALID, "the FP64 push needs double SC endpoints");
    if (nsteps < 0 || nsteps > FFA_MAX_FUSED || (size_t)nsteps * 16 > FFA_CTX_INJ_BYTES)
        return fail(c, FFA_ERR_INVALID, "nsteps %d outside [0, %d]", nsteps, (int)(FFA_CTX_INJ_BYTES / 16));
    for (int k = 0; k < nsteps; ++k)
        if (inj[2 * k] < 0 || inj[2 * k + 1] > p->n || inj[2 * k] > inj[2 * k + 1])
            return fail(c, FFA_ERR_INVALID, "inject range of step %d invalid", k);
    if (nsteps == 0) return FFA_OK;
    hipError_t e = hipMemcpyAsync(c->dev + FFA_CTX_STEPS_OFF, steps, sizeof(ffa_step) * nsteps,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dev + FFA_CTX_INJ_OFF, inj, sizeof(int64_t) * 2 * nsteps, hipMemcpyHostToDevice,
                           c->stream);
    if (e != hipSuccess) return hip_check(c, e, "ffa_push_boris_f64_fused (step table)");
    return hip_check(c, launch_boris_f64_multi(c, p, nsteps), "ffa_push_boris_f64_fused");
}

int ffa_push_boris_fused_f32(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                             int32_t nsteps) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !steps || !inj) return fail(c, FFA_ERR_INVALID, "ctx, steps or inj is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && !c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP32 tier needs a float Taylor map");
    if (int rc = check_particles(c, p)) return rc;
    if (nsteps < 0 || nsteps > FFA_MAX_FUSED || (size_t)nsteps * 16 > FFA_CTX_INJ_BYTES)
        return fail(c, FFA_ERR_INVALID, "nsteps %d outside [0, %d]", nsteps, (int)(FFA_CTX_INJ_BYTES / 16));
    for (int k = 0; k < nsteps; ++k)
        if (inj[2 * k] < 0 || inj[2 * k + 1] > p->n || inj[2 * k] > inj[2 * k + 1])
            return fail(c, FFA_ERR_INVALID, "inject range of step %d invalid", k);
    if (nsteps == 0) return FFA_OK;
    hipError_t e = hipMemcpyAsync(c->dev + FFA_CTX_STEPS_OFF, steps, sizeof(ffa_step) * nsteps,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->dev + FFA_CTX_INJ_OFF, inj, sizeof(int64_t) * 2 * nsteps, hipMemcpyHostToDevice,
                           c->stream);
    if (e != hipSuccess) return hip_check(c, e, "ffa_push_boris_fused_f32 (step table)");
    return hip_check(c, launch_boris_fused_f32(c, p, nsteps), "ffa_push_boris_fused_f32");
}

int ffa_sc_workspace_layout(int64_t n, int32_t ns, int32_t nz, int32_t nx, ffa_sc_layout *out) {
    if (!out) return fail(nullptr, FFA_ERR_INVALID, "out is NULL");
    if (n < 0 || ns < 2 || ns > 4096) return fail(nullptr, FFA_ERR_INVALID, "need n >= 0 and 2 <= ns <= 4096");
    if (!fft_size_ok(nx) || !fft_size_ok(nz))
        return fail(nullptr, FFA_ERR_INVALID, "nx, nz must be in [32, 1024] with no prime factor above 5 (got %d, %d)",
                    nx, nz);
    return ffa_sc_layout_compute(n, ns, nz, nx, out);
}

namespace {
// Arguments of a solve on a caller workspace (ffa_sc_solve, ffa_sc_solve_field
// and the push that applies a deferred gather).
int check_sc_solve(ffa_ctx_s *c, const ffa_particles *p, const ffa_sc_params *sp, const void *workspace) {
    if (!c || !sp || !workspace) return fail(c, FFA_ERR_INVALID, "ctx, params or workspace is NULL");
    if (!(sp->c > 0.0)) return fail(c, FFA_ERR_INVALID, "speed of light must be positive");
    if (int rc = check_particles(c, p)) return rc;
    if (sp->ns < 2 || sp->ns > 4096) return fail(c, FFA_ERR_INVALID, "ns must be in [2, 4096]");
    if (!fft_size_ok(sp->nx) || !fft_size_ok(sp->nz))
        return fail(c, FFA_ERR_INVALID, "nx, nz must be in [32, 1024] with no prime factor above 5 (got %d, %d)",
                    sp->nx, sp->nz);
    if (sp->bits < 1 || sp->bits > 20) return fail(c, FFA_ERR_INVALID, "bits must be in [1, 20]");
    if ((double)p->n * (double)(1LL << (2 * sp->bits)) >= 9.2e18)
        return fail(c, FFA_ERR_INVALID, "fixed-point scale 2^%d too large for %lld particles", 2 * sp->bits,
                    (long long)p->n);
    if (!sp->seo_r || !sp->seo_pr || !sp->seo_ek_axis || !sp->seo_perimeter || sp->n_ek < 2 || sp->n_fi < 2)
        return fail(c, FFA_ERR_INVALID, "SEO tables missing or smaller than 2x2");
    return FFA_OK;
}
}  // namespace

int ffa_sc_solve_field(void *ctx, const ffa_particles *p, const ffa_sc_params *sp, void *workspace) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (int rc = check_sc_solve(c, p, sp, workspace)) return rc;
    ffa_sc_layout L;
    ffa_sc_layout_compute(p->n, sp->ns, sp->nz, sp->nx, &L);
    return hip_check(c, launch_sc_solve(c, p, sp, workspace, &L, false), "ffa_sc_solve_field");
}

int ffa_push_boris_fused_f32_sc(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                                int32_t nsteps, const ffa_sc_params *sp, const void *workspace) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !steps || !inj) return fail(c, FFA_ERR_INVALID, "ctx, steps or inj is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (c->host_machine.max_order > 0 && !c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "the FP32 tier needs a float Taylor map");
    if (c->mlp_packed != nullptr)
        return fail(c, FFA_ERR_STATE, "the surrogate push takes its kicks from ffa_sc_solve (surrogate enabled)");
    if (int rc = check_sc_solve(c, p, sp, workspace)) return rc;
    if (nsteps < 0 || nsteps > FFA_MAX_FUSED || (size_t)nsteps * 16 > FFA_CTX_INJ_BYTES)
        return fail(c, FFA_ERR_INVALID, "nsteps %d outside [0, %d]", nsteps, (int)(FFA_CTX_INJ_BYTES / 16));
    for (int k = 0; k < nsteps; ++k)
        if (inj[2 * k] < 0 || inj[2 * k + 1] > p->n || inj[2 * k] > inj[2 * k + 1])
            return fail(c, FFA_ERR_INVALID, "inject range of step %d invalid", k);
    hipError_t e = hipSuccess;
    if (nsteps > 0) {
        e = hipMemcpyAsync(c->dev + FFA_CTX_STEPS_OFF, steps, sizeof(ffa_step) * nsteps, hipMemcpyHostToDevice,
                           c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->dev + FFA_CTX_INJ_OFF, inj, sizeof(int64_t) * 2 * nsteps, hipMemcpyHostToDevice,
                               c->stream);
    }
    if (e != hipSuccess) return hip_check(c, e, "ffa_push_boris_fused_f32_sc (step table)");
    ffa_sc_layout L;
    ffa_sc_layout_compute(p->n, sp->ns, sp->nz, sp->nx, &L);
    sc_gather_src g;
    const bool from_phi = sc_gather_inputs(sp, const_cast<void *>(workspace), &L, &g);
    return hip_check(c, launch_boris_fused_f32_gather(c, p, nsteps, g, from_phi), "ffa_push_boris_fused_f32_sc");
}

int ffa_sc_solve(void *ctx, const ffa_particles *p, const ffa_sc_params *sp, void *workspace) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (int rc = check_sc_solve(c, p, sp, workspace)) return rc;
    ffa_sc_layout L;
    ffa_sc_layout_compute(p->n, sp->ns, sp->nz, sp->nx, &L);
    return hip_check(c, launch_sc_solve(c, p, sp, workspace, &L), "ffa_sc_solve");
}

// ---- stage-level SC entry points (FFAG_SC_new module functions) ----------
namespace {
int check_stage_ws(ffa_ctx_s *c, const void *ws, size_t ws_bytes, int64_t n, int32_t ns, int32_t nz, int32_t nx) {
    if (!ws) return fail(c, FFA_ERR_INVALID, "workspace is NULL");
    const size_t need = sc_stage_workspace_bytes(n, ns, nz, nx);
    if (ws_bytes < need)
        return fail(c, FFA_ERR_INVALID, "workspace of %zu bytes is smaller than the %zu this call needs", ws_bytes, need);
    return FFA_OK;
}
int check_grid(ffa_ctx_s *c, int32_t ns, int32_t nz, int32_t nx) {
    if (ns < 1 || ns > 4096) return fail(c, FFA_ERR_INVALID, "ns must be in [1, 4096]");
    if (!fft_size_ok(nx) || !fft_size_ok(nz))
        return fail(c, FFA_ERR_INVALID, "nx, nz must be in [32, 1024] with no prime factor above 5 (got %d, %d)",
                    nx, nz);
    return FFA_OK;
}
}  // namespace

int ffa_sc_stage_workspace_size(int64_t n, int32_t ns, int32_t nz, int32_t nx, size_t *bytes) {
    if (!bytes) return fail(nullptr, FFA_ERR_INVALID, "bytes is NULL");
    if (n < 0 || n > 0xFFFFFFFFLL) return fail(nullptr, FFA_ERR_INVALID, "n must be in [0, 2^32)");
    if (ns < 1 || ns > 4096 || nz < 1 || nx < 1 || nz > 1024 || nx > 1024)
        return fail(nullptr, FFA_ERR_INVALID, "grid (%d, %d, %d) out of range", ns, nz, nx);
    *bytes = sc_stage_workspace_bytes(n, ns, nz, nx);
    return FFA_OK;
}

int ffa_sc_local_coords(void *ctx, const double *bunch, const uint8_t *survive, int64_t n, double speed_c, int32_t rk,
                        double *out, double *count_sum, void *workspace, size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !count_sum) return fail(c, FFA_ERR_INVALID, "ctx or count_sum is NULL");
    if (n < 0 || n > 0xFFFFFFFFLL) return fail(c, FFA_ERR_INVALID, "n must be in [0, 2^32)");
    if (n > 0 && (!bunch || !survive || !out)) return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (!(speed_c > 0.0)) return fail(c, FFA_ERR_INVALID, "speed_c must be positive");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, n, 1, 1, 1)) return rc;
    return hip_check(c, launch_stage_local_coords(c, bunch, survive, n, speed_c, rk != 0, out, count_sum, workspace),
                     "ffa_sc_local_coords");
}

int ffa_sc_divide_arc(void *ctx, const double *fi_mod, int64_t n, double *out, void *workspace, size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !out) return fail(c, FFA_ERR_INVALID, "ctx or out is NULL");
    if (n < 0) return fail(c, FFA_ERR_INVALID, "negative count");
    if (n > 0 && !fi_mod) return fail(c, FFA_ERR_INVALID, "fi_mod is NULL");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, 0, 1, 1, 1)) return rc;
    return hip_check(c, launch_stage_divide_arc(c, fi_mod, n, out, workspace), "ffa_sc_divide_arc");
}

int ffa_sc_slices(void *ctx, const double *r, const double *z, const double *fi_mod, const double *r0, int64_t n,
                  double slice_step, double fi_s, int32_t ns, double *x, double *r2d, double *z2d, double *x2d,
                  int64_t stride, int32_t *counts, int32_t *slice_idx, double *sums, void *workspace,
                  size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !counts || !sums) return fail(c, FFA_ERR_INVALID, "ctx, counts or sums is NULL");
    if (n < 0 || n > 0xFFFFFFFFLL) return fail(c, FFA_ERR_INVALID, "n must be in [0, 2^32)");
    if (ns < 1 || ns > 4096) return fail(c, FFA_ERR_INVALID, "ns must be in [1, 4096]");
    if (stride < 0) return fail(c, FFA_ERR_INVALID, "negative stride");
    if (n > 0 && (!r || !z || !fi_mod || !r0 || !x || !slice_idx || !z2d || !x2d))
        return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (!(slice_step == slice_step)) return fail(c, FFA_ERR_INVALID, "slice_step is NaN");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, n, ns, 1, 1)) return rc;
    return hip_check(c, launch_stage_slices(c, r, z, fi_mod, r0, n, slice_step, fi_s, ns, x, r2d, z2d, x2d, stride,
                                            counts, slice_idx, sums, workspace),
                     "ffa_sc_slices");
}

int ffa_sc_deposit(void *ctx, const double *z2d, const double *x2d, int64_t stride, const int32_t *counts, int32_t ns,
                   int32_t nz, int32_t nx, double q_macro, double bunch_length, double xmin, double xmax, double zmin,
                   double zmax, int32_t bits, float *rho, void *workspace, size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !counts || !rho) return fail(c, FFA_ERR_INVALID, "ctx, counts or rho is NULL");
    if (ns < 1 || ns > 4096 || nz < 2 || nx < 2 || nz > 1024 || nx > 1024)
        return fail(c, FFA_ERR_INVALID, "grid (%d, %d, %d) out of range", ns, nz, nx);
    if (bits < 1 || bits > 20) return fail(c, FFA_ERR_INVALID, "bits must be in [1, 20]");
    if (stride < 0 || (stride > 0 && (!z2d || !x2d))) return fail(c, FFA_ERR_INVALID, "slice arrays missing");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, 0, ns, nz, nx)) return rc;
    return hip_check(c, launch_stage_deposit(c, z2d, x2d, stride, counts, ns, nz, nx, q_macro, bunch_length, xmin, xmax,
                                             zmin, zmax, bits, rho, workspace),
                     "ffa_sc_deposit");
}

int ffa_sc_green(void *ctx, int32_t nx, int32_t nz, double dx, double dz, float *kernel, void *workspace,
                 size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !kernel) return fail(c, FFA_ERR_INVALID, "ctx or kernel is NULL");
    if (nx < 1 || nz < 1 || nx > 1024 || nz > 1024) return fail(c, FFA_ERR_INVALID, "nx, nz must be in [1, 1024]");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, 0, 2, nz, nx)) return rc;
    return hip_check(c, launch_stage_green(c, nx, nz, dx, dz, kernel, workspace), "ffa_sc_green");
}

int ffa_sc_poisson(void *ctx, const float *rho, int32_t ns, int32_t nz, int32_t nx, double dx, double dz, float *phi,
                   void *workspace, size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !rho || !phi) return fail(c, FFA_ERR_INVALID, "ctx, rho or phi is NULL");
    if (int rc = check_grid(c, ns, nz, nx)) return rc;
    if (!(dx > 0.0) || !(dz > 0.0)) return fail(c, FFA_ERR_INVALID, "dx and dz must be positive");
    if (int rc = check_stage_ws(c, workspace, ws_bytes, 0, ns, nz, nx)) return rc;
    return hip_check(c, launch_stage_poisson(c, rho, ns, nz, nx, dx, dz, phi, workspace), "ffa_sc_poisson");
}

int ffa_sc_field(void *ctx, const float *phi, int32_t ns, int32_t nz, int32_t nx, double dr, double dz,
                 double bunch_length, double mean_gamma_m1, float *ex, float *ez, void *workspace, size_t ws_bytes) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !phi || !ex || !ez) return fail(c, FFA_ERR_INVALID, "ctx, phi, ex or ez is NULL");
    if (ns < 1 || ns > 4096 || nz < 2 || nx < 2 || nz > 1024 || nx > 1024)
        return fail(c, FFA_ERR_INVALID, "grid (%d, %d, %d) out of range", ns, nz, nx);
    if (int rc = check_stage_ws(c, workspace, ws_bytes, 0, 1, 1, 1)) return rc;
    return hip_check(c, launch_stage_field(c, phi, ns, nz, nx, dr, dz, bunch_length, mean_gamma_m1, ex, ez, workspace),
                     "ffa_sc_field");
}

int ffa_sc_total_field_1d(void *ctx, const double *s, const double *q, int32_t n, double *E) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(c, FFA_ERR_INVALID, "ctx is NULL");
    if (n < 0) return fail(c, FFA_ERR_INVALID, "negative count");
    if (n > 0 && (!s || !q || !E)) return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    return hip_check(c, launch_stage_total_field_1d(c, s, q, n, E), "ffa_sc_total_field_1d");
}

int ffa_sc_interpolate(void *ctx, const float *ex, const float *ez, int32_t ns, int32_t nz, int32_t nx, const double *x,
                       const double *z, const double *pr0, const int32_t *slice, const int64_t *rows, int64_t n,
                       double x0, double z0, double dx, double dz, double bunch_length, double *ex_p, double *ez_p,
                       double *ef_p, double *bunch, int64_t n_rows) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(c, FFA_ERR_INVALID, "ctx is NULL");
    if (ns < 1 || nz < 2 || nx < 2) return fail(c, FFA_ERR_INVALID, "grid (%d, %d, %d) out of range", ns, nz, nx);
    if (n < 0 || n_rows < 0) return fail(c, FFA_ERR_INVALID, "negative count");
    if (n > 0 && (!ex || !ez || !x || !z || !pr0 || !slice || !ex_p || !ez_p || !ef_p))
        return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (bunch && !rows) return fail(c, FFA_ERR_INVALID, "bunch given without rows");
    return hip_check(c, launch_stage_interp(c, ex, ez, ns, nz, nx, x, z, pr0, slice, rows, n, x0, z0, dx, dz,
                                            bunch_length, ex_p, ez_p, ef_p, bunch, n_rows),
                     "ffa_sc_interpolate");
}

int ffa_sc_solve_noalloc(void *ctx, const ffa_particles *p, const ffa_sc_params *sp) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !sp) return fail(c, FFA_ERR_INVALID, "ctx or params is NULL");
    if (!c->scratch) return fail(c, FFA_ERR_STATE, "the context has no scratch address");
    if (!(sp->c > 0.0)) return fail(c, FFA_ERR_INVALID, "speed of light must be positive");
    if (int rc = check_particles(c, p)) return rc;
    if (sp->ns < 2 || sp->ns > 4096) return fail(c, FFA_ERR_INVALID, "ns must be in [2, 4096]");
    if (!fft_size_ok(sp->nx) || !fft_size_ok(sp->nz))
        return fail(c, FFA_ERR_INVALID, "nx, nz must be in [32, 1024] with no prime factor above 5 (got %d, %d)",
                    sp->nx, sp->nz);
    if (sp->bits < 1 || sp->bits > 20) return fail(c, FFA_ERR_INVALID, "bits must be in [1, 20]");
    if ((double)p->n * (double)(1LL << (2 * sp->bits)) >= 9.2e18)
        return fail(c, FFA_ERR_INVALID, "fixed-point scale 2^%d too large for %lld particles", 2 * sp->bits,
                    (long long)p->n);
    if (!sp->seo_r || !sp->seo_pr || !sp->seo_ek_axis || !sp->seo_perimeter || sp->n_ek < 2 || sp->n_fi < 2)
        return fail(c, FFA_ERR_INVALID, "SEO tables missing or smaller than 2x2");
    ffa_sc_layout L;
    ffa_sc_layout_compute(p->n, sp->ns, sp->nz, sp->nx, &L);
    if (L.total > FFA_SC_NOALLOC_BYTES)
        return fail(c, FFA_ERR_INVALID,
                    "ffa_sc_solve_noalloc: %lld particles on a (%d, %d, %d) grid need %zu bytes of workspace, above "
                    "the %zu of its static array; use ffa_sc_solve with a caller workspace",
                    (long long)p->n, sp->ns, sp->nz, sp->nx, L.total, (size_t)FFA_SC_NOALLOC_BYTES);
    return hip_check(c, launch_sc_solve(c, p, sp, c->scratch, &L), "ffa_sc_solve_noalloc");
}

int ffa_reduce_phi(void *ctx, const ffa_particles *p, double *out_dev) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !out_dev) return fail(c, FFA_ERR_INVALID, "ctx or out is NULL");
    if (int rc = check_particles(c, p)) return rc;
    return hip_check(c, launch_reduce_phi(c, p, out_dev), "ffa_reduce_phi");
}

int ffa_detect_crossing(void *ctx, const double *fi_pre, const uint8_t *flags_pre, const double *fi_post, int64_t n,
                        double target, int64_t *idx_out, int64_t *count) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !count) return fail(c, FFA_ERR_INVALID, "ctx or count is NULL");
    if (n < 0) return fail(c, FFA_ERR_INVALID, "negative particle count");
    if (n > 0 && (!fi_pre || !flags_pre || !fi_post || !idx_out))
        return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (!(target >= 0.0 && target < FFA_TWO_PI)) return fail(c, FFA_ERR_INVALID, "target azimuth must be in [0, 2 pi)");
    return hip_check(c, launch_detect_crossing(c, fi_pre, flags_pre, fi_post, n, target, idx_out, count),
                     "ffa_detect_crossing");
}

int ffa_orbit_track(void *ctx, const ffa_orbit *o) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !o) return fail(c, FFA_ERR_INVALID, "ctx or orbit description is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (o->n < 0 || o->nsteps < 0) return fail(c, FFA_ERR_INVALID, "negative particle or step count");
    if (o->n > 0 && (!o->state || !o->q_over_p)) return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (!(o->h == o->h) || !(o->t_step == o->t_step) || !(o->t0 == o->t0))
        return fail(c, FFA_ERR_INVALID, "t0, t_step and h must be numbers");
    return hip_check(c, launch_orbit_rk4(c, o), "ffa_orbit_track");
}

int ffa_spiral_gap_maps(void *ctx, const ffa_spiral_gap *g) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !g) return fail(c, FFA_ERR_INVALID, "ctx or spiral gap description is NULL");
    if (g->nr < 2 || g->nphi < 2) return fail(c, FFA_ERR_INVALID, "the map needs at least 2 x 2 nodes");
    if ((int64_t)g->nr * g->nphi > ((int64_t)1 << 31) * 64) return fail(c, FFA_ERR_INVALID, "map too large");
    if (!(g->dr > 0.0) || !(g->dphi > 0.0)) return fail(c, FFA_ERR_INVALID, "dr and dphi must be positive");
    if (!(g->r_ref > 0.0) || !(g->r_min > 0.0) || !(g->r_max > g->r_min))
        return fail(c, FFA_ERR_INVALID, "need 0 < r_min < r_max and r_ref > 0");
    if (!(g->delta > 0.0)) return fail(c, FFA_ERR_INVALID, "delta must be positive");
    if (g->n_iter < 0) return fail(c, FFA_ERR_INVALID, "n_iter must be >= 0");
    if (!g->r_axis || !g->phi_axis || !g->er || !g->ephi || !g->ez_over_z || !g->dist || !g->flag)
        return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    return hip_check(c, launch_spiral_gap(c, g), "ffa_spiral_gap_maps");
}

int ffa_taylor_maps(void *ctx, const ffa_taylor_gen *g) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !g) return fail(c, FFA_ERR_INVALID, "ctx or generator description is NULL");
    if (g->n_orders < 1 || g->n_orders > FFA_TAYLOR_MAX_ORDERS)
        return fail(c, FFA_ERR_INVALID, "n_orders must be in [1, %d]", FFA_TAYLOR_MAX_ORDERS);
    if (g->n_q != 2 * g->n_orders) return fail(c, FFA_ERR_INVALID, "n_q must be 2 * n_orders");
    if (g->n_poles < 1 || g->n_poles > FFA_TAYLOR_MAX_POLES)
        return fail(c, FFA_ERR_INVALID, "n_poles must be in [1, %d]", FFA_TAYLOR_MAX_POLES);
    if (g->layout != 0 && g->layout != 1) return fail(c, FFA_ERR_INVALID, "layout must be 0 or 1");
    if (g->layout == 1 && g->n_orders < 2) return fail(c, FFA_ERR_INVALID, "the tracking layout needs n_orders >= 2");
    if (g->out_f32 != 0 && g->out_f32 != 1) return fail(c, FFA_ERR_INVALID, "out_f32 must be 0 or 1");
    if (g->nr < 2 || g->n_phi < 2) return fail(c, FFA_ERR_INVALID, "the grid needs at least 2 x 2 points");
    if (g->nr > 65535) return fail(c, FFA_ERR_INVALID, "nr must be <= 65535");
    if (g->n_phi > ((int64_t)1 << 23)) return fail(c, FFA_ERR_INVALID, "n_phi must be <= 2^23");
    if (!g->theta || !g->poles || !g->sig_poly || !g->weights || !g->shift || !g->flutter || !g->out)
        return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    return hip_check(c, launch_taylor_maps(c, g), "ffa_taylor_maps");
}

int ffa_sort_workspace_size(int64_t count, size_t *bytes) {
    if (!bytes) return fail(nullptr, FFA_ERR_INVALID, "bytes is NULL");
    if (count < 0 || count > 0xFFFFFFFFLL) return fail(nullptr, FFA_ERR_INVALID, "count must be in [0, 2^32)");
    *bytes = sort_workspace_bytes(count);
    return FFA_OK;
}

int ffa_sort_particles(void *ctx, const ffa_particles *src, const ffa_particles *dst, int64_t count, int32_t nbins,
                       const int64_t *pid_src, int64_t *pid_dst, void *workspace) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !workspace || !pid_src || !pid_dst) return fail(c, FFA_ERR_INVALID, "a pointer argument is NULL");
    if (int rc = check_particles(c, src)) return rc;
    if (int rc = check_particles(c, dst)) return rc;
    if (!dst || dst->n != src->n || dst->esc_f32 != src->esc_f32)
        return fail(c, FFA_ERR_INVALID, "dst must have the size and element types of src");
    if (count < 0 || count > src->n || count > 0xFFFFFFFFLL)
        return fail(c, FFA_ERR_INVALID, "count %lld outside [0, n]", (long long)count);
    if (nbins < 1 || nbins > 65535) return fail(c, FFA_ERR_INVALID, "nbins must be in [1, 65535]");
    if (src->r == dst->r || pid_src == pid_dst) return fail(c, FFA_ERR_INVALID, "src and dst must not alias");
    const uint32_t *order = nullptr;
    hipError_t e = launch_sort_order(c, src, count, nbins, workspace, &order);
    if (e == hipSuccess) e = launch_gather_particles(c, src, dst, order, count, pid_src, pid_dst);
    return hip_check(c, e, "ffa_sort_particles");
}

int ffa_bfield_eval(void *ctx, const double *r, const double *fi, const double *z, int64_t n, double *out,
                    int32_t mode, double zt) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(c, FFA_ERR_INVALID, "ctx is NULL");
    if (!c->have_machine) return fail(c, FFA_ERR_STATE, "ffa_set_machine has not been called");
    if (n < 0) return fail(c, FFA_ERR_INVALID, "negative point count");
    if (mode < 0 || mode > 3) return fail(c, FFA_ERR_INVALID, "mode must be 0, 1, 2 or 3");
    if (mode == 2 && !(zt > 0.0)) return fail(c, FFA_ERR_INVALID, "mode 2 needs zt > 0");
    if (n > 0 && (!r || !fi || !z || !out)) return fail(c, FFA_ERR_INVALID, "an array pointer is NULL");
    if (mode == 3) {
        if (c->host_machine.max_order > 0 && !c->host_machine.taylor_f32)
            return fail(c, FFA_ERR_INVALID, "ffa_bfield_eval mode 3 needs the FP32 Taylor map");
        return hip_check(c, launch_bfield_eval_f32(c, r, fi, z, n, out), "ffa_bfield_eval");
    }
    if (c->host_machine.max_order > 0 && c->host_machine.taylor_f32)
        return fail(c, FFA_ERR_INVALID, "ffa_bfield_eval modes 0-2 need the FP64 Taylor map");
    return hip_check(c, launch_bfield_eval(c, r, fi, z, n, out, mode, zt), "ffa_bfield_eval");
}

}  // extern "C"

namespace {

int check_mlp_desc(ffa_ctx_s *c, const ffa_mlp_desc *d) {
    if (!d) return fail(c, FFA_ERR_INVALID, "surrogate descriptor is NULL");
    if (!(d->zt > 0.0) || !(d->inv_rh > 0.0) || !(d->inv_sector > 0.0))
        return fail(c, FFA_ERR_INVALID, "surrogate descriptor needs zt, inv_rh, inv_sector > 0");
    if (!(d->out_delta >= 0.0)) return fail(c, FFA_ERR_INVALID, "out_delta must be >= 0");
    for (int k = 0; k < FFA_MLP_NFREQ; ++k)
        if (d->m[k] != std::floor(d->m[k]))
            return fail(c, FFA_ERR_INVALID, "azimuthal harmonic m[%d] = %g is not an integer", k, (double)d->m[k]);
    return FFA_OK;
}

}  // namespace

extern "C" {

int ffa_mlp_workspace_layout(int64_t n, ffa_mlp_layout *out) {
    if (!out) return fail(nullptr, FFA_ERR_INVALID, "out is NULL");
    if (n <= 0 || n > (int64_t)1 << 26) return fail(nullptr, FFA_ERR_INVALID, "batch must be in [1, 2^26]");
    mlp::workspace_layout(n, *out);
    return FFA_OK;
}

int ffa_mlp_pack(void *ctx, const float *params, void *packed) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !params || !packed) return fail(c, FFA_ERR_INVALID, "ctx, params or packed is NULL");
    return hip_check(c, launch_mlp_pack(c, params, packed), "ffa_mlp_pack");
}

int ffa_mlp_grad(void *ctx, const ffa_mlp_desc *d, const void *packed, const float *pts, const float *targets,
                 int64_t n, void *workspace, float *grad, float *loss) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !packed || !pts || !targets || !workspace || !grad)
        return fail(c, FFA_ERR_INVALID, "a pointer argument is NULL");
    if (int rc = check_mlp_desc(c, d)) return rc;
    if (n <= 0 || n > (int64_t)1 << 26) return fail(c, FFA_ERR_INVALID, "batch must be in [1, 2^26]");
    return hip_check(c,
                     launch_mlp_grad(c, d, packed, pts, targets, n, workspace, nullptr, grad, nullptr, nullptr,
                                     const_cast<void *>(packed), nullptr, loss),
                     "ffa_mlp_grad");
}

int ffa_mlp_adam(void *ctx, float *params, const float *grad, float *m, float *v, void *packed, const ffa_adam *hp) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !params || !grad || !m || !v || !packed || !hp) return fail(c, FFA_ERR_INVALID, "a pointer is NULL");
    if (hp->step < 1) return fail(c, FFA_ERR_INVALID, "Adam step count starts at 1");
    return hip_check(c, launch_mlp_adam(c, params, grad, m, v, packed, hp), "ffa_mlp_adam");
}

int ffa_mlp_train_step(void *ctx, const ffa_mlp_desc *d, float *params, float *m, float *v, void *packed,
                       const float *pts, const float *targets, int64_t n, void *workspace, const ffa_adam *hp,
                       float *loss) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !params || !m || !v || !packed || !pts || !targets || !workspace || !hp)
        return fail(c, FFA_ERR_INVALID, "a pointer argument is NULL");
    if (int rc = check_mlp_desc(c, d)) return rc;
    if (n <= 0 || n > (int64_t)1 << 26) return fail(c, FFA_ERR_INVALID, "batch must be in [1, 2^26]");
    if (hp->step < 1) return fail(c, FFA_ERR_INVALID, "Adam step count starts at 1");
    if (c->mlp_packed == packed)
        return fail(c, FFA_ERR_STATE, "these weights are in use by the push (ffa_mlp_set); train a copy");
    return hip_check(c,
                     launch_mlp_grad(c, d, packed, pts, targets, n, workspace, params, nullptr, m, v, packed, hp,
                                     loss),
                     "ffa_mlp_train_step");
}

int ffa_mlp_eval(void *ctx, const ffa_mlp_desc *d, const void *packed, const float *pts, int64_t n, float *out,
                 int32_t mode) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c || !packed) return fail(c, FFA_ERR_INVALID, "ctx or packed is NULL");
    if (int rc = check_mlp_desc(c, d)) return rc;
    if (n < 0) return fail(c, FFA_ERR_INVALID, "negative point count");
    if (n > 0 && (!pts || !out)) return fail(c, FFA_ERR_INVALID, "pts or out is NULL");
    if (mode != 0 && mode != 1) return fail(c, FFA_ERR_INVALID, "mode must be 0 or 1");
    return hip_check(c, launch_mlp_eval(c, d, packed, pts, n, out, mode), "ffa_mlp_eval");
}

int ffa_mlp_set(void *ctx, const ffa_mlp_desc *d, const void *packed) {
    ffa_ctx_s *c = as_ctx(ctx);
    if (!c) return fail(c, FFA_ERR_INVALID, "ctx is NULL");
    if (!packed) {
        c->mlp_packed = nullptr;
        return FFA_OK;
    }
    if (int rc = check_mlp_desc(c, d)) return rc;
    hipError_t e = mlp_upload_desc(c, d);
    if (e != hipSuccess) return hip_check(c, e, "ffa_mlp_set");
    c->mlp_packed = packed;
    return FFA_OK;
}

}  // extern "C"

// The order-0 flutter of one sector for the FP32 tier's table, in FP64 from
// the same logistic-edge model as the FP64 path (the convex rows: y = y02
// |y01| s1 s2, s1 = 1 / (1 + e^(-k (x - x0))), s2 = 1 / (1 + e^(k (x - x1))),
// k = 16 / W0) and its slope F' = sum y k (s2 - s1) at the nodes x_i = i h,
// h = P / FFA_FLUT_N.  Interval i stores the cubic Hermite interpolant of
// (F, F') between x_i and x_i+1 as c0 + c1 t + c2 t^2 + c3 t^3, t in [0, 1]:
// one 16-byte load gives F and, through dF/dt / h, F'.
static void flutter_table(const ffa_machine &m, std::vector<float> &t) {
    const double P = FFA_TWO_PI / (double)m.nsectors, h = P / (double)FFA_FLUT_N;
    std::vector<double> F(FFA_FLUT_N + 1), D(FFA_FLUT_N + 1);
    for (int i = 0; i <= FFA_FLUT_N; ++i) {
        const double x = (double)i * h;
        double f = 0.0, f1 = 0.0;
        for (int j = 0; j < m.n_seg; ++j) {
            const double x0 = m.convex[j][0], x1 = m.convex[j][1], ks = 16.0 / m.convex[j][2];
            const double yc = m.convex[j][4] * std::fabs(m.convex[j][3]);
            const double s1 = 1.0 / (1.0 + std::exp(-ks * (x - x0)));
            const double s2 = 1.0 / (1.0 + std::exp(ks * (x - x1)));
            const double y = yc * s1 * s2;
            f += y;
            f1 += y * ks * (s2 - s1);
        }
        F[i] = f;
        D[i] = h * f1;                 // dF/dt
    }
    t.assign(4 * (size_t)FFA_FLUT_N, 0.0f);
    for (int i = 0; i < FFA_FLUT_N; ++i) {
        const double a = F[i], b = F[i + 1], da = D[i], db = D[i + 1];
        t[4 * (size_t)i + 0] = (float)a;
        t[4 * (size_t)i + 1] = (float)da;
        t[4 * (size_t)i + 2] = (float)(3.0 * (b - a) - 2.0 * da - db);
        t[4 * (size_t)i + 3] = (float)(2.0 * (a - b) + da + db);
    }
}

void ffa_fconst_from_machine(const ffa_machine &m, ffa_fconst &f) {
    std::memset(&f, 0, sizeof f);
    f.T = m.taylor_f32 ? reinterpret_cast<const float *>(m.taylor) : nullptr;
    f.P = (float)(FFA_TWO_PI / (double)m.nsectors);
    f.invP = (float)((double)m.nsectors / FFA_TWO_PI);
    f.B0 = (float)m.B0;
    f.kv = (float)m.k_value;
    f.n_seg = m.n_seg;
    for (int j = 0; j < m.n_seg; ++j) {
        f.seg[j][0] = (float)m.convex[j][0];
        f.seg[j][1] = (float)m.convex[j][1];
        f.seg[j][2] = (float)(16.0 / m.convex[j][2]);
        f.seg[j][3] = (float)(m.convex[j][4] * std::fabs(m.convex[j][3]));
        f.seg[j][4] = (float)(16.0 / m.convex[j][2] * 1.4426950408889634);     // ks / ln 2
    }
    // Horner from the highest degree: leading zero coefficients add exact
    // zeros, so they are dropped (a zero spiral costs no work)
    int a0 = 0, b0 = 0;
    while (a0 < m.n_spoly && (float)m.s_poly[a0] == 0.0f) ++a0;
    while (b0 < m.n_sppoly && (float)m.sp_poly[b0] == 0.0f) ++b0;
    f.n_sp = m.n_spoly - a0;
    f.n_spp = m.n_sppoly - b0;
    for (int a = a0; a < m.n_spoly; ++a) f.sp[a - a0] = (float)m.s_poly[a];
    for (int a = b0; a < m.n_sppoly; ++a) f.spp[a - b0] = (float)m.sp_poly[a];
    f.no = m.max_order;
    f.NX = m.r_size;
    f.NY = m.fi_size;
    f.r_min = (float)m.r_min;
    f.inv_rstep = m.r_step > 0.0 ? (float)(1.0 / m.r_step) : 0.0f;
    f.fi_min = (float)m.fi_min;
    f.inv_fistep = m.fi_step > 0.0 ? (float)(1.0 / m.fi_step) : 0.0f;
    f.NXf = (float)m.r_size;
    f.NYf = (float)m.fi_size;
    f.NXm2f = (float)(m.r_size - 2);
    f.NYm2f = (float)(m.fi_size - 2);
    f.n_harm = m.n_harm;
    for (int k = 0; k < m.n_harm; ++k) {
        f.harm[k][0] = (float)m.harm_m[k];
        f.harm[k][1] = (float)m.harm_amp[k];
        f.harm[k][2] = (float)m.harm_phase[k];
        f.harm[k][3] = (float)m.harm_rmin[k];
        f.harm[k][4] = (float)m.harm_rmax[k];
    }
    f.rf = m.rf_enable;
    f.n_gap = m.n_gap;
    f.acc = (float)m.acc_etot;
    for (int g = 0; g < m.n_gap; ++g) {
        for (int j = 0; j < 4; ++j) {
            f.gap_box[g][j] = (float)m.gap_box[g][j];
            f.gap_px[g][j] = (float)m.gap_px[g][j];
            f.gap_py[g][j] = (float)m.gap_py[g][j];
        }
        f.gap_thn[g] = (float)(m.gap_theta[g] + FFA_PI / 2.0);
        f.gap_az[g] = (float)m.gap_theta[g];
    }
    for (int g = 0; g < FFA_MAX_GAP; ++g) f.rf_shift[g] = (float)m.rf_shift[g];
    f.gap_model = m.gap_model;
    if (m.gap_model == 1) {
        const double phi_max = m.emap_phi0 + m.emap_dphi * (double)(m.emap_nphi - 1);
        f.emap_nr = m.emap_nr;
        f.emap_nphi = m.emap_nphi;
        f.emap_r0 = (float)m.emap_r0;
        f.emap_inv_dr = (float)(1.0 / m.emap_dr);
        f.emap_phi0 = (float)m.emap_phi0;
        f.emap_inv_dphi = (float)(1.0 / m.emap_dphi);
        f.emap_phi_max = (float)phi_max;
        f.emap_center = (float)(0.5 * (m.emap_phi0 + phi_max));
        f.emap = reinterpret_cast<const double *>(m.emap);
    }
}
