// Internal declarations shared by the translation units of libffa_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "ffa_hip.h"

#define FFA_PI 3.141592653589793
#define FFA_TWO_PI (2.0 * FFA_PI)

// FP32 constants of the field evaluation, precomputed on the host from the
// machine description (divisions done once) and read by the FP32 tier with
// uniform loads.
struct ffa_fconst {
    const float *T;                // FP32 interleaved Taylor map (device)
    float P, invP;                 // sector period, 1 / period
    float B0, kv;
    int32_t n_seg, n_sp, n_spp, no;
    float seg[FFA_MAX_SEG][5];     // x0, x1, ks = 16/W0, y02 |y01|, ks log2(e)
    float sp[FFA_MAX_POLY], spp[FFA_MAX_POLY];
    int32_t NX, NY;
    float r_min, inv_rstep, fi_min, inv_fistep;
    float NXf, NYf, NXm2f, NYm2f;  // NX, NY, NX - 2, NY - 2 (index clamps)
    int32_t n_harm, rf, n_gap, pad;
    float harm[FFA_MAX_HARM][5];   // m, amp, phase, rmin, rmax
    float acc;                     // 1 / gap_width
    float gap_box[FFA_MAX_GAP][4];
    float gap_px[FFA_MAX_GAP][4], gap_py[FFA_MAX_GAP][4];
    float gap_thn[FFA_MAX_GAP];    // gap azimuth + pi/2
    float rf_shift[FFA_MAX_GAP];
    // spiral gaps: window [emap_phi0, emap_phi_max] about each gap azimuth,
    // bilinear maps on (r, local azimuth)
    int32_t gap_model, emap_nr, emap_nphi, pad2;
    float emap_r0, emap_inv_dr, emap_phi0, emap_inv_dphi, emap_phi_max, emap_center;
    float gap_az[FFA_MAX_GAP];
    const double *emap;
    // order-0 flutter F(fs) over one sector in FFA_FLUT_N intervals of
    // width h = P / N: per interval the cubic c0 + c1 t + c2 t^2 + c3 t^3 in
    // t = fs / h - i (device; written by ffa_set_machine)
    const float *flut;
    float flut_scale;              // N / P = 1 / h
    double flut_scale_d;           // N / P in FP64 (index from the FP64 sector azimuth)
};

void ffa_fconst_from_machine(const ffa_machine &m, ffa_fconst &f);

struct ffa_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool have_machine = false;
    ffa_machine host_machine{};
    // small device block owned by the context:
    //   [0, 8K)        ffa_machine
    //   [8K, 8K+256K)  step table for fused launches
    //   [..., +64K)    injection ranges for fused launches
    //   [..., +64K)    reduction partials / scalars
    //   [..., +4K)     surrogate descriptor
    //   [..., +128K)   order-0 flutter table of the FP32 tier
    char *dev = nullptr;
    size_t dev_bytes = 0;
    const void *mlp_packed = nullptr;   // surrogate weights used by the FP32 push (null: off)
    void *scratch = nullptr;            // this device's address of the static SC workspace array
};

static constexpr size_t FFA_CTX_MACHINE_OFF = 0;
static constexpr size_t FFA_CTX_STEPS_OFF = 8192;
static constexpr size_t FFA_CTX_STEPS_BYTES = 256 * 1024;
static constexpr size_t FFA_CTX_INJ_OFF = FFA_CTX_STEPS_OFF + FFA_CTX_STEPS_BYTES;
static constexpr size_t FFA_CTX_INJ_BYTES = 64 * 1024;
static constexpr size_t FFA_CTX_RED_OFF = FFA_CTX_INJ_OFF + FFA_CTX_INJ_BYTES;
static constexpr size_t FFA_CTX_RED_BYTES = 64 * 1024;
static constexpr size_t FFA_CTX_MLP_OFF = FFA_CTX_RED_OFF + FFA_CTX_RED_BYTES;   // surrogate descriptor
static constexpr size_t FFA_CTX_MLP_BYTES = 4096;
static constexpr int FFA_FLUT_N = 8192;                // flutter table intervals per sector
static constexpr size_t FFA_CTX_FLUT_OFF = FFA_CTX_MLP_OFF + FFA_CTX_MLP_BYTES;
static constexpr size_t FFA_CTX_FLUT_BYTES = (size_t)FFA_FLUT_N * 16;
static constexpr size_t FFA_CTX_DEV_BYTES = FFA_CTX_FLUT_OFF + FFA_CTX_FLUT_BYTES;
static constexpr int FFA_MAX_FUSED = (int)(FFA_CTX_STEPS_BYTES / sizeof(ffa_step));

static constexpr size_t FFA_CTX_FCONST_OFF = 6144;   // after the machine (< 6 KB)

// MI355X dispatches workgroups round-robin over its 8 XCDs (each with its own
// L2).  This maps the hardware block index to a logical one so that every XCD
// works on one contiguous range of logical blocks: neighbouring particles
// (ordered by azimuth) then share one L2 for their map and field reads.  A
// bijection on [0, gridDim.x) whatever the dispatch order.
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned b = blockIdx.x, nb = gridDim.x;
    const unsigned x = b & 7u, k = b >> 3;
    const unsigned q = nb >> 3, r = nb & 7u;
    return x * q + (x < r ? x : r) + k;
}

inline const ffa_machine *ctx_machine_dev(ffa_ctx_s *c) {
    return reinterpret_cast<const ffa_machine *>(c->dev + FFA_CTX_MACHINE_OFF);
}

inline const ffa_fconst *ctx_fconst_dev(ffa_ctx_s *c) {
    return reinterpret_cast<const ffa_fconst *>(c->dev + FFA_CTX_FCONST_OFF);
}

// launchers implemented in the .hip files; return a hipError_t
hipError_t launch_inject(ffa_ctx_s *c, const ffa_particles *p, int64_t b, int64_t e);
hipError_t launch_boris_f64(ffa_ctx_s *c, const ffa_particles *p, const ffa_step *s);
hipError_t launch_rk4_f64(ffa_ctx_s *c, const ffa_particles *p, const ffa_step *s);
hipError_t launch_boris_fused_f32(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps);
hipError_t launch_boris_f64_multi(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps);
hipError_t launch_reduce_phi(ffa_ctx_s *c, const ffa_particles *p, double *out_dev);
hipError_t launch_bfield_eval(ffa_ctx_s *c, const double *r, const double *fi, const double *z, int64_t n,
                              double *out, int mode, double zt);
hipError_t launch_bfield_eval_f32(ffa_ctx_s *c, const double *r, const double *fi, const double *z, int64_t n,
                                  double *out);
hipError_t launch_mlp_pack(ffa_ctx_s *c, const float *params, void *packed);
hipError_t launch_mlp_grad(ffa_ctx_s *c, const ffa_mlp_desc *d, const void *packed, const float *pts,
                           const float *tgt, int64_t n, void *ws, float *params, float *grad, float *mom, float *vel,
                           void *packed_out, const ffa_adam *hp, float *loss);
hipError_t launch_mlp_adam(ffa_ctx_s *c, float *params, const float *grad, float *mom, float *vel, void *packed,
                           const ffa_adam *hp);
hipError_t launch_mlp_eval(ffa_ctx_s *c, const ffa_mlp_desc *d, const void *packed, const float *pts, int64_t n,
                           float *out, int mode);
hipError_t launch_detect_crossing(ffa_ctx_s *c, const double *fi_pre, const uint8_t *flags_pre, const double *fi_post,
                                  int64_t n, double target, int64_t *idx_out, int64_t *count);
hipError_t launch_taylor_maps(ffa_ctx_s *c, const ffa_taylor_gen *g);
hipError_t launch_orbit_rk4(ffa_ctx_s *c, const ffa_orbit *o);
hipError_t launch_spiral_gap(ffa_ctx_s *c, const ffa_spiral_gap *g);
size_t sort_workspace_bytes(int64_t count);
hipError_t launch_sort_order(ffa_ctx_s *c, const ffa_particles *p, int64_t count, int nbins, void *ws,
                             const uint32_t **order_out);
uint16_t *sort_key_buffer(void *ws);
hipError_t launch_sort_keys16(ffa_ctx_s *c, int64_t count, int max_key, void *ws, const uint32_t **order_out,
                              const uint16_t **keys_out);
hipError_t launch_gather_particles(ffa_ctx_s *c, const ffa_particles *src, const ffa_particles *dst,
                                   const uint32_t *order, int64_t count, const int64_t *pid_src, int64_t *pid_dst);
hipError_t mlp_upload_desc(ffa_ctx_s *c, const ffa_mlp_desc *d);   // -> FFA_CTX_MLP_OFF (synchronous)
namespace mlp {
size_t workspace_layout(int64_t n, ffa_mlp_layout &L);
}
size_t sc_stage_workspace_bytes(int64_t n, int32_t ns, int32_t nz, int32_t nx);
hipError_t launch_stage_local_coords(ffa_ctx_s *c, const double *bunch, const uint8_t *survive, int64_t n,
                                     double speed_c, int rk, double *out, double *count_sum, void *ws);
hipError_t launch_stage_divide_arc(ffa_ctx_s *c, const double *fi, int64_t n, double *out, void *ws);
hipError_t launch_stage_slices(ffa_ctx_s *c, const double *r, const double *z, const double *fim, const double *r0,
                               int64_t n, double step, double fs, int ns, double *x, double *r2d, double *z2d,
                               double *x2d, int64_t stride, int32_t *counts, int32_t *sl, double *sums, void *ws);
hipError_t launch_stage_deposit(ffa_ctx_s *c, const double *z2d, const double *x2d, int64_t stride,
                                const int32_t *counts, int ns, int nz, int nx, double q_macro, double bl, double xmin,
                                double xmax, double zmin, double zmax, int bits, float *rho, void *ws);
hipError_t launch_stage_green(ffa_ctx_s *c, int nx, int nz, double dx, double dz, float *out, void *ws);
hipError_t launch_stage_poisson(ffa_ctx_s *c, const float *rho, int ns, int nz, int nx, double dx, double dz, float *phi,
                                void *ws);
hipError_t launch_stage_field(ffa_ctx_s *c, const float *phi, int ns, int nz, int nx, double dr, double dz, double bl,
                              double mgm1, float *ex, float *ez, void *ws);
hipError_t launch_stage_total_field_1d(ffa_ctx_s *c, const double *s, const double *q, int n, double *E);
hipError_t launch_stage_interp(ffa_ctx_s *c, const float *ex, const float *ez, int ns, int nz, int nx, const double *x,
                               const double *z, const double *pr0, const int32_t *slice, const int64_t *rows, int64_t n,
                               double x0, double z0, double dx, double dz, double bl, double *ex_p, double *ez_p,
                               double *ef_p, double *bunch, int64_t n_rows);
hipError_t launch_sc_solve(ffa_ctx_s *c, const ffa_particles *p, const ffa_sc_params *sp,
                           void *ws, const ffa_sc_layout *L, bool gather = true);
// What the gather of a solve reads (its parameters and workspace regions),
// for the FP32 push that applies a solve's gather in its prologue.
struct sc_gather_src {
    ffa_sc_params sp;
    const ffa_sc_result *res;
    const float2 *exz;
    const float *phi;
    const double *px, *ppr0;
    const int32_t *pslice;
};
// Fills g from the solve's parameters and workspace; returns true when the
// gather differentiates phi (no field cubes on this grid).
bool sc_gather_inputs(const ffa_sc_params *sp, void *ws, const ffa_sc_layout *L, sc_gather_src *g);
hipError_t launch_boris_fused_f32_gather(ffa_ctx_s *c, const ffa_particles *p, int32_t nsteps, const sc_gather_src &g,
                                         bool from_phi);
