/*
 * ffa_hip.h -- C ABI of libffa_hip.so, the MI355X (gfx950) implementation of
 * PyFFA's multiparticle tracking hot path.
 *
 * The reference has no compiled boundary: its Python classes call Numba
 * kernels directly.  Each entry point below replaces one of those calls (or a
 * fused group of them); the citation on each names the reference interface it
 * stands in for.  All pointers inside the structs are DEVICE pointers (the
 * Python side owns the memory as torch tensors and passes data_ptr()), except
 * where a comment says host.  Every call is asynchronous on the context's
 * stream and returns 0 on success or a negative FFA_ERR_* code;
 * ffa_last_error() then describes the failure.  A context is used by one host
 * thread at a time.
 */
#ifndef FFA_HIP_H
#define FFA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FFA_OK 0
#define FFA_ERR_INVALID (-1)   /* bad argument / shape */
#define FFA_ERR_HIP (-2)       /* a HIP runtime call failed */
#define FFA_ERR_STATE (-3)     /* call sequence error (e.g. no machine set) */

#define FFA_MAX_SEG 8
#define FFA_MAX_POLY 8
#define FFA_MAX_HARM 8
#define FFA_MAX_GAP 8

/* Per-particle flag byte (replaces LocalBunch columns 7 and 8 and the
 * Malloc_SC_survive_flag_Pre/PrePre masks, src/FFAG_Bunch.py:23-62, 241-242). */
#define FFA_F_INJ 1u
#define FFA_F_ALIVE 2u
#define FFA_F_SCPRE 4u
#define FFA_F_SCPREPRE 8u

/* Particle state as structure-of-arrays (LocalBunch columns 0-5 and 9 plus
 * the SC endpoint fields Malloc_SC_E*_Pre / _Pre_Pre, src/FFAG_Bunch.py:
 * 144-149).  fidot holds dphi/dt [rad/s] for the Boris pusher and E_total [J]
 * for RK4 (src/FFAG_ParasAndConversion.py:202-269).  esc_pre / esc_prepre are
 * 3*n values, planes (E_r, E_z, E_phi); their element type is double for the
 * FP64 path and float for the FP32 tier (esc_f32 = 1). */
typedef struct {
    double *r, *vr, *z, *vz, *fi, *fidot, *rf_phase;
    uint8_t *flags;
    void *esc_pre, *esc_prepre;
    int64_t n;
    int32_t esc_f32;
} ffa_particles;

/* Fields and constants (what FFAG_GlobalParameters carries into the kernels:
 * FFAG_Bfield_analytical src/FFAG_Field.py:641-706, FFAG_EField_new
 * src/FFAG_Field.py:168-198, FFAG_BField_Error src/FFAG_Field.py:965-998,
 * constants src/FFAG_ParasAndConversion.py:13-19). */
typedef struct {
    /* Taylor maps, device, interleaved [fi_size][r_size][max_order][3] with the
     * component order (Bz, Bfi, Br); orders 1..max_order only. */
    const void *taylor;
    int32_t taylor_f32;         /* element type of taylor: 0 double, 1 float */
    int32_t max_order;
    int32_t r_size, fi_size;
    double r_min, r_step, fi_min, fi_step;
    /* analytic order-0 field */
    double B0, k_value;
    int32_t nsectors, n_seg;
    double convex[FFA_MAX_SEG][5];     /* (x0, x1, W0, y01, y02) rows */
    int32_t n_spoly, n_sppoly;
    double s_poly[FFA_MAX_POLY], sp_poly[FFA_MAX_POLY];   /* highest degree first */
    /* harmonic errors */
    int32_t n_harm;
    double harm_m[FFA_MAX_HARM], harm_amp[FFA_MAX_HARM], harm_phase[FFA_MAX_HARM];
    double harm_rmin[FFA_MAX_HARM], harm_rmax[FFA_MAX_HARM];
    /* fixed RF gaps */
    int32_t rf_enable, n_gap;
    double acc_etot;                    /* 1 / gap_width */
    double gap_theta[FFA_MAX_GAP];      /* gap azimuth [rad] */
    double gap_px[FFA_MAX_GAP][4], gap_py[FFA_MAX_GAP][4];
    double gap_box[FFA_MAX_GAP][4];     /* x_min, x_max, y_min, y_max */
    double rf_shift[FFA_MAX_GAP];
    /* spiral RF gaps (FFAG_EField_spiral, src/FFAG_Field.py:379-498):
     * gap_model 1 replaces the fixed rectangles by the first gap whose azimuth
     * window gap_theta[g] + [emap_phi0, emap_phi0 + (emap_nphi-1) emap_dphi]
     * (shifted by whole turns) holds the particle, with the fields bilinear in
     * (r, local azimuth) from three maps scaled by V0 acc_etot */
    int32_t gap_model;                  /* 0 fixed rectangles, 1 spiral maps */
    int32_t emap_nr, emap_nphi;
    double emap_r0, emap_dr, emap_phi0, emap_dphi;
    const void *emap;                   /* device double [emap_nr][emap_nphi][3]: Er, Ez / z, Ephi */
    /* aperture */
    int32_t aperture_enable;
    double aperture[4];                 /* rmin, rmax, zmin, zmax */
    /* constants */
    double q, c, m0, E0;
} ffa_machine;

/* One push step (scalars the reference computes on the host per step). */
typedef struct {
    double t;          /* time at the start of the step [s] */
    double dt;         /* step [s] */
    double dRF;        /* RF phase advance 2 pi f(t) dt (Boris) */
    double v0;         /* gap voltage V0(t) [V] */
    double sc_s;       /* SC extrapolation factor k/SC_step (0 before the 2nd solve) */
    /* RK4 only: f(t) and V0(t) at t, t+h/2, t+h */
    double rk_rate[3]; /* 2 pi f */
    double rk_v0[3];
} ffa_step;

/* SC solve parameters (FFAG_Bunch attributes read by Build_2_5D_SliceGrid_sbs,
 * src/FFAG_Bunch.py:1157-1352). */
typedef struct {
    int32_t ns, nx, nz;
    double sigma_ratio;
    double q_macro;        /* marcosize * q */
    int32_t rk;            /* gamma from E_total (GetLocalCoordinates_2p5_rk) */
    int32_t first;         /* first solve of the run (steps_i == 0) */
    double s_last;         /* extrapolation factor of the step before this solve */
    int32_t bits;          /* fixed-point fractional bits per CIC axis, 1..20 */
    int32_t cell_counts;   /* 1: also store the particles per CIC base cell in the
                              cell_count region (a checking aid; 0 leaves it unwritten) */
    int32_t charge_grid;   /* 1: store the int64 fixed-point charge grid in the grid
                              region; 0: the library chooses (grids of up to 2^23
                              cells store it, larger ones deposit the scaled float32
                              charge straight into the packed FFT buffer) */
    int32_t field_cubes;   /* 1: store the (Ex, Ez) cubes in the exz region and gather
                              the kicks from them; 0: the library chooses, and on grids
                              of up to 2^23 cells the gather differentiates phi at the
                              particle's four cells itself (same arithmetic, same kicks)
                              and leaves the exz region unwritten, while above 2^23
                              cells it stores the cubes and leaves the phi region
                              unwritten (phi never leaves LDS) */
    /* SEO tables, device, row-major (n_ek, n_fi) */
    const double *seo_r, *seo_pr;
    const double *seo_ek_axis, *seo_perimeter;   /* n_ek each */
    int32_t n_ek, n_fi;
    double ek0, dek, fi0, dfi;
    double E0_MeV;         /* rest energy used for Ek = (gamma-1) E0 */
    double c;              /* speed of light used for gamma (FFAG_Bunch.speed_c) */
} ffa_sc_params;

/* Scalars a solve leaves on the device (read back only when asked). */
typedef struct {
    double count, sum_gamma, fi_s, fi_e, slice_step, mean_gamma_m1;
    double sum_x, sum_z, sum_x2, sum_z2;
    double xmin, xmax, zmin, zmax, dx, dz;
    double bunch_length, perimeter, qscale;
    double bad_grid;
    double arc_min[4], arc_max[4];
    double pad[4];
} ffa_sc_result;

/* Device workspace of one SC solve; ffa_sc_workspace_layout() gives the
 * byte offset of each region inside one caller-allocated block. */
typedef struct {
    size_t total;
    size_t grid;        /* int64 (ns, nz, nx) fixed-point charge (when charge_grid asks) */
    size_t cell_count;  /* int32 (ns, nz, nx) particles per CIC base cell */
    size_t cbuf;        /* complex64 (ceil(ns/2), nz, nx) packed slice pairs */
    size_t phi;         /* float (ns, nz, nx) potential (not written when the library
                           chooses the field cubes itself: field_cubes = 0 above 2^23 cells) */
    size_t exz;         /* float2 (ns, nz, nx): (Ex, Ez) of each cell, interleaved */
    size_t px;          /* double n: x = r - r0 */
    size_t ppr0;        /* double n: pr0 */
    size_t pslice;      /* int32 n: slice index */
    size_t slice_count; /* int32 ns */
    size_t partials;    /* double [FFA_SC_BLOCKS][16] */
    size_t result;      /* ffa_sc_result */
    size_t records;     /* uint64 n: slice-sorted packed CIC records */
    size_t slice_start; /* int32 ns + 1: record offset of each slice */
    size_t block_offset; /* int32 [FFA_SC_BLOCKS][ns]: record offset of each block's run inside each slice */
    size_t records_by_row;  /* uint64 n: records of each slice ordered by CIC row */
    size_t row_start;   /* int32 ns * (nz + 1): record offset of each (slice, row) */
    size_t tables;      /* FFT twiddles (4 x 1024 complex64) and the Green's function (nx * nz float32) */
} ffa_sc_layout;

#define FFA_SC_BLOCKS 1024

/* Context lifecycle.  stream may be NULL (default stream). */
int ffa_create(int device, void *stream, void **ctx_out);
int ffa_destroy(void *ctx);
const char *ffa_last_error(void *ctx);
/* Later calls run on stream; the new stream first waits for the work already
 * queued on the old one (an event), so a call that rewrites the context's
 * device tables cannot overtake kernels still reading them. */
int ffa_set_stream(void *ctx, void *stream);

/* FFAG_Bfield_analytical / FFAG_EField_new / FFAG_BField_Error /
 * GlobalParameters.AddAperture -> one machine description
 * (examples/main_track.py:247-270).  Host pointer; copied. */
int ffa_set_machine(void *ctx, const ffa_machine *m);

/* FFAG_Bunch.InjectParticles, src/FFAG_Bunch.py:265-303: the host finds the
 * contiguous index range [begin, end) by searchsorted; this sets the
 * injected and alive flags of that range. */
int ffa_inject(void *ctx, const ffa_particles *p, int64_t begin, int64_t end);

/* One Boris step in FP64: the SC extrapolation of src/FFAG_track.py:449-478
 * fused with Boris_solve, src/FFAG_track.py:948-1049 (fast_mod_parallel,
 * GetBfields_Njit, Bmap_add_all_orders, harmonics, emap_kernel per gap,
 * cylindrical_to_cartesian, boris_push_cartesian_njit,
 * convert_cartesian_to_cylindrical). */
int ffa_push_boris_f64(void *ctx, const ffa_particles *p, const ffa_step *s);

/* One classic RK4 step in FP64: rk4_step_malloc + RK4_Post_Step_njit,
 * src/FFAG_track.py:51-61, src/FFAG_MathTools.py:2205-2299, with the four
 * FunctionForAccelerationBunch_dt evaluations (src/FFAG_track.py:1743-1801)
 * kept in registers. */
int ffa_push_rk4_f64(void *ctx, const ffa_particles *p, const ffa_step *s);

/* nsteps consecutive calls of ffa_inject + ffa_push_boris_f64 (rk4 == 0) or
 * ffa_push_rk4_f64 (rk4 != 0) issued from C: the reference loop body of
 * src/FFAG_track.py:331-556 for steps without an SC solve, one kernel launch
 * per step, without a host round trip per step.  steps: host array of
 * nsteps entries; inj: host array of nsteps (begin, end) pairs. */
int ffa_push_f64_multi(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                       int32_t nsteps, int32_t rk4);

/* The same nsteps Boris steps as ffa_push_f64_multi(rk4 = 0) -- the loop body
 * of src/FFAG_track.py:331-556 for steps without an SC solve, each step the
 * one ffa_push_boris_f64 replaces -- in one launch with the state held in
 * registers between steps: bit-identical results without the state's round
 * trip through HBM at every step (nsteps <= 2978, the context's step table).  Host arrays as for
 * ffa_push_f64_multi. */
int ffa_push_boris_f64_fused(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                             int32_t nsteps);

/* nsteps consecutive Boris steps of the FP32-field tier in one launch
 * (state and RF phase in FP64, field evaluation and SC kick in FP32, state
 * held in registers between steps).  steps: host array of nsteps entries.
 * inj: host array of nsteps (begin, end) pairs injected before each step
 * (begin == end for none).  Same physics as nsteps calls of
 * ffa_push_boris_f64 up to FP32 field rounding. */
int ffa_push_boris_fused_f32(void *ctx, const ffa_particles *p, const ffa_step *steps,
                             const int64_t *inj, int32_t nsteps);

/* Workspace layout of an SC solve for n particles on an (ns, nz, nx) grid. */
int ffa_sc_workspace_layout(int64_t n, int32_t ns, int32_t nz, int32_t nx, ffa_sc_layout *out);

/* FFAG_Bunch.Build_2_5D_SliceGrid_sbs, src/FFAG_Bunch.py:1157-1352, plus
 * the endpoint roll of src/FFAG_track.py:356-382: survivor coordinates,
 * covering arc, SEO lookup, slicing and moments, fixed-point CIC deposit,
 * periodic FFT Poisson solve per slice, -grad(phi)/gamma^2, bilinear gather
 * with the pr0 projection, then Pre/PrePre endpoint fields and masks rolled.
 * The roll swaps buffer roles: the new pre-step field is written into the
 * memory p->esc_prepre points to, and the memory p->esc_pre points to (left
 * untouched) holds the new prepre field, so the caller swaps its esc_pre and
 * esc_prepre pointers after the call.  workspace: device block laid out by
 * ffa_sc_workspace_layout. */
int ffa_sc_solve(void *ctx, const ffa_particles *p, const ffa_sc_params *sp, void *workspace);

/* ffa_sc_solve up to the field, with its gather and endpoint roll left to
 * the next ffa_push_boris_fused_f32_sc: the particles' endpoint fields and
 * flags are not touched and the endpoint pointers are not swapped yet.  The
 * tracking loop of the FP32 tier solves this way, so that the kicks are
 * gathered inside the push launch that uses them (one pass over the
 * particles fewer per solve).  Same checks as ffa_sc_solve.  This split has
 * no counterpart in the reference, whose solve and push are one Python loop
 * (src/FFAG_track.py:356-382, 449-478). */
int ffa_sc_solve_field(void *ctx, const ffa_particles *p, const ffa_sc_params *sp, void *workspace);

/* The gather and endpoint roll of the solve ffa_sc_solve_field left in
 * workspace (same p, sp and workspace), then nsteps steps as
 * ffa_push_boris_fused_f32 (nsteps may be 0: the gather alone).  Results
 * are bit-identical to ffa_sc_solve followed by the pointer swap and
 * ffa_push_boris_fused_f32: the gather runs per particle with the solve's
 * own FP64 arithmetic.  p carries the endpoint pointers the solve saw; the
 * caller swaps esc_pre and esc_prepre after this call, as after
 * ffa_sc_solve.  Refused while the surrogate is enabled. */
int ffa_push_boris_fused_f32_sc(void *ctx, const ffa_particles *p, const ffa_step *steps, const int64_t *inj,
                                int32_t nsteps, const ffa_sc_params *sp, const void *workspace);

/* ---- Stage-level space charge: the module functions of FFAG_SC_new -------
 * Each entry point below runs one stage of the 2.5D solve on device arrays
 * laid out as the reference's numpy arrays (the (N, 16) row-major particle
 * matrix, (ns, stride) per-slice lists, (ns, nz, nx) float32 cubes), for
 * callers that drive the stages themselves (src/FFAG_Bunch.py:1157-1352
 * calls them in this order).  Entry points with a workspace take a device
 * block of at least ffa_sc_stage_workspace_size(n, ns, nz, nx) bytes for the
 * largest n and grid they are called with, and its size ws_bytes.  Sums are
 * fixed-order tree reductions (reproducible run to run; not the
 * reference's sequential rounding). */
int ffa_sc_stage_workspace_size(int64_t n, int32_t ns, int32_t nz, int32_t nx, size_t *bytes);

/* GetLocalCoordinates_2p5 / _rk (src/FFAG_SC_new.py:74-148): the rows with
 * survive[i] != 0, in row order, into out = 6 planes of n doubles (r, fi, z,
 * gamma, fi mod 2 pi, row index); count_sum (device, 2 doubles) = (count,
 * sum of gamma).  gamma = 1/sqrt(1 - v^2/c^2) from columns 0-5 (rk = 0) or
 * E_total / E0 (rk != 0). */
int ffa_sc_local_coords(void *ctx, const double *bunch, const uint8_t *survive, int64_t n, double speed_c, int32_t rk,
                        double *out, double *count_sum, void *workspace, size_t ws_bytes);

/* divide_arc2 (src/FFAG_SC_new.py:534-566): out (device, 3 doubles) =
 * (arc start, arc end, span) of the smallest of the four shifted covers. */
int ffa_sc_divide_arc(void *ctx, const double *fi_mod, int64_t n, double *out, void *workspace, size_t ws_bytes);

/* ComputeSliceBoundsAndCoordsSBS_merged (src/FFAG_SC_new.py:449-531):
 * x = r - r0, slice s = int((fi_mod (+2 pi if below fi_s) - fi_s) / slice_step)
 * clamped to [0, ns); counts[ns], slice_idx[n], sums (device, 4 doubles) =
 * (sum x, sum z, sum x^2, sum z^2); the per-slice lists r2d (may be NULL),
 * z2d, x2d (ns rows of `stride` doubles) in particle order. */
int ffa_sc_slices(void *ctx, const double *r, const double *z, const double *fi_mod, const double *r0, int64_t n,
                  double slice_step, double fi_s, int32_t ns, double *x, double *r2d, double *z2d, double *x2d,
                  int64_t stride, int32_t *counts, int32_t *slice_idx, double *sums, void *workspace,
                  size_t ws_bytes);

/* DistributeChargeSBS_Njit (src/FFAG_SC_new.py:244-312) on the per-slice
 * lists: the reference's CIC cells and clamped fractions, rounded to `bits`
 * fractional bits and summed as integers (bit-reproducible), scaled by
 * q_macro ns / (bunch_length dx dz) into rho (ns, nz, nx) float32; zeros
 * when dx, dz or bunch_length is not positive. */
int ffa_sc_deposit(void *ctx, const double *z2d, const double *x2d, int64_t stride, const int32_t *counts, int32_t ns,
                   int32_t nz, int32_t nx, double q_macro, double bunch_length, double xmin, double xmax, double zmin,
                   double zmax, int32_t bits, float *rho, void *workspace, size_t ws_bytes);

/* make_fft_kernel_open_numba (src/FFAG_SC_new.py:731-760): the periodic
 * Green's function 1 / (eps0 k^2), (nz, nx) float32, k = 0 mode zero. */
int ffa_sc_green(void *ctx, int32_t nx, int32_t nz, double dx, double dz, float *kernel, void *workspace,
                 size_t ws_bytes);

/* The per-slice periodic Poisson solve of Build_2_5D_SliceGrid_sbs
 * (src/FFAG_Bunch.py:1330-1338, pyFFTW plans of src/FFAG_track.py:318-328):
 * phi_k = Re IFFT2(FFT2(rho_k) G) / (nz nx) for every slice, on the LDS
 * Stockham FFT (nx, nz in [32, 1024] with no prime factor above 5). */
int ffa_sc_poisson(void *ctx, const float *rho, int32_t ns, int32_t nz, int32_t nx, double dx, double dz, float *phi,
                   void *workspace, size_t ws_bytes);

/* compute_field_from_potential_numba_SBS (src/FFAG_SC_new.py:672-728) for ns
 * slices of phi: ex, ez (ns, nz, nx) float32 = -grad(phi) / gamma^2 with the
 * reference's edge stencils; zeros when dr, dz or bunch_length is 0. */
int ffa_sc_field(void *ctx, const float *phi, int32_t ns, int32_t nz, int32_t nx, double dr, double dz,
                 double bunch_length, double mean_gamma_m1, float *ex, float *ez, void *workspace, size_t ws_bytes);

/* total_field_1d_numba (src/FFAG_SC_new.py:763-782): E[i] = K sum_{j != i}
 * q_j sign(s_i - s_j) / (s_i - s_j)^2, zeros unless s[1] > s[0]. */
int ffa_sc_total_field_1d(void *ctx, const double *s, const double *q, int32_t n, double *E);

/* interpolate_fields_to_particles_SBS_numba (src/FFAG_SC_new.py:870-972):
 * bilinear ex, ez of each particle's slice at (x, z), projected with pr0,
 * into ex_p, ez_p, ef_p and, when bunch is given, columns 10-12 of row
 * rows[k] of the (n_rows, 16) matrix; zeros when dx, dz or bunch_length is
 * not positive. */
int ffa_sc_interpolate(void *ctx, const float *ex, const float *ez, int32_t ns, int32_t nz, int32_t nx, const double *x,
                       const double *z, const double *pr0, const int32_t *slice, const int64_t *rows, int64_t n,
                       double x0, double z0, double dx, double dz, double bunch_length, double *ex_p, double *ez_p,
                       double *ef_p, double *bunch, int64_t n_rows);

/* ffa_sc_solve without a caller workspace and without any allocation (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a static device array of
 * FFA_SC_NOALLOC_BYTES per GPU, held from the library's load to the end of
 * the process; problems whose ffa_sc_workspace_layout total exceeds it are
 * rejected (FFA_ERR_INVALID).  No region of the array needs clearing
 * between calls (every region a solve reads, it first writes), but all calls
 * on one GPU share the array: calls that overlap in time on one GPU, from
 * any contexts or streams, race.  Same results as ffa_sc_solve. */
#define FFA_SC_NOALLOC_BYTES ((size_t)256 << 20)
int ffa_sc_solve_noalloc(void *ctx, const ffa_particles *p, const ffa_sc_params *sp);

/* Stop-condition reduction of src/FFAG_track.py:559-577: out (device, 2
 * doubles) = (sum of non-NaN phi, count of non-NaN phi) over all particles. */
int ffa_reduce_phi(void *ctx, const ffa_particles *p, double *out_dev);

/* Position detectors: detect_azimuth_crossing (src/FFAG_MathTools.py:
 * 1633-1695) as PositionDump / PositionDumpBunch use it
 * (src/FFAG_dump.py:610-626, 893-908).  A particle alive before the step
 * (flags_pre, the reference's Step_SurviveFlag) crosses target in [0, 2 pi)
 * when its pre/post azimuths, reduced mod 2 pi, bracket target (wrapping
 * steps included).  Writes the crossing particles' positions to idx_out (n
 * int64 capacity, arbitrary order) and their number to *count (device int64). */
int ffa_detect_crossing(void *ctx, const double *fi_pre, const uint8_t *flags_pre, const double *fi_post, int64_t n,
                        double target, int64_t *idx_out, int64_t *count);

/* Particle order (no counterpart in the reference, which keeps the input
 * order): a stable sort of particles [0, count) by azimuth bin
 * floor((fi mod 2 pi) * nbins / 2 pi) (particles not injected-and-alive
 * last), written into a second set of arrays dst of the same size and
 * element types, particles [count, n) copied unchanged.  pid_src / pid_dst
 * (int64, n) carry each particle's input index so the caller can restore the
 * input order.  Neighbouring threads then read neighbouring Taylor-map cells
 * in the push and neighbouring slices in the SC gather.  nbins in
 * [1, 65535]; workspace: ffa_sort_workspace_size(count) bytes. */
int ffa_sort_workspace_size(int64_t count, size_t *bytes);
int ffa_sort_particles(void *ctx, const ffa_particles *src, const ffa_particles *dst, int64_t count, int32_t nbins,
                       const int64_t *pid_src, int64_t *pid_dst, void *workspace);

/* Static magnetic field at n points (device FP64 arrays r, fi, z; out: 3*n
 * doubles, planes Bz, Bfi, Br) from the machine set by ffa_set_machine:
 *   mode 0: GetBfields_Njit + Bmap_add_all_orders + harmonics
 *           (src/FFAG_Field.py:872-884, 1333-1426; src/FFAG_MathTools.py:
 *           1139-1188; src/FFAG_Field.py:1541-1587), the field the push uses;
 *   mode 1: the Bmap_add_all_orders part alone (orders 1..max_order);
 *   mode 2: that part in coefficient form for a trained range zt:
 *           (sum_n Bz_n z^(2n-2) zt^2, sum_n Bfi_n z^(2n-2) zt^3,
 *            sum_n Br_n z^(2n-2) zt^3), i.e. mode 1 divided by (z/zt)^2,
 *           (z/zt)^3, (z/zt)^3;
 *   mode 3: the mode-0 field as the FP32 tier's push evaluates it (FP32
 *           arithmetic on the FP32 Taylor map), widened to double; needs a
 *           machine whose Taylor map is float (taylor_f32 = 1).
 * Used to label the surrogate's training samples and to judge its field and
 * the FP32 tier's field against the FP64 one. */
int ffa_bfield_eval(void *ctx, const double *r, const double *fi, const double *z, int64_t n, double *out,
                    int32_t mode, double zt);

/* ---------------------------------------------------------------------
 * Field surrogate (no counterpart in the reference): a Fourier-feature MLP
 * that stands in for the Taylor-map gather and z extrapolation
 * Bmap_add_all_orders (src/FFAG_MathTools.py:1139-1188) inside |z| <= zt.
 *   a = (r - r_c) * inv_rh, u = frac(fi * inv_sector), c = z / zt
 *   rev_f = m_f u + wa_f a + wc_f c^2 (f < 32, revolutions)
 *   x0 = (sin 2 pi rev_0, cos 2 pi rev_0, sin 2 pi rev_1, ...)   (64)
 *   x_{l+1} = silu(x_l W_l + b_l), l = 0..3                      (128)
 *   o = x_4 W_4 + b_4                                             (3)
 *   dBz = S_z c^2 g(o_0), dBfi = S_fi c^3 g(o_1), dBr = S_r c^3 g(o_2)
 *   g(o) = o, or delta sinh(o) when out_delta = delta > 0 (the network then
 *   learns asinh(coefficient / (S delta)): relative accuracy where the
 *   Taylor coefficients are large, absolute accuracy where they are small)
 * Parameters: FP32 master vector of FFA_MLP_NPARAM floats, W_l stored
 * [in][out] row-major, in the order W0 b0 W1 b1 W2 b2 W3 b3 W4 b4.  The
 * kernels run on a packed BF16 copy (FFA_MLP_PACKED_BYTES, zero-initialised
 * by the caller once) with FP32 accumulation on the matrix cores.
 * --------------------------------------------------------------------- */
#define FFA_MLP_NFREQ 32
#define FFA_MLP_NPARAM 58243
#define FFA_MLP_PACKED_BYTES 236160

typedef struct {
    double r_c, inv_rh;        /* radial normalisation */
    double inv_sector;         /* 1 / sector period [1/rad] */
    double zt;                 /* trained |z| range [m] */
    double out_scale[3];       /* S_z, S_fi, S_r [T] */
    double out_delta;          /* 0: linear outputs; > 0: delta sinh(o) */
    float m[FFA_MLP_NFREQ];    /* integer azimuthal harmonics */
    float wa[FFA_MLP_NFREQ];   /* radial frequencies [rev] */
    float wc[FFA_MLP_NFREQ];   /* z^2 frequencies [rev] */
    int32_t fallback;          /* push: 1 = map gather for |z| > zt, 0 = order 0 only there */
    int32_t pad;
} ffa_mlp_desc;

typedef struct {
    float lr, beta1, beta2, eps;
    int32_t step;              /* 1-based Adam step count */
    int32_t pad;
} ffa_adam;

/* Workspace of one training/gradient call on a batch of n points. */
typedef struct {
    size_t total;
    int64_t batch_padded;      /* n rounded up to FFA_MLP_KCHUNK */
    size_t xfm;                /* bf16 activations, feature-major */
    size_t dzfm;               /* bf16 backward signals, feature-major */
    size_t zs;                 /* fp32 pre-activations, wave order */
    size_t partials;           /* fp32 split-K weight-gradient partials */
    size_t loss;               /* fp32 per-wave squared-error sums */
} ffa_mlp_layout;

#define FFA_MLP_KCHUNK 512

int ffa_mlp_workspace_layout(int64_t n, ffa_mlp_layout *out);

/* params (FP32 master) -> packed BF16 layouts. */
int ffa_mlp_pack(void *ctx, const float *params, void *packed);

/* Gradient of the mean squared error over the 3 n outputs against
 * targets (3*n floats, planes o_0, o_1, o_2) at points (3*n floats,
 * planes r, fi, z).  grad: FFA_MLP_NPARAM floats; loss: 1 float (device). */
int ffa_mlp_grad(void *ctx, const ffa_mlp_desc *d, const void *packed, const float *pts, const float *targets,
                 int64_t n, void *workspace, float *grad, float *loss);

/* One Adam step (Kingma & Ba 2015, Algorithm 1) on params with moments m, v
 * from grad, then repacks. */
int ffa_mlp_adam(void *ctx, float *params, const float *grad, float *m, float *v, void *packed,
                 const ffa_adam *hp);

/* ffa_mlp_grad followed by ffa_mlp_adam without materialising grad. */
int ffa_mlp_train_step(void *ctx, const ffa_mlp_desc *d, float *params, float *m, float *v, void *packed,
                       const float *pts, const float *targets, int64_t n, void *workspace, const ffa_adam *hp,
                       float *loss);

/* Network outputs at n points (pts: 3*n floats r, fi, z): out 3*n floats,
 * planes dBz, dBfi, dBr (mode 0) or raw outputs o (mode 1). */
int ffa_mlp_eval(void *ctx, const ffa_mlp_desc *d, const void *packed, const float *pts, int64_t n, float *out,
                 int32_t mode);

/* Make ffa_push_boris_fused_f32 use the surrogate (packed != NULL) in place
 * of the Taylor-map gather for particles with |z| <= zt, or stop using it
 * (packed == NULL).  The packed buffer must stay alive while enabled. */
int ffa_mlp_set(void *ctx, const ffa_mlp_desc *d, const void *packed);

/* ---- Closed-orbit tracking in azimuth ----------------------------------
 * Replaces rk4_solve / rk4_solve_vect over FunctionForSEO(Vect)
 * (src/FFAG_track.py:582-640, 1597-1695), the tracking under the closed-orbit
 * search, the tunes and the Twiss functions (FFAG_SearchSEO,
 * src/FFAG_track.py:1142-1586; FFAG_BetaFuncCalc, src/FFAG_BetaFunc.py:45).
 * Each particle's state (r, pr, z, pz), pr and pz the direction cosines of
 * the momentum P, obeys, with theta the independent variable and
 * pf = sqrt(1 - pr^2 - pz^2),
 *   dr/dth = r pr / pf,   dpr/dth = pf - (q/P) (r Bz - r pz / pf Bfi),
 *   dz/dth = r pz / pf,   dpz/dth = (q/P) (r Br - r pr / pf Bfi),
 * in the order-0 field of the machine set by ffa_set_machine (GetBfields at
 * theta mod the sector) plus its harmonic errors (at theta mod 2 pi); no
 * Taylor maps, no RF.  Classic RK4 steps of size h; step s starts at
 * t0 + s * t_step (the reference's linspace grid). */
typedef struct ffa_orbit {
    double *state;             /* device [4][n]: r, pr, z, pz (start in, end out) */
    const double *q_over_p;    /* device [n]: q / P of each particle */
    int64_t n;
    double t0, t_step, h;
    int32_t nsteps;            /* >= 0 */
    int32_t pad;
    double *traj;              /* device [nsteps + 1][4][n] states at every grid point, or NULL */
} ffa_orbit;

int ffa_orbit_track(void *ctx, const ffa_orbit *o);

/* ---- Spiral RF-gap field maps -------------------------------------------
 * Replaces spiral_distance_direction_rphi_with_flag, the Gaussian gap
 * profile, project_to_polar_components and the div E = 0 step of
 * build_spiral_map_from_config (examples/main_Emap.py:267-445, 865-931).
 * Node (i, j) is (r_axis[i], phi_axis[j]) (continuous azimuth); outputs are
 * row-major [nr][nphi].  The centre line is phi = s_ref + a ln(r / r_ref) on
 * [r_min, r_max]; flag 1 / 2 marks nodes whose nearest line point is the
 * r_min / r_max end (no field there). */
typedef struct ffa_spiral_gap {
    const double *r_axis;      /* device [nr], uniform step dr */
    const double *phi_axis;    /* device [nphi], uniform step dphi */
    int32_t nr, nphi;
    double dr, dphi;
    double r_ref, s_ref, a;    /* a = tan(spiral angle) */
    double r_min, r_max;
    double U, delta;           /* profile amplitude and Gaussian sigma (FWHM / 2.35482) */
    double side;               /* +1 left normal, -1 right normal */
    int32_t n_iter;            /* Newton iterations (>= 0) */
    int32_t pad;
    double *er, *ephi, *ez_over_z, *dist;   /* device outputs [nr][nphi] */
    int32_t *flag;                          /* device output [nr][nphi] */
} ffa_spiral_gap;

int ffa_spiral_gap_maps(void *ctx, const ffa_spiral_gap *g);

/* ---- Taylor-coefficient map generator ----------------------------------
 * Replaces the per-radius evaluation loop of Generate_Bmaps_MultiProcess /
 * compute_taylor_coefficients_group (src/FFAG_HighOrder.py:913-998, driven
 * by examples/main_Bmap.py:48-75).  For B(r, theta) = A(r) G(theta - s(r)),
 * every map value is a short sum over flutter derivatives
 *   map[c][n](r_i, theta_j) = sum_q weights[i][c][n][q] G^(q)(theta_{j - shift_i}),
 * c = Bz, Br, Bfi; the host expands the polar Laplacian, the Leibniz and
 * chain rules into the per-radius weights.  The kernel first tabulates
 * G^(q) on the theta grid: for every pole, height * sum_i C(q, i)
 * P1_i(sigma1) P2_(q-i)(sigma2) with sigma1 = 1/(1 + exp(-k (theta - x0))),
 * sigma2 = 1/(1 + exp(k (theta - x1))) and the derivative polynomials P in
 * sigma given by the caller (sigmoid_nth_derivative, :212-243; q = 0 is
 * (height sigma1) sigma2, shape_function :246-257), then fills the maps.
 * The theta index wraps periodically over the n_phi grid points (np.roll of
 * shift_function, :393-397). */
typedef struct ffa_taylor_gen {
    int32_t n_orders;          /* max_order + 1 (1..FFA_TAYLOR_MAX_ORDERS) */
    int32_t n_q;               /* flutter derivative orders 0..n_q-1; 2 * n_orders */
    int32_t n_poles;           /* 1..FFA_TAYLOR_MAX_POLES */
    int32_t layout;            /* 0: stacks [3][n_orders][nr][n_phi], planes Bz, Br, Bfi;
                                  1: tracking layout [n_phi][nr][3 (n_orders - 1)], orders 1..,
                                     (Bz, Bfi, Br) per order (the table ffa_set_machine reads) */
    int32_t out_f32;           /* 0: double output, 1: float (rounded from the double sum) */
    int32_t pad;
    int64_t nr, n_phi;
    const double *theta;       /* device [n_phi] azimuth grid (rad) */
    const double *poles;       /* device [n_poles][4]: x0, x1, k = 16 / W0, height (sign |polarity|) */
    const double *sig_poly;    /* device [n_poles][2][n_q][n_q + 1]: coefficients of P_i (highest power
                                  first, left-padded with zeros) for the rising (k) and falling (-k) edge;
                                  row i = 0 is sigma itself (0 ... 0 1 0) */
    const double *weights;     /* device [nr][3][n_orders][n_q] */
    const int32_t *shift;      /* device [nr] grid steps, taken modulo n_phi */
    double *flutter;           /* device workspace [n_q][n_phi] (overwritten) */
    void *out;                 /* device output, layout and type above */
} ffa_taylor_gen;

#define FFA_TAYLOR_MAX_ORDERS 8
#define FFA_TAYLOR_MAX_POLES 16

int ffa_taylor_maps(void *ctx, const ffa_taylor_gen *g);

#ifdef __cplusplus
}
#endif
#endif
