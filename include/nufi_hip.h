/*
 * nufi_hip.h — C ABI of the MI355X (gfx950) Numerical Flow Iteration rho path.
 *
 * Plain C types only: sizes are uint64_t, fields are double, every call
 * returns an int status (NUFI_OK == 0) and the message of the last failure on
 * a context is available from nufi_last_error(ctx).  All device memory is
 * owned by the context; the caller owns every host pointer it passes in.
 *
 * Reference interface that each entry point replaces (paulwilhelmvlasov/
 * NumericalFlowIteration, files relative to the repository root):
 *   nufi_config              nufi/config.hpp:33-53 (dim1::config_t),
 *                            nufi/config.hpp:167-194 (dim3::config_t)
 *   nufi_create              nufi/cuda_scheduler.hpp:43-85 (ctors),
 *                            nufi/cuda_kernel.cu:81-109, 468-483
 *   nufi_compute_rho         nufi/cuda_scheduler.hpp:88-111,
 *                            nufi/cuda_kernel.cu:112-132, 485-510
 *   nufi_download_rho        nufi/cuda_scheduler.hpp:113-118,
 *                            nufi/cuda_kernel.cu:134-145, 512-523
 *   nufi_upload_phi          nufi/cuda_scheduler.hpp:120-125,
 *                            nufi/cuda_kernel.cu:147-156, 525-536
 *   nufi_compute_metrics     nufi/cuda_scheduler.hpp:128-151,
 *                            nufi/cuda_kernel.cu:158-176
 *   nufi_download_metrics    nufi/cuda_scheduler.hpp:153-157,
 *                            nufi/cuda_kernel.cu:178-189
 *   nufi_poisson_solve       nufi/poisson.cpp:66-89 (dim1), :328-362 (dim3)
 *   nufi_interpolate         nufi/fields.hpp:63-142 (dim1), :352-490 (dim3)
 *   nufi_step / nufi_run     the driver step bin/test_nufi_gpu_1d.cpp:241-253
 *                            and bin/test_nufi_gpu_3d.cpp:150-162, made
 *                            device-resident (no host round trip per step)
 */
#ifndef NUFI_HIP_H
#define NUFI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum nufi_status {
    NUFI_OK              = 0,
    NUFI_ERR_INVALID     = 1, /* bad argument / config                     */
    NUFI_ERR_RANGE       = 2, /* n > Nt, q range, too many blocks
                                 (std::range_error in the reference)       */
    NUFI_ERR_DEVICE      = 3, /* HIP runtime failure (cuda::exception)     */
    NUFI_ERR_ALLOC       = 4, /* device allocation failure                 */
    NUFI_ERR_UNSUPPORTED = 5, /* order / dim / precision not built         */
    NUFI_ERR_IO          = 6  /* history file: open, read, write or format */
};

/* ---- precision of the characteristic trace ----------------------------- */
enum nufi_precision {
    NUFI_FP64  = 0, /* the plain FP64 method                                */
    NUFI_FP32C = 1  /* FP32 trace with compensated position and velocity
                       updates (hi + lo pairs, error-free Fast2Sums);
                       Poisson, fit and rho stay FP64                       */
};

/* ---- initial conditions (config_t::f0) --------------------------------- */
enum nufi_f0_kind {
    /* 1D families; M(u) = exp(-u^2/2)/sqrt(2 pi)                            */
    NUFI_F0_LANDAU_1D     = 1, /* (1 + alpha cos(k x)) M(u)                 */
    NUFI_F0_TWOSTREAM_1D  = 2, /* (1 + alpha cos(k x)) u^2 M(u)
                                  (the dim1 default, config.hpp:80-82)      */
    NUFI_F0_MULTIMODE_1D  = 3, /* (1 + sum_m a_m cos(m k x)) M(u), m=1..n_modes */
    /* 3D families; M3(v) = exp(-|v|^2/2)/(2 pi)^(3/2)                        */
    NUFI_F0_LANDAU_3D     = 11, /* (1 + alpha sum_d cos(k x_d)) M3(v)
                                   (config.hpp:233-234)                     */
    NUFI_F0_BUMPONTAIL_3D = 12, /* the dim3 default, config.hpp:244-246     */
    NUFI_F0_TWOSTREAM_3D  = 13  /* config.hpp:237-240                        */
};

#define NUFI_MAX_MODES 8

typedef struct nufi_f0 {
    int32_t kind;                  /* enum nufi_f0_kind                     */
    int32_t n_modes;               /* NUFI_F0_MULTIMODE_1D only             */
    double  alpha;                 /* perturbation amplitude                */
    double  k;                     /* perturbation wave number              */
    double  a[NUFI_MAX_MODES];     /* multimode amplitudes a_1..a_n_modes   */
} nufi_f0;

/*
 * Grid description: the fields of dim3::config_t (config.hpp:167-194); a 1D
 * config has dim = 1 and ignores the y, z, v, w fields.  Derived fields
 * (L*, d*, *_inv, du, dv, dw) are filled by nufi_config_derive in the order
 * config.hpp:67-69 / 213-218 uses.
 */
typedef struct nufi_config {
    int32_t  dim;                  /* 1 or 3                                */
    int32_t  reserved;
    uint64_t Nx, Ny, Nz;           /* grid points in physical space         */
    uint64_t Nu, Nv, Nw;           /* velocity quadrature nodes per axis    */
    uint64_t Nt;                   /* number of time steps                  */
    double   dt;
    double   x_min, x_max, y_min, y_max, z_min, z_max;
    double   u_min, u_max, v_min, v_max, w_min, w_max;
    double   dx, dx_inv, Lx, Lx_inv;
    double   dy, dy_inv, Ly, Ly_inv;
    double   dz, dz_inv, Lz, Lz_inv;
    double   du, dv, dw;
    nufi_f0  f0;
} nufi_config;

typedef struct nufi_ctx nufi_ctx;

/* ---- library ----------------------------------------------------------- */
const char *nufi_version(void);
/* Fills the derived fields of *conf from the primary ones. */
int nufi_config_derive(nufi_config *conf);
int nufi_device_count(int *count);

/* ---- context ----------------------------------------------------------- */
/*
 * Allocates, on `device`, the history of (Nt+1) coefficient slices of
 * (Nx+order-1)^dim reals each, rho and metrics buffers (cuda_kernel.cu:81-109),
 * plus per-slice tables derived from each slice as it is stored: in 1D the
 * first differences (Nx+order+3 rounded to 4 reals) and, at order 4, the
 * kick-polynomial table (3(Nx+2) rounded to 4 reals, FP32 for NUFI_FP32C);
 * in 3D FP32C an FP32 copy of the history.
 * conf_metrics may be NULL (then conf is used for metrics, as the dim3
 * scheduler does).  Errors: NUFI_ERR_UNSUPPORTED for an order or dim that is
 * not built, NUFI_ERR_ALLOC / NUFI_ERR_DEVICE.  On failure *out is NULL and
 * nufi_last_error(NULL) gives the reason.
 */
int nufi_create(const nufi_config *conf, const nufi_config *conf_metrics,
                int order, int precision, int device, nufi_ctx **out);
int nufi_destroy(nufi_ctx *ctx);
/* Message of the last failure on ctx (ctx == NULL: last failed create on
 * this thread).  Never NULL. */
const char *nufi_last_error(const nufi_ctx *ctx);
/* The HIP stream (hipStream_t) every kernel of ctx runs on. */
int nufi_get_stream(nufi_ctx *ctx, void **stream);
int nufi_synchronize(nufi_ctx *ctx);

/* ---- the reference scheduler contract ---------------------------------- */
/*
 * rho partials of nodes q in [q_begin, q_end) at time step n, from slices
 * 0..n-1 of the device history.  Node q is (ix = q / Nvel, iv = q % Nvel)
 * with Nvel = Nu (1D) or Nu*Nv*Nw with u innermost (3D).  Each node adds
 * -dv * ftilda(n, x, v) to rho[ix] (no "+1": cuda_kernel.cu:50, 425).
 * Asynchronous on the context stream.  NUFI_ERR_RANGE for n > Nt or
 * q_end > total nodes; q_begin == q_end is a no-op.
 */
int nufi_compute_rho(nufi_ctx *ctx, uint64_t n, uint64_t q_begin, uint64_t q_end);
/* Synchronises and ADDS the device rho (Nx*Ny*Nz reals) into rho_accum. */
int nufi_download_rho(nufi_ctx *ctx, double *rho_accum);
/* Copies slice n from history_base + n*stride_t (host) to the device. */
int nufi_upload_phi(nufi_ctx *ctx, uint64_t n, const double *history_base);
/*
 * Metrics [int f, int f^2, kinetic energy, entropy] of nodes [q_begin,q_end)
 * of the metrics grid at step n (with the initial half-kick, rho.hpp:63-96).
 */
int nufi_compute_metrics(nufi_ctx *ctx, uint64_t n, uint64_t q_begin, uint64_t q_end);
/* Synchronises and ADDS the 4 metrics into metrics_accum. */
int nufi_download_metrics(nufi_ctx *ctx, double *metrics_accum);

/* ---- device-resident step loop ----------------------------------------- */
/*
 * One full time step on the device: rho(n) from slices 0..n-1, periodic
 * Poisson solve, spline fit, slice n appended to the device history.
 * *energy_out (may be NULL) receives the spectral electric energy
 * (poisson.cpp:73-88) — this synchronises; with NULL the step stays async.
 */
int nufi_step(nufi_ctx *ctx, uint64_t n, double *energy_out);
/* Steps n_begin..n_end-1; energies (may be NULL) gets n_end-n_begin values.
 * The steps' launches are captured into a hipGraph on the first call for a
 * step range (and launch shape) and replayed on the next ones, unless
 * nufi_set_run_graph(ctx, 0) or the trace launches are timed
 * (nufi_set_timing); the results are the same either way. */
int nufi_run(nufi_ctx *ctx, uint64_t n_begin, uint64_t n_end, double *energies);
/* No counterpart in the reference (whose driver launches each step from the
 * host, bin/test_nufi_gpu_1d.cpp:241-253): whether nufi_run replays a graph
 * (default 1).  A context created with NUFI_RUN_GRAPH=0 in the environment
 * never replays one, whatever this is set to (graph launches are not traced
 * by the profiler runs of this build). */
int nufi_set_run_graph(nufi_ctx *ctx, int enabled);
/* Host copies of history slices [n_begin, n_end). */
int nufi_download_history(nufi_ctx *ctx, uint64_t n_begin, uint64_t n_end, double *out);
int nufi_upload_history(nufi_ctx *ctx, uint64_t n_begin, uint64_t n_end, const double *in);
/* Energies of steps [n_begin, n_end) computed by nufi_step/nufi_run. */
int nufi_download_energies(nufi_ctx *ctx, uint64_t n_begin, uint64_t n_end, double *out);

/* ---- field solves on host buffers -------------------------------------- */
/* In place: rho -> nodal phi; *energy (may be NULL) = spectral 1/2 int E^2. */
int nufi_poisson_solve(nufi_ctx *ctx, double *data, double *energy);
/* Nodal values (Nx^dim) -> padded coefficient slice ((Nx+order-1)^dim). */
int nufi_interpolate(nufi_ctx *ctx, double *slice, const double *values);

/* ---- trace kernel selection and launch shape -------------------------- */
enum nufi_trace_variant {
    NUFI_TRACE_FAST   = 0, /* FP64: cell-unit positions, FMA (order 4)       */
    NUFI_TRACE_STRICT = 1, /* FP64 in the reference's operation order
                              (rho.hpp:46-58, fields.hpp:42-60,
                              splines.hpp:98-109), no FMA contraction        */
    NUFI_TRACE_FP32C  = 2  /* set by precision NUFI_FP32C                     */
};
/* Switch an FP64 context between FAST and STRICT (no counterpart in the
 * reference, which has one kernel; used by parity tests). */
int nufi_set_trace_variant(nufi_ctx *ctx, int variant);
/* Workgroup size, nodes per thread and history slices per LDS chunk of the
 * trace kernel (what the autotuner sets); NUFI_ERR_UNSUPPORTED if the shape
 * is not built or does not fit the LDS.  1D FAST and FP32C kernels take
 * (block/64)*nodes_per_thread a multiple of 16 (whole velocity groups per
 * workgroup) and treat the chunk as an upper bound: the library uses the
 * largest chunk whose tables two LDS buffers hold (one LDS buffer per chunk
 * for wider grids).  The FAST
 * and FP32C results do not depend on the shape (fixed summation groups and a
 * re-wrap schedule fixed by the slice index), so tuning never changes a
 * result. */
int nufi_set_launch_shape(nufi_ctx *ctx, int block, int nodes_per_thread, int chunk);
int nufi_get_launch_shape(nufi_ctx *ctx, int *block, int *nodes_per_thread, int *chunk);
/* Autotuner (no counterpart in the reference, whose launch shape is fixed at
 * 64 threads, cuda_kernel.cu:123-124): times every built shape that fits the
 * LDS on the trace kernel at step n_probe over all nodes (`repeats` launches
 * each, HIP events on the context stream, history as currently on the device),
 * keeps the fastest and returns it.  Results are overwritten partials only. */
int nufi_autotune(nufi_ctx *ctx, uint64_t n_probe, int repeats, int *block, int *nodes_per_thread,
                  int *chunk, double *best_ms);

/* ---- timing of the trace kernel ---------------------------------------- */
/*
 * When enabled, every trace-kernel launch is bracketed by HIP events on the
 * context stream; nufi_trace_timing reports launches, summed kernel ms and
 * summed characteristic steps since the last reset.
 */
int nufi_set_timing(nufi_ctx *ctx, int enabled);
int nufi_trace_timing(nufi_ctx *ctx, uint64_t *launches, double *kernel_ms,
                      double *char_steps);

/* ---- adjoint ------------------------------------------------------------ */
/*
 * No counterpart in the reference.  Runs the whole forward loop (steps 0..Nt)
 * and a reverse sweep through the time steps that recomputes every trace
 * (nothing of the foot-point paths is stored), giving the gradient of
 * J = sum_{n=0}^{Nt} dt * energy_n (energy_n as returned by the Poisson solve,
 * poisson.cpp:73-88) with respect to the parameters of f0: a_1..a_n_modes for
 * NUFI_F0_MULTIMODE_1D, alpha for the Landau and two-stream families.
 * dJ_da receives n_par >= that count values; *J_out (may be NULL) receives J.
 * 1D, order 4; the history is overwritten by the forward run.  An FP32C
 * context differentiates its own forward model, the FP32-compensated trace
 * (the reverse sweep undoes its steps on the same hi + lo pairs; the adjoint
 * variables, the slice adjoints and the field adjoint are FP64).
 */
int nufi_adjoint_grad(nufi_ctx *ctx, double *dJ_da, int n_par, double *J_out);

/* ---- rho without allocation ------------------------------------------- */
/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): rho of time step n of a 3D order-4 FP64 grid, from slices
 * 0..n-1 of a device history the caller owns (the reference layout: slice m
 * at hist_dev + m*stride_t, stride_t = (Nx+3)(Ny+3)(Nz+3); hist_dev 16-byte
 * aligned, hist_len >= n*stride_t + 1 reals) into rho_dev (Nx*Ny*Nz reals,
 * device, overwritten with -du dv dw sum_v f: what nufi_compute_rho over all
 * nodes then nufi_download_rho into zeros gives, bit for bit).  Asynchronous
 * on `stream` (a hipStream_t; NULL is the default stream) of the current
 * device.  Allocates nothing: the partial sums live in one fixed device array
 * of NUFI_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; grids
 * whose partials do not fit it (ceil(Nvel/vgroup)*Nx*Ny*Nz reals, vgroup = 256
 * velocities per partial at 16^3 x 32^3) give NUFI_ERR_UNSUPPORTED.  Every call
 * on one GPU shares that array: calls that overlap on one GPU race.  Errors
 * are reported by nufi_last_error(NULL).
 */
#define NUFI_NOALLOC_SCRATCH_BYTES ( 8u << 20 )
int nufi_rho3d_noalloc(const nufi_config *conf, uint64_t n, const double *hist_dev, uint64_t hist_len,
                       double *rho_dev, void *stream);

/* ---- batched sweep: many independent 1D runs on one device --------------- */
/*
 * No counterpart in the reference (which runs one configuration per
 * programme): n_runs independent 1D configs (any Nx, Nu, domain, f0; one dt)
 * stepped together.  Each time step is one trace launch over the wave-tasks of
 * every run plus one field launch (a workgroup per run); nufi_batch_run with
 * use_graph != 0 captures the step range into a hipGraph once and replays it.
 * Each run's history and energies are bit-identical to a single-run context of
 * the same config and precision.  Order 4; runs past their own Nt idle.
 */
typedef struct nufi_batch nufi_batch;
int nufi_batch_create(const nufi_config *confs, uint64_t n_runs, int order, int precision, int device,
                      nufi_batch **out);
int nufi_batch_destroy(nufi_batch *b);
const char *nufi_batch_last_error(const nufi_batch *b);
/* Steps [n_begin, n_end) of every run (asynchronous); use_graph is ignored
 * (steps launched one by one) for a batch created with NUFI_RUN_GRAPH=0 in
 * the environment. */
int nufi_batch_run(nufi_batch *b, uint64_t n_begin, uint64_t n_end, int use_graph);
int nufi_batch_synchronize(nufi_batch *b);
int nufi_batch_energies(nufi_batch *b, uint64_t run, uint64_t n_begin, uint64_t n_end, double *out);
int nufi_batch_history(nufi_batch *b, uint64_t run, uint64_t n_begin, uint64_t n_end, double *out);
int nufi_batch_set_launch_shape(nufi_batch *b, int block, int nodes_per_thread, int chunk);
int nufi_batch_get_launch_shape(nufi_batch *b, int *block, int *nodes_per_thread, int *chunk);
/* Autotuner over the batch trace launch at step n_probe (all runs). */
int nufi_batch_autotune(nufi_batch *b, uint64_t n_probe, int repeats, int *block, int *nodes_per_thread,
                        int *chunk, double *best_ms);

/* ---- diagnostics samplers (1D) ---------------------------------------- */
/*
 * The plotting block of bin/test_nufi_gpu_1d.cpp:290-361 on the device, at
 * `count` arbitrary points (host arrays in and out), from the device history:
 *   nufi_sample_field  E = -dphi/dx and rho = -d^2phi/dx^2 of slice n
 *                      (dim1::eval<order,1|2>, fields.hpp:36-61)
 *   nufi_sample_f      f = eval_f(n, x, v) (rho.hpp:63-96: the trace with the
 *                      initial half-kick from slice n; slices 0..n are read)
 *   nufi_phase_flow    eval_phase_flow(n, x, v) (rho.hpp:98-131): the foot
 *                      point of that trace for n > 1 — for n <= 1 the point
 *                      is not moved, as in the reference — with x wrapped by
 *                      x - Lx floor(x / Lx)
 * The reference's operation order (no FMA contraction), any built order.
 * NUFI_ERR_RANGE for n > Nt, NUFI_ERR_UNSUPPORTED for 3D contexts.
 */
int nufi_sample_field(nufi_ctx *ctx, uint64_t n, const double *x, uint64_t count, double *E, double *rho);
int nufi_sample_f(nufi_ctx *ctx, uint64_t n, const double *x, const double *v, uint64_t count, double *f);
int nufi_phase_flow(nufi_ctx *ctx, uint64_t n, const double *x, const double *v, uint64_t count,
                    double *x_out, double *v_out);

/* ---- history files: checkpoint / resume and isolated steps ------------- */
/*
 * The coefficient history (slices 0..n_slices-1, each stride_t reals, the
 * layout of the device history) on disk.  Two formats:
 *   NUFI_HISTORY_BINARY  lossless: the 8 bytes "NUFIHST1", a fixed header
 *                        (little-endian: u32 dim, u32 order, u64 Nx, Ny, Nz,
 *                        Nt, n_slices, stride_t, f64 dt, u64 n_energies), the
 *                        slices as f64, then n_energies f64 energies.
 *   NUFI_HISTORY_TEXT    the layout of the reference's potential files
 *                        (bin/test_nufi_cpu_3d_isolated.cpp:128-135,
 *                        158-162, 190-211): the lines "Nt = ", "dt = ",
 *                        "Nx = ", "Ny = ", "Nz = ", "order = ", a blank line,
 *                        then one value per line — written here with 17
 *                        significant digits, so it reads back exactly; files
 *                        the reference wrote (16 digits) read to within one
 *                        unit in the 16th digit.  No energies.
 * The file functions are host-only (no device is touched).
 */
enum nufi_history_format {
    NUFI_HISTORY_BINARY = 0,
    NUFI_HISTORY_TEXT   = 1
};
typedef struct nufi_history_info {
    int      dim, order;
    uint64_t Nx, Ny, Nz;       /* Ny = Nz = 1 in 1D                         */
    uint64_t Nt;               /* the run's last step                        */
    uint64_t n_slices;         /* slices stored: 0..n_slices-1               */
    uint64_t stride_t;         /* reals per slice: prod (N_d + order - 1)    */
    double   dt;
    uint64_t n_energies;       /* energies stored (binary only)             */
} nufi_history_info;
/* Write n_slices slices (and info->n_energies energies, binary only; may be
 * NULL when n_energies is 0). */
int nufi_history_write(const char *path, int format, const nufi_history_info *info,
                       const double *slices, const double *energies);
/* Read the header (format detected from the first bytes). */
int nufi_history_read_info(const char *path, nufi_history_info *info, int *format);
/* Read the header and up to slice_capacity slices / energy_capacity energies. */
int nufi_history_read(const char *path, nufi_history_info *info, double *slices, uint64_t slice_capacity,
                      double *energies, uint64_t energy_capacity);
/* Checkpoint: slices 0..n_slices-1 of the context's device history (and the
 * energies of those steps, binary) to a file. */
int nufi_save_history(nufi_ctx *ctx, const char *path, int format, uint64_t n_slices);
/* Resume / isolated steps: load a file written for the same grid (dim,
 * order, Nx, Ny, Nz, dt; file Nt <= context Nt) into the device history (the
 * tables of every slice are rebuilt, energies restored when stored).
 * *n_slices receives the slice count: nufi_run(ctx, *n_slices, Nt + 1, ...)
 * continues the run, nufi_step(ctx, n, ...) with n < *n_slices recomputes an
 * isolated step from the stored slices 0..n-1
 * (bin/test_nufi_cpu_3d_isolated.cpp:169-228). */
int nufi_load_history(nufi_ctx *ctx, const char *path, uint64_t *n_slices);

#ifdef __cplusplus
}
#endif

#endif /* NUFI_HIP_H */
