// internal.hpp — host-side declarations shared by the translation units of
// libnufi_hip.so (context layout, kernel parameter blocks, launchers).
#ifndef NUFI_INTERNAL_HPP
#define NUFI_INTERNAL_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include <string>
#include <vector>

#include "../../include/nufi_hip.h"

namespace nufi_impl
{

// Trace kernel variants.
enum trace_variant : int
{
    TRACE_FAST   = 0, // FP64, cell-unit positions, FMA, closed-form spline
    TRACE_STRICT = 1, // FP64 in the reference's operation order, no contraction
    TRACE_FP32C  = 2  // FP32 with compensated (two-sum) position/velocity
};

// ---- 1D trace ----------------------------------------------------------------
// Node q = ix*Nu + iu (cuda_kernel.cu:39-44); nodes with q outside
// [q_begin, q_end) contribute zero.  Two decompositions:
//
// velocity-major (strict and metrics kernels): the launch is cut into
// wave-tasks; task g covers ix = ix_first + g / wpx and iu = (g % wpx)*64 +
// lane, and writes one partial sum of f to partials[(g % wpx)*Nx + ix].
//
// x-major (fast and FP32C rho kernels): the 64 lanes of a wave take xw
// consecutive x nodes (xw = min(Nx, 64)) at vpw = 64/xw velocities; lanes at
// one velocity and neighbouring x stay close along their characteristics, so
// the table reads of a wave are near-contiguous in LDS.  Workgroup b covers
// x-tile b % n_xt (x nodes ix_first + xw*(b % n_xt) + 0..xw-1) and velocity
// block b / n_xt of vb = NPT*WPB*vpw velocities; thread (wave w, lane) takes
// the velocity rows (r*WPB + w)*vpw + lane/xw, r < NPT.  The f values are
// staged in LDS and summed over fixed groups of XM_GROUP consecutive
// velocities (a sequential compensated sum in iu order) into
// partials[(iu/XM_GROUP)*Nx + ix], ppx = ceil(Nu/XM_GROUP) groups: the
// partials do not depend on the launch shape.  All rho partials are stored
// group-major (partials[j*n_space + i]), so both their writers and the
// field kernel's fixed-order sum over j read and write coalesced.
constexpr uint32_t XM_GROUP = 16;
inline uint32_t xm_xw( uint32_t Nx ) { return Nx < 64 ? Nx : 64; }
inline uint32_t xm_ppx( uint32_t Nu ) { return ( Nu + XM_GROUP - 1 ) / XM_GROUP; }

// Trace kernel outputs.
enum trace_mode : int
{
    MODE_RHO     = 0, // f without the initial half-kick (rho.hpp:31-61), one partial per wave
    MODE_METRICS = 1  // f with it (rho.hpp:63-96), four weighted partials per wave
                      // [f, f^2, |v|^2 f / 2, -f log f] (cuda_kernel.cu:53-79, 428-466)
};

struct trace1_params
{
    const double *dhist;     // extended differences of slice m at dhist + m*stride_d (FP64)
    const double *qhist;     // FP64 rho: kick-polynomial tables, slice m at qhist + m*stride_q, LDS layout
    const float  *qhist32;   // FP32C rho: the same, FP32 tables scaled by fp32c1_scale
    uint32_t stride_q, q2_off;        // LDS kick-polynomial table: slice stride, Q2 row (poly_stride, poly_q2)
    double   kick;                    // -dt/(2 dx), the scale of the kick polynomials
    uint32_t pipe;                    // x-major: tables double-buffered (poly_ready_pipeline)
    double       *partials;  // velocity-major: partials[j*Nx + ix], j < wpx; x-major: j < ppx;
                             // MODE_METRICS: partials[4*g + k]
    uint64_t n;
    uint64_t q_begin, q_end;
    uint32_t Nx, Nu, stride_d, wpx;   // Nu / wpx: of the node grid
    uint32_t ix_first, n_tasks;
    uint32_t ix_count, xw, vpw, n_xt, ppx;   // x-major: x range, tile shape, partials per x
    uint32_t chunk;          // slices per LDS chunk
    uint32_t order;
    int      mode;           // trace_mode
    double x_min, dx, dx_inv, Lx, Lx_inv, dt;    // the field grid
    double nx_min, ndx, nu_min, ndu;             // the node grid (= field grid for rho)
    double weight;           // MODE_METRICS quadrature weight
    float f32_p0, f32_w0;    // FP32C: offsets of the position and velocity hi parts (fp32c_frame)
    nufi_f0 f0;
};

// ---- 3D trace ----------------------------------------------------------------
// Node q = l*Nvel + iv with l = (iz*Ny + iy)*Nx + ix and iv = (iw*Nv + iv')*Nu
// + iu (cuda_kernel.cu:399-412).  The strict, FP32C and metrics kernels use
// the wave-task cut of the 1D case with Nu -> Nvel (partials[j*n_space + l],
// j = iv/64); the fast rho kernel the x-major layout of trace3d.hip (64
// neighbouring spatial points per wave, partials[(iv/vgroup)*n_space + l]).
// Velocities per x-major partial: 256, halved (down to 16) until the launch
// has at least 8192 waves; fixed per grid, so results do not depend on the shape.
inline uint32_t xm3_vgroup( uint64_t n_space, uint64_t n_vel )
{
    uint32_t g = 256;
    const uint64_t tiles = ( n_space + 63 ) / 64;
    while ( g > 16 && tiles * ( ( n_vel + g - 1 ) / g ) < 8192 ) g /= 2;
    return g;
}
struct trace3_params
{
    const double *hist;      // coefficient slices (FP64), stride st
    const float  *hist32;    // FP32 copy (TRACE_FP32C)
    double       *partials;  // MODE_RHO: partials[j*n_space + l]; MODE_METRICS: partials[4*g + k]
    int      mode;
    double   weight;
    uint64_t n;
    uint64_t q_begin, q_end;
    uint32_t Nx, Ny, Nz, Nu, Nv, Nw;
    uint32_t sy, sz, st;     // strides of a slice: sy = Nx+order-1, sz = sy*(Ny+order-1)
    uint32_t st32;           // FP32C: slice stride of hist32, st rounded up to 4 floats (16-byte slices)
    uint32_t dma;            // fast / FP32C rho: two LDS buffers filled by direct global->LDS loads
    uint32_t wpx;            // waves per spatial point: ceil(Nvel/64)
    uint32_t l_first, n_tasks;
    uint32_t n_count, n_st;  // x-major: spatial points of the range, 64-point tiles
    uint32_t vgroup, ppx;    // x-major: velocities per partial sum, partials per point
    uint32_t chunk;
    uint32_t order;
    double x_min, y_min, z_min, dx, dy, dz, dx_inv, dy_inv, dz_inv;
    double Lx, Ly, Lz, Lx_inv, Ly_inv, Lz_inv;
    double u_min, v_min, w_min, du, dv, dw, dt;
    double kick72[3], kick72h[3]; // fast FP64: -dt d_inv / 72 per axis (72-scaled gradient), and half of it
    float f32_p0[3], f32_w0[3];   // FP32C: offsets of the position and velocity hi parts per axis
    nufi_f0 f0;
};

// FP32C frames of one axis (period N cells, velocity nodes v_lo + (i + 1/2) dv,
// i < nv, dt * d_inv cells per step per unit velocity): W0 = 1.5 * 2^j and
// P0 = 1.5 * 2^k with 2^(j-1) >= wcap (a bound on |velocity| in cells per
// step, twice the grid's with one cell to spare for the kicks) and
// 2^(k-1) >= N + 17 wcap + 2 (a wrapped position plus the drift of the 16
// slices between re-wraps): W0 + w and P0 + s then stay in one binade each,
// [2^j, 2^(j+1)) and [2^k, 2^(k+1)), and P0 is a multiple of N when N is a
// power of two.  The 1D kernels drift the hi parts by an ordered Fast2Sum.
// shared (3D): W0 = P0, both in [2^k, 2^(k+1)): every hi part is then a
// multiple of one ulp, 2^(k-23), and the drift sh - (wh - W0) is exact (no
// error term), at a velocity grid 2^(k-j) times coarser in its hi part.  Over
// the 160 steps of 3D Landau that changes no error measurably; over the 1600
// of the 1D two-stream instability it raises the energy error about fivefold,
// so the 1D kernels keep separate offsets.
inline void fp32c_frame( double N, double cd, double v_lo, double nv, double dv, float &p0, float &w0,
                         bool shared = false )
{
    const double vmax = std::max( std::fabs( v_lo ), std::fabs( v_lo + nv * dv ) );
    const double wcap = 2.0 * cd * vmax + 1.0;
    int j = 1, k = 1;
    while ( std::ldexp( 1.0, j - 1 ) < wcap ) ++j;
    while ( std::ldexp( 1.0, k - 1 ) < N + 17.0 * wcap + 2.0 ) ++k;
    p0 = float( 1.5 * std::ldexp( 1.0, k ) );
    w0 = shared ? p0 : float( 1.5 * std::ldexp( 1.0, j ) );
}

// The field and velocity grid of a 3D trace launch from a config.
inline void fill_trace3_geometry( trace3_params &p, const nufi_config &g )
{
    p.x_min = g.x_min; p.y_min = g.y_min; p.z_min = g.z_min;
    p.dx = g.dx; p.dy = g.dy; p.dz = g.dz;
    p.dx_inv = g.dx_inv; p.dy_inv = g.dy_inv; p.dz_inv = g.dz_inv;
    p.Lx = g.Lx; p.Ly = g.Ly; p.Lz = g.Lz;
    p.Lx_inv = g.Lx_inv; p.Ly_inv = g.Ly_inv; p.Lz_inv = g.Lz_inv;
    p.u_min = g.u_min; p.v_min = g.v_min; p.w_min = g.w_min;
    p.du = g.du; p.dv = g.dv; p.dw = g.dw; p.dt = g.dt;
    const double dinv[3] = { g.dx_inv, g.dy_inv, g.dz_inv };
    const double vlo[3] = { g.u_min, g.v_min, g.w_min }, dv[3] = { g.du, g.dv, g.dw };
    const double nv[3] = { double( g.Nu ), double( g.Nv ), double( g.Nw ) }, nx[3] = { double( g.Nx ), double( g.Ny ), double( g.Nz ) };
    for ( int a = 0; a < 3; ++a )
    {
        p.kick72[a]  = -g.dt * dinv[a] / 72.0;
        p.kick72h[a] = 0.5 * p.kick72[a];
        fp32c_frame( nx[a], g.dt * dinv[a], vlo[a], nv[a], dv[a], p.f32_p0[a], p.f32_w0[a], true );
    }
    p.f0 = g.f0;
}

// NUFI_OK if the config and order describe a grid the library accepts (capi.hip).
int validate_config( const nufi_config *c, int order );
// The message nufi_last_error(NULL) returns on this thread (capi.hip).
void set_context_free_error( const std::string &msg );

// One run of a batched sweep (device-resident parameter block).
struct batch_run
{
    trace1_params p;          // n is set per step
    uint64_t Nt;
    uint32_t wg_begin, n_wg;  // the run's workgroups in the batch trace launch
};

struct launch_shape
{
    int block;   // threads per workgroup
    int npt;     // nodes (wave-task rows) per thread
    int chunk;   // history slices per LDS chunk
};

// Launchers (trace1d.hip / trace3d.hip).  Return hipSuccess or the error.
hipError_t launch_trace1d( const trace1_params &p, int variant, const launch_shape &s, hipStream_t st );
hipError_t launch_trace3d( const trace3_params &p, int variant, const launch_shape &s, hipStream_t st );
size_t trace1d_lds_bytes( const trace1_params &p, int variant, const launch_shape &s );
size_t trace1d_slice_elems( const trace1_params &p, int variant );
hipError_t launch_trace1d_batch( const batch_run *runs, const uint32_t *wg_run, uint32_t n_wg, uint64_t n,
                                 int variant, const launch_shape &s, size_t lds, hipStream_t st );
size_t trace3d_lds_bytes( const trace3_params &p, int variant, const launch_shape &s );
// 3D FP32C rho: whether two LDS buffers of `chunk` slices of st32 floats fit 160 KiB
inline bool fp32c3_dma( uint32_t chunk, uint32_t st32 )
{
    return 2ull * chunk * st32 * sizeof(float) <= 160u * 1024u;
}
// 3D fast FP64 rho: whether two LDS buffers of `chunk` slices (+ 2 doubles for
// the 8-byte slice alignment) fit 160 KiB
inline bool fast3_dma( uint32_t chunk, uint32_t st )
{
    return 2ull * ( uint64_t( chunk ) * st + 2 ) * sizeof(double) <= 160u * 1024u;
}
bool trace1d_shape_ok( int variant, int order, const launch_shape &s );
// x-major fields (xw, vpw, n_xt, ppx) from Nx, Nu and ix_count, the grid of
// a launch (x-major or wave-task decomposition, by variant and mode) and the
// slices per LDS chunk it uses (bounded by the step and, x-major, planned by
// poly_chunk_plan within lds_max bytes, *pipe telling which loop it takes)
void trace1d_xm_setup( trace1_params &p );
uint32_t trace1d_grid( const trace1_params &p, int variant, const launch_shape &s );
uint32_t trace1d_chunk( const trace1_params &p, int variant, const launch_shape &s, size_t lds_max = 160 * 1024,
                        bool *pipe = nullptr );
bool trace3d_shape_ok( int variant, int order, const launch_shape &s );

// ---- adjoint (adjoint1d.hip) ---------------------------------------------------
struct adjoint1_params
{
    const double *qhist;      // kick-polynomial tables (FP64), slice m at qhist + m*stride_q
    const float  *qhist32;    // FP32C: the FP32 tables scaled by the kick (qhist is then null)
    const double *g;          // dJ/drho_n (Nx)
    double       *lambda;     // folded slice adjoints, slice m at lambda + m*Nx
    double       *grad_part;  // per wave-task parameter-gradient accumulators [g][NUFI_MAX_MODES]
    uint64_t n;
    uint32_t Nx, Nu, stride_d, stride_q, q2_off, wpx, n_tasks, chunk, npar;
    double x_min, dx, dx_inv, dt, u_min, du;
    float f32_p0, f32_w0;     // FP32C: the trace kernel's frames (fp32c_frame)
    nufi_f0 f0;
};

struct adjoint_field_params
{
    uint32_t N, log2x;
    const double  *rho, *lambda;   // rho_n and lambda_n (Nx)
    double        *g;              // out: dJ/drho_n
    const double2 *tw, *mult;      // twiddles, step multiplier
    const double  *ewt;
    double dt;
};

hipError_t launch_adjoint1d( const adjoint1_params &p, hipStream_t st );
size_t adjoint1d_lds_bytes( const adjoint1_params &p );
hipError_t launch_adjoint_field( const adjoint_field_params &p, hipStream_t st );
hipError_t launch_adjoint_field_large( const adjoint_field_params &p, double2 *W0, double2 *W1, hipStream_t st );
hipError_t launch_reduce_grad( const double *part, uint32_t n_tasks, uint32_t npar, double *grad, hipStream_t st );

// ---- diagnostics samplers (sample1d.hip) -------------------------------------
enum sample_what : int
{
    SAMPLE_FIELD = 0, // out0 = E = -phi'(x), out1 = -phi''(x) of slice n
    SAMPLE_F     = 1, // out0 = eval_f(n, x, v) (rho.hpp:63-96)
    SAMPLE_FLOW  = 2  // (out0, out1) = eval_phase_flow(n, x, v) (rho.hpp:98-131)
};

struct sample1_params
{
    const double *dhist;     // extended differences, slice m at dhist + m*stride_d
    const double *x, *v;     // the points (device)
    double *out0, *out1;
    uint64_t n, count;
    uint32_t stride_d;
    double x_min, Lx, Lx_inv, dx_inv, dt;
    nufi_f0 f0;
};
hipError_t launch_sample1d( const sample1_params &p, int what, int order, hipStream_t st );

// ---- field solve (field.hip) -------------------------------------------------
// Input rho either from the partials (sum_j partials[j*N + i] in order j = 0..,
// times -dv), or from a plain array.  Output: the padded coefficient slice, its
// differences (1D) and FP32 copies, the energy, and optionally the nodal phi.
enum field_mode : int
{
    FIELD_STEP    = 0, // rho -> Poisson -> fit -> slice n
    FIELD_POISSON = 1, // rho -> nodal phi (in/out buffer) + energy
    FIELD_FIT     = 2  // nodal values -> slice
};

struct field_params
{
    int      dim, order, mode;
    uint32_t Nx, Ny, Nz, N;      // N = Nx*Ny*Nz
    uint32_t log2x, log2y, log2z; // 0xffffffff if not a power of two
    const double *partials;      // FIELD_STEP: partials (nullable -> use values)
    uint32_t wpx;
    double   weight;             // -dv multiplier applied to partial sums
    double  *values;             // rho / phi (FIELD_POISSON in/out, FIELD_FIT in)
    double  *rho_out;            // optional copy of rho
    const double2 *tw_x, *tw_y, *tw_z; // forward twiddles exp(-2 pi i k/N_axis)
    int      tw_lds;             // copy them into LDS first (set by the launcher)
    const double2 *mult;         // per-mode multiplier (complex), N entries
    const double  *ewt;          // per-mode energy weight, N entries
    double  *slice;              // history slice n (coefficients)
    double  *dslice;             // 1D differences of slice n (stride_d)
    double  *qslice;             // 1D FP64 kick-polynomial table of slice n (poly_stride layout) or null
    float   *qslice32;           // 1D FP32 table of slice n, scaled by kick, or null
    uint32_t q2_off;             // their Q2 row (poly_q2)
    double   kick;               // the FP32 tables' scale: 1D fp32c1_scale, 3D fp32c3_scale
    float   *slice32;            // FP32 coefficients times kick (3D FP32C) or null
    double  *energy;             // energies[n]
    uint32_t sy, sz, stride_d;
};

hipError_t launch_field( const field_params &p, hipStream_t st );
// The same solve over many workgroups for grids field_supported refuses
// (field_large.hip): W0, W1 hold N complex values each, epart
// field_large_partials(N) reals.
hipError_t launch_field_large( const field_params &p, double2 *W0, double2 *W1, double *epart, hipStream_t st );
size_t field_large_partials( uint32_t N );
// 1D lines up to this length take the multi-workgroup solve (a direct DFT per line)
constexpr uint64_t FIELD_LARGE_MAX_1D = 32768;

// One run of a batched sweep for the field kernel: f holds the bases of the
// run's history (slice, dslice at slice 0; energy at step 0).
struct batch_field
{
    field_params f;
    uint64_t Nt;
    uint32_t st, sd, sq;
};
hipError_t launch_field_batch( const batch_field *runs, uint32_t n_runs, uint64_t n, size_t lds, hipStream_t st );
// LDS of the batch field launch; *tw_lds tells whether the runs' twiddles fit it
size_t field_batch_lds_bytes( uint32_t max_N, int *tw_lds );
size_t field_lds_bytes( const field_params &p );
int field_supported( int dim, uint32_t Nx, uint32_t Ny, uint32_t Nz );

// metrics[k] = sum_g partials[4*g + k] in a fixed order (one workgroup).
hipError_t launch_reduce_metrics( const double *partials, uint32_t n_tasks, double *metrics, hipStream_t st );

// Sum of the partials into rho[ix] = weight * sum_j partials[j*n_space + ix]
// for ix in [first, first+count), zero elsewhere (n_space entries).
hipError_t launch_reduce_partials( const double *partials, uint32_t wpx, double weight,
                                   uint32_t first, uint32_t count, uint32_t n_space,
                                   double *rho, hipStream_t st );
// The 1D FP32C trace carries its velocities in cells per step (w = dt dx_inv
// u): its FP32 kick-polynomial tables are scaled in FP64 by the kick
// -dt dx_inv / 2 times dt dx_inv before their one rounding.
inline double fp32c1_scale( double dt, double dx_inv ) { return -0.5 * ( dt * dx_inv ) * ( dt * dx_inv ); }

// The 3D FP32C trace carries its velocities in cells per step (w = dt dx_inv
// u) and reads the slice as float( c * fp32c3_scale ): the x kick -dt dx_inv
// times dt dx_inv (velocity to cells per step) over 72 = 2*6*6, the factor of
// one derivative basis (2 D) and two value bases (6 B) in their
// integer-coefficient form, so the gradient the trace evaluates is the change
// of w itself and no rounded constant scales every node's kick alike.
inline double fp32c3_scale( double dt, double dx_inv ) { return -( dt * dx_inv ) * ( dt * dx_inv ) / 72.0; }

// Differences and the FP64 / FP32 kick-polynomial tables (1D) and the FP32
// copy (3D FP32C) of an uploaded slice.
hipError_t launch_slice_tables( int dim, const double *slice, uint32_t len, uint32_t Nx, double *dslice,
                                double *qslice, float *qslice32, double kick, float *slice32, hipStream_t st );

// The 1D kick-polynomial table of one slice, written with the slice (field
// and slice-table kernels) and copied into LDS by the fast, FP32C and adjoint
// kernels: for cell i (i = -1 .. Nx, periodic; entry j = i+1) the
// coefficients of 2 dphi/dt = Q0 + Q1 t + Q2 t^2 on [0,1), with Q0 = D0 + D1,
// Q1 = 2 (D1 - D0), Q2 = (D0 - D1) + (D2 - D1) from the differences
// D_l = c[i+l+1] - c[i+l] of the cell's window (the FP32 table scaled by the
// kick -dt/(2 dx)).  Layout: the pairs (Q0, Q1) at 2j, 2j+1, then the Q2 row
// at poly_q2 + j; the stride is a multiple of 4 elements so that every slice
// starts 16-byte aligned in FP64 and FP32.
__host__ __device__ inline uint32_t poly_q2( uint32_t Nx ) { return 2u * ( Nx + 2u ); }
__host__ __device__ inline uint32_t poly_stride( uint32_t Nx ) { return ( 3u * ( Nx + 2u ) + 3u ) & ~3u; }
// The table chunk of a 1D x-major launch for a requested chunk `req` >= 1:
// the largest chunk up to req whose two table buffers (`elem`-byte entries)
// fit `lds_max` bytes, double-buffered; failing that, one buffer of up to req
// slices (*pipe = false).  At least 1.
inline uint32_t poly_chunk_plan( uint32_t Nx, size_t elem, uint64_t req, size_t lds_max, bool *pipe )
{
    const uint64_t row = uint64_t( poly_stride( Nx ) ) * elem;
    uint64_t c = std::min( req, uint64_t( lds_max ) / ( 2 * row ) );
    *pipe = c >= 1;
    if ( c == 0 ) c = std::max<uint64_t>( 1, std::min<uint64_t>( req, uint64_t( lds_max ) / row ) );
    return uint32_t( std::min<uint64_t>( c, 0xffffffffu ) );
}

} // namespace nufi_impl

#endif
