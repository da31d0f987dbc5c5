// trace1d.hip — the 1D characteristic trace and rho partial sums on gfx950.
//
// Reference: dim1::eval_ftilda (nufi/rho.hpp:31-61) inside dim1::cuda_eval_rho
// (nufi/cuda_kernel.cu:31-51).  For every node (x_ix, u_iu) the characteristic
// is traced backward through slices n-1 .. 1 (drift, E from slice m, kick),
// then a final drift and half-kick with slice 0, and f0 is evaluated at the
// foot point.  The reference adds -du*f into rho[ix] with an atomic per node;
// here each wave of 64 velocity nodes of one x reduces its f values with a
// compensated butterfly and writes one partial, which the field kernel sums in
// a fixed order — deterministic, no atomics.
//
// Every workgroup streams the history through LDS in chunks of `chunk`
// slices; all nodes of the workgroup consume the same slice at the same
// time (the trace loop is uniform in m), so each slice is read from L2 once
// per workgroup and then gathered from LDS by every node.  A slice holds the
// first differences D[i] = c[i+1] - c[i] of the coefficient slice, which are
// exactly the differences splines.hpp:98-100 forms on the fly, extended
// periodically by one entry on each side (stored at offset 1: D[-1] ..
// D[Nx+order-2]) so that a cell index of -1 or Nx from a wrap that rounds
// onto the boundary needs no correction.
//
// Variants:
//   strict : the reference's operation order (unwrapped x, the wrap of
//            fields.hpp:42-46, de Boor of splines.hpp:102-106), no FMA
//            contraction — bit-level faithful up to device cos/exp.
//   fast   : cell-unit position s = (x - x_min)/dx wrapped exactly into
//            [0, Nx) every step, the derivative of the cubic B-spline in the
//            closed form  dphi/dt = D1 + (1-t)^2/2 (D0-D1) + t^2/2 (D2-D1)
//            and FMAs (order 4).
//   fp32c  : as fast, in FP32, with the position and velocity carried as
//            unevaluated sums hi + lo updated by TwoSum / TwoProd; f0 and the
//            rho sums stay FP64 (order 4).
#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"
#include "step1d.hpp"
#include "strict1d.hpp"

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

// Cooperative copy of `count` elements into LDS (16-byte vectors where the
// source allows it; chunk bases are 16-byte aligned by construction).
template <int BLOCK, typename T>
__device__ __forceinline__ void stage( T *__restrict__ dst, const T *__restrict__ src, uint32_t count )
{
    constexpr uint32_t V = 16 / sizeof(T), B = STAGE_BATCH;
    const uint32_t nv = count / V;
    const uint4 *s4 = reinterpret_cast<const uint4*>( src );
    uint4 *d4 = reinterpret_cast<uint4*>( dst );
    for ( uint32_t base = threadIdx.x; base < nv; base += B * BLOCK )   // B loads in flight per thread
    {
        uint4 v[B];
        #pragma unroll
        for ( uint32_t k = 0; k < B; ++k )             // clamped, unconditional loads stay in registers
        {
            const uint32_t i = base + k * BLOCK;
            v[k] = s4[ i < nv ? i : nv - 1 ];
        }
        #pragma unroll
        for ( uint32_t k = 0; k < B; ++k )
        {
            const uint32_t i = base + k * BLOCK;
            if ( i < nv ) d4[i] = v[k];
        }
    }
    for ( uint32_t i = nv*V + threadIdx.x; i < count; i += BLOCK ) dst[i] = src[i];
}

struct node_id
{
    uint32_t g, ix, j, iu;
    bool task, active;
};

template <int WPB, int NPT>
__device__ __forceinline__ node_id make_node( const trace1_params &p, int r, uint32_t wg )
{
    node_id id;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    id.g = ( wg * NPT + r ) * WPB + wave;
    id.task = id.g < p.n_tasks;
    const uint32_t ixl = id.task ? id.g / p.wpx : 0;
    id.j  = id.task ? id.g % p.wpx : 0;
    id.ix = p.ix_first + ixl;
    id.iu = id.j * WAVE + lane;
    const uint64_t q = uint64_t(id.ix) * p.Nu + id.iu;
    id.active = id.task && id.iu < p.Nu && q >= p.q_begin && q < p.q_end;
    return id;
}

__device__ __forceinline__ void write_partial( const trace1_params &p, const node_id &id, double f )
{
    const csum v = wave_csum( csum { f, 0.0 } );
    if ( id.task && ( threadIdx.x % WAVE ) == 0 )
        p.partials[ size_t(id.j) * p.Nx + id.ix ] = v.s + v.c;
}

// ---- x-major layout (fast and FP32C rho kernels) ---------------------------
struct xm_tile
{
    uint32_t xt, vb, xo, vr;   // x-tile, velocity block, lane's x offset and velocity row
    bool lane_ok;              // the lane maps to an x node of the range
};

__device__ __forceinline__ xm_tile make_tile( const trace1_params &p, uint32_t wg )
{
    xm_tile t;
    const uint32_t lane = threadIdx.x % WAVE;
    t.xt = wg % p.n_xt; t.vb = wg / p.n_xt;
    t.xo = lane % p.xw; t.vr = lane / p.xw;
    t.lane_ok = t.vr < p.vpw && t.xt * p.xw + t.xo < p.ix_count;
    return t;
}

struct xm_node
{
    uint32_t ix, iu, vi;       // vi: velocity row within the workgroup
    bool active;
};

template <int WPB, int NPT>
__device__ __forceinline__ xm_node make_xm_node( const trace1_params &p, const xm_tile &t, int r )
{
    xm_node id;
    const uint32_t wave = threadIdx.x / WAVE;
    id.vi = ( uint32_t( r ) * WPB + wave ) * p.vpw + t.vr;
    id.ix = t.lane_ok ? p.ix_first + t.xt * p.xw + t.xo : p.ix_first;
    id.iu = t.vb * ( uint32_t( NPT * WPB ) * p.vpw ) + id.vi;
    const uint64_t q = uint64_t( id.ix ) * p.Nu + id.iu;
    id.active = t.lane_ok && id.iu < p.Nu && q >= p.q_begin && q < p.q_end;
    return id;
}

// f of the workgroup's nodes -> LDS -> one compensated sum per (x, group of
// XM_GROUP consecutive velocities), summed in iu order, at partials[group*Nx + x].  `fbuf` may alias the
// history chunk: the barrier below ends its use.
template <int BLOCK, int NPT>
__device__ __forceinline__ void xm_reduce( const trace1_params &p, const xm_tile &t, const xm_node id[NPT],
                                           const double f[NPT], double *fbuf )
{
    constexpr uint32_t WPB = BLOCK / WAVE;
    const uint32_t groups = ( uint32_t( NPT ) * WPB * p.vpw ) / XM_GROUP;
    __syncthreads();
    if ( t.vr < p.vpw )
    {
        #pragma unroll
        for ( int r = 0; r < NPT; ++r ) fbuf[ id[r].vi * p.xw + t.xo ] = id[r].active ? f[r] : 0.0;
    }
    __syncthreads();
    for ( uint32_t k = threadIdx.x; k < groups * p.xw; k += BLOCK )
    {
        const uint32_t xo = k % p.xw, gi = k / p.xw;
        const uint32_t xl = t.xt * p.xw + xo, gg = t.vb * groups + gi;
        if ( xl >= p.ix_count || gg >= p.ppx ) continue;
        const double *col = fbuf + gi * XM_GROUP * p.xw + xo;
        csum acc { 0.0, 0.0 };
        #pragma unroll
        for ( uint32_t v = 0; v < XM_GROUP; ++v ) acc = csum_add( acc, csum { col[ v * p.xw ], 0.0 } );
        p.partials[ size_t( gg ) * p.Nx + p.ix_first + xl ] = acc.s + acc.c;
    }
}

// The four metric integrands of one node (cuda_kernel.cu:73-78), wave-summed.
__device__ __forceinline__ void write_metrics( double *partials, uint32_t g, bool task, bool active,
                                               double w, double f, double v2 )
{
    double m[4];
    m[0] = active ? w*f : 0.0;
    m[1] = active ? w*f*f : 0.0;
    m[2] = active ? w*(v2*f/2) : 0.0;
    m[3] = ( active && f > 0 ) ? -w*f*log(f) : 0.0;
    #pragma unroll
    for ( int k = 0; k < 4; ++k )
    {
        const csum v = wave_csum( csum { m[k], 0.0 } );
        if ( task && ( threadIdx.x % WAVE ) == 0 ) partials[ size_t(g) * 4 + k ] = v.s + v.c;
    }
}

// first slice of the backward loop: n-1 for rho; n (the initial half-kick) for
// metrics, none at n = 0 where both return f0 at the node
template <int MODE>
__device__ __forceinline__ int64_t top_slice( uint64_t n )
{
    if ( MODE == MODE_METRICS ) return n == 0 ? -1 : int64_t( n );
    return int64_t( n ) - 1;
}

// ---- fast FP64 (order 4) ----------------------------------------------------
template <int BLOCK, int NPT, int MODE>
__device__ __forceinline__ void trace1d_fast_body( const trace1_params &p, uint32_t wg, double *lds )
{
    constexpr int WPB = BLOCK / WAVE;

    node_id id[NPT];
    double s[NPT], u[NPT];
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        id[r] = make_node<WPB,NPT>( p, r, wg );
        // node (cuda_kernel.cu:43-44); on the field grid its cell position is exactly ix
        s[r] = ( MODE == MODE_RHO ) ? double( id[r].ix )
                                    : ( ( p.nx_min + id[r].ix*p.ndx ) - p.x_min ) * p.dx_inv;
        u[r] = p.nu_min + id[r].iu*p.ndu + p.ndu/2;
    }

    const double Nxd   = double( p.Nx );
    const double invNx = 1.0 / Nxd;
    const double cd    = p.dt * p.dx_inv;          // drift in cells per unit u
    const double ck    = -0.5 * p.dt * p.dx_inv;   // kick per unit 2 dphi/dt
    const uint32_t sd  = p.stride_d;

    for ( int64_t mhi = top_slice<MODE>( p.n ); mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        stage<BLOCK>( lds, p.dhist + size_t( mlo ) * sd, uint32_t( cnt ) * sd );
        __syncthreads();
        for ( int jj = cnt - 1; jj >= 0; --jj )
        {
            const double *D = lds + jj * sd + 1;      // D[-1] .. D[Nx+order-2]
            const uint64_t m = uint64_t( mlo + jj );
            const bool initial = ( MODE == MODE_METRICS ) && m == p.n;   // half-kick, no drift
            const double kf = ( m == 0 || initial ) ? 0.5 * ck : ck;
            #pragma unroll
            for ( int r = 0; r < NPT; ++r )
            {
                sv_step( s[r], u[r], D, cd, kf, Nxd, invNx, !initial );
            }
        }
    }

    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        const double f = id[r].active ? f0_1d( p.f0, p.x_min + s[r]*p.dx, u[r] ) : 0.0;
        if ( MODE == MODE_RHO ) write_partial( p, id[r], f );
        else
        {
            const double u0 = p.nu_min + id[r].iu*p.ndu + p.ndu/2;  // the node's velocity (cuda_kernel.cu:70)
            write_metrics( p.partials, id[r].g, id[r].task, id[r].active, p.weight, f, u0*u0 );
        }
    }
}

template <int BLOCK, int NPT, int MODE>
__global__ __launch_bounds__( BLOCK )
void trace1d_fast_kernel( trace1_params p )
{
    extern __shared__ double lds[];
    trace1d_fast_body<BLOCK,NPT,MODE>( p, blockIdx.x, lds );
}

// ---- strict FP64 (reference operation order, any order 3..8) ---------------
template <int ORDER>
__device__ __forceinline__ double efield_strict( double x, const double *D, const trace1_params &p )
{
    return -phi_deriv_strict<ORDER,1>( x, D, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
}

template <int ORDER, int BLOCK, int MODE>
__global__ __launch_bounds__( BLOCK )
void trace1d_strict_kernel( trace1_params p )
{
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    constexpr int WPB = BLOCK / WAVE;
    const node_id id = make_node<WPB,1>( p, 0, blockIdx.x );

    double x = p.nx_min + id.ix * p.ndx;                    // cuda_kernel.cu:43, 69
    const double u0 = p.nu_min + id.iu * p.ndu + p.ndu / 2; // cuda_kernel.cu:44, 70
    double u = u0;
    const uint32_t sd = p.stride_d;
    const double half_dt = 0.5 * p.dt;

    for ( int64_t mhi = top_slice<MODE>( p.n ); mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        stage<BLOCK>( lds, p.dhist + size_t( mlo ) * sd, uint32_t( cnt ) * sd );
        __syncthreads();
        for ( int jj = cnt - 1; jj >= 0; --jj )
        {
            const uint64_t m = uint64_t( mlo + jj );
            const bool initial = ( MODE == MODE_METRICS ) && m == p.n;   // rho.hpp:76-79
            if ( !initial ) x = x - p.dt * u;
            const double E = efield_strict<ORDER>( x, lds + jj * sd + 1, p );
            u = ( m == 0 || initial ) ? u + half_dt * E : u + p.dt * E;
        }
    }
    const double f = id.active ? f0_1d( p.f0, x, u ) : 0.0;
    if ( MODE == MODE_RHO ) write_partial( p, id, f );
    else write_metrics( p.partials, id.g, id.task, id.active, p.weight, f, u0*u0 );
}

// ---- fast FP64 with the kick-polynomial table (order 4, rho) ----------------
// Per char-step: drift (FMA), floor, index, one 16-byte and one 8-byte LDS read,
// two FMAs for Q0 + t (Q1 + t Q2), one FMA for the kick.  The tables are
// copied from p.qhist into LDS by direct loads, double-buffered when p.pipe is
// set (poly_ready_pipeline), else one buffer per chunk (poly_ready_direct).
template <int BLOCK, int NPT, bool POW2>
__device__ __forceinline__ void trace1d_poly_body( const trace1_params &p, uint32_t wg, double *lds )
{
    constexpr int WPB = BLOCK / WAVE;
    const xm_tile tile = make_tile( p, wg );
    xm_node id[NPT];
    double s[NPT], u[NPT];
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        id[r] = make_xm_node<WPB,NPT>( p, tile, r );
        s[r] = double( id[r].ix );
        u[r] = p.nu_min + id[r].iu*p.ndu + p.ndu/2;          // cuda_kernel.cu:44
    }
    const double Nxd = double( p.Nx ), invNx = 1.0 / Nxd;
    const int mask = int( p.Nx ) - 1;
    const double cd = p.dt * p.dx_inv, kick = p.kick;
    const uint32_t sq = p.stride_q, q2 = p.q2_off;

    // slices mlo+hi down to mlo+lo, in segments between re-wrap points so that
    // the step loop itself carries no wrap; slice 0 (the half kick) peeled
    auto steps = [&]( const double *Qc, int mlo, int hi, int lo )
    {
        int m = mlo + hi;
        const int mend = mlo + lo;
        const double *Q = Qc + hi * sq, *Q2 = Q + q2;
        while ( m >= mend )
        {
            if ( POW2 && ( m & ( POLY_WRAP - 1 ) ) == POLY_WRAP - 1 )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r ) rewrap( s[r], Nxd, invNx );
            }
            // slices m .. seg need no re-wrap (the next one is below seg)
            const int seg = POW2 ? ( ( m & ~( POLY_WRAP - 1 ) ) > mend ? ( m & ~( POLY_WRAP - 1 ) ) : mend ) : mend;
            const int s1 = seg > 0 ? seg : 1;
            #pragma unroll 2
            for ( ; m >= s1; --m, Q -= sq, Q2 -= sq )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r ) poly_step<POW2>( s[r], u[r], Q, Q2, cd, kick, mask, Nxd, invNx );
            }
            if ( seg == 0 && m == 0 )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r ) poly_step<POW2>( s[r], u[r], Q, Q2, cd, 0.5 * kick, mask, Nxd, invNx );
                --m; Q -= sq; Q2 -= sq;
            }
        }
    };
    if ( p.pipe ) poly_ready_pipeline<BLOCK>( lds, p.qhist, sq, p.n, p.chunk, steps );
    else          poly_ready_direct<BLOCK>( lds, p.qhist, sq, p.n, p.chunk, steps );

    double f[NPT];
    #pragma unroll
    for ( int r = 0; r < NPT; ++r ) f[r] = id[r].active ? f0_1d( p.f0, p.x_min + s[r]*p.dx, u[r] ) : 0.0;
    xm_reduce<BLOCK,NPT>( p, tile, id, f, lds );
}

template <int BLOCK, int NPT, bool POW2>
__global__ __launch_bounds__( BLOCK )
void trace1d_poly_kernel( trace1_params p )
{
    extern __shared__ double lds[];
    trace1d_poly_body<BLOCK,NPT,POW2>( p, blockIdx.x, lds );
}

// ---- FP32 compensated (order 4) --------------------------------------------
// The FP32 kick-polynomial table: an 8-byte (Q0, Q1) and a 4-byte Q2 LDS
// read per char-step.
template <int BLOCK, int NPT, bool POW2>
__device__ __forceinline__ void trace1d_fp32c_body( const trace1_params &p, uint32_t wg, float *ldsf )
{
    constexpr int WPB = BLOCK / WAVE;
    const xm_tile tile = make_tile( p, wg );
    xm_node id[NPT];
    float sh[NPT], sl[NPT], uh[NPT], ul[NPT];       // position (cells) + P0, velocity (cells per step) + W0
    const double cd = p.dt * p.dx_inv;
    const frame1 fr = make_frame1( p.f32_p0, p.f32_w0, p.Nx );
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        id[r] = make_xm_node<WPB,NPT>( p, tile, r );
        fp32c_start( fr, id[r].ix, cd * ( p.nu_min + id[r].iu*p.ndu + p.ndu/2 ), sh[r], sl[r], uh[r], ul[r] );
    }

    const int mask    = int( p.Nx ) - 1;
    const uint32_t sq = p.stride_q, q2 = p.q2_off;

    auto wrap = [&]( int r ) { fp32c_rewrap( fr, sh[r], sl[r], uh[r], ul[r] ); };
    auto step = [&]( const float *Q, float hf )
    {
        #pragma unroll
        for ( int r = 0; r < NPT; ++r )
        {
            fp32c_move( fr, sh[r], sl[r], uh[r], ul[r] );
            // exact periodic wrap into [P0, P0 + Nx) and renormalised pairs;
            // POW2: the cell index wraps by the mask and the position is
            // re-wrapped every POLY_WRAP slices
            if ( !POW2 ) wrap( r );
            float t;
            const int i = fp32c_locate<POW2>( fr, sh[r], sl[r], mask, t );
            fp32c_kick( uh[r], ul[r], hf * fp32c_poly( Q, q2, i, t ) );
        }
    };
    // slices mlo+hi down to mlo+lo, in segments between re-wrap points;
    // slice 0 (the half kick) peeled
    auto steps = [&]( const float *Qc, int mlo, int hi, int lo )
    {
        int m = mlo + hi;
        const int mend = mlo + lo;
        const float *Q = Qc + hi * sq;
        while ( m >= mend )
        {
            if ( POW2 && ( m & ( POLY_WRAP - 1 ) ) == POLY_WRAP - 1 )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r ) wrap( r );
            }
            const int seg = POW2 ? ( ( m & ~( POLY_WRAP - 1 ) ) > mend ? ( m & ~( POLY_WRAP - 1 ) ) : mend ) : mend;
            const int s1 = seg > 0 ? seg : 1;
            for ( ; m >= s1; --m, Q -= sq ) step( Q, 1.0f );
            if ( seg == 0 && m == 0 ) { step( Q, 0.5f ); --m; Q -= sq; }
        }
    };
    if ( p.pipe ) poly_ready_pipeline<BLOCK>( ldsf, p.qhist32, sq, p.n, p.chunk, steps );
    else          poly_ready_direct<BLOCK>( ldsf, p.qhist32, sq, p.n, p.chunk, steps );

    double f[NPT];
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        const double s = fp32c_position( fr, sh[r], sl[r] );
        const double u = fp32c_velocity( fr, uh[r], ul[r] ) / cd;
        f[r] = id[r].active ? f0_1d( p.f0, p.x_min + s*p.dx, u ) : 0.0;
    }
    xm_reduce<BLOCK,NPT>( p, tile, id, f, reinterpret_cast<double*>( ldsf ) );
}

template <int BLOCK, int NPT, bool POW2>
__global__ __launch_bounds__( BLOCK )
void trace1d_fp32c_kernel( trace1_params p )
{
    extern __shared__ float ldsf[];
    trace1d_fp32c_body<BLOCK,NPT,POW2>( p, blockIdx.x, ldsf );
}

// ---- batched sweep: one launch over the wave-tasks of many independent runs --
// Workgroup b belongs to run wg_run[b]; its parameter block lives in device
// memory (runs[r].p, with n filled in here) and its workgroup index within the
// run is b - runs[r].wg_begin.  Runs past their own Nt skip the step.
template <int BLOCK, int NPT, int VARIANT>
__global__ __launch_bounds__( BLOCK )
void trace1d_batch_kernel( const batch_run *__restrict__ runs, const uint32_t *__restrict__ wg_run, uint64_t n )
{
    extern __shared__ double lds_b[];
    const uint32_t r = wg_run[ blockIdx.x ];
    if ( n > runs[r].Nt ) return;                     // uniform over the workgroup
    trace1_params p = runs[r].p;
    p.n = n;
    const uint32_t wg = blockIdx.x - runs[r].wg_begin;
    const bool pow2 = ( p.Nx & ( p.Nx - 1 ) ) == 0;
    if ( VARIANT == TRACE_FP32C )
    {
        if ( pow2 ) trace1d_fp32c_body<BLOCK,NPT,true>( p, wg, reinterpret_cast<float*>( lds_b ) );
        else        trace1d_fp32c_body<BLOCK,NPT,false>( p, wg, reinterpret_cast<float*>( lds_b ) );
    }
    else if ( pow2 ) trace1d_poly_body<BLOCK,NPT,true>( p, wg, lds_b );
    else        trace1d_poly_body<BLOCK,NPT,false>( p, wg, lds_b );
}

template <typename K>
hipError_t launch_kernel( K kernel, uint32_t grid, int block, size_t lds, const trace1_params &p,
                          hipStream_t st )
{
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( kernel ),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    hipLaunchKernelGGL( kernel, dim3( grid ), dim3( block ), lds, st, p );
    return hipGetLastError();
}

template <int BLOCK, int NPT>
hipError_t launch_fast_or_32( const trace1_params &p, int variant, uint32_t grid, size_t lds, hipStream_t st )
{
    if ( p.mode == MODE_METRICS )
        return launch_kernel( trace1d_fast_kernel<BLOCK,NPT,MODE_METRICS>, grid, BLOCK, lds, p, st );
    if ( variant == TRACE_FP32C )
    {
        if ( ( p.Nx & ( p.Nx - 1 ) ) == 0 )
            return launch_kernel( trace1d_fp32c_kernel<BLOCK,NPT,true>, grid, BLOCK, lds, p, st );
        return launch_kernel( trace1d_fp32c_kernel<BLOCK,NPT,false>, grid, BLOCK, lds, p, st );
    }
    if ( ( p.Nx & ( p.Nx - 1 ) ) == 0 )
        return launch_kernel( trace1d_poly_kernel<BLOCK,NPT,true>, grid, BLOCK, lds, p, st );
    return launch_kernel( trace1d_poly_kernel<BLOCK,NPT,false>, grid, BLOCK, lds, p, st );
}

template <int BLOCK>
hipError_t launch_npt( const trace1_params &p, int variant, int npt, uint32_t grid, size_t lds, hipStream_t st )
{
    switch ( npt )
    {
    case 1: return launch_fast_or_32<BLOCK,1>( p, variant, grid, lds, st );
    case 2: return launch_fast_or_32<BLOCK,2>( p, variant, grid, lds, st );
    case 4: return launch_fast_or_32<BLOCK,4>( p, variant, grid, lds, st );
    default: return hipErrorInvalidValue;
    }
}

template <int ORDER, int MODE>
hipError_t launch_strict_mode( const trace1_params &p, int block, uint32_t grid, size_t lds, hipStream_t st )
{
    switch ( block )
    {
    case 64:   return launch_kernel( trace1d_strict_kernel<ORDER,64,MODE>,   grid, 64,   lds, p, st );
    case 256:  return launch_kernel( trace1d_strict_kernel<ORDER,256,MODE>,  grid, 256,  lds, p, st );
    case 1024: return launch_kernel( trace1d_strict_kernel<ORDER,1024,MODE>, grid, 1024, lds, p, st );
    default:   return hipErrorInvalidValue;
    }
}

template <int ORDER>
hipError_t launch_strict( const trace1_params &p, int block, uint32_t grid, size_t lds, hipStream_t st )
{
    if ( p.mode == MODE_METRICS ) return launch_strict_mode<ORDER,MODE_METRICS>( p, block, grid, lds, st );
    return launch_strict_mode<ORDER,MODE_RHO>( p, block, grid, lds, st );
}

template <int BLOCK, int NPT>
hipError_t launch_batch_npt( const batch_run *runs, const uint32_t *wg_run, uint32_t n_wg, uint64_t n,
                             int variant, size_t lds, hipStream_t st )
{
    if ( variant == TRACE_FP32C )
        hipLaunchKernelGGL( ( trace1d_batch_kernel<BLOCK,NPT,TRACE_FP32C> ), dim3( n_wg ), dim3( BLOCK ), lds, st,
                            runs, wg_run, n );
    else
        hipLaunchKernelGGL( ( trace1d_batch_kernel<BLOCK,NPT,TRACE_FAST> ), dim3( n_wg ), dim3( BLOCK ), lds, st,
                            runs, wg_run, n );
    return hipGetLastError();
}

template <int BLOCK>
hipError_t launch_batch_block( const batch_run *runs, const uint32_t *wg_run, uint32_t n_wg, uint64_t n,
                               int variant, int npt, size_t lds, hipStream_t st )
{
    switch ( npt )
    {
    case 1: return launch_batch_npt<BLOCK,1>( runs, wg_run, n_wg, n, variant, lds, st );
    case 2: return launch_batch_npt<BLOCK,2>( runs, wg_run, n_wg, n, variant, lds, st );
    case 4: return launch_batch_npt<BLOCK,4>( runs, wg_run, n_wg, n, variant, lds, st );
    default: return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t launch_trace1d_batch( const batch_run *runs, const uint32_t *wg_run, uint32_t n_wg, uint64_t n,
                                 int variant, const launch_shape &s, size_t lds, hipStream_t st )
{
    if ( ( variant != TRACE_FAST && variant != TRACE_FP32C ) || n >= ( uint64_t( 1 ) << 31 ) ) return hipErrorInvalidValue;
    if ( n_wg == 0 ) return hipSuccess;
    switch ( s.block )
    {
    case 64:   return launch_batch_block<64>  ( runs, wg_run, n_wg, n, variant, s.npt, lds, st );
    case 128:  return launch_batch_block<128> ( runs, wg_run, n_wg, n, variant, s.npt, lds, st );
    case 256:  return launch_batch_block<256> ( runs, wg_run, n_wg, n, variant, s.npt, lds, st );
    case 512:  return launch_batch_block<512> ( runs, wg_run, n_wg, n, variant, s.npt, lds, st );
    case 1024: return launch_batch_block<1024>( runs, wg_run, n_wg, n, variant, s.npt, lds, st );
    default:   return hipErrorInvalidValue;
    }
}

// shapes with an instantiated kernel (velocity-major launches)
static bool shape_built( int variant, int order, const launch_shape &s )
{
    if ( variant == TRACE_STRICT )
        return order >= 3 && order <= 8 && s.npt == 1 &&
               ( s.block == 64 || s.block == 256 || s.block == 1024 ) && s.chunk >= 1;
    if ( order != 4 ) return false;
    const bool bok = s.block == 64 || s.block == 128 || s.block == 256 || s.block == 512 || s.block == 1024;
    const bool nok = s.npt == 1 || s.npt == 2 || s.npt == 4;
    return bok && nok && s.chunk >= 1;
}

// x-major: the velocity rows of a workgroup are whole summation groups; and
// 1024 threads x 4 nodes exceeds the 128 VGPRs a 1024-thread workgroup allows
static bool shape_groups( const launch_shape &s )
{
    return ( uint32_t( s.block / WAVE ) * uint32_t( s.npt ) ) % XM_GROUP == 0 && !( s.block == 1024 && s.npt == 4 );
}

// a shape for the rho kernel of the variant
bool trace1d_shape_ok( int variant, int order, const launch_shape &s )
{
    return shape_built( variant, order, s ) && ( variant == TRACE_STRICT || shape_groups( s ) );
}

static bool x_major( const trace1_params &p, int variant )
{
    return p.mode == MODE_RHO && variant != TRACE_STRICT;
}

void trace1d_xm_setup( trace1_params &p )
{
    p.xw   = xm_xw( p.Nx );
    p.vpw  = WAVE / p.xw;
    p.n_xt = ( p.ix_count + p.xw - 1 ) / p.xw;
    p.ppx  = xm_ppx( p.Nu );
}

// the FP64 and FP32C rho kernels read their tables from qhist / qhist32
static bool poly_ready( const trace1_params &p, int variant )
{
    return ( variant == TRACE_FAST && p.qhist != nullptr ) || ( variant == TRACE_FP32C && p.qhist32 != nullptr );
}

uint32_t trace1d_chunk( const trace1_params &p, int variant, const launch_shape &s, size_t lds_max, bool *pipe )
{
    uint64_t c = uint64_t( s.chunk );
    bool pp = false;
    if ( x_major( p, variant ) )
        c = poly_chunk_plan( p.Nx, variant == TRACE_FP32C ? sizeof(float) : sizeof(double), c, lds_max, &pp );
    const uint64_t v = p.mode == MODE_METRICS ? ( p.n > 0 ? p.n + 1 : 0 ) : p.n;
    if ( v < c ) c = v > 0 ? v : 1;
    if ( pipe ) *pipe = pp;
    return uint32_t( c );
}

uint32_t trace1d_grid( const trace1_params &p, int variant, const launch_shape &s )
{
    if ( x_major( p, variant ) )
    {
        const uint32_t vbw = uint32_t( s.block / WAVE ) * uint32_t( s.npt ) * p.vpw;
        return p.n_xt * ( ( p.Nu + vbw - 1 ) / vbw );
    }
    const uint32_t tasks_per_block = uint32_t( s.block / WAVE ) * uint32_t( s.npt );
    return ( p.n_tasks + tasks_per_block - 1 ) / tasks_per_block;
}


// LDS row of one slice: the kick-polynomial table for the fast and FP32C rho
// kernels, the extended differences for the strict and metrics kernels
size_t trace1d_slice_elems( const trace1_params &p, int variant )
{
    return ( p.mode == MODE_RHO && variant != TRACE_STRICT ) ? p.stride_q : p.stride_d;
}

size_t trace1d_lds_bytes( const trace1_params &p, int variant, const launch_shape &s )
{
    const size_t elem = ( variant == TRACE_FP32C && p.mode == MODE_RHO ) ? sizeof(float) : sizeof(double);
    bool pipe = false;
    const uint32_t chunk = trace1d_chunk( p, variant, s, 160 * 1024, &pipe );
    const size_t hist = ( pipe ? 2 : 1 ) * size_t( chunk ) * trace1d_slice_elems( p, variant ) * elem;
    // x-major: the f values of the workgroup (at most block*npt doubles) reuse the LDS
    const size_t fbuf = x_major( p, variant ) ? size_t( s.block ) * size_t( s.npt ) * sizeof(double) : 0;
    return hist > fbuf ? hist : fbuf;
}

hipError_t launch_trace1d( const trace1_params &p_in, int variant, const launch_shape &s, hipStream_t st )
{
    if ( !shape_built( variant, int( p_in.order ), s ) || p_in.n >= ( uint64_t( 1 ) << 31 ) ) return hipErrorInvalidValue;
    trace1_params p = p_in;
    if ( x_major( p, variant ) )
    {
        if ( !shape_groups( s ) ) return hipErrorInvalidValue;
        trace1d_xm_setup( p );
        if ( p.ix_count == 0 ) return hipSuccess;
    }
    else if ( p.n_tasks == 0 ) return hipSuccess;
    // the x-major rho kernels copy the kick-polynomial tables of the history
    if ( x_major( p, variant ) && !poly_ready( p, variant ) ) return hipErrorInvalidValue;
    const uint32_t grid = trace1d_grid( p, variant, s );
    bool pipe = false;
    p.chunk = trace1d_chunk( p, variant, s, 160 * 1024, &pipe );
    p.pipe = pipe ? 1u : 0u;
    const size_t lds = trace1d_lds_bytes( p, variant, s );

    if ( variant == TRACE_STRICT )
    {
        switch ( p.order )
        {
        case 3: return launch_strict<3>( p, s.block, grid, lds, st );
        case 4: return launch_strict<4>( p, s.block, grid, lds, st );
        case 5: return launch_strict<5>( p, s.block, grid, lds, st );
        case 6: return launch_strict<6>( p, s.block, grid, lds, st );
        case 7: return launch_strict<7>( p, s.block, grid, lds, st );
        case 8: return launch_strict<8>( p, s.block, grid, lds, st );
        default: return hipErrorInvalidValue;
        }
    }
    switch ( s.block )
    {
    case 64:   return launch_npt<64>  ( p, variant, s.npt, grid, lds, st );
    case 128:  return launch_npt<128> ( p, variant, s.npt, grid, lds, st );
    case 256:  return launch_npt<256> ( p, variant, s.npt, grid, lds, st );
    case 512:  return launch_npt<512> ( p, variant, s.npt, grid, lds, st );
    case 1024: return launch_npt<1024>( p, variant, s.npt, grid, lds, st );
    default:   return hipErrorInvalidValue;
    }
}

} // namespace nufi_impl
