// step1d.hpp — the FP64 characteristic step of the 1D fast trace, shared by
// the forward trace kernel and the adjoint kernel so that both reach the same
// foot points bit for bit.
//
// Position in cell units s = (x - x_min)/dx, wrapped exactly into [0, Nx)
// (Nx*w is an integer, so s - Nx*w is exact); D points at the extended first
// differences of the slice (D[-1] .. D[Nx+2]); the derivative of the cubic
// B-spline in the cell is  2 dphi/dt = 2 D1 + (1-t)^2 (D0-D1) + t^2 (D2-D1).
#ifndef NUFI_STEP1D_HPP
#define NUFI_STEP1D_HPP

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace nufi_dev
{

// exact periodic wrap and cell location: returns the cell (in [-1, Nx]) and t
__device__ __forceinline__ int wrap_locate( double &s, double Nxd, double invNx, double &t )
{
    const double w = floor( s * invNx );
    s = fma( -Nxd, w, s );
    const double k = floor( s );
    t = s - k;
    return int( k );
}

// 2 dphi/dt at (cell i, t)
__device__ __forceinline__ double slope2( const double *D, int i, double t )
{
    const double D0 = D[i], D1 = D[i+1], D2 = D[i+2];
    const double a = 1.0 - t;
    return fma( a*a, D0 - D1, fma( t*t, D2 - D1, D1 + D1 ) );
}

// one drift + kick: s -= cd*u, u += kf * 2 dphi/dt(s)  (kf = -dt/(2 dx), halved at m = 0)
__device__ __forceinline__ void sv_step( double &s, double &u, const double *D, double cd, double kf,
                                         double Nxd, double invNx, bool drift = true )
{
    double x = drift ? fma( -cd, u, s ) : s;
    double t;
    const int i = wrap_locate( x, Nxd, invNx, t );
    u = fma( kf, slope2( D, i, t ), u );
    s = x;
}

// ---- the kick-polynomial step --------------------------------------------
// Q points at the table of one slice: the pair (Q0, Q1) of cell i at
// Q + 2*(i+1), Q2 at Q + q2 + (i+1), i = -1 .. Nx.  An FP64 table holds the
// polynomial of 2 dphi/dt itself and the step scales it:
// u += hk * (Q0 + Q1 t + Q2 t^2), hk = kick, or kick/2 on slice 0.  With POW2 the
// position is not wrapped per step (the cell index wraps by a mask); the
// caller re-wraps s before the step on every slice m with poly_wrap_due(m),
// a schedule fixed by m alone so that the result does not depend on how the
// slices are chunked.  Otherwise s is wrapped exactly each step.
constexpr int POLY_WRAP = 16;

__device__ __forceinline__ bool poly_wrap_due( int64_t m )
{
    return ( m & ( POLY_WRAP - 1 ) ) == POLY_WRAP - 1;
}

// Stepping down from slice m, the last slice before the next re-wrap point is
// the one above m' = (m & ~(POLY_WRAP-1)) - 1 (the largest m' < m with
// poly_wrap_due(m')); returns max(m', mlo - 1): step while slice > it.
__device__ __forceinline__ int64_t poly_segment_end( int64_t m, int64_t mlo )
{
    const int64_t next = ( m & ~int64_t( POLY_WRAP - 1 ) ) - 1;
    return next > mlo - 1 ? next : mlo - 1;
}

template <bool POW2>
__device__ __forceinline__ int poly_locate( double &x, int mask, double Nxd, double invNx, double &t )
{
    if ( POW2 )
    {
        const double k = floor( x );
        t = x - k;
        return int( k ) & mask;
    }
    return wrap_locate( x, Nxd, invNx, t );
}

__device__ __forceinline__ double poly_eval( const double *Q, uint32_t q2, int i, double t, double &dpdt )
{
    const double2 q01 = reinterpret_cast<const double2*>( Q )[ i + 1 ];
    const double c2 = Q[ q2 + ( i + 1 ) ];
    dpdt = fma( 2.0 * c2, t, q01.y );
    return fma( fma( c2, t, q01.y ), t, q01.x );
}

// Q: the slice's pairs, Q2: its Q2 row (Q + q2)
template <bool POW2>
__device__ __forceinline__ void poly_step( double &s, double &u, const double *Q, const double *Q2, double cd,
                                           double hk, int mask, double Nxd, double invNx )
{
    double x = fma( -cd, u, s );
    double t;
    const int i = poly_locate<POW2>( x, mask, Nxd, invNx, t );
    const double2 q01 = reinterpret_cast<const double2*>( Q )[ i + 1 ];
    const double c2 = Q2[ i + 1 ];
    u = fma( hk, fma( fma( c2, t, q01.y ), t, q01.x ), u );
    s = x;
}

// ---- the FP32-compensated step ---------------------------------------------
// Position (cell units) and velocity (cells per step, w = dt dx_inv u) as
// unevaluated pairs (hi + lo) whose hi parts carry the offsets P0 and W0 of
// fp32c_frame: every hi value of a run stays in one binade, so the offsets
// come off exactly, the drift and the kick are ordered Fast2Sums and a wrap
// by a multiple of Nx is exact.  The FP32 tables are scaled to give the
// change of w.  The FP32C trace kernels and the FP32C adjoint share these, so
// both reach the same states bit for bit.
struct frame1
{
    float P0, W0, N, invN;
    int P0i;
};

__device__ __forceinline__ frame1 make_frame1( float p0, float w0, uint32_t Nx )
{
    frame1 f;
    f.P0 = p0; f.W0 = w0; f.N = float( Nx ); f.invN = 1.0f / f.N; f.P0i = int( p0 );
    return f;
}

// drift s -= w (UNDO: s += w, the inverse step of the adjoint): a Fast2Sum of
// the hi parts (|sh| > |d|), the lo parts added to its error, left unnormalised
template <bool UNDO = false>
__device__ __forceinline__ void fp32c_move( const frame1 &f, float &sh, float &sl, float wh, float wl )
{
    const float d = UNDO ? f.W0 - wh : wh - f.W0;                  // exact
    const float y = sh - d;
    const float e = ( sh - y ) - d;
    sl = ( UNDO ? sl + wl : sl - wl ) + e;
    sh = y;
}

// wrap the position into [P0, P0 + Nx) (exact) and renormalise both pairs
__device__ __forceinline__ void fp32c_rewrap( const frame1 &f, float &sh, float &sl, float &wh, float &wl )
{
    sh = fmaf( -f.N, floorf( ( sh - f.P0 ) * f.invN ), sh );
    float s = sh + sl;
    sl = sl - ( s - sh ); sh = s;
    s = wh + wl;
    wl = wl - ( s - wh ); wh = s;
}

// the cell (in [-1, Nx]; POW2: wrapped by the mask, P0 being a multiple of
// Nx) and t of the pair.  sh >= P0 / 2 > 0, so truncation is the floor and
// the fraction sh - floor(sh) is exact (one instruction).
template <bool POW2>
__device__ __forceinline__ int fp32c_locate( const frame1 &f, float sh, float sl, int mask, float &t )
{
    t = __builtin_amdgcn_fractf( sh ) + sl;
    return POW2 ? ( int( sh ) & mask ) : int( sh ) - f.P0i;
}

// position (cells) and velocity (cells per step) of the pairs, in FP64
__device__ __forceinline__ double fp32c_position( const frame1 &f, float sh, float sl )
{
    return double( sh - f.P0 ) + double( sl );
}
__device__ __forceinline__ double fp32c_velocity( const frame1 &f, float wh, float wl )
{
    return double( wh - f.W0 ) + double( wl );
}

// the start state of node (ix, w0 cells per step)
__device__ __forceinline__ void fp32c_start( const frame1 &f, uint32_t ix, double w0, float &sh, float &sl,
                                             float &wh, float &wl )
{
    sh = f.P0 + float( ix ); sl = 0.0f;                              // exact
    const double w = double( f.W0 ) + w0;
    wh = float( w ); wl = float( w - double( wh ) );
}

// the kick q = kick dt dx_inv (Q0 + Q1 t + Q2 t^2) in cells per step from the
// FP32 table (pairs at Q, Q2 row at Q + q2)
__device__ __forceinline__ float fp32c_poly( const float *Q, uint32_t q2, int i, float t )
{
    const float2 q01 = reinterpret_cast<const float2*>( Q )[ i + 1 ];
    return fmaf( fmaf( Q[ q2 + ( i + 1 ) ], t, q01.y ), t, q01.x );
}

// w += q: a Fast2Sum, |wh| >= W0 / 2 > |q|
__device__ __forceinline__ void fp32c_kick( float &wh, float &wl, float q )
{
    const float s = wh + q;
    wl = wl + ( q - ( s - wh ) );
    wh = s;
}

// ---- kick-polynomial tables ------------------------------------------------
// Entry j (cell j-1) of a slice's table from the entries j, j+1, j+2 of the
// slice's extended difference row (entry e holds c[e] - c[e-1], periodic):
// Q0 = D0 + D1, Q1 = 2 (D1 - D0), Q2 = (D0 - D1) + (D2 - D1).  An FP32 table is
// scaled by the kick in FP64 before its one rounding to float; an FP64 table
// is not (the step applies the kick).  The field kernel writes the table of
// every slice it appends (the slice-table kernel on upload); the rho and
// adjoint kernels copy the tables into LDS.
template <typename T>
__device__ __forceinline__ void put_poly( T *pair, T *q2p, double D0, double D1, double D2, double kick )
{
    const double t1 = D1 - D0;
    double q0 = D0 + D1, q1 = t1 + t1, q2 = ( D2 - D1 ) - t1;
    if ( sizeof( T ) < sizeof( double ) ) { q0 *= kick; q1 *= kick; q2 *= kick; }
    pair[0] = T( q0 ); pair[1] = T( q1 ); *q2p = T( q2 );
}

// The backward loop over slices n-1 .. 0 in chunks of C slices, highest
// first (chunk k: slices max(n-1-kC-C+1, 0) .. n-1-kC), n < 2^31, over tables
// in global memory in the LDS layout (slice m at qhist + m*sq, sq*sizeof(T) a
// multiple of 16 bytes).  steps(Q, mlo, hi, lo) takes the steps of slices
// mlo+hi down to mlo+lo with the chunk's tables at Q (slice mlo+jj at
// Q + jj*sq).
//
// Two LDS buffers alternate: while chunk k is traced from one, chunk k+1 is
// copied into the other by direct global->LDS loads, which take no registers
// and no VALU; each wave waits for its own copies and the workgroup barrier
// publishes them at the chunk's end, when chunk k's buffer is also free again.
template <int BLOCK, typename T, typename STEPS>
__device__ __forceinline__ void poly_ready_pipeline( T *lds, const T *qhist, uint32_t sq, uint64_t n, uint32_t C,
                                                     STEPS steps )
{
    if ( n == 0 ) return;
    const int Ci = int( C ), top = int( n ) - 1;
    const int K = ( top + Ci ) / Ci;
    const uint32_t w16 = sq * uint32_t( sizeof( T ) ) / 16u;       // 16-byte words per slice
    auto lo_of  = [&]( int k ) { const int h = top - k * Ci; return h - Ci + 1 > 0 ? h - Ci + 1 : 0; };
    auto cnt_of = [&]( int k ) { return top - k * Ci - lo_of( k ) + 1; };
    T *const buf0 = lds, *const buf1 = lds + size_t( C ) * sq;
    dma_copy16<BLOCK>( buf0, qhist + size_t( lo_of( 0 ) ) * sq, uint32_t( cnt_of( 0 ) ) * w16 );
    dma_publish();
    for ( int k = 0; k < K; ++k )
    {
        T *const cur = ( k & 1 ) ? buf1 : buf0;
        T *const nxt = ( k & 1 ) ? buf0 : buf1;
        if ( k + 1 < K ) dma_copy16<BLOCK>( nxt, qhist + size_t( lo_of( k + 1 ) ) * sq, uint32_t( cnt_of( k + 1 ) ) * w16 );
        const int cnt = cnt_of( k );
        steps( cur, lo_of( k ), cnt - 1, 0 );
        dma_publish();
    }
}

// The same loop with one LDS buffer of C slices (tables too wide for two):
// each chunk is copied, published, then traced.
template <int BLOCK, typename T, typename STEPS>
__device__ __forceinline__ void poly_ready_direct( T *lds, const T *qhist, uint32_t sq, uint64_t n, uint32_t C,
                                                   STEPS steps )
{
    const uint32_t w16 = sq * uint32_t( sizeof( T ) ) / 16u;
    for ( int mhi = int( n ) - 1; mhi >= 0; mhi -= int( C ) )
    {
        const int mlo = mhi - int( C ) + 1 > 0 ? mhi - int( C ) + 1 : 0;
        const int cnt = mhi - mlo + 1;
        __syncthreads();                            // the previous chunk's reads are done
        dma_copy16<BLOCK>( lds, qhist + size_t( mlo ) * sq, uint32_t( cnt ) * w16 );
        dma_publish();
        steps( lds, mlo, cnt - 1, 0 );
    }
}

// exact re-wrap of a position into [0, Nx) (Nx*w is an integer)
__device__ __forceinline__ void rewrap( double &s, double Nxd, double invNx )
{
    s = fma( -Nxd, floor( s * invNx ), s );
}

} // namespace nufi_dev

#endif
