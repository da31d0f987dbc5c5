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

// ---- kick-polynomial tables built in LDS ----------------------------------
// Entry j (cell j-1) of a slice's table from the entries j, j+1, j+2 of the
// slice's extended difference row (entry e holds c[e] - c[e-1], periodic):
// Q0 = D0 + D1, Q1 = 2 (D1 - D0), Q2 = (D0 - D1) + (D2 - D1).  An FP32 table is
// scaled by the kick in FP64 before its one rounding to float; an FP64 table
// is not (the step applies the kick).
template <typename T>
__device__ __forceinline__ void put_poly( T *pair, T *q2p, double D0, double D1, double D2, double kick )
{
    const double t1 = D1 - D0;
    double q0 = D0 + D1, q1 = t1 + t1, q2 = ( D2 - D1 ) - t1;
    if ( sizeof( T ) < sizeof( double ) ) { q0 *= kick; q1 *= kick; q2 *= kick; }
    pair[0] = T( q0 ); pair[1] = T( q1 ); *q2p = T( q2 );
}

// The difference windows of one chunk (cnt slices from mlo, `cells` = Nx+2
// windows per slice) held in registers, MAXC windows per thread.  Window c of
// a thread is number threadIdx.x + c*BLOCK of the chunk, slice jj, cell j; its
// offsets are the same in every chunk and are computed once.
template <int BLOCK, int MAXC>
struct poly_fetch
{
    double d[MAXC][3];
    uint32_t goff[MAXC];   // jj*sd + j: first difference of the window from the chunk's first row
    uint32_t loff[MAXC];   // jj*sq + j: the pair is at loff + j, Q2 at loff + q2
    uint32_t cell[MAXC];   // j

    __device__ __forceinline__ void init( uint32_t sd, uint32_t sq, uint32_t cells )
    {
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c )
        {
            const uint32_t idx = threadIdx.x + uint32_t( c * BLOCK );
            const uint32_t jj = idx / cells, j = idx - jj * cells;
            goff[c] = jj * sd + j; loff[c] = jj * sq + j; cell[c] = j;
        }
    }

    __host__ __device__ __forceinline__ static bool fits( uint32_t cnt, uint32_t cells )
    {
        return cnt * cells <= uint32_t( BLOCK * MAXC );
    }

    // A window belongs to a chunk of cnt slices iff jj < cnt, i.e. goff <
    // cnt*sd (j < sd).  The load is unconditional (a window outside the chunk
    // reads the chunk's first window): inside a branch each load would wait
    // for every load before it (vmcnt(0) at the join).
    __device__ __forceinline__ void load_one( int c, const double *base, uint32_t lim )
    {
        const double *D = base + ( goff[c] < lim ? goff[c] : 0u );
        d[c][0] = D[0]; d[c][1] = D[1]; d[c][2] = D[2];
    }

    template <typename T>
    __device__ __forceinline__ void store_one( int c, T *Ql, uint32_t q2, uint32_t lim, double kick ) const
    {
        if ( goff[c] < lim ) put_poly( Ql + loff[c] + cell[c], Ql + loff[c] + q2, d[c][0], d[c][1], d[c][2], kick );
    }
};

// the table of cnt slices from mlo built straight from global memory
template <int BLOCK, typename T>
__device__ __forceinline__ void build_poly( T *Ql, const double *dhist, uint32_t sd, uint32_t sq, uint32_t q2,
                                            int64_t mlo, int cnt, uint32_t cells, double kick )
{
    for ( uint32_t idx = threadIdx.x; idx < uint32_t( cnt ) * cells; idx += BLOCK )
    {
        const uint32_t jj = idx / cells, j = idx - jj * cells;
        const double *D = dhist + size_t( mlo + jj ) * sd + j;
        T *Q = Ql + jj * sq;
        put_poly( Q + 2*j, Q + q2 + j, D[0], D[1], D[2], kick );
    }
}

// The backward loop over slices n-1 .. 0 in chunks of C slices, highest
// first (chunk k: slices max(n-1-kC-C+1, 0) .. n-1-kC), n < 2^31.
// steps(Q, mlo, hi, lo) takes the steps of slices mlo+hi down to mlo+lo with
// the chunk's tables at Q (slice mlo+jj at Q + jj*sq).
//
// Pipelined (fits(C, cells), LDS for 2*C slices): the tables alternate
// between two LDS buffers.  While chunk k is traced from one buffer, the
// windows of chunk k+1 (loaded during chunk k-1) are turned into the other
// buffer and those of chunk k+2 are loaded, WPP windows per thread before
// each of the chunk's PARTS = MAXC/WPP parts of steps, so that the table work
// fills issue slots the latency-bound steps leave free (fewer parts cost less
// scalar bookkeeping per step: WPP = 3 suits the FP64 step, 2 the FP32C one).  One barrier per chunk: chunk
// k+1's buffer was last read in chunk k-1.
//
// Every window's load is issued unconditionally, MAXC per thread and chunk
// (windows past the chunk, or past the last chunk, read a valid window they
// do not use): with straight-line loads the wait before turning window c into
// the table is for the load issued one chunk earlier, not for the ones just
// issued.
template <int BLOCK, int MAXC, int WPP, typename T, typename STEPS>
__device__ __forceinline__ void poly_pipeline( T *lds, const double *dhist, uint32_t sd, uint32_t sq, uint32_t q2,
                                               uint32_t cells, double kick, uint64_t n, uint32_t C, STEPS steps )
{
    if ( n == 0 ) return;
    constexpr int PARTS = ( MAXC + WPP - 1 ) / WPP;
    poly_fetch<BLOCK,MAXC> pf;
    pf.init( sd, sq, cells );
    const int Ci = int( C ), top = int( n ) - 1;
    const int K = ( top + Ci ) / Ci;
    auto lo_of  = [&]( int k ) { const int h = top - k * Ci; return h - Ci + 1 > 0 ? h - Ci + 1 : 0; };
    auto cnt_of = [&]( int k ) { return top - k * Ci - lo_of( k ) + 1; };
    const int P = ( Ci + PARTS - 1 ) / PARTS;          // steps per part
    T *const buf0 = lds, *const buf1 = lds + size_t( C ) * sq;

    {   // chunk 0 into buf0, chunk 1 on its way
        const double *b0 = dhist + size_t( lo_of( 0 ) ) * sd;
        const uint32_t l0 = uint32_t( cnt_of( 0 ) ) * sd;
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c ) pf.load_one( c, b0, l0 );
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c ) pf.store_one( c, buf0, q2, l0, kick );
        const double *b1 = dhist + size_t( K > 1 ? lo_of( 1 ) : 0 ) * sd;
        const uint32_t l1 = K > 1 ? uint32_t( cnt_of( 1 ) ) * sd : 0u;
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c ) pf.load_one( c, b1, l1 );
        __syncthreads();
    }
    for ( int k = 0; k < K; ++k )
    {
        T *const cur = ( k & 1 ) ? buf1 : buf0;
        T *const nxt = ( k & 1 ) ? buf0 : buf1;
        const int mlo = lo_of( k ), cnt = cnt_of( k );
        const uint32_t l1 = k + 1 < K ? uint32_t( cnt_of( k + 1 ) ) * sd : 0u;    // 0: no window stored
        const double *b2 = dhist + size_t( k + 2 < K ? lo_of( k + 2 ) : 0 ) * sd;
        const uint32_t l2 = k + 2 < K ? uint32_t( cnt_of( k + 2 ) ) * sd : 0u;
        #pragma unroll
        for ( int part = 0; part < PARTS; ++part )
        {
            #pragma unroll
            for ( int w = 0; w < WPP; ++w )
            {
                const int c = part * WPP + w;
                if ( c < MAXC ) { pf.store_one( c, nxt, q2, l1, kick ); pf.load_one( c, b2, l2 ); }
            }
            const int hi = cnt - 1 - part * P;
            const int lo = cnt - ( part + 1 ) * P > 0 ? cnt - ( part + 1 ) * P : 0;
            if ( hi >= lo ) steps( cur, mlo, hi, lo );
        }
        __syncthreads();
    }
}

// The same loop over tables already in global memory in the LDS layout (slice
// m at qhist + m*sq, sq*sizeof(T) a multiple of 16 bytes): while chunk k is
// traced from one LDS buffer, chunk k+1 is copied into the other by direct
// global->LDS loads, which take no registers and no VALU; each wave waits for
// its own copies and the workgroup barrier publishes them at the chunk's end,
// when chunk k's buffer is also free again.
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

// The same loop with one LDS buffer of C slices built directly per chunk
// (tables too wide for the prefetch or for two buffers).
template <int BLOCK, typename T, typename STEPS>
__device__ __forceinline__ void poly_direct( T *lds, const double *dhist, uint32_t sd, uint32_t sq, uint32_t q2,
                                             uint32_t cells, double kick, uint64_t n, uint32_t C, STEPS steps )
{
    for ( int mhi = int( n ) - 1; mhi >= 0; mhi -= int( C ) )
    {
        const int mlo = mhi - int( C ) + 1 > 0 ? mhi - int( C ) + 1 : 0;
        const int cnt = mhi - mlo + 1;
        __syncthreads();
        build_poly<BLOCK>( lds, dhist, sd, sq, q2, mlo, cnt, cells, kick );
        __syncthreads();
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
