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
// Q + 2*(i+1), Q2 at Q + q2 + (i+1), i = -1 .. Nx.
// u += h * (Q0 + Q1 t + Q2 t^2), h = 1, or 1/2 on slice 0.  With POW2 the
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

template <bool POW2>
__device__ __forceinline__ void poly_step( double &s, double &u, const double *Q, uint32_t q2, double cd, double h,
                                           int mask, double Nxd, double invNx )
{
    double x = fma( -cd, u, s );
    double t;
    const int i = poly_locate<POW2>( x, mask, Nxd, invNx, t );
    const double2 q01 = reinterpret_cast<const double2*>( Q )[ i + 1 ];
    const double c2 = Q[ q2 + ( i + 1 ) ];
    u = fma( h, fma( fma( c2, t, q01.y ), t, q01.x ), u );
    s = x;
}

// ---- kick-polynomial tables built in LDS ----------------------------------
// Entry j (cell j-1) of a slice's table from the entries j, j+1, j+2 of the
// slice's extended difference row (entry e holds c[e] - c[e-1], periodic):
// Q0 = kick (D0 + D1), Q1 = kick 2 (D1 - D0), Q2 = kick ((D0 - D1) + (D2 - D1)).
template <typename T>
__device__ __forceinline__ void put_poly( T *Q, uint32_t q2, uint32_t j, double D0, double D1, double D2, double kick )
{
    Q[2*j]     = T( kick * ( D0 + D1 ) );
    Q[2*j + 1] = T( kick * ( 2.0 * ( D1 - D0 ) ) );
    Q[q2 + j]  = T( kick * ( ( D0 - D1 ) + ( D2 - D1 ) ) );
}

// The difference windows of one chunk (cnt slices from mlo, `cells` = Nx+2
// entries per slice) held in registers, MAXC windows per thread: loaded while
// the previous chunk is traced, turned into the LDS table after its barrier.
template <int BLOCK, int MAXC>
struct poly_fetch
{
    double d[MAXC][3];

    __device__ __forceinline__ static bool fits( int cnt, uint32_t cells )
    {
        return uint32_t( cnt ) * cells <= uint32_t( BLOCK * MAXC );
    }

    __device__ __forceinline__ void load( const double *dhist, uint32_t sd, int64_t mlo, int cnt, uint32_t cells )
    {
        const uint32_t total = uint32_t( cnt ) * cells;
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c )
        {
            const uint32_t idx = threadIdx.x + uint32_t( c * BLOCK );
            if ( idx < total )
            {
                const uint32_t jj = idx / cells, j = idx - jj * cells;
                const double *D = dhist + size_t( mlo + jj ) * sd + j;
                d[c][0] = D[0]; d[c][1] = D[1]; d[c][2] = D[2];
            }
        }
    }

    template <typename T>
    __device__ __forceinline__ void store( T *Ql, uint32_t sq, uint32_t q2, int cnt, uint32_t cells, double kick ) const
    {
        const uint32_t total = uint32_t( cnt ) * cells;
        #pragma unroll
        for ( int c = 0; c < MAXC; ++c )
        {
            const uint32_t idx = threadIdx.x + uint32_t( c * BLOCK );
            if ( idx < total )
            {
                const uint32_t jj = idx / cells, j = idx - jj * cells;
                put_poly( Ql + jj * sq, q2, j, d[c][0], d[c][1], d[c][2], kick );
            }
        }
    }
};

// the same table built straight from global memory (no prefetch)
template <int BLOCK, typename T>
__device__ __forceinline__ void build_poly( T *Ql, const double *dhist, uint32_t sd, uint32_t sq, uint32_t q2,
                                            int64_t mlo, int cnt, uint32_t cells, double kick )
{
    for ( uint32_t idx = threadIdx.x; idx < uint32_t( cnt ) * cells; idx += BLOCK )
    {
        const uint32_t jj = idx / cells, j = idx - jj * cells;
        const double *D = dhist + size_t( mlo + jj ) * sd + j;
        put_poly( Ql + jj * sq, q2, j, D[0], D[1], D[2], kick );
    }
}

// Table chunks of the backward loop, highest slices first: build the chunk
// (from the prefetch when it fits), then start loading the next one.
template <int BLOCK, int MAXC, typename T>
__device__ __forceinline__ void poly_chunk_begin( poly_fetch<BLOCK,MAXC> &pf, T *Ql, const double *dhist,
                                                  uint32_t sd, uint32_t sq, uint32_t q2, uint32_t cells,
                                                  double kick, int64_t mhi, uint32_t chunk )
{
    const int64_t mlo = mhi - int64_t( chunk ) + 1 > 0 ? mhi - int64_t( chunk ) + 1 : 0;
    const int cnt = int( mhi - mlo + 1 );
    __syncthreads();                                   // the previous chunk is no longer read
    if ( pf.fits( cnt, cells ) ) pf.store( Ql, sq, q2, cnt, cells, kick );
    else build_poly<BLOCK>( Ql, dhist, sd, sq, q2, mlo, cnt, cells, kick );
    __syncthreads();
    const int64_t nhi = mlo - 1;
    if ( nhi >= 0 )
    {
        const int64_t nlo = nhi - int64_t( chunk ) + 1 > 0 ? nhi - int64_t( chunk ) + 1 : 0;
        const int ncnt = int( nhi - nlo + 1 );
        if ( pf.fits( ncnt, cells ) ) pf.load( dhist, sd, nlo, ncnt, cells );
    }
}

template <int BLOCK, int MAXC>
__device__ __forceinline__ void poly_prefetch_first( poly_fetch<BLOCK,MAXC> &pf, const double *dhist, uint32_t sd,
                                                     uint32_t cells, int64_t mhi, uint32_t chunk )
{
    if ( mhi < 0 ) return;
    const int64_t mlo = mhi - int64_t( chunk ) + 1 > 0 ? mhi - int64_t( chunk ) + 1 : 0;
    const int cnt = int( mhi - mlo + 1 );
    if ( pf.fits( cnt, cells ) ) pf.load( dhist, sd, mlo, cnt, cells );
}

// exact re-wrap of a position into [0, Nx) (Nx*w is an integer)
__device__ __forceinline__ void rewrap( double &s, double Nxd, double invNx )
{
    s = fma( -Nxd, floor( s * invNx ), s );
}

} // namespace nufi_dev

#endif
