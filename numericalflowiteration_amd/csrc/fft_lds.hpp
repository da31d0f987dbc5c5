// fft_lds.hpp — batched Stockham FFT over LDS buffers (radix 4, with one
// radix-2 stage for odd log2 L; a direct DFT for an axis whose length is not a
// power of two), shared by the field kernels and the adjoint field kernel.  All
// FIELD_BLOCK threads of the workgroup take part; every stage ends with a
// barrier.
#ifndef NUFI_FFT_LDS_HPP
#define NUFI_FFT_LDS_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nufi_impl
{

constexpr int FIELD_BLOCK = 512;

__device__ __forceinline__ double2 cmul( double2 a, double2 b )
{
    return make_double2( a.x*b.x - a.y*b.y, a.x*b.y + a.y*b.x );
}

// twiddles an axis of length L reads: exp(-2 pi i k/L) for k < L/2 (radix 2)
// or k < L (direct DFT); none for L <= 1
__host__ __device__ inline uint32_t twiddle_entries( uint32_t L, uint32_t log2L )
{
    if ( L <= 1 ) return 0;
    return log2L == 0xffffffffu ? L : L / 2;
}

struct axis_desc
{
    uint32_t L, log2L, S;       // length, log2 (or ~0u), stride
    uint32_t nlines;
    uint32_t mode;              // 0: x lines, 1: y lines, 2: z lines
    uint32_t Nx, Ny;
};

__device__ __forceinline__ uint32_t line_base( const axis_desc &a, uint32_t line )
{
    if ( a.mode == 0 ) return line * a.Nx;
    if ( a.mode == 1 ) return ( line / a.Nx ) * a.Nx * a.Ny + ( line % a.Nx );
    return line;
}

// Transform all lines along one axis; returns the buffer holding the result.
__device__ inline double2 *fft_axis( double2 *in, double2 *out, const axis_desc &a, const double2 *tw, bool inverse )
{
    const uint32_t L = a.L;
    if ( L <= 1 ) return in;
    if ( a.log2L == 0xffffffffu )
    {
        // direct DFT, O(L^2) per line, exact angle reduction j*k mod L
        const uint32_t total = a.nlines * L;
        for ( uint32_t e = threadIdx.x; e < total; e += FIELD_BLOCK )
        {
            const uint32_t line = e / L, k = e % L;
            const uint32_t base = line_base( a, line );
            double2 acc = make_double2( 0.0, 0.0 );
            uint32_t m = 0;
            for ( uint32_t j = 0; j < L; ++j )
            {
                double2 w = tw[m];
                if ( inverse ) w.y = -w.y;
                const double2 v = cmul( in[ base + j*a.S ], w );
                acc.x += v.x; acc.y += v.y;
                m += k; if ( m >= L ) m -= L;
            }
            out[ base + k*a.S ] = acc;
        }
        __syncthreads();
        return out;
    }
    // Stockham: one radix-2 stage when log2 L is odd, then radix-4 stages
    // (half the barriers of radix 2).  The table holds exp(-2 pi i m/L) for
    // m < L/2; m >= L/2 reads -tw[m - L/2].
    auto twid = [&]( uint32_t m )
    {
        const uint32_t hl = L >> 1;
        double2 w = m < hl ? tw[m] : make_double2( -tw[m - hl].x, -tw[m - hl].y );
        if ( inverse ) w.y = -w.y;
        return w;
    };
    uint32_t Ns = 1, log2Ns = 0;
    if ( a.log2L & 1u )
    {
        const uint32_t half = L >> 1, lh = a.log2L - 1;
        for ( uint32_t e = threadIdx.x; e < a.nlines * half; e += FIELD_BLOCK )
        {
            const uint32_t line = e >> lh, j = e & ( half - 1 );
            const uint32_t base = line_base( a, line );
            const double2 x0 = in[ base + j*a.S ], x1 = in[ base + ( j + half )*a.S ];   // Ns = 1: no twiddle
            out[ base + ( 2*j )*a.S ]     = make_double2( x0.x + x1.x, x0.y + x1.y );
            out[ base + ( 2*j + 1 )*a.S ] = make_double2( x0.x - x1.x, x0.y - x1.y );
        }
        __syncthreads();
        double2 *t = in; in = out; out = t;
        Ns = 2; log2Ns = 1;
    }
    const uint32_t quarter = L >> 2, lq = a.log2L - 2;
    for ( ; Ns < L; Ns <<= 2, log2Ns += 2 )
    {
        const uint32_t twstep = L >> ( log2Ns + 2 );       // L / (4 Ns)
        for ( uint32_t e = threadIdx.x; e < a.nlines * quarter; e += FIELD_BLOCK )
        {
            const uint32_t line = e >> lq, j = e & ( quarter - 1 );
            const uint32_t base = line_base( a, line );
            const uint32_t k = j & ( Ns - 1 );
            const double2 b0 = in[ base + j*a.S ];
            const double2 b1 = cmul( in[ base + ( j + quarter )*a.S ],     twid( k * twstep ) );
            const double2 b2 = cmul( in[ base + ( j + 2*quarter )*a.S ],   twid( 2 * k * twstep ) );
            const double2 b3 = cmul( in[ base + ( j + 3*quarter )*a.S ],   twid( 3 * k * twstep ) );
            const double2 s02 = make_double2( b0.x + b2.x, b0.y + b2.y ), d02 = make_double2( b0.x - b2.x, b0.y - b2.y );
            const double2 s13 = make_double2( b1.x + b3.x, b1.y + b3.y ), d13 = make_double2( b1.x - b3.x, b1.y - b3.y );
            // forward: -i (b1 - b3); inverse: +i (b1 - b3)
            const double2 rot = inverse ? make_double2( -d13.y, d13.x ) : make_double2( d13.y, -d13.x );
            const uint32_t o = ( ( j >> log2Ns ) << ( log2Ns + 2 ) ) + k;
            out[ base + o*a.S ]            = make_double2( s02.x + s13.x, s02.y + s13.y );
            out[ base + ( o + Ns )*a.S ]   = make_double2( d02.x + rot.x, d02.y + rot.y );
            out[ base + ( o + 2*Ns )*a.S ] = make_double2( s02.x - s13.x, s02.y - s13.y );
            out[ base + ( o + 3*Ns )*a.S ] = make_double2( d02.x - rot.x, d02.y - rot.y );
        }
        __syncthreads();
        double2 *t = in; in = out; out = t;
    }
    return in;
}

} // namespace nufi_impl

#endif
