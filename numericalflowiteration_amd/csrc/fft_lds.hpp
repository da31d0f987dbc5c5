// fft_lds.hpp — batched radix-2 Stockham FFT over LDS buffers (a direct DFT
// for an axis whose length is not a power of two), shared by the field kernels
// and the adjoint field kernel.  All FIELD_BLOCK threads of the workgroup take
// part; every stage ends with a barrier.
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
    const uint32_t half = L >> 1, lh = a.log2L - 1;
    const uint32_t total = a.nlines * half;
    uint32_t log2Ns = 0;
    for ( uint32_t Ns = 1; Ns < L; Ns <<= 1, ++log2Ns )
    {
        const uint32_t twstep = L >> ( log2Ns + 1 );       // L / (2 Ns)
        for ( uint32_t e = threadIdx.x; e < total; e += FIELD_BLOCK )
        {
            const uint32_t line = e >> lh, j = e & ( half - 1 );
            const uint32_t base = line_base( a, line );
            const double2 x0 = in[ base + j*a.S ];
            double2 x1 = in[ base + ( j + half )*a.S ];
            const uint32_t k = j & ( Ns - 1 );
            double2 w = tw[ k * twstep ];
            if ( inverse ) w.y = -w.y;
            x1 = cmul( x1, w );
            const uint32_t o = ( ( j >> log2Ns ) << ( log2Ns + 1 ) ) + k;
            out[ base + o*a.S ]        = make_double2( x0.x + x1.x, x0.y + x1.y );
            out[ base + ( o + Ns )*a.S ] = make_double2( x0.x - x1.x, x0.y - x1.y );
        }
        __syncthreads();
        double2 *t = in; in = out; out = t;
    }
    return in;
}

} // namespace nufi_impl

#endif
