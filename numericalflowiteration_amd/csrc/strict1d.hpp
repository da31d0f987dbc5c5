// strict1d.hpp — the 1D spline evaluation in the reference's operation order
// (dim1::eval, nufi/fields.hpp:36-61, and splines1d::eval,
// nufi/splines.hpp:81-110), shared by the strict trace kernel and the
// diagnostics samplers.  No FMA contraction.
//
// D points at the extended first differences of a slice stored at offset 1:
// D[k] = c[k+1] - c[k] for k = -1 .. Nx+order-2 (periodic) — exactly the
// first differencing pass of splines.hpp:98-100.  A wrapped position that
// rounds onto Lx gives knot Nx and one that rounds below 0 gives -1; both read
// the periodic copies, so no read leaves the row.
#ifndef NUFI_STRICT1D_HPP
#define NUFI_STRICT1D_HPP

#include <hip/hip_runtime.h>

namespace nufi_dev
{

template <int K>
__host__ __device__ constexpr double inv_fact()
{
    double f = 1.0;
    for ( int i = 2; i <= K; ++i ) f *= double(i);
    return 1.0 / f;
}

// d^DERIV phi / dx^DERIV at x (DERIV = 1 or 2) from slice differences D
template <int ORDER, int DERIV>
__device__ __forceinline__ double phi_deriv_strict( double x, const double *D, double x_min, double Lx,
                                                    double Lx_inv, double dx_inv )
{
#pragma clang fp contract(off)
    static_assert( DERIV >= 1 && DERIV < ORDER, "derivative 1 .. order-1" );
    x = x - x_min;
    x = x - Lx * floor( x * Lx_inv );
    const double knot = floor( x * dx_inv );
    const int ii = int( knot );
    x = x * dx_inv - knot;
    double factor = 1.0;
    #pragma unroll
    for ( int i = 0; i < DERIV; ++i ) factor = factor * dx_inv;

    constexpr int n = ORDER, d = DERIV;
    double c[ORDER];
    #pragma unroll
    for ( int j = 1; j < n; ++j ) c[j] = D[ ii + j - 1 ];          // the first pass
    #pragma unroll
    for ( int j = 2; j <= d; ++j )                                 // further passes
        #pragma unroll
        for ( int i = n - 1; i >= j; --i ) c[i] = c[i] - c[i-1];
    #pragma unroll
    for ( int j = 1; j < n - d; ++j )
    {
        #pragma unroll
        for ( int i = n - d - 1; i >= j; --i )
        {
            const double lo = ( ( ( x + double(n) ) - double(d) ) - 1.0 ) - double(i);
            const double hi = double( i - j + 1 ) - x;
            c[d+i] = lo * c[d+i] + hi * c[d+i-1];
        }
    }
    return factor * ( inv_fact<n - d - 1>() * c[n-1] );
}

} // namespace nufi_dev

#endif
