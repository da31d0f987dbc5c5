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

} // namespace nufi_dev

#endif
