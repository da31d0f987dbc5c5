// host_tables.hpp — host-side precomputation (long double) of the Fourier
// tables of the field solve: per-axis twiddles, the combined Poisson x fit
// multiplier 1/(N kappa^2 s_k), the Poisson-only and fit-only multipliers and
// the spectral energy weights (poisson.cpp:73-88).
#ifndef NUFI_HOST_TABLES_HPP
#define NUFI_HOST_TABLES_HPP

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/nufi_hip.h"

namespace nufi_impl
{

// B-spline values at t = 0 of the `order` bases that are non-zero on [0,1)
// (the collocation weights of fields.hpp:70-75), by the Cox-de Boor recurrence.
inline std::vector<long double> basis_at_zero( int order )
{
    std::vector<long double> w( order, 0.0L );
    w[order-1] = 1.0L;
    const long double t = 0.0L;
    for ( int m = 1; m < order; ++m )
    {
        w[order-m-1] = ( 1.0L - t ) * w[order-m];
        for ( int s = 1 - m; s < 0; ++s )
            w[order-1+s] = ( t - s ) * w[order-1+s] + ( ( m + 1 + s ) - t ) * w[order+s];
        w[order-1] *= t;
    }
    long double f = 1.0L;
    for ( int i = 2; i <= order - 1; ++i ) f *= i;
    for ( auto &v : w ) v /= f;
    return w;
}

inline std::vector<double2> twiddles( uint64_t L )
{
    std::vector<double2> tw( L ? L : 1 );
    const long double two_pi = 6.283185307179586476925286766559L;
    for ( uint64_t m = 0; m < L; ++m )
    {
        const long double a = -two_pi * (long double)m / (long double)L;
        tw[m] = make_double2( (double)std::cos( a ), (double)std::sin( a ) );
    }
    return tw;
}

struct fourier_tables
{
    std::vector<double2> ms, mp, mf;    // step (Poisson x fit), Poisson-only, fit-only
    std::vector<double>  ew;            // energy weights
    std::vector<double2> tx, ty, tz;    // forward twiddles per axis
};

inline fourier_tables make_tables( const nufi_config &g, int dim, int order )
{
    const uint64_t Nx = g.Nx, Ny = dim == 3 ? g.Ny : 1, Nz = dim == 3 ? g.Nz : 1;
    const uint64_t N = Nx * Ny * Nz;
    const std::vector<long double> B = basis_at_zero( order );
    const long double two_pi = 6.283185307179586476925286766559L;
    using cld = std::complex<long double>;
    auto symbol = [&]( uint64_t k, uint64_t L ) {
        cld s = 0;
        for ( int j = 0; j < order; ++j )
        {
            const long double a = two_pi * (long double)( ( uint64_t( j ) * k ) % L ) / (long double)L;
            s += B[j] * cld( std::cos( a ), std::sin( a ) );
        }
        return s;
    };
    auto fold = []( uint64_t i, uint64_t L ) { return (long double)( ( 2*i < L ) ? i : L - i ); };
    const long double Lx = g.Lx, Ly = dim == 3 ? g.Ly : 1.0L, Lz = dim == 3 ? g.Lz : 1.0L;
    const long double fx = two_pi / Lx, fy = two_pi / Ly, fz = two_pi / Lz;
    const long double vol = Lx * Ly * Lz;

    std::vector<double2> ms( N ), mp( N ), mf( N );
    std::vector<double> ew( N );
    for ( uint64_t k = 0; k < Nz; ++k )
    for ( uint64_t j = 0; j < Ny; ++j )
    for ( uint64_t i = 0; i < Nx; ++i )
    {
        const uint64_t idx = ( k * Ny + j ) * Nx + i;
        const cld s = symbol( i, Nx ) * symbol( j, Ny ) * symbol( k, Nz );
        const long double ii = fold( i, Nx ) * fx, jj = fold( j, Ny ) * fy, kk = fold( k, Nz ) * fz;
        const long double kap2 = ii*ii + jj*jj + kk*kk;
        const bool singular = std::abs( s ) < 1e-12L;
        const cld fit = singular ? cld( 0 ) : cld( 1 ) / ( (long double)N * s );
        mf[idx] = make_double2( (double)fit.real(), (double)fit.imag() );
        if ( idx == 0 )
        {
            ms[idx] = make_double2( 0.0, 0.0 );
            mp[idx] = make_double2( 0.0, 0.0 );
            ew[idx] = 0.0;
            continue;
        }
        const long double pois = 1.0L / ( (long double)N * kap2 );
        mp[idx] = make_double2( (double)pois, 0.0 );
        const cld st = singular ? cld( 0 ) : cld( 1 ) / ( (long double)N * kap2 * s );
        ms[idx] = make_double2( (double)st.real(), (double)st.imag() );
        ew[idx] = (double)( ( vol / 2.0L ) / ( (long double)N * (long double)N * kap2 ) );
    }
    fourier_tables t;
    t.ms = std::move( ms ); t.mp = std::move( mp ); t.mf = std::move( mf ); t.ew = std::move( ew );
    t.tx = twiddles( Nx ); t.ty = twiddles( Ny ); t.tz = twiddles( Nz );
    return t;
}

} // namespace nufi_impl

#endif
