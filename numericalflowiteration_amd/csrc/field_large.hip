// field_large.hip — the field solve for grids whose two complex buffers do not
// fit one workgroup's LDS (more than 5104 points): the same operator as
// field.hip (rho -> spectral energy -> Poisson x fit multiplier -> slice; or
// the Poisson-only / fit-only multipliers) with the transforms done as one
// direct DFT pass per axis between two global complex buffers, one thread per
// output value (O(N_axis) per value: 3D axes are short, and a 1D line is
// capped at 32768 points), over many workgroups.
//
//   load      W0 = rho (from the partial sums, or the given values)
//   axis x,y,z (forward, e^{-2 pi i jk/L})             W0 -> W1 -> W0 -> W1
//   spectrum  energy partials sum_k w_k |F_k|^2 per workgroup, F_k *= m_k
//   energy    the partials summed in workgroup order (compensated), one workgroup
//   axis x,y,z (inverse, unnormalised: m_k carries 1/N)
//   store     the slice (periodically padded), 1D differences and kick-
//             polynomial tables, the 3D FP32 copy; or the nodal values
//
// The 1D adjoint's field step (adjoint1d.hip) uses the same passes.
#include "device_common.hpp"
#include "fft_lds.hpp"
#include "internal.hpp"
#include "step1d.hpp"

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

constexpr int LB = 256;   // threads per workgroup of every pass

__device__ __forceinline__ double sum_partials_l( const double *__restrict__ partials, uint32_t wpx, size_t n,
                                                  uint32_t i )
{
    csum acc { 0.0, 0.0 };
    for ( uint32_t j = 0; j < wpx; ++j ) acc = csum_add( acc, csum { partials[ size_t( j ) * n + i ], 0.0 } );
    return acc.s + acc.c;
}

__global__ __launch_bounds__( LB ) void large_load_kernel( field_params p, double2 *W )
{
    const uint32_t i = blockIdx.x * LB + threadIdx.x;
    if ( i >= p.N ) return;
    const double v = ( p.mode == FIELD_STEP && p.partials != nullptr ) ? p.weight * sum_partials_l( p.partials, p.wpx, p.N, i )
                                                                      : p.values[i];
    if ( p.rho_out ) p.rho_out[i] = v;
    W[i] = make_double2( v, 0.0 );
}

// out[base + k*stride] = sum_j in[base + j*stride] w^(jk), w = tw[1] (inverse: conj)
__global__ __launch_bounds__( LB ) void large_dft_kernel( const double2 *__restrict__ in, double2 *__restrict__ out,
                                                          uint32_t N, uint32_t L, uint32_t stride,
                                                          const double2 *__restrict__ tw, int inverse )
{
    const uint32_t idx = blockIdx.x * LB + threadIdx.x;
    if ( idx >= N ) return;
    const uint32_t k = ( idx / stride ) % L;
    const uint32_t base = idx - k * stride;
    const double sg = inverse ? -1.0 : 1.0;
    double re = 0.0, im = 0.0;
    uint32_t m = 0;                                       // j*k mod L
    for ( uint32_t j = 0; j < L; ++j )
    {
        const double2 a = in[ base + j * stride ];
        const double2 w = tw[m];
        const double wi = sg * w.y;
        re = fma( a.x, w.x, fma( -a.y, wi, re ) );
        im = fma( a.x, wi, fma( a.y, w.x, im ) );
        m += k;
        if ( m >= L ) m -= L;
    }
    out[idx] = make_double2( re, im );
}

// scale = 0: F_k *= mult_k, energy partial sum_k ewt_k |F_k|^2 (before the scaling);
// scale = 1: F_k *= 2 dt ewt_k (no energy); scale = 2: F_k *= conj(mult_k)
__global__ __launch_bounds__( LB ) void large_spectrum_kernel( double2 *F, uint32_t N, const double2 *mult,
                                                               const double *ewt, double dt, int scale, double *epart )
{
    __shared__ double red[LB / WAVE];
    const uint32_t k = blockIdx.x * LB + threadIdx.x;
    double e = 0.0;
    if ( k < N )
    {
        const double2 f = F[k];
        if ( scale == 0 )
        {
            if ( ewt ) e = ewt[k] * ( f.x*f.x + f.y*f.y );
            F[k] = cmul( f, mult[k] );
        }
        else if ( scale == 1 )
        {
            const double w = 2.0 * dt * ewt[k];
            F[k] = make_double2( w * f.x, w * f.y );
        }
        else F[k] = cmul( f, make_double2( mult[k].x, -mult[k].y ) );
    }
    if ( epart == nullptr ) return;
    e = wave_sum( e );
    if ( threadIdx.x % WAVE == 0 ) red[ threadIdx.x / WAVE ] = e;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        double s = 0.0;
        for ( int w = 0; w < LB / WAVE; ++w ) s += red[w];
        epart[ blockIdx.x ] = s;
    }
}

__global__ __launch_bounds__( LB ) void large_energy_kernel( const double *epart, uint32_t nb, double *energy )
{
    __shared__ double rs[LB], rc[LB];
    csum acc { 0.0, 0.0 };
    for ( uint32_t b = threadIdx.x; b < nb; b += LB ) acc = csum_add( acc, csum { epart[b], 0.0 } );
    rs[threadIdx.x] = acc.s; rc[threadIdx.x] = acc.c;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        csum t { 0.0, 0.0 };
        for ( int i = 0; i < LB; ++i ) t = csum_add( t, csum { rs[i], rc[i] } );
        *energy = t.s + t.c;
    }
}

// the outputs of field_body's last stage, one thread per output entry
__global__ __launch_bounds__( LB ) void large_store_kernel( field_params p, const double2 *__restrict__ R )
{
    const uint32_t e2 = blockIdx.x * LB + threadIdx.x;
    if ( p.mode == FIELD_POISSON )
    {
        if ( e2 < p.N ) p.values[e2] = R[e2].x;
        return;
    }
    const uint32_t o1 = uint32_t( p.order ) - 1;
    if ( p.dim == 1 )
    {
        const uint32_t Nx = p.Nx, len = Nx + o1;
        if ( e2 < len ) p.slice[e2] = R[ e2 % Nx ].x;
        if ( p.dslice && e2 < p.stride_d )
            p.dslice[e2] = ( e2 < len + 1 ) ? R[ e2 % Nx ].x - R[ ( e2 + Nx - 1 ) % Nx ].x : 0.0;
        if ( ( p.qslice || p.qslice32 ) && e2 < Nx + 2 )
        {
            double D[3];
            #pragma unroll
            for ( uint32_t l = 0; l < 3; ++l ) D[l] = R[ ( e2 + l ) % Nx ].x - R[ ( e2 + l + Nx - 1 ) % Nx ].x;
            if ( p.qslice )   put_poly( p.qslice + 2*e2, p.qslice + p.q2_off + e2, D[0], D[1], D[2], p.kick );
            if ( p.qslice32 ) put_poly( p.qslice32 + 2*e2, p.qslice32 + p.q2_off + e2, D[0], D[1], D[2], p.kick );
        }
        return;
    }
    const uint32_t ex = p.Nx + o1, ey = p.Ny + o1, ez = p.Nz + o1;
    if ( e2 >= ex * ey * ez ) return;
    const uint32_t i = e2 % ex, j = ( e2 / ex ) % ey, k = e2 / ( ex * ey );
    const double c = R[ ( ( k % p.Nz ) * p.Ny + ( j % p.Ny ) ) * p.Nx + ( i % p.Nx ) ].x;
    p.slice[ k * p.sz + j * p.sy + i ] = c;
    if ( p.slice32 ) p.slice32[ k * p.sz + j * p.sy + i ] = float( c * p.kick );
}

// g = R.x (accumulate: g += R.x)
__global__ __launch_bounds__( LB ) void large_real_kernel( const double2 *R, uint32_t N, double *g, int accumulate )
{
    const uint32_t i = blockIdx.x * LB + threadIdx.x;
    if ( i < N ) g[i] = accumulate ? g[i] + R[i].x : R[i].x;
}

inline uint32_t blocks( uint32_t n ) { return ( n + LB - 1 ) / LB; }

// the axes as (length, stride, twiddles): x fastest
struct axes3
{
    uint32_t L[3], S[3];
    const double2 *tw[3];
    int n;
};

axes3 axes_of( const field_params &p )
{
    axes3 a {};
    a.n = p.dim == 1 ? 1 : 3;
    a.L[0] = p.Nx; a.S[0] = 1;           a.tw[0] = p.tw_x;
    a.L[1] = p.Ny; a.S[1] = p.Nx;        a.tw[1] = p.tw_y;
    a.L[2] = p.Nz; a.S[2] = p.Nx * p.Ny; a.tw[2] = p.tw_z;
    return a;
}

// the transform of every axis; returns the buffer holding the result
double2 *dft_all( const axes3 &a, uint32_t N, double2 *in, double2 *other, int inverse, hipStream_t st )
{
    double2 *src = in, *dst = other;
    for ( int d = 0; d < a.n; ++d )
    {
        if ( a.L[d] == 1 ) continue;
        hipLaunchKernelGGL( large_dft_kernel, dim3( blocks( N ) ), dim3( LB ), 0, st, src, dst, N, a.L[d], a.S[d],
                            a.tw[d], inverse );
        double2 *t = src; src = dst; dst = t;
    }
    return src;
}

} // namespace

size_t field_large_partials( uint32_t N ) { return blocks( N ); }

hipError_t launch_field_large( const field_params &p, double2 *W0, double2 *W1, double *epart, hipStream_t st )
{
    const uint32_t N = p.N;
    const axes3 a = axes_of( p );
    hipLaunchKernelGGL( large_load_kernel, dim3( blocks( N ) ), dim3( LB ), 0, st, p, W0 );
    double2 *F = dft_all( a, N, W0, W1, 0, st );
    const bool energy = p.mode != FIELD_FIT && p.energy != nullptr;
    hipLaunchKernelGGL( large_spectrum_kernel, dim3( blocks( N ) ), dim3( LB ), 0, st, F, N, p.mult,
                        p.mode != FIELD_FIT ? p.ewt : nullptr, 0.0, 0, energy ? epart : nullptr );
    if ( energy ) hipLaunchKernelGGL( large_energy_kernel, dim3( 1 ), dim3( LB ), 0, st, epart, blocks( N ), p.energy );
    double2 *R = dft_all( a, N, F, F == W0 ? W1 : W0, 1, st );
    const uint32_t o1 = uint32_t( p.order ) - 1;
    uint32_t outs = N;
    if ( p.mode != FIELD_POISSON )
        outs = p.dim == 1 ? std::max( p.Nx + o1 + 1, p.stride_d )
                          : ( p.Nx + o1 ) * ( p.Ny + o1 ) * ( p.Nz + o1 );
    hipLaunchKernelGGL( large_store_kernel, dim3( blocks( outs ) ), dim3( LB ), 0, st, p, R );
    return hipGetLastError();
}

hipError_t launch_adjoint_field_large( const adjoint_field_params &q, double2 *W0, double2 *W1, hipStream_t st )
{
    // g = 2 dt Re F^H( w . F rho ) + Re F^H( conj(m) . F lambda ), N = Nx
    field_params p {};
    p.dim = 1; p.Nx = q.N; p.Ny = 1; p.Nz = 1; p.N = q.N; p.tw_x = q.tw;
    const axes3 a = axes_of( p );
    for ( int pass = 0; pass < 2; ++pass )
    {
        field_params l = p;
        l.mode = FIELD_POISSON; l.values = const_cast<double*>( pass == 0 ? q.rho : q.lambda );
        hipLaunchKernelGGL( large_load_kernel, dim3( blocks( q.N ) ), dim3( LB ), 0, st, l, W0 );
        double2 *F = dft_all( a, q.N, W0, W1, 0, st );
        hipLaunchKernelGGL( large_spectrum_kernel, dim3( blocks( q.N ) ), dim3( LB ), 0, st, F, q.N, q.mult, q.ewt,
                            q.dt, pass == 0 ? 1 : 2, static_cast<double*>( nullptr ) );
        double2 *R = dft_all( a, q.N, F, F == W0 ? W1 : W0, 1, st );
        hipLaunchKernelGGL( large_real_kernel, dim3( blocks( q.N ) ), dim3( LB ), 0, st, R, q.N, q.g, pass );
    }
    return hipGetLastError();
}

} // namespace nufi_impl
