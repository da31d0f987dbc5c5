// adjoint1d.hip — reverse-mode gradient of J = sum_n dt * energy_n (the time-
// integrated spectral electric energy) with respect to the parameters of the
// 1D initial condition, on the device, for the FP64 fast trace.
//
// Forward model (one time step n; the same arithmetic as trace1d.hip):
//   rho_n[i] = -du sum_u f0( foot_n(x_i, u; S_0..S_{n-1}); a )
//   S_n      = M rho_n          (Poisson + fit + periodic padding, linear)
//   e_n      = sum_k w_k |F rho_n|_k^2
// Reverse sweep, n = Nt .. 0, with lambda_m = dJ/dS_m (folded mod Nx):
//   g_n = 2 dt Re F^H( w . F rho_n ) + Re F^H( conj(m) . F lambda_n )  (= dJ/d rho_n)
//   per node: fbar = -du g_n[i]; trace forward to the foot; dJ/da += fbar df0/da;
//   then walk back through slices 0 .. n-1, undoing each Stormer-Verlet step
//   as it was taken (the states are recomputed, never stored), with the
//   polynomial p(t) = Q0 + Q1 t + Q2 t^2 of 2 dphi/dt of the forward kernel
//   (hk = h kick, h = 1, or 1/2 on slice 0):
//     forward  s_a = s_b - cd u_b,  u_a = u_b + hk p(s_a; S_m)
//     reverse  pbar = hk ubar;  sbar += pbar dp/dt;  lambda_m += pbar dp/dS_m;
//              u_b = u_a - hk p;  ubar += -cd sbar;  s_b = s_a + cd u_b
// lambda_m accumulates in LDS per workgroup (all nodes of a workgroup visit the
// same slice at the same time) and is flushed with global atomics once per
// chunk of slices; its summation order therefore varies at the rounding level.
//
// FP32C contexts differentiate the FP32-compensated trace, whose velocity w is
// in cells per step (w = cd u): the forward sweep repeats the FP32C kernel's
// steps (the shared fp32c_* helpers of step1d.hpp, so the foot points are
// those of the forward run), the reverse sweep undoes them on the same pairs
// (w_b = w_a - h q(s_a), s_b = s_a + w_b, both with TwoSum), and the adjoints
// sbar, wbar, lambda and the parameter gradients are FP64, as is the field
// adjoint.  The FP32 tables hold kick cd (Q0, Q1, Q2), so pbar = h kick cd wbar,
// sbar += h wbar dq/dt, wbar -= sbar, and the seed is wbar = fbar df0/du / cd.
#include "device_common.hpp"
#include "fft_lds.hpp"
#include "internal.hpp"
#include "step1d.hpp"

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

// d f0 / d(x, u) and d f0 / d(parameters) for the 1D families.  Parameters:
// alpha for Landau and two-stream; a_1..a_n for the multimode family.
__device__ inline void f0_grad_1d( const nufi_f0 &p, double x, double u, double &fx, double &fu,
                                   double *fa, int npar )
{
    const double e = exp( -u*u/2.0 );
    switch ( p.kind )
    {
    case NUFI_F0_LANDAU_1D:
    {
        const double c = cos( p.k*x ), s = sin( p.k*x );
        fx = kInvSqrt2Pi * ( -p.alpha * p.k * s ) * e;
        fu = kInvSqrt2Pi * ( 1.0 + p.alpha*c ) * e * ( -u );
        if ( npar > 0 ) fa[0] = kInvSqrt2Pi * c * e;
        break;
    }
    case NUFI_F0_TWOSTREAM_1D:
    {
        const double c = cos( p.k*x ), s = sin( p.k*x );
        fx = kInvSqrt2Pi * ( -p.alpha * p.k * s ) * e * u*u;
        fu = kInvSqrt2Pi * ( 1.0 + p.alpha*c ) * e * u * ( 2.0 - u*u );
        if ( npar > 0 ) fa[0] = kInvSqrt2Pi * c * e * u*u;
        break;
    }
    case NUFI_F0_MULTIMODE_1D:
    {
        double sm = 1.0, dsm = 0.0;
        #pragma unroll
        for ( int m = 1; m <= NUFI_MAX_MODES; ++m )     // static indices: no private-memory copy
        {
            if ( m > p.n_modes ) continue;
            const double km = p.k * m;
            const double c = cos( km * x ), s = sin( km * x );
            sm += p.a[m-1] * c;
            dsm -= p.a[m-1] * km * s;
            if ( m <= npar ) fa[m-1] = kInvSqrt2Pi * c * e;
        }
        fx = kInvSqrt2Pi * dsm * e;
        fu = kInvSqrt2Pi * sm * e * ( -u );
        break;
    }
    default:
        fx = fu = 0.0;
        #pragma unroll
        for ( int m = 0; m < NUFI_MAX_MODES; ++m ) if ( m < npar ) fa[m] = 0.0;
    }
}

template <int BLOCK, bool POW2>
__global__ __launch_bounds__( BLOCK )
void adjoint1d_kernel( adjoint1_params p )
{
    extern __shared__ double lds[];
    double *Ql  = lds;                                   // chunk of kick-polynomial slices
    double *lam = lds + size_t( p.chunk ) * p.stride_q;  // chunk of folded slice adjoints
    constexpr int WPB = BLOCK / WAVE;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const uint32_t g = blockIdx.x * WPB + wave;
    const bool task = g < p.n_tasks;
    const uint32_t ix = task ? g / p.wpx : 0, j = task ? g % p.wpx : 0;
    const uint32_t iu = j * WAVE + lane;
    const bool active = task && iu < p.Nu;

    const int Nx = int( p.Nx ), mask = Nx - 1;
    const double Nxd = p.Nx, invNx = 1.0 / Nxd;
    const double cd = p.dt * p.dx_inv, kick = -0.5 * p.dt * p.dx_inv;
    const uint32_t sq = p.stride_q, q2 = p.q2_off;

    double s = double( ix );
    double u = p.u_min + iu*p.du + p.du/2;

    // forward: slices n-1 .. 0, the step of trace1d_poly_body
    for ( int64_t mhi = int64_t( p.n ) - 1; mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        dma_copy16<BLOCK>( Ql, p.qhist + size_t( mlo ) * sq, uint32_t( cnt ) * sq / 2u );
        dma_publish();
        for ( int jj = cnt - 1; jj >= 0; )          // segments between re-wrap points
        {
            if ( POW2 && poly_wrap_due( mlo + jj ) ) rewrap( s, Nxd, invNx );
            const int jstop = POW2 ? int( poly_segment_end( mlo + jj, mlo ) - mlo ) : -1;
            for ( ; jj > jstop; --jj )
                poly_step<POW2>( s, u, Ql + jj * sq, Ql + jj * sq + q2, cd, ( mlo + jj == 0 ) ? 0.5 * kick : kick, mask,
                                 Nxd, invNx );
        }
    }

    // seed at the foot
    const double fbar = active ? -p.du * p.g[ix] : 0.0;
    double fx, fu, fa[NUFI_MAX_MODES];
    f0_grad_1d( p.f0, p.x_min + s*p.dx, u, fx, fu, fa, p.npar );
    double sbar = fbar * fx * p.dx;     // d/ds = dx d/dx
    double ubar = fbar * fu;
    #pragma unroll
    for ( int m = 0; m < NUFI_MAX_MODES; ++m )          // npar is uniform: the branch is too
    {
        if ( m >= int( p.npar ) ) continue;
        const csum v = wave_csum( csum { fbar * fa[m], 0.0 } );
        if ( task && lane == 0 ) p.grad_part[ size_t( g ) * NUFI_MAX_MODES + m ] += v.s + v.c;
    }

    // reverse: slices 0 .. n-1
    for ( int64_t mlo = 0; mlo < int64_t( p.n ); mlo += p.chunk )
    {
        const int64_t mhi = ( mlo + p.chunk - 1 < int64_t( p.n ) - 1 ) ? mlo + p.chunk - 1 : int64_t( p.n ) - 1;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        dma_copy16<BLOCK>( Ql, p.qhist + size_t( mlo ) * sq, uint32_t( cnt ) * sq / 2u );
        for ( uint32_t q = threadIdx.x; q < uint32_t( cnt ) * p.Nx; q += BLOCK ) lam[q] = 0.0;
        dma_publish();
        for ( int jj = 0; jj < cnt; ++jj )
        {
            const double hk = ( mlo + jj == 0 ) ? 0.5 * kick : kick;
            // the cell and t the forward step used at this position
            const double k = floor( s );
            const double t = s - k;
            const int i = POW2 ? ( int( k ) & mask ) : int( k );
            double dpdt;
            const double pv = poly_eval( Ql + jj * sq, q2, i, t, dpdt );
            const double pbar = hk * ubar;
            sbar = fma( pbar, dpdt, sbar );
            if ( active )
            {
                // adjoints of q0 = D0 + D1, q1 = 2 (D1 - D0), q2 = D0 - 2 D1 + D2
                const double q0 = pbar, q1 = pbar * t, q2 = pbar * ( t*t );
                const double d0 = q0 - 2.0*q1 + q2, d1 = q0 + 2.0*q1 - 2.0*q2, d2 = q2;
                double *L = lam + jj * p.Nx;
                const int i0 = ( i + Nx ) % Nx;
                atomicAdd( L + i0, -d0 );
                atomicAdd( L + ( i0 + 1 ) % Nx, d0 - d1 );
                atomicAdd( L + ( i0 + 2 ) % Nx, d1 - d2 );
                atomicAdd( L + ( i0 + 3 ) % Nx, d2 );
            }
            // undo the step
            u = fma( -hk, pv, u );
            ubar = fma( -cd, sbar, ubar );
            s = fma( cd, u, s );
            if ( !POW2 || poly_wrap_due( mlo + jj ) ) rewrap( s, Nxd, invNx );
        }
        __syncthreads();
        for ( uint32_t q = threadIdx.x; q < uint32_t( cnt ) * p.Nx; q += BLOCK )
        {
            const double v = lam[q];
            if ( v != 0.0 ) atomicAdd( p.lambda + size_t( mlo ) * p.Nx + q, v );
        }
    }
}

// adjoints of q0 = D0 + D1, q1 = 2 (D1 - D0), q2 = D0 - 2 D1 + D2 folded into
// the slice adjoints of cell i (periodic, Nx points)
__device__ __forceinline__ void scatter_poly_adjoint( double *L, int i, int Nx, double pbar, double t )
{
    const double q0 = pbar, q1 = pbar * t, q2 = pbar * ( t*t );
    const double d0 = q0 - 2.0*q1 + q2, d1 = q0 + 2.0*q1 - 2.0*q2, d2 = q2;
    const int i0 = ( i + Nx ) % Nx;
    atomicAdd( L + i0, -d0 );
    atomicAdd( L + ( i0 + 1 ) % Nx, d0 - d1 );
    atomicAdd( L + ( i0 + 2 ) % Nx, d1 - d2 );
    atomicAdd( L + ( i0 + 3 ) % Nx, d2 );
}

template <int BLOCK, bool POW2>
__global__ __launch_bounds__( BLOCK )
void adjoint1d_fp32c_kernel( adjoint1_params p )
{
    extern __shared__ double lds[];
    float  *Ql  = reinterpret_cast<float*>( lds );                       // chunk of FP32 tables
    double *lam = lds + ( size_t( p.chunk ) * p.stride_q + 1 ) / 2;      // chunk of folded slice adjoints
    constexpr int WPB = BLOCK / WAVE;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const uint32_t g = blockIdx.x * WPB + wave;
    const bool task = g < p.n_tasks;
    const uint32_t ix = task ? g / p.wpx : 0, j = task ? g % p.wpx : 0;
    const uint32_t iu = j * WAVE + lane;
    const bool active = task && iu < p.Nu;

    const int Nx = int( p.Nx ), mask = Nx - 1;
    const double cd = p.dt * p.dx_inv, kick = -0.5 * p.dt * p.dx_inv;
    const uint32_t sq = p.stride_q, q2 = p.q2_off;
    const frame1 fr = make_frame1( p.f32_p0, p.f32_w0, p.Nx );

    // the FP32C trace kernel's start state (trace1d_fp32c_body)
    float sh, sl, uh, ul;
    fp32c_start( fr, ix, cd * ( p.u_min + iu*p.du + p.du/2 ), sh, sl, uh, ul );

    // forward: slices n-1 .. 0
    for ( int64_t mhi = int64_t( p.n ) - 1; mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        dma_copy16<BLOCK>( Ql, p.qhist32 + size_t( mlo ) * sq, uint32_t( cnt ) * sq / 4u );
        dma_publish();
        for ( int jj = cnt - 1; jj >= 0; --jj )
        {
            const int64_t m = mlo + jj;
            if ( POW2 && poly_wrap_due( m ) ) fp32c_rewrap( fr, sh, sl, uh, ul );
            fp32c_move( fr, sh, sl, uh, ul );
            if ( !POW2 ) fp32c_rewrap( fr, sh, sl, uh, ul );
            float t;
            const int i = fp32c_locate<POW2>( fr, sh, sl, mask, t );
            fp32c_kick( uh, ul, ( m == 0 ? 0.5f : 1.0f ) * fp32c_poly( Ql + jj * sq, q2, i, t ) );
        }
    }

    // seed at the foot
    const double fbar = active ? -p.du * p.g[ix] : 0.0;
    double fx, fu, fa[NUFI_MAX_MODES];
    f0_grad_1d( p.f0, p.x_min + fp32c_position( fr, sh, sl ) * p.dx, fp32c_velocity( fr, uh, ul ) / cd,
                fx, fu, fa, p.npar );
    double sbar = fbar * fx * p.dx;     // d/ds = dx d/dx
    double ubar = fbar * fu / cd;       // d/dw = d/du / cd
    #pragma unroll
    for ( int m = 0; m < NUFI_MAX_MODES; ++m )
    {
        if ( m >= int( p.npar ) ) continue;
        const csum v = wave_csum( csum { fbar * fa[m], 0.0 } );
        if ( task && lane == 0 ) p.grad_part[ size_t( g ) * NUFI_MAX_MODES + m ] += v.s + v.c;
    }

    // reverse: slices 0 .. n-1, undoing each step on the pairs
    for ( int64_t mlo = 0; mlo < int64_t( p.n ); mlo += p.chunk )
    {
        const int64_t mhi = ( mlo + p.chunk - 1 < int64_t( p.n ) - 1 ) ? mlo + p.chunk - 1 : int64_t( p.n ) - 1;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        dma_copy16<BLOCK>( Ql, p.qhist32 + size_t( mlo ) * sq, uint32_t( cnt ) * sq / 4u );
        for ( uint32_t q = threadIdx.x; q < uint32_t( cnt ) * p.Nx; q += BLOCK ) lam[q] = 0.0;
        dma_publish();
        for ( int jj = 0; jj < cnt; ++jj )
        {
            const int64_t m = mlo + jj;
            const float hf = m == 0 ? 0.5f : 1.0f;
            const float *Q = Ql + jj * sq;
            float t;
            const int i = fp32c_locate<POW2>( fr, sh, sl, mask, t );
            const float q = fp32c_poly( Q, q2, i, t );
            // dq/dt of the scaled table, in FP64
            const float2 q01 = reinterpret_cast<const float2*>( Q )[ i + 1 ];
            const double dqdt = fma( 2.0 * double( Q[ q2 + ( i + 1 ) ] ), double( t ), double( q01.y ) );
            const double pbar = double( hf ) * kick * cd * ubar;
            sbar = fma( double( hf ) * ubar, dqdt, sbar );
            if ( active ) scatter_poly_adjoint( lam + jj * p.Nx, i, Nx, pbar, double( t ) );
            // undo the kick, then the drift
            fp32c_kick( uh, ul, -( hf * q ) );
            ubar = ubar - sbar;
            fp32c_move<true>( fr, sh, sl, uh, ul );
            if ( !POW2 || poly_wrap_due( m ) ) fp32c_rewrap( fr, sh, sl, uh, ul );
        }
        __syncthreads();
        for ( uint32_t q = threadIdx.x; q < uint32_t( cnt ) * p.Nx; q += BLOCK )
        {
            const double v = lam[q];
            if ( v != 0.0 ) atomicAdd( p.lambda + size_t( mlo ) * p.Nx + q, v );
        }
    }
}

// g = 2 dt Re F^H( w . F rho ) + Re F^H( conj(m) . F lambda )   (one workgroup)
__global__ __launch_bounds__( FIELD_BLOCK )
void adjoint_field_kernel( adjoint_field_params p )
{
    extern __shared__ double2 ldsa[];
    double2 *A = ldsa, *B = ldsa + p.N;
    const uint32_t N = p.N;
    axis_desc ax { N, p.log2x, 1u, 1u, 0u, N, 1u };

    for ( int pass = 0; pass < 2; ++pass )
    {
        const double *in = pass == 0 ? p.rho : p.lambda;
        for ( uint32_t i = threadIdx.x; i < N; i += FIELD_BLOCK ) A[i] = make_double2( in[i], 0.0 );
        __syncthreads();
        double2 *F = fft_axis( A, B, ax, p.tw, false );
        double2 *G = ( F == A ) ? B : A;
        for ( uint32_t k = threadIdx.x; k < N; k += FIELD_BLOCK )
        {
            if ( pass == 0 ) { const double w = 2.0 * p.dt * p.ewt[k]; F[k] = make_double2( w * F[k].x, w * F[k].y ); }
            else { const double2 m = p.mult[k]; F[k] = cmul( F[k], make_double2( m.x, -m.y ) ); }
        }
        __syncthreads();
        double2 *R = fft_axis( F, G, ax, p.tw, true );
        for ( uint32_t i = threadIdx.x; i < N; i += FIELD_BLOCK )
            p.g[i] = ( pass == 0 ) ? R[i].x : p.g[i] + R[i].x;
        __syncthreads();
    }
}

__global__ __launch_bounds__( 256 )
void reduce_grad_kernel( const double *part, uint32_t n_tasks, uint32_t npar, double *grad )
{
    __shared__ double red_s[256], red_c[256];
    for ( uint32_t m = 0; m < npar; ++m )
    {
        csum acc { 0, 0 };
        for ( uint32_t g = threadIdx.x; g < n_tasks; g += 256 )
            acc = csum_add( acc, csum { part[ size_t( g ) * NUFI_MAX_MODES + m ], 0.0 } );
        red_s[threadIdx.x] = acc.s; red_c[threadIdx.x] = acc.c;
        __syncthreads();
        if ( threadIdx.x == 0 )
        {
            csum t { 0, 0 };
            for ( int i = 0; i < 256; ++i ) t = csum_add( t, csum { red_s[i], red_c[i] } );
            grad[m] = t.s + t.c;
        }
        __syncthreads();
    }
}

} // namespace

size_t adjoint1d_lds_bytes( const adjoint1_params &p )
{
    if ( p.qhist32 ) return ( size_t( p.chunk ) * p.stride_q + 1 ) / 2 * sizeof(double) + size_t( p.chunk ) * p.Nx * sizeof(double);
    return size_t( p.chunk ) * ( p.stride_q + p.Nx ) * sizeof(double);
}

hipError_t launch_adjoint1d( const adjoint1_params &p, hipStream_t st )
{
    if ( p.n_tasks == 0 ) return hipSuccess;
    constexpr int BLOCK = 256;
    const uint32_t grid = ( p.n_tasks + BLOCK / WAVE - 1 ) / ( BLOCK / WAVE );
    const size_t lds = adjoint1d_lds_bytes( p );
    const bool pow2 = ( p.Nx & ( p.Nx - 1 ) ) == 0;
    const bool f32 = p.qhist32 != nullptr;
    const void *k = f32 ? ( pow2 ? reinterpret_cast<const void*>( adjoint1d_fp32c_kernel<BLOCK,true> )
                                 : reinterpret_cast<const void*>( adjoint1d_fp32c_kernel<BLOCK,false> ) )
                        : ( pow2 ? reinterpret_cast<const void*>( adjoint1d_kernel<BLOCK,true> )
                                 : reinterpret_cast<const void*>( adjoint1d_kernel<BLOCK,false> ) );
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( k, hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    if ( f32 && pow2 )  hipLaunchKernelGGL( ( adjoint1d_fp32c_kernel<BLOCK,true> ), dim3( grid ), dim3( BLOCK ), lds, st, p );
    else if ( f32 )     hipLaunchKernelGGL( ( adjoint1d_fp32c_kernel<BLOCK,false> ), dim3( grid ), dim3( BLOCK ), lds, st, p );
    else if ( pow2 )    hipLaunchKernelGGL( ( adjoint1d_kernel<BLOCK,true> ), dim3( grid ), dim3( BLOCK ), lds, st, p );
    else                hipLaunchKernelGGL( ( adjoint1d_kernel<BLOCK,false> ), dim3( grid ), dim3( BLOCK ), lds, st, p );
    return hipGetLastError();
}

hipError_t launch_adjoint_field( const adjoint_field_params &p, hipStream_t st )
{
    const size_t lds = size_t( p.N ) * 2 * sizeof(double2);
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( adjoint_field_kernel ),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    hipLaunchKernelGGL( adjoint_field_kernel, dim3( 1 ), dim3( FIELD_BLOCK ), lds, st, p );
    return hipGetLastError();
}

hipError_t launch_reduce_grad( const double *part, uint32_t n_tasks, uint32_t npar, double *grad, hipStream_t st )
{
    hipLaunchKernelGGL( reduce_grad_kernel, dim3( 1 ), dim3( 256 ), 0, st, part, n_tasks, npar, grad );
    return hipGetLastError();
}

} // namespace nufi_impl
