// field.hip — the per-step field solve on the device: rho -> periodic Poisson
// solve -> spline fit -> coefficient slice n, in one LDS-resident workgroup.
//
// Reference: poisson<double>::solve (nufi/poisson.cpp:66-89 dim1, :328-362
// dim3) runs a DHT, scales mode k by 1/(N |kappa_k|^2) with the folded index
// (2i < N ? i : N-i), accumulates the spectral energy, zeroes mode 0 and
// transforms back; dim{1,3}::interpolate (nufi/fields.hpp:63-142, 352-490)
// then solves the circulant collocation system sum_j B_j(0) c[i+j] = phi_i
// by LSMR and pads the slice periodically.
//
// Both operators are diagonal in Fourier space, so here they are one
// multiplier: c_hat_k = rho_hat_k / (N kappa_k^2 s_k) with the collocation
// symbol s_k = prod_axes sum_j B_j(0) e^{+2 pi i j k / N_axis} (exact solve;
// the pseudo-inverse where s_k = 0, which is what LSMR from the zero-filled
// slot converges to).  A real, folded multiplier commutes with the Hartley
// kernel, so DHT-scale-DHT equals FFT-scale-IFFT and the spectral energy
// sum_k H_k^2 w_k equals sum_k |F_k|^2 w_k.  The transform is a batched
// radix-4 Stockham FFT along each axis (a direct DFT for an axis whose length
// is not a power of two), ping-ponged between two LDS buffers.
#include "device_common.hpp"
#include "internal.hpp"
#include "fft_lds.hpp"
#include "step1d.hpp"

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

// Compensated sum over j = 0 .. wpx-1 of partials[j*n + i], in j order; the
// loads go out RB at a time (clamped, unconditional) so that one point's sum
// waits for ceil(wpx/RB) round trips instead of wpx.
__device__ __forceinline__ double sum_partials( const double *__restrict__ partials, uint32_t wpx, size_t n,
                                                uint32_t i )
{
    constexpr uint32_t RB = 64;
    csum acc { 0.0, 0.0 };
    for ( uint32_t j0 = 0; j0 < wpx; j0 += RB )
    {
        double v[RB];
        #pragma unroll
        for ( uint32_t k = 0; k < RB; ++k )
        {
            const uint32_t j = j0 + k < wpx ? j0 + k : wpx - 1;
            v[k] = partials[ size_t( j ) * n + i ];
        }
        #pragma unroll
        for ( uint32_t k = 0; k < RB; ++k )
            if ( j0 + k < wpx ) acc = csum_add( acc, csum { v[k], 0.0 } );
    }
    return acc.s + acc.c;
}

__device__ double2 *fft3( double2 *A, double2 *B, const field_params &p, const double2 *tx, const double2 *ty,
                          const double2 *tz, bool inverse )
{
    axis_desc ax { p.Nx, p.log2x, 1u,          p.N / p.Nx, 0u, p.Nx, p.Ny };
    axis_desc ay { p.Ny, p.log2y, p.Nx,        p.N / p.Ny, 1u, p.Nx, p.Ny };
    axis_desc az { p.Nz, p.log2z, p.Nx * p.Ny, p.N / p.Nz, 2u, p.Nx, p.Ny };
    double2 *r = fft_axis( A, B, ax, tx, inverse );
    double2 *o = ( r == A ) ? B : A;
    r = fft_axis( r, o, ay, ty, inverse );
    o = ( r == A ) ? B : A;
    r = fft_axis( r, o, az, tz, inverse );
    return r;
}

__device__ __forceinline__ void field_body( const field_params &p, double2 *lds2 )
{
    double2 *A = lds2, *B = lds2 + p.N;
    double *red = reinterpret_cast<double*>( lds2 + 2*p.N );
    const uint32_t N = p.N;

    // the twiddle tables in LDS (one load per kernel instead of one L2 round
    // trip per FFT stage), when they fit beside the buffers
    const double2 *tx = p.tw_x, *ty = p.tw_y, *tz = p.tw_z;
    if ( p.tw_lds )
    {
        double2 *t = lds2 + 2*p.N + ( FIELD_BLOCK / WAVE ) / 2;
        const uint32_t lx = twiddle_entries( p.Nx, p.log2x ), ly = twiddle_entries( p.Ny, p.log2y );
        const uint32_t lz = twiddle_entries( p.Nz, p.log2z );
        for ( uint32_t i = threadIdx.x; i < lx; i += FIELD_BLOCK ) t[i] = p.tw_x[i];
        for ( uint32_t i = threadIdx.x; i < ly; i += FIELD_BLOCK ) t[lx + i] = p.tw_y[i];
        for ( uint32_t i = threadIdx.x; i < lz; i += FIELD_BLOCK ) t[lx + ly + i] = p.tw_z[i];
        tx = t; ty = t + lx; tz = t + lx + ly;
    }

    // 1. rho (or the nodal values) into A
    for ( uint32_t i = threadIdx.x; i < N; i += FIELD_BLOCK )
    {
        double v;
        if ( p.mode == FIELD_STEP && p.partials != nullptr )
            v = p.weight * sum_partials( p.partials, p.wpx, N, i );
        else
        {
            v = p.values[i];
        }
        if ( p.rho_out ) p.rho_out[i] = v;
        A[i] = make_double2( v, 0.0 );
    }
    __syncthreads();

    // 2. forward transform
    double2 *F = fft3( A, B, p, tx, ty, tz, false );
    double2 *G = ( F == A ) ? B : A;

    // 3. spectral energy and the multiplier
    double e = 0.0;
    if ( p.mode != FIELD_FIT )
        for ( uint32_t k = threadIdx.x; k < N; k += FIELD_BLOCK )
        {
            const double2 f = F[k];
            e += p.ewt[k] * ( f.x*f.x + f.y*f.y );
        }
    e = wave_sum( e );
    if ( threadIdx.x % WAVE == 0 ) red[ threadIdx.x / WAVE ] = e;
    for ( uint32_t k = threadIdx.x; k < N; k += FIELD_BLOCK )
        F[k] = cmul( F[k], p.mult[k] );
    __syncthreads();
    if ( threadIdx.x == 0 && p.mode != FIELD_FIT && p.energy != nullptr )
    {
        double s = 0.0;
        for ( int w = 0; w < FIELD_BLOCK / WAVE; ++w ) s += red[w];
        *p.energy = s;
    }

    // 4. inverse transform
    double2 *R = fft3( F, G, p, tx, ty, tz, true );

    // 5. outputs
    if ( p.mode == FIELD_POISSON )
    {
        for ( uint32_t i = threadIdx.x; i < N; i += FIELD_BLOCK ) p.values[i] = R[i].x;
        return;
    }
    const uint32_t o1 = uint32_t( p.order ) - 1;
    if ( p.dim == 1 )
    {
        const uint32_t len = p.Nx + o1;
        for ( uint32_t i = threadIdx.x; i < len; i += FIELD_BLOCK ) p.slice[i] = R[ i % p.Nx ].x;
        // extended differences: entry j holds D[j-1] = c[j] - c[j-1] (periodic)
        if ( p.dslice )
            for ( uint32_t j = threadIdx.x; j < p.stride_d; j += FIELD_BLOCK )
                p.dslice[j] = ( j < len + 1 ) ? R[ j % p.Nx ].x - R[ ( j + p.Nx - 1 ) % p.Nx ].x : 0.0;
        // the kick-polynomial tables from the same differences: entry j uses D[j-1], D[j], D[j+1]
        if ( p.qslice || p.qslice32 )
            for ( uint32_t j = threadIdx.x; j < p.Nx + 2; j += FIELD_BLOCK )
            {
                double D[3];
                #pragma unroll
                for ( uint32_t l = 0; l < 3; ++l ) D[l] = R[ ( j + l ) % p.Nx ].x - R[ ( j + l + p.Nx - 1 ) % p.Nx ].x;
                if ( p.qslice )   put_poly( p.qslice + 2*j, p.qslice + p.q2_off + j, D[0], D[1], D[2], p.kick );
                if ( p.qslice32 ) put_poly( p.qslice32 + 2*j, p.qslice32 + p.q2_off + j, D[0], D[1], D[2], p.kick );
            }
    }
    else
    {
        const uint32_t ex = p.Nx + o1, ey = p.Ny + o1, ez = p.Nz + o1;
        const uint32_t total = ex * ey * ez;
        for ( uint32_t e2 = threadIdx.x; e2 < total; e2 += FIELD_BLOCK )
        {
            const uint32_t i = e2 % ex, j = ( e2 / ex ) % ey, k = e2 / ( ex * ey );
            const double c = R[ ( ( k % p.Nz ) * p.Ny + ( j % p.Ny ) ) * p.Nx + ( i % p.Nx ) ].x;
            p.slice[ k * p.sz + j * p.sy + i ] = c;
            if ( p.slice32 ) p.slice32[ k * p.sz + j * p.sy + i ] = float( c * p.kick );
        }
    }
}

__global__ __launch_bounds__( FIELD_BLOCK )
void field_kernel( field_params p )
{
    extern __shared__ double2 lds2[];
    field_body( p, lds2 );
}

// Batched sweep: workgroup r solves run r's field at step n (runs past their
// Nt skip it); the per-step outputs are the run's slice n, its differences,
// its kick-polynomial table and energies[n].
__global__ __launch_bounds__( FIELD_BLOCK )
void field_batch_kernel( const batch_field *__restrict__ runs, uint64_t n )
{
    extern __shared__ double2 lds2b[];
    const batch_field &b = runs[ blockIdx.x ];
    if ( n > b.Nt ) return;
    field_params f = b.f;
    f.slice = b.f.slice + n * b.st;
    if ( f.dslice )   f.dslice   = b.f.dslice + n * b.sd;
    if ( f.qslice )   f.qslice   = b.f.qslice + n * b.sq;
    if ( f.qslice32 ) f.qslice32 = b.f.qslice32 + n * b.sq;
    f.energy = b.f.energy + n;
    field_body( f, lds2b );
}

__global__ __launch_bounds__( 64 ) void reduce_partials_kernel( const double *partials, uint32_t wpx, double weight,
                                        uint32_t first, uint32_t count, uint32_t n_space, double *rho )
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n_space ) return;
    double v = 0.0;
    if ( i >= first && i < first + count ) v = weight * sum_partials( partials, wpx, n_space, i );
    rho[i] = v;
}

// FP32 copy of an uploaded slice of `len` reals (3D: scaled by `kick`, which
// is then fp32c3_scale) and, in 1D (period Nx), the
// extended differences dslice[j] = c[j mod Nx] - c[(j-1) mod Nx], j <= len,
// and the kick-polynomial table entries j < Nx+2 from dslice[j .. j+2].
__global__ void slice_tables_kernel( int dim, const double *slice, uint32_t len, uint32_t Nx, double *dslice,
                                     double *qslice, float *qslice32, double kick, float *slice32 )
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i > len ) return;
    if ( slice32 && i < len ) slice32[i] = float( slice[i] * kick );
    if ( dim == 1 && dslice ) dslice[i] = slice[ i % Nx ] - slice[ ( i + Nx - 1 ) % Nx ];
    if ( dim == 1 && ( qslice || qslice32 ) && i < Nx + 2 )
    {
        double D[3];
        #pragma unroll
        for ( uint32_t l = 0; l < 3; ++l ) D[l] = slice[ ( i + l ) % Nx ] - slice[ ( i + l + Nx - 1 ) % Nx ];
        if ( qslice )   put_poly( qslice + 2*i, qslice + poly_q2( Nx ) + i, D[0], D[1], D[2], kick );
        if ( qslice32 ) put_poly( qslice32 + 2*i, qslice32 + poly_q2( Nx ) + i, D[0], D[1], D[2], kick );
    }
}

// metrics[k] = sum over tasks of partials[4 g + k], compensated, fixed order
__global__ __launch_bounds__( 256 )
void reduce_metrics_kernel( const double *partials, uint32_t n_tasks, double *metrics )
{
    __shared__ double red_s[4][256], red_c[4][256];
    csum acc[4] = { { 0, 0 }, { 0, 0 }, { 0, 0 }, { 0, 0 } };
    for ( uint32_t g = threadIdx.x; g < n_tasks; g += 256 )
        #pragma unroll
        for ( int k = 0; k < 4; ++k ) acc[k] = csum_add( acc[k], csum { partials[ size_t( g ) * 4 + k ], 0.0 } );
    #pragma unroll
    for ( int k = 0; k < 4; ++k ) { red_s[k][threadIdx.x] = acc[k].s; red_c[k][threadIdx.x] = acc[k].c; }
    __syncthreads();
    if ( threadIdx.x < 4 )
    {
        const int k = threadIdx.x;
        csum t { 0, 0 };
        for ( int i = 0; i < 256; ++i ) t = csum_add( t, csum { red_s[k][i], red_c[k][i] } );
        metrics[k] = t.s + t.c;
    }
}

} // namespace

hipError_t launch_reduce_metrics( const double *partials, uint32_t n_tasks, double *metrics, hipStream_t st )
{
    hipLaunchKernelGGL( reduce_metrics_kernel, dim3( 1 ), dim3( 256 ), 0, st, partials, n_tasks, metrics );
    return hipGetLastError();
}

int field_supported( int, uint32_t Nx, uint32_t Ny, uint32_t Nz )
{
    const uint64_t N = uint64_t( Nx ) * Ny * Nz;
    if ( N == 0 ) return 0;
    // two complex double buffers of N entries must fit the 160 KiB LDS
    return ( N * 2 * sizeof(double2) + 64 * sizeof(double) ) <= 160u * 1024u ? 1 : 0;
}

// buffers + reduction scratch, and the twiddle tables when they fit
static size_t field_base_bytes( uint32_t N )
{
    return size_t( N ) * 2 * sizeof(double2) + ( FIELD_BLOCK / WAVE ) * sizeof(double);
}

static size_t field_twiddle_bytes( const field_params &p )
{
    return sizeof(double2) * ( size_t( twiddle_entries( p.Nx, p.log2x ) ) + twiddle_entries( p.Ny, p.log2y ) +
                               twiddle_entries( p.Nz, p.log2z ) );
}

size_t field_lds_bytes( const field_params &p )
{
    const size_t base = field_base_bytes( p.N ), all = base + field_twiddle_bytes( p );
    return all <= 160u * 1024u ? all : base;
}

hipError_t launch_field( const field_params &p_in, hipStream_t st )
{
    field_params p = p_in;
    const size_t lds = field_lds_bytes( p );
    p.tw_lds = lds > field_base_bytes( p.N ) ? 1 : 0;
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( field_kernel ),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    hipLaunchKernelGGL( field_kernel, dim3( 1 ), dim3( FIELD_BLOCK ), lds, st, p );
    return hipGetLastError();
}

hipError_t launch_field_batch( const batch_field *runs, uint32_t n_runs, uint64_t n, size_t lds, hipStream_t st )
{
    if ( n_runs == 0 ) return hipSuccess;
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( field_batch_kernel ),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    hipLaunchKernelGGL( field_batch_kernel, dim3( n_runs ), dim3( FIELD_BLOCK ), lds, st, runs, n );
    return hipGetLastError();
}

size_t field_batch_lds_bytes( uint32_t max_N, int *tw_lds )
{
    // 1D runs: at most max_N twiddles (max_N/2 for a power of two)
    const size_t all = field_base_bytes( max_N ) + size_t( max_N ) * sizeof(double2);
    *tw_lds = all <= 160u * 1024u ? 1 : 0;
    return *tw_lds ? all : field_base_bytes( max_N );
}

hipError_t launch_reduce_partials( const double *partials, uint32_t wpx, double weight,
                                   uint32_t first, uint32_t count, uint32_t n_space,
                                   double *rho, hipStream_t st )
{
    const uint32_t block = 64, grid = ( n_space + block - 1 ) / block;     // one point per thread, many CUs
    hipLaunchKernelGGL( reduce_partials_kernel, dim3( grid ), dim3( block ), 0, st,
                        partials, wpx, weight, first, count, n_space, rho );
    return hipGetLastError();
}

hipError_t launch_slice_tables( int dim, const double *slice, uint32_t len, uint32_t Nx, double *dslice,
                                double *qslice, float *qslice32, double kick, float *slice32, hipStream_t st )
{
    const uint32_t block = 256, grid = ( len + 1 + block - 1 ) / block;
    hipLaunchKernelGGL( slice_tables_kernel, dim3( grid ), dim3( block ), 0, st,
                        dim, slice, len, Nx, dslice, qslice, qslice32, kick, slice32 );
    return hipGetLastError();
}

} // namespace nufi_impl
