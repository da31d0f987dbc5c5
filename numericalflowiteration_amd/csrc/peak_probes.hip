// peak_probes.hip — microbenchmarks of the MI355X peaks the rooflines cite,
// built as libnufi_probes.so (not part of the rho-path library):
//   nufi_probe_fma  VALU FMA throughput (FP64 or non-packed FP32): 8
//                   independent FMA chains per lane, 1024 threads per CU, and
//                   the core clock during the kernel from the shader clock
//                   counter (s_memtime) against the 100 MHz real-time counter
//                   (s_memrealtime);
//   nufi_probe_int  32-bit integer VALU issue (xor + add chains), the same shape;
//   nufi_probe_lds  LDS read bandwidth: 8- or 16-byte reads, lanes on
//                   consecutive entries or hashed over the array, every index
//                   masked into a 32 KiB LDS array.
// All return through out-parameters and 0 on success (a hipError_t
// otherwise); they allocate and free their own small device buffers.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace
{

constexpr int FMA_BLOCK = 1024;
constexpr int FMA_CHAINS = 8;
constexpr int LDS_BLOCK = 256;
constexpr uint32_t LDS_N = 4096;          // doubles in the LDS array (32 KiB), a power of two
constexpr int LDS_UNROLL = 8;

// start and end stamps of each workgroup (thread 0, after a barrier at the end)
__device__ __forceinline__ void stamp( uint64_t *clk, uint64_t c0, uint64_t r0 )
{
    const uint64_t c1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    clk[ 2 * blockIdx.x ] = c1 - c0;
    clk[ 2 * blockIdx.x + 1 ] = r1 - r0;
}

template <typename T>
__global__ __launch_bounds__( FMA_BLOCK ) void fma_probe( T *out, uint64_t *clk, int iters, T a, T b )
{
    uint64_t c0 = 0, r0 = 0;
    if ( threadIdx.x == 0 ) { c0 = __builtin_amdgcn_s_memtime(); r0 = __builtin_amdgcn_s_memrealtime(); }
    T x[FMA_CHAINS];
    #pragma unroll
    for ( int k = 0; k < FMA_CHAINS; ++k ) x[k] = T( threadIdx.x + k );
    for ( int it = 0; it < iters; ++it )
    {
        #pragma unroll
        for ( int k = 0; k < FMA_CHAINS; ++k ) x[k] = x[k] * a + b;
    }
    T s = 0;
    #pragma unroll
    for ( int k = 0; k < FMA_CHAINS; ++k ) s += x[k];
    out[ blockIdx.x * FMA_BLOCK + threadIdx.x ] = s;
    __syncthreads();
    if ( threadIdx.x == 0 ) stamp( clk, c0, r0 );
}

__global__ __launch_bounds__( FMA_BLOCK ) void int_probe( uint32_t *out, uint64_t *clk, int iters, uint32_t a )
{
    uint64_t c0 = 0, r0 = 0;
    if ( threadIdx.x == 0 ) { c0 = __builtin_amdgcn_s_memtime(); r0 = __builtin_amdgcn_s_memrealtime(); }
    uint32_t x[FMA_CHAINS];
    #pragma unroll
    for ( int k = 0; k < FMA_CHAINS; ++k ) x[k] = threadIdx.x + uint32_t( k );
    for ( int it = 0; it < iters; ++it )
    {
        #pragma unroll
        for ( int k = 0; k < FMA_CHAINS; ++k ) x[k] = ( x[k] ^ a ) + 0x9e37u;
    }
    uint32_t s = 0;
    #pragma unroll
    for ( int k = 0; k < FMA_CHAINS; ++k ) s += x[k];
    out[ blockIdx.x * FMA_BLOCK + threadIdx.x ] = s;
    __syncthreads();
    if ( threadIdx.x == 0 ) stamp( clk, c0, r0 );
}

// lane start index: consecutive entries per wave, or a multiplicative hash
__device__ __forceinline__ uint32_t lds_start( int scattered, uint32_t lane, uint32_t wave )
{
    return scattered ? ( ( lane * 2654435761u ) >> 20 ) : ( lane + wave * 64u );
}

template <int WIDTH>
__global__ __launch_bounds__( LDS_BLOCK ) void lds_probe( double *out, uint64_t *clk, int iters, int scattered )
{
    __shared__ double buf[LDS_N];
    uint64_t c0 = 0, r0 = 0;
    if ( threadIdx.x == 0 ) { c0 = __builtin_amdgcn_s_memtime(); r0 = __builtin_amdgcn_s_memrealtime(); }
    for ( uint32_t i = threadIdx.x; i < LDS_N; i += LDS_BLOCK ) buf[i] = double( i );
    __syncthreads();
    const uint32_t lane = threadIdx.x % 64u, wave = threadIdx.x / 64u;
    uint32_t idx = lds_start( scattered, lane, wave ) & ( LDS_N - 1 );
    double acc0 = 0.0, acc1 = 0.0;
    for ( int it = 0; it < iters; ++it )
    {
        #pragma unroll
        for ( int u = 0; u < LDS_UNROLL; ++u )
        {
            const uint32_t a = ( idx + uint32_t( u ) * 512u ) & ( LDS_N - 1 );
            if ( WIDTH == 8 ) acc0 += buf[a];
            else
            {
                const double2 v = *reinterpret_cast<const double2*>( buf + ( a & ~1u ) );
                acc0 += v.x; acc1 += v.y;
            }
        }
        idx = ( idx + 8u ) & ( LDS_N - 1 );
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        out[ blockIdx.x ] = acc0 + acc1;
        stamp( clk, c0, r0 );
    }
}

// mean core clock (GHz) over the workgroups' stamps
double clock_ghz( const uint64_t *clk, int grid )
{
    double cyc = 0.0, rt = 0.0;
    for ( int i = 0; i < grid; ++i ) { cyc += double( clk[ 2 * i ] ); rt += double( clk[ 2 * i + 1 ] ); }
    return rt > 0.0 ? cyc / ( rt / 100e6 ) / 1e9 : 0.0;
}

int cu_count()
{
    int dev = 0, cus = 0;
    if ( hipGetDevice( &dev ) != hipSuccess ) return 0;
    if ( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) != hipSuccess ) return 0;
    return cus;
}

// times one launch (after a short warm-up launch) and reads the stamps back
template <typename LAUNCH>
hipError_t timed( LAUNCH launch, uint64_t *d_clk, uint64_t *h_clk, int grid, float *ms )
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e;
    if ( ( e = hipEventCreate( &e0 ) ) != hipSuccess ) return e;
    if ( ( e = hipEventCreate( &e1 ) ) != hipSuccess ) { (void)hipEventDestroy( e0 ); return e; }
    launch( true );
    if ( ( e = hipGetLastError() ) == hipSuccess && ( e = hipDeviceSynchronize() ) == hipSuccess &&
         ( e = hipEventRecord( e0, nullptr ) ) == hipSuccess )
    {
        launch( false );
        if ( ( e = hipGetLastError() ) == hipSuccess && ( e = hipEventRecord( e1, nullptr ) ) == hipSuccess &&
             ( e = hipEventSynchronize( e1 ) ) == hipSuccess && ( e = hipEventElapsedTime( ms, e0, e1 ) ) == hipSuccess )
            e = hipMemcpy( h_clk, d_clk, size_t( grid ) * 2 * sizeof( uint64_t ), hipMemcpyDeviceToHost );
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    return e;
}

} // namespace

extern "C" int nufi_probe_fma( int fp64, int iters, double *tflops, double *ghz )
{
    const int grid = cu_count();
    if ( grid <= 0 || iters <= 0 || tflops == nullptr || ghz == nullptr ) return int( hipErrorInvalidValue );
    void *out = nullptr;
    uint64_t *clk = nullptr;
    hipError_t e = hipMalloc( &out, size_t( grid ) * FMA_BLOCK * sizeof( double ) );
    if ( e == hipSuccess ) e = hipMalloc( reinterpret_cast<void**>( &clk ), size_t( grid ) * 2 * sizeof( uint64_t ) );
    uint64_t *h = new uint64_t[ size_t( grid ) * 2 ];
    float ms = 0.0f;
    if ( e == hipSuccess )
        e = timed( [&]( bool warm )
                   {
                       const int n = warm ? 100 : iters;
                       if ( fp64 ) hipLaunchKernelGGL( fma_probe<double>, dim3( grid ), dim3( FMA_BLOCK ), 0, nullptr,
                                                       static_cast<double*>( out ), clk, n, 0.999, 0.001 );
                       else hipLaunchKernelGGL( fma_probe<float>, dim3( grid ), dim3( FMA_BLOCK ), 0, nullptr,
                                                static_cast<float*>( out ), clk, n, 0.999f, 0.001f );
                   }, clk, h, grid, &ms );
    if ( e == hipSuccess )
    {
        const double flops = 2.0 * FMA_CHAINS * double( iters ) * double( grid ) * FMA_BLOCK;
        *tflops = flops / ( double( ms ) * 1e-3 ) / 1e12;
        *ghz = clock_ghz( h, grid );
    }
    delete[] h;
    if ( clk ) (void)hipFree( clk );
    if ( out ) (void)hipFree( out );
    return int( e );
}

// wave instructions per SIMD per clock of the core clock measured in the kernel
extern "C" int nufi_probe_int( int iters, double *insts_per_simd_clock, double *ghz )
{
    const int grid = cu_count();
    if ( grid <= 0 || iters <= 0 || insts_per_simd_clock == nullptr || ghz == nullptr ) return int( hipErrorInvalidValue );
    uint32_t *out = nullptr;
    uint64_t *clk = nullptr;
    hipError_t e = hipMalloc( reinterpret_cast<void**>( &out ), size_t( grid ) * FMA_BLOCK * sizeof( uint32_t ) );
    if ( e == hipSuccess ) e = hipMalloc( reinterpret_cast<void**>( &clk ), size_t( grid ) * 2 * sizeof( uint64_t ) );
    uint64_t *h = new uint64_t[ size_t( grid ) * 2 ];
    float ms = 0.0f;
    if ( e == hipSuccess )
        e = timed( [&]( bool warm )
                   {
                       hipLaunchKernelGGL( int_probe, dim3( grid ), dim3( FMA_BLOCK ), 0, nullptr, out, clk,
                                           warm ? 100 : iters, 0x1234u );
                   }, clk, h, grid, &ms );
    if ( e == hipSuccess )
    {
        *ghz = clock_ghz( h, grid );
        const double insts = 2.0 * FMA_CHAINS * double( iters ) * double( grid ) * FMA_BLOCK / 64.0;   // wave instructions
        *insts_per_simd_clock = *ghz > 0.0 ? insts / ( double( ms ) * 1e-3 ) / ( 4.0 * grid ) / ( *ghz * 1e9 ) : 0.0;
    }
    delete[] h;
    if ( clk ) (void)hipFree( clk );
    if ( out ) (void)hipFree( out );
    return int( e );
}

extern "C" int nufi_probe_lds( int width, int scattered, int iters, double *tbs, double *ghz )
{
    const int cus = cu_count();
    if ( cus <= 0 || iters <= 0 || ( width != 8 && width != 16 ) || tbs == nullptr || ghz == nullptr )
        return int( hipErrorInvalidValue );
    const int grid = cus * 4;                 // 4 workgroups of 256 threads (32 KiB LDS each) per CU
    double *out = nullptr;
    uint64_t *clk = nullptr;
    hipError_t e = hipMalloc( reinterpret_cast<void**>( &out ), size_t( grid ) * sizeof( double ) );
    if ( e == hipSuccess ) e = hipMalloc( reinterpret_cast<void**>( &clk ), size_t( grid ) * 2 * sizeof( uint64_t ) );
    uint64_t *h = new uint64_t[ size_t( grid ) * 2 ];
    float ms = 0.0f;
    if ( e == hipSuccess )
        e = timed( [&]( bool warm )
                   {
                       const int n = warm ? 100 : iters;
                       if ( width == 8 ) hipLaunchKernelGGL( lds_probe<8>, dim3( grid ), dim3( LDS_BLOCK ), 0, nullptr,
                                                             out, clk, n, scattered );
                       else hipLaunchKernelGGL( lds_probe<16>, dim3( grid ), dim3( LDS_BLOCK ), 0, nullptr,
                                                out, clk, n, scattered );
                   }, clk, h, grid, &ms );
    if ( e == hipSuccess )
    {
        const double bytes = double( width ) * LDS_UNROLL * double( iters ) * double( grid ) * LDS_BLOCK;
        *tbs = bytes / ( double( ms ) * 1e-3 ) / 1e12;
        *ghz = clock_ghz( h, grid );
    }
    delete[] h;
    if ( clk ) (void)hipFree( clk );
    if ( out ) (void)hipFree( out );
    return int( e );
}
