// device_common.hpp — device helpers shared by the gfx950 kernels of the NuFI
// rho path: the initial conditions f0, error-free transformations and the
// deterministic wave / workgroup reductions.
#ifndef NUFI_DEVICE_COMMON_HPP
#define NUFI_DEVICE_COMMON_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/nufi_hip.h"

namespace nufi_dev
{

constexpr int WAVE = 64;

// Global loads each thread keeps in flight when it copies a history chunk
// into LDS (the copy is otherwise one L2 round trip per element).
constexpr unsigned STAGE_BATCH = 8;

// 1/sqrt(2 pi), (2 pi)^(-3/2) and the 3D two-stream constant, as the literals
// of nufi/config.hpp:82, 233, 238.
constexpr double kInvSqrt2Pi = 0.39894228040143267793994;
constexpr double kInv2Pi32   = 0.06349363593424096978576330493464;
constexpr double kTwoStream3 = 0.03174681796712048489288165246732;

// ---- initial conditions (config.hpp:72-84, 221-247) -----------------------
// The same operation order as the host oracle, so that in FP64 the two differ
// only by the rounding of the device cos/exp.
__device__ inline double f0_1d( const nufi_f0 &p, double x, double u )
{
#pragma clang fp contract(off)
    switch ( p.kind )
    {
    case NUFI_F0_LANDAU_1D:
        return kInvSqrt2Pi * ( 1.0 + p.alpha*cos(p.k*x) ) * exp( -u*u/2.0 );
    case NUFI_F0_TWOSTREAM_1D:
        return kInvSqrt2Pi * ( 1.0 + p.alpha*cos(p.k*x) ) * exp( -u*u/2.0 ) * u*u;
    case NUFI_F0_MULTIMODE_1D:
    {
        double s = 1.0;
        #pragma unroll
        for ( int m = 1; m <= NUFI_MAX_MODES; ++m )     // static indices: no private-memory copy
            if ( m <= p.n_modes ) s = s + p.a[m-1] * cos( ( p.k*m ) * x );
        return kInvSqrt2Pi * s * exp( -u*u/2.0 );
    }
    default:
        return 0.0;
    }
}

__device__ inline double f0_3d( const nufi_f0 &p, double x, double y, double z,
                                double u, double v, double w )
{
#pragma clang fp contract(off)
    switch ( p.kind )
    {
    case NUFI_F0_LANDAU_3D:
        return kInv2Pi32 * ( 1.0 + p.alpha*cos(p.k*x) + p.alpha*cos(p.k*y) + p.alpha*cos(p.k*z) )
               * exp( -(u*u+v*v+w*w)/2.0 );
    case NUFI_F0_TWOSTREAM_3D:
    {
        const double v0 = 2.4;
        return kTwoStream3 * ( ( exp(-(v-v0)*(v-v0)/2.0) + exp(-(v+v0)*(v+v0)/2.0) ) )
               * exp( -(u*u+w*w)/2.0 )
               * ( 1.0 + p.alpha * ( cos(p.k*x) + cos(p.k*y) + cos(p.k*z) ) );
    }
    case NUFI_F0_BUMPONTAIL_3D:
        return kInv2Pi32 * ( 0.9*exp(-0.5*u*u) + 0.2*exp(-2.0*(u-4.5)*(u-4.5)) )
               * exp( -0.5 * (v*v + w*w) )
               * ( 1.0 + 0.03*( cos(0.3*x) + cos(0.3*y) + cos(0.3*z) ) );
    default:
        return 0.0;
    }
}

// ---- error-free transformations (Knuth TwoSum, Dekker Fast2Sum) ----------
template <typename T>
__device__ __forceinline__ void two_sum( T a, T b, T &s, T &e )
{
    s = a + b;
    const T bb = s - a;
    e = ( a - ( s - bb ) ) + ( b - bb );
}

template <typename T>
__device__ __forceinline__ void fast_two_sum( T a, T b, T &s, T &e )
{
    s = a + b;
    e = b - ( s - a );
}

// A compensated FP64 value (sum + running error), combined pairwise; every
// pairwise combination is commutative bit for bit, so butterfly reductions
// leave the same result in every lane.
struct csum
{
    double s, c;
};

__device__ __forceinline__ csum csum_add( csum a, csum b )
{
    double s, e;
    two_sum( a.s, b.s, s, e );
    return csum { s, ( a.c + b.c ) + e };
}

__device__ __forceinline__ csum wave_csum( csum v )
{
    #pragma unroll
    for ( int off = WAVE/2; off > 0; off >>= 1 )
    {
        csum o;
        o.s = __shfl_xor( v.s, off );
        o.c = __shfl_xor( v.c, off );
        v = csum_add( v, o );
    }
    return v;
}

__device__ __forceinline__ double wave_sum( double v )
{
    #pragma unroll
    for ( int off = WAVE/2; off > 0; off >>= 1 )
        v += __shfl_xor( v, off );
    return v;
}

// Direct global->LDS copy of n16 16-byte words (gfx950 global_load_lds_dwordx4):
// lane l of a wave writes word q0 + l of dst, the wave's words starting at
// q0 (M0 holds the wave's LDS address); the loads hold no registers and are
// waited for (vmcnt) before the barrier that publishes the buffer (dma_publish).
// dst and src are 16-byte aligned.  The instruction is issued from inline
// assembly: seen as an LDS write, the compiler would make every LDS read that
// follows wait for all copies in flight (vmcnt(0)), including the copy into
// the other buffer that the reads do not touch.  The compiler's own waits for
// its loads stay correct with these extra loads outstanding: vector memory
// loads complete in issue order.
template <int BLOCK>
__device__ __forceinline__ void dma_copy16( void *dst, const void *src, uint32_t n16 )
{
    typedef __attribute__((address_space(3))) char lds_char;
    const uint32_t lane = threadIdx.x % WAVE;
    const uint32_t w = __builtin_amdgcn_readfirstlane( threadIdx.x / WAVE );     // the wave index, scalar
    const uint32_t base = uint32_t( reinterpret_cast<uintptr_t>( (lds_char*)dst ) );
    for ( uint32_t q0 = w * WAVE; q0 < n16; q0 += BLOCK )                        // uniform in the wave
    {
        if ( q0 + lane < n16 )
        {
            const char *g = static_cast<const char*>( src ) + size_t( q0 + lane ) * 16;
            const uint32_t m0 = base + q0 * 16u;
            asm volatile( "s_nop 0\n\tglobal_load_lds_dwordx4 %0, off" :: "v"( g ), "{m0}"( m0 ) : "memory" );
        }
    }
}

// wait for this wave's direct LDS loads (vmcnt(0)), then the workgroup barrier
__device__ __forceinline__ void dma_publish()
{
    __builtin_amdgcn_s_waitcnt( 0x0F70 );     // vmcnt(0); expcnt and lgkmcnt not waited for
    __syncthreads();
}

} // namespace nufi_dev

#endif
