// noalloc.hip — nufi_rho3d_noalloc: rho of one 3D time step from a device
// history the caller owns into a device rho the caller owns, with no device
// allocation.  It runs the kernels of nufi_compute_rho on a 3D FP64 context
// (the x-major fast trace of trace3d.hip, then the fixed-order partial sum of
// field.hip), so its result equals compute_rho(n, 0, all nodes) followed by a
// download into zeros, bit for bit.
//
// The rho partial sums (ppx partials per spatial point) go to `scratch`, one
// file-scope device array per GPU, instead of context memory.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a stand-in,
// not a device address): the first call on a device, with that device
// current, looks up its address with hipGetSymbolAddress and keeps it per
// device.  Every call on a GPU shares that GPU's array, so calls that overlap
// on one GPU (two streams, two threads) race; the trace kernel writes every
// partial the sum reads, so nothing has to be cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.hpp"

namespace
{

constexpr size_t SCRATCH_DOUBLES = NUFI_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__( 256 ) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double *g_scratch[MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int fail( int code, const std::string &msg )
{
    nufi_impl::set_context_free_error( msg );
    return code;
}

// The current device's copy of `scratch` (looked up once per device).
int device_scratch( double **out )
{
    int dev = 0;
    hipError_t e = hipGetDevice( &dev );
    if ( e != hipSuccess ) return fail( NUFI_ERR_DEVICE, std::string( "rho3d_noalloc: hipGetDevice: " ) + hipGetErrorString( e ) );
    if ( dev < 0 || dev >= MAX_DEVICES ) return fail( NUFI_ERR_UNSUPPORTED, "rho3d_noalloc: device index above 63" );
    std::lock_guard<std::mutex> lock( g_scratch_mu );
    if ( g_scratch[dev] == nullptr )
    {
        void *p = nullptr;
        e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
        if ( e != hipSuccess || p == nullptr )
            return fail( NUFI_ERR_DEVICE, std::string( "rho3d_noalloc: hipGetSymbolAddress: " ) + hipGetErrorString( e ) );
        g_scratch[dev] = static_cast<double*>( p );
    }
    *out = g_scratch[dev];
    return NUFI_OK;
}

} // namespace

extern "C" int nufi_rho3d_noalloc( const nufi_config *conf, uint64_t n, const double *hist_dev, uint64_t hist_len,
                                   double *rho_dev, void *stream )
{
    using namespace nufi_impl;
    constexpr int order = 4;
    if ( validate_config( conf, order ) != NUFI_OK ) return fail( NUFI_ERR_INVALID, "rho3d_noalloc: invalid config" );
    if ( conf->dim != 3 ) return fail( NUFI_ERR_UNSUPPORTED, "rho3d_noalloc: 3D configs only" );
    if ( n > conf->Nt ) return fail( NUFI_ERR_RANGE, "rho3d_noalloc: time step out of range" );
    if ( rho_dev == nullptr ) return fail( NUFI_ERR_INVALID, "rho3d_noalloc: rho is NULL" );

    const uint64_t sy = conf->Nx + order - 1, sz = sy * ( conf->Ny + order - 1 ), st = sz * ( conf->Nz + order - 1 );
    const uint64_t n_space = conf->Nx * conf->Ny * conf->Nz, n_vel = conf->Nu * conf->Nv * conf->Nw;
    const uint32_t vgroup = xm3_vgroup( n_space, n_vel );
    const uint64_t ppx = ( n_vel + vgroup - 1 ) / vgroup;
    if ( ppx * n_space > SCRATCH_DOUBLES )
        return fail( NUFI_ERR_UNSUPPORTED, "rho3d_noalloc: this grid needs " + std::to_string( ppx * n_space * 8 ) +
                                           " bytes of partial sums, above the " +
                                           std::to_string( NUFI_NOALLOC_SCRATCH_BYTES ) + "-byte scratch array" );
    if ( st * sizeof(double) > 160 * 1024 )
        return fail( NUFI_ERR_UNSUPPORTED, "rho3d_noalloc: one slice of this grid exceeds the LDS" );
    const uint64_t wpx = ( n_vel + 63 ) / 64;
    if ( n_space * wpx >= ( uint64_t( 1 ) << 32 ) ) return fail( NUFI_ERR_RANGE, "rho3d_noalloc: grid too large for one launch" );
    if ( n > 0 )
    {
        // the trace copies slices 0..n-1 as the 16-byte words covering them:
        // the history must start 16-byte aligned and hold one real past slice n-1
        if ( hist_dev == nullptr ) return fail( NUFI_ERR_INVALID, "rho3d_noalloc: history is NULL" );
        if ( reinterpret_cast<uintptr_t>( hist_dev ) % 16 != 0 )
            return fail( NUFI_ERR_INVALID, "rho3d_noalloc: the history must be 16-byte aligned" );
        if ( hist_len < n * st + 1 )
            return fail( NUFI_ERR_INVALID, "rho3d_noalloc: the history must hold n slices and one more real" );
    }
    double *part = nullptr;
    int rc = device_scratch( &part );
    if ( rc ) return rc;

    trace3_params p {};
    p.hist = hist_dev; p.hist32 = nullptr; p.partials = part;
    p.mode = MODE_RHO;
    p.n = n; p.q_begin = 0; p.q_end = n_space * n_vel;
    p.Nx = uint32_t( conf->Nx ); p.Ny = uint32_t( conf->Ny ); p.Nz = uint32_t( conf->Nz );
    p.Nu = uint32_t( conf->Nu ); p.Nv = uint32_t( conf->Nv ); p.Nw = uint32_t( conf->Nw );
    p.sy = uint32_t( sy ); p.sz = uint32_t( sz ); p.st = uint32_t( st ); p.st32 = uint32_t( ( st + 3 ) & ~uint64_t( 3 ) );
    p.wpx = uint32_t( wpx );
    p.l_first = 0; p.n_tasks = uint32_t( n_space * wpx ); p.n_count = uint32_t( n_space );
    p.vgroup = vgroup; p.ppx = uint32_t( ppx );
    p.order = order;
    fill_trace3_geometry( p, *conf );
    const hipStream_t s = static_cast<hipStream_t>( stream );
    hipError_t e = launch_trace3d( p, TRACE_FAST, launch_shape { 512, 4, 1 }, s );
    if ( e != hipSuccess ) return fail( NUFI_ERR_DEVICE, std::string( "rho3d_noalloc: trace launch: " ) + hipGetErrorString( e ) );
    e = launch_reduce_partials( part, uint32_t( ppx ), -( conf->du * conf->dv * conf->dw ), 0, uint32_t( n_space ),
                                uint32_t( n_space ), rho_dev, s );
    if ( e != hipSuccess ) return fail( NUFI_ERR_DEVICE, std::string( "rho3d_noalloc: reduce launch: " ) + hipGetErrorString( e ) );
    return NUFI_OK;
}
