// capi.hip — the C ABI of libnufi_hip.so (declared in include/nufi_hip.h).
//
// A context owns, on one device and one HIP stream (default priority):
//   hist     (Nt+1) coefficient slices of (Nx+order-1)^dim reals   FP64
//   dhist    1D: first differences of every slice (stride_d)        FP64
//            (the fast, FP32C and adjoint kernels build the per-cell kick
//            polynomials from it in LDS)
//   hist32   3D FP32C: the coefficients rounded to FP32 (slice stride st32: st rounded up to 4)
//   partials compensated partial sums of f: per wave of 64 velocity nodes
//            (strict), per x-major velocity group (fast, FP32C)
//   rho, energies, the FFT twiddles and Fourier multipliers.
// Kernels: trace1d.hip / trace3d.hip (the characteristic trace), field.hip
// (Poisson + fit).  Reference semantics are cited at each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "host_tables.hpp"
#include "internal.hpp"

using namespace nufi_impl;

struct nufi_ctx
{
    nufi_config conf {}, conf_m {};
    int order = 4, precision = NUFI_FP64, device = 0, dim = 1;
    int variant = TRACE_FAST;
    hipStream_t stream = nullptr;
    uint32_t st = 0, sd = 0, sy = 0, sz = 0;
    uint32_t st32 = 0;         // hist32 slice stride: st rounded up to 4 floats
    uint64_t n_space = 0, n_vel = 0, n_nodes = 0;
    uint32_t wpx = 0;          // wave-task partials per spatial point (strict 1D, 3D)
    uint32_t ppx = 0;          // x-major partials per spatial point (fast; 1D FP32C)
    uint32_t vgroup = 0;       // 3D x-major: velocities per partial
    int field_ok = 0;          // 1: one-workgroup field solve; 2: the multi-workgroup one; 0: none
    double2 *lw0 = nullptr, *lw1 = nullptr;   // multi-workgroup solve: two complex buffers
    double *lpart = nullptr;                  // and its energy partials
    int n_cu = 256;
    size_t lds_cap = 65536;
    launch_shape shape { 256, 1, 1 };

    double *hist = nullptr, *dhist = nullptr;
    double *qhist = nullptr;   // 1D FP64: kick-polynomial tables of every slice (stride sq), or null
    float *qhist32 = nullptr;  // 1D FP32C: the FP32 tables scaled by the kick (stride sq), or null
    float *hist32 = nullptr;
    uint32_t sq = 0;           // 1D: LDS kick-polynomial table stride (poly_stride)
    double *partials = nullptr, *rho = nullptr, *energies = nullptr, *values = nullptr;
    double *scratch_slice = nullptr, *metrics = nullptr, *mpartials = nullptr, *energy_tmp = nullptr;
    double *rho_hist = nullptr, *lambda = nullptr, *gvec = nullptr, *grad_part = nullptr, *grad = nullptr;
    uint64_t m_space = 0, m_vel = 0, m_nodes = 0;   // metrics node grid
    uint32_t m_wpx = 0;
    double2 *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr;
    double2 *mult_step = nullptr, *mult_pois = nullptr, *mult_fit = nullptr;
    double *ewt = nullptr;
    double *samp = nullptr;    // sampler points and outputs: 4 x samp_cap reals (grown on demand)
    uint64_t samp_cap = 0;

    std::vector<double> host_rho;
    std::string err;

    // nufi_run replays a hipGraph of its step range (captured once per range,
    // launch shape and variant) unless disabled or the launches are timed
    bool use_graph = true;
    bool graph_allowed = true;  // false when created with NUFI_RUN_GRAPH=0 in the environment
    hipGraphExec_t run_graph = nullptr;
    uint64_t graph_b = 0, graph_e = 0;
    launch_shape graph_shape { 0, 0, 0 };
    int graph_variant = -1;

    bool timing = false;
    std::vector<std::pair<hipEvent_t,hipEvent_t>> ev;
    size_t ev_used = 0;
    uint64_t launches = 0;
    double csteps = 0.0;
};

namespace
{

thread_local std::string g_create_error = "no error";

int fail( nufi_ctx *c, int code, const std::string &msg )
{
    if ( c ) c->err = msg; else g_create_error = msg;
    return code;
}

int hip_fail( nufi_ctx *c, hipError_t e, const char *where )
{
    return fail( c, e == hipErrorOutOfMemory ? NUFI_ERR_ALLOC : NUFI_ERR_DEVICE,
                 std::string( where ) + ": " + hipGetErrorString( e ) );
}

#define HIP_TRY( ctx, expr ) do { hipError_t e_ = ( expr ); if ( e_ != hipSuccess ) return hip_fail( ( ctx ), e_, #expr ); } while ( 0 )

uint32_t log2_or_none( uint64_t n )
{
    if ( n == 0 || ( n & ( n - 1 ) ) ) return 0xffffffffu;
    uint32_t l = 0;
    while ( ( uint64_t( 1 ) << l ) < n ) ++l;
    return l;
}

template <typename T>
int dev_alloc( nufi_ctx *c, T **p, size_t count, const char *what )
{
    if ( count == 0 ) count = 1;
    hipError_t e = hipMalloc( reinterpret_cast<void**>( p ), count * sizeof(T) );
    if ( e != hipSuccess ) return hip_fail( c, e, what );
    // zeroed and waited for: a plain hipMemset goes to the null stream, which
    // the context's non-blocking stream does not wait for, so a late memset
    // could clear what the first kernels on the context stream write
    e = hipMemsetAsync( *p, 0, count * sizeof(T), c->stream );
    if ( e == hipSuccess ) e = hipStreamSynchronize( c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, what );
    return NUFI_OK;
}

void free_all( nufi_ctx *c )
{
    void *ptrs[] = { c->hist, c->dhist, c->qhist, c->qhist32, c->hist32, c->partials, c->rho, c->energies,
                     c->values, c->scratch_slice, c->metrics, c->mpartials, c->energy_tmp,
                     c->rho_hist, c->lambda, c->gvec, c->grad_part, c->grad, c->tw_x, c->tw_y, c->tw_z,
                     c->mult_step, c->mult_pois, c->mult_fit, c->ewt, c->samp, c->lw0, c->lw1, c->lpart };
    if ( c->run_graph ) (void)hipGraphExecDestroy( c->run_graph );
    c->run_graph = nullptr;
    for ( void *p : ptrs ) if ( p ) (void)hipFree( p );
    for ( auto &pr : c->ev ) { (void)hipEventDestroy( pr.first ); (void)hipEventDestroy( pr.second ); }
    c->ev.clear();
    if ( c->stream ) (void)hipStreamDestroy( c->stream );
}

int validate( const nufi_config *c, int order )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( c->dim != 1 && c->dim != 3 ) return NUFI_ERR_UNSUPPORTED;
    if ( c->Nx == 0 || c->Nu == 0 || !( c->dt > 0 ) || !( c->x_max > c->x_min ) || !( c->u_max > c->u_min ) )
        return NUFI_ERR_INVALID;
    if ( c->dim == 3 && ( c->Ny == 0 || c->Nz == 0 || c->Nv == 0 || c->Nw == 0 ||
                          !( c->y_max > c->y_min ) || !( c->z_max > c->z_min ) ||
                          !( c->v_max > c->v_min ) || !( c->w_max > c->w_min ) ) )
        return NUFI_ERR_INVALID;
    if ( order < 3 || order > 8 ) return NUFI_ERR_UNSUPPORTED;
    if ( c->Nx > ( 1u << 24 ) || c->Ny > ( 1u << 24 ) || c->Nz > ( 1u << 24 ) ) return NUFI_ERR_UNSUPPORTED;
    return NUFI_OK;
}

// The LDS that one history slice takes in the trace kernel of `variant` (the
// smallest staging any of its launch shapes uses): 1D fast / FP32C one row of
// kick-polynomial tables, 1D strict one difference row, 3D one slice (FP32
// for FP32C).  Host-only, so grids no kernel can stage are refused before any
// device work.
constexpr size_t LDS_LIMIT = 160 * 1024;    // gfx950 LDS per workgroup

size_t slice_lds_bytes( const nufi_config *c, int order, int variant )
{
    const uint64_t o1 = uint64_t( order ) - 1;
    if ( c->dim == 1 )
    {
        if ( variant == TRACE_STRICT ) return size_t( ( c->Nx + order + 3 ) & ~uint64_t( 3 ) ) * sizeof(double);
        const size_t elem = variant == TRACE_FP32C ? sizeof(float) : sizeof(double);
        return size_t( poly_stride( uint32_t( c->Nx ) ) ) * elem;
    }
    const uint64_t st = ( c->Nx + o1 ) * ( c->Ny + o1 ) * ( c->Nz + o1 );
    if ( variant == TRACE_FP32C ) return size_t( ( st + 3 ) & ~uint64_t( 3 ) ) * sizeof(float);
    return size_t( st ) * sizeof(double);
}

// Host precomputation of the Fourier tables (long double), uploaded once.
int build_tables( nufi_ctx *c )
{
    const fourier_tables t = make_tables( c->conf, c->dim, c->order );
    const size_t N = t.ms.size();
    int rc;
    if ( ( rc = dev_alloc( c, &c->tw_x, t.tx.size(), "twiddles" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->tw_y, t.ty.size(), "twiddles" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->tw_z, t.tz.size(), "twiddles" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->mult_step, N, "multipliers" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->mult_pois, N, "multipliers" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->mult_fit, N, "multipliers" ) ) ) return rc;
    if ( ( rc = dev_alloc( c, &c->ewt, N, "energy weights" ) ) ) return rc;
    HIP_TRY( c, hipMemcpy( c->tw_x, t.tx.data(), t.tx.size() * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->tw_y, t.ty.data(), t.ty.size() * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->tw_z, t.tz.data(), t.tz.size() * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->mult_step, t.ms.data(), N * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->mult_pois, t.mp.data(), N * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->mult_fit, t.mf.data(), N * sizeof(double2), hipMemcpyHostToDevice ) );
    HIP_TRY( c, hipMemcpy( c->ewt, t.ew.data(), N * sizeof(double), hipMemcpyHostToDevice ) );
    return NUFI_OK;
}

// partials per spatial point of the rho kernel in use
uint32_t rho_ppx( const nufi_ctx *c )
{
    const bool xm = c->variant == TRACE_FAST || ( c->dim == 1 && c->variant == TRACE_FP32C );
    return xm ? c->ppx : c->wpx;
}

// The chunk a rho launch of shape s uses and its LDS: 1D x-major launches
// plan the requested chunk within 160 KiB (poly_chunk_plan).
int effective_chunk( const nufi_ctx *c, const launch_shape &s, size_t *lds = nullptr )
{
    const size_t elem = ( c->variant == TRACE_FP32C ) ? sizeof(float) : sizeof(double);
    const size_t row = c->dim == 1 ? ( c->variant == TRACE_STRICT ? c->sd : c->sq ) : c->st;
    if ( c->dim == 3 && c->variant == TRACE_FP32C )
    {
        if ( lds ) *lds = ( fp32c3_dma( uint32_t( s.chunk ), c->st32 ) ? 2 : 1 ) * size_t( s.chunk ) * c->st32 * elem;
        return s.chunk;
    }
    if ( c->dim == 3 && c->variant == TRACE_FAST )
    {
        if ( lds ) *lds = fast3_dma( uint32_t( s.chunk ), c->st ) ? 2 * ( size_t( s.chunk ) * c->st + 2 ) * elem
                                                                   : size_t( s.chunk ) * c->st * elem;
        return s.chunk;
    }
    if ( c->dim != 1 || c->variant == TRACE_STRICT )
    {
        if ( lds ) *lds = size_t( s.chunk ) * row * elem;
        return s.chunk;
    }
    bool pipe = false;
    const uint32_t ch = poly_chunk_plan( uint32_t( c->conf.Nx ), elem, uint64_t( s.chunk ), 160 * 1024, &pipe );
    if ( lds ) *lds = ( pipe ? 2 : 1 ) * size_t( ch ) * row * elem;
    return int( ch );
}

// Default launch shapes (the autotuner refines them): 1D x-major 1024 threads
// x 1 node (16 velocity rows per workgroup), the largest chunk two table
// buffers in 160 KiB of LDS hold; 3D FP64 1024 threads x 2 nodes, FP32C 512 x 4.
void default_shape( nufi_ctx *c )
{
    const uint64_t waves = c->n_space * c->wpx;
    const size_t elem = ( c->variant == TRACE_FP32C ) ? sizeof(float) : sizeof(double);
    if ( c->dim == 1 )
    {
        const size_t budget = std::min<size_t>( c->lds_cap, 64 * 1024 );
        const size_t row = ( c->variant == TRACE_STRICT ) ? c->sd : c->sq;
        int chunk = int( budget / ( row * elem ) );
        chunk = std::max( 1, std::min( chunk, 64 ) );
        if ( c->variant == TRACE_STRICT )
        {
            int block = 1024;
            while ( block > 64 && waves / uint64_t( block / 64 ) < uint64_t( c->n_cu ) ) block /= 2;
            block = block >= 1024 ? 1024 : ( block >= 256 ? 256 : 64 );
            c->shape = launch_shape { block, 1, chunk };
        }
        else    // x-major: 16 velocity rows per workgroup, the largest chunk planned
        {
            bool pipe = false;
            c->shape = launch_shape { 1024, 1, int( poly_chunk_plan( uint32_t( c->conf.Nx ), elem, 1u << 20,
                                                                    std::min<size_t>( c->lds_cap, 160 * 1024 ), &pipe ) ) };
        }
    }
    else
    {
        const size_t per = size_t( c->st ) * elem;
        int chunk = int( std::min<size_t>( c->lds_cap, 64 * 1024 ) / per );
        chunk = std::max( 1, std::min( chunk, 1 ) );
        // the autotuner's picks on 16^3 x 32^3 (two 1-slice buffers each); eight
        // FP64 nodes a lane trace twice the node-steps per slice barrier of four
        if ( c->variant == TRACE_STRICT ) c->shape = launch_shape { 256, 1, chunk };
        else if ( c->variant == TRACE_FP32C ) c->shape = launch_shape { 512, 4, 1 };
        else c->shape = launch_shape { 512, 8, chunk };
    }
}

int begin_timing( nufi_ctx *c, hipEvent_t *stop )
{
    *stop = nullptr;
    if ( !c->timing ) return NUFI_OK;
    if ( c->ev_used == c->ev.size() )
    {
        hipEvent_t a, b;
        HIP_TRY( c, hipEventCreate( &a ) );
        HIP_TRY( c, hipEventCreate( &b ) );
        c->ev.emplace_back( a, b );
    }
    auto &pr = c->ev[ c->ev_used++ ];
    HIP_TRY( c, hipEventRecord( pr.first, c->stream ) );
    *stop = pr.second;
    return NUFI_OK;
}

// rho partials of nodes [q_begin, q_end) at step n (no reduction).
int launch_trace( nufi_ctx *c, uint64_t n, uint64_t q_begin, uint64_t q_end, uint64_t *first, uint64_t *count )
{
    const uint64_t l0 = q_begin / c->n_vel, l1 = ( q_end - 1 ) / c->n_vel;
    *first = l0; *count = l1 - l0 + 1;
    const uint64_t tasks = ( l1 - l0 + 1 ) * c->wpx;
    if ( tasks >= ( uint64_t( 1 ) << 32 ) ) return fail( c, NUFI_ERR_RANGE, "compute_rho: too many wave tasks for one launch" );
    hipEvent_t stop;
    int rc = begin_timing( c, &stop );
    if ( rc ) return rc;
    hipError_t e;
    if ( c->dim == 1 )
    {
        trace1_params p {};
        p.dhist = c->dhist; p.partials = c->partials;
        p.qhist = c->variant == TRACE_FAST ? c->qhist : nullptr;
        p.qhist32 = c->variant == TRACE_FP32C ? c->qhist32 : nullptr;
        p.stride_q = c->sq; p.q2_off = poly_q2( uint32_t( c->conf.Nx ) ); p.kick = -0.5 * c->conf.dt * c->conf.dx_inv;
        p.n = n; p.q_begin = q_begin; p.q_end = q_end;
        p.Nx = uint32_t( c->conf.Nx ); p.Nu = uint32_t( c->conf.Nu ); p.stride_d = c->sd; p.wpx = c->wpx;
        p.ix_first = uint32_t( l0 ); p.n_tasks = uint32_t( tasks ); p.ix_count = uint32_t( l1 - l0 + 1 );
        p.order = uint32_t( c->order );
        p.mode = MODE_RHO;
        p.x_min = c->conf.x_min; p.dx = c->conf.dx; p.dx_inv = c->conf.dx_inv;
        p.Lx = c->conf.Lx; p.Lx_inv = c->conf.Lx_inv; p.dt = c->conf.dt;
        p.nx_min = c->conf.x_min; p.ndx = c->conf.dx; p.nu_min = c->conf.u_min; p.ndu = c->conf.du;
        fp32c_frame( double( c->conf.Nx ), c->conf.dt * c->conf.dx_inv, c->conf.u_min, double( c->conf.Nu ),
                     c->conf.du, p.f32_p0, p.f32_w0 );
        p.f0 = c->conf.f0;
        e = launch_trace1d( p, c->variant, c->shape, c->stream );
    }
    else
    {
        trace3_params p {};
        p.hist = c->hist; p.hist32 = c->hist32; p.partials = c->partials;
        p.mode = MODE_RHO;
        p.n = n; p.q_begin = q_begin; p.q_end = q_end;
        p.Nx = uint32_t( c->conf.Nx ); p.Ny = uint32_t( c->conf.Ny ); p.Nz = uint32_t( c->conf.Nz );
        p.Nu = uint32_t( c->conf.Nu ); p.Nv = uint32_t( c->conf.Nv ); p.Nw = uint32_t( c->conf.Nw );
        p.sy = c->sy; p.sz = c->sz; p.st = c->st; p.st32 = c->st32; p.wpx = c->wpx;
        p.l_first = uint32_t( l0 ); p.n_tasks = uint32_t( tasks );
        p.n_count = uint32_t( l1 - l0 + 1 ); p.vgroup = c->vgroup; p.ppx = c->ppx;
        p.order = uint32_t( c->order );
        fill_trace3_geometry( p, c->conf );
        e = launch_trace3d( p, c->variant, c->shape, c->stream );
    }
    if ( e != hipSuccess ) return hip_fail( c, e, "trace kernel launch" );
    if ( stop ) HIP_TRY( c, hipEventRecord( stop, c->stream ) );
    c->launches += 1;
    c->csteps += double( q_end - q_begin ) * double( n );
    return NUFI_OK;
}

// The scale the field and slice-table kernels apply to the FP32 per-slice
// tables (FP64 tables are not scaled): 1D fp32c1_scale, 3D fp32c3_scale.
double table_kick( const nufi_ctx *c )
{
    if ( c->dim == 1 ) return fp32c1_scale( c->conf.dt, c->conf.dx_inv );
    return fp32c3_scale( c->conf.dt, c->conf.dx_inv );
}

double rho_weight( const nufi_ctx *c )
{
    // -du per node (cuda_kernel.cu:47-50); 3D: -du*dv*dw (cuda_kernel.cu:419-425)
    if ( c->dim == 1 ) return -c->conf.du;
    return -( c->conf.du * c->conf.dv * c->conf.dw );
}

field_params base_field( nufi_ctx *c )
{
    field_params f {};
    f.dim = c->dim; f.order = c->order;
    f.Nx = uint32_t( c->conf.Nx );
    f.Ny = c->dim == 3 ? uint32_t( c->conf.Ny ) : 1u;
    f.Nz = c->dim == 3 ? uint32_t( c->conf.Nz ) : 1u;
    f.N = f.Nx * f.Ny * f.Nz;
    f.log2x = log2_or_none( f.Nx ); f.log2y = log2_or_none( f.Ny ); f.log2z = log2_or_none( f.Nz );
    f.wpx = rho_ppx( c );
    f.weight = rho_weight( c );
    f.tw_x = c->tw_x; f.tw_y = c->tw_y; f.tw_z = c->tw_z;
    f.ewt = c->ewt;
    f.sy = c->sy; f.sz = c->sz; f.stride_d = c->sd;
    return f;
}

// The field solve of f on the context's stream: one workgroup, or the
// multi-workgroup passes for grids one workgroup's LDS cannot hold.
hipError_t run_field( nufi_ctx *c, const field_params &f )
{
    if ( c->field_ok == 2 ) return launch_field_large( f, c->lw0, c->lw1, c->lpart, c->stream );
    return launch_field( f, c->stream );
}

} // namespace

int nufi_impl::validate_config( const nufi_config *c, int order ) { return validate( c, order ); }

void nufi_impl::set_context_free_error( const std::string &msg ) { g_create_error = msg; }

extern "C" {

const char *nufi_version( void ) { return "nufi-mi355x 0.1 (gfx950)"; }

int nufi_config_derive( nufi_config *c )
{
    if ( c == nullptr || c->Nx == 0 || c->Nu == 0 ) return NUFI_ERR_INVALID;
    // config.hpp:67-69 / 213-218
    c->Lx = c->x_max - c->x_min; c->Lx_inv = 1 / c->Lx;
    c->dx = c->Lx / c->Nx;       c->dx_inv = 1 / c->dx;
    c->du = ( c->u_max - c->u_min ) / c->Nu;
    if ( c->dim == 3 )
    {
        if ( c->Ny == 0 || c->Nz == 0 || c->Nv == 0 || c->Nw == 0 ) return NUFI_ERR_INVALID;
        c->Ly = c->y_max - c->y_min; c->Ly_inv = 1 / c->Ly;
        c->Lz = c->z_max - c->z_min; c->Lz_inv = 1 / c->Lz;
        c->dy = c->Ly / c->Ny;       c->dy_inv = 1 / c->dy;
        c->dz = c->Lz / c->Nz;       c->dz_inv = 1 / c->dz;
        c->dv = ( c->v_max - c->v_min ) / c->Nv;
        c->dw = ( c->w_max - c->w_min ) / c->Nw;
    }
    return NUFI_OK;
}

int nufi_device_count( int *count )
{
    if ( count == nullptr ) return NUFI_ERR_INVALID;
    int n = 0;
    hipError_t e = hipGetDeviceCount( &n );
    if ( e != hipSuccess ) { *count = 0; return fail( nullptr, NUFI_ERR_DEVICE, hipGetErrorString( e ) ); }
    *count = n;
    return NUFI_OK;
}

int nufi_create( const nufi_config *conf, const nufi_config *conf_metrics, int order, int precision,
                 int device, nufi_ctx **out )
{
    if ( out == nullptr ) return fail( nullptr, NUFI_ERR_INVALID, "nufi_create: out is NULL" );
    *out = nullptr;
    int rc = validate( conf, order );
    if ( rc ) return fail( nullptr, rc, "nufi_create: invalid or unsupported config/order" );
    if ( conf_metrics && validate( conf_metrics, order ) )
        return fail( nullptr, NUFI_ERR_INVALID, "nufi_create: invalid metrics config" );
    if ( precision != NUFI_FP64 && precision != NUFI_FP32C )
        return fail( nullptr, NUFI_ERR_UNSUPPORTED, "nufi_create: unknown precision" );
    if ( precision == NUFI_FP32C && order != 4 )
        return fail( nullptr, NUFI_ERR_UNSUPPORTED, "nufi_create: FP32C is built for order 4 only" );
    if ( conf->dim == 3 && order != 4 )
        return fail( nullptr, NUFI_ERR_UNSUPPORTED, "nufi_create: 3D is built for order 4 only" );

    // the trace kernels stage whole slices in LDS: 1D FP64 grids too wide for
    // one kick-polynomial row take the strict kernel (one difference row)
    int variant = precision == NUFI_FP32C ? TRACE_FP32C : ( ( conf->dim == 1 && order == 4 ) || conf->dim == 3 ? TRACE_FAST : TRACE_STRICT );
    if ( variant == TRACE_FAST && conf->dim == 1 && slice_lds_bytes( conf, order, TRACE_FAST ) > LDS_LIMIT )
        variant = TRACE_STRICT;
    if ( slice_lds_bytes( conf, order, variant ) > LDS_LIMIT )
        return fail( nullptr, NUFI_ERR_UNSUPPORTED,
                     "nufi_create: one history slice of this grid (" + std::to_string( slice_lds_bytes( conf, order, variant ) ) +
                     " bytes in the trace kernel's layout) exceeds the 160 KiB of LDS the kernel stages it in" );

    int ndev = 0;
    hipError_t he = hipGetDeviceCount( &ndev );
    if ( he != hipSuccess || ndev == 0 )
        return fail( nullptr, NUFI_ERR_DEVICE, "nufi_create: no HIP device" );
    if ( device < 0 || device >= ndev ) return fail( nullptr, NUFI_ERR_INVALID, "nufi_create: bad device index" );

    nufi_ctx *c = new ( std::nothrow ) nufi_ctx;
    if ( c == nullptr ) return fail( nullptr, NUFI_ERR_ALLOC, "nufi_create: host allocation" );
    c->conf = *conf;
    c->conf_m = conf_metrics ? *conf_metrics : *conf;
    c->order = order; c->precision = precision; c->device = device; c->dim = conf->dim;
    c->variant = variant;
    // NUFI_RUN_GRAPH=0 in the environment: nufi_run always launches each step
    {
        const char *g = std::getenv( "NUFI_RUN_GRAPH" );
        c->graph_allowed = !( g && g[0] == '0' );
        c->use_graph = c->graph_allowed;
    }

    auto bail = [&]( int code ) { std::string m = c->err; free_all( c ); delete c; return fail( nullptr, code, m ); };

    if ( ( he = hipSetDevice( device ) ) != hipSuccess ) { hip_fail( c, he, "hipSetDevice" ); return bail( NUFI_ERR_DEVICE ); }
    if ( ( he = hipStreamCreateWithFlags( &c->stream, hipStreamNonBlocking ) ) != hipSuccess )
    { hip_fail( c, he, "hipStreamCreate" ); return bail( NUFI_ERR_DEVICE ); }
    hipDeviceProp_t prop;
    if ( hipGetDeviceProperties( &prop, device ) == hipSuccess )
    {
        c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        int lds = 0;
        if ( hipDeviceGetAttribute( &lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device ) == hipSuccess && lds > 0 )
            c->lds_cap = size_t( lds );
    }

    const uint64_t o1 = uint64_t( order ) - 1;
    if ( c->dim == 1 )
    {
        c->st = uint32_t( conf->Nx + o1 );
        // extended differences D[-1] .. D[Nx+order-2]: Nx+order entries, padded
        // to a multiple of 4 so every slice starts 16-byte aligned
        c->sd = uint32_t( ( conf->Nx + order + 3 ) & ~uint64_t( 3 ) );
        c->sy = c->st; c->sz = c->st;
        c->n_space = conf->Nx; c->n_vel = conf->Nu;
    }
    else
    {
        c->sy = uint32_t( conf->Nx + o1 );
        c->sz = uint32_t( c->sy * ( conf->Ny + o1 ) );
        c->st = uint32_t( c->sz * ( conf->Nz + o1 ) );
        c->sd = c->st;
        c->n_space = conf->Nx * conf->Ny * conf->Nz;
        c->n_vel = conf->Nu * conf->Nv * conf->Nw;
    }
    c->n_nodes = c->n_space * c->n_vel;
    c->wpx = uint32_t( ( c->n_vel + 63 ) / 64 );
    if ( c->dim == 1 ) c->ppx = xm_ppx( uint32_t( conf->Nu ) );
    else
    {
        c->vgroup = xm3_vgroup( c->n_space, c->n_vel );
        c->ppx = uint32_t( ( c->n_vel + c->vgroup - 1 ) / c->vgroup );
    }
    c->field_ok = field_supported( c->dim, uint32_t( conf->Nx ), c->dim == 3 ? uint32_t( conf->Ny ) : 1u,
                                   c->dim == 3 ? uint32_t( conf->Nz ) : 1u );
    if ( !c->field_ok && ( c->dim == 3 || conf->Nx <= FIELD_LARGE_MAX_1D ) && c->n_space < ( uint64_t( 1 ) << 31 ) )
    {
        c->field_ok = 2;
        if ( ( rc = dev_alloc( c, &c->lw0, c->n_space, "field buffers" ) ) ) return bail( rc );
        if ( ( rc = dev_alloc( c, &c->lw1, c->n_space, "field buffers" ) ) ) return bail( rc );
        if ( ( rc = dev_alloc( c, &c->lpart, field_large_partials( uint32_t( c->n_space ) ), "field partials" ) ) )
            return bail( rc );
    }
    const uint64_t slices = conf->Nt + 1;
    if ( ( rc = dev_alloc( c, &c->hist, slices * c->st + 8, "history" ) ) ) return bail( rc );
    if ( c->dim == 1 && ( rc = dev_alloc( c, &c->dhist, slices * c->sd + 8, "difference history" ) ) ) return bail( rc );
    if ( c->dim == 1 ) c->sq = poly_stride( uint32_t( conf->Nx ) );
    // the FP64 rho kernel copies these tables instead of building them (3 reals per cell and slice)
    if ( c->dim == 1 && precision == NUFI_FP64 && order == 4 &&
         ( rc = dev_alloc( c, &c->qhist, slices * c->sq + 8, "table history" ) ) ) return bail( rc );
    if ( c->dim == 1 && precision == NUFI_FP32C &&
         ( rc = dev_alloc( c, &c->qhist32, slices * c->sq + 8, "FP32 table history" ) ) ) return bail( rc );
    c->st32 = ( c->st + 3u ) & ~3u;
    if ( precision == NUFI_FP32C && c->dim == 3 && ( rc = dev_alloc( c, &c->hist32, slices * c->st32 + 8, "FP32 history" ) ) )
        return bail( rc );
    if ( ( rc = dev_alloc( c, &c->partials, c->n_space * std::max( c->wpx, c->ppx ), "partials" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->rho, c->n_space, "rho" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->values, c->n_space, "values" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->energies, slices, "energies" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->scratch_slice, c->st + 8, "scratch slice" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->metrics, 8, "metrics" ) ) ) return bail( rc );
    if ( ( rc = dev_alloc( c, &c->energy_tmp, 8, "energy scratch" ) ) ) return bail( rc );
    // metrics node grid: conf_metrics in 1D (cuda_kernel.cu:97-109), conf in 3D
    if ( c->dim == 1 ) { c->m_space = c->conf_m.Nx; c->m_vel = c->conf_m.Nu; }
    else { c->m_space = c->n_space; c->m_vel = c->n_vel; }
    c->m_nodes = c->m_space * c->m_vel;
    c->m_wpx = uint32_t( ( c->m_vel + 63 ) / 64 );
    if ( ( rc = dev_alloc( c, &c->mpartials, 4 * c->m_space * c->m_wpx, "metric partials" ) ) ) return bail( rc );
    if ( c->field_ok && ( rc = build_tables( c ) ) ) return bail( rc );
    c->host_rho.assign( c->n_space, 0.0 );
    default_shape( c );
    *out = c;
    return NUFI_OK;
}

int nufi_destroy( nufi_ctx *c )
{
    if ( c == nullptr ) return NUFI_OK;
    (void)hipSetDevice( c->device );
    if ( c->stream ) (void)hipStreamSynchronize( c->stream );
    free_all( c );
    delete c;
    return NUFI_OK;
}

const char *nufi_last_error( const nufi_ctx *c )
{
    if ( c == nullptr ) return g_create_error.c_str();
    return c->err.empty() ? "no error" : c->err.c_str();
}

int nufi_get_stream( nufi_ctx *c, void **stream )
{
    if ( c == nullptr || stream == nullptr ) return NUFI_ERR_INVALID;
    *stream = reinterpret_cast<void*>( c->stream );
    return NUFI_OK;
}

int nufi_synchronize( nufi_ctx *c )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

// cuda_kernel.cu:112-132 / 485-510
int nufi_compute_rho( nufi_ctx *c, uint64_t n, uint64_t q_begin, uint64_t q_end )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( n > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "Time-step out of range." );
    if ( q_begin > q_end || q_end > c->n_nodes ) return fail( c, NUFI_ERR_RANGE, "compute_rho: node range out of range" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    if ( q_begin == q_end )
    {
        // empty range: the device rho becomes zero, so download_rho adds nothing
        hipError_t e = launch_reduce_partials( c->partials, rho_ppx( c ), 0.0, 0, 0, uint32_t( c->n_space ), c->rho, c->stream );
        if ( e != hipSuccess ) return hip_fail( c, e, "reduce launch" );
        return NUFI_OK;
    }
    uint64_t first, count;
    int rc = launch_trace( c, n, q_begin, q_end, &first, &count );
    if ( rc ) return rc;
    hipError_t e = launch_reduce_partials( c->partials, rho_ppx( c ), rho_weight( c ), uint32_t( first ), uint32_t( count ),
                                           uint32_t( c->n_space ), c->rho, c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, "reduce launch" );
    return NUFI_OK;
}

// cuda_kernel.cu:134-145 — synchronising copy, then rho[i] += partial
int nufi_download_rho( nufi_ctx *c, double *rho_accum )
{
    if ( c == nullptr || rho_accum == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipMemcpyAsync( c->host_rho.data(), c->rho, c->n_space * sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    for ( uint64_t i = 0; i < c->n_space; ++i ) rho_accum[i] += c->host_rho[i];
    return NUFI_OK;
}

// cuda_kernel.cu:147-156 — slice n from history_base + n*stride_t
int nufi_upload_phi( nufi_ctx *c, uint64_t n, const double *history_base )
{
    if ( c == nullptr || history_base == nullptr ) return NUFI_ERR_INVALID;
    if ( n > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "upload_phi: time step out of range" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    double *dst = c->hist + n * c->st;
    HIP_TRY( c, hipMemcpyAsync( dst, history_base + n * c->st, c->st * sizeof(double), hipMemcpyHostToDevice, c->stream ) );
    hipError_t e = launch_slice_tables( c->dim, dst, c->st, uint32_t( c->conf.Nx ),
                                        c->dhist ? c->dhist + n * c->sd : nullptr,
                                        c->qhist ? c->qhist + n * c->sq : nullptr,
                                        c->qhist32 ? c->qhist32 + n * c->sq : nullptr,
                                        table_kick( c ),
                                        c->hist32 ? c->hist32 + n * c->st32 : nullptr, c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, "slice table launch" );
    // the copy reads host memory the caller may reuse: complete it before returning
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

// cuda_kernel.cu:158-176 / 538-559: eval_f (with the initial half-kick from
// slice n, rho.hpp:63-96) on the metrics grid, four weighted sums.  FP64 in
// every precision (a diagnostic); the strict kernel in strict contexts.
int nufi_compute_metrics( nufi_ctx *c, uint64_t n, uint64_t q_begin, uint64_t q_end )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( n > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "compute_metrics: time step out of range" );
    if ( q_begin > q_end || q_end > c->m_nodes ) return fail( c, NUFI_ERR_RANGE, "compute_metrics: node range out of range" );
    if ( slice_lds_bytes( &c->conf, c->order, c->dim == 1 || c->variant == TRACE_STRICT ? TRACE_STRICT : TRACE_FAST ) > LDS_LIMIT )
        return fail( c, NUFI_ERR_UNSUPPORTED, "compute_metrics: one FP64 slice of this grid exceeds the LDS" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    uint64_t tasks = 0;
    if ( q_begin < q_end )
    {
        const uint64_t l0 = q_begin / c->m_vel, l1 = ( q_end - 1 ) / c->m_vel;
        tasks = ( l1 - l0 + 1 ) * c->m_wpx;
        const int variant = c->variant == TRACE_STRICT ? TRACE_STRICT : TRACE_FAST;
        hipError_t e;
        if ( c->dim == 1 )
        {
            const nufi_config &g = c->conf, &m = c->conf_m;
            trace1_params p {};
            p.dhist = c->dhist; p.partials = c->mpartials; p.mode = MODE_METRICS;
            p.n = n; p.q_begin = q_begin; p.q_end = q_end;
            p.Nx = uint32_t( g.Nx ); p.Nu = uint32_t( m.Nu ); p.stride_d = c->sd; p.wpx = c->m_wpx;
            p.ix_first = uint32_t( l0 ); p.n_tasks = uint32_t( tasks ); p.order = uint32_t( c->order );
            p.x_min = g.x_min; p.dx = g.dx; p.dx_inv = g.dx_inv; p.Lx = g.Lx; p.Lx_inv = g.Lx_inv; p.dt = g.dt;
            p.nx_min = m.x_min; p.ndx = m.dx; p.nu_min = m.u_min; p.ndu = m.du;
            p.weight = m.du * m.dx;                                // cuda_kernel.cu:71
            p.f0 = g.f0;
            const int chunk = int( std::max<size_t>( 1, std::min<size_t>( 64, std::min<size_t>( c->lds_cap, 65536 ) / ( size_t( c->sd ) * 8 ) ) ) );
            e = launch_trace1d( p, variant, launch_shape { 256, 1, chunk }, c->stream );
        }
        else
        {
            const nufi_config &g = c->conf;
            trace3_params p {};
            p.hist = c->hist; p.partials = c->mpartials; p.mode = MODE_METRICS;
            p.weight = g.du * g.dv * g.dw;                         // cuda_kernel.cu:451 (no dx dy dz)
            p.n = n; p.q_begin = q_begin; p.q_end = q_end;
            p.Nx = uint32_t( g.Nx ); p.Ny = uint32_t( g.Ny ); p.Nz = uint32_t( g.Nz );
            p.Nu = uint32_t( g.Nu ); p.Nv = uint32_t( g.Nv ); p.Nw = uint32_t( g.Nw );
            p.sy = c->sy; p.sz = c->sz; p.st = c->st; p.wpx = c->m_wpx;
            p.l_first = uint32_t( l0 ); p.n_tasks = uint32_t( tasks ); p.order = uint32_t( c->order );
            fill_trace3_geometry( p, g );
            e = launch_trace3d( p, variant, launch_shape { 256, 1, 1 }, c->stream );
        }
        if ( e != hipSuccess ) return hip_fail( c, e, "metrics trace launch" );
    }
    hipError_t e = launch_reduce_metrics( c->mpartials, uint32_t( tasks ), c->metrics, c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, "metrics reduce launch" );
    return NUFI_OK;
}

// cuda_kernel.cu:178-189 — synchronise, metrics[k] += device metrics
int nufi_download_metrics( nufi_ctx *c, double *metrics )
{
    if ( c == nullptr || metrics == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    double m[4] = { 0, 0, 0, 0 };
    HIP_TRY( c, hipMemcpyAsync( m, c->metrics, sizeof(m), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    for ( int k = 0; k < 4; ++k ) metrics[k] += m[k];
    return NUFI_OK;
}

static int step_async( nufi_ctx *c, uint64_t n, double *rho_out = nullptr );

int adjoint_npar( const nufi_f0 &f )
{
    if ( f.kind == NUFI_F0_MULTIMODE_1D ) return f.n_modes;
    if ( f.kind == NUFI_F0_LANDAU_1D || f.kind == NUFI_F0_TWOSTREAM_1D ) return 1;
    return 0;
}

static int step_async( nufi_ctx *c, uint64_t n, double *rho_out )
{
    if ( n > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "step: time step out of range" );
    if ( !c->field_ok ) return fail( c, NUFI_ERR_UNSUPPORTED, "step: grid too large for the single-workgroup field solve" );
    uint64_t first, count;
    int rc = launch_trace( c, n, 0, c->n_nodes, &first, &count );
    if ( rc ) return rc;
    field_params f = base_field( c );
    f.mode = FIELD_STEP;
    f.partials = c->partials;
    if ( uint64_t( f.wpx ) * c->n_space > ( uint64_t( 1 ) << 16 ) )
    {
        // many partials (3D): one thread per point over many workgroups sums them
        // in the same order, and the single field workgroup starts from rho
        double *dst = rho_out ? rho_out : c->rho;
        hipError_t e = launch_reduce_partials( c->partials, f.wpx, f.weight, 0, uint32_t( c->n_space ),
                                               uint32_t( c->n_space ), dst, c->stream );
        if ( e != hipSuccess ) return hip_fail( c, e, "reduce launch" );
        f.partials = nullptr;
        f.values = dst;
    }
    f.mult = c->mult_step;
    f.slice = c->hist + n * c->st;
    f.dslice = c->dhist ? c->dhist + n * c->sd : nullptr;
    f.qslice = c->qhist ? c->qhist + n * c->sq : nullptr;
    f.qslice32 = c->qhist32 ? c->qhist32 + n * c->sq : nullptr;
    f.q2_off = poly_q2( uint32_t( c->conf.Nx ) );
    f.kick = table_kick( c );
    f.slice32 = c->hist32 ? c->hist32 + n * c->st32 : nullptr;
    f.energy = c->energies + n;
    f.rho_out = rho_out ? rho_out : c->rho;
    hipError_t e = run_field( c, f );
    if ( e != hipSuccess ) return hip_fail( c, e, "field kernel launch" );
    return NUFI_OK;
}

int nufi_step( nufi_ctx *c, uint64_t n, double *energy_out )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    int rc = step_async( c, n );
    if ( rc ) return rc;
    if ( energy_out )
    {
        HIP_TRY( c, hipMemcpyAsync( energy_out, c->energies + n, sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
        HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    }
    return NUFI_OK;
}

int nufi_run( nufi_ctx *c, uint64_t n_begin, uint64_t n_end, double *energies )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( n_end < n_begin || ( n_end > 0 && n_end - 1 > c->conf.Nt ) ) return fail( c, NUFI_ERR_RANGE, "run: step range out of range" );
    if ( n_end > n_begin && !c->field_ok )
        return fail( c, NUFI_ERR_UNSUPPORTED, "step: grid too large for the single-workgroup field solve" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    if ( c->use_graph && !c->timing && n_end > n_begin )
    {
        // the trace and field launches of every step, captured once and replayed
        // (no memset or event in the captured range: every buffer a step reads
        // is written by a kernel of that step or an earlier one)
        const bool same = c->run_graph && c->graph_b == n_begin && c->graph_e == n_end &&
                          c->graph_variant == c->variant && c->graph_shape.block == c->shape.block &&
                          c->graph_shape.npt == c->shape.npt && c->graph_shape.chunk == c->shape.chunk;
        if ( !same )
        {
            if ( c->run_graph ) { HIP_TRY( c, hipGraphExecDestroy( c->run_graph ) ); c->run_graph = nullptr; }
            HIP_TRY( c, hipStreamBeginCapture( c->stream, hipStreamCaptureModeThreadLocal ) );
            int rc = NUFI_OK;
            for ( uint64_t n = n_begin; n < n_end && rc == NUFI_OK; ++n ) rc = step_async( c, n );
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture( c->stream, &g );
            if ( rc ) { if ( g ) (void)hipGraphDestroy( g ); return rc; }
            if ( e != hipSuccess ) return hip_fail( c, e, "hipStreamEndCapture" );
            e = hipGraphInstantiate( &c->run_graph, g, nullptr, nullptr, 0 );
            (void)hipGraphDestroy( g );
            if ( e != hipSuccess ) { c->run_graph = nullptr; return hip_fail( c, e, "hipGraphInstantiate" ); }
            c->graph_b = n_begin; c->graph_e = n_end;
            c->graph_variant = c->variant; c->graph_shape = c->shape;
        }
        HIP_TRY( c, hipGraphLaunch( c->run_graph, c->stream ) );
    }
    else
    {
        for ( uint64_t n = n_begin; n < n_end; ++n )
        {
            int rc = step_async( c, n );
            if ( rc ) return rc;
        }
    }
    if ( energies && n_end > n_begin )
    {
        HIP_TRY( c, hipMemcpyAsync( energies, c->energies + n_begin, ( n_end - n_begin ) * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream ) );
        HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    }
    return NUFI_OK;
}

// Gradient of J = sum_{n=0}^{Nt} dt * energy_n with respect to the initial
// condition's parameters (adjoint1d.hip): forward run storing rho_n, then the
// reverse sweep.  1D, order 4, FP64.
int nufi_adjoint_grad( nufi_ctx *c, double *dJ_da, int n_par, double *J_out )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( c->dim != 1 || c->order != 4 )
        return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: built for 1D, order 4" );
    const int npar = adjoint_npar( c->conf.f0 );
    if ( npar == 0 ) return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: no parameters for this f0 family" );
    if ( n_par < npar || dJ_da == nullptr ) return fail( c, NUFI_ERR_INVALID, "adjoint_grad: output too short" );
    if ( !c->field_ok ) return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: grid too large for the field solve" );
    // one slice's kick-polynomial table and its slice adjoints in LDS
    if ( size_t( c->sq ) * ( c->precision == NUFI_FP32C ? sizeof(float) : sizeof(double) ) +
         c->conf.Nx * sizeof(double) + 16 > LDS_LIMIT )
        return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: one slice's tables and adjoints exceed the LDS" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    const uint64_t Nt = c->conf.Nt, Nx = c->conf.Nx;
    const uint32_t n_tasks = uint32_t( c->n_space * c->wpx );
    int rc;
    if ( c->rho_hist == nullptr )
    {
        if ( ( rc = dev_alloc( c, &c->rho_hist, ( Nt + 1 ) * Nx, "adjoint rho history" ) ) ) return rc;
        if ( ( rc = dev_alloc( c, &c->lambda, ( Nt + 1 ) * Nx, "adjoint slice adjoints" ) ) ) return rc;
        if ( ( rc = dev_alloc( c, &c->gvec, Nx, "adjoint g" ) ) ) return rc;
        if ( ( rc = dev_alloc( c, &c->grad_part, size_t( n_tasks ) * NUFI_MAX_MODES, "adjoint partials" ) ) ) return rc;
        if ( ( rc = dev_alloc( c, &c->grad, NUFI_MAX_MODES, "adjoint gradient" ) ) ) return rc;
    }
    // forward run (the FP64 fast or the FP32C trace, as nufi_run), keeping
    // rho_n; a STRICT context's launch shape is not one the fast kernel takes,
    // so the fast default shape runs it and both are restored afterwards
    const int saved_variant = c->variant;
    const launch_shape saved_shape = c->shape;
    if ( saved_variant == TRACE_STRICT )
    {
        if ( slice_lds_bytes( &c->conf, c->order, TRACE_FAST ) > LDS_LIMIT )
            return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: grid too large for the fast trace" );
        c->variant = TRACE_FAST;
        default_shape( c );
    }
    for ( uint64_t n = 0; n <= Nt && ( rc = step_async( c, n, c->rho_hist + n * Nx ) ) == NUFI_OK; ++n ) {}
    c->variant = saved_variant;
    c->shape = saved_shape;
    if ( rc ) return rc;
    HIP_TRY( c, hipMemsetAsync( c->lambda, 0, ( Nt + 1 ) * Nx * sizeof(double), c->stream ) );
    HIP_TRY( c, hipMemsetAsync( c->grad_part, 0, size_t( n_tasks ) * NUFI_MAX_MODES * sizeof(double), c->stream ) );

    adjoint_field_params af {};
    af.N = uint32_t( Nx ); af.log2x = log2_or_none( Nx );
    af.g = c->gvec; af.tw = c->tw_x; af.mult = c->mult_step; af.ewt = c->ewt; af.dt = c->conf.dt;
    adjoint1_params ap {};
    const bool f32 = c->precision == NUFI_FP32C;
    ap.qhist = f32 ? nullptr : c->qhist; ap.qhist32 = f32 ? c->qhist32 : nullptr;
    ap.g = c->gvec; ap.lambda = c->lambda; ap.grad_part = c->grad_part;
    ap.Nx = uint32_t( Nx ); ap.Nu = uint32_t( c->conf.Nu ); ap.stride_d = c->sd;
    ap.stride_q = c->sq; ap.q2_off = poly_q2( uint32_t( Nx ) );
    ap.wpx = c->wpx;
    ap.n_tasks = n_tasks; ap.npar = uint32_t( npar );
    ap.x_min = c->conf.x_min; ap.dx = c->conf.dx; ap.dx_inv = c->conf.dx_inv; ap.dt = c->conf.dt;
    ap.u_min = c->conf.u_min; ap.du = c->conf.du; ap.f0 = c->conf.f0;
    fp32c_frame( double( Nx ), c->conf.dt * c->conf.dx_inv, c->conf.u_min, double( c->conf.Nu ), c->conf.du,
                 ap.f32_p0, ap.f32_w0 );
    const size_t per_slice = size_t( c->sq ) * ( f32 ? sizeof(float) : sizeof(double) ) + Nx * sizeof(double);
    ap.chunk = uint32_t( std::max<size_t>( 1, std::min<size_t>( 32, std::min<size_t>( c->lds_cap, 65536 ) / per_slice ) ) );
    if ( adjoint1d_lds_bytes( ap ) > LDS_LIMIT )
        return fail( c, NUFI_ERR_UNSUPPORTED, "adjoint_grad: one slice's tables and adjoints exceed the LDS" );
    for ( uint64_t n = Nt + 1; n-- > 0; )
    {
        af.rho = c->rho_hist + n * Nx; af.lambda = c->lambda + n * Nx;
        hipError_t e = c->field_ok == 2 ? launch_adjoint_field_large( af, c->lw0, c->lw1, c->stream )
                                        : launch_adjoint_field( af, c->stream );
        if ( e != hipSuccess ) return hip_fail( c, e, "adjoint field launch" );
        ap.n = n;
        e = launch_adjoint1d( ap, c->stream );
        if ( e != hipSuccess ) return hip_fail( c, e, "adjoint trace launch" );
    }
    hipError_t e = launch_reduce_grad( c->grad_part, n_tasks, uint32_t( npar ), c->grad, c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, "adjoint reduce launch" );
    std::vector<double> en( Nt + 1 );
    HIP_TRY( c, hipMemcpyAsync( dJ_da, c->grad, npar * sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipMemcpyAsync( en.data(), c->energies, ( Nt + 1 ) * sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    if ( J_out )
    {
        double J = 0;
        for ( double v : en ) J += c->conf.dt * v;
        *J_out = J;
    }
    return NUFI_OK;
}

int nufi_download_history( nufi_ctx *c, uint64_t n_begin, uint64_t n_end, double *out )
{
    if ( c == nullptr || out == nullptr ) return NUFI_ERR_INVALID;
    if ( n_end < n_begin || n_end > c->conf.Nt + 1 ) return fail( c, NUFI_ERR_RANGE, "download_history: range" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipMemcpyAsync( out, c->hist + n_begin * c->st, ( n_end - n_begin ) * c->st * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

int nufi_upload_history( nufi_ctx *c, uint64_t n_begin, uint64_t n_end, const double *in )
{
    if ( c == nullptr || in == nullptr ) return NUFI_ERR_INVALID;
    if ( n_end < n_begin || n_end > c->conf.Nt + 1 ) return fail( c, NUFI_ERR_RANGE, "upload_history: range" );
    for ( uint64_t n = n_begin; n < n_end; ++n )
    {
        int rc = nufi_upload_phi( c, n, in - n_begin * c->st );
        if ( rc ) return rc;
    }
    return NUFI_OK;
}

int nufi_download_energies( nufi_ctx *c, uint64_t n_begin, uint64_t n_end, double *out )
{
    if ( c == nullptr || out == nullptr ) return NUFI_ERR_INVALID;
    if ( n_end < n_begin || n_end > c->conf.Nt + 1 ) return fail( c, NUFI_ERR_RANGE, "download_energies: range" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipMemcpyAsync( out, c->energies + n_begin, ( n_end - n_begin ) * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

// ---- diagnostics samplers (bin/test_nufi_gpu_1d.cpp:290-361) ---------------
namespace
{

int sample_1d( nufi_ctx *c, int what, uint64_t n, const double *x, const double *v, uint64_t count,
               double *o0, double *o1 )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( c->dim != 1 ) return fail( c, NUFI_ERR_UNSUPPORTED, "samplers: 1D contexts only" );
    if ( n > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "samplers: time step out of range" );
    if ( count == 0 ) return NUFI_OK;
    if ( x == nullptr || o0 == nullptr || ( what != SAMPLE_FIELD && v == nullptr ) ||
         ( what != SAMPLE_F && o1 == nullptr ) )
        return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    if ( c->samp_cap < count )
    {
        HIP_TRY( c, hipStreamSynchronize( c->stream ) );
        if ( c->samp ) HIP_TRY( c, hipFree( c->samp ) );
        c->samp = nullptr; c->samp_cap = 0;
        if ( int rc = dev_alloc( c, &c->samp, 4 * count, "sampler buffers" ) ) return rc;
        c->samp_cap = count;
    }
    double *dx = c->samp, *dv = c->samp + count, *d0 = c->samp + 2 * count, *d1 = c->samp + 3 * count;
    const size_t bytes = count * sizeof(double);
    HIP_TRY( c, hipMemcpyAsync( dx, x, bytes, hipMemcpyHostToDevice, c->stream ) );
    if ( what != SAMPLE_FIELD ) HIP_TRY( c, hipMemcpyAsync( dv, v, bytes, hipMemcpyHostToDevice, c->stream ) );
    sample1_params p {};
    p.dhist = c->dhist; p.x = dx; p.v = dv; p.out0 = d0; p.out1 = d1;
    p.n = n; p.count = count; p.stride_d = c->sd;
    p.x_min = c->conf.x_min; p.Lx = c->conf.Lx; p.Lx_inv = c->conf.Lx_inv; p.dx_inv = c->conf.dx_inv;
    p.dt = c->conf.dt; p.f0 = c->conf.f0;
    hipError_t e = launch_sample1d( p, what, c->order, c->stream );
    if ( e != hipSuccess ) return hip_fail( c, e, "sampler launch" );
    HIP_TRY( c, hipMemcpyAsync( o0, d0, bytes, hipMemcpyDeviceToHost, c->stream ) );
    if ( what != SAMPLE_F ) HIP_TRY( c, hipMemcpyAsync( o1, d1, bytes, hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

} // namespace

int nufi_sample_field( nufi_ctx *c, uint64_t n, const double *x, uint64_t count, double *E, double *rho )
{
    return sample_1d( c, SAMPLE_FIELD, n, x, nullptr, count, E, rho );
}

int nufi_sample_f( nufi_ctx *c, uint64_t n, const double *x, const double *v, uint64_t count, double *f )
{
    return sample_1d( c, SAMPLE_F, n, x, v, count, f, nullptr );
}

int nufi_phase_flow( nufi_ctx *c, uint64_t n, const double *x, const double *v, uint64_t count,
                     double *x_out, double *v_out )
{
    return sample_1d( c, SAMPLE_FLOW, n, x, v, count, x_out, v_out );
}

// ---- history files ---------------------------------------------------------
int nufi_save_history( nufi_ctx *c, const char *path, int format, uint64_t n_slices )
{
    if ( c == nullptr || path == nullptr ) return NUFI_ERR_INVALID;
    if ( n_slices > c->conf.Nt + 1 ) return fail( c, NUFI_ERR_RANGE, "save_history: more slices than the history holds" );
    std::vector<double> h( n_slices * c->st ), e( n_slices );
    int rc = nufi_download_history( c, 0, n_slices, h.data() );
    if ( rc == NUFI_OK ) rc = nufi_download_energies( c, 0, n_slices, e.data() );
    if ( rc ) return rc;
    nufi_history_info info {};
    info.dim = c->dim; info.order = c->order;
    info.Nx = c->conf.Nx;
    info.Ny = c->dim == 3 ? c->conf.Ny : 1;
    info.Nz = c->dim == 3 ? c->conf.Nz : 1;
    info.Nt = c->conf.Nt; info.n_slices = n_slices; info.stride_t = c->st; info.dt = c->conf.dt;
    info.n_energies = format == NUFI_HISTORY_BINARY ? n_slices : 0;
    rc = nufi_history_write( path, format, &info, h.data(), e.data() );
    if ( rc ) return fail( c, rc, std::string( "save_history: cannot write " ) + path );
    return NUFI_OK;
}

int nufi_load_history( nufi_ctx *c, const char *path, uint64_t *n_slices )
{
    if ( c == nullptr || path == nullptr ) return NUFI_ERR_INVALID;
    nufi_history_info info {};
    int fmt = 0;
    int rc = nufi_history_read_info( path, &info, &fmt );
    if ( rc ) return fail( c, rc, std::string( "load_history: cannot read " ) + path );
    const bool same = info.dim == c->dim && info.order == c->order && info.Nx == c->conf.Nx &&
                      info.Ny == ( c->dim == 3 ? c->conf.Ny : 1 ) && info.Nz == ( c->dim == 3 ? c->conf.Nz : 1 ) &&
                      info.stride_t == c->st;
    // the text header carries dt with six decimals (std::to_string)
    const bool dt_ok = fmt == NUFI_HISTORY_BINARY ? info.dt == c->conf.dt
                                                  : std::fabs( info.dt - c->conf.dt ) <= 5e-7;
    if ( !same || !dt_ok ) return fail( c, NUFI_ERR_INVALID, "load_history: the file is for another grid" );
    if ( info.Nt > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "load_history: the file's run is longer than Nt" );
    std::vector<double> h( info.n_slices * c->st ), e( info.n_energies );
    rc = nufi_history_read( path, &info, h.data(), info.n_slices, e.data(), info.n_energies );
    if ( rc ) return fail( c, rc, std::string( "load_history: cannot read " ) + path );
    rc = nufi_upload_history( c, 0, info.n_slices, h.data() );
    if ( rc ) return rc;
    if ( info.n_energies )
    {
        HIP_TRY( c, hipMemcpyAsync( c->energies, e.data(), std::min( info.n_energies, info.n_slices ) * sizeof(double),
                                    hipMemcpyHostToDevice, c->stream ) );
        HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    }
    if ( n_slices ) *n_slices = info.n_slices;
    return NUFI_OK;
}

// poisson.cpp:66-89 / 328-362 on a host buffer
int nufi_poisson_solve( nufi_ctx *c, double *data, double *energy )
{
    if ( c == nullptr || data == nullptr ) return NUFI_ERR_INVALID;
    if ( !c->field_ok ) return fail( c, NUFI_ERR_UNSUPPORTED, "poisson_solve: grid too large" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipMemcpyAsync( c->values, data, c->n_space * sizeof(double), hipMemcpyHostToDevice, c->stream ) );
    field_params f = base_field( c );
    f.mode = FIELD_POISSON; f.values = c->values; f.mult = c->mult_pois; f.energy = c->energy_tmp;
    hipError_t e = run_field( c, f );
    if ( e != hipSuccess ) return hip_fail( c, e, "field kernel launch" );
    HIP_TRY( c, hipMemcpyAsync( data, c->values, c->n_space * sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    double en = 0;
    HIP_TRY( c, hipMemcpyAsync( &en, c->energy_tmp, sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    if ( energy ) *energy = en;
    return NUFI_OK;
}

// fields.hpp:63-142 / 352-490 on host buffers
int nufi_interpolate( nufi_ctx *c, double *slice, const double *values )
{
    if ( c == nullptr || slice == nullptr || values == nullptr ) return NUFI_ERR_INVALID;
    if ( !c->field_ok ) return fail( c, NUFI_ERR_UNSUPPORTED, "interpolate: grid too large" );
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipMemcpyAsync( c->values, values, c->n_space * sizeof(double), hipMemcpyHostToDevice, c->stream ) );
    field_params f = base_field( c );
    f.mode = FIELD_FIT; f.values = c->values; f.mult = c->mult_fit; f.slice = c->scratch_slice;
    hipError_t e = run_field( c, f );
    if ( e != hipSuccess ) return hip_fail( c, e, "field kernel launch" );
    HIP_TRY( c, hipMemcpyAsync( slice, c->scratch_slice, c->st * sizeof(double), hipMemcpyDeviceToHost, c->stream ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    return NUFI_OK;
}

int nufi_set_trace_variant( nufi_ctx *c, int variant )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( c->precision == NUFI_FP32C )
        return variant == TRACE_FP32C ? NUFI_OK : fail( c, NUFI_ERR_UNSUPPORTED, "FP32C context: variant is fixed" );
    if ( variant != TRACE_FAST && variant != TRACE_STRICT )
        return fail( c, NUFI_ERR_UNSUPPORTED, "set_trace_variant: FP64 contexts take FAST or STRICT" );
    if ( variant == TRACE_FAST && c->order != 4 )
        return fail( c, NUFI_ERR_UNSUPPORTED, "set_trace_variant: FAST is built for order 4 only" );
    if ( slice_lds_bytes( &c->conf, c->order, variant ) > LDS_LIMIT )
        return fail( c, NUFI_ERR_UNSUPPORTED, "set_trace_variant: one slice of this grid exceeds the LDS in that kernel's layout" );
    c->variant = variant;
    default_shape( c );
    return NUFI_OK;
}

int nufi_set_launch_shape( nufi_ctx *c, int block, int npt, int chunk )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    const launch_shape s { block, npt, chunk };
    const bool ok = c->dim == 1 ? trace1d_shape_ok( c->variant, c->order, s )
                                : trace3d_shape_ok( c->variant, c->order, s );
    size_t lds = 0;
    if ( ok ) (void)effective_chunk( c, s, &lds );
    if ( !ok || lds > c->lds_cap || lds > 160u * 1024u )
        return fail( c, NUFI_ERR_UNSUPPORTED, "set_launch_shape: shape not built or LDS too large" );
    c->shape = s;
    return NUFI_OK;
}

int nufi_get_launch_shape( nufi_ctx *c, int *block, int *npt, int *chunk )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( block ) *block = c->shape.block;
    if ( npt ) *npt = c->shape.npt;
    if ( chunk ) *chunk = c->shape.chunk;
    return NUFI_OK;
}

int nufi_autotune( nufi_ctx *c, uint64_t n_probe, int repeats, int *block, int *npt, int *chunk, double *best_ms )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    if ( n_probe > c->conf.Nt ) return fail( c, NUFI_ERR_RANGE, "autotune: probe step out of range" );
    if ( repeats < 1 ) repeats = 1;
    HIP_TRY( c, hipSetDevice( c->device ) );
    std::vector<launch_shape> cand;
    const int blocks1[] = { 256, 512, 1024 }, npts[] = { 1, 2, 4 };
    if ( c->dim == 1 )
    {
        if ( c->variant == TRACE_STRICT )
        { for ( int b : { 256, 1024 } ) for ( int ch : { 16, 32, 64 } ) cand.push_back( { b, 1, ch } ); }
        else    // x-major: trace1d_shape_ok keeps the shapes with whole summation groups
            for ( int b : blocks1 ) for ( int np : npts ) for ( int ch : { 4, 8, 16, 32, 1 << 20 } ) cand.push_back( { b, np, ch } );
    }
    else
    {
        if ( c->variant == TRACE_STRICT ) cand.push_back( { 256, 1, 1 } );
        else
        {
            for ( int b : blocks1 ) for ( int np : npts ) for ( int ch : { 1, 2 } ) cand.push_back( { b, np, ch } );
            if ( c->variant == TRACE_FAST ) for ( int ch : { 1, 2 } ) cand.push_back( { 512, 8, ch } );
        }
    }
    const launch_shape saved = c->shape;
    const bool saved_timing = c->timing;
    c->timing = false;
    hipEvent_t e0, e1;
    HIP_TRY( c, hipEventCreate( &e0 ) );
    HIP_TRY( c, hipEventCreate( &e1 ) );
    double best = 1e300;
    launch_shape best_s = saved;
    int rc = NUFI_OK;
    std::vector<launch_shape> tried;
    for ( const launch_shape &sh : cand )
    {
        const bool ok = c->dim == 1 ? trace1d_shape_ok( c->variant, c->order, sh ) : trace3d_shape_ok( c->variant, c->order, sh );
        size_t lds = 0;
        if ( ok ) (void)effective_chunk( c, sh, &lds );
        if ( !ok || lds > std::min<size_t>( c->lds_cap, 160 * 1024 ) ) continue;
        bool seen = false;      // x-major chunks above the plan run as the plan: time each launch once
        for ( const launch_shape &t : tried )
            seen = seen || ( t.block == sh.block && t.npt == sh.npt && effective_chunk( c, t ) == effective_chunk( c, sh ) );
        if ( seen ) continue;
        tried.push_back( sh );
        c->shape = sh;
        uint64_t f, n;
        if ( ( rc = launch_trace( c, n_probe, 0, c->n_nodes, &f, &n ) ) ) break;    // warm-up
        hipError_t he = hipEventRecord( e0, c->stream );
        for ( int r = 0; r < repeats && he == hipSuccess && rc == NUFI_OK; ++r ) rc = launch_trace( c, n_probe, 0, c->n_nodes, &f, &n );
        if ( he == hipSuccess ) he = hipEventRecord( e1, c->stream );
        if ( he == hipSuccess ) he = hipEventSynchronize( e1 );
        float ms = 0;
        if ( he == hipSuccess ) he = hipEventElapsedTime( &ms, e0, e1 );
        if ( he != hipSuccess ) { rc = hip_fail( c, he, "autotune timing" ); break; }
        if ( rc ) break;
        if ( ms / repeats < best ) { best = ms / repeats; best_s = sh; }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    c->timing = saved_timing;
    c->launches = 0; c->csteps = 0.0; c->ev_used = 0;
    c->shape = ( rc == NUFI_OK ) ? best_s : saved;
    if ( rc ) return rc;
    if ( block ) *block = c->shape.block;
    if ( npt ) *npt = c->shape.npt;
    if ( chunk ) *chunk = c->shape.chunk;
    if ( best_ms ) *best_ms = best;
    return NUFI_OK;
}

int nufi_set_run_graph( nufi_ctx *c, int enabled )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    c->use_graph = enabled != 0 && c->graph_allowed;
    return NUFI_OK;
}

int nufi_set_timing( nufi_ctx *c, int enabled )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    c->timing = enabled != 0;
    c->ev_used = 0; c->launches = 0; c->csteps = 0.0;
    return NUFI_OK;
}

int nufi_trace_timing( nufi_ctx *c, uint64_t *launches, double *kernel_ms, double *char_steps )
{
    if ( c == nullptr ) return NUFI_ERR_INVALID;
    HIP_TRY( c, hipSetDevice( c->device ) );
    HIP_TRY( c, hipStreamSynchronize( c->stream ) );
    double ms = 0.0;
    for ( size_t i = 0; i < c->ev_used; ++i )
    {
        float t = 0.0f;
        HIP_TRY( c, hipEventElapsedTime( &t, c->ev[i].first, c->ev[i].second ) );
        ms += t;
    }
    if ( launches ) *launches = c->launches;
    if ( kernel_ms ) *kernel_ms = ms;
    if ( char_steps ) *char_steps = c->csteps;
    return NUFI_OK;
}

} // extern "C"
