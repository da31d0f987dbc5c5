// batch.hip — many independent 1D runs (uneven grids and parameters) stepped
// together on one device: per time step one trace launch covers the wave-tasks
// of every run (each workgroup belongs to one run and streams that run's
// history through LDS) and one field launch solves every run's field (one
// workgroup per run).  A range of steps is captured once into a hipGraph and
// replayed.  The per-run arithmetic is the single-run kernels' (the same device
// bodies), so each run's history is bit-identical to a single-run context's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "host_tables.hpp"
#include "internal.hpp"

using namespace nufi_impl;

struct nufi_batch
{
    int device = 0, order = 4, precision = NUFI_FP64, variant = TRACE_FAST;
    hipStream_t stream = nullptr;
    std::vector<nufi_config> confs;
    uint64_t n_runs = 0, Nt_max = 0;
    std::vector<size_t> off_hist, off_dhist, off_q, off_part, off_en, off_tab;
    std::vector<uint32_t> st, sd, sq, wpx;

    double *hist = nullptr, *dhist = nullptr, *partials = nullptr, *energies = nullptr;
    double *qhist = nullptr;       // FP64: every run's ready kick-polynomial tables (stride sq)
    float *qhist32 = nullptr;      // FP32C: the same in FP32, scaled by the run's fp32c1_scale
    double2 *tw = nullptr, *mult = nullptr;
    double *ewt = nullptr;
    batch_run *d_runs = nullptr;
    batch_field *d_fields = nullptr;
    uint32_t *d_wg_run = nullptr;
    size_t d_wg_cap = 0;
    std::vector<batch_run> h_runs;
    uint32_t n_wg = 0, max_N = 0, max_sq = 0;
    launch_shape shape { 1024, 1, 16 };
    size_t trace_lds = 0, field_lds = 0;

    hipGraphExec_t graph = nullptr;
    bool graph_ok = true;      // false when created with NUFI_RUN_GRAPH=0 in the environment
    uint64_t graph_b = 0, graph_e = 0;
    int graph_block = 0, graph_npt = 0, graph_chunk = 0;
    std::string err;
};

namespace
{

thread_local std::string g_batch_error = "no error";

int bfail( nufi_batch *b, int code, const std::string &m )
{
    if ( b ) b->err = m; else g_batch_error = m;
    return code;
}

int bhip( nufi_batch *b, hipError_t e, const char *where )
{
    return bfail( b, e == hipErrorOutOfMemory ? NUFI_ERR_ALLOC : NUFI_ERR_DEVICE,
                  std::string( where ) + ": " + hipGetErrorString( e ) );
}

#define BHIP( b, expr ) do { hipError_t e_ = ( expr ); if ( e_ != hipSuccess ) return bhip( ( b ), e_, #expr ); } while ( 0 )

size_t align_up( size_t v, size_t a ) { return ( v + a - 1 ) / a * a; }

template <typename T>
int balloc( nufi_batch *b, T **p, size_t count, const char *what )
{
    if ( count == 0 ) count = 1;
    hipError_t e = hipMalloc( reinterpret_cast<void**>( p ), count * sizeof(T) );
    if ( e != hipSuccess ) return bhip( b, e, what );
    // zeroed and waited for (a null-stream memset is not ordered before the
    // work of the batch's non-blocking stream)
    e = hipMemsetAsync( *p, 0, count * sizeof(T), b->stream );
    if ( e == hipSuccess ) e = hipStreamSynchronize( b->stream );
    if ( e != hipSuccess ) return bhip( b, e, what );
    return NUFI_OK;
}

void bfree_all( nufi_batch *b )
{
    if ( b->graph ) (void)hipGraphExecDestroy( b->graph );
    void *ptrs[] = { b->hist, b->dhist, b->qhist, b->qhist32, b->partials, b->energies, b->tw, b->mult, b->ewt,
                     b->d_runs, b->d_fields, b->d_wg_run };
    for ( void *p : ptrs ) if ( p ) (void)hipFree( p );
    if ( b->stream ) (void)hipStreamDestroy( b->stream );
}

// (Re)build the workgroup table and the run descriptors for the current shape.
int apply_shape( nufi_batch *b, const launch_shape &s )
{
    const size_t elem = b->variant == TRACE_FP32C ? sizeof(float) : sizeof(double);
    if ( !trace1d_shape_ok( b->variant, b->order, s ) )
        return bfail( b, NUFI_ERR_UNSUPPORTED, "batch: launch shape not built" );
    // per run: the table chunk (bounded by the register prefetch) and its LDS,
    // or the x-major f buffer of the workgroup, whichever is larger
    size_t lds = size_t( s.block ) * size_t( s.npt ) * sizeof(double);
    std::vector<uint32_t> wg_run;
    for ( uint64_t r = 0; r < b->n_runs; ++r )
    {
        batch_run &h = b->h_runs[r];
        trace1_params q = h.p;
        q.n = h.Nt;                                    // the chunk bound of the longest step
        bool pipe = false;
        h.p.chunk = trace1d_chunk( q, b->variant, s, 65536, &pipe );
        h.p.pipe = pipe ? 1u : 0u;
        lds = std::max( lds, ( pipe ? 2 : 1 ) * size_t( h.p.chunk ) * b->sq[r] * elem );
        h.wg_begin = uint32_t( wg_run.size() );
        h.n_wg = trace1d_grid( h.p, b->variant, s );
        wg_run.insert( wg_run.end(), h.n_wg, uint32_t( r ) );
    }
    if ( lds > 65536 ) return bfail( b, NUFI_ERR_UNSUPPORTED, "batch: launch shape needs more than 64 KiB of LDS" );
    if ( wg_run.size() >= ( size_t( 1 ) << 31 ) ) return bfail( b, NUFI_ERR_RANGE, "batch: too many workgroups" );
    if ( wg_run.size() > b->d_wg_cap )
    {
        if ( b->d_wg_run ) BHIP( b, hipFree( b->d_wg_run ) );
        b->d_wg_run = nullptr;
        int rc = balloc( b, &b->d_wg_run, wg_run.size(), "batch workgroup table" );
        if ( rc ) return rc;
        b->d_wg_cap = wg_run.size();
    }
    BHIP( b, hipMemcpy( b->d_wg_run, wg_run.data(), wg_run.size() * sizeof(uint32_t), hipMemcpyHostToDevice ) );
    BHIP( b, hipMemcpy( b->d_runs, b->h_runs.data(), b->n_runs * sizeof(batch_run), hipMemcpyHostToDevice ) );
    b->n_wg = uint32_t( wg_run.size() );
    b->shape = s;
    b->trace_lds = lds;
    return NUFI_OK;
}

int launch_step( nufi_batch *b, uint64_t n )
{
    hipError_t e = launch_trace1d_batch( b->d_runs, b->d_wg_run, b->n_wg, n, b->variant, b->shape, b->trace_lds,
                                         b->stream );
    if ( e != hipSuccess ) return bhip( b, e, "batch trace launch" );
    e = launch_field_batch( b->d_fields, uint32_t( b->n_runs ), n, b->field_lds, b->stream );
    if ( e != hipSuccess ) return bhip( b, e, "batch field launch" );
    return NUFI_OK;
}

} // namespace

extern "C" {

int nufi_batch_create( const nufi_config *confs, uint64_t n_runs, int order, int precision, int device,
                       nufi_batch **out )
{
    if ( out == nullptr ) return bfail( nullptr, NUFI_ERR_INVALID, "batch_create: out is NULL" );
    *out = nullptr;
    if ( confs == nullptr || n_runs == 0 ) return bfail( nullptr, NUFI_ERR_INVALID, "batch_create: no runs" );
    if ( order != 4 ) return bfail( nullptr, NUFI_ERR_UNSUPPORTED, "batch_create: order 4 only" );
    if ( precision != NUFI_FP64 && precision != NUFI_FP32C )
        return bfail( nullptr, NUFI_ERR_UNSUPPORTED, "batch_create: unknown precision" );
    for ( uint64_t r = 0; r < n_runs; ++r )
    {
        const nufi_config &c = confs[r];
        if ( c.dim != 1 || c.Nx == 0 || c.Nu == 0 || !( c.dt > 0 ) || !( c.x_max > c.x_min ) || !( c.u_max > c.u_min ) )
            return bfail( nullptr, NUFI_ERR_INVALID, "batch_create: run " + std::to_string( r ) + " is not a valid 1D config" );
        if ( !field_supported( 1, uint32_t( c.Nx ), 1, 1 ) || c.Nx > ( 1u << 20 ) )
            return bfail( nullptr, NUFI_ERR_UNSUPPORTED, "batch_create: grid too large for the field solve" );
    }
    int ndev = 0;
    if ( hipGetDeviceCount( &ndev ) != hipSuccess || ndev == 0 ) return bfail( nullptr, NUFI_ERR_DEVICE, "batch_create: no HIP device" );
    if ( device < 0 || device >= ndev ) return bfail( nullptr, NUFI_ERR_INVALID, "batch_create: bad device index" );

    nufi_batch *b = new ( std::nothrow ) nufi_batch;
    if ( b == nullptr ) return bfail( nullptr, NUFI_ERR_ALLOC, "batch_create: host allocation" );
    auto bail = [&]( int code ) { std::string m = b->err; bfree_all( b ); delete b; return bfail( nullptr, code, m ); };
    b->device = device; b->order = order; b->precision = precision;
    {
        const char *g = std::getenv( "NUFI_RUN_GRAPH" );
        b->graph_ok = !( g && g[0] == '0' );
    }
    b->variant = precision == NUFI_FP32C ? TRACE_FP32C : TRACE_FAST;
    b->confs.assign( confs, confs + n_runs );
    b->n_runs = n_runs;
    hipError_t he;
    if ( ( he = hipSetDevice( device ) ) != hipSuccess ) { bhip( b, he, "hipSetDevice" ); return bail( NUFI_ERR_DEVICE ); }
    if ( ( he = hipStreamCreateWithFlags( &b->stream, hipStreamNonBlocking ) ) != hipSuccess )
    { bhip( b, he, "hipStreamCreate" ); return bail( NUFI_ERR_DEVICE ); }

    // packed per-run layout, 256-byte aligned sub-buffers
    size_t n_hist = 0, n_dhist = 0, n_q = 0, n_part = 0, n_en = 0, n_tab = 0;
    for ( uint64_t r = 0; r < n_runs; ++r )
    {
        const nufi_config &c = confs[r];
        const uint32_t st = uint32_t( c.Nx + order - 1 );
        const uint32_t sd = uint32_t( ( c.Nx + order + 3 ) & ~uint64_t( 3 ) );
        const uint32_t wpx = xm_ppx( uint32_t( c.Nu ) );     // x-major partials per x
        const uint32_t sq = poly_stride( uint32_t( c.Nx ) );
        b->st.push_back( st ); b->sd.push_back( sd ); b->sq.push_back( sq ); b->wpx.push_back( wpx );
        b->off_hist.push_back( n_hist );  n_hist  = align_up( n_hist + ( c.Nt + 1 ) * st + 8, 32 );
        b->off_dhist.push_back( n_dhist ); n_dhist = align_up( n_dhist + ( c.Nt + 1 ) * sd + 8, 32 );
        b->off_q.push_back( n_q );        n_q     = align_up( n_q + ( c.Nt + 1 ) * sq + 8, 32 );
        b->off_part.push_back( n_part );  n_part  = align_up( n_part + c.Nx * wpx, 32 );
        b->off_en.push_back( n_en );      n_en    = align_up( n_en + c.Nt + 1, 32 );
        b->off_tab.push_back( n_tab );    n_tab   = align_up( n_tab + c.Nx, 16 );
        b->Nt_max = std::max<uint64_t>( b->Nt_max, c.Nt );
        b->max_N = std::max<uint32_t>( b->max_N, uint32_t( c.Nx ) );
        b->max_sq = std::max<uint32_t>( b->max_sq, sq );
    }
    int rc;
    if ( ( rc = balloc( b, &b->hist, n_hist, "batch history" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->dhist, n_dhist, "batch differences" ) ) ) return bail( rc );
    if ( precision == NUFI_FP32C ) { if ( ( rc = balloc( b, &b->qhist32, n_q, "batch FP32 tables" ) ) ) return bail( rc ); }
    else if ( ( rc = balloc( b, &b->qhist, n_q, "batch tables" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->partials, n_part, "batch partials" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->energies, n_en, "batch energies" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->tw, n_tab, "batch twiddles" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->mult, n_tab, "batch multipliers" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->ewt, n_tab, "batch energy weights" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->d_runs, n_runs, "batch run table" ) ) ) return bail( rc );
    if ( ( rc = balloc( b, &b->d_fields, n_runs, "batch field table" ) ) ) return bail( rc );

    int tw_lds = 0;
    b->field_lds = field_batch_lds_bytes( b->max_N, &tw_lds );
    std::vector<double2> tw( n_tab ), mult( n_tab );
    std::vector<double> ewt( n_tab );
    std::vector<batch_field> fields( n_runs );
    b->h_runs.resize( n_runs );
    for ( uint64_t r = 0; r < n_runs; ++r )
    {
        const nufi_config &c = confs[r];
        const fourier_tables t = make_tables( c, 1, order );
        std::copy( t.tx.begin(), t.tx.end(), tw.begin() + b->off_tab[r] );
        std::copy( t.ms.begin(), t.ms.end(), mult.begin() + b->off_tab[r] );
        std::copy( t.ew.begin(), t.ew.end(), ewt.begin() + b->off_tab[r] );

        trace1_params &p = b->h_runs[r].p;
        p = trace1_params {};
        p.dhist = b->dhist + b->off_dhist[r];
        p.qhist = b->qhist ? b->qhist + b->off_q[r] : nullptr;
        p.qhist32 = b->qhist32 ? b->qhist32 + b->off_q[r] : nullptr;
        p.stride_q = b->sq[r]; p.q2_off = poly_q2( uint32_t( c.Nx ) ); p.kick = -0.5 * c.dt * c.dx_inv;
        p.partials = b->partials + b->off_part[r];
        p.q_begin = 0; p.q_end = c.Nx * c.Nu;
        p.Nx = uint32_t( c.Nx ); p.Nu = uint32_t( c.Nu ); p.stride_d = b->sd[r];
        p.ix_first = 0; p.ix_count = uint32_t( c.Nx );
        p.order = uint32_t( order ); p.mode = MODE_RHO;
        trace1d_xm_setup( p );
        p.x_min = c.x_min; p.dx = c.dx; p.dx_inv = c.dx_inv; p.Lx = c.Lx; p.Lx_inv = c.Lx_inv; p.dt = c.dt;
        p.nx_min = c.x_min; p.ndx = c.dx; p.nu_min = c.u_min; p.ndu = c.du;
        fp32c_frame( double( c.Nx ), c.dt * c.dx_inv, c.u_min, double( c.Nu ), c.du, p.f32_p0, p.f32_w0 );
        p.f0 = c.f0;
        b->h_runs[r].Nt = c.Nt;

        field_params &f = fields[r].f;
        f = field_params {};
        f.dim = 1; f.order = order; f.mode = FIELD_STEP;
        f.Nx = uint32_t( c.Nx ); f.Ny = 1; f.Nz = 1; f.N = uint32_t( c.Nx );
        uint32_t l2 = 0; while ( ( uint64_t( 1 ) << l2 ) < c.Nx ) ++l2;
        f.log2x = ( ( c.Nx & ( c.Nx - 1 ) ) == 0 ) ? l2 : 0xffffffffu;
        f.log2y = 0; f.log2z = 0;
        f.partials = p.partials; f.wpx = b->wpx[r]; f.weight = -c.du; f.tw_lds = tw_lds;
        f.tw_x = b->tw + b->off_tab[r]; f.tw_y = f.tw_x; f.tw_z = f.tw_x;
        f.mult = b->mult + b->off_tab[r]; f.ewt = b->ewt + b->off_tab[r];
        f.slice = b->hist + b->off_hist[r];
        f.dslice = b->dhist + b->off_dhist[r];
        f.qslice = b->qhist ? b->qhist + b->off_q[r] : nullptr;
        f.qslice32 = b->qhist32 ? b->qhist32 + b->off_q[r] : nullptr;
        f.q2_off = poly_q2( uint32_t( c.Nx ) ); f.kick = fp32c1_scale( c.dt, c.dx_inv );
        f.energy = b->energies + b->off_en[r];
        f.sy = b->st[r]; f.sz = b->st[r]; f.stride_d = b->sd[r];
        fields[r].Nt = c.Nt; fields[r].st = b->st[r]; fields[r].sd = b->sd[r]; fields[r].sq = b->sq[r];
    }
    if ( ( he = hipMemcpy( b->tw, tw.data(), n_tab * sizeof(double2), hipMemcpyHostToDevice ) ) != hipSuccess ||
         ( he = hipMemcpy( b->mult, mult.data(), n_tab * sizeof(double2), hipMemcpyHostToDevice ) ) != hipSuccess ||
         ( he = hipMemcpy( b->ewt, ewt.data(), n_tab * sizeof(double), hipMemcpyHostToDevice ) ) != hipSuccess ||
         ( he = hipMemcpy( b->d_fields, fields.data(), n_runs * sizeof(batch_field), hipMemcpyHostToDevice ) ) != hipSuccess )
    { bhip( b, he, "batch table upload" ); return bail( NUFI_ERR_DEVICE ); }

    const size_t elem = b->variant == TRACE_FP32C ? sizeof(float) : sizeof(double);
    const int chunk = int( std::max<size_t>( 1, std::min<size_t>( 64, 65536 / ( size_t( b->max_sq ) * elem ) ) ) );
    if ( ( rc = apply_shape( b, launch_shape { 1024, 1, chunk } ) ) ) return bail( rc );
    *out = b;
    return NUFI_OK;
}

int nufi_batch_destroy( nufi_batch *b )
{
    if ( b == nullptr ) return NUFI_OK;
    (void)hipSetDevice( b->device );
    if ( b->stream ) (void)hipStreamSynchronize( b->stream );
    bfree_all( b );
    delete b;
    return NUFI_OK;
}

const char *nufi_batch_last_error( const nufi_batch *b )
{
    if ( b == nullptr ) return g_batch_error.c_str();
    return b->err.empty() ? "no error" : b->err.c_str();
}

int nufi_batch_run( nufi_batch *b, uint64_t n_begin, uint64_t n_end, int use_graph )
{
    if ( b == nullptr ) return NUFI_ERR_INVALID;
    if ( n_end < n_begin || ( n_end > 0 && n_end - 1 > b->Nt_max ) ) return bfail( b, NUFI_ERR_RANGE, "batch_run: step range" );
    if ( n_end == n_begin ) return NUFI_OK;
    BHIP( b, hipSetDevice( b->device ) );
    if ( !use_graph || !b->graph_ok )
    {
        for ( uint64_t n = n_begin; n < n_end; ++n ) { int rc = launch_step( b, n ); if ( rc ) return rc; }
        return NUFI_OK;
    }
    const bool same = b->graph && b->graph_b == n_begin && b->graph_e == n_end && b->graph_block == b->shape.block &&
                      b->graph_npt == b->shape.npt && b->graph_chunk == b->shape.chunk;
    if ( !same )
    {
        if ( b->graph ) { BHIP( b, hipGraphExecDestroy( b->graph ) ); b->graph = nullptr; }
        BHIP( b, hipStreamBeginCapture( b->stream, hipStreamCaptureModeThreadLocal ) );
        int rc = NUFI_OK;
        for ( uint64_t n = n_begin; n < n_end && rc == NUFI_OK; ++n ) rc = launch_step( b, n );
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture( b->stream, &g );
        if ( rc ) { if ( g ) (void)hipGraphDestroy( g ); return rc; }
        if ( e != hipSuccess ) return bhip( b, e, "hipStreamEndCapture" );
        e = hipGraphInstantiate( &b->graph, g, nullptr, nullptr, 0 );
        (void)hipGraphDestroy( g );
        if ( e != hipSuccess ) { b->graph = nullptr; return bhip( b, e, "hipGraphInstantiate" ); }
        b->graph_b = n_begin; b->graph_e = n_end;
        b->graph_block = b->shape.block; b->graph_npt = b->shape.npt; b->graph_chunk = b->shape.chunk;
    }
    BHIP( b, hipGraphLaunch( b->graph, b->stream ) );
    return NUFI_OK;
}

int nufi_batch_synchronize( nufi_batch *b )
{
    if ( b == nullptr ) return NUFI_ERR_INVALID;
    BHIP( b, hipSetDevice( b->device ) );
    BHIP( b, hipStreamSynchronize( b->stream ) );
    return NUFI_OK;
}

int nufi_batch_energies( nufi_batch *b, uint64_t run, uint64_t n_begin, uint64_t n_end, double *out )
{
    if ( b == nullptr || out == nullptr ) return NUFI_ERR_INVALID;
    if ( run >= b->n_runs || n_end < n_begin || n_end > b->confs[run].Nt + 1 ) return bfail( b, NUFI_ERR_RANGE, "batch_energies: range" );
    BHIP( b, hipSetDevice( b->device ) );
    BHIP( b, hipMemcpyAsync( out, b->energies + b->off_en[run] + n_begin, ( n_end - n_begin ) * sizeof(double),
                             hipMemcpyDeviceToHost, b->stream ) );
    BHIP( b, hipStreamSynchronize( b->stream ) );
    return NUFI_OK;
}

int nufi_batch_history( nufi_batch *b, uint64_t run, uint64_t n_begin, uint64_t n_end, double *out )
{
    if ( b == nullptr || out == nullptr ) return NUFI_ERR_INVALID;
    if ( run >= b->n_runs || n_end < n_begin || n_end > b->confs[run].Nt + 1 ) return bfail( b, NUFI_ERR_RANGE, "batch_history: range" );
    BHIP( b, hipSetDevice( b->device ) );
    BHIP( b, hipMemcpyAsync( out, b->hist + b->off_hist[run] + n_begin * b->st[run],
                             ( n_end - n_begin ) * b->st[run] * sizeof(double), hipMemcpyDeviceToHost, b->stream ) );
    BHIP( b, hipStreamSynchronize( b->stream ) );
    return NUFI_OK;
}

int nufi_batch_set_launch_shape( nufi_batch *b, int block, int npt, int chunk )
{
    if ( b == nullptr ) return NUFI_ERR_INVALID;
    BHIP( b, hipSetDevice( b->device ) );
    BHIP( b, hipStreamSynchronize( b->stream ) );
    return apply_shape( b, launch_shape { block, npt, chunk } );
}

int nufi_batch_autotune( nufi_batch *b, uint64_t n_probe, int repeats, int *block, int *npt, int *chunk,
                         double *best_ms )
{
    if ( b == nullptr ) return NUFI_ERR_INVALID;
    if ( n_probe > b->Nt_max ) return bfail( b, NUFI_ERR_RANGE, "batch_autotune: probe step out of range" );
    if ( repeats < 1 ) repeats = 1;
    BHIP( b, hipSetDevice( b->device ) );
    BHIP( b, hipStreamSynchronize( b->stream ) );
    const size_t elem = b->variant == TRACE_FP32C ? sizeof(float) : sizeof(double);
    const int cmax = int( std::max<size_t>( 1, std::min<size_t>( 64, 65536 / ( size_t( b->max_sq ) * elem ) ) ) );
    const launch_shape saved = b->shape;
    hipEvent_t e0, e1;
    BHIP( b, hipEventCreate( &e0 ) );
    BHIP( b, hipEventCreate( &e1 ) );
    double best = 1e300;
    launch_shape best_s = saved;
    int rc = NUFI_OK;
    for ( int bl : { 256, 512, 1024 } )
    for ( int np : { 1, 2, 4 } )
    for ( int ch : { std::max( 1, cmax / 4 ), std::max( 1, cmax / 2 ), cmax } )
    {
        if ( ( rc = apply_shape( b, launch_shape { bl, np, ch } ) ) ) { rc = NUFI_OK; continue; }
        hipError_t he = launch_trace1d_batch( b->d_runs, b->d_wg_run, b->n_wg, n_probe, b->variant, b->shape,
                                             b->trace_lds, b->stream );
        if ( he == hipSuccess ) he = hipEventRecord( e0, b->stream );
        for ( int r = 0; r < repeats && he == hipSuccess; ++r )
            he = launch_trace1d_batch( b->d_runs, b->d_wg_run, b->n_wg, n_probe, b->variant, b->shape,
                                       b->trace_lds, b->stream );
        if ( he == hipSuccess ) he = hipEventRecord( e1, b->stream );
        if ( he == hipSuccess ) he = hipEventSynchronize( e1 );
        float ms = 0;
        if ( he == hipSuccess ) he = hipEventElapsedTime( &ms, e0, e1 );
        if ( he != hipSuccess ) { rc = bhip( b, he, "batch autotune" ); break; }
        if ( ms / repeats < best ) { best = ms / repeats; best_s = b->shape; }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    const int rc2 = apply_shape( b, rc == NUFI_OK ? best_s : saved );
    if ( rc ) return rc;
    if ( rc2 ) return rc2;
    if ( block ) *block = b->shape.block;
    if ( npt ) *npt = b->shape.npt;
    if ( chunk ) *chunk = b->shape.chunk;
    if ( best_ms ) *best_ms = best;
    return NUFI_OK;
}

int nufi_batch_get_launch_shape( nufi_batch *b, int *block, int *npt, int *chunk )
{
    if ( b == nullptr ) return NUFI_ERR_INVALID;
    if ( block ) *block = b->shape.block;
    if ( npt ) *npt = b->shape.npt;
    if ( chunk ) *chunk = b->shape.chunk;
    return NUFI_OK;
}

} // extern "C"
