/*
 * oracle_capi.cpp — TEST INFRASTRUCTURE: extern "C" entry points of the CPU
 * oracle (nufi_oracle.hpp) for the ctypes checkers in tests/ and for the
 * cpu_baseline leg of bench.py.  Built by oracle/Makefile into
 * oracle/_build/liboracle.so with the reference's own flags
 * (configure.ac:13-15: C++17, -O2, OpenMP; no -march, no FP contraction).
 *
 * History arrays passed in hold slices m = 0.. at hist + m*stride_t with
 * stride_t = (Nx+order-1)^dim and must extend `order` reals past the last
 * slice read (the t = 0 edge read of fields.hpp:49-58 may touch it).
 */
#include "nufi_oracle.hpp"
#include "dual.hpp"
#include "opcount.hpp"

#include <omp.h>
#include <cstdio>

using namespace oracle;

namespace
{

template <template <int> class F, typename... Args>
int dispatch_order( int order, Args&&... args )
{
    switch ( order )
    {
    case 3: return F<3>::run( args... );
    case 4: return F<4>::run( args... );
    case 5: return F<5>::run( args... );
    case 6: return F<6>::run( args... );
    case 7: return F<7>::run( args... );
    case 8: return F<8>::run( args... );
    default: return NUFI_ERR_UNSUPPORTED;
    }
}

size_t stride1( const nufi_config *c, int order ) { return c->Nx + order - 1; }
size_t stride3( const nufi_config *c, int order )
{
    return (c->Nx + order - 1)*(c->Ny + order - 1)*(c->Nz + order - 1);
}

/* ---- basis / de Boor --------------------------------------------------- */
template <int order>
struct basis_op
{
    static int run( int deriv, double t, double *out )
    {
        switch ( deriv )
        {
        case 0: basis<double,order,0>( t, out ); return 0;
        case 1: basis<double,order,1>( t, out ); return 0;
        case 2: basis<double,order,2>( t, out ); return 0;
        default: return NUFI_ERR_UNSUPPORTED;
        }
    }
};
template <int order>
struct deboor_op
{
    static int run( int deriv, double t, const double *c, double *out )
    {
        switch ( deriv )
        {
        case 0: *out = deboor<double,order,0>( t, c ); return 0;
        case 1: *out = deboor<double,order,1>( t, c ); return 0;
        case 2: *out = deboor<double,order,2>( t, c ); return 0;
        default: return NUFI_ERR_UNSUPPORTED;
        }
    }
};

/* ---- fields ------------------------------------------------------------ */
template <int order>
struct field1d_op
{
    static int run( const nufi_config *c, int deriv, double x, const double *slice, double *out )
    {
        grid1<double> g( *c );
        switch ( deriv )
        {
        case 0: *out = field1d<double,order,0>( x, slice, g ); return 0;
        case 1: *out = field1d<double,order,1>( x, slice, g ); return 0;
        case 2: *out = field1d<double,order,2>( x, slice, g ); return 0;   // rho = -phi'' plots
        default: return NUFI_ERR_UNSUPPORTED;
        }
    }
};
template <int order>
struct field3d_op
{
    static int run( const nufi_config *c, int dx, int dy, int dz, double x, double y, double z,
                    const double *slice, double *out )
    {
        grid3<double> g( *c );
        const int code = dx*100 + dy*10 + dz;
        switch ( code )
        {
        case   0: *out = field3d<double,order,0,0,0>( x, y, z, slice, g ); return 0;
        case 100: *out = field3d<double,order,1,0,0>( x, y, z, slice, g ); return 0;
        case  10: *out = field3d<double,order,0,1,0>( x, y, z, slice, g ); return 0;
        case   1: *out = field3d<double,order,0,0,1>( x, y, z, slice, g ); return 0;
        default: return NUFI_ERR_UNSUPPORTED;
        }
    }
};

/* ---- trace ------------------------------------------------------------- */
template <int order>
struct trace1d_op
{
    static int run( const nufi_config *c, int with_half_kick, uint64_t n, double x, double u,
                    const double *hist, double *out )
    {
        grid1<double> g( *c );
        *out = with_half_kick ? f1d<double,order>( n, x, u, hist, g )
                              : ftilda1d<double,order>( n, x, u, hist, g );
        return 0;
    }
};
template <int order>
struct phase_flow1d_op
{
    static int run( const nufi_config *c, uint64_t n, double x, double u, const double *hist,
                    double *x_out, double *u_out )
    {
        grid1<double> g( *c );
        phase_flow1d<double,order>( n, x, u, hist, g );
        *x_out = x; *u_out = u;
        return 0;
    }
};
template <int order>
struct trace3d_op
{
    static int run( const nufi_config *c, int with_half_kick, uint64_t n, const double *xv,
                    const double *hist, double *out )
    {
        grid3<double> g( *c );
        *out = with_half_kick ? f3d<double,order>( n, xv[0], xv[1], xv[2], xv[3], xv[4], xv[5], hist, g )
                              : ftilda3d<double,order>( n, xv[0], xv[1], xv[2], xv[3], xv[4], xv[5], hist, g );
        return 0;
    }
};

/* ---- rho over the whole grid (OpenMP over x only, as
 *      bin/test_nufi_cpu_1d.cpp:65-70 / test_nufi_cpu_3d.cpp:68-72) ------ */
template <int order>
struct rho1d_op
{
    static int run( const nufi_config *c, uint64_t n, const double *hist, double *rho, int nthreads,
                    int exact_sum = 0 )
    {
        grid1<double> g( *c );
        const long long N = (long long)g.Nx;
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for ( long long i = 0; i < N; ++i )
            rho[i] = exact_sum ? rho1d<double,order,long double>( n, (size_t)i, hist, g )
                               : rho1d<double,order>( n, (size_t)i, hist, g );
        return 0;
    }
};
template <int order>
struct rho3d_op
{
    static int run( const nufi_config *c, uint64_t n, const double *hist, double *rho, int nthreads,
                    int exact_sum = 0 )
    {
        grid3<double> g( *c );
        const long long N = (long long)( g.Nx*g.Ny*g.Nz );
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for ( long long l = 0; l < N; ++l )
            rho[l] = exact_sum ? rho3d<double,order,long double>( n, (size_t)l, hist, g )
                               : rho3d<double,order>( n, (size_t)l, hist, g );
        return 0;
    }
};

/* rho at a list of spatial points (the x loop of rho3d_op restricted to
 * them, OpenMP over the points as bin/test_nufi_cpu_3d.cpp:68-72 is over x):
 * isolated steps on a full-size 3D grid at a sample of points. */
template <int order>
struct rho3d_points_op
{
    static int run( const nufi_config *c, uint64_t n, const double *hist, const uint64_t *points,
                    uint64_t count, double *rho, int nthreads, int exact_sum )
    {
        grid3<double> g( *c );
        const uint64_t N = g.Nx*g.Ny*g.Nz;
        for ( uint64_t i = 0; i < count; ++i ) if ( points[i] >= N ) return NUFI_ERR_RANGE;
        #pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for ( long long i = 0; i < (long long)count; ++i )
            rho[i] = exact_sum ? rho3d<double,order,long double>( n, (size_t)points[i], hist, g )
                               : rho3d<double,order>( n, (size_t)points[i], hist, g );
        return 0;
    }
};

/* ---- fit ---------------------------------------------------------------- */
template <int order>
struct fit1d_op
{
    static int run( const nufi_config *c, double *slice, const double *values )
    {
        fit1d<order>( grid1<double>( *c ), slice, values ); return 0;
    }
};
template <int order>
struct fit3d_op
{
    static int run( const nufi_config *c, double *slice, const double *values )
    {
        fit3d<order>( grid3<double>( *c ), slice, values ); return 0;
    }
};

/* ---- one step / whole run (test_nufi_cpu_1d.cpp:60-75) ----------------- */
template <int order>
struct step1d_op
{
    static int run( const nufi_config *c, uint64_t n, double *hist, double *energy, double *rho_out,
                    int nthreads )
    {
        grid1<double> g( *c );
        std::vector<double> rho( g.Nx );
        rho1d_op<order>::run( c, n, hist, rho.data(), nthreads );
        if ( rho_out ) std::memcpy( rho_out, rho.data(), sizeof(double)*g.Nx );
        const double e = poisson1d( g, rho.data() );
        if ( energy ) *energy = e;
        fit1d<order>( g, hist + n*stride1( c, order ), rho.data() );
        return 0;
    }
};
template <int order>
struct step3d_op
{
    static int run( const nufi_config *c, uint64_t n, double *hist, double *energy, double *rho_out,
                    int nthreads )
    {
        grid3<double> g( *c );
        const size_t N = g.Nx*g.Ny*g.Nz;
        std::vector<double> rho( N );
        rho3d_op<order>::run( c, n, hist, rho.data(), nthreads );
        if ( rho_out ) std::memcpy( rho_out, rho.data(), sizeof(double)*N );
        const double e = poisson3d( g, rho.data() );
        if ( energy ) *energy = e;
        fit3d<order>( g, hist + n*stride3( c, order ), rho.data() );
        return 0;
    }
};

/* ---- metrics (cuda_kernel.cu:53-79, 428-466) --------------------------- */
template <int order>
struct metrics1d_op
{
    static int run( const nufi_config *c, const nufi_config *cm, uint64_t n, const double *hist,
                    double *metrics, int nthreads )
    {
        grid1<double> g( *c ), gm( *cm );
        const long long Nq = (long long)( gm.Nx*gm.Nu );
        double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        #pragma omp parallel for num_threads(nthreads) reduction(+:m0,m1,m2,m3) schedule(static)
        for ( long long q = 0; q < Nq; ++q )
        {
            const size_t ix = (size_t)q / gm.Nu, iu = (size_t)q % gm.Nu;
            const double x = gm.x_min + ix*gm.dx;
            const double u = gm.u_min + iu*gm.du + gm.du/2;
            const double f = f1d<double,order>( n, x, u, hist, g );
            const double w = gm.du*gm.dx;
            m0 += w*f; m1 += w*f*f; m2 += w*(u*u*f/2);
            m3 += ( f > 0 ) ? -w*f*std::log(f) : 0;
        }
        metrics[0] = m0; metrics[1] = m1; metrics[2] = m2; metrics[3] = m3;
        return 0;
    }
};
template <int order>
struct metrics3d_op
{
    static int run( const nufi_config *c, const nufi_config *, uint64_t n, const double *hist,
                    double *metrics, int nthreads )
    {
        grid3<double> g( *c );
        const size_t Nvel = g.Nu*g.Nv*g.Nw;
        const long long Nq = (long long)( g.Nx*g.Ny*g.Nz*Nvel );
        double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        #pragma omp parallel for num_threads(nthreads) reduction(+:m0,m1,m2,m3) schedule(static)
        for ( long long q = 0; q < Nq; ++q )
        {
            size_t t = (size_t)q;
            const size_t iz = t / ( g.Ny*g.Nx*Nvel ); t %= g.Ny*g.Nx*Nvel;
            const size_t iy = t / ( g.Nx*Nvel );      t %= g.Nx*Nvel;
            const size_t ix = t / Nvel;               t %= Nvel;
            const size_t iw = t / ( g.Nv*g.Nu );      t %= g.Nv*g.Nu;
            const size_t iv = t / g.Nu, iu = t % g.Nu;
            const double x = g.x_min + ix*g.dx, y = g.y_min + iy*g.dy, z = g.z_min + iz*g.dz;
            const double u = g.u_min + iu*g.du + g.du/2;
            const double v = g.v_min + iv*g.dv + g.dv/2;
            const double w = g.w_min + iw*g.dw + g.dw/2;
            const double f = f3d<double,order>( n, x, y, z, u, v, w, hist, g );
            const double wt = g.du*g.dv*g.dw;
            m0 += wt*f; m1 += wt*f*f; m2 += wt*(u*u+v*v+w*w)*f/2;
            m3 += ( f > 0 ) ? -wt*f*std::log(f) : 0;
        }
        metrics[0] = m0; metrics[1] = m1; metrics[2] = m2; metrics[3] = m3;
        return 0;
    }
};

/* ---- forward-mode tangent of the 1D time loop --------------------------
 * J = sum_{n} dt * energy_n and dJ/d(parameter) for up to NUFI_MAX_MODES
 * directions (multimode a_m; alpha for Landau / two-stream), by dual numbers
 * through the trace and componentwise (linear) Poisson solve and fit. */
using D8 = dual<NUFI_MAX_MODES>;

/* scaled DHT coefficients d_i = H_i / (N kappa_i^2) of one real array (poisson.cpp:68-82) */
std::vector<double> scaled_spectrum( const grid1<double> &g, std::vector<double> v )
{
    std::vector<long double> tmp;
    cas_table T( g.Nx );
    dht_line( T, v.data(), 1, tmp );
    const double fac_N = double(1) / double( g.Nx ), fac_x = ( 2*M_PI*g.Lx_inv );
    for ( size_t i = 1; i < g.Nx; ++i )
    {
        const double ii = ( 2*i < g.Nx ) ? double(i) : double(g.Nx - i);
        v[i] *= fac_N / ( ii*ii*fac_x*fac_x );
    }
    v[0] = 0;
    return v;
}

int tangent_run1d( const nufi_config *c, uint64_t n_end, double *J, double *dJ, int nthreads )
{
    constexpr int K = NUFI_MAX_MODES;
    constexpr int order = 4;
    grid1<double> g( *c );
    grid1<D8> gd( *c );
    int npar = 0;
    if ( c->f0.kind == NUFI_F0_MULTIMODE_1D )
    { npar = c->f0.n_modes; for ( int m = 0; m < npar; ++m ) gd.a[m].d[m] = 1.0; }
    else if ( c->f0.kind == NUFI_F0_LANDAU_1D || c->f0.kind == NUFI_F0_TWOSTREAM_1D )
    { npar = 1; gd.alpha.d[0] = 1.0; }
    else return NUFI_ERR_UNSUPPORTED;

    const size_t st = g.Nx + order - 1, N = g.Nx;
    std::vector<D8> hist( ( n_end + 1 ) * st + order + 1 );
    std::vector<D8> rho( N );
    double Jv = 0, dJv[K] = {};
    for ( uint64_t n = 0; n < n_end; ++n )
    {
        const long long NN = (long long)N;
        #pragma omp parallel for num_threads(nthreads) schedule(static)
        for ( long long i = 0; i < NN; ++i ) rho[i] = rho1d<D8,order>( n, (size_t)i, hist.data(), gd );
        // value: exactly the oracle's step (poisson1d + fit1d)
        std::vector<double> v( N );
        for ( size_t i = 0; i < N; ++i ) v[i] = rho[i].v;
        std::vector<double> phi = v;
        const double e = poisson1d( g, phi.data() );
        std::vector<double> sl( st );
        fit1d<order>( g, sl.data(), phi.data() );
        for ( size_t i = 0; i < st; ++i ) hist[n*st + i] = D8( sl[i] );
        // tangents: the same linear maps per component, energy by the product rule
        const std::vector<double> d0 = scaled_spectrum( g, v );
        const double fac_x = ( 2*M_PI*g.Lx_inv );
        for ( int k = 0; k < npar; ++k )
        {
            std::vector<double> vk( N );
            for ( size_t i = 0; i < N; ++i ) vk[i] = rho[i].d[k];
            const std::vector<double> dk = scaled_spectrum( g, vk );
            double de = 0;
            for ( size_t i = 1; i < N; ++i )
            {
                const double ii = ( 2*i < N ) ? double(i) : double(N - i);
                de += 2.0 * ( ii*fac_x*d0[i] ) * ( ii*fac_x*dk[i] );
            }
            de *= g.Lx/2;
            dJv[k] += g.dt * de;
            std::vector<double> phik = vk;
            poisson1d( g, phik.data() );
            std::vector<double> slk( st );
            fit1d<order>( g, slk.data(), phik.data() );
            for ( size_t i = 0; i < st; ++i ) hist[n*st + i].d[k] = slk[i];
        }
        Jv += g.dt * e;
    }
    *J = Jv;
    for ( int k = 0; k < npar; ++k ) dJ[k] = dJv[k];
    return 0;
}

bool bad( const nufi_config *c, int dim )
{
    return c == nullptr || c->dim != dim || c->Nx == 0 || c->Nu == 0 ||
           ( dim == 3 && ( c->Ny == 0 || c->Nz == 0 || c->Nv == 0 || c->Nw == 0 ) );
}

} // namespace

extern "C" {

int oracle_basis( int order, int deriv, double t, double *out )
{ return dispatch_order<basis_op>( order, deriv, t, out ); }

int oracle_deboor( int order, int deriv, double t, const double *c, double *out )
{ return dispatch_order<deboor_op>( order, deriv, t, c, out ); }

int oracle_field1d( const nufi_config *c, int order, int deriv, double x, const double *slice, double *out )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<field1d_op>( order, c, deriv, x, slice, out ); }

int oracle_field3d( const nufi_config *c, int order, int dx, int dy, int dz,
                    double x, double y, double z, const double *slice, double *out )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<field3d_op>( order, c, dx, dy, dz, x, y, z, slice, out ); }

int oracle_trace1d( const nufi_config *c, int order, int with_half_kick, uint64_t n, double x, double u,
                    const double *hist, double *out )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<trace1d_op>( order, c, with_half_kick, n, x, u, hist, out ); }

int oracle_phase_flow1d( const nufi_config *c, int order, uint64_t n, double x, double u, const double *hist,
                         double *x_out, double *u_out )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<phase_flow1d_op>( order, c, n, x, u, hist, x_out, u_out ); }

int oracle_trace3d( const nufi_config *c, int order, int with_half_kick, uint64_t n, const double *xv,
                    const double *hist, double *out )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<trace3d_op>( order, c, with_half_kick, n, xv, hist, out ); }

int oracle_rho1d( const nufi_config *c, int order, uint64_t n, const double *hist, double *rho, int nthreads )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<rho1d_op>( order, c, n, hist, rho, nthreads ); }

int oracle_rho3d( const nufi_config *c, int order, uint64_t n, const double *hist, double *rho, int nthreads )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<rho3d_op>( order, c, n, hist, rho, nthreads ); }

/* rho with the v-sum accumulated in long double: the reference's arithmetic
 * per node, without the rounding of its sequential FP64 sum (rho.hpp:140-143). */
int oracle_rho1d_exact_sum( const nufi_config *c, int order, uint64_t n, const double *hist, double *rho,
                            int nthreads )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<rho1d_op>( order, c, n, hist, rho, nthreads, 1 ); }

int oracle_rho3d_exact_sum( const nufi_config *c, int order, uint64_t n, const double *hist, double *rho,
                            int nthreads )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<rho3d_op>( order, c, n, hist, rho, nthreads, 1 ); }

int oracle_rho3d_points( const nufi_config *c, int order, uint64_t n, const double *hist, const uint64_t *points,
                         uint64_t count, double *rho, int nthreads, int exact_sum )
{
    if ( bad( c, 3 ) || ( count && ( points == nullptr || rho == nullptr ) ) ) return NUFI_ERR_INVALID;
    return dispatch_order<rho3d_points_op>( order, c, n, hist, points, count, rho, nthreads, exact_sum );
}

int oracle_poisson1d( const nufi_config *c, double *data, double *energy )
{
    if ( bad( c, 1 ) ) return NUFI_ERR_INVALID;
    const double e = poisson1d( grid1<double>( *c ), data );
    if ( energy ) *energy = e;
    return 0;
}

int oracle_poisson3d( const nufi_config *c, double *data, double *energy )
{
    if ( bad( c, 3 ) ) return NUFI_ERR_INVALID;
    const double e = poisson3d( grid3<double>( *c ), data );
    if ( energy ) *energy = e;
    return 0;
}

int oracle_fit1d( const nufi_config *c, int order, double *slice, const double *values )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<fit1d_op>( order, c, slice, values ); }

int oracle_fit3d( const nufi_config *c, int order, double *slice, const double *values )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<fit3d_op>( order, c, slice, values ); }

int oracle_step1d( const nufi_config *c, int order, uint64_t n, double *hist, double *energy,
                   double *rho_out, int nthreads )
{ if ( bad( c, 1 ) ) return NUFI_ERR_INVALID; return dispatch_order<step1d_op>( order, c, n, hist, energy, rho_out, nthreads ); }

int oracle_step3d( const nufi_config *c, int order, uint64_t n, double *hist, double *energy,
                   double *rho_out, int nthreads )
{ if ( bad( c, 3 ) ) return NUFI_ERR_INVALID; return dispatch_order<step3d_op>( order, c, n, hist, energy, rho_out, nthreads ); }

/* Steps n_begin .. n_end-1 of the time loop into hist / energies[n - n_begin]. */
int oracle_run1d( const nufi_config *c, int order, uint64_t n_begin, uint64_t n_end, double *hist,
                  double *energies, int nthreads )
{
    if ( bad( c, 1 ) ) return NUFI_ERR_INVALID;
    for ( uint64_t n = n_begin; n < n_end; ++n )
    {
        const int rc = oracle_step1d( c, order, n, hist, energies ? energies + (n - n_begin) : nullptr,
                                      nullptr, nthreads );
        if ( rc ) return rc;
    }
    return 0;
}

int oracle_run3d( const nufi_config *c, int order, uint64_t n_begin, uint64_t n_end, double *hist,
                  double *energies, int nthreads )
{
    if ( bad( c, 3 ) ) return NUFI_ERR_INVALID;
    for ( uint64_t n = n_begin; n < n_end; ++n )
    {
        const int rc = oracle_step3d( c, order, n, hist, energies ? energies + (n - n_begin) : nullptr,
                                      nullptr, nthreads );
        if ( rc ) return rc;
    }
    return 0;
}

int oracle_metrics1d( const nufi_config *c, const nufi_config *cm, int order, uint64_t n,
                      const double *hist, double *metrics, int nthreads )
{
    if ( bad( c, 1 ) ) return NUFI_ERR_INVALID;
    if ( cm == nullptr ) cm = c;
    return dispatch_order<metrics1d_op>( order, c, cm, n, hist, metrics, nthreads );
}

int oracle_metrics3d( const nufi_config *c, int order, uint64_t n, const double *hist, double *metrics,
                      int nthreads )
{
    if ( bad( c, 3 ) ) return NUFI_ERR_INVALID;
    return dispatch_order<metrics3d_op>( order, c, c, n, hist, metrics, nthreads );
}

int oracle_max_threads( void ) { return omp_get_max_threads(); }

/* Arithmetic of one characteristic step (drift, E, kick) in the reference's
 * formulation, order 4, counted on opcount (opcount.hpp). */
int oracle_count_step_ops( const nufi_config *c, long *ops )
{
    if ( c == nullptr || ops == nullptr ) return NUFI_ERR_INVALID;
    if ( c->dim == 1 )
    {
        const grid1<opcount> g( *c );
        std::vector<opcount> slice( c->Nx + 3 );
        for ( size_t i = 0; i < slice.size(); ++i ) slice[i] = opcount::variable( 0.01 * double( i % 7 ) );
        opcount x = opcount::variable( c->x_min + 0.37 * c->Lx ), u = opcount::variable( 0.7 );
        opcount::n = 0;
        x = x - g.dt*u;
        const opcount E = -field1d<opcount,4,1>( x, slice.data(), g );
        u = u + g.dt*E;
        *ops = opcount::n;
        return 0;
    }
    if ( c->dim == 3 )
    {
        const grid3<opcount> g( *c );
        std::vector<opcount> slice( ( c->Nx + 3 ) * ( c->Ny + 3 ) * ( c->Nz + 3 ) );
        for ( size_t i = 0; i < slice.size(); ++i ) slice[i] = opcount::variable( 0.01 * double( i % 7 ) );
        opcount x = opcount::variable( c->x_min + 0.37 * c->Lx ), y = opcount::variable( c->y_min + 0.61 * c->Ly ),
                z = opcount::variable( c->z_min + 0.23 * c->Lz );
        opcount u = opcount::variable( 0.7 ), v = opcount::variable( -0.4 ), w = opcount::variable( 1.1 );
        opcount Ex, Ey, Ez;
        opcount::n = 0;
        x = x - g.dt*u; y = y - g.dt*v; z = z - g.dt*w;
        efield3d<opcount,4>( x, y, z, slice.data(), g, Ex, Ey, Ez );
        u = u + g.dt*Ex; v = v + g.dt*Ey; w = w + g.dt*Ez;
        *ops = opcount::n;
        return 0;
    }
    return NUFI_ERR_UNSUPPORTED;
}

/* J = sum_{n<n_end} dt*energy_n and its forward-mode tangent dJ/d(parameters). */
int oracle_tangent_run1d( const nufi_config *c, uint64_t n_end, double *J, double *dJ, int nthreads )
{
    if ( bad( c, 1 ) || J == nullptr || dJ == nullptr ) return NUFI_ERR_INVALID;
    return tangent_run1d( c, n_end, J, dJ, nthreads );
}

} // extern "C"
