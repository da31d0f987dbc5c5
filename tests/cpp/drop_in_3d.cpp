// The reference 3D driver's step loop (bin/test_nufi_gpu_3d.cpp:150-162)
// written against the drop-in nufi::dim3::gpu_scheduler: compute_rho ->
// download_rho -> Poisson -> fit -> upload_phi with a host round trip per
// step, checked against the device-resident loop (nufi_run) of a second
// context, plus the metrics and range_error contracts.
// Usage: drop_in_3d [float]   (exit 0 on agreement)
#include <nufi/gpu_scheduler.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

template <typename real>
int run( double tol )
{
    nufi::dim3::config_t<real> conf;                 // 3D Landau on a small grid
    conf.Nx = conf.Ny = conf.Nz = 8;
    conf.Nu = conf.Nv = conf.Nw = 16;
    conf.x_min = conf.y_min = conf.z_min = 0;
    conf.x_max = conf.y_max = conf.z_max = real( 4*M_PI );
    conf.u_min = conf.v_min = conf.w_min = -6;
    conf.u_max = conf.v_max = conf.w_max = 6;
    conf.dt = real( 1./8. ); conf.Nt = 12;
    conf.derive();
    conf.f0_spec.kind = NUFI_F0_LANDAU_3D; conf.f0_spec.alpha = 0.01; conf.f0_spec.k = 0.5;
    const size_t order = 4, N = conf.Nx*conf.Ny*conf.Nz;

    nufi::dim3::gpu_scheduler<real,order> sched { conf };
    const size_t st = sched.stride_t();
    std::vector<real> coeffs( ( conf.Nt + 1 ) * st, real( 0 ) );
    std::vector<real> rho( N );
    std::vector<double> e_loop( conf.Nt + 1 ), buf( N ), slice( st );
    for ( size_t n = 0; n <= conf.Nt; ++n )
    {
        std::memset( rho.data(), 0, sizeof(real)*N );
        sched.compute_rho( n, 0, N*conf.Nu*conf.Nv*conf.Nw );
        sched.download_rho( rho.data() );
        for ( size_t i = 0; i < N; ++i ) buf[i] = 1.0 + double( rho[i] );       // rho = 1 - dv sum f
        nufi::detail::check( nufi_poisson_solve( sched.handle(), buf.data(), &e_loop[n] ), sched.handle() );
        nufi::detail::check( nufi_interpolate( sched.handle(), slice.data(), buf.data() ), sched.handle() );
        for ( size_t i = 0; i < st; ++i ) coeffs[n*st + i] = real( slice[i] );
        sched.upload_phi( n, coeffs.data() );
    }
    real metrics[4] = { 0, 0, 0, 0 };
    sched.compute_metrics( conf.Nt, 0, N*conf.Nu*conf.Nv*conf.Nw );
    sched.download_metrics( metrics );

    nufi::dim3::gpu_scheduler<real,order> dev { conf };
    std::vector<double> e_dev( conf.Nt + 1 );
    nufi::detail::check( nufi_run( dev.handle(), 0, conf.Nt + 1, e_dev.data() ), dev.handle() );

    double worst = 0;
    for ( size_t n = 0; n <= conf.Nt; ++n )
        worst = std::fmax( worst, std::fabs( e_loop[n] - e_dev[n] ) / e_dev[n] );

    bool range_ok = false, upload_range_ok = false;
    try { sched.compute_rho( conf.Nt + 1, 0, 1 ); }
    catch ( const std::range_error & ) { range_ok = true; }
    try { sched.upload_phi( conf.Nt + 1, coeffs.data() ); }
    catch ( const std::range_error & ) { upload_range_ok = true; }

    // mass: the dim3 metrics weight du dv dw (cuda_kernel.cu:457) gives Nx Ny Nz times the velocity integral
    const double mass_rel = std::fabs( double( metrics[0] ) / double( N ) - 1.0 );
    std::printf( "drop_in_3d<%s>: steps 0..%zu, max energy rel diff host-loop vs device-loop %.3e, mass %.6f "
                 "(rel %.2e), range_error %s/%s\n", sizeof(real) == 8 ? "double" : "float", conf.Nt, worst,
                 double( metrics[0] ), mass_rel, range_ok ? "ok" : "MISSING", upload_range_ok ? "ok" : "MISSING" );
    return ( worst <= tol && range_ok && upload_range_ok && mass_rel < 1e-3 ) ? 0 : 1;
}

int main( int argc, char **argv )
{
    // float: rho and the coefficients cross the boundary as float each step
    // (the reference's float path), ~1e-6 of the perturbation per step
    if ( argc > 1 && std::strcmp( argv[1], "float" ) == 0 ) return run<float>( 1e-3 );
    return run<double>( 1e-12 );
}
