/*
 * nufi/gpu_scheduler.hpp — header-only C++ drop-in for the reference's GPU rho
 * scheduler, on top of the C ABI in nufi_hip.h (link with -lnufi_hip).
 *
 * Replaces nufi/cuda_scheduler.hpp (dim1: :33-164, dim3: :287-392) and the
 * config structs of nufi/config.hpp (:34-84, :168-247) for code that drives
 * the GPU rho path:
 *
 *     nufi::dim1::config_t<double> conf;                 // same fields and defaults
 *     nufi::dim1::gpu_scheduler<double,4> sched { conf };
 *     sched.compute_rho( n, 0, conf.Nx*conf.Nu );        // async
 *     sched.download_rho( rho );                         // sync, rho[i] += ...
 *     sched.upload_phi( n, coeffs );                     // slice n of the host history
 *
 * Errors map back to the reference's exception types: NUFI_ERR_RANGE ->
 * std::range_error (cuda_kernel.cu:115-116, 502-505), everything else ->
 * nufi::gpu_error (a std::runtime_error, as cuda::exception is).
 *
 * real = double runs the FP64 method; real = float runs the FP32-compensated
 * trace (the reference's float instantiation runs FP32 throughout); host data
 * crosses the boundary as the caller's real type.
 *
 * Differences from the reference class: one device per scheduler (device
 * index argument, default 0) instead of every visible device; f0 is a runtime
 * family (config_t::f0_spec) instead of a static member function.
 */
#ifndef NUFI_GPU_SCHEDULER_HPP
#define NUFI_GPU_SCHEDULER_HPP

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../nufi_hip.h"

namespace nufi
{

class gpu_error : public std::runtime_error
{
public:
    gpu_error( int code, const std::string &msg ): std::runtime_error( msg ), code_ { code } {}
    int code() const noexcept { return code_; }
private:
    int code_;
};

namespace detail
{
inline void check( int rc, const nufi_ctx *ctx )
{
    if ( rc == NUFI_OK ) return;
    const std::string msg = nufi_last_error( ctx );
    if ( rc == NUFI_ERR_RANGE ) throw std::range_error( msg );
    throw gpu_error( rc, msg );
}
}

namespace dim1
{

// Fields and defaults of dim1::config_t (config.hpp:34-70).
template <typename real>
struct config_t
{
    size_t Nx, Nu, Nt;
    real   dt;
    real   x_min, x_max;
    real   u_min, u_max;
    real   dx, dx_inv, Lx, Lx_inv;
    real   du;
    nufi_f0 f0_spec;               // runtime initial condition (config.hpp:72-84 hard-codes it)

    config_t() noexcept
    {
        Nx = 256; Nu = 512;
        u_min = -10; u_max = 10;
        x_min = 0; x_max = 4*M_PI;
        dt = 1./16.; Nt = 100/dt;
        derive();
        f0_spec = nufi_f0 {};
        f0_spec.kind = NUFI_F0_TWOSTREAM_1D; f0_spec.alpha = 0.01; f0_spec.k = 0.5;   // config.hpp:80-82
    }

    // Recompute the dependent fields (config.hpp:67-69) after editing the primary ones.
    void derive() noexcept
    {
        Lx = x_max - x_min; Lx_inv = 1/Lx;
        dx = Lx/Nx; dx_inv = 1/dx;
        du = (u_max - u_min)/Nu;
    }

    nufi_config c_config() const noexcept
    {
        nufi_config c {};
        c.dim = 1; c.Nx = Nx; c.Ny = 1; c.Nz = 1; c.Nu = Nu; c.Nv = 1; c.Nw = 1; c.Nt = Nt;
        c.dt = dt; c.x_min = x_min; c.x_max = x_max; c.u_min = u_min; c.u_max = u_max;
        c.dx = dx; c.dx_inv = dx_inv; c.Lx = Lx; c.Lx_inv = Lx_inv; c.du = du;
        c.f0 = f0_spec;
        return c;
    }
};

template <typename real, size_t order>
class gpu_scheduler
{
public:
    gpu_scheduler() = delete;
    gpu_scheduler( const gpu_scheduler& ) = delete;
    gpu_scheduler& operator=( const gpu_scheduler& ) = delete;
    gpu_scheduler( gpu_scheduler &&rhs ) noexcept: conf_ { rhs.conf_ }, ctx_ { rhs.ctx_ }, up_ { std::move( rhs.up_ ) }
    { rhs.ctx_ = nullptr; }
    gpu_scheduler& operator=( gpu_scheduler &&rhs ) noexcept
    {
        if ( this != &rhs )
        {
            nufi_destroy( ctx_ ); conf_ = rhs.conf_; ctx_ = rhs.ctx_; up_ = std::move( rhs.up_ ); rhs.ctx_ = nullptr;
        }
        return *this;
    }

    explicit gpu_scheduler( const config_t<real> &conf, int device = 0 ): conf_ { conf }
    {
        const nufi_config c = conf.c_config();
        detail::check( nufi_create( &c, nullptr, int(order), precision(), device, &ctx_ ), nullptr );
    }

    gpu_scheduler( const config_t<real> &conf, const config_t<real> &conf_metrics, int device = 0 ): conf_ { conf }
    {
        const nufi_config c = conf.c_config(), cm = conf_metrics.c_config();
        detail::check( nufi_create( &c, &cm, int(order), precision(), device, &ctx_ ), nullptr );
    }

    ~gpu_scheduler() { nufi_destroy( ctx_ ); }

    // cuda_scheduler.hpp:88-111
    void compute_rho( size_t n, size_t q_begin, size_t q_end )
    { detail::check( nufi_compute_rho( ctx_, n, q_begin, q_end ), ctx_ ); }

    // cuda_scheduler.hpp:113-118: synchronise, rho[i] += device rho
    void download_rho( real *rho )
    {
        if ( std::is_same<real,double>::value )
        { detail::check( nufi_download_rho( ctx_, reinterpret_cast<double*>( rho ) ), ctx_ ); return; }
        std::vector<double> tmp( conf_.Nx, 0.0 );
        detail::check( nufi_download_rho( ctx_, tmp.data() ), ctx_ );
        for ( size_t i = 0; i < conf_.Nx; ++i ) rho[i] += real( tmp[i] );
    }

    // cuda_scheduler.hpp:120-125: slice n of the host history at coeffs + n*stride_t
    void upload_phi( size_t n, const real *coeffs )
    {
        if ( std::is_same<real,double>::value )
        { detail::check( nufi_upload_phi( ctx_, n, reinterpret_cast<const double*>( coeffs ) ), ctx_ ); return; }
        // a float history crosses as doubles: slice n is widened into a
        // history-sized buffer made once, so each call costs O(stride_t)
        if ( n > conf_.Nt ) throw std::range_error( "upload_phi: time step out of range" );
        const size_t st = stride_t();
        if ( up_.empty() ) up_.assign( ( conf_.Nt + 1 ) * st, 0.0 );
        for ( size_t i = 0; i < st; ++i ) up_[n*st + i] = double( coeffs[n*st + i] );
        detail::check( nufi_upload_phi( ctx_, n, up_.data() ), ctx_ );
    }

    // cuda_scheduler.hpp:128-157
    void compute_metrics( size_t n, size_t q_begin, size_t q_end )
    { detail::check( nufi_compute_metrics( ctx_, n, q_begin, q_end ), ctx_ ); }

    void download_metrics( real *metrics )
    {
        double m[4] = { 0, 0, 0, 0 };
        detail::check( nufi_download_metrics( ctx_, m ), ctx_ );
        for ( int i = 0; i < 4; ++i ) metrics[i] += real( m[i] );
    }

    // device-resident step: rho(n) -> Poisson -> fit -> slice n; returns the electric energy
    double step( size_t n )
    {
        double e = 0;
        detail::check( nufi_step( ctx_, n, &e ), ctx_ );
        return e;
    }

    size_t stride_t() const noexcept { return conf_.Nx + order - 1; }
    nufi_ctx *handle() noexcept { return ctx_; }

private:
    static int precision() noexcept { return std::is_same<real,float>::value ? NUFI_FP32C : NUFI_FP64; }
    config_t<real> conf_;
    nufi_ctx *ctx_ = nullptr;
    std::vector<double> up_;       // float: the widened host history upload_phi hands over
};

} // namespace dim1

namespace dim3
{

// Fields and defaults of dim3::config_t (config.hpp:168-219).
template <typename real>
struct config_t
{
    size_t Nx, Ny, Nz;
    size_t Nu, Nv, Nw;
    size_t Nt;
    real   dt;
    real   x_min, x_max, y_min, y_max, z_min, z_max;
    real   u_min, u_max, v_min, v_max, w_min, w_max;
    real   dx, dx_inv, Lx, Lx_inv;
    real   dy, dy_inv, Ly, Ly_inv;
    real   dz, dz_inv, Lz, Lz_inv;
    real   du, dv, dw;
    nufi_f0 f0_spec;

    config_t() noexcept
    {
        Nx = Ny = Nz = 8;
        Nu = Nv = Nw = 8;
        u_min = v_min = w_min = -9;
        u_max = v_max = w_max = 0;
        x_min = y_min = z_min = 0;
        x_max = y_max = z_max = 20*M_PI/3.0;
        dt = 1./10.; Nt = 5/dt;
        derive();
        f0_spec = nufi_f0 {};
        f0_spec.kind = NUFI_F0_BUMPONTAIL_3D;                                         // config.hpp:244-246
    }

    void derive() noexcept
    {
        Lx = x_max - x_min; Lx_inv = 1/Lx;
        Ly = y_max - y_min; Ly_inv = 1/Ly;
        Lz = z_max - z_min; Lz_inv = 1/Lz;
        dx = Lx/Nx; dx_inv = 1/dx;
        dy = Ly/Ny; dy_inv = 1/dy;
        dz = Lz/Nz; dz_inv = 1/dz;
        du = (u_max - u_min)/Nu;
        dv = (v_max - v_min)/Nv;
        dw = (w_max - w_min)/Nw;
    }

    nufi_config c_config() const noexcept
    {
        nufi_config c {};
        c.dim = 3; c.Nx = Nx; c.Ny = Ny; c.Nz = Nz; c.Nu = Nu; c.Nv = Nv; c.Nw = Nw; c.Nt = Nt; c.dt = dt;
        c.x_min = x_min; c.x_max = x_max; c.y_min = y_min; c.y_max = y_max; c.z_min = z_min; c.z_max = z_max;
        c.u_min = u_min; c.u_max = u_max; c.v_min = v_min; c.v_max = v_max; c.w_min = w_min; c.w_max = w_max;
        c.dx = dx; c.dx_inv = dx_inv; c.Lx = Lx; c.Lx_inv = Lx_inv;
        c.dy = dy; c.dy_inv = dy_inv; c.Ly = Ly; c.Ly_inv = Ly_inv;
        c.dz = dz; c.dz_inv = dz_inv; c.Lz = Lz; c.Lz_inv = Lz_inv;
        c.du = du; c.dv = dv; c.dw = dw;
        c.f0 = f0_spec;
        return c;
    }
};

template <typename real, size_t order>
class gpu_scheduler
{
public:
    gpu_scheduler() = delete;
    gpu_scheduler( const gpu_scheduler& ) = delete;
    gpu_scheduler& operator=( const gpu_scheduler& ) = delete;
    gpu_scheduler( gpu_scheduler &&rhs ) noexcept: conf_ { rhs.conf_ }, ctx_ { rhs.ctx_ }, up_ { std::move( rhs.up_ ) }
    { rhs.ctx_ = nullptr; }
    gpu_scheduler& operator=( gpu_scheduler &&rhs ) noexcept
    {
        if ( this != &rhs )
        {
            nufi_destroy( ctx_ ); conf_ = rhs.conf_; ctx_ = rhs.ctx_; up_ = std::move( rhs.up_ ); rhs.ctx_ = nullptr;
        }
        return *this;
    }

    explicit gpu_scheduler( const config_t<real> &conf, int device = 0 ): conf_ { conf }
    {
        const nufi_config c = conf.c_config();
        detail::check( nufi_create( &c, nullptr, int(order), precision(), device, &ctx_ ), nullptr );
    }

    ~gpu_scheduler() { nufi_destroy( ctx_ ); }

    void compute_rho( size_t n, size_t q_begin, size_t q_end )
    { detail::check( nufi_compute_rho( ctx_, n, q_begin, q_end ), ctx_ ); }

    void download_rho( real *rho )
    {
        const size_t N = conf_.Nx*conf_.Ny*conf_.Nz;
        if ( std::is_same<real,double>::value )
        { detail::check( nufi_download_rho( ctx_, reinterpret_cast<double*>( rho ) ), ctx_ ); return; }
        std::vector<double> tmp( N, 0.0 );
        detail::check( nufi_download_rho( ctx_, tmp.data() ), ctx_ );
        for ( size_t i = 0; i < N; ++i ) rho[i] += real( tmp[i] );
    }

    void upload_phi( size_t n, const real *coeffs )
    {
        if ( std::is_same<real,double>::value )
        { detail::check( nufi_upload_phi( ctx_, n, reinterpret_cast<const double*>( coeffs ) ), ctx_ ); return; }
        // a float history crosses as doubles: slice n is widened into a
        // history-sized buffer made once, so each call costs O(stride_t)
        if ( n > conf_.Nt ) throw std::range_error( "upload_phi: time step out of range" );
        const size_t st = stride_t();
        if ( up_.empty() ) up_.assign( ( conf_.Nt + 1 ) * st, 0.0 );
        for ( size_t i = 0; i < st; ++i ) up_[n*st + i] = double( coeffs[n*st + i] );
        detail::check( nufi_upload_phi( ctx_, n, up_.data() ), ctx_ );
    }

    void compute_metrics( size_t n, size_t q_begin, size_t q_end )
    { detail::check( nufi_compute_metrics( ctx_, n, q_begin, q_end ), ctx_ ); }

    void download_metrics( real *metrics )
    {
        double m[4] = { 0, 0, 0, 0 };
        detail::check( nufi_download_metrics( ctx_, m ), ctx_ );
        for ( int i = 0; i < 4; ++i ) metrics[i] += real( m[i] );
    }

    double step( size_t n )
    {
        double e = 0;
        detail::check( nufi_step( ctx_, n, &e ), ctx_ );
        return e;
    }

    size_t stride_t() const noexcept
    {
        return ( conf_.Nx + order - 1 )*( conf_.Ny + order - 1 )*( conf_.Nz + order - 1 );
    }
    nufi_ctx *handle() noexcept { return ctx_; }

private:
    static int precision() noexcept { return std::is_same<real,float>::value ? NUFI_FP32C : NUFI_FP64; }
    config_t<real> conf_;
    nufi_ctx *ctx_ = nullptr;
    std::vector<double> up_;       // float: the widened host history upload_phi hands over
};

} // namespace dim3

} // namespace nufi

#endif
