// sample1d.hip — diagnostics samplers of the 1D path: the field, f and the
// phase flow at arbitrary points, from the device history.
//
// Reference: the plotting block of bin/test_nufi_gpu_1d.cpp:290-361 evaluates,
// on the host, E = -dphi/dx and rho = -d^2phi/dx^2 of slice n
// (dim1::eval<order,1|2>, nufi/fields.hpp:36-61), f = eval_f(n, x, v)
// (nufi/rho.hpp:63-96: the trace with the initial half-kick from slice n) and
// the phase flow eval_phase_flow(n, x, v) (nufi/rho.hpp:98-131: the same
// trace's foot point for n > 1, the point itself for n <= 1, then x wrapped
// by x - Lx floor(x / Lx)).  Here one thread takes one point and reads the
// slices straight from the (L2-resident) difference history; the arithmetic
// is the reference's operation order (strict1d.hpp), f0 as in the trace
// kernels.
#include "device_common.hpp"
#include "internal.hpp"
#include "strict1d.hpp"

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

template <int ORDER>
__global__ __launch_bounds__( 256 )
void sample_field_kernel( sample1_params p )
{
#pragma clang fp contract(off)
    const uint64_t i = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i >= p.count ) return;
    const double *D = p.dhist + p.n * p.stride_d + 1;
    const double x = p.x[i];
    p.out0[i] = -phi_deriv_strict<ORDER,1>( x, D, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
    p.out1[i] = -phi_deriv_strict<ORDER,2>( x, D, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
}

// FLOW = false: f = eval_f(n, x, v); FLOW = true: the phase flow (x, v)
template <int ORDER, bool FLOW>
__global__ __launch_bounds__( 256 )
void sample_trace_kernel( sample1_params p )
{
#pragma clang fp contract(off)
    const uint64_t i = uint64_t( blockIdx.x ) * blockDim.x + threadIdx.x;
    if ( i >= p.count ) return;
    double x = p.x[i], u = p.v[i];
    const double half_dt = 0.5 * p.dt;
    uint64_t n = p.n;
    const bool trace = FLOW ? n > 1 : n > 0;
    if ( trace )
    {
        const double *D = p.dhist + n * p.stride_d + 1;
        u = u + half_dt * -phi_deriv_strict<ORDER,1>( x, D, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
        while ( --n )
        {
            x = x - p.dt * u;
            D = p.dhist + n * p.stride_d + 1;
            u = u + p.dt * -phi_deriv_strict<ORDER,1>( x, D, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
        }
        x = x - p.dt * u;
        u = u + half_dt * -phi_deriv_strict<ORDER,1>( x, p.dhist + 1, p.x_min, p.Lx, p.Lx_inv, p.dx_inv );
    }
    if ( FLOW )
    {
        p.out0[i] = x - p.Lx * floor( x * p.Lx_inv );
        p.out1[i] = u;
    }
    else p.out0[i] = f0_1d( p.f0, x, u );
}

template <int ORDER>
hipError_t launch_order( const sample1_params &p, int what, hipStream_t st )
{
    const uint32_t grid = uint32_t( ( p.count + 255 ) / 256 );
    switch ( what )
    {
    case SAMPLE_FIELD: hipLaunchKernelGGL( sample_field_kernel<ORDER>, dim3( grid ), dim3( 256 ), 0, st, p ); break;
    case SAMPLE_F:     hipLaunchKernelGGL( ( sample_trace_kernel<ORDER,false> ), dim3( grid ), dim3( 256 ), 0, st, p ); break;
    case SAMPLE_FLOW:  hipLaunchKernelGGL( ( sample_trace_kernel<ORDER,true> ), dim3( grid ), dim3( 256 ), 0, st, p ); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace

hipError_t launch_sample1d( const sample1_params &p, int what, int order, hipStream_t st )
{
    if ( p.count == 0 ) return hipSuccess;
    if ( p.count > uint64_t( 0xffffffffu ) * 256u ) return hipErrorInvalidValue;
    switch ( order )
    {
    case 3: return launch_order<3>( p, what, st );
    case 4: return launch_order<4>( p, what, st );
    case 5: return launch_order<5>( p, what, st );
    case 6: return launch_order<6>( p, what, st );
    case 7: return launch_order<7>( p, what, st );
    case 8: return launch_order<8>( p, what, st );
    default: return hipErrorInvalidValue;
    }
}

} // namespace nufi_impl
