// trace3d.hip — the 3D3V characteristic trace and rho partial sums on gfx950.
//
// Reference: dim3::eval_ftilda (nufi/rho.hpp:318-367) inside dim3::cuda_eval_rho
// (nufi/cuda_kernel.cu:392-426).  Per characteristic step: drift x,y,z; E =
// -grad phi_m from three tensor-spline evaluations (fields.hpp:308-350 ->
// splines.hpp:144-172); kick u,v,w.  After the loop a final drift and
// half-kick with slice 0, and f0 at the foot.
//
// The workgroup streams coefficient slices through LDS (one slice is
// (Nx+3)(Ny+3)(Nz+3) reals: 6859 = 54.9 KB FP64 at 16^3) and every node
// gathers its 4x4x4 window from LDS.  The fast FP64 rho kernel lays a wave's
// lanes over 64 neighbouring spatial nodes at one velocity (x-major, below),
// so a wave's windows overlap; each lane sums f over a fixed group of
// consecutive velocities.  The FP32C, strict and metrics kernels keep the
// reference's velocity-major order: a wave takes 64 velocity nodes of one
// spatial point and reduces them with a compensated butterfly.  (With half
// the LDS bytes per gather, the FP32C trace is bound by the VALU, not the
// LDS, and velocity-major stages each slice once per workgroup instead of
// once per velocity pass.)  The partial sums are added in a fixed order by
// the field kernel and do not depend on the launch shape.
//
// Variants:
//   strict : the reference's arithmetic (three separate tensor evaluations,
//            Cox-de Boor bases, de Boor in z, unwrapped positions), no FMA
//            contraction.
//   fast   : cell-unit positions wrapped exactly (every 16 slices on power-of-
//            two grids, the cell index by a mask in between); the value and
//            derivative bases of each axis in closed form; the three gradient
//            components share one pass over the 64 coefficients:
//              A_jk = sum_i c_ijk Bx_i,  X_jk = sum_i c_ijk Dx_i    (128 FMA)
//              P_k = sum_j A_jk By_j, Q_k = sum_j A_jk Dy_j,
//              R_k = sum_j X_jk By_j                                  (48 FMA)
//              Ex ~ sum_k R_k Bz_k, Ey ~ sum_k Q_k Bz_k, Ez ~ sum_k P_k Dz_k (12)
//   fp32c  : as fast in FP32 from an FP32 copy of the slice, positions (cell
//            units) and velocities (cells per step, w = dt dx_inv u) as hi + lo
//            pairs held at one offset (one binade for every hi part): exact
//            drifts of the hi parts, ordered Fast2Sum kicks; f0 and the rho
//            sums in FP64.  The copy is scaled in
//            FP64 by the kick in cells per step over 72 before its one rounding
//            and the bases are 6B and 2D, so no FP32-rounded constant (1/6,
//            -dt/dx) multiplies the kicks: such a constant biases every node's
//            kick by the same relative error, which the sum over velocities
//            does not average away (for 3D Landau it shifted the phase of the
//            damped field enough to give 2e-5 relative energy error near a
//            minimum of the energy).  For the same reason the four FP32 slope
//            bases sum to exactly zero (cubic_bases_x6).
#include "device_common.hpp"
#include "internal.hpp"

#include <type_traits>

namespace nufi_impl
{

using namespace nufi_dev;

namespace
{

template <int BLOCK, typename T>
__device__ __forceinline__ void stage3( T *__restrict__ dst, const T *__restrict__ src, uint32_t count )
{
    // slice bases are not 16-byte aligned in general (odd strides): scalar
    // copies, STAGE_BATCH loads in flight per thread before their LDS stores
    constexpr uint32_t B = STAGE_BATCH;
    for ( uint32_t base = threadIdx.x; base < count; base += B * BLOCK )
    {
        T v[B];
        #pragma unroll
        for ( uint32_t k = 0; k < B; ++k )             // clamped, unconditional loads stay in registers
        {
            const uint32_t i = base + k * BLOCK;
            v[k] = src[ i < count ? i : count - 1 ];
        }
        #pragma unroll
        for ( uint32_t k = 0; k < B; ++k )
        {
            const uint32_t i = base + k * BLOCK;
            if ( i < count ) dst[i] = v[k];
        }
    }
}

// The FP64 slices start 8-byte aligned (odd strides): copy the 16-byte words
// that cover `count` doubles from src and return where src[0] lands in dst
// (dst + 0 or + 1).  The history allocation has slack past its last slice.
template <int BLOCK>
__device__ __forceinline__ const double *dma_doubles( double *dst, const double *src, uint32_t count )
{
    const uint32_t mis = uint32_t( reinterpret_cast<uintptr_t>( src ) & 15u ) / 8u;
    dma_copy16<BLOCK>( dst, src - mis, ( count + mis + 1 ) / 2 );
    return dst + mis;
}

struct node3
{
    uint32_t g, l, j, iv;
    bool task, active;
};

template <int WPB, int NPT>
__device__ __forceinline__ node3 make_node3( const trace3_params &p, int r )
{
    node3 id;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    id.g = ( blockIdx.x * NPT + r ) * WPB + wave;
    id.task = id.g < p.n_tasks;
    const uint32_t ll = id.task ? id.g / p.wpx : 0;
    id.j  = id.task ? id.g % p.wpx : 0;
    id.l  = p.l_first + ll;
    id.iv = id.j * WAVE + lane;
    const uint64_t nvel = uint64_t( p.Nu ) * p.Nv * p.Nw;
    const uint64_t q = uint64_t( id.l ) * nvel + id.iv;
    id.active = id.task && id.iv < nvel && q >= p.q_begin && q < p.q_end;
    return id;
}

__device__ __forceinline__ void write_partial3( const trace3_params &p, const node3 &id, double f )
{
    const csum v = wave_csum( csum { f, 0.0 } );
    if ( id.task && ( threadIdx.x % WAVE ) == 0 )
        p.partials[ size_t( id.j ) * ( size_t( p.Nx ) * p.Ny * p.Nz ) + id.l ] = v.s + v.c;
}

// the four metric integrands (cuda_kernel.cu:453-465), wave-summed
__device__ __forceinline__ void write_metrics3( const trace3_params &p, const node3 &id, double f, double v2 )
{
    const bool a = id.active;
    const double w = p.weight;
    double m[4];
    m[0] = a ? w*f : 0.0;
    m[1] = a ? w*f*f : 0.0;
    m[2] = a ? w*v2*f/2 : 0.0;
    m[3] = ( a && f > 0 ) ? -w*f*log(f) : 0.0;
    #pragma unroll
    for ( int k = 0; k < 4; ++k )
    {
        const csum v = wave_csum( csum { m[k], 0.0 } );
        if ( id.task && ( threadIdx.x % WAVE ) == 0 ) p.partials[ size_t( id.g ) * 4 + k ] = v.s + v.c;
    }
}

template <int MODE>
__device__ __forceinline__ int64_t top_slice3( uint64_t n )
{
    if ( MODE == MODE_METRICS ) return n == 0 ? -1 : int64_t( n );
    return int64_t( n ) - 1;
}

// node coordinates, cuda_kernel.cu:402-418
__device__ __forceinline__ void node_coords( const trace3_params &p, const node3 &id,
                                             uint32_t &ix, uint32_t &iy, uint32_t &iz,
                                             double &u, double &v, double &w )
{
    ix = id.l % p.Nx; iy = ( id.l / p.Nx ) % p.Ny; iz = id.l / ( p.Nx * p.Ny );
    const uint32_t iu = id.iv % p.Nu, ivv = ( id.iv / p.Nu ) % p.Nv, iw = id.iv / ( p.Nu * p.Nv );
    u = p.u_min + iu  * p.du + p.du / 2;
    v = p.v_min + ivv * p.dv + p.dv / 2;
    w = p.w_min + iw  * p.dw + p.dw / 2;
}

// ---- closed-form cubic B-spline bases on [0,1) ------------------------------
// values  B = ((1-t)^3, 3t^3-6t^2+4, -3t^3+3t^2+3t+1, t^3)/6
// slopes  D = (-(1-t)^2, 3t^2-4t, -3t^2+2t+1, t^2)/2
// Both kernels evaluate them times 6 (values) and 2 (slopes): integer-
// coefficient polynomials, with -(1-t)^2 and t^2 left to the contraction as a
// negation and a reuse.  The gradient from these bases is 72 times the
// gradient in reference-cell units: the FP64 kick folds 1/72 into its FP64
// constant; the FP32C kernel reads slices pre-scaled by fp32c3_scale, so its
// gradient is the kick itself.
//
// FP64 form: every constant is one the FP64 VALU takes inline (1, 2, 4, -4),
// so none occupies registers; the second value and slope come from the first
// by the symmetry B_2(t) = B_1(1-t), D_2(t) = -D_1(1-t):
//   B_1 = 3t^2 (t-2) + 4,  D_1 = 3t^2 - 4t,  3t^2 = 2t^2 + t^2.
__device__ __forceinline__ void cubic_bases_x6( double t, double B[4], double D[4] )
{
    const double a = 1.0 - t, t2 = t*t, a2 = a*a;
    const double t23 = fma( 2.0, t2, t2 ), a23 = fma( 2.0, a2, a2 );
    B[0] = a2 * a;
    B[3] = t2 * t;
    B[1] = fma( t23, t - 2.0, 4.0 );
    B[2] = fma( a23, a - 2.0, 4.0 );
    D[0] = -a2;
    D[3] = t2;
    D[1] = fma( -4.0, t, t23 );
    D[2] = fma( 4.0, a, -a23 );
}

// FP32 form.  The four slopes must sum to exactly zero as the FP32 numbers
// they are: a gradient contraction sum_i c_i D_i then carries no term in the
// level of the coefficients, only in their differences.  Rounded slopes leave
// a residual r(t) = sum_i D_i of about 2^-24 that multiplies the local level
// of phi, not its gradient; r depends on t alone, so it does not average out
// over cells or velocities, and on 3D Landau it gave 2e-6 relative energy
// error where the field passes near zero.  So (1-t)^2, t^2 and 3t^2 - 4t are
// each rounded once onto multiples of 2^-23 (an FMA that adds 1 or -1, which
// puts the result at magnitude [1,2] or a little above, then that 1 taken back
// exactly), and the third slope (a^2 - t^2) - D_1, whose partial results stay
// below 2 in magnitude, is exact.  The values need no such care (the slopes
// of the other axes cancel their level): B_2 is B_1 at 1 - t.
__device__ __forceinline__ void cubic_bases_x6( float t, float B[4], float D[4] )
{
    const float a = 1.0f - t;
    const float a2 = fmaf( a, a, 1.0f ) - 1.0f;                  // (1-t)^2 on the 2^-23 grid
    const float t2 = fmaf( t, t, 1.0f ) - 1.0f;                  // t^2 on the 2^-23 grid
    B[0] = a2 * a;
    B[3] = t2 * t;
    B[1] = fmaf( t2, fmaf( 3.0f, t, -6.0f ), 4.0f );            // 3t^3 - 6t^2 + 4
    B[2] = fmaf( a2, fmaf( 3.0f, a, -6.0f ), 4.0f );            // B_1(1-t)
    D[0] = -a2;
    D[3] = t2;
    D[1] = fmaf( t, fmaf( 3.0f, t, -4.0f ), -1.0f ) + 1.0f;     // 3t^2 - 4t on the 2^-23 grid
    D[2] = ( a2 - t2 ) - D[1];                                   // exact
}

// Shared-basis gradient of the 4x4x4 window at c (72 times the gradient in
// reference-cell units with the bases above): per row (j,k)
//   A_jk = sum_i c_ijk Bx_i,  X_jk = sum_i c_ijk Dx_i,
// per plane k  P_k = sum_j A_jk By_j, Q_k = sum_j A_jk Dy_j, R_k = sum_j X_jk By_j,
// and g = (sum_k R_k Bz_k, sum_k Q_k Bz_k, sum_k P_k Dz_k).
//
// The window gather is software-pipelined (FP64 and FP32C kernels): the 16 rows
// (4 values each) of the 4x4x4 window are read PF rows ahead of the FMAs that
// use them, and the first PF rows before the bases are computed, so each row's
// LDS latency runs under earlier work.  The scheduling barriers keep the
// compiler from sinking the reads back next to their uses (it does so to save
// registers, and the waves a SIMD holds then cannot cover the latency).
// Measured at 512 x 4 on 3D Landau (FP64): PF = 1 is the fastest, 2 and 3 no
// faster, 4 slower.
constexpr int GATHER_PF = 1;

// The rows of a node's 4x4x4 window: row idx (j = idx & 3, k = idx >> 2) is
// the four reals from c + k*sz + j*sy.
template <typename T>
struct window_rt
{
    const T *c;
    uint32_t sy, sz;
    __device__ __forceinline__ const T *row( int idx ) const { return c + ( idx >> 2 ) * sz + ( idx & 3 ) * sy; }
};

// The same rows in the slice of a G^3 grid (strides G+3 and (G+3)^2): one LDS
// address per z plane, held in a register, and the rows of a plane at
// immediate offsets of the reads (at most 3(G+3)+3 reals, within the 255
// elements of ds_read2_b64 / ds_read2_b32 offsets for G <= 77).  Without the
// opaque register the compiler folds every row into a base address of its own.
template <int G, typename T = double>
struct window_g
{
    using lds_ptr = const __attribute__(( address_space( 3 ) )) T *;
    lds_ptr pl[4];
    __device__ __forceinline__ explicit window_g( const T *w )
    {
        pl[0] = (lds_ptr)w;
        #pragma unroll
        for ( int k = 1; k < 4; ++k )
        {
            uint32_t a = uint32_t( size_t( (lds_ptr)( w + k * ( G + 3 ) * ( G + 3 ) ) ) );
            asm volatile( "" : "+v"( a ) );
            pl[k] = (lds_ptr)size_t( a );
        }
    }
    __device__ __forceinline__ lds_ptr row( int idx ) const { return pl[ idx >> 2 ] + ( idx & 3 ) * ( G + 3 ); }
};

template <typename Win, typename T>
__device__ __forceinline__ void load_row( const Win &w, int idx, T r[4] )
{
    const auto q = w.row( idx );
    r[0] = q[0]; r[1] = q[1]; r[2] = q[2]; r[3] = q[3];
}

// the first PF rows of the window, before the bases are computed
template <int PF, typename Win, typename T>
__device__ __forceinline__ void grad_prefetch( const Win &w, T buf[PF + 1][4] )
{
    #pragma unroll
    for ( int d = 0; d < PF; ++d ) load_row( w, d, buf[d] );
    __builtin_amdgcn_sched_barrier( 0 );
}

template <int PF, typename Win, typename T>
__device__ __forceinline__ void grad_rows( const Win &w, T buf[PF + 1][4],
                                           const T Bx[4], const T Dx[4], const T By[4],
                                           const T Dy[4], const T Bz[4], const T Dz[4],
                                           T &gx, T &gy, T &gz )
{
    T P = T(0), Q = T(0), R = T(0);
    gx = T(0); gy = T(0); gz = T(0);
    #pragma unroll
    for ( int idx = 0; idx < 16; ++idx )
    {
        if ( idx + PF < 16 ) load_row( w, idx + PF, buf[ ( idx + PF ) % ( PF + 1 ) ] );
        __builtin_amdgcn_sched_barrier( 0 );
        const T *r = buf[ idx % ( PF + 1 ) ];
        const int j = idx & 3, k = idx >> 2;
        const T A = fma( r[3], Bx[3], fma( r[2], Bx[2], fma( r[1], Bx[1], r[0]*Bx[0] ) ) );
        const T X = fma( r[3], Dx[3], fma( r[2], Dx[2], fma( r[1], Dx[1], r[0]*Dx[0] ) ) );
        P = j == 0 ? A * By[0] : fma( A, By[j], P );
        Q = j == 0 ? A * Dy[0] : fma( A, Dy[j], Q );
        R = j == 0 ? X * By[0] : fma( X, By[j], R );
        if ( j == 3 )
        {
            gx = fma( R, Bz[k], gx );
            gy = fma( Q, Bz[k], gy );
            gz = fma( P, Dz[k], gz );
        }
    }
}

// one axis of the fast FP64 step: drift (unless `still`), exact wrap, locate.
// MASK (power-of-two axis): no wrap here, the cell index wraps by a mask and
// the caller re-wraps the position exactly before every slice m with
// wrap3_due(m), a schedule fixed by m alone.
constexpr int WRAP3 = 16;
__device__ __forceinline__ bool wrap3_due( int64_t m ) { return ( m & ( WRAP3 - 1 ) ) == WRAP3 - 1; }

template <bool MASK>
__device__ __forceinline__ int advance_axis( double &s, double vel, double cd, double N, double invN, int Ni,
                                             double &t, bool still = false )
{
    double x = still ? s : fma( -cd, vel, s );
    if ( !MASK ) x = fma( -N, floor( x * invN ), x );
    const double k = floor( x );
    t = x - k;
    int i = int( k );
    if ( MASK ) i &= Ni - 1;
    else
    {
        i = ( i < 0 ) ? i + Ni : i;
        i = ( i >= Ni ) ? i - Ni : i;
    }
    s = x;
    return i;
}

// exact periodic re-wrap (N*floor(s/N) is an integer)
__device__ __forceinline__ void rewrap3( double &s, double N, double invN )
{
    s = fma( -N, floor( s * invN ), s );
}

// one fast FP64 characteristic step of one node with slice C: drift (unless
// `still`), exact wrap, E = -grad phi from the shared-basis window, kick by hf*dt
// (`half`: hf = 1/2, the kick constants are selected per slice, not multiplied)
struct fast3_consts
{
    double Nxd, Nyd, Nzd, iNx, iNy, iNz, cdx, cdy, cdz;
    double kx, ky, kz, kxh, kyh, kzh;     // -dt d_inv / 72 (times 72-scaled gradients), and half of it
    int Nxi, Nyi, Nzi;
    uint32_t sy, sz;
};

__device__ __forceinline__ fast3_consts make_fast3( const trace3_params &p )
{
    fast3_consts k;
    k.Nxd = p.Nx; k.Nyd = p.Ny; k.Nzd = p.Nz;
    k.iNx = 1.0 / k.Nxd; k.iNy = 1.0 / k.Nyd; k.iNz = 1.0 / k.Nzd;
    k.Nxi = int( p.Nx ); k.Nyi = int( p.Ny ); k.Nzi = int( p.Nz );
    k.cdx = p.dt * p.dx_inv; k.cdy = p.dt * p.dy_inv; k.cdz = p.dt * p.dz_inv;
    k.kx = p.kick72[0]; k.ky = p.kick72[1]; k.kz = p.kick72[2];           // kernel arguments: scalar registers
    k.kxh = p.kick72h[0]; k.kyh = p.kick72h[1]; k.kzh = p.kick72h[2];
    k.sy = p.sy; k.sz = p.sz;
    return k;
}

template <int G, typename T>
__device__ __forceinline__ typename std::enable_if<( G > 0 ), window_g<G, T>>::type
make_window( const T *w, uint32_t, uint32_t ) { return window_g<G, T>( w ); }

template <int G, typename T>
__device__ __forceinline__ typename std::enable_if<( G == 0 ), window_rt<T>>::type
make_window( const T *w, uint32_t sy, uint32_t sz ) { return window_rt<T> { w, sy, sz }; }

// G > 0: a cubic G^3 grid whose slice strides (G+3, (G+3)^2) are compile-time
// constants, so the window's row offsets are immediates of the LDS reads and
// only the window base is computed per node-step; G = 0 takes the strides from
// the kernel arguments (24-bit products: a slice stays far below 2^24 reals).
template <bool MASK = false, int PF = GATHER_PF, int G = 0>
__device__ __forceinline__ void step3_fast( const fast3_consts &k, const double *C, bool half, bool still,
                                            double &sx, double &sy_, double &sz_, double &vu, double &vv, double &vw )
{
    double tx, ty, tz;
    const int i = advance_axis<MASK>( sx,  vu, k.cdx, k.Nxd, k.iNx, k.Nxi, tx, still );
    const int j = advance_axis<MASK>( sy_, vv, k.cdy, k.Nyd, k.iNy, k.Nyi, ty, still );
    const int l = advance_axis<MASK>( sz_, vw, k.cdz, k.Nzd, k.iNz, k.Nzi, tz, still );
    constexpr uint32_t GY = G > 0 ? uint32_t( G + 3 ) : 0u, GZ = GY * GY;
    const uint32_t sy = G > 0 ? GY : k.sy, sz = G > 0 ? GZ : k.sz;
    const double *W = C + __umul24( uint32_t( l ), sz ) + __umul24( uint32_t( j ), sy ) + i;
    using Win = typename std::conditional<( G > 0 ), window_g<G>, window_rt<double>>::type;
    Win win = make_window<G>( W, sy, sz );
    double buf[PF + 1][4];
    grad_prefetch<PF>( win, buf );
    double Bx[4], Dx[4], By[4], Dy[4], Bz[4], Dz[4];
    cubic_bases_x6( tx, Bx, Dx );
    cubic_bases_x6( ty, By, Dy );
    cubic_bases_x6( tz, Bz, Dz );
    double gx, gy, gz;
    grad_rows<PF>( win, buf, Bx, Dx, By, Dy, Bz, Dz, gx, gy, gz );
    vu = fma( half ? k.kxh : k.kx, gx, vu );
    vv = fma( half ? k.kyh : k.ky, gy, vv );
    vw = fma( half ? k.kzh : k.kz, gz, vw );
}

template <int BLOCK, int NPT, int MODE>
__global__ __launch_bounds__( BLOCK )
void trace3d_fast_kernel( trace3_params p )
{
    extern __shared__ double lds[];
    constexpr int WPB = BLOCK / WAVE;

    node3 id[NPT];
    double sx[NPT], sy_[NPT], sz_[NPT], vu[NPT], vv[NPT], vw[NPT];
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        id[r] = make_node3<WPB,NPT>( p, r );
        uint32_t ix, iy, iz;
        node_coords( p, id[r], ix, iy, iz, vu[r], vv[r], vw[r] );
        sx[r] = double( ix ); sy_[r] = double( iy ); sz_[r] = double( iz );
    }
    const fast3_consts k3 = make_fast3( p );
    const uint32_t st = p.st;

    for ( int64_t mhi = top_slice3<MODE>( p.n ); mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        stage3<BLOCK>( lds, p.hist + size_t( mlo ) * st, uint32_t( cnt ) * st );
        __syncthreads();
        for ( int jj = cnt - 1; jj >= 0; --jj )
        {
            const double *C = lds + size_t( jj ) * st;
            const uint64_t m = uint64_t( mlo + jj );
            const bool initial = ( MODE == MODE_METRICS ) && m == p.n;     // rho.hpp:386-394
            const bool half = m == 0 || initial;
            #pragma unroll
            for ( int r = 0; r < NPT; ++r )
                step3_fast( k3, C, half, initial, sx[r], sy_[r], sz_[r], vu[r], vv[r], vw[r] );
        }
    }

    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        const double f = id[r].active
            ? f0_3d( p.f0, p.x_min + sx[r]*p.dx, p.y_min + sy_[r]*p.dy, p.z_min + sz_[r]*p.dz,
                     vu[r], vv[r], vw[r] )
            : 0.0;
        if ( MODE == MODE_RHO ) write_partial3( p, id[r], f );
        else
        {
            uint32_t ix, iy, iz; double u0, v0, w0;
            node_coords( p, id[r], ix, iy, iz, u0, v0, w0 );
            write_metrics3( p, id[r], f, u0*u0 + v0*v0 + w0*w0 );
        }
    }
}

// ---- x-major layout (fast and FP32C rho kernels) ---------------------------
// The 64 lanes of a wave take 64 consecutive spatial nodes l (x fastest) at
// one velocity, so the 4x4x4 windows of a wave overlap and its LDS gathers
// are near-contiguous.  Workgroup b covers spatial tile b % n_st (l = l_first
// + 64*(b % n_st) + lane) and the velocity groups (b / n_st)*WPB + w of its
// waves w; a group is vgroup consecutive velocities iv.  Each lane traces the
// velocities of its group NPT at a time in iv order and sums f in a
// compensated accumulator, written to partials[group*n_space + l]: the
// partial sums do not depend on the launch shape.
struct xm3_lane
{
    uint32_t l, g, ix, iy, iz;
    bool ok;
};

template <int WPB>
__device__ __forceinline__ xm3_lane make_xm3( const trace3_params &p )
{
    xm3_lane a;
    const uint32_t lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    const uint32_t ll = ( blockIdx.x % p.n_st ) * WAVE + lane;
    a.g  = ( blockIdx.x / p.n_st ) * WPB + wave;
    a.ok = ll < p.n_count && a.g < p.ppx;
    a.l  = p.l_first + ( ll < p.n_count ? ll : 0u );
    a.ix = a.l % p.Nx; a.iy = ( a.l / p.Nx ) % p.Ny; a.iz = a.l / ( p.Nx * p.Ny );
    return a;
}

// velocity of node iv of the group (cuda_kernel.cu:402-418) and whether it is traced
__device__ __forceinline__ bool xm3_velocity( const trace3_params &p, const xm3_lane &a, uint32_t k,
                                              double &u, double &v, double &w )
{
    const uint64_t nvel = uint64_t( p.Nu ) * p.Nv * p.Nw;
    const uint64_t iv = uint64_t( a.g ) * p.vgroup + k;
    const uint64_t q = uint64_t( a.l ) * nvel + iv;
    const uint32_t ivc = iv < nvel ? uint32_t( iv ) : 0u;
    const uint32_t iu = ivc % p.Nu, ivv = ( ivc / p.Nu ) % p.Nv, iw = ivc / ( p.Nu * p.Nv );
    u = p.u_min + iu  * p.du + p.du / 2;
    v = p.v_min + ivv * p.dv + p.dv / 2;
    w = p.w_min + iw  * p.dw + p.dw / 2;
    return a.ok && iv < nvel && q >= p.q_begin && q < p.q_end;
}

// The NPT nodes a lane traces in one velocity pass of the x-major FP64 kernel
// (G: the compile-time grid edge of step3_fast, 0 for any grid).
template <int NPT, bool MASK, int G>
struct xm3_nodes
{
    double sx[NPT], sy[NPT], sz[NPT], vu[NPT], vv[NPT], vw[NPT];
    bool act[NPT];

    __device__ __forceinline__ void start( const trace3_params &p, const xm3_lane &a, uint32_t pass )
    {
        #pragma unroll
        for ( int r = 0; r < NPT; ++r )
        {
            act[r] = xm3_velocity( p, a, pass + r, vu[r], vv[r], vw[r] );
            sx[r] = double( a.ix ); sy[r] = double( a.iy ); sz[r] = double( a.iz );
        }
    }

    // slices mlo+cnt-1 .. mlo with the chunk's slices at Cb + jj*st; slice 0
    // (the half kick) is peeled off the loop
    __device__ __forceinline__ void steps( const fast3_consts &k3, const double *Cb, uint32_t st, int64_t mlo, int cnt )
    {
        const int jlast = mlo == 0 ? 1 : 0;
        for ( int jj = cnt - 1; jj >= jlast; --jj )
        {
            const double *C = Cb + size_t( jj ) * st;
            if ( MASK && wrap3_due( mlo + jj ) )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r )
                {
                    rewrap3( sx[r], k3.Nxd, k3.iNx ); rewrap3( sy[r], k3.Nyd, k3.iNy ); rewrap3( sz[r], k3.Nzd, k3.iNz );
                }
            }
            #pragma unroll
            for ( int r = 0; r < NPT; ++r )
                step3_fast<MASK, GATHER_PF, G>( k3, C, false, false, sx[r], sy[r], sz[r], vu[r], vv[r], vw[r] );
        }
        if ( mlo == 0 )                  // slice 0: wrap3_due(0) is false
        {
            #pragma unroll
            for ( int r = 0; r < NPT; ++r )
                step3_fast<MASK, GATHER_PF, G>( k3, Cb, true, false, sx[r], sy[r], sz[r], vu[r], vv[r], vw[r] );
        }
    }

    // f0 at the feet, added to the lane's compensated sum
    __device__ __forceinline__ void finish( const trace3_params &p, csum &acc ) const
    {
        #pragma unroll
        for ( int r = 0; r < NPT; ++r )
        {
            const double f = act[r]
                ? f0_3d( p.f0, p.x_min + sx[r]*p.dx, p.y_min + sy[r]*p.dy, p.z_min + sz[r]*p.dz, vu[r], vv[r], vw[r] )
                : 0.0;
            acc = csum_add( acc, csum { f, 0.0 } );
        }
    }
};

// 1024-thread workgroups run four waves per SIMD (128 VGPRs: the f0 evaluation
// after each velocity pass spills, the step loop does not); 512 threads, two
// (about 230 VGPRs, no spills; with four nodes a lane, the default shape)
template <int BLOCK, int NPT, bool MASK, int G = 0>
__global__ __launch_bounds__( BLOCK ) __attribute__(( amdgpu_waves_per_eu( BLOCK >= 1024 ? 4 : 1 ) ))
void trace3d_xm_fast_kernel( trace3_params p )
{
    extern __shared__ double lds[];
    constexpr int WPB = BLOCK / WAVE;
    const xm3_lane a = make_xm3<WPB>( p );
    const fast3_consts k3 = make_fast3( p );
    const uint32_t st = p.st;
    csum acc { 0.0, 0.0 };
    xm3_nodes<NPT,MASK,G> nd;
    const uint32_t Cn = p.chunk;
    const int64_t top = int64_t( p.n ) - 1;
    const int64_t K = ( top + int64_t( Cn ) ) / int64_t( Cn );      // chunks per pass (0 at n = 0)
    auto lo_of = [&]( int64_t k ) { const int64_t h = top - k * Cn; return h - int64_t( Cn ) + 1 > 0 ? h - int64_t( Cn ) + 1 : 0; };
    const uint32_t P = p.vgroup / NPT;                                // passes
    if ( p.dma && K > 0 )
    {
        // The (pass, chunk) sequence in one loop: the chunks alternate between
        // two LDS buffers (chunk*st + 2 doubles each); the next chunk, of this
        // pass or the next, is copied straight into the other buffer while this
        // one is traced, and one barrier per chunk publishes it.
        double *const buf0 = lds, *const buf1 = lds + size_t( Cn ) * st + 2;
        const int64_t T = int64_t( P ) * K;
        const double *C0;
        {
            const int64_t l0 = lo_of( 0 );
            C0 = dma_doubles<BLOCK>( buf0, p.hist + size_t( l0 ) * st, uint32_t( top - l0 + 1 ) * st );
            dma_publish();
        }
        const double *cur = C0;
        for ( int64_t i = 0; i < T; ++i )
        {
            const int64_t k = i % K;
            if ( k == 0 ) nd.start( p, a, uint32_t( i / K ) * NPT );
            const int64_t mlo = lo_of( k );
            const int cnt = int( top - k * Cn - mlo + 1 );
            const double *next = nullptr;
            if ( i + 1 < T )
            {
                const int64_t k1 = ( k + 1 ) % K, l1 = lo_of( k1 );
                next = dma_doubles<BLOCK>( ( i & 1 ) ? buf0 : buf1, p.hist + size_t( l1 ) * st,
                                           uint32_t( top - k1 * Cn - l1 + 1 ) * st );
            }
            nd.steps( k3, cur, st, mlo, cnt );
            if ( k == K - 1 ) nd.finish( p, acc );
            dma_publish();
            cur = next;
        }
    }
    else
    {
        for ( uint32_t pass = 0; pass < P; ++pass )
        {
            nd.start( p, a, pass * NPT );
            for ( int64_t k = 0; k < K; ++k )
            {
                const int64_t mlo = lo_of( k );
                const int cnt = int( top - k * Cn - mlo + 1 );
                __syncthreads();
                stage3<BLOCK>( lds, p.hist + size_t( mlo ) * st, uint32_t( cnt ) * st );
                __syncthreads();
                nd.steps( k3, lds, st, mlo, cnt );
            }
            nd.finish( p, acc );
        }
    }
    if ( a.ok ) p.partials[ size_t( a.g ) * ( size_t( p.Nx ) * p.Ny * p.Nz ) + a.l ] = acc.s + acc.c;
}

// ---- FP32 compensated --------------------------------------------------------
// Positions (cells) and velocities (cells per step) are unevaluated pairs
// hi + lo whose hi parts carry a per-axis offset, P0 + s and W0 + w (W0 = P0,
// fp32c_frame's shared form), chosen on the host so that every hi value of the
// run stays inside one binade, the same for both.  Then the offsets come off exactly, all hi parts are
// multiples of one ulp, so the drift of the hi parts is exact (sh - d stays in
// the binade), the kick is an ordered Fast2Sum (3 operations, exact) rather
// than a TwoSum (6), and a wrap by a multiple of N is exact.  The lo parts move
// unnormalised; both pairs are renormalised every WRAP3 slices.
struct frame32
{
    float P0, W0, N, invN;
    int Ni, P0i;
};

// drift s -= w, wrap into [P0, P0 + N) (unless MASK: P0 is then a multiple of
// N, the cell index wraps by the mask and frame32_rewrap wraps the position
// every WRAP3 slices), locate
template <bool MASK>
__device__ __forceinline__ int drift_locate32( const frame32 &f, float &sh, float &sl, float wh, float wl, float &t )
{
    const float d = wh - f.W0;                       // exact: wh and W0 share a binade
    sh = sh - d;                                     // exact: sh, d on the frames' one ulp grid
    sl = sl - wl;
    if ( !MASK ) sh = fmaf( -f.N, floorf( ( sh - f.P0 ) * f.invN ), sh );   // exact
    t = __builtin_amdgcn_fractf( sh ) + sl;          // sh > 0: the fraction is exact and
    int i = int( sh );                               // truncation is the floor
    if ( MASK ) i &= f.Ni - 1;
    else
    {
        i -= f.P0i;
        i = ( i < 0 ) ? i + f.Ni : i;
        i = ( i >= f.Ni ) ? i - f.Ni : i;
    }
    return i;
}

// every WRAP3 slices: wrap (MASK) and renormalise the position and velocity pairs
template <bool MASK>
__device__ __forceinline__ void frame32_rewrap( const frame32 &f, float &sh, float &sl, float &wh, float &wl )
{
    if ( MASK ) sh = fmaf( -f.N, floorf( ( sh - f.P0 ) * f.invN ), sh );      // exact
    float s = sh + sl;
    sl = sl - ( s - sh ); sh = s;
    s = wh + wl;
    wl = wl - ( s - wh ); wh = s;
}

// w += q: Fast2Sum, |wh| >= W0 / 2 > |q|
__device__ __forceinline__ void kick32( float &wh, float &wl, float q )
{
    const float s = wh + q;
    wl = wl + ( q - ( s - wh ) );
    wh = s;
}

// velocity-major: node r of a lane is make_node3's (wave task g, lane), the
// wave's 64 lanes are 64 consecutive velocities of one spatial point
// (G: compile-time grid edge as in step3_fast; CUBIC: dx = dy = dz, the kicks
// need no axis ratio)
constexpr int KICK_GENERAL = 0, KICK_ONE = 1, KICK_HALF = 2;

template <int BLOCK, int NPT, bool MASK, int G = 0, bool CUBIC = false>
__global__ __launch_bounds__( BLOCK ) __attribute__(( amdgpu_waves_per_eu( 4 ) ))
void trace3d_fp32c_kernel( trace3_params p )
{
    extern __shared__ float ldsf[];
    constexpr int WPB = BLOCK / WAVE;

    node3 id[NPT];
    float xh[NPT], xl[NPT], yh[NPT], yl[NPT], zh[NPT], zl[NPT];     // positions (cells) + P0
    float uh[NPT], ul[NPT], vh[NPT], vl[NPT], wh[NPT], wl[NPT];     // velocities (cells per step) + W0
    const double cdx = p.dt * p.dx_inv, cdy = p.dt * p.dy_inv, cdz = p.dt * p.dz_inv;
    frame32 fx, fy, fz;
    {
        const uint32_t Ns[3] = { p.Nx, p.Ny, p.Nz };
        frame32 *fs[3] = { &fx, &fy, &fz };
        #pragma unroll
        for ( int a = 0; a < 3; ++a )
        {
            fs[a]->P0 = p.f32_p0[a]; fs[a]->W0 = p.f32_w0[a];
            fs[a]->N = float( Ns[a] ); fs[a]->invN = 1.0f / fs[a]->N;
            fs[a]->Ni = int( Ns[a] ); fs[a]->P0i = int( p.f32_p0[a] );
        }
    }
    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        id[r] = make_node3<WPB,NPT>( p, r );
        uint32_t ix, iy, iz; double u, v, w;
        node_coords( p, id[r], ix, iy, iz, u, v, w );
        u = double( fx.W0 ) + u * cdx; v = double( fy.W0 ) + v * cdy; w = double( fz.W0 ) + w * cdz;
        xh[r] = fx.P0 + float( ix ); xl[r] = 0.0f;                   // exact
        yh[r] = fy.P0 + float( iy ); yl[r] = 0.0f;
        zh[r] = fz.P0 + float( iz ); zl[r] = 0.0f;
        uh[r] = float( u ); ul[r] = float( u - double( uh[r] ) );
        vh[r] = float( v ); vl[r] = float( v - double( vh[r] ) );
        wh[r] = float( w ); wl[r] = float( w - double( wh[r] ) );
    }
    // the slices carry the x kick in cells per step (fp32c3_scale); y and z
    // scale it by (dy_inv/dx_inv)^2 and (dz_inv/dx_inv)^2, split hi + lo
    // (exactly 1 + 0 on cubic cells)
    const double ry = ( p.dy_inv / p.dx_inv ) * ( p.dy_inv / p.dx_inv ), rz = ( p.dz_inv / p.dx_inv ) * ( p.dz_inv / p.dx_inv );
    const float ryh = float( ry ), ryl = float( ry - double( ryh ) );
    const float rzh = float( rz ), rzl = float( rz - double( rzh ) );
    const uint32_t st = p.st32, sy = G > 0 ? uint32_t( G + 3 ) : p.sy, sz = G > 0 ? uint32_t( ( G + 3 ) * ( G + 3 ) ) : p.sz;
    const uint32_t C = p.chunk;

    // one slice of all NPT nodes of the lane: drift, locate, gradient, kick
    auto trace_slice = [&]( const float *Cs, float hf, auto kick_form ) __attribute__(( always_inline ))
    {
        constexpr int KICK = decltype( kick_form )::value;
        #pragma unroll
        for ( int r = 0; r < NPT; ++r )
        {
            float tx, ty, tz;
            const int i = drift_locate32<MASK>( fx, xh[r], xl[r], uh[r], ul[r], tx );
            const int j = drift_locate32<MASK>( fy, yh[r], yl[r], vh[r], vl[r], ty );
            const int k = drift_locate32<MASK>( fz, zh[r], zl[r], wh[r], wl[r], tz );
            const float *W = Cs + __umul24( uint32_t( k ), sz ) + __umul24( uint32_t( j ), sy ) + i;
            float buf[GATHER_PF + 1][4];
            using Win = typename std::conditional<( G > 0 ), window_g<G, float>, window_rt<float>>::type;
            const Win win = make_window<G>( W, sy, sz );
            grad_prefetch<GATHER_PF>( win, buf );
            float Bx[4], Dx[4], By[4], Dy[4], Bz[4], Dz[4];
            cubic_bases_x6( tx, Bx, Dx );
            cubic_bases_x6( ty, By, Dy );
            cubic_bases_x6( tz, Bz, Dz );
            float gx, gy, gz;
            grad_rows<GATHER_PF>( win, buf, Bx, Dx, By, Dy, Bz, Dz, gx, gy, gz );
            if ( KICK == KICK_ONE )
            {
                kick32( uh[r], ul[r], gx ); kick32( vh[r], vl[r], gy ); kick32( wh[r], wl[r], gz );
            }
            else if ( KICK == KICK_HALF )
            {
                kick32( uh[r], ul[r], 0.5f * gx ); kick32( vh[r], vl[r], 0.5f * gy ); kick32( wh[r], wl[r], 0.5f * gz );
            }
            else
            {
                const float hyh = hf * ryh, hyl = hf * ryl, hzh = hf * rzh, hzl = hf * rzl;
                kick32( uh[r], ul[r], hf * gx );
                kick32( vh[r], vl[r], fmaf( gy, hyh, gy * hyl ) );
                kick32( wh[r], wl[r], fmaf( gz, hzh, gz * hzl ) );
            }
        }
    };

    // chunk k: slices max(top - kC - C + 1, 0) .. top - kC.  With p.dma the
    // chunks alternate between two LDS buffers: chunk k+1 is copied straight
    // into the other buffer while chunk k is traced, and one barrier per chunk
    // publishes it (the other buffer was last read in chunk k-1).
    const int64_t top = int64_t( p.n ) - 1;
    const int64_t K = ( top + int64_t( C ) ) / int64_t( C );
    auto lo_of = [&]( int64_t k ) { const int64_t h = top - k * C; return h - int64_t( C ) + 1 > 0 ? h - int64_t( C ) + 1 : 0; };
    float *const buf0 = ldsf, *const buf1 = ldsf + size_t( C ) * st;
    if ( p.dma && K > 0 )
    {
        const int64_t l0 = lo_of( 0 );
        dma_copy16<BLOCK>( buf0, p.hist32 + size_t( l0 ) * st, uint32_t( top - l0 + 1 ) * st / 4 );
        dma_publish();
    }
    for ( int64_t k = 0; k < K; ++k )
    {
        const int64_t mhi = top - k * C, mlo = lo_of( k );
        const int cnt = int( mhi - mlo + 1 );
        float *cur = buf0;
        if ( p.dma )
        {
            cur = ( k & 1 ) ? buf1 : buf0;
            if ( k + 1 < K )
            {
                const int64_t nlo = lo_of( k + 1 );
                dma_copy16<BLOCK>( ( k & 1 ) ? buf0 : buf1, p.hist32 + size_t( nlo ) * st,
                                   uint32_t( mlo - nlo ) * st / 4 );
            }
        }
        else
        {
            __syncthreads();
            stage3<BLOCK>( ldsf, p.hist32 + size_t( mlo ) * st, uint32_t( cnt ) * st );
            __syncthreads();
        }
        // slice 0 takes the half kick: hf = 1/2, a power of two, so hf * q is
        // exact.  On cubic cells ry = rz = 1 (hi 1, lo 0) and every kick is hf
        // times the gradient exactly, so slice 0 is peeled off the loop and the
        // slices m > 0 multiply by nothing (the same numbers as the general form).
        const int jlast = ( CUBIC && mlo == 0 ) ? 1 : 0;
        for ( int jj = cnt - 1; jj >= jlast; --jj )
        {
            const float *Cs = cur + size_t( jj ) * st;
            if ( wrap3_due( mlo + jj ) )
            {
                #pragma unroll
                for ( int r = 0; r < NPT; ++r )
                {
                    frame32_rewrap<MASK>( fx, xh[r], xl[r], uh[r], ul[r] );
                    frame32_rewrap<MASK>( fy, yh[r], yl[r], vh[r], vl[r] );
                    frame32_rewrap<MASK>( fz, zh[r], zl[r], wh[r], wl[r] );
                }
            }
            if ( CUBIC ) trace_slice( Cs, 1.0f, std::integral_constant<int, KICK_ONE>() );
            else trace_slice( Cs, ( mlo + jj == 0 ) ? 0.5f : 1.0f, std::integral_constant<int, KICK_GENERAL>() );
        }
        if ( jlast == 1 ) trace_slice( cur, 0.5f, std::integral_constant<int, KICK_HALF>() );     // wrap3_due(0) is false
        if ( p.dma ) dma_publish();
    }

    #pragma unroll
    for ( int r = 0; r < NPT; ++r )
    {
        double f = 0.0;
        if ( id[r].active )
        {
            const double x = p.x_min + ( double( xh[r] - fx.P0 ) + double( xl[r] ) ) * p.dx;   // hi - P0 exact
            const double y = p.y_min + ( double( yh[r] - fy.P0 ) + double( yl[r] ) ) * p.dy;
            const double z = p.z_min + ( double( zh[r] - fz.P0 ) + double( zl[r] ) ) * p.dz;
            f = f0_3d( p.f0, x, y, z, ( double( uh[r] - fx.W0 ) + double( ul[r] ) ) / cdx,
                       ( double( vh[r] - fy.W0 ) + double( vl[r] ) ) / cdy, ( double( wh[r] - fz.W0 ) + double( wl[r] ) ) / cdz );
        }
        write_partial3( p, id[r], f );
    }
}

// ---- strict FP64 (the reference's arithmetic, order 4) ----------------------
template <int ORDER, int DERIV>
__device__ __forceinline__ void basis_strict( double t, double out[ORDER] )
{
#pragma clang fp contract(off)
    double w[ORDER];
    w[ORDER-1] = 1.0;
    #pragma unroll
    for ( int m = 1; m < ORDER - DERIV; ++m )
    {
        w[ORDER-m-1] = ( 1.0 - t ) * w[ORDER-m];
        #pragma unroll
        for ( int s = 1 - m; s < 0; ++s )
            w[ORDER-1+s] = ( t - double(s) ) * w[ORDER-1+s] + ( double(m+1+s) - t ) * w[ORDER+s];
        w[ORDER-1] = w[ORDER-1] * t;
    }
    #pragma unroll
    for ( int j = DERIV - 1; j >= 0; --j )
    {
        w[j] = -w[j+1];
        #pragma unroll
        for ( int i = j + 1; i < ORDER - 1; ++i ) w[i] = w[i] - w[i+1];
    }
    double f = 1.0;
    #pragma unroll
    for ( int i = 2; i <= ORDER - DERIV - 1; ++i ) f *= double(i);
    const double inv = 1.0 / f;
    #pragma unroll
    for ( int i = 0; i < ORDER; ++i ) out[i] = w[i] * inv;
}

template <int ORDER, int DERIV>
__device__ __forceinline__ double deboor_strict( double t, double c[ORDER] )
{
#pragma clang fp contract(off)
    constexpr int n = ORDER, d = DERIV;
    #pragma unroll
    for ( int j = 1; j <= d; ++j )
        #pragma unroll
        for ( int i = n - 1; i >= j; --i ) c[i] = c[i] - c[i-1];
    #pragma unroll
    for ( int j = 1; j < n - d; ++j )
        #pragma unroll
        for ( int i = n - d - 1; i >= j; --i )
        {
            const double lo = ( ( ( t + double(n) ) - double(d) ) - 1.0 ) - double(i);
            const double hi = double( i - j + 1 ) - t;
            c[d+i] = lo * c[d+i] + hi * c[d+i-1];
        }
    double f = 1.0;
    #pragma unroll
    for ( int i = 2; i <= n - d - 1; ++i ) f *= double(i);
    return ( 1.0 / f ) * c[n-1];
}

template <int ORDER, int DX, int DY, int DZ>
__device__ __forceinline__ double field3_strict( double x, double y, double z, const double *C,
                                                 const trace3_params &p )
{
#pragma clang fp contract(off)
    x = x - p.x_min; y = y - p.y_min; z = z - p.z_min;
    x = x - p.Lx * floor( x * p.Lx_inv );
    y = y - p.Ly * floor( y * p.Ly_inv );
    z = z - p.Lz * floor( z * p.Lz_inv );
    const double xk = floor( x * p.dx_inv ), yk = floor( y * p.dy_inv ), zk = floor( z * p.dz_inv );
    int ii = int( xk ), jj = int( yk ), kk = int( zk );
    x = x * p.dx_inv - xk; y = y * p.dy_inv - yk; z = z * p.dz_inv - zk;
    ii = ( ii >= int( p.Nx ) ) ? ii - int( p.Nx ) : ii; ii = ( ii < 0 ) ? ii + int( p.Nx ) : ii;
    jj = ( jj >= int( p.Ny ) ) ? jj - int( p.Ny ) : jj; jj = ( jj < 0 ) ? jj + int( p.Ny ) : jj;
    kk = ( kk >= int( p.Nz ) ) ? kk - int( p.Nz ) : kk; kk = ( kk < 0 ) ? kk + int( p.Nz ) : kk;
    double factor = 1.0;
    if ( DX ) factor = factor * p.dx_inv;
    if ( DY ) factor = factor * p.dy_inv;
    if ( DZ ) factor = factor * p.dz_inv;

    const double *c = C + kk * p.sz + jj * p.sy + ii;
    double czy[ORDER*ORDER], cz[ORDER], B[ORDER];
    #pragma unroll
    for ( int i = 0; i < ORDER*ORDER; ++i ) czy[i] = 0.0;
    #pragma unroll
    for ( int i = 0; i < ORDER; ++i ) cz[i] = 0.0;
    basis_strict<ORDER,DX>( x, B );
    #pragma unroll
    for ( int k = 0; k < ORDER; ++k )
    #pragma unroll
    for ( int j = 0; j < ORDER; ++j )
    #pragma unroll
    for ( int i = 0; i < ORDER; ++i )
        czy[k*ORDER+j] = czy[k*ORDER+j] + c[k*p.sz + j*p.sy + i] * B[i];
    basis_strict<ORDER,DY>( y, B );
    #pragma unroll
    for ( int k = 0; k < ORDER; ++k )
    #pragma unroll
    for ( int j = 0; j < ORDER; ++j )
        cz[k] = cz[k] + czy[k*ORDER+j] * B[j];
    return factor * deboor_strict<ORDER,DZ>( z, cz );
}

template <int BLOCK, int MODE>
__global__ __launch_bounds__( BLOCK )
void trace3d_strict_kernel( trace3_params p )
{
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    constexpr int WPB = BLOCK / WAVE;
    const node3 id = make_node3<WPB,1>( p, 0 );
    uint32_t ix, iy, iz; double u, v, w;
    node_coords( p, id, ix, iy, iz, u, v, w );
    const double v2 = u*u + v*v + w*w;
    double x = p.x_min + ix * p.dx, y = p.y_min + iy * p.dy, z = p.z_min + iz * p.dz;
    const uint32_t st = p.st;
    const double half_dt = 0.5 * p.dt;

    for ( int64_t mhi = top_slice3<MODE>( p.n ); mhi >= 0; mhi -= p.chunk )
    {
        const int64_t mlo = mhi - int64_t( p.chunk ) + 1 > 0 ? mhi - int64_t( p.chunk ) + 1 : 0;
        const int cnt = int( mhi - mlo + 1 );
        __syncthreads();
        stage3<BLOCK>( lds, p.hist + size_t( mlo ) * st, uint32_t( cnt ) * st );
        __syncthreads();
        for ( int jj = cnt - 1; jj >= 0; --jj )
        {
            const double *C = lds + size_t( jj ) * st;
            const uint64_t m = uint64_t( mlo + jj );
            const bool initial = ( MODE == MODE_METRICS ) && m == p.n;
            if ( !initial ) { x = x - p.dt * u; y = y - p.dt * v; z = z - p.dt * w; }
            const double Ex = -field3_strict<4,1,0,0>( x, y, z, C, p );
            const double Ey = -field3_strict<4,0,1,0>( x, y, z, C, p );
            const double Ez = -field3_strict<4,0,0,1>( x, y, z, C, p );
            const double h = ( m == 0 || initial ) ? half_dt : p.dt;
            u = u + h * Ex; v = v + h * Ey; w = w + h * Ez;
        }
    }
    const double f = id.active ? f0_3d( p.f0, x, y, z, u, v, w ) : 0.0;
    if ( MODE == MODE_RHO ) write_partial3( p, id, f );
    else write_metrics3( p, id, f, v2 );
}

template <typename K>
hipError_t launch3( K kernel, uint32_t grid, int block, size_t lds, const trace3_params &p, hipStream_t st )
{
    if ( lds > 65536 )
    {
        hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( kernel ),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int( lds ) );
        if ( e != hipSuccess ) return e;
    }
    hipLaunchKernelGGL( kernel, dim3( grid ), dim3( block ), lds, st, p );
    return hipGetLastError();
}

template <int BLOCK>
hipError_t launch3_npt( const trace3_params &p, int variant, int npt, uint32_t grid, size_t lds, hipStream_t st )
{
    if ( p.mode == MODE_METRICS )
    {
        switch ( npt )
        {
        case 1: return launch3( trace3d_fast_kernel<BLOCK,1,MODE_METRICS>, grid, BLOCK, lds, p, st );
        case 2: return launch3( trace3d_fast_kernel<BLOCK,2,MODE_METRICS>, grid, BLOCK, lds, p, st );
        case 4: return launch3( trace3d_fast_kernel<BLOCK,4,MODE_METRICS>, grid, BLOCK, lds, p, st );
        default: return hipErrorInvalidValue;
        }
    }
    // power-of-two axes: the cell index wraps by a mask, positions every 16 slices
    const bool mask = ( p.Nx & ( p.Nx - 1 ) ) == 0 && ( p.Ny & ( p.Ny - 1 ) ) == 0 && ( p.Nz & ( p.Nz - 1 ) ) == 0;
    const bool g16 = mask && ( npt == 2 || npt == 4 || ( npt == 8 && variant == TRACE_FAST ) ) && p.Nx == 16 && p.Ny == 16 && p.Nz == 16 && p.sy == 19 && p.sz == 19 * 19;
    if ( variant == TRACE_FP32C )
    {
        const bool cubic = p.dx_inv == p.dy_inv && p.dx_inv == p.dz_inv;
        if ( g16 && cubic && npt == 4 ) return launch3( trace3d_fp32c_kernel<BLOCK,4,true,16,true>, grid, BLOCK, lds, p, st );
        if ( g16 && cubic ) return launch3( trace3d_fp32c_kernel<BLOCK,2,true,16,true>, grid, BLOCK, lds, p, st );
        switch ( npt * 2 + ( mask ? 1 : 0 ) )
        {
        case 2: return launch3( trace3d_fp32c_kernel<BLOCK,1,false>, grid, BLOCK, lds, p, st );
        case 3: return launch3( trace3d_fp32c_kernel<BLOCK,1,true>,  grid, BLOCK, lds, p, st );
        case 4: return launch3( trace3d_fp32c_kernel<BLOCK,2,false>, grid, BLOCK, lds, p, st );
        case 5: return launch3( trace3d_fp32c_kernel<BLOCK,2,true>,  grid, BLOCK, lds, p, st );
        case 8: return launch3( trace3d_fp32c_kernel<BLOCK,4,false>, grid, BLOCK, lds, p, st );
        case 9: return launch3( trace3d_fp32c_kernel<BLOCK,4,true>,  grid, BLOCK, lds, p, st );
        default: return hipErrorInvalidValue;
        }
    }
    // the 16^3 grid with its strides compiled in (two and four nodes a lane)
    if ( g16 && npt == 8 ) return launch3( trace3d_xm_fast_kernel<BLOCK,8,true,16>, grid, BLOCK, lds, p, st );
    if ( g16 && npt == 4 ) return launch3( trace3d_xm_fast_kernel<BLOCK,4,true,16>, grid, BLOCK, lds, p, st );
    if ( g16 ) return launch3( trace3d_xm_fast_kernel<BLOCK,2,true,16>, grid, BLOCK, lds, p, st );
    // eight nodes a lane exist for 16^3 only; elsewhere four (the x-major grid and the
    // partial sums do not depend on the nodes a lane holds, and 4 divides vgroup too)
    if ( npt == 8 ) npt = 4;
    switch ( npt * 2 + ( mask ? 1 : 0 ) )
    {
    case 2: return launch3( trace3d_xm_fast_kernel<BLOCK,1,false>, grid, BLOCK, lds, p, st );
    case 3: return launch3( trace3d_xm_fast_kernel<BLOCK,1,true>,  grid, BLOCK, lds, p, st );
    case 4: return launch3( trace3d_xm_fast_kernel<BLOCK,2,false>, grid, BLOCK, lds, p, st );
    case 5: return launch3( trace3d_xm_fast_kernel<BLOCK,2,true>,  grid, BLOCK, lds, p, st );
    case 8: return launch3( trace3d_xm_fast_kernel<BLOCK,4,false>, grid, BLOCK, lds, p, st );
    case 9: return launch3( trace3d_xm_fast_kernel<BLOCK,4,true>,  grid, BLOCK, lds, p, st );
    default: return hipErrorInvalidValue;
    }
}

} // namespace

bool trace3d_shape_ok( int variant, int order, const launch_shape &s )
{
    if ( order != 4 || s.chunk < 1 ) return false;
    if ( variant == TRACE_STRICT )
        return s.npt == 1 && ( s.block == 64 || s.block == 256 );
    return ( s.block == 256 || s.block == 512 || s.block == 1024 ) &&
           ( s.npt == 1 || s.npt == 2 || s.npt == 4 || ( s.npt == 8 && variant == TRACE_FAST ) );
}

static uint64_t visited3( const trace3_params &p )
{
    return p.mode == MODE_METRICS ? ( p.n > 0 ? p.n + 1 : 0 ) : p.n;
}

size_t trace3d_lds_bytes( const trace3_params &p, int variant, const launch_shape &s )
{
    const uint64_t v = visited3( p );
    const uint64_t c = v < uint64_t( s.chunk ) ? ( v > 0 ? v : 1 ) : uint64_t( s.chunk );
    if ( variant == TRACE_FP32C && p.mode == MODE_RHO )        // the buffers of the FP32 slices
        return ( fp32c3_dma( uint32_t( s.chunk ), p.st32 ) ? 2 : 1 ) * size_t( c ) * p.st32 * sizeof(float);
    if ( variant == TRACE_FAST && p.mode == MODE_RHO && fast3_dma( uint32_t( s.chunk ), p.st ) )
        return 2 * ( size_t( c ) * p.st + 2 ) * sizeof(double);
    return size_t( c ) * p.st * sizeof(double);
}

hipError_t launch_trace3d( const trace3_params &p_in, int variant, const launch_shape &s, hipStream_t st )
{
    if ( !trace3d_shape_ok( variant, int( p_in.order ), s ) ) return hipErrorInvalidValue;
    trace3_params p = p_in;
    uint32_t grid;
    if ( p.mode == MODE_RHO && variant == TRACE_FAST )
    {
        // x-major: spatial tiles x velocity-group blocks; NPT divides vgroup
        if ( p.n_count == 0 ) return hipSuccess;
        if ( p.vgroup == 0 || p.vgroup % uint32_t( s.npt ) != 0 ) return hipErrorInvalidValue;
        const uint32_t wpb = uint32_t( s.block / WAVE );
        p.n_st = ( p.n_count + WAVE - 1 ) / WAVE;
        grid = p.n_st * ( ( p.ppx + wpb - 1 ) / wpb );
    }
    else
    {
        if ( p.n_tasks == 0 ) return hipSuccess;
        const uint32_t tasks_per_block = uint32_t( s.block / WAVE ) * uint32_t( s.npt );
        grid = ( p.n_tasks + tasks_per_block - 1 ) / tasks_per_block;
    }
    p.chunk = uint32_t( s.chunk );
    const uint64_t v = visited3( p );
    if ( v < p.chunk ) p.chunk = v > 0 ? uint32_t( v ) : 1u;
    p.dma = p.mode == MODE_RHO && ( ( variant == TRACE_FP32C && fp32c3_dma( uint32_t( s.chunk ), p.st32 ) ) ||
                                    ( variant == TRACE_FAST && fast3_dma( uint32_t( s.chunk ), p.st ) ) ) ? 1u : 0u;
    const size_t lds = trace3d_lds_bytes( p, variant, s );
    if ( lds > 160u * 1024u ) return hipErrorInvalidValue;

    if ( variant == TRACE_STRICT )
    {
        if ( p.mode == MODE_METRICS )
        {
            if ( s.block == 64 ) return launch3( trace3d_strict_kernel<64,MODE_METRICS>, grid, 64, lds, p, st );
            return launch3( trace3d_strict_kernel<256,MODE_METRICS>, grid, 256, lds, p, st );
        }
        if ( s.block == 64 )  return launch3( trace3d_strict_kernel<64,MODE_RHO>,  grid, 64,  lds, p, st );
        return launch3( trace3d_strict_kernel<256,MODE_RHO>, grid, 256, lds, p, st );
    }
    switch ( s.block )
    {
    case 256:  return launch3_npt<256> ( p, variant, s.npt, grid, lds, st );
    case 512:  return launch3_npt<512> ( p, variant, s.npt, grid, lds, st );
    case 1024: return launch3_npt<1024>( p, variant, s.npt, grid, lds, st );
    default:   return hipErrorInvalidValue;
    }
}

} // namespace nufi_impl
