/*
 * dual.hpp — TEST INFRASTRUCTURE: forward-mode dual numbers (value + K
 * directional derivatives) so that the oracle's templated trace computes the
 * exact tangent of the discrete forward model.  floor and index conversion
 * act on the value (their derivative is zero almost everywhere).
 */
#ifndef NUFI_ORACLE_DUAL_HPP
#define NUFI_ORACLE_DUAL_HPP

#include <cmath>
#include <cstddef>

namespace oracle
{

template <int K>
struct dual
{
    double v;
    double d[K];
    constexpr dual(): v( 0 ), d {} {}
    constexpr dual( double x ): v( x ), d {} {}
    explicit operator double() const { return v; }
};

#define NUFI_DUAL_BIN( OP, VAL, DER )                                                    \
template <int K> inline dual<K> operator OP( const dual<K> &a, const dual<K> &b )        \
{ dual<K> r; r.v = VAL; for ( int i = 0; i < K; ++i ) r.d[i] = DER; return r; }          \
template <int K> inline dual<K> operator OP( const dual<K> &a, double bv )               \
{ return a OP dual<K>( bv ); }                                                           \
template <int K> inline dual<K> operator OP( double av, const dual<K> &b )               \
{ return dual<K>( av ) OP b; }

NUFI_DUAL_BIN( +, a.v + b.v, a.d[i] + b.d[i] )
NUFI_DUAL_BIN( -, a.v - b.v, a.d[i] - b.d[i] )
NUFI_DUAL_BIN( *, a.v * b.v, a.d[i] * b.v + a.v * b.d[i] )
NUFI_DUAL_BIN( /, a.v / b.v, ( a.d[i] * b.v - a.v * b.d[i] ) / ( b.v * b.v ) )
#undef NUFI_DUAL_BIN

template <int K> inline dual<K> operator-( const dual<K> &a )
{ dual<K> r; r.v = -a.v; for ( int i = 0; i < K; ++i ) r.d[i] = -a.d[i]; return r; }

template <int K> inline bool operator<( const dual<K> &a, const dual<K> &b ) { return a.v < b.v; }
template <int K> inline bool operator>( const dual<K> &a, const dual<K> &b ) { return a.v > b.v; }

template <int K> inline dual<K> floor( const dual<K> &a ) { return dual<K>( std::floor( a.v ) ); }
template <int K> inline dual<K> cos( const dual<K> &a )
{ dual<K> r( std::cos( a.v ) ); const double s = -std::sin( a.v ); for ( int i = 0; i < K; ++i ) r.d[i] = s * a.d[i]; return r; }
template <int K> inline dual<K> sin( const dual<K> &a )
{ dual<K> r( std::sin( a.v ) ); const double c = std::cos( a.v ); for ( int i = 0; i < K; ++i ) r.d[i] = c * a.d[i]; return r; }
template <int K> inline dual<K> exp( const dual<K> &a )
{ dual<K> r( std::exp( a.v ) ); for ( int i = 0; i < K; ++i ) r.d[i] = r.v * a.d[i]; return r; }

template <int K> inline size_t to_index( const dual<K> &x ) { return static_cast<size_t>( x.v ); }

} // namespace oracle

#endif
