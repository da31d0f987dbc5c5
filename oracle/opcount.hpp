/*
 * opcount.hpp — TEST INFRASTRUCTURE: a scalar for the oracle's templates that
 * counts the arithmetic of one characteristic step in the reference's own
 * formulation (nufi_oracle.hpp's field1d / efield3d, the restatement of
 * fields.hpp:36-61, 308-350 and splines.hpp:81-172).
 *
 * A value is step-variant (positions, velocities, the coefficients a step
 * loads, and whatever is computed from them) or invariant (literals and the
 * grid's configuration values).  +, -, *, / and floor count 1 when an operand
 * is step-variant; operations among invariants are hoistable and count 0, as
 * do sign flips, comparisons and index conversions.
 */
#ifndef NUFI_ORACLE_OPCOUNT_HPP
#define NUFI_ORACLE_OPCOUNT_HPP

#include <cmath>
#include <cstddef>

namespace oracle
{

struct opcount
{
    double v;
    bool var;
    static inline long n = 0;

    opcount(): v( 0 ), var( false ) {}
    opcount( double x ): v( x ), var( false ) {}
    static opcount variable( double x ) { opcount r( x ); r.var = true; return r; }
    explicit operator double() const { return v; }
};

inline opcount counted( double v, bool var )
{
    opcount r( v );
    r.var = var;
    if ( var ) ++opcount::n;
    return r;
}

#define NUFI_OPCOUNT_BIN( OP )                                                                       \
inline opcount operator OP( const opcount &a, const opcount &b ) { return counted( a.v OP b.v, a.var || b.var ); } \
inline opcount operator OP( const opcount &a, double b ) { return a OP opcount( b ); }              \
inline opcount operator OP( double a, const opcount &b ) { return opcount( a ) OP b; }

NUFI_OPCOUNT_BIN( + )
NUFI_OPCOUNT_BIN( - )
NUFI_OPCOUNT_BIN( * )
NUFI_OPCOUNT_BIN( / )
#undef NUFI_OPCOUNT_BIN

inline opcount operator-( const opcount &a ) { opcount r( -a.v ); r.var = a.var; return r; }
inline bool operator<( const opcount &a, const opcount &b ) { return a.v < b.v; }
inline bool operator>( const opcount &a, const opcount &b ) { return a.v > b.v; }
inline opcount floor( const opcount &a ) { return counted( std::floor( a.v ), a.var ); }
inline size_t to_index( const opcount &x ) { return static_cast<size_t>( x.v ); }

} // namespace oracle

#endif
