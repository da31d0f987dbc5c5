// history_io.cpp — the coefficient history on disk (host only, no device).
//
// Binary: "NUFIHST1", a fixed little-endian header, the slices (f64), then
// the energies (f64).  Text: the layout of the reference's potential files,
// bin/test_nufi_cpu_3d_isolated.cpp:128-135 (seven header lines), 158-162
// (one value per line) and 190-211 (the reader skips the seven lines and
// reads (Nt+1)*stride_t values) — written with 17 significant digits so the
// values read back exactly.
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include <sys/types.h>

#include "../../include/nufi_hip.h"

namespace
{

const char MAGIC[8] = { 'N', 'U', 'F', 'I', 'H', 'S', 'T', '1' };

struct file_closer
{
    FILE *f;
    ~file_closer() { if ( f ) std::fclose( f ); }
};

uint64_t expected_stride( const nufi_history_info &h )
{
    const uint64_t o1 = uint64_t( h.order ) - 1;
    return ( h.Nx + o1 ) * ( h.dim == 3 ? ( h.Ny + o1 ) * ( h.Nz + o1 ) : 1 );
}

bool info_ok( const nufi_history_info &h )
{
    if ( h.dim != 1 && h.dim != 3 ) return false;
    if ( h.order < 1 || h.Nx == 0 || h.Ny == 0 || h.Nz == 0 ) return false;
    if ( h.dim == 1 && ( h.Ny != 1 || h.Nz != 1 ) ) return false;
    if ( h.n_slices > h.Nt + 1 || h.stride_t != expected_stride( h ) ) return false;
    return std::isfinite( h.dt ) && h.dt > 0;
}

bool put( FILE *f, const void *p, size_t n ) { return std::fwrite( p, 1, n, f ) == n; }
bool get( FILE *f, void *p, size_t n ) { return std::fread( p, 1, n, f ) == n; }

// (The build targets little-endian x86-64 hosts; fields are written as stored.)
bool write_header( FILE *f, const nufi_history_info &h )
{
    const uint32_t dim = uint32_t( h.dim ), order = uint32_t( h.order );
    return put( f, MAGIC, 8 ) && put( f, &dim, 4 ) && put( f, &order, 4 ) && put( f, &h.Nx, 8 ) &&
           put( f, &h.Ny, 8 ) && put( f, &h.Nz, 8 ) && put( f, &h.Nt, 8 ) && put( f, &h.n_slices, 8 ) &&
           put( f, &h.stride_t, 8 ) && put( f, &h.dt, 8 ) && put( f, &h.n_energies, 8 );
}

bool read_header( FILE *f, nufi_history_info &h )
{
    uint32_t dim = 0, order = 0;
    if ( !( get( f, &dim, 4 ) && get( f, &order, 4 ) && get( f, &h.Nx, 8 ) && get( f, &h.Ny, 8 ) &&
            get( f, &h.Nz, 8 ) && get( f, &h.Nt, 8 ) && get( f, &h.n_slices, 8 ) && get( f, &h.stride_t, 8 ) &&
            get( f, &h.dt, 8 ) && get( f, &h.n_energies, 8 ) ) )
        return false;
    h.dim = int( dim ); h.order = int( order );
    return true;
}

// "<key> = <value>" (std::to_string of the reference's header fields)
bool text_field( FILE *f, const char *key, std::string &value )
{
    char line[256];
    if ( !std::fgets( line, sizeof line, f ) ) return false;
    const size_t kl = std::strlen( key );
    if ( std::strncmp( line, key, kl ) != 0 || std::strncmp( line + kl, " = ", 3 ) != 0 ) return false;
    value.assign( line + kl + 3 );
    while ( !value.empty() && ( value.back() == '\n' || value.back() == '\r' ) ) value.pop_back();
    return !value.empty();
}

bool parse_u64( const std::string &s, uint64_t &v )
{
    char *end = nullptr;
    errno = 0;
    const unsigned long long x = std::strtoull( s.c_str(), &end, 10 );
    if ( errno || end == s.c_str() || *end != '\0' ) return false;
    v = uint64_t( x );
    return true;
}

bool parse_f64( const std::string &s, double &v )
{
    char *end = nullptr;
    errno = 0;
    v = std::strtod( s.c_str(), &end );
    return !errno && end != s.c_str() && *end == '\0';
}

// The seven header lines; the slice count is whatever the file holds, counted
// by the caller (the reference writes all Nt+1 slices).
bool read_text_header( FILE *f, nufi_history_info &h )
{
    std::string v;
    uint64_t order = 0;
    if ( !text_field( f, "Nt", v ) || !parse_u64( v, h.Nt ) ) return false;
    if ( !text_field( f, "dt", v ) || !parse_f64( v, h.dt ) ) return false;
    if ( !text_field( f, "Nx", v ) || !parse_u64( v, h.Nx ) ) return false;
    if ( !text_field( f, "Ny", v ) || !parse_u64( v, h.Ny ) ) return false;
    if ( !text_field( f, "Nz", v ) || !parse_u64( v, h.Nz ) ) return false;
    if ( !text_field( f, "order", v ) || !parse_u64( v, order ) ) return false;
    char blank[8];
    if ( !std::fgets( blank, sizeof blank, f ) ) return false;
    h.order = int( order );
    h.dim = ( h.Ny == 1 && h.Nz == 1 ) ? 1 : 3;
    h.stride_t = expected_stride( h );
    h.n_energies = 0;
    return true;
}

int detect( FILE *f, int &format )
{
    char head[8];
    if ( !get( f, head, 8 ) ) return NUFI_ERR_IO;
    if ( std::memcmp( head, MAGIC, 8 ) == 0 ) { format = NUFI_HISTORY_BINARY; return NUFI_OK; }
    if ( std::memcmp( head, "Nt = ", 5 ) == 0 )
    {
        format = NUFI_HISTORY_TEXT;
        std::rewind( f );
        return NUFI_OK;
    }
    return NUFI_ERR_IO;
}

// Header of a text file and the number of whole slices it holds.
int text_info( FILE *f, nufi_history_info &h )
{
    if ( !read_text_header( f, h ) ) return NUFI_ERR_IO;
    const off_t body = ftello( f );
    uint64_t values = 0;
    char line[128];
    while ( std::fgets( line, sizeof line, f ) )
        if ( line[0] != '\n' && line[0] != '\r' ) ++values;
    h.n_slices = h.stride_t ? values / h.stride_t : 0;
    if ( h.n_slices > h.Nt + 1 ) h.n_slices = h.Nt + 1;
    if ( fseeko( f, body, SEEK_SET ) != 0 ) return NUFI_ERR_IO;
    return info_ok( h ) ? NUFI_OK : NUFI_ERR_IO;
}

} // namespace

extern "C" {

int nufi_history_write( const char *path, int format, const nufi_history_info *info, const double *slices,
                        const double *energies )
{
    if ( path == nullptr || info == nullptr || !info_ok( *info ) ) return NUFI_ERR_INVALID;
    if ( info->n_slices > 0 && slices == nullptr ) return NUFI_ERR_INVALID;
    if ( format == NUFI_HISTORY_BINARY && info->n_energies > 0 && energies == nullptr ) return NUFI_ERR_INVALID;
    if ( format != NUFI_HISTORY_BINARY && format != NUFI_HISTORY_TEXT ) return NUFI_ERR_INVALID;
    file_closer fc { std::fopen( path, format == NUFI_HISTORY_BINARY ? "wb" : "w" ) };
    if ( fc.f == nullptr ) return NUFI_ERR_IO;
    const uint64_t count = info->n_slices * info->stride_t;
    if ( format == NUFI_HISTORY_BINARY )
    {
        if ( !write_header( fc.f, *info ) ) return NUFI_ERR_IO;
        if ( count && std::fwrite( slices, sizeof(double), count, fc.f ) != count ) return NUFI_ERR_IO;
        if ( info->n_energies &&
             std::fwrite( energies, sizeof(double), info->n_energies, fc.f ) != info->n_energies )
            return NUFI_ERR_IO;
    }
    else
    {
        // std::to_string( dt ) prints six decimals (the reference's header);
        // the exact dt is not needed to read the values back
        const std::string dt = std::to_string( info->dt );
        if ( std::fprintf( fc.f, "Nt = %" PRIu64 "\ndt = %s\nNx = %" PRIu64 "\nNy = %" PRIu64 "\nNz = %" PRIu64
                                 "\norder = %d\n\n",
                           info->Nt, dt.c_str(), info->Nx, info->Ny, info->Nz, info->order ) < 0 )
            return NUFI_ERR_IO;
        for ( uint64_t i = 0; i < count; ++i )
            if ( std::fprintf( fc.f, "%.17g\n", slices[i] ) < 0 ) return NUFI_ERR_IO;
    }
    if ( std::fflush( fc.f ) != 0 ) return NUFI_ERR_IO;
    return NUFI_OK;
}

int nufi_history_read_info( const char *path, nufi_history_info *info, int *format )
{
    if ( path == nullptr || info == nullptr ) return NUFI_ERR_INVALID;
    file_closer fc { std::fopen( path, "rb" ) };
    if ( fc.f == nullptr ) return NUFI_ERR_IO;
    int fmt = 0;
    if ( int rc = detect( fc.f, fmt ) ) return rc;
    nufi_history_info h {};
    if ( fmt == NUFI_HISTORY_BINARY )
    {
        if ( !read_header( fc.f, h ) || !info_ok( h ) ) return NUFI_ERR_IO;
    }
    else if ( int rc = text_info( fc.f, h ) ) return rc;
    *info = h;
    if ( format ) *format = fmt;
    return NUFI_OK;
}

int nufi_history_read( const char *path, nufi_history_info *info, double *slices, uint64_t slice_capacity,
                       double *energies, uint64_t energy_capacity )
{
    if ( path == nullptr || info == nullptr ) return NUFI_ERR_INVALID;
    file_closer fc { std::fopen( path, "rb" ) };
    if ( fc.f == nullptr ) return NUFI_ERR_IO;
    int fmt = 0;
    if ( int rc = detect( fc.f, fmt ) ) return rc;
    nufi_history_info h {};
    if ( fmt == NUFI_HISTORY_BINARY )
    {
        if ( !read_header( fc.f, h ) || !info_ok( h ) ) return NUFI_ERR_IO;
    }
    else if ( int rc = text_info( fc.f, h ) ) return rc;
    const uint64_t ns = h.n_slices < slice_capacity ? h.n_slices : slice_capacity;
    const uint64_t count = ns * h.stride_t;
    if ( count && slices == nullptr ) return NUFI_ERR_INVALID;
    if ( fmt == NUFI_HISTORY_BINARY )
    {
        if ( count && std::fread( slices, sizeof(double), count, fc.f ) != count ) return NUFI_ERR_IO;
        const uint64_t ne = h.n_energies < energy_capacity ? h.n_energies : energy_capacity;
        if ( ne )
        {
            if ( energies == nullptr ) return NUFI_ERR_INVALID;
            if ( fseeko( fc.f, off_t( ( h.n_slices - ns ) * h.stride_t * sizeof(double) ), SEEK_CUR ) != 0 )
                return NUFI_ERR_IO;
            if ( std::fread( energies, sizeof(double), ne, fc.f ) != ne ) return NUFI_ERR_IO;
        }
    }
    else
    {
        char line[128];
        uint64_t i = 0;
        while ( i < count && std::fgets( line, sizeof line, fc.f ) )
        {
            if ( line[0] == '\n' || line[0] == '\r' ) continue;
            char *end = nullptr;
            slices[i] = std::strtod( line, &end );
            if ( end == line ) return NUFI_ERR_IO;
            ++i;
        }
        if ( i != count ) return NUFI_ERR_IO;
    }
    *info = h;
    return NUFI_OK;
}

} // extern "C"
