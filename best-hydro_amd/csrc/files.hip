// BEST's files (host code): the mesh it reads and the .be file it writes.
//
//   e_edit / f_edit / i_edit   Ew.d, Fw.d and Iw output editing (best/output.f's FORMATs 340, 350,
//                              352, 403-416, 420, 401, 402, 410, 182, 134, 980, 990-999; the mesh
//                              file's 9e17.8, rcoal/scoalesce.f:895-899)
//   format_be                  the records in OUTPUT's order (best/output.f:36-168); list-directed
//                              records (WRITE(9,*)) start with a blank
//   bhg_format_be, bhg_write_be   the C ABI of the .be file
//   read_best                  the mesh as best/best.f:189-220 reads it: list-directed READs, one
//                              for the header (NAR, and TRUEAREA TRUEVOL with the area correction)
//                              and one per triangle (9 values); each READ starts on a new line, may
//                              run over several and skips the rest of its last
//   format_best                the mesh as rcoal writes it (rcoal/scoalesce.f:891-899): the
//                              list-directed header, then one 9e17.8 line per triangle
//   bhg_read_best, bhg_write_best   their C ABI
//
// The eigenvalues, traces and anisotropies are computed here from the tensors with JACOBI
// (jacobi3) and the arithmetic of best/output.f:99-106 and :121-128, as OUTPUT does; non-finite
// tensors or viscosity values are refused with best.f's message (best/best.f:483-496,
// best/compon.f90:214-226) and nothing is written.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "best_hip.h"
#include "files.h"
#include "transport.h"

namespace bhg {
namespace {

std::string fit( const std::string& t, int w ) {
   if ( (int)t.size() > w ) return std::string( w, '*' );
   return std::string( w - t.size(), ' ' ) + t;
}

// Ew.d: [-]0.x1...xd E+zz (|exponent| <= 99) or [-]0.x1...xd+zzz (<= 999), right-justified; the
// leading zero dropped when the field has no room for it; Inf and NaN right-justified.
std::string e_edit( double x, int w, int d ) {
   if ( std::isnan( x ) ) return fit( "NaN", w );
   const bool neg = std::signbit( x );
   if ( std::isinf( x ) ) return fit( neg ? "-Inf" : "Inf", w );
   std::string digits;
   int e = 0;
   if ( x == 0.0 ) {
      digits.assign( d, '0' );
   } else {
      // d significant digits, correctly rounded: m.mmmme+XX, and the exponent of 0.mmmmm is XX + 1
      char buf[ 64 ];
      std::snprintf( buf, sizeof buf, "%.*e", d - 1, std::fabs( x ) );
      const char* p = std::strchr( buf, 'e' );
      for ( const char* q = buf; q < p; ++q )
         if ( *q != '.' ) digits.push_back( *q );
      e = std::atoi( p + 1 ) + 1;
   }
   char ex[ 8 ];
   if ( std::abs( e ) <= 99 ) std::snprintf( ex, sizeof ex, "E%c%02d", e < 0 ? '-' : '+', std::abs( e ) );
   else if ( std::abs( e ) <= 999 ) std::snprintf( ex, sizeof ex, "%c%03d", e < 0 ? '-' : '+', std::abs( e ) );
   else return std::string( w, '*' );
   const std::string sign = neg ? "-" : "";
   std::string t = sign + "0." + digits + ex;
   if ( (int)t.size() > w ) t = sign + "." + digits + ex;
   return fit( t, w );
}

// Fw.d: [-]i.f...f rounded to d decimals, right-justified; the zero before the point dropped
// when the field has no room for it.
std::string f_edit( double x, int w, int d ) {
   char buf[ 400 ];
   std::snprintf( buf, sizeof buf, "%.*f", d, std::fabs( x ) );
   const std::string sign = std::signbit( x ) ? "-" : "";
   std::string t = sign + buf;
   if ( (int)t.size() > w && buf[ 0 ] == '0' && buf[ 1 ] == '.' ) t = sign + ( buf + 1 );
   return fit( t, w );
}

std::string i_edit( long n, int w ) { return fit( std::to_string( n ), w ); }

bool finite_all( const double* a, int n ) {
   for ( int i = 0; i < n; ++i )
      if ( !std::isfinite( a[ i ] ) ) return false;
   return true;
}

std::string format_be( const bhg_result& r, const bhg_be_params& p ) {
   if ( !( finite_all( r.dtt, 9 ) && finite_all( r.dtr, 9 ) && finite_all( r.drr, 9 ) && finite_all( r.urc, 3 ) &&
           finite_all( r.udr, 3 ) ) )
      throw std::runtime_error( "the diffusion tensors are not finite (NaN or Inf); nothing was written to the .be file" );
   if ( p.ivisco && !( std::isfinite( r.chiprime ) && finite_all( r.dvisc, 3 ) ) )
      throw std::runtime_error( "the viscosity factor are not finite (NaN or Inf); nothing was written to the .be file" );
   std::string o;
   auto rec = [&]( const std::string& s ) { o += s, o += '\n'; };
   auto ld = [&]( const std::string& s ) { rec( " " + s ); };  // list-directed record
   auto e125 = [&]( double x ) { return e_edit( x, 12, 5 ); };
   auto f340 = [&]( double a, double b, double c ) { rec( "   " + e125( a ) + "   " + e125( b ) + "   " + e125( c ) ); };
   auto tensor = [&]( const double* t ) {  // rows j: t(j,1), t(j,2), t(j,3), column-major storage
      for ( int j = 0; j < 3; ++j ) f340( t[ j ], t[ j + 3 ], t[ j + 6 ] );
   };
   const double temperature = 293.15, eta = 0.01002, pi = 3.141592653;
   const double tunit = ( 1.38066E-16 ) * temperature * 1.0E8 / ( 8.0 * eta * pi );
   const double runit = tunit * 1.0e16;

   std::string ver = p.ver ? p.ver : "";
   ver.resize( 16, ' ' );
   ld( "BEST Ver. " + ver );
   rec( "" ), rec( " BE - MACROMOLECULE TRANSPORT TENSORS" ), rec( "" );
   rec( "     Area (A^2)            " + e125( r.totarea ) );
   if ( p.iarea == 1 ) {
      rec( "     True Area (A^2)       " + e125( p.truearea ) );
      rec( "     Areafactor                 " + f_edit( r.afac, 7, 4 ) );
      rec( "     True Volume (A^3)     " + e125( p.truevol ) );
   }
   if ( p.mw == 1 ) rec( "     Molecular Weight (g/mol)  " + f_edit( p.wm, 8, 1 ) );
   rec( "     Temperature (K)           " + f_edit( temperature, 8, 3 ) );
   rec( "     Solvent Viscosity (Poise) " + f_edit( eta, 8, 5 ) );
   rec( "     Number of Surface Patches    " + i_edit( p.nar, 5 ) );
   ld( " " );
   if ( p.cg == 0 ) ld( " PRE-AVERAGING USED FOR THIS CALCULATION" );
   ld( " " );
   ld( " Centroid Coordinates (A)" );
   f340( r.centroid[ 0 ], r.centroid[ 1 ], r.centroid[ 2 ] );
   ld( " Vector between Center of Resistance/Centroid " );
   f340( r.urc[ 0 ], r.urc[ 1 ], r.urc[ 2 ] );
   ld( " Vector between Centers of Diffusion/Resistance" );
   f340( r.udr[ 0 ], r.udr[ 1 ], r.udr[ 2 ] );
   ld( " " ), ld( " Diffusion Tensors in Center of Mobility" ), ld( " " ), ld( " Dtt Tensor (1/A)" );
   tensor( r.dtt );
   ld( " " ), ld( " Symmetric Dtr Tensor (1/A^2)" );
   tensor( r.dtr );
   ld( " " ), ld( "  Drr TENSOR (1/A^3)" );
   tensor( r.drr );

   // eigenvalues, 1/3 trace and anisotropy as best/output.f:99-106
   auto eig = [&]( const double* t, double d[ 3 ], double& trace, double& del ) {
      jacobi3( t, d );
      trace = ( d[ 0 ] + d[ 1 ] + d[ 2 ] ) / 3.;
      del = std::sqrt(
         std::fabs( d[ 0 ] * d[ 0 ] + d[ 1 ] * d[ 1 ] + d[ 2 ] * d[ 2 ] - d[ 0 ] * d[ 1 ] - d[ 0 ] * d[ 2 ] - d[ 1 ] * d[ 2 ] ) );
   };
   double d[ 3 ], trace, del;
   eig( r.drr, d, trace, del );
   ld( " " ), ld( "EIGENVALUES OF Drr TENSOR (1/A^3)" );
   f340( d[ 0 ], d[ 1 ], d[ 2 ] );
   rec( "  1/3 Trace = " + e125( trace ) );
   rec( "  Anisotropy=" + e125( del ) );
   ld( " " ), ld( "EIGENVALUES OF Drr TENSOR (1/s)" );
   f340( runit * d[ 0 ], runit * d[ 1 ], runit * d[ 2 ] );
   rec( "  1/3 Trace DRR(EXACT) =" + e_edit( runit * trace, 15, 5 ) + " 1/s" ), rec( "" );
   eig( r.dtt, d, trace, del );
   ld( "EIGENVALUES OF Dtt TENSOR (1/A)" );
   f340( d[ 0 ], d[ 1 ], d[ 2 ] );
   rec( "  1/3 Trace = " + e125( trace ) );
   rec( "  Anisotropy=" + e125( del ) );
   ld( " " ), ld( "EIGENVALUES OF Dtt TENSOR (cm^2/s)" );
   f340( tunit * d[ 0 ], tunit * d[ 1 ], tunit * d[ 2 ] );
   rec( "  1/3 Trace DTT(EXACT) =" + e_edit( tunit * trace, 15, 5 ) + " cm^2/s" ), rec( "" );

   if ( p.ivisco ) {
      const std::string what = p.mw == 1 ? "INTRINSIC VISCOSITY" : "VISCOSITY FACTOR";
      if ( p.icov == 1 ) {
         ld( what + " AT THE CENTER OF VISCOSITY" );
         ld( " Vector between Center of Viscosity/Centroid" );
         rec( "     Dx = " + e125( r.dvisc[ 0 ] ) + "  Dy = " + e125( r.dvisc[ 1 ] ) + "  Dz = " + e125( r.dvisc[ 2 ] ) );
      } else {
         ld( what + " AT THE CENTROID" );
      }
      rec( "" );
      if ( p.iarea == 1 ) rec( "     CHI               " + e_edit( r.chiprime / p.truevol, 16, 6 ) );
      rec( "     CHI*VOLUME (A^3)  " + e_edit( r.chiprime, 16, 6 ) ), rec( "" );
      if ( p.mw == 1 ) {
         rec( "     ETA (cm^3/g)      " + e_edit( r.chiprime * 0.602214 / p.wm, 16, 6 ) );
         rec( "" );
      }
   }
   return o;
}

// ------------------------------------------------------------------ the mesh file

// The list-directed READs of best/best.f:189-220 over the file's lines.  A READ takes its values
// from the current line and the lines after it as needed, then the next READ starts on the line
// after the one its last value came from.  Values are separated by blanks, tabs or commas; r*v
// repeats v r times; reals may carry an E, D or Q exponent or a signed exponent with no letter.
class ListReader {
 public:
   ListReader( const std::string& path ) : path_( path ), in_( path ) {
      if ( !in_ ) throw std::runtime_error( "cannot open " + path );
   }
   // one READ of k values
   std::vector<std::string> read( int k ) {
      std::vector<std::string> out;
      while ( (int)out.size() < k ) {
         if ( pos_ >= tok_.size() ) {  // the READ goes on to the next line
            if ( !next_line() ) {
               std::ostringstream m;
               m << path_ << ": end of file at line " << line_ << " (" << out.size() << " of " << k << " values read)";
               throw std::runtime_error( m.str() );
            }
            continue;
         }
         std::string t = tok_[ pos_++ ];
         const size_t star = t.find( '*' );
         if ( star != std::string::npos ) {
            const long r = std::strtol( t.substr( 0, star ).c_str(), nullptr, 10 );
            const std::string v = t.substr( star + 1 );
            if ( r <= 0 || v.empty() ) bad( t );
            for ( long j = 0; j < r && (int)out.size() < k; ++j ) out.push_back( v );
         } else {
            out.push_back( t );
         }
      }
      // the rest of this line is skipped: the next READ starts on a new line
      pos_ = tok_.size();
      return out;
   }
   [[noreturn]] void bad( const std::string& t ) const {
      std::ostringstream m;
      m << path_ << ", line " << line_ << ": cannot read \"" << t << "\"";
      throw std::runtime_error( m.str() );
   }
   long as_int( const std::string& t ) const {
      char* e = nullptr;
      errno = 0;
      const long v = std::strtol( t.c_str(), &e, 10 );
      if ( errno || e == t.c_str() || *e ) bad( t );
      return v;
   }
   double as_real( const std::string& t ) const {
      // Fortran real input: the exponent letter may be D or Q, or absent before a signed exponent
      std::string u;
      for ( size_t j = 0; j < t.size(); ++j ) {
         char c = t[ j ];
         if ( c == 'D' || c == 'd' || c == 'Q' || c == 'q' ) c = 'E';
         if ( ( c == '+' || c == '-' ) && j > 0 && ( std::isdigit( (unsigned char)t[ j - 1 ] ) || t[ j - 1 ] == '.' ) ) u += 'E';
         u += c;
      }
      char* e = nullptr;
      errno = 0;
      const double v = std::strtod( u.c_str(), &e );  // correctly rounded, as the Fortran runtimes read
      if ( e == u.c_str() || *e ) bad( t );
      return v;
   }

 private:
   bool next_line() {
      std::string s;
      if ( !std::getline( in_, s ) ) return false;
      ++line_;
      tok_.clear();
      pos_ = 0;
      std::string cur;
      for ( char c : s ) {
         if ( c == ' ' || c == '\t' || c == ',' || c == '\r' ) {
            if ( !cur.empty() ) tok_.push_back( cur ), cur.clear();
         } else {
            cur += c;
         }
      }
      if ( !cur.empty() ) tok_.push_back( cur );
      return true;
   }
   std::string path_;
   std::ifstream in_;
   std::vector<std::string> tok_;
   size_t pos_ = 0;
   long line_ = 0;
};

// The header of rcoal's files: NAR as list-directed output writes an integer (a blank, then I12),
// then TRUEAREA and TRUEVOL each in a field of 26 characters: 16 significant digits in F form
// right-justified in 21 and 5 blanks for magnitudes from 0.1 to 1e16, else 0.d...dE+xx with 16
// digits right-justified in 26 (the layout of the reference's own mesh files, e.g.
// examples/best/2eu7_02562).
std::string header_real( double x ) {
   const double a = std::fabs( x );
   if ( std::isfinite( x ) && a >= 0.1 && a < 1e16 ) {
      char buf[ 64 ];
      std::snprintf( buf, sizeof buf, "%.15e", a );
      // digits before the point after rounding to 16 significant digits (0 below 1)
      const int k = std::atoi( std::strchr( buf, 'e' ) + 1 ) + 1;
      std::snprintf( buf, sizeof buf, "%.*f", 16 - k, a );
      std::string t = std::string( std::signbit( x ) ? "-" : "" ) + buf;
      return fit( t, 21 ) + "     ";
   }
   return e_edit( x, 26, 16 );
}

std::string format_best( int nar, double truearea, double truevol, const double* v1, const double* v2, const double* v3 ) {
   std::string o = " " + i_edit( nar, 12 ) + header_real( truearea ) + header_real( truevol ) + "\n";
   o.reserve( o.size() + (size_t)nar * 154 );
   const double* v[ 3 ] = { v1, v2, v3 };
   for ( int t = 0; t < nar; ++t ) {
      for ( int k = 0; k < 3; ++k )
         for ( int c = 0; c < 3; ++c ) o += e_edit( v[ k ][ (size_t)c * nar + t ], 17, 8 );
      o += '\n';
   }
   return o;
}

}  // namespace

BestMesh read_best( const std::string& path, int iarea ) {
   ListReader r( path );
   BestMesh m;
   const std::vector<std::string> h = r.read( iarea == 1 ? 3 : 1 );
   const long nar = r.as_int( h[ 0 ] );
   if ( nar < 1 || nar > ( 1L << 28 ) ) throw std::runtime_error( path + ": NAR = " + h[ 0 ] + " is not a number of triangles" );
   m.nar = (int)nar;
   if ( iarea == 1 ) m.truearea = r.as_real( h[ 1 ] ), m.truevol = r.as_real( h[ 2 ] );
   m.v1.resize( 3 * (size_t)nar ), m.v2.resize( 3 * (size_t)nar ), m.v3.resize( 3 * (size_t)nar );
   for ( long t = 0; t < nar; ++t ) {
      const std::vector<std::string> v = r.read( 9 );
      for ( int c = 0; c < 3; ++c ) {
         m.v1[ (size_t)c * nar + t ] = r.as_real( v[ c ] );
         m.v2[ (size_t)c * nar + t ] = r.as_real( v[ 3 + c ] );
         m.v3[ (size_t)c * nar + t ] = r.as_real( v[ 6 + c ] );
      }
   }
   return m;
}

}  // namespace bhg

extern "C" {

long bhg_format_be( const bhg_result* r, const bhg_be_params* p, char* buf, long len, char* err, int errlen ) {
   try {
      if ( !r || !p ) throw std::runtime_error( "bhg_format_be: no result or parameters" );
      const std::string s = bhg::format_be( *r, *p );
      if ( buf && len > 0 ) {
         const long k = std::min<long>( (long)s.size(), len - 1 );
         std::memcpy( buf, s.data(), k );
         buf[ k ] = '\0';
      }
      return (long)s.size();
   } catch ( const std::exception& e ) {
      if ( err && errlen > 0 ) std::snprintf( err, errlen, "%s", e.what() );
      return -1;
   }
}

int bhg_write_be( const char* path, const bhg_result* r, const bhg_be_params* p, char* err, int errlen ) {
   try {
      if ( !path || !r || !p ) throw std::runtime_error( "bhg_write_be: no path, result or parameters" );
      const std::string s = bhg::format_be( *r, *p );
      FILE* f = std::fopen( path, "wb" );
      if ( !f ) throw std::runtime_error( std::string( "cannot open " ) + path );
      const size_t w = std::fwrite( s.data(), 1, s.size(), f );
      if ( std::fclose( f ) != 0 || w != s.size() ) throw std::runtime_error( std::string( "cannot write " ) + path );
      return 0;
   } catch ( const std::exception& e ) {
      if ( err && errlen > 0 ) std::snprintf( err, errlen, "%s", e.what() );
      return 1;
   }
}

int bhg_read_best( const char* path, int iarea, int cap, int* nar, double* truearea, double* truevol, double* v1, double* v2,
                   double* v3, char* err, int errlen ) {
   try {
      if ( !path || !nar ) throw std::runtime_error( "bhg_read_best: no path" );
      const bhg::BestMesh m = bhg::read_best( path, iarea );
      *nar = m.nar;
      if ( truearea ) *truearea = m.truearea;
      if ( truevol ) *truevol = m.truevol;
      if ( !v1 || !v2 || !v3 ) return 0;  // sizing call
      if ( cap < m.nar ) throw std::runtime_error( "bhg_read_best: room for " + std::to_string( cap ) + " triangles, the file has " +
                                                    std::to_string( m.nar ) );
      const size_t b = 3 * (size_t)m.nar * sizeof( double );
      std::memcpy( v1, m.v1.data(), b ), std::memcpy( v2, m.v2.data(), b ), std::memcpy( v3, m.v3.data(), b );
      return 0;
   } catch ( const std::exception& e ) {
      if ( err && errlen > 0 ) std::snprintf( err, errlen, "%s", e.what() );
      return 1;
   }
}

int bhg_write_best( const char* path, int nar, double truearea, double truevol, const double* v1, const double* v2,
                    const double* v3, char* err, int errlen ) {
   try {
      if ( !path || nar < 1 || !v1 || !v2 || !v3 ) throw std::runtime_error( "bhg_write_best: no path, triangles or vertices" );
      const std::string s = bhg::format_best( nar, truearea, truevol, v1, v2, v3 );
      FILE* f = std::fopen( path, "wb" );
      if ( !f ) throw std::runtime_error( std::string( "cannot open " ) + path );
      const size_t w = std::fwrite( s.data(), 1, s.size(), f );
      if ( std::fclose( f ) != 0 || w != s.size() ) throw std::runtime_error( std::string( "cannot write " ) + path );
      return 0;
   } catch ( const std::exception& e ) {
      if ( err && errlen > 0 ) std::snprintf( err, errlen, "%s", e.what() );
      return 1;
   }
}

}  // extern "C"
