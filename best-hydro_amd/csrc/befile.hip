// The .be file: the records BEST's OUTPUT writes (best/output.f:36-168) from one mesh solve's
// results, with the Fortran edit descriptors they use written out here (host code).
//
//   e_edit / f_edit / i_edit   Ew.d, Fw.d and Iw output editing (best/output.f's FORMATs 340, 350,
//                              352, 403-416, 420, 401, 402, 410, 182, 134, 980, 990-999)
//   format_be                  the records in OUTPUT's order; list-directed records (WRITE(9,*))
//                              start with a blank
//   bhg_format_be, bhg_write_be   the C ABI
//
// The eigenvalues, traces and anisotropies are computed here from the tensors with JACOBI
// (jacobi3) and the arithmetic of best/output.f:99-106 and :121-128, as OUTPUT does; non-finite
// tensors or viscosity values are refused with best.f's message (best/best.f:483-496,
// best/compon.f90:214-226) and nothing is written.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "best_hip.h"
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

}  // namespace
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

}  // extern "C"
