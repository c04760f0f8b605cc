// The off-diagonal 3x3 block of BEST's GR for one source triangle and one collocation point, as
// a device function shared by the dense assembly and the H-matrix kernels.  Files that include it
// are compiled with -ffp-contract=off: the association order and the literals are the
// reference's, so the entries are its bits wherever the device libm's log and sqrt agree.
#pragma once

#include "bh_common.h"

namespace bhg {

namespace {

// 21-point Gauss-Legendre nodes on [0, 1] and the [-1, 1] weights, as best/triangleint.f:66-108
// writes them (the final 0.5 * sums undo the interval change).
__constant__ double c_node[ 21 ] = {
    3.1239146898052499e-3, 1.6386580716846853e-2, 3.9950332924799586e-2, 7.3318317708341358e-2,
    0.1157800182621610,    0.1664305979012938,    0.2241905820563901,    0.2878289398962806,
    0.3559893415987995,    0.4272190729195525,    0.5,                   0.5727809270804476,
    0.6440106584012005,    0.7121710601037194,    0.7758094179436099,    0.8335694020987062,
    0.8842199817378390,    0.9266816822916586,    0.9600496670752004,    0.9836134192831531,
    0.9968760853101948 };
__constant__ double c_weight[ 21 ] = {
    1.601722825777433e-2, 3.695378977085249e-2, 5.713442542685720e-2, 7.610011362837930e-2,
    9.344442345603390e-2, 0.1087972991671484,   0.1218314160537285,   0.1322689386333375,
    0.1398873947910731,   0.1445244039899701,   0.1460811336496904,   0.1445244039899701,
    0.1398873947910731,   0.1322689386333375,   0.1218314160537285,   0.1087972991671484,
    9.344442345603390e-2, 7.610011362837930e-2, 5.713442542685720e-2, 3.695378977085249e-2,
    1.601722825777433e-2 };

__device__ inline double dot3( const double* p, const double* q ) { return p[ 0 ] * q[ 0 ] + p[ 1 ] * q[ 1 ] + p[ 2 ] * q[ 2 ]; }

// a / b and sqrt(x) with the instructions the compiler emits for the IEEE operations, less their
// range handling.  The quotient: the reciprocal (v_rcp_f64) refined by two Newton steps, the
// product with a and one correction of it; the compiler's sequence runs the same steps on
// operands that v_div_scale_f64 first scales when an exponent is near the ends of the range (and
// v_div_fixup_f64 then handles zeros, infinities and NaNs).  The root: v_rsq_f64 refined by the
// same coupled iteration for sqrt(x) and 1 / (2 sqrt(x)) the compiler runs after scaling x below
// 2^-767 up by 2^256 (and before passing 0 and +inf through).  Every operand here is a length, a
// squared distance or a product of a few, far from both ends of the double range and not zero
// (the collocation point is strictly inside its own triangle, off the source triangle's edges),
// so the scalings are identities and the results are the IEEE operations' bits, with a third
// fewer instructions per quotient and half per root.
__device__ __forceinline__ double qdiv( double a, double b ) {
   double r = __builtin_amdgcn_rcp( b );
   double e = fma( -b, r, 1.0 );
   r = fma( r, e, r );
   e = fma( -b, r, 1.0 );
   r = fma( r, e, r );
   const double q = a * r;
   const double rem = fma( -b, q, a );
   return fma( rem, r, q );
}

__device__ __forceinline__ double qsqrt( double x ) {
   const double y = __builtin_amdgcn_rsq( x );
   double g = x * y;
   double h = y * 0.5;
   const double r = fma( -h, g, 0.5 );
   g = fma( g, r, g );
   const double d0 = fma( -g, g, x );
   h = fma( h, r, h );
   g = fma( d0, h, g );
   const double d1 = fma( -g, g, x );
   return fma( d1, h, g );
}

// The off-diagonal block: source triangle record s (uniform across the wave), collocation point
// c (this lane's incenter).  Writes the 6 distinct entries g11 g21 g22 g31 g32 g33.
// best/triangleint.f:17-188 with the integrands h, hx, hy, hxy of :198-247; sqrt(d2121) and the
// square roots that h, hy and hxy spell identically are shared, which leaves the bits unchanged.
__device__ inline void pair_block( const double* __restrict__ s, const double c[ 3 ], double cg, double g[ 6 ] ) {
   double v1c[ 3 ], v2c[ 3 ], v3c[ 3 ];
   const double* v21 = s + kV21;
   const double* v31 = s + kV31;
   const double* v32 = s + kV32;
#pragma unroll
   for ( int k = 0; k < 3; ++k ) {
      v1c[ k ] = s[ kP1 + k ] - c[ k ];
      v2c[ k ] = s[ kP2 + k ] - c[ k ];
      v3c[ k ] = s[ kP3 + k ] - c[ k ];
   }
   const double d2121 = s[ kD2121 ], d2131 = s[ kD2131 ], d2132 = s[ kD2132 ];
   const double d3131 = s[ kD3131 ], d3232 = s[ kD3232 ], d3231 = s[ kD3231 ];
   const double s21 = s[ kS21 ];
   const double d211c = dot3( v21, v1c ), d212c = dot3( v21, v2c ), d322c = dot3( v32, v2c );
   const double d1c1c = dot3( v1c, v1c ), d1c31 = dot3( v1c, v31 ), d2c2c = dot3( v2c, v2c );
   const double d2c1c = dot3( v2c, v1c ), d3c31 = dot3( v3c, v31 ), d3c32 = dot3( v3c, v32 );
   const double d3c3c = dot3( v3c, v3c );

   double t0 = 0.0, s0 = 0.0, sx = 0.0, sy = 0.0, sxy = 0.0, syy = 0.0, sxx = 0.0;
#pragma unroll 3
   for ( int j = 0; j < 21; ++j ) {
      const double y = c_node[ j ];
      const double w = c_weight[ j ];
      const double q1 = qsqrt( d1c1c + 2.0 * y * d1c31 + y * y * d3131 );
      const double q2 = qsqrt( d2c2c + 2.0 * y * d322c + y * y * d3232 );
      // h(y)
      const double fh = qdiv( log( qdiv( d212c + d2132 * y + s21 * q2,
                                         d211c + d2131 * y + s21 * qsqrt( d1c1c + 2.0 * d1c31 * y + d3131 * y * y ) ) ),
                              s21 );
      // hx(y)
      double a = qdiv( -d1c31 - y * d2131, qsqrt( d1c1c + 2.0 * y * d211c + y * y * d2121 ) );
      double b = qdiv( -y * d3231 + d3c31, qsqrt( y * y * d3232 - 2.0 * y * d3c32 + d3c3c ) );
      double q = ( -d1c31 * d1c31 - 2.0 * y * d1c31 * d2131 - y * y * d2131 * d2131 + d1c1c * d3131 +
                   2.0 * y * d211c * d3131 + y * y * d2121 * d3131 );
      const double fhx = qdiv( a + b, q );
      // hy(y)
      a = qdiv( d211c + y * d2131, q1 );
      b = qdiv( y * d2132 + d212c, q2 );
      q = ( d211c * d211c - d1c1c * d2121 - 2.0 * y * d1c31 * d2121 + 2.0 * y * d211c * d2131 + y * y * d2131 * d2131 -
            y * y * d2121 * d3131 );
      const double fhy = qdiv( a - b, q );
      // hxy(y)
      b = qdiv( y * d211c - d2c1c - 2.0 * y * d1c31 - y * d2131 - y * y * d3231, q2 );
      q = ( -d211c * d211c + d1c1c * d2121 + 2.0 * y * d1c31 * d2121 - 2.0 * y * d211c * d2131 - y * y * ( d2131 * d2131 ) +
            y * y * d2121 * d3131 );
      const double fhxy = qdiv( q1 + b, q );
      t0 = t0 + w * fh;
      s0 = s0 + w * fhy;
      sx = sx + w * fhxy;
      sy = sy + w * y * fhy;
      sxy = sxy + w * y * fhxy;
      syy = syy + w * y * y * fhy;
      sxx = sxx + w * y * y * fhx;
   }
   t0 = 0.5 * t0;
   s0 = 0.5 * s0;
   sx = 0.5 * sx;
   sy = 0.5 * sy;
   sxy = 0.5 * sxy;
   syy = 0.5 * syy;
   sxx = 0.5 * sxx;
   auto tab = [&]( int p, int r ) {
      return v1c[ p ] * v1c[ r ] * s0 + ( v1c[ p ] * v21[ r ] + v21[ p ] * v1c[ r ] ) * sx +
             ( v31[ p ] * v1c[ r ] + v1c[ p ] * v31[ r ] ) * sy + ( v21[ p ] * v31[ r ] + v31[ p ] * v21[ r ] ) * sxy +
             v21[ p ] * v21[ r ] * sxx + v31[ p ] * v31[ r ] * syy;
   };
   const double ar = s[ kArea ];
   g[ 0 ] = 2.0 * ar * ( t0 + cg * tab( 0, 0 ) );
   g[ 1 ] = 2.0 * ar * tab( 0, 1 ) * cg;
   g[ 2 ] = 2.0 * ar * ( t0 + cg * tab( 1, 1 ) );
   g[ 3 ] = 2.0 * ar * tab( 0, 2 ) * cg;
   g[ 4 ] = 2.0 * ar * tab( 1, 2 ) * cg;
   g[ 5 ] = 2.0 * ar * ( t0 + cg * tab( 2, 2 ) );
}

// Row-major 3x3 from the 6 distinct entries of the symmetric pair block.
__device__ inline double sym_at( const double g[ 6 ], int a, int b ) {
   const int lo = a < b ? a : b, hi = a < b ? b : a;
   // (0,0)->0 (1,0)->1 (1,1)->2 (2,0)->3 (2,1)->4 (2,2)->5
   return g[ hi * ( hi + 1 ) / 2 + lo ];
}

}  // namespace

}  // namespace bhg
