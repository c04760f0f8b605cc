// The adjoint of BEST's dense solve: the gradient of D_t, D_r and CHI*V with respect to the
// vertex coordinates.  For an output Q of the solutions F of GR F = U,
//
//   dQ/dtheta = dQ/dtheta|_F + sum_k lambda_k^T ( dU_k/dtheta - dGR/dtheta F_k ),  GR^T lambda_k = dQ/dF_k,
//
// so the gradient costs one more solve, with the same LU factors transposed in place (GR^T =
// U^T L^T P), and one pass over the triangle pairs.
// This file holds the device parts:
//
//   k_adj_rhs      dQ/dF: the reductions of best/difco.f:73-98 and best/visco.f:191-205 are
//                  linear in F, so each triangle's adjoint right-hand side is the reductions'
//                  weights contracted with dQ/d(reductions)
//   k_transpose    in-place transpose of square blocks: the factors, the super-block inverses
//                  and the FP64 matrix of the refinement, for the adjoint solve
//   k_adj_pairs    sum_k lambda_k,i^T (dG_ij/dtheta) F_k,j over every pair, theta the source
//                  triangle's 9 vertex coordinates, its corrected area (G is linear in it) and the
//                  collocation point: reverse mode through the pair integral of
//                  best/triangleint.f (pair_adjoint), the three outputs' contractions at once
//   k_adj_tri      per triangle: the self block (best/tgetgr.f:35-145), incenter, area and XMU
//                  with their derivatives (forward mode over the 9 coordinates and the area
//                  factor), and the direct terms of the right-hand sides and reductions
//
// The host (solver.hip) runs the solves and the chain rule through the area correction, the
// deflation weights and the centroid shift.  Compiled with -ffp-contract=off like the assembly.
#include "adjoint.h"
#include "pair_kernel.h"

namespace bhg {

namespace {

// ------------------------------------------------------------------ forward-mode scalar

struct D1 {
   double v, d;  // value and derivative along one direction
};
__device__ inline D1 mk( double v, double d = 0.0 ) { return D1 { v, d }; }
__device__ inline D1 operator+( D1 a, D1 b ) { return { a.v + b.v, a.d + b.d }; }
__device__ inline D1 operator-( D1 a, D1 b ) { return { a.v - b.v, a.d - b.d }; }
__device__ inline D1 operator-( D1 a ) { return { -a.v, -a.d }; }
__device__ inline D1 operator*( D1 a, D1 b ) { return { a.v * b.v, a.d * b.v + a.v * b.d }; }
__device__ inline D1 operator/( D1 a, D1 b ) {
   const double q = a.v / b.v;
   return { q, ( a.d - q * b.d ) / b.v };
}
__device__ inline D1 operator+( D1 a, double b ) { return { a.v + b, a.d }; }
__device__ inline D1 operator+( double a, D1 b ) { return { a + b.v, b.d }; }
__device__ inline D1 operator-( D1 a, double b ) { return { a.v - b, a.d }; }
__device__ inline D1 operator-( double a, D1 b ) { return { a - b.v, -b.d }; }
__device__ inline D1 operator*( D1 a, double b ) { return { a.v * b, a.d * b }; }
__device__ inline D1 operator*( double a, D1 b ) { return { a * b.v, a * b.d }; }
__device__ inline D1 operator/( D1 a, double b ) { return { a.v / b, a.d / b }; }
__device__ inline D1 operator/( double a, D1 b ) {
   const double q = a / b.v;
   return { q, -q * b.d / b.v };
}

__device__ inline double tsqrt( double x ) { return sqrt( x ); }
__device__ inline double tlog( double x ) { return log( x ); }
__device__ inline double tacos( double x ) { return acos( x ); }
__device__ inline double tsin( double x ) { return sin( x ); }
__device__ inline double tcos( double x ) { return cos( x ); }
__device__ inline D1 tsqrt( D1 x ) {
   const double s = sqrt( x.v );
   return { s, 0.5 * x.d / s };
}
__device__ inline D1 tlog( D1 x ) { return { log( x.v ), x.d / x.v }; }
__device__ inline D1 tacos( D1 x ) { return { acos( x.v ), -x.d / sqrt( 1.0 - x.v * x.v ) }; }
__device__ inline D1 tsin( D1 x ) { return { sin( x.v ), cos( x.v ) * x.d }; }
__device__ inline D1 tcos( D1 x ) { return { cos( x.v ), -sin( x.v ) * x.d }; }
__device__ inline double val( double x ) { return x; }
__device__ inline double val( D1 x ) { return x.v; }
__device__ inline double der( double ) { return 0.0; }
__device__ inline double der( D1 x ) { return x.d; }

template <typename T>
__device__ inline T tdot( const T* p, const T* q ) {
   return p[ 0 ] * q[ 0 ] + p[ 1 ] * q[ 1 ] + p[ 2 ] * q[ 2 ];
}

// ------------------------------------------------------------------ the functions differentiated

template <typename T>
__device__ inline T tf_si( T x ) {
   return 2.0 * ( tlog( ( 1.0 + tsqrt( 1.0 - x * x ) ) / x ) - tsqrt( 1.0 - x * x ) );
}
template <typename T>
__device__ inline T tf_csi( T x ) {
   return 3.0 - 1.0 / x - 2.0 * x;
}

// The triangle's own terms as functions of its vertices and the area factor: incenter x,
// geometric area A, XMU = afac A n (n BEST's unit normal) and the self block S (row-major), as
// k_self computes them (best/tgetgr.f:35-145, best/best.f:387-399).
template <typename T>
__device__ inline void self_geom_t( const T p1[ 3 ], const T p2[ 3 ], const T p3[ 3 ], T afac, double cg, T x[ 3 ], T& A, T xmu[ 3 ],
                                    T S[ 9 ] ) {
   const double pi = 3.1415926535897932385;
   T e21[ 3 ], e31[ 3 ], e32[ 3 ];
   for ( int k = 0; k < 3; ++k ) e21[ k ] = p2[ k ] - p1[ k ], e31[ k ] = p3[ k ] - p1[ k ], e32[ k ] = p3[ k ] - p2[ k ];
   const T a = tsqrt( tdot( e21, e21 ) ), b = tsqrt( tdot( e31, e31 ) ), c = tsqrt( tdot( e32, e32 ) );
   const T cosC = tdot( e21, e31 ) / a / b;
   const T t = ( a + b - c ) / ( 2. * ( 1. + cosC ) );
   for ( int k = 0; k < 3; ++k ) x[ k ] = p1[ k ] + t * ( e21[ k ] / a + e31[ k ] / b );
   const T cxn = e31[ 1 ] * e21[ 2 ] - e31[ 2 ] * e21[ 1 ];
   const T cyn = e31[ 2 ] * e21[ 0 ] - e31[ 0 ] * e21[ 2 ];
   const T czn = e31[ 0 ] * e21[ 1 ] - e31[ 1 ] * e21[ 0 ];
   const T pl = tsqrt( cxn * cxn + cyn * cyn + czn * czn );
   A = 0.5 * pl;
   xmu[ 0 ] = afac * A * cxn / pl;
   xmu[ 1 ] = afac * A * cyn / pl;
   xmu[ 2 ] = afac * A * czn / pl;
   const T nx = 0.5 * cxn / A, ny = 0.5 * cyn / A, nz = 0.5 * czn / A;
   const T cc = tacos( cosC );
   const T r = t * tsin( cc );
   const T cosB = -tdot( e21, e32 ) / a / c;
   const T bb = tacos( cosB );
   const T aa = pi - bb - cc;
   const T t2 = 1. / ( 2. * tcos( 0.5 * cc ) );
   const T dc = ( a + b - c ) * t2;
   const T db = tsqrt( a * a + ( b - c ) * ( b - c ) - 2. * a * ( b - c ) * cosC ) * t2;
   const T da = tsqrt( b * b + ( a - c ) * ( a - c ) - 2. * b * ( a - c ) * cosC ) * t2;
   const T ha = r / da, hb = r / db, hc = r / dc;
   const T dpa = ( pi - aa ) * ( da - r ) - 2.0 * ( da * tacos( ha ) - r * tlog( ( 1. + tsqrt( 1. - ha * ha ) ) / ha ) );
   const T dpb = ( pi - bb ) * ( db - r ) - 2.0 * ( db * tacos( hb ) - r * tlog( ( 1. + tsqrt( 1. - hb * hb ) ) / hb ) );
   const T dpc = ( pi - cc ) * ( dc - r ) - 2.0 * ( dc * tacos( hc ) - r * tlog( ( 1. + tsqrt( 1. - hc * hc ) ) / hc ) );
   const T sca = 0.5 * r * ( ( tcos( 2.0 * cc ) + tcos( 2.0 * bb ) ) * tf_si( ha ) - ( tsin( 2.0 * cc ) + tsin( 2.0 * bb ) ) * tf_csi( ha ) );
   const T scb = 0.5 * r * ( ( 1. + tcos( 2.0 * bb ) ) * tf_si( hb ) + tsin( 2.0 * bb ) * tf_csi( hb ) );
   const T scc = 0.5 * r * ( ( 1. + tcos( 2.0 * cc ) ) * tf_si( hc ) + tsin( 2.0 * cc ) * tf_csi( hc ) );
   const T cca = -tsin( bb - cc ) * r * ( tcos( aa ) * tf_si( ha ) + tsin( aa ) * tf_csi( ha ) );
   const T ccb = tsin( bb ) * r * ( tcos( bb ) * tf_si( hb ) + tsin( bb ) * tf_csi( hb ) );
   const T ccc = -tsin( cc ) * r * ( tcos( cc ) * tf_si( hc ) + tsin( cc ) * tf_csi( hc ) );
   const T cor = tsqrt( afac );
   T g[ 3 ][ 3 ];
   g[ 0 ][ 0 ] = cor * ( ( 2.0 + cg ) * pi * r + ( 1.0 + 0.5 * cg ) * cg * ( dpa + dpb + dpc ) + cg * 0.25 * ( sca + scb + scc ) );
   g[ 0 ][ 1 ] = 0.5 * cor * ( cca + ccb + ccc ) * cg;
   g[ 1 ][ 0 ] = g[ 0 ][ 1 ];
   g[ 1 ][ 1 ] = g[ 0 ][ 0 ] - 0.5 * cor * cg * ( sca + scb + scc );
   g[ 2 ][ 2 ] = cor * ( 2.0 * pi * r + ( dpa + dpb + dpc ) );
   g[ 0 ][ 2 ] = g[ 1 ][ 2 ] = g[ 2 ][ 0 ] = g[ 2 ][ 1 ] = T {};
   T rr[ 3 ][ 3 ];
   rr[ 0 ][ 0 ] = e21[ 0 ] / a, rr[ 0 ][ 1 ] = e21[ 1 ] / a, rr[ 0 ][ 2 ] = e21[ 2 ] / a;
   rr[ 2 ][ 0 ] = nx, rr[ 2 ][ 1 ] = ny, rr[ 2 ][ 2 ] = nz;
   rr[ 1 ][ 0 ] = ( ny * e21[ 2 ] - nz * e21[ 1 ] ) / a;
   rr[ 1 ][ 1 ] = ( nz * e21[ 0 ] - nx * e21[ 2 ] ) / a;
   rr[ 1 ][ 2 ] = ( nx * e21[ 1 ] - ny * e21[ 0 ] ) / a;
   for ( int p = 0; p < 3; ++p )
      for ( int q = 0; q < 3; ++q ) {
         T acc = T {};
         for ( int k = 0; k < 3; ++k )
            acc = acc + rr[ k ][ p ] * ( g[ k ][ 0 ] * rr[ 0 ][ q ] + g[ k ][ 1 ] * rr[ 1 ][ q ] + g[ k ][ 2 ] * rr[ 2 ][ q ] );
         S[ 3 * p + q ] = acc;
      }
}

// The 11 right-hand sides at incenter x (best/difco.f:28-44, best/visco.f:216-243), u[k][c].
template <typename T>
__device__ inline void rhs_t( const T x[ 3 ], T u[ 11 ][ 3 ] ) {
   for ( int k = 0; k < 11; ++k ) u[ k ][ 0 ] = u[ k ][ 1 ] = u[ k ][ 2 ] = T {};
   u[ 0 ][ 0 ] = T {} + 1.0, u[ 1 ][ 1 ] = T {} + 1.0, u[ 2 ][ 2 ] = T {} + 1.0;
   u[ 3 ][ 1 ] = -x[ 2 ], u[ 3 ][ 2 ] = x[ 1 ];
   u[ 4 ][ 0 ] = x[ 2 ], u[ 4 ][ 2 ] = -x[ 0 ];
   u[ 5 ][ 0 ] = -x[ 1 ], u[ 5 ][ 1 ] = x[ 0 ];
   u[ 6 ][ 0 ] = 0.5 * x[ 1 ], u[ 6 ][ 1 ] = 0.5 * x[ 0 ];
   u[ 7 ][ 0 ] = 0.5 * x[ 2 ], u[ 7 ][ 2 ] = 0.5 * x[ 0 ];
   u[ 8 ][ 1 ] = 0.5 * x[ 2 ], u[ 8 ][ 2 ] = 0.5 * x[ 1 ];
   u[ 9 ][ 0 ] = 0.5 * x[ 0 ], u[ 9 ][ 1 ] = -0.5 * x[ 1 ];
   u[ 10 ][ 0 ] = 0.5 * x[ 0 ], u[ 10 ][ 2 ] = -0.5 * x[ 2 ];
}

// One triangle's share of the reductions of k_sums, slots 0..44 (DIFCO's forces and torques, the
// CHI moments at the centroid), as a function of its incenter x, corrected area ar and solutions
// f[k][c].
template <typename T, typename TF>
__device__ inline void tri_sums_t( const T x[ 3 ], T ar, const TF f[ 11 ][ 3 ], T out[ kAdjSlots ] ) {
   for ( int mx = 0; mx < 6; ++mx ) {
      const TF* g = f[ mx ];
      const T q1 = x[ 1 ] * g[ 2 ] - x[ 2 ] * g[ 1 ], q2 = x[ 2 ] * g[ 0 ] - x[ 0 ] * g[ 2 ], q3 = x[ 0 ] * g[ 1 ] - x[ 1 ] * g[ 0 ];
      out[ 6 * mx + 0 ] = -( ar * g[ 0 ] );
      out[ 6 * mx + 1 ] = -( ar * g[ 1 ] );
      out[ 6 * mx + 2 ] = -( ar * g[ 2 ] );
      out[ 6 * mx + 3 ] = -( ar * q1 );
      out[ 6 * mx + 4 ] = -( ar * q2 );
      out[ 6 * mx + 5 ] = -( ar * q3 );
   }
   const TF* w0 = f[ 6 ];
   const TF* w1 = f[ 7 ];
   const TF* w2 = f[ 8 ];
   const TF* w3 = f[ 9 ];
   const TF* w4 = f[ 10 ];
   out[ 36 ] = ar * ( w0[ 0 ] * x[ 1 ] + w0[ 1 ] * x[ 0 ] );
   out[ 37 ] = ar * ( w1[ 0 ] * x[ 2 ] + w1[ 2 ] * x[ 0 ] );
   out[ 38 ] = ar * ( w2[ 1 ] * x[ 2 ] + w2[ 2 ] * x[ 1 ] );
   out[ 39 ] = ar * w3[ 0 ] * x[ 0 ];
   out[ 40 ] = ar * w3[ 1 ] * x[ 1 ];
   out[ 41 ] = ar * w3[ 2 ] * x[ 2 ];
   out[ 42 ] = ar * w4[ 0 ] * x[ 0 ];
   out[ 43 ] = ar * w4[ 1 ] * x[ 1 ];
   out[ 44 ] = ar * w4[ 2 ] * x[ 2 ];
}

// ------------------------------------------------------------------ the pair term in reverse mode

// The pointer as a value the compiler cannot see through, so that loads through it are issued
// again rather than their results held in registers (the source record is wave-uniform: SGPRs).
__device__ inline const double* reread( const double* q ) {
   asm volatile( "" : "+s"( q ) );
   return q;
}
__host__ inline const double* reread( const double* q ) { return q; }

// Reverse mode of the pair block's contractions L_o = sum_e W[o][e] g_e (o = 0, 1, 2; g the 6
// distinct entries of best/triangleint.f's block, as pair_block writes them) with respect to the
// source vertices p (p1, p2, p3) and the collocation point c, for a fixed corrected area ar:
//
//   L_o = 2 ar [ t0 tr W_o + cg ( s0 a11 + 2 sx a12 + 2 sy a13 + 2 sxy a23 + sxx a22 + syy a33 ) ],
//
// a_uv = u^T S_o v over the vectors v1c, v21, v31 (S_o the symmetric matrix with W_o's diagonal
// entries on its diagonal and half its off-diagonal ones off it).  The adjoints of the seven
// moments are known before the quadrature, so each of the 21 nodes is evaluated and then
// differentiated at once (no tape), accumulating the adjoints of the 17 dot products; after the
// loop they and the direct terms of tab go to the six difference vectors and to p and c.
// gp[o][3 v + k] = dL_o / dp_v,k, gc[o][k] = dL_o / dc_k, lar[o] = L_o / ar, for NO outputs at once.
// p (the 9 coordinates) is read from memory again after the quadrature: the six difference
// vectors are not held in registers through it.
template <int NO>
__host__ __device__ inline void pair_adjoint( const double* p, const double c[ 3 ], double ar, double cg, const double ( *W )[ 6 ],
                                              const double* node, const double* weight, double ( *gp )[ 9 ], double ( *gc )[ 3 ],
                                              double* lar ) {
   double v21[ 3 ], v31[ 3 ], v32[ 3 ], v1c[ 3 ], v2c[ 3 ], v3c[ 3 ];
   auto vectors = [&]( const double* q ) {
      for ( int k = 0; k < 3; ++k ) {
         v21[ k ] = q[ 3 + k ] - q[ k ];
         v31[ k ] = q[ 6 + k ] - q[ k ];
         v32[ k ] = q[ 6 + k ] - q[ 3 + k ];
         v1c[ k ] = q[ k ] - c[ k ];
         v2c[ k ] = q[ 3 + k ] - c[ k ];
         v3c[ k ] = q[ 6 + k ] - c[ k ];
      }
   };
   vectors( p );
   auto dot = []( const double* a, const double* b ) { return a[ 0 ] * b[ 0 ] + a[ 1 ] * b[ 1 ] + a[ 2 ] * b[ 2 ]; };
   const double d2121 = dot( v21, v21 ), d2131 = dot( v21, v31 ), d2132 = dot( v21, v32 );
   const double d3131 = dot( v31, v31 ), d3232 = dot( v32, v32 ), d3231 = dot( v32, v31 );
   const double s21 = sqrt( d2121 ), is21 = 1.0 / s21;
   const double d211c = dot( v21, v1c ), d212c = dot( v21, v2c ), d322c = dot( v32, v2c );
   const double d1c1c = dot( v1c, v1c ), d1c31 = dot( v1c, v31 ), d2c2c = dot( v2c, v2c );
   const double d2c1c = dot( v2c, v1c ), d3c31 = dot( v3c, v31 ), d3c32 = dot( v3c, v32 );
   const double d3c3c = dot( v3c, v3c );

   // S_o v for the three vectors (S_o from W_o), the quadratic forms, the moments' adjoints
   double Sv1c[ 3 ], Sv21[ 3 ], Sv31[ 3 ];
   auto svec = [&]( const double* w ) {
      const double S[ 3 ][ 3 ] = { { w[ 0 ], 0.5 * w[ 1 ], 0.5 * w[ 3 ] },
                                   { 0.5 * w[ 1 ], w[ 2 ], 0.5 * w[ 4 ] },
                                   { 0.5 * w[ 3 ], 0.5 * w[ 4 ], w[ 5 ] } };
      for ( int r = 0; r < 3; ++r ) {
         Sv1c[ r ] = dot( S[ r ], v1c );
         Sv21[ r ] = dot( S[ r ], v21 );
         Sv31[ r ] = dot( S[ r ], v31 );
      }
   };
   double bt0[ NO ], bs0[ NO ], bsx[ NO ], bsy[ NO ], bsxy[ NO ], bsxx[ NO ], bsyy[ NO ];
   for ( int o = 0; o < NO; ++o ) {
      const double* w = W[ o ];
      svec( w );
      const double a11 = dot( v1c, Sv1c ), a12 = dot( v1c, Sv21 ), a13 = dot( v1c, Sv31 );
      const double a22 = dot( v21, Sv21 ), a23 = dot( v21, Sv31 ), a33 = dot( v31, Sv31 );
      const double f = 2.0 * ar * cg;
      bt0[ o ] = 2.0 * ar * ( w[ 0 ] + w[ 2 ] + w[ 5 ] );
      bs0[ o ] = f * a11;
      bsx[ o ] = 2.0 * f * a12;
      bsy[ o ] = 2.0 * f * a13;
      bsxy[ o ] = 2.0 * f * a23;
      bsxx[ o ] = f * a22;
      bsyy[ o ] = f * a33;
   }

   // adjoints of the dot products (and of s21), per output
   double b2131[ NO ] = {}, b2132[ NO ] = {}, b3131[ NO ] = {}, b3232[ NO ] = {}, b3231[ NO ] = {}, bs21[ NO ] = {}, b2121[ NO ] = {};
   double b211c[ NO ] = {}, b212c[ NO ] = {}, b322c[ NO ] = {}, b1c1c[ NO ] = {}, b1c31[ NO ] = {}, b2c2c[ NO ] = {};
   double b2c1c[ NO ] = {}, b3c31[ NO ] = {}, b3c32[ NO ] = {}, b3c3c[ NO ] = {};
   double t0 = 0.0, s0 = 0.0, sx = 0.0, sy = 0.0, sxy = 0.0, syy = 0.0, sxx = 0.0;
#pragma unroll 1
   for ( int j = 0; j < 21; ++j ) {
      const double y = node[ j ], w = weight[ j ], yy = y * y;
      // the node's values
      const double A1 = d1c1c + 2.0 * y * d1c31 + yy * d3131, q1 = sqrt( A1 ), iq1 = 1.0 / q1;
      const double A2 = d2c2c + 2.0 * y * d322c + yy * d3232, q2 = sqrt( A2 ), iq2 = 1.0 / q2;
      const double Nn = d212c + d2132 * y + s21 * q2, Dd = d211c + d2131 * y + s21 * q1;
      const double iNn = 1.0 / Nn, iDd = 1.0 / Dd;
      const double fh = log( Nn * iDd ) * is21;
      const double B1 = d1c1c + 2.0 * y * d211c + yy * d2121, ir1 = 1.0 / sqrt( B1 );
      const double B2 = yy * d3232 - 2.0 * y * d3c32 + d3c3c, ir2 = 1.0 / sqrt( B2 );
      const double a = ( -d1c31 - y * d2131 ) * ir1, b = ( -y * d3231 + d3c31 ) * ir2;
      const double Q = -d1c31 * d1c31 - 2.0 * y * d1c31 * d2131 - yy * d2131 * d2131 + d1c1c * d3131 + 2.0 * y * d211c * d3131 +
                       yy * d2121 * d3131;
      const double iQ = 1.0 / Q, fhx = ( a + b ) * iQ;
      const double a2 = ( d211c + y * d2131 ) * iq1, b2 = ( y * d2132 + d212c ) * iq2;
      const double Q2 = d211c * d211c - d1c1c * d2121 - 2.0 * y * d1c31 * d2121 + 2.0 * y * d211c * d2131 + yy * d2131 * d2131 -
                        yy * d2121 * d3131;
      const double iQ2 = 1.0 / Q2, fhy = ( a2 - b2 ) * iQ2;
      const double b3 = ( y * d211c - d2c1c - 2.0 * y * d1c31 - y * d2131 - yy * d3231 ) * iq2;
      // hxy's denominator is -Q2 (the kernel spells it out; the same value to rounding)
      const double fhxy = -( q1 + b3 ) * iQ2;
      t0 += w * fh, s0 += w * fhy, sx += w * fhxy, sy += w * y * fhy, sxy += w * y * fhxy, syy += w * yy * fhy, sxx += w * yy * fhx;
      const double hw = 0.5 * w;
      for ( int o = 0; o < NO; ++o ) {
         // seeds: the moments' adjoints times d(moment)/d(integrand) = 0.5 w y^k
         const double gfh = hw * bt0[ o ];
         const double gfhy = hw * ( bs0[ o ] + y * bsy[ o ] + yy * bsyy[ o ] );
         const double gfhxy = hw * ( bsx[ o ] + y * bsxy[ o ] );
         const double gfhx = hw * yy * bsxx[ o ];
         // fh = (log Nn - log Dd) / s21
         const double gNn = gfh * iNn * is21, gDd = -gfh * iDd * is21;
         bs21[ o ] += -gfh * fh * is21 + q2 * gNn + q1 * gDd;
         b212c[ o ] += gNn, b2132[ o ] += y * gNn;
         b211c[ o ] += gDd, b2131[ o ] += y * gDd;
         double gq1 = s21 * gDd, gq2 = s21 * gNn;
         // fhx = (a + b) / Q
         const double ga = gfhx * iQ, gb = gfhx * iQ, gQ = -gfhx * fhx * iQ;
         b1c31[ o ] -= ga * ir1, b2131[ o ] -= y * ga * ir1;
         const double gB1 = -0.5 * ga * a * ir1 * ir1;
         b1c1c[ o ] += gB1, b211c[ o ] += 2.0 * y * gB1, b2121[ o ] += yy * gB1;
         b3231[ o ] -= y * gb * ir2, b3c31[ o ] += gb * ir2;
         const double gB2 = -0.5 * gb * b * ir2 * ir2;
         b3232[ o ] += yy * gB2, b3c32[ o ] -= 2.0 * y * gB2, b3c3c[ o ] += gB2;
         b1c31[ o ] += gQ * ( -2.0 * d1c31 - 2.0 * y * d2131 );
         b2131[ o ] += gQ * ( -2.0 * y * d1c31 - 2.0 * yy * d2131 );
         b1c1c[ o ] += gQ * d3131;
         b3131[ o ] += gQ * B1;
         b211c[ o ] += gQ * 2.0 * y * d3131;
         b2121[ o ] += gQ * yy * d3131;
         // fhy = (a2 - b2) / Q2 and fhxy = -(q1 + b3) / Q2
         const double ga2 = gfhy * iQ2, gb2 = -gfhy * iQ2;
         gq1 += -gfhxy * iQ2;
         const double gb3 = -gfhxy * iQ2;
         const double gQ2 = -( gfhy * fhy + gfhxy * fhxy ) * iQ2;
         b211c[ o ] += ga2 * iq1, b2131[ o ] += y * ga2 * iq1, gq1 += -ga2 * a2 * iq1;
         b2132[ o ] += y * gb2 * iq2, b212c[ o ] += gb2 * iq2, gq2 += -gb2 * b2 * iq2;
         const double gN3 = gb3 * iq2;
         gq2 += -gb3 * b3 * iq2;
         b211c[ o ] += y * gN3, b2c1c[ o ] -= gN3, b1c31[ o ] -= 2.0 * y * gN3, b2131[ o ] -= y * gN3, b3231[ o ] -= yy * gN3;
         b211c[ o ] += gQ2 * ( 2.0 * d211c + 2.0 * y * d2131 );
         b1c1c[ o ] += gQ2 * ( -d2121 );
         b2121[ o ] += gQ2 * ( -A1 );
         b1c31[ o ] += gQ2 * ( -2.0 * y * d2121 );
         b2131[ o ] += gQ2 * ( 2.0 * y * d211c + 2.0 * yy * d2131 );
         b3131[ o ] += gQ2 * ( -yy * d2121 );
         // q1 = sqrt(A1), q2 = sqrt(A2)
         const double gA1 = 0.5 * gq1 * iq1, gA2 = 0.5 * gq2 * iq2;
         b1c1c[ o ] += gA1, b1c31[ o ] += 2.0 * y * gA1, b3131[ o ] += yy * gA1;
         b2c2c[ o ] += gA2, b322c[ o ] += 2.0 * y * gA2, b3232[ o ] += yy * gA2;
      }
   }
   t0 *= 0.5, s0 *= 0.5, sx *= 0.5, sy *= 0.5, sxy *= 0.5, syy *= 0.5, sxx *= 0.5;

   vectors( reread( p ) );  // the coordinates again, from memory
   for ( int o = 0; o < NO; ++o ) {
      const double* w = W[ o ];
      svec( w );
      const double a11 = dot( v1c, Sv1c ), a12 = dot( v1c, Sv21 ), a13 = dot( v1c, Sv31 );
      const double a22 = dot( v21, Sv21 ), a23 = dot( v21, Sv31 ), a33 = dot( v31, Sv31 );
      lar[ o ] = 2.0 * ( t0 * ( w[ 0 ] + w[ 2 ] + w[ 5 ] ) +
                         cg * ( s0 * a11 + 2.0 * sx * a12 + 2.0 * sy * a13 + 2.0 * sxy * a23 + sxx * a22 + syy * a33 ) );
      // the vectors' adjoints: tab's direct terms, then the dot products
      const double f = 4.0 * ar * cg;
      double g21[ 3 ], g31[ 3 ], g32[ 3 ], g1c[ 3 ], g2c[ 3 ], g3c[ 3 ];
      const double bb2121 = b2121[ o ] + 0.5 * bs21[ o ] * is21;
      for ( int k = 0; k < 3; ++k ) {
         g1c[ k ] = f * ( s0 * Sv1c[ k ] + sx * Sv21[ k ] + sy * Sv31[ k ] );
         g21[ k ] = f * ( sx * Sv1c[ k ] + sxy * Sv31[ k ] + sxx * Sv21[ k ] );
         g31[ k ] = f * ( sy * Sv1c[ k ] + sxy * Sv21[ k ] + syy * Sv31[ k ] );
         g32[ k ] = 0.0, g2c[ k ] = 0.0, g3c[ k ] = 0.0;
         g21[ k ] += 2.0 * bb2121 * v21[ k ] + b2131[ o ] * v31[ k ] + b2132[ o ] * v32[ k ] + b211c[ o ] * v1c[ k ] + b212c[ o ] * v2c[ k ];
         g31[ k ] += b2131[ o ] * v21[ k ] + 2.0 * b3131[ o ] * v31[ k ] + b3231[ o ] * v32[ k ] + b1c31[ o ] * v1c[ k ] + b3c31[ o ] * v3c[ k ];
         g32[ k ] += b2132[ o ] * v21[ k ] + 2.0 * b3232[ o ] * v32[ k ] + b3231[ o ] * v31[ k ] + b322c[ o ] * v2c[ k ] + b3c32[ o ] * v3c[ k ];
         g1c[ k ] += b211c[ o ] * v21[ k ] + 2.0 * b1c1c[ o ] * v1c[ k ] + b1c31[ o ] * v31[ k ] + b2c1c[ o ] * v2c[ k ];
         g2c[ k ] += b212c[ o ] * v21[ k ] + b322c[ o ] * v32[ k ] + 2.0 * b2c2c[ o ] * v2c[ k ] + b2c1c[ o ] * v1c[ k ];
         g3c[ k ] += b3c31[ o ] * v31[ k ] + b3c32[ o ] * v32[ k ] + 2.0 * b3c3c[ o ] * v3c[ k ];
         gp[ o ][ k ] = -g21[ k ] - g31[ k ] + g1c[ k ];
         gp[ o ][ 3 + k ] = g21[ k ] - g32[ k ] + g2c[ k ];
         gp[ o ][ 6 + k ] = g31[ k ] + g32[ k ] + g3c[ k ];
         gc[ o ][ k ] = -( g1c[ k ] + g2c[ k ] + g3c[ k ] );
      }
   }
}

// ------------------------------------------------------------------ kernels

// y_o(3 i + c, k) = sum_s ghat[o][s] d(sum_s)/dF(3 i + c, k): the reductions are linear in F,
// so each weight is the reduction of a unit solution.
__global__ void k_adj_rhs( int n, const double* __restrict__ xc, const double* __restrict__ area, const double* __restrict__ ghat,
                           double* __restrict__ y, long ld ) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   const double x[ 3 ] = { xc[ i ], xc[ n + i ], xc[ 2 * n + i ] };
   const double ar = area[ i ];
   double f[ 11 ][ 3 ] = {};
   for ( int k = 0; k < 11; ++k )
      for ( int c = 0; c < 3; ++c ) {
         f[ k ][ c ] = 1.0;
         double out[ kAdjSlots ];
         tri_sums_t<double, double>( x, ar, f, out );
         f[ k ][ c ] = 0.0;
         for ( int o = 0; o < kAdjOutputs; ++o ) {
            double s = 0.0;
            for ( int q = 0; q < kAdjSlots; ++q ) s += ghat[ o * kAdjSlots + q ] * out[ q ];
            y[ (long)o * 11 * ld + (long)k * ld + 3L * i + c ] = s;
         }
      }
}

// In-place transpose of square blocks of order 32 nt: workgroup (t, b) swaps the 32 x 32 tiles
// (bi, bj) and (bj, bi) of block b transposed, t = the pair's index in the upper triangle (bi <= bj)
// row by row.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_transpose( int nt, T* __restrict__ a, long ld, long stride ) {
   __shared__ T s1[ 32 ][ 33 ], s2[ 32 ][ 33 ];
   long t = blockIdx.x;
   int bi = 0;
   while ( t >= nt - bi ) t -= nt - bi, ++bi;
   const int bj = bi + (int)t;
   const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
   T* base = a + (long)blockIdx.y * stride;
   T* A = base + (long)bj * 32 * ld + (long)bi * 32;  // tile (bi, bj)
   T* B = base + (long)bi * 32 * ld + (long)bj * 32;  // tile (bj, bi)
   for ( int r = ty; r < 32; r += 8 ) {
      s1[ r ][ tx ] = A[ (long)r * ld + tx ];
      s2[ r ][ tx ] = B[ (long)r * ld + tx ];
   }
   __syncthreads();
   for ( int r = ty; r < 32; r += 8 ) {
      A[ (long)r * ld + tx ] = s2[ tx ][ r ];
      B[ (long)r * ld + tx ] = s1[ tx ][ r ];
   }
}

// lambda_o,k at triangle i component a: lam[o * 11 * ld + k * ld + 3 i + a]; F likewise without o.
// Per pair (n1 collocation = lane, n2 source = uniform over the wave) and output o: M_o =
// sum_k lambda_o,k,n1 F_k,n2^T, contracted with the pair block G (symmetric: 6 weights), and the
// gradient of that contraction with respect to the source's 9 coordinates and the collocation
// point by pair_adjoint, plus M_o : G / area.  Every lane evaluates (so the wave sums need no
// divergent branch): on the diagonal pair and on idle lanes the weights are 0 and the point is
// moved off the triangle, so their terms are exact zeros.  gsrc: per (collocation block, n2) the
// 30 sums over the block's lanes (o-major: 9 coordinates, then the area); gcol: per (source
// strip, n1) 9 values (o-major, 3 coordinates).
__global__ void __launch_bounds__( 256 ) __attribute__( ( amdgpu_waves_per_eu( 2 ) ) ) k_adj_pairs( int n, int strip, double cg, const double* __restrict__ src,
                                                     const double* __restrict__ xc, const double* __restrict__ lam,
                                                     const double* __restrict__ F, long ld, double* __restrict__ gsrc,
                                                     double* __restrict__ gcol ) {
   __shared__ double red[ 4 ][ kAdjOutputs * 10 ];
   const int n1 = blockIdx.x * blockDim.x + threadIdx.x;
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int n2lo = blockIdx.y * strip, n2hi = min( n, n2lo + strip );
   const bool act = n1 < n;
   const int m1 = act ? n1 : 0;
   const double c[ 3 ] = { xc[ m1 ], xc[ n + m1 ], xc[ 2 * n + m1 ] };
   double gcs[ kAdjOutputs ][ 3 ] = {};
   for ( int n2 = n2lo; n2 < n2hi; ++n2 ) {
      const double* s = src + (size_t)kSrcStride * n2;
      const bool live = act && n2 != n1;
      const double* p = s + kP1;
      // a point off the source triangle for the lanes whose terms are zero
      const double off = live ? 0.0 : 4.0 * ( fabs( p[ 3 ] - p[ 0 ] ) + fabs( p[ 6 ] - p[ 0 ] ) + fabs( p[ 4 ] - p[ 1 ] ) +
                                              fabs( p[ 7 ] - p[ 1 ] ) + 1.0 );
      const double ce[ 3 ] = { live ? c[ 0 ] : p[ 0 ] + off, live ? c[ 1 ] : p[ 1 ] + off, live ? c[ 2 ] : p[ 2 ] + off };
#pragma unroll 1
      for ( int o = 0; o < kAdjOutputs; ++o ) {
         // M_o (symmetric part: G is symmetric), the contraction weights of the 6 distinct entries
         double M[ 3 ][ 3 ] = {};
         for ( int k = 0; k < 11; ++k ) {
            const double* l = lam + (long)o * 11 * ld + (long)k * ld + 3L * m1;
            const double* f = F + (long)k * ld + 3L * n2;
            const double l0 = l[ 0 ], l1 = l[ 1 ], l2 = l[ 2 ], f0 = f[ 0 ], f1 = f[ 1 ], f2 = f[ 2 ];
            M[ 0 ][ 0 ] += l0 * f0, M[ 0 ][ 1 ] += l0 * f1, M[ 0 ][ 2 ] += l0 * f2;
            M[ 1 ][ 0 ] += l1 * f0, M[ 1 ][ 1 ] += l1 * f1, M[ 1 ][ 2 ] += l1 * f2;
            M[ 2 ][ 0 ] += l2 * f0, M[ 2 ][ 1 ] += l2 * f1, M[ 2 ][ 2 ] += l2 * f2;
         }
         const double z = live ? 1.0 : 0.0;
         const double W[ 1 ][ 6 ] = { { z * M[ 0 ][ 0 ], z * ( M[ 1 ][ 0 ] + M[ 0 ][ 1 ] ), z * M[ 1 ][ 1 ],
                                        z * ( M[ 2 ][ 0 ] + M[ 0 ][ 2 ] ), z * ( M[ 2 ][ 1 ] + M[ 1 ][ 2 ] ), z * M[ 2 ][ 2 ] } };
         double gp[ 1 ][ 9 ], gc[ 1 ][ 3 ], lar[ 1 ];
         pair_adjoint<1>( p, ce, s[ kArea ], cg, W, c_node, c_weight, gp, gc, lar );
         for ( int k = 0; k < 3; ++k )  // selects: o is not a compile-time index here
            for ( int oo = 0; oo < kAdjOutputs; ++oo ) gcs[ oo ][ k ] += oo == o ? gc[ 0 ][ k ] : 0.0;
         // the 10 terms summed over the wave (a fixed order)
#pragma unroll
         for ( int q = 0; q < 10; ++q ) {
            double v = q < 9 ? gp[ 0 ][ q ] : lar[ 0 ];
#pragma unroll
            for ( int sh = 32; sh > 0; sh >>= 1 ) v += __shfl_xor( v, sh );
            if ( lane == 0 ) red[ wave ][ o * 10 + q ] = v;
         }
      }
      __syncthreads();
      if ( threadIdx.x < kAdjOutputs * 10 )
         gsrc[ ( (long)blockIdx.x * n + n2 ) * ( kAdjOutputs * 10 ) + threadIdx.x ] =
             ( red[ 0 ][ threadIdx.x ] + red[ 1 ][ threadIdx.x ] ) + ( red[ 2 ][ threadIdx.x ] + red[ 3 ][ threadIdx.x ] );
      __syncthreads();
   }
   if ( act )
      for ( int o = 0; o < kAdjOutputs; ++o )
         for ( int k = 0; k < 3; ++k ) gcol[ ( (long)blockIdx.y * n + n1 ) * ( kAdjOutputs * 3 ) + o * 3 + k ] = gcs[ o ][ k ];
}

// out[i * width + q] = sum over the nparts partials in order.
__global__ void k_adj_sum_parts( long count, int nparts, long stride, const double* __restrict__ part, double* __restrict__ out ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t >= count ) return;
   double a = 0.0;
   for ( int p = 0; p < nparts; ++p ) a += part[ (long)p * stride + t ];
   out[ t ] = a;
}

// Per triangle i, directions theta = its 9 vertex coordinates (v1 x y z, v2, v3) and the area
// factor (10): the derivatives of its incenter, geometric area and XMU, the self block contracted
// with sum_k lambda_o,k,i F_k,i^T, the trace of the self block; and the direct terms at fixed F:
// sum_k lambda_o,k^T dU_k/dx and d(reductions)/dx, d(reductions)/d(area), contracted with ghat.
__global__ void k_adj_tri( int n, const double* __restrict__ v1, const double* __restrict__ v2, const double* __restrict__ v3,
                           const double* __restrict__ afac_p, double cg, const double* __restrict__ lam,
                           const double* __restrict__ F, long ld, const double* __restrict__ ghat, AdjTri* __restrict__ out ) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   const double afac = *afac_p;
   const double* vv[ 3 ] = { v1, v2, v3 };
   AdjTri R {};
   double Fi[ 11 ][ 3 ], L[ kAdjOutputs ][ 11 ][ 3 ];
   for ( int k = 0; k < 11; ++k )
      for ( int c = 0; c < 3; ++c ) {
         Fi[ k ][ c ] = F[ (long)k * ld + 3L * i + c ];
         for ( int o = 0; o < kAdjOutputs; ++o ) L[ o ][ k ][ c ] = lam[ (long)o * 11 * ld + (long)k * ld + 3L * i + c ];
      }
   // M_o = sum_k lambda F^T (3 x 3) for the self block
   double M[ kAdjOutputs ][ 9 ] = {};
   for ( int o = 0; o < kAdjOutputs; ++o )
      for ( int k = 0; k < 11; ++k )
         for ( int a = 0; a < 3; ++a )
            for ( int b = 0; b < 3; ++b ) M[ o ][ 3 * a + b ] += L[ o ][ k ][ a ] * Fi[ k ][ b ];
   for ( int dir = 0; dir < 10; ++dir ) {
      D1 p[ 3 ][ 3 ];
      for ( int v = 0; v < 3; ++v )
         for ( int k = 0; k < 3; ++k ) p[ v ][ k ] = mk( vv[ v ][ (size_t)k * n + i ], dir == 3 * v + k ? 1.0 : 0.0 );
      const D1 af = mk( afac, dir == 9 ? 1.0 : 0.0 );
      D1 x[ 3 ], A, xmu[ 3 ], S[ 9 ];
      self_geom_t<D1>( p[ 0 ], p[ 1 ], p[ 2 ], af, cg, x, A, xmu, S );
      for ( int k = 0; k < 3; ++k ) {
         R.dx[ k ][ dir ] = x[ k ].d;
         R.dxmu[ k ][ dir ] = xmu[ k ].d;
         if ( dir == 0 ) R.x[ k ] = x[ k ].v, R.xmu[ k ] = xmu[ k ].v;
      }
      R.dA[ dir ] = A.d;
      if ( dir == 0 ) R.A = A.v;
      R.dtr[ dir ] = S[ 0 ].d + S[ 4 ].d + S[ 8 ].d;
      for ( int o = 0; o < kAdjOutputs; ++o ) {
         double t = 0.0;
         for ( int e = 0; e < 9; ++e ) t += M[ o ][ e ] * S[ e ].d;
         R.gself[ o ][ dir ] = t;
      }
   }
   // direct terms at fixed F: right-hand sides and reductions, along x (3) and the area (1)
   const double x0[ 3 ] = { R.x[ 0 ], R.x[ 1 ], R.x[ 2 ] };
   const double ar0 = afac * R.A;
   for ( int dir = 0; dir < 4; ++dir ) {
      D1 x[ 3 ], ar = mk( ar0, dir == 3 ? 1.0 : 0.0 );
      for ( int k = 0; k < 3; ++k ) x[ k ] = mk( x0[ k ], dir == k ? 1.0 : 0.0 );
      D1 u[ 11 ][ 3 ];
      rhs_t<D1>( x, u );
      D1 sums[ kAdjSlots ];
      tri_sums_t<D1, double>( x, ar, Fi, sums );
      for ( int o = 0; o < kAdjOutputs; ++o ) {
         double tu = 0.0, ts = 0.0;
         for ( int k = 0; k < 11; ++k )
            for ( int c = 0; c < 3; ++c ) tu += L[ o ][ k ][ c ] * u[ k ][ c ].d;
         for ( int q = 0; q < kAdjSlots; ++q ) ts += ghat[ o * kAdjSlots + q ] * sums[ q ].d;
         if ( dir < 3 ) {
            R.grhs[ o ][ dir ] = tu;
            R.gsx[ o ][ dir ] = ts;
         } else {
            R.gsa[ o ] = ts;
         }
      }
   }
   out[ i ] = R;
}

}  // namespace

void launch_adj_rhs( hipStream_t st, int n, const double* xc, const double* area, const double* ghat, double* y, long ld ) {
   hipLaunchKernelGGL( k_adj_rhs, dim3( ( n + 127 ) / 128 ), dim3( 128 ), 0, st, n, xc, area, ghat, y, ld );
}

template <typename T>
void launch_transpose( hipStream_t st, int m, T* a, long ld, int nb, long stride ) {
   if ( m % 32 ) throw std::runtime_error( "transpose: order not a multiple of 32" );
   const int nt = m / 32;
   hipLaunchKernelGGL( k_transpose<T>, dim3( (unsigned)( (long)nt * ( nt + 1 ) / 2 ), (unsigned)nb ), dim3( 256 ), 0, st, nt, a,
                       ld, stride );
}
template void launch_transpose<double>( hipStream_t, int, double*, long, int, long );
template void launch_transpose<float>( hipStream_t, int, float*, long, int, long );

void launch_adj_pairs( hipStream_t st, int n, int strip, double cg, const double* src, const double* xc, const double* lam,
                       const double* F, long ld, double* gsrc_part, double* gcol_part, double* gsrc, double* gcol ) {
   const int nb = ( n + 255 ) / 256, ns = ( n + strip - 1 ) / strip;
   hipLaunchKernelGGL( k_adj_pairs, dim3( nb, ns ), dim3( 256 ), 0, st, n, strip, cg, src, xc, lam, F, ld, gsrc_part, gcol_part );
   const long cs = (long)n * kAdjOutputs * 10, cc = (long)n * kAdjOutputs * 3;
   hipLaunchKernelGGL( k_adj_sum_parts, dim3( (unsigned)( ( cs + 255 ) / 256 ) ), dim3( 256 ), 0, st, cs, nb, cs, gsrc_part, gsrc );
   hipLaunchKernelGGL( k_adj_sum_parts, dim3( (unsigned)( ( cc + 255 ) / 256 ) ), dim3( 256 ), 0, st, cc, ns, cc, gcol_part, gcol );
}

void launch_adj_tri( hipStream_t st, int n, const double* v1, const double* v2, const double* v3, const double* afac, double cg,
                     const double* lam, const double* F, long ld, const double* ghat, AdjTri* out ) {
   hipLaunchKernelGGL( k_adj_tri, dim3( ( n + 63 ) / 64 ), dim3( 64 ), 0, st, n, v1, v2, v3, afac, cg, lam, F, ld, ghat, out );
}

}  // namespace bhg
