// The right-hand sides BEST solves for, the reductions of its solutions and the 3x3 algebra that
// turns the reductions into BEST's tensors, all on the GPU (FP64 throughout; the algebra is also
// callable on the host, where the adjoint differentiates it).
//
//   k_rhs           best/difco.f:28-44 (six rigid-body flows) and best/visco.f:216-243 (STRAIN,
//                   five strain flows), as 11 columns of one right-hand-side block
//   k_sums          best/difco.f:73-98 (forces, torques), best/visco.f:76-129 (BB, C of the centre
//                   of viscosity) and :191-205 (the CHI moments, with the shift to the centre of
//                   viscosity applied on the fly): per-workgroup partial sums
//   difco_algebra   best/difco.f:106-401 with MATIN3 (best/matin3.f)
//   visco_chiprime  best/visco.f:118-211
//   output_values   best/output.f:16-29, :98-167 with JACOBI (best/jacobi.f)
//   k_transport     the partial sums added in workgroup order, then the algebra above, on one
//                   thread: the whole of DIFCO/VISCO/OUTPUT's arithmetic without a host round trip
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "bh_common.h"
#include "transport.h"

namespace bhg {

// Column layout of the right-hand-side block.
constexpr int kRhsRigid = 0;   // 6 columns
constexpr int kRhsStrain = 6;  // 5 columns

__global__ void k_rhs( int n, const double* __restrict__ xc, double* __restrict__ b, long ldb, int with_strain ) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   const double x1 = xc[ i ], x2 = xc[ n + i ], x3 = xc[ 2 * n + i ];
   const long i0 = 3 * (long)i;
   double* c[ 11 ];
   for ( int k = 0; k < 11; ++k ) c[ k ] = b + (long)k * ldb + i0;
   for ( int k = 0; k < 6; ++k ) c[ k ][ 0 ] = c[ k ][ 1 ] = c[ k ][ 2 ] = 0.0;
   c[ 0 ][ 0 ] = 1.0;
   c[ 1 ][ 1 ] = 1.0;
   c[ 2 ][ 2 ] = 1.0;
   c[ 3 ][ 1 ] = -x3;
   c[ 3 ][ 2 ] = x2;
   c[ 4 ][ 0 ] = x3;
   c[ 4 ][ 2 ] = -x1;
   c[ 5 ][ 0 ] = -x2;
   c[ 5 ][ 1 ] = x1;
   if ( !with_strain ) return;
   double** s = c + kRhsStrain;
   s[ 0 ][ 0 ] = 0.5 * x2, s[ 0 ][ 1 ] = 0.5 * x1, s[ 0 ][ 2 ] = 0.0;
   s[ 1 ][ 0 ] = 0.5 * x3, s[ 1 ][ 1 ] = 0.0, s[ 1 ][ 2 ] = 0.5 * x1;
   s[ 2 ][ 0 ] = 0.0, s[ 2 ][ 1 ] = 0.5 * x3, s[ 2 ][ 2 ] = 0.5 * x2;
   s[ 3 ][ 0 ] = 0.5 * x1, s[ 3 ][ 1 ] = -0.5 * x2, s[ 3 ][ 2 ] = 0.0;
   s[ 4 ][ 0 ] = 0.5 * x1, s[ 4 ][ 1 ] = 0.0, s[ 4 ][ 2 ] = -0.5 * x3;
}

// Partial sums, one set per workgroup (kSumSlots doubles each):
//   [0, 36)   DIFCO: flow mx -> 6 mx + {0..2: -sum area F_k, 3..5: -sum area (x cross F)_k}
//   [36, 45)  CHI112, CHI213, CHI323, CHI411, CHI422, CHI433, CHI511, CHI522, CHI533 with x and
//             the strain solutions moved to the centre of viscosity d (d = 0: the centroid)
//   [45, 54)  BB(1,1), BB(1,2), BB(1,3), BB(2,2), BB(2,3), BB(3,3), C1, C2, C3 (before scaling)
__global__ void __launch_bounds__( 256 ) k_sums( int n, const double* __restrict__ xc, const double* __restrict__ area,
                                                const double* __restrict__ b, long ldb, int nstrain, double d1, double d2,
                                                double d3, const double* __restrict__ dptr, int want_bb,
                                                double* __restrict__ partial ) {
   __shared__ double red[ 4 * kSumSlots ];
   double acc[ kSumSlots ];
   if ( dptr ) d1 = dptr[ 0 ], d2 = dptr[ 1 ], d3 = dptr[ 2 ];  // the centre computed on the device
   for ( int k = 0; k < kSumSlots; ++k ) acc[ k ] = 0.0;
   const int per = ( n + gridDim.x - 1 ) / gridDim.x;
   const int lo = blockIdx.x * per, hi = min( n, lo + per );
   for ( int i = lo + threadIdx.x; i < hi; i += blockDim.x ) {
      const double x1 = xc[ i ], x2 = xc[ n + i ], x3 = xc[ 2 * n + i ];
      const double ar = area[ i ];
      const long i0 = 3 * (long)i;
      for ( int mx = 0; mx < 6; ++mx ) {
         const double* f = b + (long)( kRhsRigid + mx ) * ldb + i0;
         const double q1 = ( x2 * f[ 2 ] - x3 * f[ 1 ] ), q2 = ( x3 * f[ 0 ] - x1 * f[ 2 ] ), q3 = ( x1 * f[ 1 ] - x2 * f[ 0 ] );
         acc[ 6 * mx + 0 ] = acc[ 6 * mx + 0 ] - ar * f[ 0 ];
         acc[ 6 * mx + 1 ] = acc[ 6 * mx + 1 ] - ar * f[ 1 ];
         acc[ 6 * mx + 2 ] = acc[ 6 * mx + 2 ] - ar * f[ 2 ];
         acc[ 6 * mx + 3 ] = acc[ 6 * mx + 3 ] - ar * q1;
         acc[ 6 * mx + 4 ] = acc[ 6 * mx + 4 ] - ar * q2;
         acc[ 6 * mx + 5 ] = acc[ 6 * mx + 5 ] - ar * q3;
      }
      if ( nstrain < 5 ) continue;
      // W(l, c) at this triangle after the move to d (best/visco.f:145-166).
      double w[ 5 ][ 3 ];
      const double* e1p = b + (long)( kRhsRigid + 0 ) * ldb + i0;
      const double* e2p = b + (long)( kRhsRigid + 1 ) * ldb + i0;
      const double* e3p = b + (long)( kRhsRigid + 2 ) * ldb + i0;
      for ( int c = 0; c < 3; ++c ) {
         const double e1 = e1p[ c ], e2 = e2p[ c ], e3 = e3p[ c ];
         const double* v = b + (long)kRhsStrain * ldb + i0 + c;
         w[ 0 ][ c ] = v[ 0 * ldb ] - 0.5 * ( d2 * e1 + d1 * e2 );
         w[ 1 ][ c ] = v[ 1 * ldb ] - 0.5 * ( d3 * e1 + d1 * e3 );
         w[ 2 ][ c ] = v[ 2 * ldb ] - 0.5 * ( d3 * e2 + d2 * e3 );
         w[ 3 ][ c ] = v[ 3 * ldb ] - 0.5 * ( d1 * e1 - d2 * e2 );
         w[ 4 ][ c ] = v[ 4 * ldb ] - 0.5 * ( d1 * e1 - d3 * e3 );
      }
      const double y1 = x1 - d1, y2 = x2 - d2, y3 = x3 - d3;
      acc[ 36 ] = acc[ 36 ] + ar * ( w[ 0 ][ 0 ] * y2 + w[ 0 ][ 1 ] * y1 );
      acc[ 37 ] = acc[ 37 ] + ar * ( w[ 1 ][ 0 ] * y3 + w[ 1 ][ 2 ] * y1 );
      acc[ 38 ] = acc[ 38 ] + ar * ( w[ 2 ][ 1 ] * y3 + w[ 2 ][ 2 ] * y2 );
      acc[ 39 ] = acc[ 39 ] + ar * w[ 3 ][ 0 ] * y1;
      acc[ 40 ] = acc[ 40 ] + ar * w[ 3 ][ 1 ] * y2;
      acc[ 41 ] = acc[ 41 ] + ar * w[ 3 ][ 2 ] * y3;
      acc[ 42 ] = acc[ 42 ] + ar * w[ 4 ][ 0 ] * y1;
      acc[ 43 ] = acc[ 43 ] + ar * w[ 4 ][ 1 ] * y2;
      acc[ 44 ] = acc[ 44 ] + ar * w[ 4 ][ 2 ] * y3;
      if ( !want_bb ) continue;
      // G(l,k) = 2 V0(3n+k, l) and H(l,k,m) from the unit-translation solutions (visco.f:36-61).
      double H[ 5 ][ 3 ][ 3 ], G[ 5 ][ 3 ];
      for ( int k = 0; k < 3; ++k ) {
         const double e1 = e1p[ k ], e2 = e2p[ k ], e3 = e3p[ k ];
         for ( int l = 0; l < 5; ++l ) G[ l ][ k ] = 2.0 * b[ (long)( kRhsStrain + l ) * ldb + i0 + k ];
         H[ 0 ][ k ][ 0 ] = e2, H[ 0 ][ k ][ 1 ] = e1, H[ 0 ][ k ][ 2 ] = 0.0;
         H[ 1 ][ k ][ 0 ] = e3, H[ 1 ][ k ][ 1 ] = 0.0, H[ 1 ][ k ][ 2 ] = e1;
         H[ 2 ][ k ][ 0 ] = 0.0, H[ 2 ][ k ][ 1 ] = e3, H[ 2 ][ k ][ 2 ] = e2;
         H[ 3 ][ k ][ 0 ] = e1, H[ 3 ][ k ][ 1 ] = -e2, H[ 3 ][ k ][ 2 ] = 0.0;
         H[ 4 ][ k ][ 0 ] = e1, H[ 4 ][ k ][ 1 ] = 0.0, H[ 4 ][ k ][ 2 ] = -e3;
      }
#define H_( l, k, m ) H[ ( l )-1 ][ ( k )-1 ][ ( m )-1 ]
#define G_( l, k ) G[ ( l )-1 ][ ( k )-1 ]
      acc[ 45 ] = acc[ 45 ] + ar * ( 3 * ( H_( 1, 2, 1 ) + H_( 2, 3, 1 ) ) + 2 * ( H_( 4, 1, 1 ) + H_( 5, 1, 1 ) ) );
      acc[ 46 ] = acc[ 46 ] + ar * ( -4 * H_( 4, 2, 1 ) + 3 * ( H_( 1, 1, 1 ) + H_( 1, 2, 2 ) + H_( 2, 3, 2 ) + H_( 3, 3, 1 ) ) +
                                     2 * ( H_( 4, 1, 2 ) + H_( 5, 1, 2 ) + H_( 5, 2, 1 ) ) );
      acc[ 47 ] = acc[ 47 ] + ar * ( -4 * H_( 5, 3, 1 ) + 3 * ( H_( 1, 2, 3 ) + H_( 2, 1, 1 ) + H_( 2, 3, 3 ) + H_( 3, 2, 1 ) ) +
                                     2 * ( H_( 4, 1, 3 ) + H_( 4, 3, 1 ) + H_( 5, 1, 3 ) ) );
      acc[ 48 ] = acc[ 48 ] + ar * ( -4 * H_( 4, 2, 2 ) + 3 * ( H_( 1, 1, 2 ) + H_( 3, 3, 2 ) ) + 2 * H_( 5, 2, 2 ) );
      acc[ 49 ] = acc[ 49 ] + ar * ( -4 * H_( 4, 2, 3 ) + 3 * ( H_( 3, 2, 2 ) + H_( 3, 3, 3 ) + H_( 1, 1, 3 ) + H_( 2, 1, 2 ) ) +
                                     2 * ( H_( 4, 3, 2 ) + H_( 5, 2, 3 ) ) );
      acc[ 50 ] = acc[ 50 ] + ar * ( -4 * H_( 5, 3, 3 ) + 3 * ( H_( 2, 1, 3 ) + H_( 3, 2, 3 ) ) + 2 * H_( 4, 3, 3 ) );
      acc[ 51 ] = acc[ 51 ] + ar * ( -3 * G_( 1, 2 ) - 3 * G_( 2, 3 ) - 2 * G_( 4, 1 ) - 2 * G_( 5, 1 ) -
                                     x1 * ( 3 * ( H_( 1, 2, 1 ) + H_( 2, 3, 1 ) ) + 2 * ( H_( 4, 1, 1 ) + H_( 5, 1, 1 ) ) ) -
                                     x2 * ( 3 * ( H_( 1, 1, 1 ) + H_( 3, 3, 1 ) ) - 4 * H_( 4, 2, 1 ) + 2 * H_( 5, 2, 1 ) ) -
                                     x3 * ( 3 * ( H_( 2, 1, 1 ) + H_( 3, 2, 1 ) ) - 4 * H_( 5, 3, 1 ) + 2 * H_( 4, 3, 1 ) ) );
      acc[ 52 ] = acc[ 52 ] + ar * ( -3 * G_( 1, 1 ) - 3 * G_( 3, 3 ) + 4 * G_( 4, 2 ) - 2 * G_( 5, 2 ) -
                                     x1 * ( 3 * ( H_( 1, 2, 2 ) + H_( 2, 3, 2 ) ) + 2 * ( H_( 4, 1, 2 ) + H_( 5, 1, 2 ) ) ) -
                                     x2 * ( 3 * ( H_( 1, 1, 2 ) + H_( 3, 3, 2 ) ) - 4 * H_( 4, 2, 2 ) + 2 * H_( 5, 2, 2 ) ) -
                                     x3 * ( 3 * ( H_( 2, 1, 2 ) + H_( 3, 2, 2 ) ) - 4 * H_( 5, 3, 2 ) + 2 * H_( 4, 3, 2 ) ) );
      acc[ 53 ] = acc[ 53 ] + ar * ( -3 * G_( 2, 1 ) - 3 * G_( 3, 2 ) - 2 * G_( 4, 3 ) + 4 * G_( 5, 3 ) -
                                     x1 * ( 3 * ( H_( 1, 2, 3 ) + H_( 2, 3, 3 ) ) + 2 * ( H_( 4, 1, 3 ) + H_( 5, 1, 3 ) ) ) -
                                     x2 * ( 3 * ( H_( 1, 1, 3 ) + H_( 3, 3, 3 ) ) - 4 * H_( 4, 2, 3 ) + 2 * H_( 5, 2, 3 ) ) -
                                     x3 * ( 3 * ( H_( 2, 1, 3 ) + H_( 3, 2, 3 ) ) - 4 * H_( 5, 3, 3 ) + 2 * H_( 4, 3, 3 ) ) );
#undef H_
#undef G_
   }
   // Fixed-order reduction over the workgroup: each slot by a butterfly within each wave (the same
   // lanes added in the same order every time), the four waves' sums added in wave order after one
   // barrier.
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for ( int k = 0; k < kSumSlots; ++k ) {
      double v = acc[ k ];
#pragma unroll
      for ( int o = 32; o > 0; o >>= 1 ) v = v + __shfl_xor( v, o, 64 );
      if ( lane == 0 ) red[ wave * kSumSlots + k ] = v;
   }
   __syncthreads();
   if ( threadIdx.x < kSumSlots ) {
      const int k = threadIdx.x;
      partial[ (long)blockIdx.x * kSumSlots + k ] = ( red[ k ] + red[ kSumSlots + k ] ) + ( red[ 2 * kSumSlots + k ] + red[ 3 * kSumSlots + k ] );
   }
}

void launch_rhs( hipStream_t st, int n, const double* xc, double* b, long ldb, int with_strain ) {
   hipLaunchKernelGGL( k_rhs, dim3( ( n + 255 ) / 256 ), dim3( 256 ), 0, st, n, xc, b, ldb, with_strain );
}

void launch_sums( hipStream_t st, int n, const double* xc, const double* area, const double* b, long ldb, int nstrain,
                  const double d[ 3 ], int want_bb, double* partial ) {
   hipLaunchKernelGGL( k_sums, dim3( kSumGroups ), dim3( 256 ), 0, st, n, xc, area, b, ldb, nstrain, d[ 0 ], d[ 1 ], d[ 2 ],
                       nullptr, want_bb, partial );
}

// ------------------------------------------------------------------ 3x3 algebra (device and host)

namespace {

// No contraction of a * b + c into fused multiply-adds in the algebra, on the device as on the
// host: the two give the same bits from the same sums.
#pragma clang fp contract( off )

#define BH_HD __host__ __device__

// A value and one derivative (forward mode), for the exact derivatives of the algebra below.
struct Dual {
   double v, d;
   BH_HD Dual( double x = 0.0, double dx = 0.0 ) : v( x ), d( dx ) {}
};
BH_HD inline Dual operator+( Dual a, Dual b ) { return Dual( a.v + b.v, a.d + b.d ); }
BH_HD inline Dual operator-( Dual a, Dual b ) { return Dual( a.v - b.v, a.d - b.d ); }
BH_HD inline Dual operator-( Dual a ) { return Dual( -a.v, -a.d ); }
BH_HD inline Dual operator*( Dual a, Dual b ) { return Dual( a.v * b.v, a.d * b.v + a.v * b.d ); }
BH_HD inline Dual operator/( Dual a, Dual b ) {
   const double q = a.v / b.v;
   return Dual( q, ( a.d - q * b.d ) / b.v );
}
BH_HD inline double val( double x ) { return x; }
BH_HD inline double val( Dual x ) { return x.v; }
BH_HD inline double tabs( double x ) { return fabs( x ); }
BH_HD inline Dual tabs( Dual x ) { return x.v < 0.0 ? -x : x; }

// 3x3 matrices column-major: m[i + 3 j] is (i, j), 0-based.
template <typename T>
BH_HD inline T& at( T* m, int i, int j ) {
   return m[ i + 3 * j ];
}
template <typename T>
BH_HD inline T at( const T* m, int i, int j ) {
   return m[ i + 3 * j ];
}
template <typename T>
BH_HD inline void copy9( T* d, const T* s ) {
   for ( int k = 0; k < 9; ++k ) d[ k ] = s[ k ];
}

// The inverse of a 3x3 matrix by LU with implicit (scaled) partial pivoting and back
// substitution per unit column (best/matin3.f).  a is overwritten.  Returns nonzero when a row is
// all zeros.  The pivots are chosen on the values (T = Dual carries the derivatives along).
template <typename T>
BH_HD int inverse3( T* a, T* y ) {
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) at( y, i, j ) = T( i == j ? 1.0 : 0.0 );
   double scale[ 3 ];
   for ( int i = 0; i < 3; ++i ) {
      double big = 0.0;
      for ( int j = 0; j < 3; ++j ) big = fabs( val( at( a, i, j ) ) ) > big ? fabs( val( at( a, i, j ) ) ) : big;
      if ( big == 0.0 ) return 1;
      scale[ i ] = 1.0 / big;
   }
   int perm[ 3 ];
   for ( int j = 0; j < 3; ++j ) {
      for ( int i = 1; i < j; ++i ) {
         T s = at( a, i, j );
         for ( int k = 0; k < i; ++k ) s = s - at( a, i, k ) * at( a, k, j );
         at( a, i, j ) = s;
      }
      double big = 0.0;
      int imax = j;
      for ( int i = j; i < 3; ++i ) {
         T s = at( a, i, j );
         for ( int k = 0; k < j; ++k ) s = s - at( a, i, k ) * at( a, k, j );
         if ( j > 0 ) at( a, i, j ) = s;
         const double r = scale[ i ] * fabs( val( s ) );
         if ( r >= big ) big = r, imax = i;
      }
      if ( imax != j ) {
         for ( int k = 0; k < 3; ++k ) {
            const T t = at( a, imax, k );
            at( a, imax, k ) = at( a, j, k );
            at( a, j, k ) = t;
         }
         scale[ imax ] = scale[ j ];
      }
      perm[ j ] = imax;
      if ( j != 2 ) {
         if ( val( at( a, j, j ) ) == 0.0 ) at( a, j, j ) = T( 1e-20 );
         const T r = T( 1.0 ) / at( a, j, j );
         for ( int i = j + 1; i < 3; ++i ) at( a, i, j ) = at( a, i, j ) * r;
      }
   }
   if ( val( at( a, 2, 2 ) ) == 0.0 ) at( a, 2, 2 ) = T( 1e-20 );
   for ( int c = 0; c < 3; ++c ) {
      T* v = y + 3 * c;
      int first = 0;
      for ( int i = 0; i < 3; ++i ) {
         T s = v[ perm[ i ] ];
         v[ perm[ i ] ] = v[ i ];
         if ( first != 0 ) {
            for ( int j = first - 1; j < i; ++j ) s = s - at( a, i, j ) * v[ j ];
         } else if ( val( s ) != 0.0 ) {
            first = 1;
         }
         v[ i ] = s;
      }
      for ( int i = 2; i >= 0; --i ) {
         T s = v[ i ];
         for ( int j = i + 1; j < 3; ++j ) s = s - at( a, i, j ) * v[ j ];
         v[ i ] = s / at( a, i, i );
      }
   }
   return 0;
}

// Cyclic Jacobi eigenvalues of a symmetric 3x3 matrix read from its upper triangle
// (best/jacobi.f).  a is destroyed.
BH_HD void eigen3( double* a, double d[ 3 ] ) {
   double b[ 3 ], z[ 3 ];
   for ( int p = 0; p < 3; ++p ) b[ p ] = d[ p ] = at( a, p, p ), z[ p ] = 0.0;
   for ( int sweep = 1; sweep <= 50; ++sweep ) {
      double off = 0.0;
      for ( int p = 0; p < 2; ++p )
         for ( int q = p + 1; q < 3; ++q ) off = off + fabs( at( a, p, q ) );
      if ( off == 0.0 ) return;
      const double tresh = sweep < 4 ? 0.2 * off / 9.0 : 0.0;
      for ( int p = 0; p < 2; ++p )
         for ( int q = p + 1; q < 3; ++q ) {
            const double apq = at( a, p, q );
            const double g = 100.0 * fabs( apq );
            if ( sweep > 4 && fabs( d[ p ] ) + g == fabs( d[ p ] ) && fabs( d[ q ] ) + g == fabs( d[ q ] ) ) {
               at( a, p, q ) = 0.0;
               continue;
            }
            if ( !( fabs( apq ) > tresh ) ) continue;
            double h = d[ q ] - d[ p ], t;
            if ( fabs( h ) + g == fabs( h ) ) {
               t = apq / h;
            } else {
               const double theta = 0.5 * h / apq;
               t = 1.0 / ( fabs( theta ) + sqrt( 1.0 + theta * theta ) );
               if ( theta < 0.0 ) t = -t;
            }
            const double c = 1.0 / sqrt( 1 + t * t ), s = t * c, tau = s / ( 1.0 + c );
            h = t * apq;
            z[ p ] = z[ p ] - h, z[ q ] = z[ q ] + h;
            d[ p ] = d[ p ] - h, d[ q ] = d[ q ] + h;
            at( a, p, q ) = 0.0;
            auto rot = [&]( double& x, double& y ) {
               const double gx = x, hy = y;
               x = gx - s * ( hy + gx * tau );
               y = hy + s * ( gx - hy * tau );
            };
            for ( int j = 0; j < p; ++j ) rot( at( a, j, p ), at( a, j, q ) );
            for ( int j = p + 1; j < q; ++j ) rot( at( a, p, j ), at( a, j, q ) );
            for ( int j = q + 1; j < 3; ++j ) rot( at( a, p, j ), at( a, q, j ) );
         }
      for ( int p = 0; p < 3; ++p ) b[ p ] = b[ p ] + z[ p ], d[ p ] = b[ p ], z[ p ] = 0.0;
   }
}

// uu = the matrix of the centre vector r of best/difco.f:180-188 (the cross-product matrix,
// with r(1) at (3,2)).
template <typename T>
BH_HD void cross_matrix( const T r[ 3 ], T* uu ) {
   at( uu, 2, 0 ) = -r[ 1 ];
   at( uu, 2, 1 ) = r[ 0 ];
   at( uu, 0, 1 ) = -r[ 2 ];
   at( uu, 1, 0 ) = -at( uu, 0, 1 );
   at( uu, 1, 2 ) = -at( uu, 2, 1 );
   at( uu, 0, 2 ) = -at( uu, 2, 0 );
   at( uu, 0, 0 ) = at( uu, 1, 1 ) = at( uu, 2, 2 ) = T( 0.0 );
}

}  // namespace

// DIFCO's algebra (best/difco.f:106-401) on the 36 force and torque sums, into the tensors of o.
template <typename R>
struct DifcoOut {
   R dtt[ 9 ], dtr[ 9 ], drr[ 9 ], urc[ 3 ], udr[ 3 ], ktt[ 9 ], ktr[ 9 ], krt[ 9 ], krr[ 9 ];
};

template <typename R>
BH_HD void difco_t( const R* s, double cg, DifcoOut<R>& o ) {
   R ktt[ 9 ], ktr[ 9 ], krt[ 9 ], krr[ 9 ], uu[ 9 ], aux[ 9 ], kinv[ 9 ], work[ 9 ];
   // Friction tensors: column i of Ktt/Ktr from translation i, of Krt/Krr from rotation i.
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         at( ktt, j, i ) = -s[ 6 * i + j ] * R( 1.0 / 1.0 );
         at( ktr, j, i ) = -s[ 6 * i + 3 + j ] * R( 1.0 / 1.0 );
         at( krt, j, i ) = -s[ 6 * ( 3 + i ) + j ] * R( 1.0 / 1.0 );
         at( krr, j, i ) = -s[ 6 * ( 3 + i ) + 3 + j ] * R( 1.0 / 1.0 );
      }
   copy9( o.ktr, ktr );
   // Centre of resistance (difco.f:141-174).
   at( uu, 0, 0 ) = at( ktt, 1, 1 ) + at( ktt, 2, 2 );
   at( uu, 1, 1 ) = at( ktt, 0, 0 ) + at( ktt, 2, 2 );
   at( uu, 2, 2 ) = at( ktt, 1, 1 ) + at( ktt, 0, 0 );
   at( uu, 0, 1 ) = -at( ktt, 0, 1 );
   at( uu, 0, 2 ) = -at( ktt, 0, 2 );
   at( uu, 2, 1 ) = -at( ktt, 1, 2 );
   at( uu, 1, 0 ) = at( uu, 0, 1 );
   at( uu, 1, 2 ) = at( uu, 2, 1 );
   at( uu, 2, 0 ) = at( uu, 0, 2 );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) at( krt, i, j ) = ( at( krt, i, j ) + at( ktr, j, i ) ) * R( 0.5 );
   const R sur1 = -at( krt, 2, 1 ) + at( krt, 1, 2 );
   const R sur2 = -at( krt, 0, 2 ) + at( krt, 2, 0 );
   const R sur3 = -at( krt, 1, 0 ) + at( krt, 0, 1 );
   inverse3( uu, aux );
   for ( int j = 0; j < 3; ++j ) o.urc[ j ] = ( at( aux, j, 0 ) * sur1 + at( aux, j, 1 ) * sur2 + at( aux, j, 2 ) * sur3 );
   cross_matrix( o.urc, uu );
   // Krr' = Krr - uu Ktt uu + Krt^T uu - uu Krt,  Krt' = Krt + Ktt uu  (difco.f:196-214).
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         R t = R( 0.0 );
         for ( int k = 0; k < 3; ++k )
            t = t + at( uu, i, k ) * ( at( ktt, k, 0 ) * at( uu, 0, j ) + at( ktt, k, 1 ) * at( uu, 1, j ) + at( ktt, k, 2 ) * at( uu, 2, j ) );
         const R t1 = at( krt, 0, i ) * at( uu, 0, j ) + at( krt, 1, i ) * at( uu, 1, j ) + at( krt, 2, i ) * at( uu, 2, j );
         const R t2 = at( uu, i, 0 ) * at( krt, 0, j ) + at( uu, i, 1 ) * at( krt, 1, j ) + at( uu, i, 2 ) * at( krt, 2, j );
         at( krr, i, j ) = at( krr, i, j ) - t + t1 - t2;
      }
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         at( krt, i, j ) = at( krt, i, j ) + at( ktt, i, 0 ) * at( uu, 0, j ) + at( ktt, i, 1 ) * at( uu, 1, j ) + at( ktt, i, 2 ) * at( uu, 2, j );
   copy9( o.ktt, ktt );
   copy9( o.krt, krt );
   copy9( o.krr, krr );
   // Dtt, Dtr, Drr with the coupling (difco.f:244-330).
   copy9( work, ktt );
   inverse3( work, kinv );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         R t = R( 0.0 );
         for ( int k = 0; k < 3; ++k )
            t = t + at( krt, i, k ) * ( at( kinv, k, 0 ) * at( krt, j, 0 ) + at( kinv, k, 1 ) * at( krt, j, 1 ) + at( kinv, k, 2 ) * at( krt, j, 2 ) );
         at( work, i, j ) = at( krr, i, j ) - t;
      }
   inverse3( work, o.drr );
   if ( cg == 0 )
      for ( int k = 0; k < 9; ++k ) o.drr[ k ] = R( 4.0 ) * o.drr[ k ] / R( 3.0 );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         R t = R( 0.0 );
         for ( int k = 0; k < 3; ++k )
            t = t + at( kinv, i, k ) * ( at( krt, 0, k ) * at( o.drr, 0, j ) + at( krt, 1, k ) * at( o.drr, 1, j ) + at( krt, 2, k ) * at( o.drr, 2, j ) );
         at( o.dtr, i, j ) = -t;
         if ( cg == 0 ) at( o.dtr, i, j ) = R( 4.0 ) * at( o.dtr, i, j ) / R( 3.0 );
      }
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         R t = R( 0.0 );
         for ( int k = 0; k < 3; ++k )
            t = t + at( kinv, i, k ) * ( at( krt, 0, k ) * at( o.dtr, j, 0 ) + at( krt, 1, k ) * at( o.dtr, j, 1 ) + at( krt, 2, k ) * at( o.dtr, j, 2 ) );
         at( o.dtt, i, j ) = at( kinv, i, j ) - t;
         if ( cg == 0 ) at( o.dtt, i, j ) = R( 4.0 ) * at( o.dtt, i, j ) / R( 3.0 );
      }
   // Centre of diffusion (difco.f:344-401).
   const R* drr = o.drr;
   at( uu, 0, 0 ) = at( drr, 1, 1 ) + at( drr, 2, 2 );
   at( uu, 1, 1 ) = at( drr, 0, 0 ) + at( drr, 2, 2 );
   at( uu, 2, 2 ) = at( drr, 1, 1 ) + at( drr, 0, 0 );
   at( uu, 0, 1 ) = -at( drr, 0, 1 );
   at( uu, 0, 2 ) = -at( drr, 0, 2 );
   at( uu, 2, 1 ) = -at( drr, 2, 1 );
   at( uu, 1, 0 ) = at( uu, 0, 1 );
   at( uu, 1, 2 ) = at( uu, 2, 1 );
   at( uu, 2, 0 ) = at( uu, 0, 2 );
   const R s1 = at( o.dtr, 2, 1 ) - at( o.dtr, 1, 2 );
   const R s2 = at( o.dtr, 0, 2 ) - at( o.dtr, 2, 0 );
   const R s3 = at( o.dtr, 1, 0 ) - at( o.dtr, 0, 1 );
   inverse3( uu, aux );
   for ( int j = 0; j < 3; ++j ) o.udr[ j ] = ( at( aux, j, 0 ) * s1 + at( aux, j, 1 ) * s2 + at( aux, j, 2 ) * s3 );
   cross_matrix( o.udr, uu );
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j ) {
         R t = R( 0.0 );
         for ( int k = 0; k < 3; ++k )
            t = t + at( uu, i, k ) * at( drr, k, 0 ) * at( uu, 0, j ) + at( uu, i, k ) * at( drr, k, 1 ) * at( uu, 1, j ) +
                at( uu, i, k ) * at( drr, k, 2 ) * at( uu, 2, j );
         const R t1 = at( o.dtr, i, 0 ) * at( uu, 0, j ) + at( o.dtr, i, 1 ) * at( uu, 1, j ) + at( o.dtr, i, 2 ) * at( uu, 2, j );
         const R t2 = -at( uu, i, 0 ) * at( o.dtr, j, 0 ) - at( uu, i, 1 ) * at( o.dtr, j, 1 ) - at( uu, i, 2 ) * at( o.dtr, j, 2 );
         at( o.dtt, i, j ) = at( o.dtt, i, j ) - t + t1 + t2;
      }
   for ( int i = 0; i < 3; ++i )
      for ( int j = 0; j < 3; ++j )
         at( o.dtr, i, j ) = at( o.dtr, i, j ) - at( uu, i, 0 ) * at( drr, 0, j ) - at( uu, i, 1 ) * at( drr, 1, j ) -
                             at( uu, i, 2 ) * at( drr, 2, j );
}

BH_HD void difco_algebra( const double* s, double cg, Transport& T ) {
   DifcoOut<double> o;
   difco_t<double>( s, cg, o );
   copy9( T.dtt, o.dtt ), copy9( T.dtr, o.dtr ), copy9( T.drr, o.drr );
   for ( int k = 0; k < 3; ++k ) T.urc[ k ] = o.urc[ k ], T.udr[ k ] = o.udr[ k ];
   copy9( T.ktt, o.ktt ), copy9( T.ktr, o.ktr ), copy9( T.krt, o.krt ), copy9( T.krr, o.krr );
}

BH_HD void visco_centre( const double* s, double d[ 3 ] ) {
   double bb[ 9 ], bbi[ 9 ];
   at( bb, 0, 0 ) = s[ 45 ] / 15.;
   at( bb, 0, 1 ) = s[ 46 ] / 30.;
   at( bb, 0, 2 ) = s[ 47 ] / 30.;
   at( bb, 1, 0 ) = at( bb, 0, 1 );
   at( bb, 1, 1 ) = s[ 48 ] / 15.;
   at( bb, 1, 2 ) = s[ 49 ] / 30.;
   at( bb, 2, 0 ) = at( bb, 0, 2 );
   at( bb, 2, 1 ) = at( bb, 1, 2 );
   at( bb, 2, 2 ) = s[ 50 ] / 15.;
   const double c1 = s[ 51 ] / 30., c2 = s[ 52 ] / 30., c3 = s[ 53 ] / 30.;
   inverse3( bb, bbi );
   d[ 0 ] = at( bbi, 0, 0 ) * c1 + at( bbi, 0, 1 ) * c2 + at( bbi, 0, 2 ) * c3;
   d[ 1 ] = at( bbi, 1, 0 ) * c1 + at( bbi, 1, 1 ) * c2 + at( bbi, 1, 2 ) * c3;
   d[ 2 ] = at( bbi, 2, 0 ) * c1 + at( bbi, 2, 1 ) * c2 + at( bbi, 2, 2 ) * c3;
}

template <typename R>
BH_HD R chiprime_t( const R* s, double cg ) {
   const double pi = 3.1415926535897932385;
   const R c112 = 0.5 * s[ 36 ], c213 = 0.5 * s[ 37 ], c323 = 0.5 * s[ 38 ];
   R chip = 8 * pi * ( 0.2 * ( c112 + c213 + c323 ) + ( s[ 39 ] - 2 * s[ 40 ] + s[ 41 ] + s[ 42 ] + s[ 43 ] - 2 * s[ 44 ] ) / 15.0 );
   if ( cg == 0 ) chip = 0.75 * chip;
   return chip;
}

BH_HD double visco_chiprime( const double* s, double cg ) { return chiprime_t<double>( s, cg ); }

BH_HD void output_values( Transport& T ) {
   const double temperature = 293.15, eta = 0.01002, pi = 3.141592653;
   const double tunit = ( 1.38066E-16 ) * temperature * 1.0E8 / ( 8.0 * eta * pi );
   const double runit = tunit * 1.0e16;
   double a[ 9 ];
   copy9( a, T.drr );
   eigen3( a, T.drr_eig );
   T.dr = runit * ( ( T.drr_eig[ 0 ] + T.drr_eig[ 1 ] + T.drr_eig[ 2 ] ) / 3. );
   copy9( a, T.dtt );
   eigen3( a, T.dtt_eig );
   T.dtrans = tunit * ( ( T.dtt_eig[ 0 ] + T.dtt_eig[ 1 ] + T.dtt_eig[ 2 ] ) / 3. );
}

// d(D_t, D_r, CHI * V)/d(sum q) for the first nslots sums, exactly (forward mode through the
// algebra: one pass per sum).  D_t and D_r as output_values prints them are 1/3 of the traces of
// Dtt and Drr in units (JACOBI's eigenvalues add up to the trace it reads), so their derivatives
// are those of the traces; CHI * V is linear in the sums.  ghat[o * nslots + q].
void algebra_gradient( const double* sums, double cg, int nslots, double* ghat ) {
   const double temperature = 293.15, eta = 0.01002, pi = 3.141592653;
   const double tunit = ( 1.38066E-16 ) * temperature * 1.0E8 / ( 8.0 * eta * pi );
   const double runit = tunit * 1.0e16;
   Dual s[ kSumSlots ];
   for ( int k = 0; k < kSumSlots; ++k ) s[ k ] = Dual( sums[ k ] );
   for ( int q = 0; q < nslots; ++q ) {
      s[ q ].d = 1.0;
      DifcoOut<Dual> o;
      difco_t<Dual>( s, cg, o );
      const Dual chip = chiprime_t<Dual>( s, cg );
      ghat[ 0 * nslots + q ] = tunit * ( ( o.dtt[ 0 ].d + o.dtt[ 4 ].d + o.dtt[ 8 ].d ) / 3. );
      ghat[ 1 * nslots + q ] = runit * ( ( o.drr[ 0 ].d + o.drr[ 4 ].d + o.drr[ 8 ].d ) / 3. );
      ghat[ 2 * nslots + q ] = chip.d;
      s[ q ].d = 0.0;
   }
}

// JACOBI's eigenvalues of a symmetric 3x3 matrix (upper triangle of a, column-major; a is not
// modified), in JACOBI's order.
void jacobi3( const double* a, double d[ 3 ] ) {
   double w[ 9 ];
   copy9( w, a );
   eigen3( w, d );
}

// The kSumSlots sums from the workgroups' partials, each added in workgroup order (the order
// the host used), one slot per thread of the workgroup, into LDS.
__device__ void total_sums( const double* __restrict__ partial, double* s ) {
   const int k = threadIdx.x;
   if ( k < kSumSlots ) {
      double t = 0.0;
      for ( int g = 0; g < kSumGroups; ++g ) t = t + partial[ (long)g * kSumSlots + k ];
      s[ k ] = t;
   }
   __syncthreads();
}

// DIFCO, VISCO (at the centroid, or the centre of viscosity when icov: the chiprime then comes
// from k_transport_chi after the sums at that centre) and OUTPUT: the sums on 64 threads, the
// algebra on one.
__global__ void __launch_bounds__( 64 ) k_transport( const double* __restrict__ partial, int ivisco, int icov, double cg, Transport* out ) {
   __shared__ double sh[ kSumSlots ];
   total_sums( partial, sh );
   if ( threadIdx.x ) return;
   double s[ kSumSlots ];
   for ( int k = 0; k < kSumSlots; ++k ) s[ k ] = sh[ k ];
   Transport T {};
   difco_algebra( s, cg, T );
   T.chiprime = 0.0;
   if ( ivisco ) {
      if ( icov ) visco_centre( s, T.dvisc );
      else T.chiprime = visco_chiprime( s, cg );
   }
   output_values( T );
   *out = T;
}

__global__ void __launch_bounds__( 64 ) k_transport_chi( const double* __restrict__ partial, double cg, Transport* out ) {
   __shared__ double sh[ kSumSlots ];
   total_sums( partial, sh );
   if ( threadIdx.x ) return;
   double s[ kSumSlots ];
   for ( int k = 0; k < kSumSlots; ++k ) s[ k ] = sh[ k ];
   out->chiprime = visco_chiprime( s, cg );
}

void launch_transport( hipStream_t st, int n, const double* xc, const double* area, const double* b, long ldb, int ivisco,
                       int icov, double cg, double* partial, Transport* d_out ) {
   hipLaunchKernelGGL( k_sums, dim3( kSumGroups ), dim3( 256 ), 0, st, n, xc, area, b, ldb, ivisco ? 5 : 0, 0.0, 0.0, 0.0,
                       nullptr, icov ? 1 : 0, partial );
   hipLaunchKernelGGL( k_transport, dim3( 1 ), dim3( 64 ), 0, st, partial, ivisco, icov, cg, d_out );
   if ( ivisco && icov ) {
      const double* d_centre = reinterpret_cast<const double*>( reinterpret_cast<const char*>( d_out ) + offsetof( Transport, dvisc ) );
      hipLaunchKernelGGL( k_sums, dim3( kSumGroups ), dim3( 256 ), 0, st, n, xc, area, b, ldb, 5, 0.0, 0.0, 0.0, d_centre, 0,
                          partial );
      hipLaunchKernelGGL( k_transport_chi, dim3( 1 ), dim3( 64 ), 0, st, partial, cg, d_out );
   }
}

}  // namespace bhg
