// Assembly of BEST's boundary-element matrix GR on the GPU (FP64 VALU).
//
//   k_prep        best/best.f:261-315   centroid shift, triangle areas, area factor
//   k_self        best/tgetgr.f:35-145  incenters, exact self blocks, the normal field XMU
//                                       (best/best.f:387-399), and the per-source records
//   k_deflation_weights  best/best.f:384-405, best/compon.f90:185-197  DIAGM and GDEF(K),
//                 summed in the reference's order by one thread so the weights are its bits
//   k_pairs       best/tgetgr.f:156-173 + best/triangleint.f:1-247  the off-diagonal 3x3 blocks,
//                 with the deflation GDEF(K) XMU XMU^T (best/best.f:405-426) fused into the store
//   k_entries     hmatrix/best_hmatrix.h:38-40  a sub-block of GR before deflation
//
// This file is compiled with -ffp-contract=off and keeps the reference's association order and
// literals, so entries match the reference to the last bit wherever the device libm's log,
// acos, sin, cos agree with the host's (they differ by at most an ulp where they do not).
//
// Layout: vertices and incenters are BEST's NAR x 3 column-major arrays (x[c*n + i]); GR is
// column-major with leading dimension ld; unknown 3*i + a is component a on triangle i.
#include "bh_common.h"
#include "pair_kernel.h"

namespace bhg {

namespace {

// best/tgetgr.f:177-187
__device__ inline double fn_si( double x ) { return 2.0 * ( log( ( 1.0 + sqrt( 1.0 - x * x ) ) / x ) - sqrt( 1.0 - x * x ) ); }
__device__ inline double fn_csi( double x ) { return 3.0 - 1.0 / x - 2.0 * x; }

}  // namespace

// Fixed-order tree sum over a 1024-thread workgroup (every thread gets the result).
__device__ inline double block_sum( double v, double* red ) {
   red[ threadIdx.x ] = v;
   __syncthreads();
   for ( int w = 512; w > 0; w >>= 1 ) {
      if ( threadIdx.x < w ) red[ threadIdx.x ] = red[ threadIdx.x ] + red[ threadIdx.x + w ];
      __syncthreads();
   }
   const double t = red[ 0 ];
   __syncthreads();
   return t;
}

// best/best.f:261-315 in one workgroup of 1024 threads: centroid of the triangle centroids,
// shift, areas, total area and the area factor.  The two sums are the reference's running sums,
// in its order and bit for bit: the workgroup computes a chunk of the terms (each exactly as the
// reference computes it) into LDS, then one lane per coordinate adds them in triangle order.  The
// pair integrals of nearly aligned triangle pairs are sensitive to the last bits of the shifted
// coordinates (an entry moved by 2e-8 relative when the centroid differed by 6e-15 A), so the
// shift must be the reference's exactly.
// run + t[0] + t[1] + ... + t[m - 1], added one at a time in that order (the running sum), with
// the LDS reads of 16 terms issued ahead of their additions.
__device__ inline double ordered_sum( double run, const double* t, int m ) {
   int q = 0;
   for ( ; q + 16 <= m; q += 16 ) {
      double v[ 16 ];
#pragma unroll
      for ( int u = 0; u < 16; ++u ) v[ u ] = t[ q + u ];
#pragma unroll
      for ( int u = 0; u < 16; ++u ) run = run + v[ u ];
   }
   for ( ; q < m; ++q ) run = run + t[ q ];
   return run;
}

__global__ void __launch_bounds__( 1024 ) k_prep( int n, double* v1, double* v2, double* v3, int iarea, double truearea,
                                                 double* area, double* scal /* [0]=afac [1]=totarea [2..4]=centroid */ ) {
   constexpr int kChunk = 1024;
   __shared__ double term[ 3 * kChunk ];
   __shared__ double shared_sum[ 3 ];
   double run = 0.0;  // lanes 0..2: the running sum of coordinate threadIdx.x
   for ( int j0 = 0; j0 < n; j0 += kChunk ) {
      const int j = j0 + threadIdx.x;
      if ( j < n )
         for ( int k = 0; k < 3; ++k ) term[ k * kChunk + threadIdx.x ] = ( v1[ k * n + j ] + v2[ k * n + j ] + v3[ k * n + j ] ) / 3.0;
      __syncthreads();
      if ( threadIdx.x < 3 ) run = ordered_sum( run, term + threadIdx.x * kChunk, min( kChunk, n - j0 ) );
      __syncthreads();
   }
   if ( threadIdx.x < 3 ) shared_sum[ threadIdx.x ] = run / (double)n;
   __syncthreads();
   const double cen[ 3 ] = { shared_sum[ 0 ], shared_sum[ 1 ], shared_sum[ 2 ] };
   double tot = 0.0;  // lane 0: the running total area
   for ( int i0 = 0; i0 < n; i0 += kChunk ) {
      const int i = i0 + threadIdx.x;
      if ( i < n ) {
         for ( int k = 0; k < 3; ++k ) {
            v1[ k * n + i ] = v1[ k * n + i ] - cen[ k ];
            v2[ k * n + i ] = v2[ k * n + i ] - cen[ k ];
            v3[ k * n + i ] = v3[ k * n + i ] - cen[ k ];
         }
         const double ax = v2[ i ] - v1[ i ], ay = v2[ n + i ] - v1[ n + i ], az = v2[ 2 * n + i ] - v1[ 2 * n + i ];
         const double bx = v3[ i ] - v1[ i ], by = v3[ n + i ] - v1[ n + i ], bz = v3[ 2 * n + i ] - v1[ 2 * n + i ];
         const double px = by * az - bz * ay, py = bz * ax - bx * az, pz = bx * ay - by * ax;
         const double ar = sqrt( px * px + py * py + pz * pz ) * 0.5;
         area[ i ] = ar;
         term[ threadIdx.x ] = ar;
      }
      __syncthreads();
      if ( threadIdx.x == 0 ) tot = ordered_sum( tot, term, min( kChunk, n - i0 ) );
      __syncthreads();
   }
   if ( threadIdx.x == 0 ) shared_sum[ 0 ] = tot;
   __syncthreads();
   tot = shared_sum[ 0 ];
   const double afac = iarea == 1 ? truearea / tot : 1.0;
   if ( threadIdx.x == 0 ) {
      scal[ 0 ] = afac;
      scal[ 1 ] = tot;
      scal[ 2 ] = cen[ 0 ];
      scal[ 3 ] = cen[ 1 ];
      scal[ 4 ] = cen[ 2 ];
   }
   if ( iarea == 1 )
      for ( int i = threadIdx.x; i < n; i += blockDim.x ) area[ i ] = afac * area[ i ];
}

// best/tgetgr.f:35-145 for every triangle, plus XMU (best/best.f:387-399) and the source record.
__global__ void __launch_bounds__( 256 ) k_self( int n, const double* __restrict__ afac_p, double cg, const double* __restrict__ v1,
                                                const double* __restrict__ v2, const double* __restrict__ v3,
                                                const double* __restrict__ area, double* __restrict__ xc,
                                                double* __restrict__ selfb, double* __restrict__ xmu,
                                                double* __restrict__ src ) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   const double afac = *afac_p;
   const double pi = 3.1415926535897932385;
   const double p1x = v1[ i ], p1y = v1[ n + i ], p1z = v1[ 2 * n + i ];
   const double p2x = v2[ i ], p2y = v2[ n + i ], p2z = v2[ 2 * n + i ];
   const double p3x = v3[ i ], p3y = v3[ n + i ], p3z = v3[ 2 * n + i ];
   const double e21x = p2x - p1x, e21y = p2y - p1y, e21z = p2z - p1z;
   const double e31x = p3x - p1x, e31y = p3y - p1y, e31z = p3z - p1z;
   const double e32x = p3x - p2x, e32y = p3y - p2y, e32z = p3z - p2z;
   const double a = sqrt( e21x * e21x + e21y * e21y + e21z * e21z );
   const double b = sqrt( e31x * e31x + e31y * e31y + e31z * e31z );
   const double c = sqrt( e32x * e32x + e32y * e32y + e32z * e32z );
   const double cosC = ( e21x * e31x + e21y * e31y + e21z * e31z ) / a / b;
   const double t = ( a + b - c ) / ( 2. * ( 1. + cosC ) );
   xc[ i ] = p1x + t * ( e21x / a + e31x / b );
   xc[ n + i ] = p1y + t * ( e21y / a + e31y / b );
   xc[ 2 * n + i ] = p1z + t * ( e21z / a + e31z / b );
   const double cxn = e31y * e21z - e31z * e21y;
   const double cyn = e31z * e21x - e31x * e21z;
   const double czn = e31x * e21y - e31y * e21x;
   const double area2 = sqrt( cxn * cxn + cyn * cyn + czn * czn ) / 2.;
   const double nx = .5 * cxn / area2, ny = .5 * cyn / area2, nz = .5 * czn / area2;
   const double cc = acos( cosC );
   const double r = t * sin( cc );
   const double cosB = -( e21x * e32x + e21y * e32y + e21z * e32z ) / a / c;
   const double bb = acos( cosB );
   const double aa = pi - bb - cc;
   const double t2 = 1. / ( 2. * cos( 0.5 * cc ) );
   const double dc = ( a + b - c ) * t2;
   const double db = sqrt( a * a + ( b - c ) * ( b - c ) - 2. * a * ( b - c ) * cosC ) * t2;
   const double da = sqrt( b * b + ( a - c ) * ( a - c ) - 2. * b * ( a - c ) * cosC ) * t2;
   const double ha = r / da, hb = r / db, hc = r / dc;
   const double dpa = ( pi - aa ) * ( da - r ) - 2.0 * ( da * acos( ha ) - r * log( ( 1. + sqrt( 1. - ha * ha ) ) / ha ) );
   const double dpb = ( pi - bb ) * ( db - r ) - 2.0 * ( db * acos( hb ) - r * log( ( 1. + sqrt( 1. - hb * hb ) ) / hb ) );
   const double dpc = ( pi - cc ) * ( dc - r ) - 2.0 * ( dc * acos( hc ) - r * log( ( 1. + sqrt( 1. - hc * hc ) ) / hc ) );
   const double sca = 0.5 * r * ( ( cos( 2.0 * cc ) + cos( 2.0 * bb ) ) * fn_si( ha ) - ( sin( 2.0 * cc ) + sin( 2.0 * bb ) ) * fn_csi( ha ) );
   const double scb = 0.5 * r * ( ( 1. + cos( 2.0 * bb ) ) * fn_si( hb ) + sin( 2.0 * bb ) * fn_csi( hb ) );
   const double scc = 0.5 * r * ( ( 1. + cos( 2.0 * cc ) ) * fn_si( hc ) + sin( 2.0 * cc ) * fn_csi( hc ) );
   const double cca = -sin( bb - cc ) * r * ( cos( aa ) * fn_si( ha ) + sin( aa ) * fn_csi( ha ) );
   const double ccb = sin( bb ) * r * ( cos( bb ) * fn_si( hb ) + sin( bb ) * fn_csi( hb ) );
   const double ccc = -sin( cc ) * r * ( cos( cc ) * fn_si( hc ) + sin( cc ) * fn_csi( hc ) );
   const double cor = sqrt( afac );
   double g[ 3 ][ 3 ];
   g[ 0 ][ 0 ] = cor * ( ( 2.0 + cg ) * pi * r + ( 1.0 + 0.5 * cg ) * cg * ( dpa + dpb + dpc ) + cg * 0.25 * ( sca + scb + scc ) );
   g[ 0 ][ 1 ] = 0.5 * cor * ( cca + ccb + ccc ) * cg;
   g[ 1 ][ 0 ] = g[ 0 ][ 1 ];
   g[ 1 ][ 1 ] = g[ 0 ][ 0 ] - 0.5 * cor * cg * ( sca + scb + scc );
   g[ 2 ][ 2 ] = cor * ( 2.0 * pi * r + ( dpa + dpb + dpc ) );
   g[ 0 ][ 2 ] = g[ 1 ][ 2 ] = g[ 2 ][ 0 ] = g[ 2 ][ 1 ] = 0.0;
   double rr[ 3 ][ 3 ];
   rr[ 0 ][ 0 ] = e21x / a;
   rr[ 0 ][ 1 ] = e21y / a;
   rr[ 0 ][ 2 ] = e21z / a;
   rr[ 2 ][ 0 ] = nx;
   rr[ 2 ][ 1 ] = ny;
   rr[ 2 ][ 2 ] = nz;
   rr[ 1 ][ 0 ] = ( ny * e21z - nz * e21y ) / a;
   rr[ 1 ][ 1 ] = ( nz * e21x - nx * e21z ) / a;
   rr[ 1 ][ 2 ] = ( nx * e21y - ny * e21x ) / a;
   for ( int p = 0; p < 3; ++p )
      for ( int q = 0; q < 3; ++q ) {
         double acc = 0.0;
         for ( int k = 0; k < 3; ++k ) acc = acc + rr[ k ][ p ] * ( g[ k ][ 0 ] * rr[ 0 ][ q ] + g[ k ][ 1 ] * rr[ 1 ][ q ] + g[ k ][ 2 ] * rr[ 2 ][ q ] );
         selfb[ 9 * (size_t)i + 3 * p + q ] = acc;
      }
   // XMU: area times BEST's unit normal (v3 - v1) x (v2 - v1), best/best.f:387-399.
   {
      const double px = e31y * e21z - e31z * e21y;
      const double py = e31z * e21x - e31x * e21z;
      const double pz = e31x * e21y - e31y * e21x;
      const double pl = sqrt( px * px + py * py + pz * pz );
      const double ar = area[ i ];
      xmu[ 3 * i ] = ar * px / pl;
      xmu[ 3 * i + 1 ] = ar * py / pl;
      xmu[ 3 * i + 2 ] = ar * pz / pl;
   }
   // Source record for the pair kernel (best/triangleint.f:31-62 computes these per pair).
   double* s = src + (size_t)kSrcStride * i;
   s[ kP1 ] = p1x, s[ kP1 + 1 ] = p1y, s[ kP1 + 2 ] = p1z;
   s[ kP2 ] = p2x, s[ kP2 + 1 ] = p2y, s[ kP2 + 2 ] = p2z;
   s[ kP3 ] = p3x, s[ kP3 + 1 ] = p3y, s[ kP3 + 2 ] = p3z;
   s[ kArea ] = area[ i ];
   const double v21[ 3 ] = { e21x, e21y, e21z }, v31[ 3 ] = { e31x, e31y, e31z }, v32[ 3 ] = { e32x, e32y, e32z };
   for ( int k = 0; k < 3; ++k ) {
      s[ kV21 + k ] = v21[ k ];
      s[ kV31 + k ] = v31[ k ];
      s[ kV32 + k ] = v32[ k ];
   }
   const double d2121 = dot3( v21, v21 );
   s[ kD2121 ] = d2121;
   s[ kD2131 ] = dot3( v21, v31 );
   s[ kD2132 ] = dot3( v21, v32 );
   s[ kD3131 ] = dot3( v31, v31 );
   s[ kD3232 ] = dot3( v32, v32 );
   s[ kD3231 ] = dot3( v32, v31 );
   s[ kS21 ] = sqrt( d2121 );
}

// DIAGM = mean diagonal entry of GR (best/best.f:386-403), and the deflation weights: ndef = 1,
// GDEF = DIAGM / |XMU|^2 (:404); ndef > 1, GDEF(K) = DIAGM / |XMU_K|^2 (DEFGAM,
// best/compon.f90:185-197).  One workgroup of 1024 threads with fixed-order tree sums.
__global__ void __launch_bounds__( 1024 ) k_deflation_weights( int n, int ndef, const double* __restrict__ selfb,
                                                              const double* __restrict__ xmu, const int* __restrict__ comp,
                                                              double* __restrict__ gdef, double* __restrict__ diagm_out,
                                                              int exact_order ) {
   __shared__ double red[ 1024 ];
   if ( exact_order ) {
      // The reference's running sums, by one thread: DIAGM and |XMU|^2 bit for bit (the value
      // bhm_diag_mean returns).  The data is staged through LDS so the loads are not serialized.
      __shared__ double sdiag, sx2;
      if ( threadIdx.x == 0 ) sdiag = 0.0, sx2 = 0.0;
      for ( int c0 = 0; c0 < n; c0 += 256 ) {
         const int m = min( 256, n - c0 );
         __syncthreads();
         if ( (int)threadIdx.x < m ) {
            const int i = c0 + threadIdx.x;
            const double* g = selfb + 9 * (size_t)i;
            red[ 4 * threadIdx.x + 0 ] = g[ 0 ];
            red[ 4 * threadIdx.x + 1 ] = g[ 4 ];
            red[ 4 * threadIdx.x + 2 ] = g[ 8 ];
         }
         __syncthreads();
         if ( threadIdx.x == 0 ) {
            double d = sdiag;
            for ( int k = 0; k < m; ++k ) d = d + red[ 4 * k ] + red[ 4 * k + 1 ] + red[ 4 * k + 2 ];
            sdiag = d;
         }
         __syncthreads();
         if ( (int)threadIdx.x < m ) {
            const int i = c0 + threadIdx.x;
            red[ 4 * threadIdx.x + 0 ] = xmu[ 3 * i ];
            red[ 4 * threadIdx.x + 1 ] = xmu[ 3 * i + 1 ];
            red[ 4 * threadIdx.x + 2 ] = xmu[ 3 * i + 2 ];
         }
         __syncthreads();
         if ( threadIdx.x == 0 ) {
            double x = sx2;
            for ( int k = 0; k < m; ++k ) {
               const double a = red[ 4 * k ], b = red[ 4 * k + 1 ], c = red[ 4 * k + 2 ];
               x = x + a * a + b * b + c * c;
            }
            sx2 = x;
         }
      }
      __syncthreads();
      if ( threadIdx.x != 0 ) return;
      const double diagm = sdiag / (double)( 3 * n );
      *diagm_out = diagm;
      if ( ndef <= 1 ) {
         gdef[ 0 ] = diagm / sx2;
         return;
      }
      for ( int k = 0; k < ndef; ++k ) gdef[ k ] = 0.0;
      for ( int i = 0; i < n; ++i ) {
         const double a = xmu[ 3 * i ], b = xmu[ 3 * i + 1 ], c = xmu[ 3 * i + 2 ];
         double& g = gdef[ comp[ i ] - 1 ];
         g = g + a * a + b * b + c * c;
      }
      for ( int k = 0; k < ndef; ++k ) gdef[ k ] = diagm / gdef[ k ];
      return;
   }
   double dsum = 0.0;
   for ( int i = threadIdx.x; i < n; i += blockDim.x ) {
      const double* g = selfb + 9 * (size_t)i;
      dsum = dsum + g[ 0 ] + g[ 4 ] + g[ 8 ];
   }
   const double diagm = block_sum( dsum, red ) / (double)( 3 * n );
   if ( threadIdx.x == 0 ) *diagm_out = diagm;
   const int nk = ndef < 1 ? 1 : ndef;
   for ( int k = 1; k <= nk; ++k ) {
      double x2 = 0.0;
      for ( int i = threadIdx.x; i < n; i += blockDim.x ) {
         if ( ndef > 1 && comp[ i ] != k ) continue;
         const double a = xmu[ 3 * i ], b = xmu[ 3 * i + 1 ], c = xmu[ 3 * i + 2 ];
         x2 = x2 + a * a + b * b + c * c;
      }
      x2 = block_sum( x2, red );
      if ( threadIdx.x == 0 ) gdef[ k - 1 ] = diagm / x2;
   }
}

// The off-diagonal blocks (and the self blocks) of GR, with the deflation fused: every lane owns
// one collocation triangle n1 (rows 3 n1 .. 3 n1 + 2) and walks a strip of source triangles n2
// (columns 3 n2 .. 3 n2 + 2), so a wave's stores cover 192 contiguous doubles of each column.
// The source record is uniform across the wave and loads as scalars.
// deflate: 0 none; 1 add gdef[comp(n2)] xmu(j) xmu(i) where comp(n1) == comp(n2).
// F32: the entries are stored a second time, rounded to FP32, into gr32 (same layout), the copy
// the FP32 factorization consumes.
template <bool F32>
__global__ void __launch_bounds__( 256, 4 ) k_pairs( int n, int strip, double cg, const double* __restrict__ src,
                                                 const double* __restrict__ xc, const double* __restrict__ selfb,
                                                 const double* __restrict__ xmu, const int* __restrict__ comp,
                                                 const double* __restrict__ gdef, int deflate, double* __restrict__ gr,
                                                 float* __restrict__ gr32, long ld ) {
   const int n1 = blockIdx.x * blockDim.x + threadIdx.x;
   const int n2lo = blockIdx.y * strip;
   const int n2hi = min( n, n2lo + strip );
   if ( n1 >= n ) return;
   const double c[ 3 ] = { xc[ n1 ], xc[ n + n1 ], xc[ 2 * n + n1 ] };
   double xm[ 3 ] = { 0.0, 0.0, 0.0 };
   int comp1 = 0;
   if ( deflate ) {
      xm[ 0 ] = xmu[ 3 * n1 ], xm[ 1 ] = xmu[ 3 * n1 + 1 ], xm[ 2 ] = xmu[ 3 * n1 + 2 ];
      comp1 = comp[ n1 ];
   }
   double* out = gr + 3 * (size_t)n1;
   for ( int n2 = n2lo; n2 < n2hi; ++n2 ) {
      double gm[ 9 ];
      if ( n2 != n1 ) {
         double g6[ 6 ];
         pair_block( src + (size_t)kSrcStride * n2, c, cg, g6 );
#pragma unroll
         for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b ) gm[ 3 * a + b ] = sym_at( g6, a, b );
      } else {
#pragma unroll
         for ( int k = 0; k < 9; ++k ) gm[ k ] = selfb[ 9 * (size_t)n1 + k ];
      }
      if ( deflate && comp[ n2 ] == comp1 ) {
         const double gk = gdef[ comp1 - 1 ];
#pragma unroll
         for ( int b = 0; b < 3; ++b ) {
            const double gdj = gk * xmu[ 3 * n2 + b ];
#pragma unroll
            for ( int a = 0; a < 3; ++a ) gm[ 3 * a + b ] = gm[ 3 * a + b ] + gdj * xm[ a ];
         }
      }
      double* col = out + (size_t)( 3 * n2 ) * ld;
#pragma unroll
      for ( int b = 0; b < 3; ++b ) {
         col[ (size_t)b * ld + 0 ] = gm[ b ];
         col[ (size_t)b * ld + 1 ] = gm[ 3 + b ];
         col[ (size_t)b * ld + 2 ] = gm[ 6 + b ];
      }
      if constexpr ( F32 ) {
         float* col32 = gr32 + 3 * (size_t)n1 + (size_t)( 3 * n2 ) * ld;
#pragma unroll
         for ( int b = 0; b < 3; ++b ) {
            col32[ (size_t)b * ld + 0 ] = (float)gm[ b ];
            col32[ (size_t)b * ld + 1 ] = (float)gm[ 3 + b ];
            col32[ (size_t)b * ld + 2 ] = (float)gm[ 6 + b ];
         }
      }
   }
}

// Rows and columns n3 .. npad-1 of a padded matrix: identity, so the padded LU is
// block-diagonal with the true factors in the leading block.
template <typename T>
__global__ void k_pad_identity( int n3, int npad, T* __restrict__ a, long ld ) {
   const long total = (long)npad * ( npad - n3 ) * 2;
   for ( long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x ) {
      long i, j;
      if ( t < (long)npad * ( npad - n3 ) ) {  // columns >= n3, all rows
         j = n3 + t / npad;
         i = t % npad;
      } else {  // rows >= n3 in columns < n3
         const long u = t - (long)npad * ( npad - n3 );
         j = u / ( npad - n3 );
         i = n3 + u % ( npad - n3 );
         if ( j >= n3 ) continue;
      }
      a[ j * ld + i ] = ( i == j ) ? T( 1 ) : T( 0 );
   }
}

// GR(i0 : i0+m-1, j0 : j0+nc-1) before deflation into a (column-major, lda).  One thread per
// covered triangle pair.
__global__ void __launch_bounds__( 256 ) k_entries( int n, double cg, const double* __restrict__ src,
                                                   const double* __restrict__ xc, const double* __restrict__ selfb,
                                                   int i0, int m, int j0, int nc, double* __restrict__ a, long lda ) {
   const int t1lo = i0 / 3, t1n = ( i0 + m - 1 ) / 3 - t1lo + 1;
   const int t2lo = j0 / 3, t2n = ( j0 + nc - 1 ) / 3 - t2lo + 1;
   const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( idx >= (long)t1n * t2n ) return;
   const int n1 = t1lo + (int)( idx % t1n ), n2 = t2lo + (int)( idx / t1n );
   double gm[ 9 ];
   if ( n1 != n2 ) {
      const double c[ 3 ] = { xc[ n1 ], xc[ n + n1 ], xc[ 2 * n + n1 ] };
      double g6[ 6 ];
      pair_block( src + (size_t)kSrcStride * n2, c, cg, g6 );
      for ( int p = 0; p < 3; ++p )
         for ( int q = 0; q < 3; ++q ) gm[ 3 * p + q ] = sym_at( g6, p, q );
   } else {
      for ( int k = 0; k < 9; ++k ) gm[ k ] = selfb[ 9 * (size_t)n1 + k ];
   }
   for ( int p = 0; p < 3; ++p )
      for ( int q = 0; q < 3; ++q ) {
         const int i = 3 * n1 + p - i0, j = 3 * n2 + q - j0;
         if ( i >= 0 && i < m && j >= 0 && j < nc ) a[ (size_t)j * lda + i ] = gm[ 3 * p + q ];
      }
}

// ------------------------------------------------------------------ host launchers

void launch_prep( hipStream_t st, int n, double* v1, double* v2, double* v3, int iarea, double truearea, double* area,
                  double* scal ) {
   hipLaunchKernelGGL( k_prep, dim3( 1 ), dim3( 1024 ), 0, st, n, v1, v2, v3, iarea, truearea, area, scal );
}

void launch_self( hipStream_t st, int n, const double* afac, double cg, const double* v1, const double* v2,
                  const double* v3, const double* area, double* xc, double* selfb, double* xmu, double* src ) {
   hipLaunchKernelGGL( k_self, dim3( ( n + 255 ) / 256 ), dim3( 256 ), 0, st, n, afac, cg, v1, v2, v3, area, xc, selfb,
                       xmu, src );
}

void launch_deflation_weights( hipStream_t st, int n, int ndef, const double* selfb, const double* xmu, const int* comp,
                               double* gdef, double* diagm, int exact_order ) {
   hipLaunchKernelGGL( k_deflation_weights, dim3( 1 ), dim3( 1024 ), 0, st, n, ndef, selfb, xmu, comp, gdef, diagm,
                       exact_order );
}

void launch_pairs( hipStream_t st, int n, int strip, double cg, const double* src, const double* xc, const double* selfb,
                   const double* xmu, const int* comp, const double* gdef, int deflate, double* gr, long ld, float* gr32 ) {
   dim3 grid( ( n + 255 ) / 256, ( n + strip - 1 ) / strip );
   // 3,457 flops per pair (3,120 add/mul + 106 sqrt + 210 div + 21 log, special functions as 1).
   ProfScope prof( kKPairs, 3457.0 * n * ( n - 1.0 ) );
   if ( gr32 )
      hipLaunchKernelGGL( k_pairs<true>, grid, dim3( 256 ), 0, st, n, strip, cg, src, xc, selfb, xmu, comp, gdef, deflate, gr,
                          gr32, ld );
   else
      hipLaunchKernelGGL( k_pairs<false>, grid, dim3( 256 ), 0, st, n, strip, cg, src, xc, selfb, xmu, comp, gdef, deflate, gr,
                          nullptr, ld );
}

void launch_pad_identity( hipStream_t st, int n3, int npad, double* a, long ld ) {
   if ( npad > n3 ) hipLaunchKernelGGL( k_pad_identity<double>, dim3( 1024 ), dim3( 256 ), 0, st, n3, npad, a, ld );
}

void launch_pad_identity32( hipStream_t st, int n3, int npad, float* a, long ld ) {
   if ( npad > n3 ) hipLaunchKernelGGL( k_pad_identity<float>, dim3( 1024 ), dim3( 256 ), 0, st, n3, npad, a, ld );
}

void launch_entries( hipStream_t st, int n, double cg, const double* src, const double* xc, const double* selfb, int i0,
                     int m, int j0, int nc, double* a, long lda ) {
   const long t1n = ( i0 + m - 1 ) / 3 - i0 / 3 + 1, t2n = ( j0 + nc - 1 ) / 3 - j0 / 3 + 1;
   const long total = t1n * t2n;
   hipLaunchKernelGGL( k_entries, dim3( (unsigned)( ( total + 255 ) / 256 ) ), dim3( 256 ), 0, st, n, cg, src, xc, selfb, i0,
                       m, j0, nc, a, lda );
}

}  // namespace bhg
