// Dense LU with partial pivoting, triangular solves and FP64 iterative refinement on MI355X.
//
// Replaces LAPACK DGETRF / DGETRS as BEST calls them (best/best.f:444, best/besolve.f90:133).
//
// The factorization is blocked and right-looking with kNB = 128 columns per step:
//   k_lu_diag      the 128 x 128 diagonal block, factored in LDS by one workgroup, then both of
//                  its triangles inverted in place (LAPACK dtrti2 order);
//   k_gemm         128 x 128 MFMA tiles (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32,
//                  FP32 accumulation): the panel below (A U^-1), the row panel right (L^-1 A) and
//                  the trailing update (A -= L U).
// GR from BEST's meshes is column diagonally dominant in the elimination sense: partial
// pivoting never exchanges rows (every multiplier |l_ij| <= 1).  The fast path therefore
// factors without exchanges and checks exactly that condition (|a_ij| <= |a_jj| at the moment
// column j is eliminated, inside the diagonal block and for every row below it): when it holds,
// the factors are those of partial pivoting, with IPIV = identity.  When any multiplier exceeds
// 1 the caller reruns the factorization on the slow path (k_piv_* kernels), which is true
// column-by-column partial pivoting (first row of largest |a|, as IDAMAX chooses).
//
// Matrices are column-major, padded to a multiple of 128 with an identity block.
#include <algorithm>

#include "bh_common.h"

namespace bhg {

namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
   typedef double v4 __attribute__( ( ext_vector_type( 4 ) ) );
   static __device__ inline v4 run( double a, double b, v4 c ) { return __builtin_amdgcn_mfma_f64_16x16x4f64( a, b, c, 0, 0, 0 ); }
};
template <>
struct Mfma<float> {
   typedef float v4 __attribute__( ( ext_vector_type( 4 ) ) );
   static __device__ inline v4 run( float a, float b, v4 c ) { return __builtin_amdgcn_mfma_f32_16x16x4f32( a, b, c, 0, 0, 0 ); }
};

template <typename T>
__device__ inline T tabs( T v ) {
   return v < T( 0 ) ? -v : v;
}

}  // namespace

// C (M x N) = alpha A (M x K) B (K x N) + beta C, column-major; M, N multiples of 128, K of BK.
// 256 threads: four waves in a 2 x 2 arrangement, each owning a 64 x 64 block of the 128 x 128
// tile as 4 x 4 MFMA 16 x 16 accumulators.  A and B slices of BK columns pass through LDS in two
// stages: while one stage feeds the MFMAs, the next slice is already in flight from global memory
// (one barrier per slice).  In LDS each wave's 64 rows (columns for B) are stored permuted, row
// 16 t + i at slot 4 i + t, so the four fragments a lane needs for one k come as one 16-byte
// (FP32) or two 16-byte (FP64) reads.  With beta != 0 the C tile is loaded into registers before
// the main loop.  C may alias A (when N == 128: each workgroup reads its whole row strip before
// it writes) or B (when M == 128).  check != nullptr: set *check when any |C| > 1 (the pivot test
// of the panel below the diagonal block).
// Batched over blockIdx.z: batch z = zdiv * q + t offsets A by q sa.hi + t sa.lo, and likewise B, C.
struct BatchStride {
   long hi = 0, lo = 0;
};

template <typename T, int BK>
__global__ void __launch_bounds__( 256 ) k_gemm( int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb,
                                                T beta, T* C, long ldc, int* check, int zdiv, BatchStride sa,
                                                BatchStride sb, BatchStride sc ) {
   constexpr int BM = 128, BN = 128, PAD = 8;
   constexpr int LDW = BM + PAD;  // LDS row length (elements)
   __shared__ __attribute__( ( aligned( 16 ) ) ) T As[ 2 ][ BK ][ LDW ];
   __shared__ __attribute__( ( aligned( 16 ) ) ) T Bs[ 2 ][ BK ][ LDW ];
   using v4 = typename Mfma<T>::v4;
   {
      const long q = blockIdx.z / zdiv, t = blockIdx.z % zdiv;
      A += q * sa.hi + t * sa.lo;
      B += q * sb.hi + t * sb.lo;
      C += q * sc.hi + t * sc.lo;
   }
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave & 1, wn = wave >> 1;
   const int bm = blockIdx.x, bn = blockIdx.y;
   const T* Ab = A + (long)bm * BM;
   const T* Bb = B + (long)bn * BN * ldb;
   constexpr int AE = BM * BK / 256;  // A elements per thread, contiguous along m
   constexpr int ATPK = BM / AE;      // threads per k column of A
   constexpr int BE = BK * BN / 256;  // B elements per thread, contiguous along k
   constexpr int BTPN = BK / BE;      // threads per column of B
   const int a_k = tid / ATPK, a_m = ( tid % ATPK ) * AE;
   const int b_n = tid / BTPN, b_k = ( tid % BTPN ) * BE;
   // permuted slot of row (column) m of the tile: wave block m / 64, t = (m % 64) / 16, i = m % 16
   auto slot = []( int m ) { return ( m & ~63 ) + ( ( m & 15 ) << 2 ) + ( ( m >> 4 ) & 3 ); };
   const int b_slot = slot( b_n );
   T ra[ AE ], rb[ BE ];
   v4 acc[ 4 ][ 4 ];
#pragma unroll
   for ( int t = 0; t < 4; ++t )
#pragma unroll
      for ( int u = 0; u < 4; ++u ) acc[ t ][ u ] = v4{ 0, 0, 0, 0 };
   // C tile of this lane, loaded ahead (element r of accumulator (t, u)).
   const int li = lane & 15, lk = lane >> 4;
   auto crow = [&]( int t, int r ) { return wm * 64 + t * 16 + ( sizeof( T ) == 8 ? lk + 4 * r : 4 * lk + r ); };
   // FP64 keeps its 64 accumulators per lane in AGPRs only without the prefetched tile: there it
   // reads C in the epilogue.
   constexpr bool kPre = sizeof( T ) == 4;
   T cpre[ 4 ][ 4 ][ 4 ];
   if ( kPre && beta != T( 0 ) ) {
#pragma unroll
      for ( int t = 0; t < 4; ++t )
#pragma unroll
         for ( int u = 0; u < 4; ++u ) {
            const T* c = C + (long)( bn * BN + wn * 64 + u * 16 + li ) * ldc + bm * BM;
#pragma unroll
            for ( int r = 0; r < 4; ++r ) cpre[ t ][ u ][ r ] = c[ crow( t, r ) ];
         }
   }
   const int nk = K / BK;
   auto gload = [&]( int kt ) {
      const T* pa = Ab + (long)( kt * BK + a_k ) * lda + a_m;
#pragma unroll
      for ( int e = 0; e < AE; ++e ) ra[ e ] = pa[ e ];
      const T* pb = Bb + (long)b_n * ldb + kt * BK + b_k;
#pragma unroll
      for ( int e = 0; e < BE; ++e ) rb[ e ] = pb[ e ];
   };
   auto sstore = [&]( int stg ) {
#pragma unroll
      for ( int e = 0; e < AE; ++e ) As[ stg ][ a_k ][ slot( a_m + e ) ] = ra[ e ];
#pragma unroll
      for ( int e = 0; e < BE; ++e ) Bs[ stg ][ b_k + e ][ b_slot ] = rb[ e ];
   };
   gload( 0 );
   sstore( 0 );
   __syncthreads();
   for ( int kt = 0; kt < nk; ++kt ) {
      const int stg = kt & 1;
      if ( kt + 1 < nk ) gload( kt + 1 );
#pragma unroll
      for ( int kk = 0; kk < BK; kk += 4 ) {
         const int kr = kk + lk;
         const T* pa = &As[ stg ][ kr ][ wm * 64 + li * 4 ];
         const T* pb = &Bs[ stg ][ kr ][ wn * 64 + li * 4 ];
         T af[ 4 ], bf[ 4 ];
#pragma unroll
         for ( int t = 0; t < 4; ++t ) af[ t ] = pa[ t ], bf[ t ] = pb[ t ];
#pragma unroll
         for ( int t = 0; t < 4; ++t )
#pragma unroll
            for ( int u = 0; u < 4; ++u ) acc[ t ][ u ] = Mfma<T>::run( af[ t ], bf[ u ], acc[ t ][ u ] );
      }
      if ( kt + 1 < nk ) sstore( stg ^ 1 );
      __syncthreads();
   }
   bool big = false;
#pragma unroll
   for ( int t = 0; t < 4; ++t )
#pragma unroll
      for ( int u = 0; u < 4; ++u ) {
         T* c = C + (long)( bn * BN + wn * 64 + u * 16 + li ) * ldc + bm * BM;
#pragma unroll
         for ( int r = 0; r < 4; ++r ) {
            T v = alpha * acc[ t ][ u ][ r ];
            if ( beta != T( 0 ) ) v = v + beta * ( kPre ? cpre[ t ][ u ][ r ] : c[ crow( t, r ) ] );
            c[ crow( t, r ) ] = v;
            big |= tabs( v ) > T( 1 );
         }
      }
   if ( check && big ) atomicOr( check, 1 );
}

// The diagonal block at A (leading dimension lda): mode 0 factors it without exchanges (flagging
// any |a_ij| > |a_jj| at the step that eliminates column j) and then inverts both triangles;
// mode 1 only inverts an already-factored block.  inv receives L^-1 (unit lower) and U^-1
// (upper) as two dense 128 x 128 column-major matrices.
//
// One workgroup of 256 threads; thread (ty, tx) of a 16 x 16 grid keeps the 8 x 8 tile at rows
// 8 ty.., columns 8 tx.. in registers.  The step loop runs over tile columns jb and, unrolled,
// over the offset jo inside the tile, so every register index is a compile-time constant.  Each
// step broadcasts pivot row and column through LDS (double-buffered: one barrier per step) and
// applies the rank-1 update to the whole tile with multipliers and pivot-row entries set to zero
// outside the active part, so the update itself has no per-element conditions.  The multipliers
// are a_ij * (1 / a_jj), as DGETF2 scales.  The inversions follow in the same way: U^-1 by rows
// from the bottom (X(j,:) = Y(j,:) / u_jj, then Y(i,:) -= u_ij X(j,:) for i < j, where a position
// still holding u_ic becomes X's entry when column c is reached) and L^-1 by rows from the top
// (unit diagonal).  The two touch disjoint halves of the block.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_lu_diag( int mode, T* A, long lda, T* __restrict__ inv, int* check ) {
   constexpr int NB = kNB, RB = 8;
   __shared__ T colb[ 2 ][ NB ], rowb[ 2 ][ NB ], colL[ 2 ][ NB ], rowL[ 2 ][ NB ];
   const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
   const int r0 = ty * RB, c0 = tx * RB;
   T a[ RB ][ RB ];
#pragma unroll
   for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
      for ( int ii = 0; ii < RB; ++ii ) a[ ii ][ jj ] = A[ (long)( c0 + jj ) * lda + r0 + ii ];
   if ( mode == 0 ) {
      bool big = false;
      for ( int jb = 0; jb < NB / RB; ++jb ) {
#pragma unroll
         for ( int jo = 0; jo < RB; ++jo ) {
            const int j = jb * RB + jo, bsel = jo & 1;
            if ( tx == jb ) {
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii ) colb[ bsel ][ r0 + ii ] = a[ ii ][ jo ];
            }
            if ( ty == jb ) {
#pragma unroll
               for ( int jj = 0; jj < RB; ++jj ) rowb[ bsel ][ c0 + jj ] = a[ jo ][ jj ];
            }
            __syncthreads();
            if ( r0 + RB - 1 > j ) {  // this tile has rows below the pivot
               const T piv = rowb[ bsel ][ j ];
               const T rp = T( 1 ) / piv;
               T l[ RB ], u[ RB ];
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii ) {
                  const T v = colb[ bsel ][ r0 + ii ];
                  const bool below = r0 + ii > j;
                  big |= below && !( tabs( v ) <= tabs( piv ) );
                  l[ ii ] = below ? v * rp : T( 0 );
               }
#pragma unroll
               for ( int jj = 0; jj < RB; ++jj ) u[ jj ] = c0 + jj > j ? rowb[ bsel ][ c0 + jj ] : T( 0 );
#pragma unroll
               for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
                  for ( int ii = 0; ii < RB; ++ii ) a[ ii ][ jj ] = a[ ii ][ jj ] - l[ ii ] * u[ jj ];
               if ( tx == jb ) {
#pragma unroll
                  for ( int ii = 0; ii < RB; ++ii )
                     if ( r0 + ii > j ) a[ ii ][ jo ] = l[ ii ];
               }
            }
         }
      }
      if ( big ) atomicOr( check, 1 );
#pragma unroll
      for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
         for ( int ii = 0; ii < RB; ++ii ) A[ (long)( c0 + jj ) * lda + r0 + ii ] = a[ ii ][ jj ];
   }
   __syncthreads();
   for ( int sb = 0; sb < NB / RB; ++sb ) {
#pragma unroll
      for ( int so = 0; so < RB; ++so ) {
         const int bsel = so & 1;
         // U^-1 works on row ju = 127 - (8 sb + so): tile row/column offset 7 - so, tile index 15 - sb.
         const int ju = NB - 1 - ( sb * RB + so ), uo = RB - 1 - so, ub = NB / RB - 1 - sb;
         // L^-1 works on row jl = 8 sb + so: offset so, tile index sb.
         const int jl = sb * RB + so;
         if ( tx == ub ) {
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) colb[ bsel ][ r0 + ii ] = a[ ii ][ uo ];
         }
         if ( ty == ub ) {
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) rowb[ bsel ][ c0 + jj ] = a[ uo ][ jj ];
         }
         if ( tx == sb ) {
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) colL[ bsel ][ r0 + ii ] = a[ ii ][ so ];
         }
         if ( ty == sb ) {
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) rowL[ bsel ][ c0 + jj ] = a[ so ][ jj ];
         }
         __syncthreads();
         const T d = T( 1 ) / rowb[ bsel ][ ju ];
         // U^-1: rows i < ju get (i, ju) <- -u_{i,ju} d and (i, c > ju) -= u_{i,ju} X(ju, c).
         if ( r0 < ju ) {
            T uc[ RB ], xr[ RB ];
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) uc[ ii ] = r0 + ii < ju ? colb[ bsel ][ r0 + ii ] : T( 0 );
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) xr[ jj ] = c0 + jj > ju ? rowb[ bsel ][ c0 + jj ] * d : T( 0 );
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii ) a[ ii ][ jj ] = a[ ii ][ jj ] - uc[ ii ] * xr[ jj ];
            if ( tx == ub ) {
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii )
                  if ( r0 + ii < ju ) a[ ii ][ uo ] = -uc[ ii ] * d;
            }
         }
         // row ju itself: X(ju, ju) = d, X(ju, c > ju) = Y(ju, c) d
         if ( ty == ub ) {
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) {
               const int c = c0 + jj;
               a[ uo ][ jj ] = c == ju ? d : ( c > ju ? a[ uo ][ jj ] * d : a[ uo ][ jj ] );
            }
         }
         // L^-1: rows i > jl get (i, jl) <- -l_{i,jl} and (i, c < jl) -= l_{i,jl} X(jl, c).
         if ( r0 + RB - 1 > jl ) {
            T lc[ RB ], xl[ RB ];
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) lc[ ii ] = r0 + ii > jl ? colL[ bsel ][ r0 + ii ] : T( 0 );
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) xl[ jj ] = c0 + jj < jl ? rowL[ bsel ][ c0 + jj ] : T( 0 );
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii ) a[ ii ][ jj ] = a[ ii ][ jj ] - lc[ ii ] * xl[ jj ];
            if ( tx == sb ) {
#pragma unroll
               for ( int ii = 0; ii < RB; ++ii )
                  if ( r0 + ii > jl ) a[ ii ][ so ] = -lc[ ii ];
            }
         }
      }
   }
   T* Li = inv;
   T* Ui = inv + NB * NB;
#pragma unroll
   for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
      for ( int ii = 0; ii < RB; ++ii ) {
         const int r = r0 + ii, c = c0 + jj;
         Li[ c * NB + r ] = r > c ? a[ ii ][ jj ] : ( r == c ? T( 1 ) : T( 0 ) );
         Ui[ c * NB + r ] = r <= c ? a[ ii ][ jj ] : T( 0 );
      }
}

// ------------------------------------------------------------------ partial pivoting (slow path)

// Stage 1 of the pivot search in column j over rows j .. npad-1: per-workgroup (|max|, first row).
template <typename T>
__global__ void __launch_bounds__( 256 ) k_piv_search( int npad, int j, const T* __restrict__ A, long lda, double* pval,
                                                      int* pidx ) {
   __shared__ double sv[ 256 ];
   __shared__ int si[ 256 ];
   const T* col = A + (long)j * lda;
   double best = -1.0;
   int bi = npad;
   for ( int i = j + blockIdx.x * 256 + threadIdx.x; i < npad; i += gridDim.x * 256 ) {
      const double v = (double)tabs( col[ i ] );
      if ( v > best ) best = v, bi = i;
   }
   sv[ threadIdx.x ] = best;
   si[ threadIdx.x ] = bi;
   __syncthreads();
   for ( int w = 128; w > 0; w >>= 1 ) {
      if ( threadIdx.x < w ) {
         const double v2 = sv[ threadIdx.x + w ];
         const int i2 = si[ threadIdx.x + w ];
         if ( v2 > sv[ threadIdx.x ] || ( v2 == sv[ threadIdx.x ] && i2 < si[ threadIdx.x ] ) ) sv[ threadIdx.x ] = v2, si[ threadIdx.x ] = i2;
      }
      __syncthreads();
   }
   if ( threadIdx.x == 0 ) pval[ blockIdx.x ] = sv[ 0 ], pidx[ blockIdx.x ] = si[ 0 ];
}

// Stage 2 + the interchange: picks the pivot row p (IDAMAX: first of the largest), records
// ipiv[j] = p + 1 and swaps rows j and p across all columns.  One workgroup per column slab.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_piv_swap( int npad, int j, int nparts, const double* __restrict__ pval,
                                                    const int* __restrict__ pidx, T* A, long lda, int* ipiv, int* info ) {
   double best = -1.0;
   int p = npad;
   for ( int k = 0; k < nparts; ++k )
      if ( pval[ k ] > best || ( pval[ k ] == best && pidx[ k ] < p ) ) best = pval[ k ], p = pidx[ k ];
   if ( blockIdx.x == 0 && threadIdx.x == 0 ) {
      ipiv[ j ] = p + 1;
      if ( best == 0.0 && *info == 0 ) *info = j + 1;
   }
   if ( p == j ) return;
   for ( int c = blockIdx.x * 256 + threadIdx.x; c < npad; c += gridDim.x * 256 ) {
      T* col = A + (long)c * lda;
      const T t = col[ j ];
      col[ j ] = col[ p ];
      col[ p ] = t;
   }
}

// Column j of the panel [kk, kk + 128): scale below the diagonal and update the panel columns to
// its right (rows j+1 .. npad-1).
template <typename T>
__global__ void __launch_bounds__( 256 ) k_piv_update( int npad, int j, int kend, T* A, long lda ) {
   const int i = j + 1 + blockIdx.x * 256 + threadIdx.x;
   if ( i >= npad ) return;
   T* cj = A + (long)j * lda;
   const T piv = cj[ j ];
   const T l = piv != T( 0 ) ? cj[ i ] / piv : cj[ i ];
   cj[ i ] = l;
   for ( int c = j + 1; c < kend; ++c ) {
      T* cc = A + (long)c * lda;
      cc[ i ] = cc[ i ] - l * cc[ j ];
   }
}

// ------------------------------------------------------------------ solves

// Partial products of the triangular solves on the matrix cores: P[kc] = A (M x K) X for the
// kc-th slice of kper columns of K, with A column-major and X, P row-interleaved (element (i, r)
// at 16 i + r, 16 right-hand-side columns).  Workgroup (row block of 64, slice kc); its four waves
// split the slice, each running 16 x 16 x 4 MFMAs over 4 row tiles (the X fragment of a 4-column
// step is one coalesced load, 64 consecutive values), and add their tiles in LDS.  Splitting K
// keeps enough waves resident to cover the memory latency; k_skinny_reduce adds the slices in a
// fixed order.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_skinny_part( int M, int K, int kper, const T* __restrict__ A, long lda,
                                                       const T* __restrict__ X, T* __restrict__ P ) {
   using v4 = typename Mfma<T>::v4;
   constexpr int PF = 4;
   __shared__ T red[ 3 ][ 64 * 16 ];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int row0 = blockIdx.x * 64;
   const int kc = blockIdx.y;
   const int kw = kper / 4;
   const int kb = kc * kper + w * kw;
   const int kn = max( 0, min( kw, K - kb ) );
   v4 acc[ 4 ];
#pragma unroll
   for ( int t = 0; t < 4; ++t ) acc[ t ] = v4{ 0, 0, 0, 0 };
   const int li = lane & 15, lk = lane >> 4;
   const T* a = A + (long)( kb + lk ) * lda + row0 + li;
   const T* x = X + (long)( kb + lk ) * 16 + li;
   for ( int k = 0; k < kn; k += 4 * PF ) {
      T fa[ PF ][ 4 ], fb[ PF ];
#pragma unroll
      for ( int u = 0; u < PF; ++u ) {
         const bool in = k + 4 * u < kn;
         const long off = (long)( k + 4 * u ) * lda;
#pragma unroll
         for ( int t = 0; t < 4; ++t ) fa[ u ][ t ] = in ? a[ off + 16 * t ] : T( 0 );
         fb[ u ] = in ? x[ (long)( k + 4 * u ) * 16 ] : T( 0 );
      }
#pragma unroll
      for ( int u = 0; u < PF; ++u )
#pragma unroll
         for ( int t = 0; t < 4; ++t ) acc[ t ] = Mfma<T>::run( fa[ u ][ t ], fb[ u ], acc[ t ] );
   }
   // tile t, accumulator element r of this lane: row 16 t + (FP32: 4 (lane/16) + r, FP64:
   // lane/16 + 4 r), right-hand side lane % 16
   auto rowof = [&]( int t, int r ) { return 16 * t + ( sizeof( T ) == 8 ? lk + 4 * r : 4 * lk + r ); };
   if ( w > 0 ) {
#pragma unroll
      for ( int t = 0; t < 4; ++t )
#pragma unroll
         for ( int r = 0; r < 4; ++r ) red[ w - 1 ][ rowof( t, r ) * 16 + li ] = acc[ t ][ r ];
   }
   __syncthreads();
   if ( w == 0 ) {
      T* out = P + (long)kc * M * 16 + (long)row0 * 16;
#pragma unroll
      for ( int t = 0; t < 4; ++t )
#pragma unroll
         for ( int r = 0; r < 4; ++r ) {
            const int e = rowof( t, r ) * 16 + li;
            out[ e ] = ( acc[ t ][ r ] + red[ 0 ][ e ] ) + ( red[ 1 ][ e ] + red[ 2 ][ e ] );
         }
   }
}

// Y = beta Y + alpha sum_kc P[kc] (row-interleaved, M x 16), the slices added in order.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_skinny_reduce( long n16, int nkc, T alpha, const T* __restrict__ P, T beta,
                                                         T* __restrict__ Y ) {
   for ( long e = blockIdx.x * 256L + threadIdx.x; e < n16; e += (long)gridDim.x * 256 ) {
      T s = P[ e ];
      for ( int c = 1; c < nkc; ++c ) s = s + P[ (long)c * n16 + e ];
      Y[ e ] = beta == T( 0 ) ? alpha * s : beta * Y[ e ] + alpha * s;
   }
}

// Column-major (ld) <-> row-interleaved (16 per row) blocks of npad rows; columns >= nrhs are zero.
template <typename S, typename D>
__global__ void k_to_interleaved( int npad, int nrhs, const S* __restrict__ b, long ldb, D* __restrict__ v ) {
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)npad * 16; e += (long)gridDim.x * blockDim.x ) {
      const long i = e >> 4;
      const int r = (int)( e & 15 );
      v[ e ] = r < nrhs ? (D)b[ (long)r * ldb + i ] : D( 0 );
   }
}

template <typename S, typename D>
__global__ void k_from_interleaved( int n, int nrhs, const S* __restrict__ v, D* __restrict__ b, long ldb ) {
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * nrhs; e += (long)gridDim.x * blockDim.x ) {
      const long i = e % n;
      const int r = (int)( e / n );
      b[ (long)r * ldb + i ] = (D)v[ i * 16 + r ];
   }
}

// Row interchanges of the factorization applied in order to a row-interleaved block.
template <typename T>
__global__ void k_apply_ipiv_il( int n, const int* __restrict__ ipiv, T* v ) {
   const int r = threadIdx.x;
   if ( r >= 16 ) return;
   for ( int k = 0; k < n; ++k ) {
      const int p = ipiv[ k ] - 1;
      if ( p != k ) {
         const T t = v[ (long)k * 16 + r ];
         v[ (long)k * 16 + r ] = v[ (long)p * 16 + r ];
         v[ (long)p * 16 + r ] = t;
      }
   }
}

// The diagonal blocks of the super-block inverses: XL and XU hold one S x S matrix per super-block
// q (columns q S .. q S + S - 1, leading dimension S); this writes the 128 x 128 inverses of the
// factorization on their diagonals and zeros elsewhere.
template <typename T>
__global__ void k_super_init( int S, int nsb, const T* __restrict__ inv, T* __restrict__ XL, T* __restrict__ XU ) {
   const long total = (long)S * S * nsb;
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x ) {
      const long q = e / ( (long)S * S );
      const int rem = (int)( e % ( (long)S * S ) );
      const int c = rem / S, r = rem % S;
      T l = T( 0 ), u = T( 0 );
      if ( r / kNB == c / kNB ) {
         const long k = q * ( S / kNB ) + c / kNB;
         const T* blk = inv + k * 2 * kNB * kNB;
         l = blk[ ( c % kNB ) * kNB + r % kNB ];
         u = blk[ kNB * kNB + ( c % kNB ) * kNB + r % kNB ];
      }
      XL[ e ] = l;
      XU[ e ] = u;
   }
}

// Row interchanges of the factorization applied to b, in order (DLASWP forward).
template <typename T>
__global__ void k_apply_ipiv( int n, int nrhs, const int* __restrict__ ipiv, T* b, long ldb ) {
   const int r = threadIdx.x;
   if ( r >= nrhs ) return;
   T* x = b + (long)r * ldb;
   for ( int k = 0; k < n; ++k ) {
      const int p = ipiv[ k ] - 1;
      if ( p != k ) {
         const T t = x[ k ];
         x[ k ] = x[ p ];
         x[ p ] = t;
      }
   }
}

// ------------------------------------------------------------------ refinement

// Partial products of A x over column chunks: part[c][rhs][row] = sum_{j in chunk c} A(row, j) x(j, rhs).
// Each thread carries four rows (row0 + 256 m) so every broadcast read of x from LDS feeds four
// loads of A.
__global__ void __launch_bounds__( 256 ) k_resid_part( int npad, int chunk, int nrhs, const double* __restrict__ A,
                                                      long lda, const double* __restrict__ x, long ldx,
                                                      double* __restrict__ part ) {
   extern __shared__ __attribute__( ( aligned( 16 ) ) ) double xs[];  // chunk x 16
   const int c0 = blockIdx.y * chunk;
   for ( int idx = threadIdx.x; idx < chunk * 16; idx += 256 ) {
      const int r = idx / chunk, j = idx % chunk;
      xs[ j * 16 + r ] = r < nrhs ? x[ (long)r * ldx + c0 + j ] : 0.0;
   }
   __syncthreads();
   const int row0 = blockIdx.x * 1024 + threadIdx.x;
   double acc[ 4 ][ 16 ];
#pragma unroll
   for ( int m = 0; m < 4; ++m )
#pragma unroll
      for ( int r = 0; r < 16; ++r ) acc[ m ][ r ] = 0.0;
   bool ok[ 4 ];
#pragma unroll
   for ( int m = 0; m < 4; ++m ) ok[ m ] = row0 + 256 * m < npad;
   const double* a = A + (long)c0 * lda + row0;
   for ( int j = 0; j < chunk; ++j ) {
      double v[ 4 ];
#pragma unroll
      for ( int m = 0; m < 4; ++m ) v[ m ] = ok[ m ] ? a[ (long)j * lda + 256 * m ] : 0.0;
      const double* xr = xs + j * 16;
#pragma unroll
      for ( int r = 0; r < 16; ++r ) {
         const double xv = xr[ r ];
#pragma unroll
         for ( int m = 0; m < 4; ++m ) acc[ m ][ r ] = acc[ m ][ r ] + v[ m ] * xv;
      }
   }
   double* out = part + (long)blockIdx.y * 16 * npad;
#pragma unroll
   for ( int m = 0; m < 4; ++m )
#pragma unroll
      for ( int r = 0; r < 16; ++r )
         if ( ok[ m ] && r < nrhs ) out[ (long)r * npad + row0 + 256 * m ] = acc[ m ][ r ];
}

__device__ inline void atomic_max_nonneg( double* addr, double v ) {
   // Non-negative doubles order as their bit patterns do.
   atomicMax( reinterpret_cast<unsigned long long*>( addr ), (unsigned long long)__double_as_longlong( v ) );
}

// r = b - sum of the chunks (in chunk order); norms[r] = max |r(:, rhs)| over rows < n3.
__global__ void __launch_bounds__( 256 ) k_resid_final( int npad, int n3, int nchunks, int nrhs,
                                                       const double* __restrict__ part, const double* __restrict__ b,
                                                       long ldb, double* __restrict__ r, long ldr, double* norms ) {
   const int row = blockIdx.x * 256 + threadIdx.x;
   if ( row >= npad ) return;
   for ( int k = 0; k < nrhs; ++k ) {
      double s = 0.0;
      for ( int c = 0; c < nchunks; ++c ) s = s + part[ (long)c * 16 * npad + (long)k * npad + row ];
      const double v = row < n3 ? b[ (long)k * ldb + row ] - s : 0.0;
      r[ (long)k * ldr + row ] = v;
      atomic_max_nonneg( &norms[ k ], fabs( v ) );
   }
}

// x += d (d in FP32 from the low-precision solve); norms[0..nrhs) = max|d|, norms[16..) = max|x|.
__global__ void __launch_bounds__( 256 ) k_ir_update( int n3, int nrhs, double* __restrict__ x, long ldx,
                                                     const float* __restrict__ d, long ldd, double* norms ) {
   const int row = blockIdx.x * 256 + threadIdx.x;
   if ( row >= n3 ) return;
   for ( int k = 0; k < nrhs; ++k ) {
      const double dv = (double)d[ (long)row * 16 + k ];
      const double xv = x[ (long)k * ldx + row ] + dv;
      x[ (long)k * ldx + row ] = xv;
      atomic_max_nonneg( &norms[ k ], fabs( dv ) );
      atomic_max_nonneg( &norms[ 16 + k ], fabs( xv ) );
   }
}

__global__ void k_d2f( long n, const double* __restrict__ a, float* __restrict__ b ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) b[ i ] = (float)a[ i ];
}

__global__ void k_f2d( long n, const float* __restrict__ a, double* __restrict__ b ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) b[ i ] = (double)a[ i ];
}

// ------------------------------------------------------------------ host side

template <typename T>
void launch_gemm_batched( hipStream_t st, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta,
                          T* C, long ldc, int* check, int batch, int zdiv, BatchStride sa, BatchStride sb,
                          BatchStride sc ) {
   if ( M <= 0 || N <= 0 || batch <= 0 ) return;
   dim3 grid( M / 128, N / 128, batch );
   ProfScope prof( sizeof( T ) == 8 ? kKGemm64 : kKGemm32, 2.0 * M * N * K * batch, st );
   if constexpr ( sizeof( T ) == 8 )
      hipLaunchKernelGGL( ( k_gemm<double, 16> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check,
                          zdiv, sa, sb, sc );
   else
      hipLaunchKernelGGL( ( k_gemm<float, 32> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check,
                          zdiv, sa, sb, sc );
}

template <typename T>
void launch_gemm( hipStream_t st, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta, T* C,
                  long ldc, int* check ) {
   launch_gemm_batched<T>( st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, 1, 1, {}, {}, {} );
}
template void launch_gemm<double>( hipStream_t, int, int, int, double, const double*, long, const double*, long, double,
                                   double*, long, int* );
template void launch_gemm<float>( hipStream_t, int, int, int, float, const float*, long, const float*, long, float, float*,
                                  long, int* );

// The blocked factorization without exchanges (the fast path), with a look-ahead of one block
// column: the trailing update of step k is split into the next block column (on st, where the
// diagonal block and the panels of step k+1 follow at once) and the rest (on st2), so the
// one-workgroup diagonal factorization of step k+1 runs while the GPU is busy with the bulk of
// step k's update.  ev_u and ev_g order the two streams (U panel k before its update on st2; the
// update on st2 before U panel k+1, which reads the rows it writes).  check is set when partial
// pivoting would have exchanged rows somewhere; the caller must then rerun lu_pivoted on the
// original matrix.
template <typename T>
void lu_nopiv( hipStream_t st, hipStream_t st2, hipEvent_t ev_u, hipEvent_t ev_g, int npad, T* A, long lda, T* inv,
               int* check ) {
   const int nb = kNB, nblk = npad / nb;
   auto diag = [&]( int k ) {
      const int kk = k * nb;
      ProfScope prof( kKDiag, 2.0 * nb * nb * nb, st );  // LU 2/3 nb^3 + two triangular inverses 2/3 nb^3
      hipLaunchKernelGGL( k_lu_diag<T>, dim3( 1 ), dim3( 256 ), 0, st, 0, A + (long)kk * lda + kk, lda,
                          inv + (long)k * 2 * nb * nb, check );
   };
   auto lpanel = [&]( int k ) {  // A_below <- A_below U^-1 (checks |l| <= 1)
      const int kk = k * nb, rest = npad - kk - nb;
      T* Akk = A + (long)kk * lda + kk;
      launch_gemm<T>( st, rest, nb, nb, T( 1 ), Akk + nb, lda, inv + (long)k * 2 * nb * nb + nb * nb, nb, T( 0 ), Akk + nb, lda,
                      check );
   };
   auto upanel = [&]( int k ) {  // A_right <- L^-1 A_right
      const int kk = k * nb, rest = npad - kk - nb;
      T* Akk = A + (long)kk * lda + kk;
      launch_gemm<T>( st, nb, rest, nb, T( 1 ), inv + (long)k * 2 * nb * nb, nb, Akk + (long)nb * lda, lda, T( 0 ),
                      Akk + (long)nb * lda, lda, nullptr );
   };
   diag( 0 );
   if ( nblk == 1 ) return;
   lpanel( 0 );
   upanel( 0 );
   for ( int k = 0; k + 1 < nblk; ++k ) {
      const int kk = k * nb, rest = npad - kk - nb;
      T* Akk = A + (long)kk * lda + kk;
      T* L = Akk + nb;
      T* U = Akk + (long)nb * lda;
      // next block column of the trailing matrix, on st
      launch_gemm<T>( st, rest, nb, nb, T( -1 ), L, lda, U, lda, T( 1 ), U + nb, lda, nullptr );
      // the rest of the trailing update, on st2, after U panel k
      if ( rest > nb ) {
         BH_HIP( hipEventRecord( ev_u, st ) );
         BH_HIP( hipStreamWaitEvent( st2, ev_u, 0 ) );
         launch_gemm<T>( st2, rest, rest - nb, nb, T( -1 ), L, lda, U + (long)nb * lda, lda, T( 1 ), U + (long)nb * lda + nb,
                         lda, nullptr );
         BH_HIP( hipEventRecord( ev_g, st2 ) );
      }
      diag( k + 1 );
      if ( k + 2 < nblk ) {
         lpanel( k + 1 );
         BH_HIP( hipStreamWaitEvent( st, ev_g, 0 ) );
         upanel( k + 1 );
      }
   }
}
template void lu_nopiv<double>( hipStream_t, hipStream_t, hipEvent_t, hipEvent_t, int, double*, long, double*, int* );
template void lu_nopiv<float>( hipStream_t, hipStream_t, hipEvent_t, hipEvent_t, int, float*, long, float*, int* );

// True partial pivoting (slow path): the panel is factored column by column with a grid-wide
// pivot search; interchanges are applied across the whole row at once.
template <typename T>
void lu_pivoted( hipStream_t st, int npad, T* A, long lda, T* inv, int* ipiv, int* info, double* pval, int* pidx ) {
   const int nb = kNB, nblk = npad / nb;
   const int nparts = 64;
   for ( int k = 0; k < nblk; ++k ) {
      const int kk = k * nb;
      for ( int j = kk; j < kk + nb; ++j ) {
         hipLaunchKernelGGL( k_piv_search<T>, dim3( nparts ), dim3( 256 ), 0, st, npad, j, A, lda, pval, pidx );
         hipLaunchKernelGGL( k_piv_swap<T>, dim3( ( npad + 255 ) / 256 ), dim3( 256 ), 0, st, npad, j, nparts, pval, pidx,
                             A, lda, ipiv, info );
         hipLaunchKernelGGL( k_piv_update<T>, dim3( ( npad + 255 ) / 256 ), dim3( 256 ), 0, st, npad, j, kk + nb, A, lda );
      }
      T* Akk = A + (long)kk * lda + kk;
      T* invk = inv + (long)k * 2 * nb * nb;
      hipLaunchKernelGGL( k_lu_diag<T>, dim3( 1 ), dim3( 256 ), 0, st, 1, Akk, lda, invk, nullptr );
      const int rest = npad - kk - nb;
      if ( rest <= 0 ) break;
      T* Abelow = Akk + nb;
      T* Aright = Akk + (long)nb * lda;
      launch_gemm<T>( st, nb, rest, nb, T( 1 ), invk, nb, Aright, lda, T( 0 ), Aright, lda, nullptr );
      launch_gemm<T>( st, rest, rest, nb, T( -1 ), Abelow, lda, Aright, lda, T( 1 ), Aright + nb, lda, nullptr );
   }
}
template void lu_pivoted<double>( hipStream_t, int, double*, long, double*, int*, int*, double*, int* );
template void lu_pivoted<float>( hipStream_t, int, float*, long, float*, int*, int*, double*, int* );

constexpr int kSkinnySlice = 256;  // columns of K per workgroup of k_skinny_part

// Y = beta Y + alpha A X (A column-major M x K, X and Y row-interleaved); P is scratch of
// (K / kSkinnySlice) M 16 elements.  M must be a multiple of 64, K of kSkinnySlice.
template <typename T>
void launch_skinny( hipStream_t st, int M, int K, T alpha, const T* A, long lda, const T* X, T beta, T* Y, T* P ) {
   if ( M <= 0 ) return;
   const int nkc = ( K + kSkinnySlice - 1 ) / kSkinnySlice;
   ProfScope prof( kKTrsm, 2.0 * M * K * 11, st );
   hipLaunchKernelGGL( k_skinny_part<T>, dim3( M / 64, nkc ), dim3( 256 ), 0, st, M, K, kSkinnySlice, A, lda, X, P );
   const long n16 = (long)M * 16;
   hipLaunchKernelGGL( k_skinny_reduce<T>, dim3( (unsigned)std::min<long>( 2048, ( n16 + 255 ) / 256 ) ), dim3( 256 ), 0, st,
                       n16, nkc, alpha, P, beta, Y );
}

// Inverses of the S x S diagonal super-blocks of L (unit lower) and U, from the 128 x 128 block
// inverses by recursive doubling: with A, B the inverses of two adjacent s x s diagonal blocks and
// C the block between them, inv([A 0; C B]) has -B^-1 C A^-1 below (and inv of upper: -A^-1 C B^-1
// right of the diagonal).  Every level is two batched GEMMs per triangle over all pairs of all
// super-blocks.  tmp holds npad * S / 2 elements.
template <typename T>
void build_super_inverses( hipStream_t st, int npad, int S, const T* LU, long lda, const T* inv, T* XL, T* XU, T* tmp ) {
   const int nsb = npad / S;
   hipLaunchKernelGGL( k_super_init<T>, dim3( 1024 ), dim3( 256 ), 0, st, S, nsb, inv, XL, XU );
   const long SS = (long)S * S;
   for ( int s = kNB; s < S; s *= 2 ) {
      const int zdiv = S / ( 2 * s ), batch = nsb * zdiv;
      const long ss = (long)s * s;
      BatchStride t_str { zdiv * ss, ss };
      // L: T = C A^-1, X21 = -B^-1 T
      BatchStride c_l { (long)S + (long)S * lda, 2L * s + 2L * s * lda };
      BatchStride a_x { SS, 2L * s * ( S + 1 ) };
      launch_gemm_batched<T>( st, s, s, s, T( 1 ), LU + s, lda, XL, S, T( 0 ), tmp, s, nullptr, batch, zdiv, c_l, a_x, t_str );
      BatchStride x21 { SS, 2L * s * ( S + 1 ) };
      launch_gemm_batched<T>( st, s, s, s, T( -1 ), XL + (long)s * ( S + 1 ), S, tmp, s, T( 0 ), XL + s, S, nullptr, batch, zdiv,
                              a_x, t_str, x21 );
      // U: T = C B^-1, X12 = -A^-1 T
      BatchStride c_u { (long)S + (long)S * lda, 2L * s + 2L * s * lda };
      launch_gemm_batched<T>( st, s, s, s, T( 1 ), LU + (long)s * lda, lda, XU + (long)s * ( S + 1 ), S, T( 0 ), tmp, s, nullptr,
                              batch, zdiv, c_u, a_x, t_str );
      launch_gemm_batched<T>( st, s, s, s, T( -1 ), XU, S, tmp, s, T( 0 ), XU + (long)s * S, S, nullptr, batch, zdiv, a_x, t_str,
                              x21 );
   }
}
template void build_super_inverses<double>( hipStream_t, int, int, const double*, long, const double*, double*, double*,
                                            double* );
template void build_super_inverses<float>( hipStream_t, int, int, const float*, long, const float*, float*, float*, float* );

// Solves L U x = P b with the factors and the super-block inverses.  b, y, x are row-interleaved
// npad x 16 blocks: b is consumed (row-exchanged when pivoted, then overwritten by the forward
// updates), y is work space, the solution goes to x.  Two launches per super-block and triangle:
// the inverse times the block, then the update of the rest of the right-hand side.
template <typename T>
void lu_solve( hipStream_t st, int npad, int n3, int S, const T* LU, long lda, const T* XL, const T* XU, const int* ipiv,
               bool pivoted, T* b, T* y, T* x, T* P ) {
   const int nsb = npad / S;
   if ( pivoted ) hipLaunchKernelGGL( k_apply_ipiv_il<T>, dim3( 1 ), dim3( 64 ), 0, st, n3, ipiv, b );
   for ( int q = 0; q < nsb; ++q ) {
      const long q0 = (long)q * S;
      launch_skinny<T>( st, S, S, T( 1 ), XL + q * (long)S * S, S, b + q0 * 16, T( 0 ), y + q0 * 16, P );
      launch_skinny<T>( st, npad - (int)q0 - S, S, T( -1 ), LU + q0 * lda + q0 + S, lda, y + q0 * 16, T( 1 ),
                        b + ( q0 + S ) * 16, P );
   }
   for ( int q = nsb - 1; q >= 0; --q ) {
      const long q0 = (long)q * S;
      launch_skinny<T>( st, S, S, T( 1 ), XU + q * (long)S * S, S, y + q0 * 16, T( 0 ), x + q0 * 16, P );
      launch_skinny<T>( st, (int)q0, S, T( -1 ), LU + q0 * lda, lda, x + q0 * 16, T( 1 ), y, P );
   }
}
template void lu_solve<double>( hipStream_t, int, int, int, const double*, long, const double*, const double*, const int*,
                                bool, double*, double*, double*, double* );
template void lu_solve<float>( hipStream_t, int, int, int, const float*, long, const float*, const float*, const int*, bool,
                               float*, float*, float*, float* );

template <typename S, typename D>
void to_interleaved( hipStream_t st, int npad, int nrhs, const S* b, long ldb, D* v ) {
   hipLaunchKernelGGL( ( k_to_interleaved<S, D> ), dim3( 1024 ), dim3( 256 ), 0, st, npad, nrhs, b, ldb, v );
}
template <typename S, typename D>
void from_interleaved( hipStream_t st, int n, int nrhs, const S* v, D* b, long ldb ) {
   hipLaunchKernelGGL( ( k_from_interleaved<S, D> ), dim3( 1024 ), dim3( 256 ), 0, st, n, nrhs, v, b, ldb );
}
template void to_interleaved<double, double>( hipStream_t, int, int, const double*, long, double* );
template void to_interleaved<double, float>( hipStream_t, int, int, const double*, long, float* );
template void from_interleaved<double, double>( hipStream_t, int, int, const double*, double*, long );

// Column chunk of the residual product: the largest of 512, 256, 128 that divides npad.
static int resid_chunk( int npad ) {
   for ( int c = 512; c > 128; c /= 2 )
      if ( npad % c == 0 ) return c;
   return 128;
}

// r = b - A x over rows < n3 (FP64), norms[0..nrhs) <- max |r| (norms must be zeroed).
void residual( hipStream_t st, int npad, int n3, int nrhs, const double* A, long lda, const double* x, long ldx,
               const double* b, long ldb, double* r, long ldr, double* part, double* norms ) {
   const int chunk = resid_chunk( npad );
   const int nchunks = npad / chunk;
   dim3 grid( ( npad + 1023 ) / 1024, nchunks );
   ProfScope prof( kKResid, 2.0 * npad * npad * nrhs, st );
   hipLaunchKernelGGL( k_resid_part, grid, dim3( 256 ), chunk * 16 * sizeof( double ), st, npad, chunk, nrhs, A, lda, x,
                       ldx, part );
   hipLaunchKernelGGL( k_resid_final, dim3( ( npad + 255 ) / 256 ), dim3( 256 ), 0, st, npad, n3, nchunks, nrhs, part, b,
                       ldb, r, ldr, norms );
}

size_t residual_part_elems( int npad ) { return (size_t)( npad / resid_chunk( npad ) ) * 16 * npad; }

void ir_update( hipStream_t st, int n3, int nrhs, double* x, long ldx, const float* d, long ldd, double* norms ) {
   hipLaunchKernelGGL( k_ir_update, dim3( ( n3 + 255 ) / 256 ), dim3( 256 ), 0, st, n3, nrhs, x, ldx, d, ldd, norms );
}

void convert_d2f( hipStream_t st, long n, const double* a, float* b ) {
   hipLaunchKernelGGL( k_d2f, dim3( 2048 ), dim3( 256 ), 0, st, n, a, b );
}

void convert_f2d( hipStream_t st, long n, const float* a, double* b ) {
   hipLaunchKernelGGL( k_f2d, dim3( 2048 ), dim3( 256 ), 0, st, n, a, b );
}

}  // namespace bhg
