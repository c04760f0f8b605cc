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
// 256 threads: four waves in a 2 x 2 arrangement, each owning a 64 x 64 block of the tile as
// 4 x 4 MFMA 16 x 16 accumulators.  A and B slices go through LDS ([k][m] and [k][n], so a
// wave's 16 lanes of one k read 16 consecutive elements); the next slice is fetched into
// registers while the current one is multiplied.  C may alias A (when N == 128: each workgroup
// reads its whole row strip before it writes) or B (when M == 128).  check != nullptr: set
// *check when any |C| > 1 (the pivot test of the panel below the diagonal block).
template <typename T, int BK>
__global__ void __launch_bounds__( 256 ) k_gemm( int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb,
                                                T beta, T* C, long ldc, int* check ) {
   constexpr int BM = 128, BN = 128, PAD = 4;
   __shared__ T As[ BK ][ BM + PAD ];
   __shared__ T Bs[ BK ][ BN + PAD ];
   using v4 = typename Mfma<T>::v4;
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
   T ra[ AE ], rb[ BE ];
   v4 acc[ 4 ][ 4 ];
#pragma unroll
   for ( int t = 0; t < 4; ++t )
#pragma unroll
      for ( int u = 0; u < 4; ++u ) acc[ t ][ u ] = v4{ 0, 0, 0, 0 };
   const int nk = K / BK;
   auto gload = [&]( int kt ) {
      const T* pa = Ab + (long)( kt * BK + a_k ) * lda + a_m;
#pragma unroll
      for ( int e = 0; e < AE; ++e ) ra[ e ] = pa[ e ];
      const T* pb = Bb + (long)b_n * ldb + kt * BK + b_k;
#pragma unroll
      for ( int e = 0; e < BE; ++e ) rb[ e ] = pb[ e ];
   };
   gload( 0 );
   for ( int kt = 0; kt < nk; ++kt ) {
      __syncthreads();
#pragma unroll
      for ( int e = 0; e < AE; ++e ) As[ a_k ][ a_m + e ] = ra[ e ];
#pragma unroll
      for ( int e = 0; e < BE; ++e ) Bs[ b_k + e ][ b_n ] = rb[ e ];
      __syncthreads();
      if ( kt + 1 < nk ) gload( kt + 1 );
#pragma unroll
      for ( int kk = 0; kk < BK; kk += 4 ) {
         T af[ 4 ], bf[ 4 ];
         const int kr = kk + ( lane >> 4 ), c16 = lane & 15;
#pragma unroll
         for ( int t = 0; t < 4; ++t ) af[ t ] = As[ kr ][ wm * 64 + t * 16 + c16 ];
#pragma unroll
         for ( int u = 0; u < 4; ++u ) bf[ u ] = Bs[ kr ][ wn * 64 + u * 16 + c16 ];
#pragma unroll
         for ( int t = 0; t < 4; ++t )
#pragma unroll
            for ( int u = 0; u < 4; ++u ) acc[ t ][ u ] = Mfma<T>::run( af[ t ], bf[ u ], acc[ t ][ u ] );
      }
   }
   bool big = false;
#pragma unroll
   for ( int t = 0; t < 4; ++t )
#pragma unroll
      for ( int u = 0; u < 4; ++u ) {
         const int col = bn * BN + wn * 64 + u * 16 + ( lane & 15 );
         T* c = C + (long)col * ldc + bm * BM + wm * 64 + t * 16;
#pragma unroll
         for ( int r = 0; r < 4; ++r ) {
            // Accumulator element r of this lane: row 4 (lane / 16) + r for the FP32 instruction,
            // row lane / 16 + 4 r for the FP64 one (measured lane maps of the two instructions).
            const int row = sizeof( T ) == 8 ? ( lane >> 4 ) + 4 * r : 4 * ( lane >> 4 ) + r;
            T v = alpha * acc[ t ][ u ][ r ];
            if ( beta != T( 0 ) ) v = v + beta * c[ row ];
            c[ row ] = v;
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
// 8 ty.., columns 8 tx.. in registers.  Each elimination step broadcasts pivot row and column
// through LDS (double-buffered, so one barrier per step) and every thread applies the rank-1
// update to its own tile.  The multipliers are a_ij * (1 / a_jj), as DGETF2 scales.
// The inversions run in place in the same loop: U^-1 by rows from the bottom (step s works on row
// 127 - s: X(j,:) = Y(j,:) / u_jj, then Y(i,:) -= u_ij X(j,:) for i < j, where a position still
// holding u_ic becomes X's entry when column c is reached) and L^-1 by rows from the top (unit
// diagonal).  The two touch disjoint halves of the tile grid.
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
      for ( int j = 0; j < NB; ++j ) {
         const int bsel = j & 1, jb = j >> 3, jo = j & 7;
         if ( tx == jb ) {
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) colb[ bsel ][ r0 + ii ] = a[ ii ][ jo ];
         }
         if ( ty == jb ) {
#pragma unroll
            for ( int jj = 0; jj < RB; ++jj ) rowb[ bsel ][ c0 + jj ] = a[ jo ][ jj ];
         }
         __syncthreads();
         if ( r0 + RB - 1 <= j ) continue;  // no rows below the pivot in this tile
         const T piv = rowb[ bsel ][ j ];
         const T rp = T( 1 ) / piv;
         T l[ RB ];
#pragma unroll
         for ( int ii = 0; ii < RB; ++ii ) {
            const int i = r0 + ii;
            const T v = colb[ bsel ][ i ];
            const bool below = i > j;
            if ( below ) big |= !( tabs( v ) <= tabs( piv ) );
            l[ ii ] = below ? v * rp : T( 0 );
         }
         if ( tx == jb ) {
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii )
               if ( r0 + ii > j ) a[ ii ][ jo ] = l[ ii ];
         }
         if ( c0 + RB - 1 <= j ) continue;
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) {
            if ( c0 + jj <= j ) continue;
            const T u = rowb[ bsel ][ c0 + jj ];
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) a[ ii ][ jj ] = a[ ii ][ jj ] - l[ ii ] * u;
         }
      }
      if ( big ) atomicOr( check, 1 );
#pragma unroll
      for ( int jj = 0; jj < RB; ++jj )
#pragma unroll
         for ( int ii = 0; ii < RB; ++ii ) A[ (long)( c0 + jj ) * lda + r0 + ii ] = a[ ii ][ jj ];
   }
   __syncthreads();
   for ( int st = 0; st < NB; ++st ) {
      const int bsel = st & 1;
      const int ju = NB - 1 - st, jl = st;
      // U^-1 step on row ju: broadcast u_{i,ju} (i < ju) and Y(ju, c >= ju) with u_{ju,ju}.
      if ( tx == ( ju >> 3 ) ) {
#pragma unroll
         for ( int ii = 0; ii < RB; ++ii ) colb[ bsel ][ r0 + ii ] = a[ ii ][ ju & 7 ];
      }
      if ( ty == ( ju >> 3 ) ) {
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) rowb[ bsel ][ c0 + jj ] = a[ ju & 7 ][ jj ];
      }
      // L^-1 step on row jl: broadcast l_{i,jl} (i > jl) and X(jl, c < jl).
      if ( tx == ( jl >> 3 ) ) {
#pragma unroll
         for ( int ii = 0; ii < RB; ++ii ) colL[ bsel ][ r0 + ii ] = a[ ii ][ jl & 7 ];
      }
      if ( ty == ( jl >> 3 ) ) {
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) rowL[ bsel ][ c0 + jj ] = a[ jl & 7 ][ jj ];
      }
      __syncthreads();
      const T d = T( 1 ) / rowb[ bsel ][ ju ];
      // row ju itself: X(ju, ju) = d, X(ju, c) = Y(ju, c) d for c > ju
      if ( ty == ( ju >> 3 ) ) {
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) {
            const int c = c0 + jj;
            if ( c == ju ) a[ ju & 7 ][ jj ] = d;
            else if ( c > ju ) a[ ju & 7 ][ jj ] = a[ ju & 7 ][ jj ] * d;
         }
      }
      // rows i < ju: position (i, ju) <- -u_{i,ju} d; (i, c > ju) -= u_{i,ju} X(ju, c)
      if ( r0 < ju && c0 + RB - 1 >= ju ) {
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) {
            const int c = c0 + jj;
            if ( c < ju ) continue;
            const T xj = c == ju ? d : rowb[ bsel ][ c ] * d;
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) {
               const int i = r0 + ii;
               if ( i >= ju ) continue;
               const T uij = colb[ bsel ][ i ];
               a[ ii ][ jj ] = c == ju ? -uij * d : a[ ii ][ jj ] - uij * xj;
            }
         }
      }
      // rows i > jl: position (i, jl) <- -l_{i,jl}; (i, c < jl) -= l_{i,jl} X(jl, c)
      if ( r0 + RB - 1 > jl && c0 <= jl ) {
#pragma unroll
         for ( int jj = 0; jj < RB; ++jj ) {
            const int c = c0 + jj;
            if ( c > jl ) continue;
            const T xj = rowL[ bsel ][ c ];
#pragma unroll
            for ( int ii = 0; ii < RB; ++ii ) {
               const int i = r0 + ii;
               if ( i <= jl ) continue;
               const T lij = colL[ bsel ][ i ];
               a[ ii ][ jj ] = c == jl ? -lij : a[ ii ][ jj ] - lij * xj;
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

// One launch solves L Y = B (lower = 1: unit lower factor, block rows top to bottom) or U X = Y
// (lower = 0: upper factor, bottom to top) for the nrhs <= 16 columns of b (npad x nrhs, in
// place), using the factor panels of LU and the inverted diagonal blocks inv.
//
// Workgroup t (in the order workgroups start, from the ticket counter) owns the t-th block row
// in solve order.  It accumulates the products with every block of the solution that precedes
// it as soon as that block's flag is raised, then multiplies by its inverted diagonal block,
// publishes its block (agent-scope release) and raises its own flag.  A workgroup only waits for
// workgroups that drew earlier tickets, which are already running, so the chain always advances;
// each wait is bounded (spin_limit), after which *err is set and the kernel ends normally.
// done[] and *ticket must be zero at launch.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_trsm_chain( int nblk, int lower, int nrhs, const T* __restrict__ LU, long lda,
                                                      const T* __restrict__ inv, T* b, long ldb, int* ticket, int* done,
                                                      int* err, int spin_limit ) {
   constexpr int NB = kNB;
   __shared__ T xs[ NB ][ 17 ];
   __shared__ T red[ 2 ][ NB ][ 17 ];
   __shared__ int s_t;
   const int tid = threadIdx.x;
   if ( tid == 0 ) s_t = atomicAdd( ticket, 1 );
   __syncthreads();
   const int t = s_t;
   const int w = lower ? t : nblk - 1 - t;
   const int i = tid & ( NB - 1 ), kh = tid >> 7;
   T acc[ 16 ];
#pragma unroll
   for ( int r = 0; r < 16; ++r ) acc[ r ] = T( 0 );
   for ( int q = 0; q < t; ++q ) {
      const int j = lower ? q : nblk - 1 - q;
      if ( tid == 0 ) {
         int spins = 0;
         while ( __hip_atomic_load( &done[ j ], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT ) == 0 ) {
            __builtin_amdgcn_s_sleep( 1 );
            if ( ++spins > spin_limit ) {
               atomicOr( err, 1 );
               break;
            }
         }
      }
      __syncthreads();
      for ( int idx = tid; idx < NB * nrhs; idx += 256 ) {
         const int k = idx & ( NB - 1 ), r = idx >> 7;
         xs[ k ][ r ] = __hip_atomic_load( &b[ (long)r * ldb + (long)j * NB + k ], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
      }
      __syncthreads();
      const T* Lp = LU + (long)( j * NB + kh * 64 ) * lda + (long)w * NB + i;
#pragma unroll 8
      for ( int k = 0; k < 64; ++k ) {
         const T l = Lp[ (long)k * lda ];
#pragma unroll
         for ( int r = 0; r < 16; ++r ) acc[ r ] = acc[ r ] + l * xs[ kh * 64 + k ][ r ];
      }
      __syncthreads();
   }
   // b_w - sum, then the diagonal block's inverse.
#pragma unroll
   for ( int r = 0; r < 16; ++r ) red[ kh ][ i ][ r ] = acc[ r ];
   __syncthreads();
   for ( int idx = tid; idx < NB * nrhs; idx += 256 ) {
      const int k = idx & ( NB - 1 ), r = idx >> 7;
      xs[ k ][ r ] = b[ (long)r * ldb + (long)w * NB + k ] - ( red[ 0 ][ k ][ r ] + red[ 1 ][ k ][ r ] );
   }
   __syncthreads();
   const T* M = inv + (long)w * 2 * NB * NB + ( lower ? 0 : NB * NB );
   T out[ 8 ];
#pragma unroll
   for ( int r = 0; r < 8; ++r ) out[ r ] = T( 0 );
   const int rb = kh * 8;
#pragma unroll 8
   for ( int k = 0; k < NB; ++k ) {
      const T m = M[ k * NB + i ];
#pragma unroll
      for ( int r = 0; r < 8; ++r ) out[ r ] = out[ r ] + m * xs[ k ][ rb + r ];
   }
   for ( int r = 0; r < 8; ++r )
      if ( rb + r < nrhs ) b[ (long)( rb + r ) * ldb + (long)w * NB + i ] = out[ r ];
   __threadfence();
   __syncthreads();
   if ( tid == 0 ) __hip_atomic_store( &done[ w ], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT );
}

__global__ void k_zero_ints( int n, int* p ) {
   for ( int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x ) p[ i ] = 0;
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
__global__ void __launch_bounds__( 256 ) k_resid_part( int npad, int chunk, int nrhs, const double* __restrict__ A,
                                                      long lda, const double* __restrict__ x, long ldx,
                                                      double* __restrict__ part ) {
   extern __shared__ double xs[];  // chunk x 16
   const int c0 = blockIdx.y * chunk;
   for ( int idx = threadIdx.x; idx < chunk * nrhs; idx += 256 ) {
      const int r = idx / chunk, j = idx % chunk;
      xs[ j * 16 + r ] = x[ (long)r * ldx + c0 + j ];
   }
   __syncthreads();
   const int row = blockIdx.x * 256 + threadIdx.x;
   if ( row >= npad ) return;
   double acc[ 16 ];
#pragma unroll
   for ( int r = 0; r < 16; ++r ) acc[ r ] = 0.0;
   const double* a = A + (long)c0 * lda + row;
   for ( int j = 0; j < chunk; ++j ) {
      const double v = a[ (long)j * lda ];
#pragma unroll
      for ( int r = 0; r < 16; ++r )
         if ( r < nrhs ) acc[ r ] = acc[ r ] + v * xs[ j * 16 + r ];
   }
   double* out = part + (long)blockIdx.y * 16 * npad;
   for ( int r = 0; r < nrhs; ++r ) out[ (long)r * npad + row ] = acc[ r ];
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
      const double dv = (double)d[ (long)k * ldd + row ];
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
void launch_gemm( hipStream_t st, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta, T* C,
                  long ldc, int* check ) {
   if ( M <= 0 || N <= 0 ) return;
   dim3 grid( M / 128, N / 128 );
   ProfScope prof( sizeof( T ) == 8 ? kKGemm64 : kKGemm32, 2.0 * M * N * K );
   if constexpr ( sizeof( T ) == 8 )
      hipLaunchKernelGGL( ( k_gemm<double, 16> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check );
   else
      hipLaunchKernelGGL( ( k_gemm<float, 32> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check );
}
template void launch_gemm<double>( hipStream_t, int, int, int, double, const double*, long, const double*, long, double,
                                   double*, long, int* );
template void launch_gemm<float>( hipStream_t, int, int, int, float, const float*, long, const float*, long, float, float*,
                                  long, int* );

// The blocked factorization without exchanges (the fast path).  check is set when partial
// pivoting would have exchanged rows somewhere; the factors are then not LAPACK's and the
// caller must rerun getrf_pivoted on the original matrix.
template <typename T>
void lu_nopiv( hipStream_t st, int npad, T* A, long lda, T* inv, int* check ) {
   const int nb = kNB, nblk = npad / nb;
   for ( int k = 0; k < nblk; ++k ) {
      const int kk = k * nb;
      T* Akk = A + (long)kk * lda + kk;
      T* invk = inv + (long)k * 2 * nb * nb;
      {
         ProfScope prof( kKDiag, 2.0 * nb * nb * nb );  // LU 2/3 nb^3 + two triangular inverses 2/3 nb^3
         hipLaunchKernelGGL( k_lu_diag<T>, dim3( 1 ), dim3( 256 ), 0, st, 0, Akk, lda, invk, check );
      }
      const int rest = npad - kk - nb;
      if ( rest <= 0 ) break;
      T* Abelow = Akk + nb;                 // rows kk+nb.., columns kk..kk+nb
      T* Aright = Akk + (long)nb * lda;     // rows kk..kk+nb, columns kk+nb..
      T* Atrail = Aright + nb;
      // L panel: A_below <- A_below U^-1 (checks |l| <= 1)
      launch_gemm<T>( st, rest, nb, nb, T( 1 ), Abelow, lda, invk + nb * nb, nb, T( 0 ), Abelow, lda, check );
      // U panel: A_right <- L^-1 A_right
      launch_gemm<T>( st, nb, rest, nb, T( 1 ), invk, nb, Aright, lda, T( 0 ), Aright, lda, nullptr );
      // trailing update
      launch_gemm<T>( st, rest, rest, nb, T( -1 ), Abelow, lda, Aright, lda, T( 1 ), Atrail, lda, nullptr );
   }
}
template void lu_nopiv<double>( hipStream_t, int, double*, long, double*, int* );
template void lu_nopiv<float>( hipStream_t, int, float*, long, float*, int* );

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

// Solves with the factors: b (npad x nrhs) in place.  pivoted: apply ipiv first.  sync holds
// nblk + 2 ints of scratch (flags, ticket, error); the error flag is left at sync[nblk + 1].
template <typename T>
void lu_solve( hipStream_t st, int npad, int n3, const T* LU, long lda, const T* inv, const int* ipiv, bool pivoted,
               int nrhs, T* b, long ldb, int* sync ) {
   const int nblk = npad / kNB;
   const int spin_limit = 1 << 24;
   if ( pivoted ) hipLaunchKernelGGL( k_apply_ipiv<T>, dim3( 1 ), dim3( 64 ), 0, st, n3, nrhs, ipiv, b, ldb );
   int* done = sync;
   int* ticket = sync + nblk;
   int* err = sync + nblk + 1;
   for ( int lower = 1; lower >= 0; --lower ) {
      hipLaunchKernelGGL( k_zero_ints, dim3( 1 ), dim3( 256 ), 0, st, nblk + 1, sync );
      ProfScope prof( kKTrsm, 1.0 * npad * npad * nrhs );  // half of the 2 n^2 nrhs of both triangles
      hipLaunchKernelGGL( k_trsm_chain<T>, dim3( nblk ), dim3( 256 ), 0, st, nblk, lower, nrhs, LU, lda, inv, b, ldb, ticket,
                          done, err, spin_limit );
   }
}
template void lu_solve<double>( hipStream_t, int, int, const double*, long, const double*, const int*, bool, int, double*,
                                long, int* );
template void lu_solve<float>( hipStream_t, int, int, const float*, long, const float*, const int*, bool, int, float*,
                               long, int* );

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
   dim3 grid( ( npad + 255 ) / 256, nchunks );
   ProfScope prof( kKResid, 2.0 * npad * npad * nrhs );
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
