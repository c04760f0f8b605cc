// Dense LU with partial pivoting, triangular solves and FP64 iterative refinement on MI355X.
//
// Replaces LAPACK DGETRF / DGETRS as BEST calls them (best/best.f:444, best/besolve.f90:133).
//
// The factorization is blocked and right-looking with kNB = 128 columns per step:
//   k_lu_diag      the 128 x 128 diagonal block, factored in LDS by one workgroup, then both of
//                  its triangles inverted in place (LAPACK dtrti2 order);
//   k_gemm         128 x 128 MFMA tiles (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32,
//                  FP32 accumulation): the panel below (A U^-1), the row panel right (L^-1 A) and
//                  the trailing update (A -= L U); smaller tiles for the small products, and
//                  k_gemm_pair / k_gemm_multi for two to four products in one launch.
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

// C (M x N) = alpha A (M x K) B (K x N) + beta C, column-major; M, N multiples of the tile (BM x BN,
// each 64 or 128), K of BK.  256 threads: four waves in a 2 x 2 arrangement, each owning a
// BM / 2 x BN / 2 block of the tile as MFMA 16 x 16 accumulators (4 x 4 for the 128 x 128 tile).
// A and B slices of BK columns pass through LDS in two stages: while one stage feeds the MFMAs,
// the next slice is already in flight from global memory (one barrier per slice).  In LDS each
// wave's rows (columns for B) are stored permuted, row 16 t + i at slot (W / 16) i + t, so the
// fragments a lane needs for one k come as one read (16 bytes for FP32 and a 64-row wave block).
// With beta != 0 the C tile is loaded into registers before the main loop.  C may alias A (when
// N == BN: each workgroup reads its whole row strip before it writes) or B (when M == BM).  check != nullptr: set *check when any |C| > thr or is not
// finite (the pivot test of the panel below the diagonal block: thr = 1 is partial pivoting's).
// Batched over blockIdx.z: batch z = zdiv * q + t offsets A by q sa.hi + t sa.lo, and likewise B, C.
struct BatchStride {
   long hi = 0, lo = 0;
};

// Waves per SIMD the register allocation must allow: three for FP32 with 16-column slices (LDS then
// holds three workgroups per CU, and the accumulators move to ordinary VGPRs without spills), two
// for FP64 (the 64 FP64 accumulators of a 128 x 128 tile in 128 VGPRs, 195 in all, no spills, and
// two workgroups' LDS per CU), else no constraint.
template <typename T, int BK>
struct GemmOcc {
   static constexpr int waves = ( sizeof( T ) == 4 && BK == 16 ) ? 3 : sizeof( T ) == 8 ? 2 : 1;
};

// The operands of one GEMM problem (as k_gemm's arguments).
template <typename T>
struct GemmProb {
   int M, N, K;
   T alpha;
   const T* A;
   long lda;
   const T* B;
   long ldb;
   T beta;
   T* C;
   long ldc;
   int* check;
   T thr;
   // The K range of each tile when an operand is triangular (the rest of its K is zeros, whose
   // products are skipped: the sums keep their bits): 0 all of K; 1 from the tile's first column
   // (B lower triangular); 2 to its last row (A lower); 3 to its last column (B upper); 4 from its
   // first row (A upper).
   int kmode = 0;
};

constexpr int kGemmPad = 8;
// LDS elements of one workgroup's two-stage A and B slices.
template <int BK, int BM, int BN>
constexpr int gemm_lds_elems() {
   return 2 * BK * ( BM + kGemmPad ) + 2 * BK * ( BN + kGemmPad );
}

// Tile (bm, bn) of problem g, by the 256 threads of the workgroup; smem holds gemm_lds_elems().
template <typename T, int BK, int BM, int BN>
__device__ inline void gemm_tile( int bm, int bn, const GemmProb<T>& g, T* smem ) {
   constexpr int PAD = kGemmPad;
   constexpr int WM = BM / 2, WN = BN / 2;  // each wave's block of the tile
   constexpr int TM = WM / 16, TN = WN / 16;  // its MFMA 16 x 16 blocks
   T* As = smem;                            // [2][BK][BM + PAD]
   T* Bs = smem + 2 * BK * ( BM + PAD );    // [2][BK][BN + PAD]
   auto as = [&]( int stg, int k, int m ) -> T& { return As[ ( stg * BK + k ) * ( BM + PAD ) + m ]; };
   auto bs = [&]( int stg, int k, int n ) -> T& { return Bs[ ( stg * BK + k ) * ( BN + PAD ) + n ]; };
   using v4 = typename Mfma<T>::v4;
   const int K = g.K;
   const long lda = g.lda, ldb = g.ldb, ldc = g.ldc;
   const T alpha = g.alpha, beta = g.beta;
   T* C = g.C;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int wm = wave & 1, wn = wave >> 1;
   const int kb = g.kmode == 1 ? bn * BN : g.kmode == 4 ? bm * BM : 0;
   const int ke = g.kmode == 2 ? min( K, ( bm + 1 ) * BM ) : g.kmode == 3 ? min( K, ( bn + 1 ) * BN ) : K;
   const T* Ab = g.A + (long)bm * BM + (long)kb * lda;
   const T* Bb = g.B + (long)bn * BN * ldb + kb;
   constexpr int AE = BM * BK / 256;  // A elements per thread, contiguous along m
   constexpr int ATPK = BM / AE;      // threads per k column of A
   constexpr int BE = BK * BN / 256;  // B elements per thread, contiguous along k
   constexpr int BTPN = BK / BE;      // threads per column of B
   const int a_k = tid / ATPK, a_m = ( tid % ATPK ) * AE;
   const int b_n = tid / BTPN, b_k = ( tid % BTPN ) * BE;
   // permuted slot of row (column) m of the tile within its wave block of W rows: t = (m % W) / 16,
   // i = m % 16 go to slot (W / 16) i + t, so the W / 16 fragments a lane needs for one k are
   // contiguous (one 16-byte read for FP32 and W = 64)
   auto slot = []( int m, int W ) { return ( m / W ) * W + ( m & 15 ) * ( W / 16 ) + ( ( m % W ) >> 4 ); };
   const int b_slot = slot( b_n, WN );
   T ra[ AE ], rb[ BE ];
   v4 acc[ TM ][ TN ];
#pragma unroll
   for ( int t = 0; t < TM; ++t )
#pragma unroll
      for ( int u = 0; u < TN; ++u ) acc[ t ][ u ] = v4{ 0, 0, 0, 0 };
   // C tile of this lane, loaded ahead (element r of accumulator (t, u)).
   const int li = lane & 15, lk = lane >> 4;
   auto crow = [&]( int t, int r ) { return wm * WM + t * 16 + ( sizeof( T ) == 8 ? lk + 4 * r : 4 * lk + r ); };
   // FP64 keeps its 64 accumulators per lane in AGPRs only without the prefetched tile: there it
   // reads C in the epilogue.
   constexpr bool kPre = sizeof( T ) == 4;
   T cpre[ TM ][ TN ][ 4 ];
   const int nk = ( ke - kb ) / BK;
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
      for ( int e = 0; e < AE; ++e ) as( stg, a_k, slot( a_m + e, WM ) ) = ra[ e ];
#pragma unroll
      for ( int e = 0; e < BE; ++e ) bs( stg, b_k + e, b_slot ) = rb[ e ];
   };
   gload( 0 );
   // the C tile after the first slice, so that storing the slice to LDS need not wait for it
   if ( kPre && beta != T( 0 ) ) {
#pragma unroll
      for ( int t = 0; t < TM; ++t )
#pragma unroll
         for ( int u = 0; u < TN; ++u ) {
            const T* c = C + (long)( bn * BN + wn * WN + u * 16 + li ) * ldc + bm * BM;
#pragma unroll
            for ( int r = 0; r < 4; ++r ) cpre[ t ][ u ][ r ] = c[ crow( t, r ) ];
         }
   }
   sstore( 0 );
   __syncthreads();
   for ( int kt = 0; kt < nk; ++kt ) {
      const int stg = kt & 1;
      if ( kt + 1 < nk ) gload( kt + 1 );
#pragma unroll
      for ( int kk = 0; kk < BK; kk += 4 ) {
         const int kr = kk + lk;
         const T* pa = &as( stg, kr, wm * WM + li * TM );
         const T* pb = &bs( stg, kr, wn * WN + li * TN );
         T af[ TM ], bf[ TN ];
#pragma unroll
         for ( int t = 0; t < TM; ++t ) af[ t ] = pa[ t ];
#pragma unroll
         for ( int u = 0; u < TN; ++u ) bf[ u ] = pb[ u ];
#pragma unroll
         for ( int t = 0; t < TM; ++t )
#pragma unroll
            for ( int u = 0; u < TN; ++u ) acc[ t ][ u ] = Mfma<T>::run( af[ t ], bf[ u ], acc[ t ][ u ] );
      }
      if ( kt + 1 < nk ) sstore( stg ^ 1 );
      __syncthreads();
   }
   bool big = false;
#pragma unroll
   for ( int t = 0; t < TM; ++t )
#pragma unroll
      for ( int u = 0; u < TN; ++u ) {
         T* c = C + (long)( bn * BN + wn * WN + u * 16 + li ) * ldc + bm * BM;
#pragma unroll
         for ( int r = 0; r < 4; ++r ) {
            T v = alpha * acc[ t ][ u ][ r ];
            if ( beta != T( 0 ) ) v = v + beta * ( kPre ? cpre[ t ][ u ][ r ] : c[ crow( t, r ) ] );
            c[ crow( t, r ) ] = v;
            big |= !( tabs( v ) <= g.thr );
         }
      }
   if ( g.check && big ) atomicOr( g.check, 1 );
}

template <typename T, int BK, int BM = 128, int BN = 128>
__global__ void __launch_bounds__( 256 ) __attribute__( ( amdgpu_waves_per_eu( GemmOcc<T, BK>::waves ) ) ) k_gemm( int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb,
                                                T beta, T* C, long ldc, int* check, T thr, int zdiv, BatchStride sa,
                                                BatchStride sb, BatchStride sc, int kmode ) {
   __shared__ __attribute__( ( aligned( 16 ) ) ) T smem[ gemm_lds_elems<BK, BM, BN>() ];
   // the smaller tiles carry the factorization's critical path: their waves take issue
   // priority over the wide trailing updates' waves sharing a CU
   if constexpr ( BM < 128 || BN < 128 ) __builtin_amdgcn_s_setprio( 2 );
   {
      const long q = blockIdx.z / zdiv, t = blockIdx.z % zdiv;
      A += q * sa.hi + t * sa.lo;
      B += q * sb.hi + t * sb.lo;
      C += q * sc.hi + t * sc.lo;
   }
   // Tile order.  Workgroups are dispatched round-robin over the 8 XCDs (workgroup id % 8), each
   // XCD with its own L2: give each XCD a contiguous run of the tiles, and order the tiles in
   // groups of kGroupM block rows, column by column, so that the tiles running together on one
   // XCD share a few row strips of A and column strips of B in its L2.
   int bm, bn;
   {
      constexpr int kXcd = 8, kGroupM = 8;
      const int nx = gridDim.x, ntile = nx * gridDim.y;
      const int id0 = blockIdx.y * nx + blockIdx.x;
      const int xcd = id0 % kXcd, loc = id0 / kXcd, base = ntile / kXcd, rem = ntile % kXcd;
      const int id = ( xcd < rem ? xcd * ( base + 1 ) : rem * ( base + 1 ) + ( xcd - rem ) * base ) + loc;
      const int per = kGroupM * gridDim.y, first = ( id / per ) * kGroupM, gm = min( nx - first, kGroupM );
      bm = first + ( id % per ) % gm;
      bn = ( id % per ) / gm;
   }
   gemm_tile<T, BK, BM, BN>( bm, bn, GemmProb<T> { M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, thr, kmode }, smem );
}

// Two independent products in one launch, for the factorization's critical path: p0 on 64 x 128
// tiles (the panel below a diagonal block, written over A: N == 128) and p1 on 128 x 64 tiles (a U
// row panel, written over B: M == 128).  Workgroups 0 .. t0 - 1 take p0's tiles, the rest p1's.
template <typename T, int BK>
__global__ void __launch_bounds__( 256 ) __attribute__( ( amdgpu_waves_per_eu( GemmOcc<T, BK>::waves ) ) ) k_gemm_pair( GemmProb<T> p0, GemmProb<T> p1 ) {
   static_assert( gemm_lds_elems<BK, 64, 128>() == gemm_lds_elems<BK, 128, 64>(), "one LDS size for both shapes" );
   __shared__ __attribute__( ( aligned( 16 ) ) ) T smem[ gemm_lds_elems<BK, 64, 128>() ];
   __builtin_amdgcn_s_setprio( 2 );  // the critical path (as the smaller tiles of k_gemm)
   const int t0 = p0.M / 64;
   if ( (int)blockIdx.x < t0 )
      gemm_tile<T, BK, 64, 128>( blockIdx.x, 0, p0, smem );
   else
      gemm_tile<T, BK, 128, 64>( 0, blockIdx.x - t0, p1, smem );
}

// Up to four independent products on 64 x 64 tiles in one launch (the narrow parts of a pair's
// trailing update: one dependent launch instead of two or four on the update's stream).
// Workgroups 0 .. end[0] - 1 take problem 0's tiles, the next ones problem 1's, and so on.
template <typename T>
struct GemmMulti {
   GemmProb<T> p[ 4 ];
   int end[ 4 ];  // running tile counts
   int n;
};

template <typename T, int BK>
__global__ void __launch_bounds__( 256 ) __attribute__( ( amdgpu_waves_per_eu( GemmOcc<T, BK>::waves ) ) ) k_gemm_multi( GemmMulti<T> mp ) {
   __shared__ __attribute__( ( aligned( 16 ) ) ) T smem[ gemm_lds_elems<BK, 64, 64>() ];
   __builtin_amdgcn_s_setprio( 2 );  // the next pair's chain waits for these (as the smaller tiles of k_gemm)
   const int b = blockIdx.x;
   int q = 0;
   while ( q + 1 < mp.n && b >= mp.end[ q ] ) ++q;
   const int t = b - ( q > 0 ? mp.end[ q - 1 ] : 0 ), mt = mp.p[ q ].M / 64;
   gemm_tile<T, BK, 64, 64>( t % mt, t / mt, mp.p[ q ], smem );
}

// ------------------------------------------------------------------ the diagonal block

constexpr int kLd = kNB + 1;  // leading dimension of the diagonal block in LDS

// v (a wave-uniform lane's value) broadcast to the whole wave.
template <typename T>
__device__ inline T bcast( T v, int lane );
template <>
__device__ inline float bcast<float>( float v, int lane ) {
   return __int_as_float( __builtin_amdgcn_readlane( __float_as_int( v ), lane ) );
}
template <>
__device__ inline double bcast<double>( double v, int lane ) {
   const long long x = __double_as_longlong( v );
   const unsigned lo = (unsigned)__builtin_amdgcn_readlane( (int)x, lane );
   const unsigned hi = (unsigned)__builtin_amdgcn_readlane( (int)( x >> 32 ), lane );
   return __longlong_as_double( (long long)( ( (unsigned long long)hi << 32 ) | lo ) );
}

// How a 32 x 32 tile operand is read: whole, as the unit lower triangle of its strict lower part,
// or as its upper triangle (diagonal included).
enum TileOp { kOpFull, kOpLowerUnit, kOpUpper };

template <int OP, typename T>
__device__ __forceinline__ T tile_elem( T v, int r, int c ) {
   if ( OP == kOpLowerUnit ) return r > c ? v : ( r == c ? T( 1 ) : T( 0 ) );
   if ( OP == kOpUpper ) return r <= c ? v : T( 0 );
   return v;
}

// acc (a 32 x 32 tile as 2 x 2 MFMA 16 x 16 blocks, one wave) += sa op(A) op(B) over k = 0..31;
// A and B are 32 x 32 tiles in LDS (leading dimension kLd); sa is +1 or -1.
template <typename T, int OA, int OB>
__device__ __forceinline__ void tile_mma( typename Mfma<T>::v4 ( &acc )[ 2 ][ 2 ], const T* A, const T* B, T sa, int lane ) {
   const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
   for ( int kk = 0; kk < 8; ++kk ) {
      const int k = 4 * kk + lk;
      T a[ 2 ], b[ 2 ];
#pragma unroll
      for ( int h = 0; h < 2; ++h ) {
         const int r = 16 * h + lr;
         a[ h ] = sa * tile_elem<OA>( A[ k * kLd + r ], r, k );
         b[ h ] = tile_elem<OB>( B[ r * kLd + k ], k, r );
      }
#pragma unroll
      for ( int bi = 0; bi < 2; ++bi )
#pragma unroll
         for ( int bj = 0; bj < 2; ++bj ) acc[ bi ][ bj ] = Mfma<T>::run( a[ bi ], b[ bj ], acc[ bi ][ bj ] );
   }
}

// Row of accumulator element r of an MFMA 16 x 16 block held by a lane (column: lane % 16).
template <typename T>
__device__ __forceinline__ int acc_row( int lane, int r ) {
   return sizeof( T ) == 8 ? ( lane >> 4 ) + 4 * r : 4 * ( lane >> 4 ) + r;
}

template <typename T>
__device__ __forceinline__ void tile_zero( typename Mfma<T>::v4 ( &acc )[ 2 ][ 2 ] ) {
#pragma unroll
   for ( int bi = 0; bi < 2; ++bi )
#pragma unroll
      for ( int bj = 0; bj < 2; ++bj )
#pragma unroll
         for ( int r = 0; r < 4; ++r ) acc[ bi ][ bj ][ r ] = T( 0 );
}

template <typename T>
__device__ __forceinline__ void tile_load( typename Mfma<T>::v4 ( &acc )[ 2 ][ 2 ], const T* C, int lane ) {
#pragma unroll
   for ( int bi = 0; bi < 2; ++bi )
#pragma unroll
      for ( int bj = 0; bj < 2; ++bj )
#pragma unroll
         for ( int r = 0; r < 4; ++r )
            acc[ bi ][ bj ][ r ] = C[ ( 16 * bj + ( lane & 15 ) ) * kLd + 16 * bi + acc_row<T>( lane, r ) ];
}

// Stores the tile; returns true if any |entry| > thr (or is not finite) when check is set.
template <typename T>
__device__ __forceinline__ bool tile_store( const typename Mfma<T>::v4 ( &acc )[ 2 ][ 2 ], T* C, int lane, bool check,
                                            T thr = T( 1 ) ) {
   bool big = false;
#pragma unroll
   for ( int bi = 0; bi < 2; ++bi )
#pragma unroll
      for ( int bj = 0; bj < 2; ++bj )
#pragma unroll
         for ( int r = 0; r < 4; ++r ) {
            const T v = acc[ bi ][ bj ][ r ];
            if ( check ) big |= !( tabs( v ) <= thr );
            C[ ( 16 * bj + ( lane & 15 ) ) * kLd + 16 * bi + acc_row<T>( lane, r ) ] = v;
         }
   return big;
}

// v, made opaque to the optimizer (held in a vector register).  The leaf routines below are
// inlined into the diagonal block's loop over its four column blocks; from there the compiler
// would hoist the per-step lane masks (i > j, and the triangle masks of the tile products) out of
// the loop, and they outgrow the scalar registers: spilled to vector-register lanes, every use
// costs readlanes and hazard waits.  An opaque lane index keeps each mask a single compare where
// it is used.
template <typename V>
__device__ __forceinline__ V opaque_v( V v ) {
   asm volatile( "" : "+v"( v ) );
   return v;
}

// 1 / v for the leaf factorization's multipliers: FP32 takes the hardware reciprocal (v_rcp_f32,
// within 1 ulp; the FP32 factors only seed the FP64 refinement, and the correctly rounded quotient
// costs ten dependent instructions per column); FP64 the correctly rounded quotient, as DGETF2
// scales by ONE / A(J,J).
template <typename T>
__device__ inline T leaf_recip( T v );
template <>
__device__ inline float leaf_recip<float>( float v ) {
   return __builtin_amdgcn_rcpf( v );
}
template <>
__device__ inline double leaf_recip<double>( double v ) {
   return 1.0 / v;
}

// One wave factors the 32 x 32 tile held as rows in registers (lane i and i + 32: row i, a[c] =
// entry (i, c)) without exchanges: multipliers a_ij * (1 / a_jj), as DGETF2 scales; pivot-row
// entries come from the pivot lane by readlane, so the 32 steps need no barrier.  Returns true
// if some |a_ij| > thr |a_jj| at the step that eliminates column j (thr = 1: partial pivoting
// would exchange rows).
template <typename T>
__device__ __forceinline__ bool leaf_lu( T ( &a )[ 32 ], int i, T thr ) {
   i = opaque_v( i );
   bool big = false;
#pragma unroll
   for ( int j = 0; j < 32; ++j ) {
      const T piv = bcast( a[ j ], j );
      const T rp = leaf_recip( piv );
      const bool below = i > j;
      big |= below && !( tabs( a[ j ] ) <= thr * tabs( piv ) );
      const T l = below ? a[ j ] * rp : T( 0 );
      // the pivot row first, then the updates: the broadcasts land in distinct scalar registers
      // and issue back to back instead of alternating with the FMAs that read them
      T prow[ 32 ];
#pragma unroll
      for ( int c = j + 1; c < 32; ++c ) prow[ c ] = bcast( a[ c ], j );
#pragma unroll
      for ( int c = j + 1; c < 32; ++c ) a[ c ] = a[ c ] - l * prow[ c ];
      a[ j ] = below ? l : a[ j ];
   }
   return big;
}

// L^-1 of the unit lower triangle in the strict lower part of the rows, in place (the upper part
// is untouched): for k = 0..31, rows i > k get X(i, c) -= l_ik X(k, c) for c < k and X(i, k) = -l_ik.
template <typename T>
__device__ __forceinline__ void leaf_inv_lower( T ( &a )[ 32 ], int i ) {
   i = opaque_v( i );
#pragma unroll
   for ( int k = 0; k < 32; ++k ) {
      const T li = i > k ? a[ k ] : T( 0 );
      T prow[ 32 ];
#pragma unroll
      for ( int c = 0; c < k; ++c ) prow[ c ] = bcast( a[ c ], k );
#pragma unroll
      for ( int c = 0; c < k; ++c ) a[ c ] = a[ c ] - li * prow[ c ];
      a[ k ] = i > k ? -li : a[ k ];
   }
}

// U^-1 of the upper triangle (diagonal included) of the rows, in place (the strict lower part is
// untouched), as U1^-1 D^-1 with U = D U1: each row is scaled by its own 1 / u_ii, the unit upper
// triangle U1 is inverted by rows from the bottom (for k = 31..0, rows i < k get X(i, c) -= u1_ik
// X(k, c) for c > k and X(i, k) = -u1_ik), and column c of the result is scaled by 1 / u_cc.
// dg is this row's diagonal entry u_ii, read by the caller from LDS: selecting it from a[] by the
// lane's row index would make the compiler send a[] through scratch memory.
template <typename T>
__device__ __forceinline__ void leaf_inv_upper( T ( &a )[ 32 ], int i, T dg ) {
   i = opaque_v( i );
   const T rd = T( 1 ) / dg;
#pragma unroll
   for ( int c = 1; c < 32; ++c ) a[ c ] = c > i ? a[ c ] * rd : a[ c ];
#pragma unroll
   for ( int k = 31; k > 0; --k ) {
      const T ui = i < k ? a[ k ] : T( 0 );
      T prow[ 32 ];
#pragma unroll
      for ( int c = k + 1; c < 32; ++c ) prow[ c ] = bcast( a[ c ], k );
#pragma unroll
      for ( int c = k + 1; c < 32; ++c ) a[ c ] = a[ c ] - ui * prow[ c ];
      a[ k ] = i < k ? -ui : a[ k ];
   }
   T rdc[ 32 ];
#pragma unroll
   for ( int c = 0; c < 32; ++c ) rdc[ c ] = bcast( rd, c );
#pragma unroll
   for ( int c = 0; c < 32; ++c ) {
      const T x = i == c ? T( 1 ) : a[ c ];
      a[ c ] = c >= i ? x * rdc[ c ] : a[ c ];
   }
}

// Half of the columns of leaf_inv_lower / leaf_inv_upper, so that two waves share each inverse:
// part 0 owns columns 0-7 and 24-31, part 1 columns 8-23, equal work (column c of L^-1 takes
// 31 - c steps, of U^-1 c - 1).  Each wave holds the whole rows (the multipliers of every step)
// and updates only its own columns; the column sets are compile-time, so no branches remain.
template <int P>
__device__ __forceinline__ constexpr bool leaf_owns( int c ) {
   return P == 0 ? ( c < 8 || c >= 24 ) : ( c >= 8 && c < 24 );
}

template <int P, typename T>
__device__ __forceinline__ void leaf_inv_lower_part( T ( &a )[ 32 ], int i ) {
   i = opaque_v( i );
#pragma unroll
   for ( int k = 0; k < 32; ++k ) {
      const T li = i > k ? a[ k ] : T( 0 );
      T prow[ 32 ];
#pragma unroll
      for ( int c = 0; c < k; ++c )
         if ( leaf_owns<P>( c ) ) prow[ c ] = bcast( a[ c ], k );
#pragma unroll
      for ( int c = 0; c < k; ++c )
         if ( leaf_owns<P>( c ) ) a[ c ] = a[ c ] - li * prow[ c ];
      if ( leaf_owns<P>( k ) ) a[ k ] = i > k ? -li : a[ k ];
   }
}

template <int P, typename T>
__device__ __forceinline__ void leaf_inv_upper_part( T ( &a )[ 32 ], int i, T dg ) {
   i = opaque_v( i );
   const T rd = T( 1 ) / dg;
   // every column is scaled: the multipliers u1_ik of the other part's columns are needed too
#pragma unroll
   for ( int c = 1; c < 32; ++c ) a[ c ] = c > i ? a[ c ] * rd : a[ c ];
#pragma unroll
   for ( int k = 31; k > 0; --k ) {
      const T ui = i < k ? a[ k ] : T( 0 );
      T prow[ 32 ];
#pragma unroll
      for ( int c = k + 1; c < 32; ++c )
         if ( leaf_owns<P>( c ) ) prow[ c ] = bcast( a[ c ], k );
#pragma unroll
      for ( int c = k + 1; c < 32; ++c )
         if ( leaf_owns<P>( c ) ) a[ c ] = a[ c ] - ui * prow[ c ];
      if ( leaf_owns<P>( k ) ) a[ k ] = i < k ? -ui : a[ k ];
   }
   T rdc[ 32 ];
#pragma unroll
   for ( int c = 0; c < 32; ++c )
      if ( leaf_owns<P>( c ) ) rdc[ c ] = bcast( rd, c );
#pragma unroll
   for ( int c = 0; c < 32; ++c ) {
      if ( !leaf_owns<P>( c ) ) continue;
      const T x = i == c ? T( 1 ) : a[ c ];
      a[ c ] = c >= i ? x * rdc[ c ] : a[ c ];
   }
}

// The diagonal block at A (leading dimension lda): mode 0 factors it without exchanges (flagging
// any |a_ij| > |a_jj| at the step that eliminates column j, and any multiplier above 1 in
// magnitude) and then inverts both triangles; mode 1 only inverts an already-factored block.
// inv receives L^-1 (unit lower) and U^-1 (upper) as two dense 128 x 128 column-major matrices.
//
// One workgroup of four waves with the block in LDS as a 4 x 4 grid of 32 x 32 tiles; blocked
// right-looking LU with 32-column steps.  Step p: wave 0 factors the diagonal tile in registers
// (leaf_lu), then wave 0 inverts its L and wave 1 its U in place (the two write disjoint halves
// of the tile); the tiles below become A U_pp^-1 and those right L_pp^-1 A, and the trailing
// tiles get A -= L U, all as 32 x 32 MFMA tile products spread over the waves.  The LU factors
// go to A as they are final.  The inverses of the whole block follow from the tile inverses by
// recursive doubling, inv([A 0; C B]) = [A^-1 0; -B^-1 C A^-1  B^-1] (and the upper analogue),
// in place in LDS: 32 -> 64 (four independent tile tasks), then 64 -> 128.
template <typename T>
__global__ void __launch_bounds__( 256 ) k_lu_diag( int mode, T* A, long lda, T* __restrict__ inv, int* check, T thr ) {
   typedef typename Mfma<T>::v4 v4;
   __shared__ T S[ kNB * kLd ];
   __builtin_amdgcn_s_setprio( 3 );  // the head of the critical path: issue priority on its CU
   const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, i = lane & 31;
   auto tile = [&]( int I, int J ) { return S + 32 * J * kLd + 32 * I; };
   {
      // 16-byte loads, all in flight before the first LDS store
      constexpr int V = 16 / sizeof( T ), NV = kNB * kNB / V / 256;
      typedef T vT __attribute__( ( ext_vector_type( V ) ) );
      vT buf[ NV ];
#pragma unroll
      for ( int q = 0; q < NV; ++q ) {
         const int v = q * 256 + tid, c = v / ( kNB / V ), r = ( v % ( kNB / V ) ) * V;
         buf[ q ] = *reinterpret_cast<const vT*>( A + (long)c * lda + r );
      }
#pragma unroll
      for ( int q = 0; q < NV; ++q ) {
         const int v = q * 256 + tid, c = v / ( kNB / V ), r = ( v % ( kNB / V ) ) * V;
#pragma unroll
         for ( int u = 0; u < V; ++u ) S[ c * kLd + r + u ] = buf[ q ][ u ];
      }
   }
   __syncthreads();
   bool big = false;
   T a[ 32 ];
   if ( mode == 0 ) {
      for ( int p = 0; p < 4; ++p ) {
         // the lane and its row, opaque per block (see opaque_v)
         const int ln = opaque_v( lane ), ri = ln & 31;
         T* Spp = tile( p, p );
         if ( wave == 0 ) {
#pragma unroll
            for ( int c = 0; c < 32; ++c ) a[ c ] = Spp[ c * kLd + ri ];
            big |= leaf_lu( a, ri, thr );
            if ( ln < 32 ) {
#pragma unroll
               for ( int c = 0; c < 32; ++c ) Spp[ c * kLd + ri ] = a[ c ];
            }
         }
         __syncthreads();
         {
            // the factored tile to A before its inverses overwrite it: 16-byte stores over the
            // whole workgroup, beside the row reads below
            constexpr int V = 16 / sizeof( T ), NV = 32 * 32 / V / 256;
            typedef T vT __attribute__( ( ext_vector_type( V ) ) );
            T* Ap = A + (long)( 32 * p ) * lda + 32 * p;
#pragma unroll
            for ( int q = 0; q < NV; ++q ) {
               const int v = q * 256 + tid, c = v / ( 32 / V ), r = ( v % ( 32 / V ) ) * V;
               vT x;
#pragma unroll
               for ( int u = 0; u < V; ++u ) x[ u ] = Spp[ c * kLd + r + u ];
               *reinterpret_cast<vT*>( Ap + (long)c * lda + r ) = x;
            }
         }
         // L^-1 on waves 0 and 1, U^-1 on waves 2 and 3, half of the columns each: every wave reads
         // the factored tile's rows first (wave 0 still holds them), then all write their columns
         if ( wave == 1 ) {
#pragma unroll
            for ( int c = 0; c < 32; ++c ) a[ c ] = Spp[ c * kLd + ri ];
         } else if ( wave >= 2 ) {
#pragma unroll
            for ( int c = 0; c < 32; ++c ) a[ c ] = c >= ri ? Spp[ c * kLd + ri ] : T( 0 );
         }
         const T dg = Spp[ ri * kLd + ri ];
         __syncthreads();
         if ( wave < 2 ) {
            if ( wave == 0 )
               leaf_inv_lower_part<0>( a, ri );
            else
               leaf_inv_lower_part<1>( a, ri );
            if ( ln < 32 ) {
#pragma unroll
               for ( int c = 0; c < 31; ++c )
                  if ( c < ri && ( wave == 0 ? leaf_owns<0>( c ) : leaf_owns<1>( c ) ) ) Spp[ c * kLd + ri ] = a[ c ];
            }
         } else {
            if ( wave == 2 )
               leaf_inv_upper_part<0>( a, ri, dg );
            else
               leaf_inv_upper_part<1>( a, ri, dg );
            if ( ln < 32 ) {
#pragma unroll
               for ( int c = 0; c < 32; ++c )
                  if ( c >= ri && ( wave == 2 ? leaf_owns<0>( c ) : leaf_owns<1>( c ) ) ) Spp[ c * kLd + ri ] = a[ c ];
            }
         }
         __syncthreads();
         const int nr = 3 - p;
         for ( int t = wave; t < 2 * nr; t += 4 ) {
            v4 acc[ 2 ][ 2 ];
            tile_zero<T>( acc );
            if ( t < nr ) {  // L panel: A U_pp^-1 (the multipliers: checked against 1)
               T* C = tile( p + 1 + t, p );
               tile_mma<T, kOpFull, kOpUpper>( acc, C, Spp, T( 1 ), ln );
               big |= tile_store<T>( acc, C, ln, true, thr );
            } else {  // U row: L_pp^-1 A
               T* C = tile( p, p + 1 + t - nr );
               tile_mma<T, kOpLowerUnit, kOpFull>( acc, Spp, C, T( 1 ), ln );
               tile_store<T>( acc, C, ln, false );
            }
         }
         __syncthreads();
         for ( int t = wave; t < nr * nr; t += 4 ) {
            const int I = p + 1 + t % nr, J = p + 1 + t / nr;
            v4 acc[ 2 ][ 2 ];
            tile_load<T>( acc, tile( I, J ), ln );
            tile_mma<T, kOpFull, kOpFull>( acc, tile( I, p ), tile( p, J ), T( -1 ), ln );
            tile_store<T>( acc, tile( I, J ), ln, false );
         }
         __syncthreads();
      }
      // (the diagonal tiles went out as they were factored)
      {
         // the off-diagonal tiles of the factors, 16-byte stores
         constexpr int V = 16 / sizeof( T ), NV = kNB * kNB / V / 256;
         typedef T vT __attribute__( ( ext_vector_type( V ) ) );
#pragma unroll 4
         for ( int q = 0; q < NV; ++q ) {
            const int v = q * 256 + tid, c = v / ( kNB / V ), r = ( v % ( kNB / V ) ) * V;
            if ( ( r >> 5 ) == ( c >> 5 ) ) continue;
            vT x;
#pragma unroll
            for ( int u = 0; u < V; ++u ) x[ u ] = S[ c * kLd + r + u ];
            *reinterpret_cast<vT*>( A + (long)c * lda + r ) = x;
         }
      }
      if ( big ) atomicOr( check, 1 );
   } else {
      T* Spp = tile( wave, wave );
#pragma unroll
      for ( int c = 0; c < 32; ++c ) a[ c ] = Spp[ c * kLd + i ];
      leaf_inv_lower( a, i );
      leaf_inv_upper( a, i, Spp[ i * kLd + i ] );
      if ( lane < 32 ) {
#pragma unroll
         for ( int c = 0; c < 32; ++c ) Spp[ c * kLd + i ] = a[ c ];
      }
   }
   __syncthreads();
   // 32 -> 64: waves 0, 1 the L blocks of tile pairs (0, 1) and (2, 3), waves 2, 3 the U blocks.
   {
      const int a0 = 2 * ( wave & 1 ), b0 = a0 + 1;
      v4 acc[ 2 ][ 2 ];
      if ( wave < 2 ) {
         T* C = tile( b0, a0 );
         tile_zero<T>( acc );
         tile_mma<T, kOpFull, kOpLowerUnit>( acc, C, tile( a0, a0 ), T( 1 ), lane );
         tile_store<T>( acc, C, lane, false );
         tile_zero<T>( acc );
         tile_mma<T, kOpLowerUnit, kOpFull>( acc, tile( b0, b0 ), C, T( -1 ), lane );
         tile_store<T>( acc, C, lane, false );
      } else {
         T* C = tile( a0, b0 );
         tile_zero<T>( acc );
         tile_mma<T, kOpFull, kOpUpper>( acc, C, tile( b0, b0 ), T( 1 ), lane );
         tile_store<T>( acc, C, lane, false );
         tile_zero<T>( acc );
         tile_mma<T, kOpUpper, kOpFull>( acc, tile( a0, a0 ), C, T( -1 ), lane );
         tile_store<T>( acc, C, lane, false );
      }
   }
   __syncthreads();
   // 64 -> 128.  Wave w: L tile (IL, JL) = (2 + w / 2, w % 2) and U tile (IU, JU) = (w / 2, 2 + w % 2).
   {
      const int IL = 2 + ( wave >> 1 ), JL = wave & 1, IU = wave >> 1, JU = 2 + ( wave & 1 );
      v4 accL[ 2 ][ 2 ], accU[ 2 ][ 2 ];
      // T = C A^-1 (L), T = C B^-1 (U)
      tile_zero<T>( accL );
      tile_zero<T>( accU );
      if ( JL == 0 ) {
         tile_mma<T, kOpFull, kOpLowerUnit>( accL, tile( IL, 0 ), tile( 0, 0 ), T( 1 ), lane );
         tile_mma<T, kOpFull, kOpFull>( accL, tile( IL, 1 ), tile( 1, 0 ), T( 1 ), lane );
         tile_mma<T, kOpFull, kOpUpper>( accU, tile( IU, 2 ), tile( 2, 2 ), T( 1 ), lane );
      } else {
         tile_mma<T, kOpFull, kOpLowerUnit>( accL, tile( IL, 1 ), tile( 1, 1 ), T( 1 ), lane );
         tile_mma<T, kOpFull, kOpFull>( accU, tile( IU, 2 ), tile( 2, 3 ), T( 1 ), lane );
         tile_mma<T, kOpFull, kOpUpper>( accU, tile( IU, 3 ), tile( 3, 3 ), T( 1 ), lane );
      }
      __syncthreads();
      tile_store<T>( accL, tile( IL, JL ), lane, false );
      tile_store<T>( accU, tile( IU, JU ), lane, false );
      __syncthreads();
      // X = -B^-1 T (L), X = -A^-1 T (U)
      tile_zero<T>( accL );
      tile_zero<T>( accU );
      if ( IL == 2 ) {
         tile_mma<T, kOpLowerUnit, kOpFull>( accL, tile( 2, 2 ), tile( 2, JL ), T( -1 ), lane );
         tile_mma<T, kOpUpper, kOpFull>( accU, tile( 0, 0 ), tile( 0, JU ), T( -1 ), lane );
         tile_mma<T, kOpFull, kOpFull>( accU, tile( 0, 1 ), tile( 1, JU ), T( -1 ), lane );
      } else {
         tile_mma<T, kOpFull, kOpFull>( accL, tile( 3, 2 ), tile( 2, JL ), T( -1 ), lane );
         tile_mma<T, kOpLowerUnit, kOpFull>( accL, tile( 3, 3 ), tile( 3, JL ), T( -1 ), lane );
         tile_mma<T, kOpUpper, kOpFull>( accU, tile( 1, 1 ), tile( 1, JU ), T( -1 ), lane );
      }
      __syncthreads();
      tile_store<T>( accL, tile( IL, JL ), lane, false );
      tile_store<T>( accU, tile( IU, JU ), lane, false );
   }
   __syncthreads();
   T* Li = inv;
   T* Ui = inv + kNB * kNB;
   constexpr int V = 16 / sizeof( T ), NV = kNB * kNB / V / 256;
   typedef T vT __attribute__( ( ext_vector_type( V ) ) );
#pragma unroll 4
   for ( int q = 0; q < NV; ++q ) {
      const int v = q * 256 + tid, c = v / ( kNB / V ), r = ( v % ( kNB / V ) ) * V;
      vT xl, xu;
#pragma unroll
      for ( int u = 0; u < V; ++u ) {
         const T x = S[ c * kLd + r + u ];
         xl[ u ] = r + u > c ? x : ( r + u == c ? T( 1 ) : T( 0 ) );
         xu[ u ] = r + u <= c ? x : T( 0 );
      }
      *reinterpret_cast<vT*>( Li + c * kNB + r ) = xl;
      *reinterpret_cast<vT*>( Ui + c * kNB + r ) = xu;
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

// Y = beta Y + alpha A X in one launch, no partial sums in memory: A column-major M x K, X and Y
// row-interleaved; M a multiple of 16, K of 64.  Workgroup: 16 rows and all of K; its sixteen
// waves split K, each one 16 x 16 x 4 MFMA tile with every load of its K range in flight at once
// (for K = 1024: 16 per lane), and wave 0 adds the sixteen waves' tiles from LDS in a fixed order
// (eight waves of 32 loads each: the solve phase 0.03 ms longer).  For the diagonal super-block steps (M = K = 1024) this is 64 workgroups and one launch
// where the split-K form needs a product and a reduce launch.  A workgroup's rows are 64 bytes of
// each column, half a 128-byte line: when the row blocks are a multiple of 16, blocks 2j and
// 2j + 1 go to workgroups 8 apart, which the dispatcher places on the same XCD (consecutive
// workgroups go to the 8 XCDs in turn), so both halves of a line meet in one L2.
template <typename T>
__global__ void __launch_bounds__( 1024 ) k_skinny_rows( int K, T alpha, const T* __restrict__ A, long lda,
                                                        const T* __restrict__ X, T beta, T* __restrict__ Y ) {
   using v4 = typename Mfma<T>::v4;
   constexpr int W = 16, PF = 16;
   __shared__ T red[ W - 1 ][ 256 ];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   int rb = blockIdx.x;
   if ( gridDim.x % 16 == 0 ) rb = ( rb & ~15 ) + 2 * ( rb & 7 ) + ( ( rb >> 3 ) & 1 );
   const int row0 = rb * 16;
   const int kw = K / W;
   const int li = lane & 15, lk = lane >> 4;
   const T* a = A + (long)( w * kw + lk ) * lda + row0 + li;
   const T* x = X + (long)( w * kw + lk ) * 16 + li;
   v4 acc = v4{ 0, 0, 0, 0 };
   for ( int k = 0; k < kw; k += 4 * PF ) {
      T fa[ PF ], fb[ PF ];
#pragma unroll
      for ( int u = 0; u < PF; ++u ) {
         const bool in = k + 4 * u < kw;
         fa[ u ] = in ? a[ (long)( k + 4 * u ) * lda ] : T( 0 );
         fb[ u ] = in ? x[ (long)( k + 4 * u ) * 16 ] : T( 0 );
      }
#pragma unroll
      for ( int u = 0; u < PF; ++u ) acc = Mfma<T>::run( fa[ u ], fb[ u ], acc );
   }
   // accumulator element r of this lane: row (FP32) 4 (lane/16) + r or (FP64) lane/16 + 4 r,
   // right-hand side lane % 16
   auto rowof = [&]( int r ) { return sizeof( T ) == 8 ? lk + 4 * r : 4 * lk + r; };
   if ( w > 0 ) {
#pragma unroll
      for ( int r = 0; r < 4; ++r ) red[ w - 1 ][ rowof( r ) * 16 + li ] = acc[ r ];
   }
   __syncthreads();
   if ( w == 0 ) {
#pragma unroll
      for ( int r = 0; r < 4; ++r ) {
         const int e = rowof( r ) * 16 + li;
         const T s = ( ( ( acc[ r ] + red[ 0 ][ e ] ) + ( red[ 1 ][ e ] + red[ 2 ][ e ] ) ) +
                       ( ( red[ 3 ][ e ] + red[ 4 ][ e ] ) + ( red[ 5 ][ e ] + red[ 6 ][ e ] ) ) ) +
                     ( ( ( red[ 7 ][ e ] + red[ 8 ][ e ] ) + ( red[ 9 ][ e ] + red[ 10 ][ e ] ) ) +
                       ( ( red[ 11 ][ e ] + red[ 12 ][ e ] ) + ( red[ 13 ][ e ] + red[ 14 ][ e ] ) ) );
         T* y = Y + (long)row0 * 16 + e;
         *y = beta == T( 0 ) ? alpha * s : beta * *y + alpha * s;
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

// Row interchanges of the factorization applied to a row-interleaved block: in order (P v), or
// undone in reverse order (P^T v, the last step of the transposed solve).
template <typename T>
__global__ void k_apply_ipiv_il( int n, const int* __restrict__ ipiv, T* v, int reverse ) {
   const int r = threadIdx.x;
   if ( r >= 16 ) return;
   for ( int j = 0; j < n; ++j ) {
      const int k = reverse ? n - 1 - j : j;
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
// Each thread carries kResidRows rows (row0 + 256 m) so every broadcast read of x from LDS feeds
// that many loads of A (four: one sweep at 5,120 triangles took 0.33 ms; one, two and eight rows
// per thread made the refinement's two sweeps 0.5, 0.13 and 0.4 ms slower).
constexpr int kResidRows = 4;
template <int RM>
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
   const int row0 = blockIdx.x * 256 * RM + threadIdx.x;
   double acc[ RM ][ 16 ];
#pragma unroll
   for ( int m = 0; m < RM; ++m )
#pragma unroll
      for ( int r = 0; r < 16; ++r ) acc[ m ][ r ] = 0.0;
   bool ok[ RM ];
#pragma unroll
   for ( int m = 0; m < RM; ++m ) ok[ m ] = row0 + 256 * m < npad;
   const double* a = A + (long)c0 * lda + row0;
   // four columns' loads issued before their products (16 loads in flight per lane); the
   // accumulation order over j is unchanged
   constexpr int U = 4;
   for ( int j = 0; j < chunk; j += U ) {
      double v[ U ][ RM ];
#pragma unroll
      for ( int u = 0; u < U; ++u )
#pragma unroll
         for ( int m = 0; m < RM; ++m ) v[ u ][ m ] = ok[ m ] ? a[ (long)( j + u ) * lda + 256 * m ] : 0.0;
#pragma unroll
      for ( int u = 0; u < U; ++u ) {
         const double* xr = xs + ( j + u ) * 16;
#pragma unroll
         for ( int r = 0; r < 16; ++r ) {
            const double xv = xr[ r ];
#pragma unroll
            for ( int m = 0; m < RM; ++m ) acc[ m ][ r ] = acc[ m ][ r ] + v[ u ][ m ] * xv;
         }
      }
   }
   double* out = part + (long)blockIdx.y * 16 * npad;
#pragma unroll
   for ( int m = 0; m < RM; ++m )
#pragma unroll
      for ( int r = 0; r < 16; ++r )
         if ( ok[ m ] && r < nrhs ) out[ (long)r * npad + row0 + 256 * m ] = acc[ m ][ r ];
}

// *addr = max(*addr, v) over a whole wave for non-negative doubles, which order as their bit
// patterns do (every lane must call it): the largest bit pattern by shuffles, then one atomic per
// wave.  A NaN (sign bit clear, as fabs leaves it) orders above every finite value and infinity,
// so it still reaches *addr.
__device__ inline void wave_atomic_max_nonneg( double* addr, double v ) {
   unsigned long long b = (unsigned long long)__double_as_longlong( v );
#pragma unroll
   for ( int o = 32; o > 0; o >>= 1 ) {
      const unsigned long long c = __shfl_xor( b, o );
      b = c > b ? c : b;
   }
   if ( ( threadIdx.x & 63 ) == 0 ) atomicMax( reinterpret_cast<unsigned long long*>( addr ), b );
}

// r = b - sum of the chunks (in chunk order); norms[r] = max |r(:, rhs)| over rows < n3.  One
// thread per row and right-hand side (blockIdx.y), so the chunk sums of all columns run at once.
__global__ void __launch_bounds__( 256 ) k_resid_final( int npad, int n3, int nchunks, int nrhs,
                                                       const double* __restrict__ part, const double* __restrict__ b,
                                                       long ldb, double* __restrict__ r, long ldr, double* norms ) {
   const int row = blockIdx.x * 256 + threadIdx.x;
   const int k = blockIdx.y;
   double v = 0.0;
   if ( row < npad ) {  // no early return: the wave reduction needs every lane
      const double* p = part + (long)k * npad + row;
      const long cs = 16L * npad;
      double s = 0.0;
      int c = 0;
      for ( ; c + 8 <= nchunks; c += 8 ) {  // eight loads in flight, added in chunk order
         double t[ 8 ];
#pragma unroll
         for ( int u = 0; u < 8; ++u ) t[ u ] = p[ ( c + u ) * cs ];
#pragma unroll
         for ( int u = 0; u < 8; ++u ) s = s + t[ u ];
      }
      for ( ; c < nchunks; ++c ) s = s + p[ c * cs ];
      v = row < n3 ? b[ (long)k * ldb + row ] - s : 0.0;
      r[ (long)k * ldr + row ] = v;
   }
   wave_atomic_max_nonneg( &norms[ k ], fabs( v ) );
}

// x += d (d in FP32 from the low-precision solve); norms[0..nrhs) = max|d|, norms[16..) = max|x|.
// One thread per row and right-hand side (blockIdx.y).
__global__ void __launch_bounds__( 256 ) k_ir_update( int n3, int nrhs, double* __restrict__ x, long ldx,
                                                     const float* __restrict__ d, long ldd, double* norms ) {
   const int row = blockIdx.x * 256 + threadIdx.x;
   const int k = blockIdx.y;
   double dv = 0.0, xv = 0.0;
   if ( row < n3 ) {  // no early return: the wave reductions need every lane
      dv = (double)d[ (long)row * 16 + k ];
      xv = x[ (long)k * ldx + row ] + dv;
      x[ (long)k * ldx + row ] = xv;
   }
   wave_atomic_max_nonneg( &norms[ k ], fabs( dv ) );
   wave_atomic_max_nonneg( &norms[ 16 + k ], fabs( xv ) );
}

__global__ void k_d2f( long n, const double* __restrict__ a, float* __restrict__ b ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) b[ i ] = (float)a[ i ];
}


// ------------------------------------------------------------------ host side

// K slice of the FP32 GEMM: 16, 32 or 64 for every tile shape, or 0 for 16 on the 128 x 128 tiles
// of the trailing updates (three workgroups per CU) and 32 on the smaller ones.  Chosen per mesh
// by the series autotuner; every launch made on this host thread uses it.
thread_local int g_gemm_bk_f32 = 0;

// The largest multiplier the exchange-free factorization accepts before the caller reruns it with
// row exchanges: 1 is partial pivoting's own test (the factors are then DGETRF's); the FP32
// factorization that iterative refinement corrects accepts more.  Set per call on this thread.
thread_local double g_growth = 1.0;

template <typename T, int BM, int BN>
void launch_gemm_shape( hipStream_t st, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta, T* C,
                        long ldc, int* check, int batch, int zdiv, BatchStride sa, BatchStride sb, BatchStride sc, int kmode ) {
   dim3 grid( M / BM, N / BN, batch );
   if constexpr ( sizeof( T ) == 8 )
      hipLaunchKernelGGL( ( k_gemm<double, 16, BM, BN> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                          check, T( g_growth ), zdiv, sa, sb, sc, kmode );
   else if ( g_gemm_bk_f32 == 16 || K % 32 != 0 || ( g_gemm_bk_f32 == 0 && BM == 128 && BN == 128 ) )
      hipLaunchKernelGGL( ( k_gemm<float, 16, BM, BN> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                          check, T( g_growth ), zdiv, sa, sb, sc, kmode );
   else if ( g_gemm_bk_f32 == 64 && K % 64 == 0 )
      hipLaunchKernelGGL( ( k_gemm<float, 64, BM, BN> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                          check, T( g_growth ), zdiv, sa, sb, sc, kmode );
   else
      hipLaunchKernelGGL( ( k_gemm<float, 32, BM, BN> ), grid, dim3( 256 ), 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                          check, T( g_growth ), zdiv, sa, sb, sc, kmode );
}

// Tiles of 128 x 128 where they fill the chip; the small products of the factorization's critical
// path (fewer 128 x 128 tiles than CUs) take 64-row and/or 64-column tiles instead, two to four
// times the workgroups for the same work.  A product written in place keeps its whole aliased
// dimension in one tile: over A (C == A) all N columns (N == 128), over B (C == B) all M rows.
template <typename T>
void launch_gemm_batched( hipStream_t st, int M, int N, int K, T alpha, const T* A, long lda, const T* B, long ldb, T beta,
                          T* C, long ldc, int* check, int batch, int zdiv, BatchStride sa, BatchStride sb,
                          BatchStride sc, int kmode = 0 ) {
   if ( M <= 0 || N <= 0 || batch <= 0 ) return;
   const long tiles = (long)( M / 128 ) * ( N / 128 ) * batch;
   constexpr long kSmall = 256;  // CUs
   const int shape = tiles >= kSmall ? 0 : C == A && N == 128 ? 1 : C == B && M == 128 ? 2 : C != A && C != B ? 3 : 0;
   ProfScope prof( sizeof( T ) == 8 ? kKGemm64 : shape == 0 ? kKGemm32T : shape == 3 ? kKGemm32S : kKGemm32,
                   2.0 * M * N * K * batch, st );
   if ( shape == 1 )
      launch_gemm_shape<T, 64, 128>( st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, batch, zdiv, sa, sb, sc, kmode );
   else if ( shape == 2 )
      launch_gemm_shape<T, 128, 64>( st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, batch, zdiv, sa, sb, sc, kmode );
   else if ( shape == 3 )
      launch_gemm_shape<T, 64, 64>( st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, batch, zdiv, sa, sb, sc, kmode );
   else
      launch_gemm_shape<T, 128, 128>( st, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, check, batch, zdiv, sa, sb, sc, kmode );
}

// k_gemm_pair with the K slice launch_gemm_shape would choose for these tile shapes.
template <typename T>
void launch_gemm_pair( hipStream_t st, const GemmProb<T>& p0, const GemmProb<T>& p1 ) {
   if ( p0.N != 128 || p1.M != 128 || p0.M % 64 != 0 || p1.N % 64 != 0 || p0.K != p1.K || p0.K % 64 != 0 )
      throw std::runtime_error( "k_gemm_pair: shapes outside its tiles" );
   const dim3 grid( p0.M / 64 + p1.N / 64 );
   if ( grid.x == 0 ) return;
   if constexpr ( sizeof( T ) == 8 )
      hipLaunchKernelGGL( ( k_gemm_pair<double, 16> ), grid, dim3( 256 ), 0, st, p0, p1 );
   else if ( g_gemm_bk_f32 == 16 )
      hipLaunchKernelGGL( ( k_gemm_pair<float, 16> ), grid, dim3( 256 ), 0, st, p0, p1 );
   else if ( g_gemm_bk_f32 == 64 )
      hipLaunchKernelGGL( ( k_gemm_pair<float, 64> ), grid, dim3( 256 ), 0, st, p0, p1 );
   else
      hipLaunchKernelGGL( ( k_gemm_pair<float, 32> ), grid, dim3( 256 ), 0, st, p0, p1 );
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

// The products p[0 .. n) in one k_gemm_multi launch when each would run on 64 x 64 tiles anyway
// (fewer 128 x 128 tiles than CUs), else one launch_gemm each, in order.
template <typename T>
void launch_gemm_multi( hipStream_t st, const GemmProb<T>* p, int n ) {
   constexpr long kSmall = 256;
   GemmMulti<T> mp {};
   bool one = n >= 1 && n <= 4;
   int tiles = 0;
   for ( int q = 0; q < n && one; ++q ) {
      const GemmProb<T>& g = p[ q ];
      if ( g.M <= 0 || g.N <= 0 ) continue;
      one = g.M % 64 == 0 && g.N % 64 == 0 && g.K % 64 == 0 && g.C != g.A && g.C != g.B && g.check == nullptr &&
            (long)( g.M / 128 ) * ( g.N / 128 ) < kSmall;
      mp.p[ mp.n ] = g;
      tiles += ( g.M / 64 ) * ( g.N / 64 );
      mp.end[ mp.n++ ] = tiles;
   }
   if ( !one ) {
      for ( int q = 0; q < n; ++q )
         launch_gemm<T>( st, p[ q ].M, p[ q ].N, p[ q ].K, p[ q ].alpha, p[ q ].A, p[ q ].lda, p[ q ].B, p[ q ].ldb, p[ q ].beta,
                         p[ q ].C, p[ q ].ldc, p[ q ].check );
      return;
   }
   if ( tiles == 0 ) return;
   double flops = 0.0;
   for ( int q = 0; q < mp.n; ++q ) flops += 2.0 * mp.p[ q ].M * mp.p[ q ].N * mp.p[ q ].K;
   ProfScope prof( sizeof( T ) == 8 ? kKGemm64 : kKGemm32S, flops, st );
   if constexpr ( sizeof( T ) == 8 )
      hipLaunchKernelGGL( ( k_gemm_multi<double, 16> ), dim3( tiles ), dim3( 256 ), 0, st, mp );
   else if ( g_gemm_bk_f32 == 16 )
      hipLaunchKernelGGL( ( k_gemm_multi<float, 16> ), dim3( tiles ), dim3( 256 ), 0, st, mp );
   else if ( g_gemm_bk_f32 == 64 )
      hipLaunchKernelGGL( ( k_gemm_multi<float, 64> ), dim3( tiles ), dim3( 256 ), 0, st, mp );
   else
      hipLaunchKernelGGL( ( k_gemm_multi<float, 32> ), dim3( tiles ), dim3( 256 ), 0, st, mp );
}


// The blocked factorization without exchanges (the fast path).  Block columns are taken in pairs
// (npad is a multiple of 256): the two 128-column blocks of a pair are factored one after the other
// (diagonal block, the panel below as A U^-1, and in between the update of the second block
// column by the first), then the two U row panels are formed and the trailing matrix gets one
// rank-256 update, which halves the passes over the trailing matrix of a rank-128 schedule.  The
// trailing update is split: the next pair of block columns on st, so the next pair's diagonal
// blocks start at once, and the rest on st2, overlapping them.  st2 takes the rest in three
// launches: first the next pair's two block rows (A), then the two block columns after the next
// pair (N), then the remainder (B); the next pair waits for A and N only, so its whole
// factorization runs beside B and st2 goes on to the next update without a gap.  The U row panels
// are off the critical path on st3: U(k0, k1) beside the panel below k0, the row panels right of
// the pair (which wait for part A of the previous pair's update, the rows they read and write)
// beside the second block column's factorization, so st runs only diagonal block, panel, update,
// diagonal block, panel, next-columns update.  check is set when partial pivoting would have exchanged
// rows somewhere; the caller must then rerun lu_pivoted on the original matrix.
// Trailing sizes (rows after the pair) from which part B of the update is deferred every other
// pair (rank-512 updates).
constexpr int kDeferMinRows = 6144;

template <typename T>
void lu_nopiv( const LuStreams& S, int npad, T* A, long lda, T* inv, int* check ) {
   const int nb = kNB, nblk = npad / nb;
   hipStream_t st = S.st, st2 = S.st2, st3 = S.st3;
   auto blk = [&]( int r, int c ) { return A + (long)c * nb * lda + (long)r * nb; };
   auto inv_l = [&]( int k ) { return inv + (long)k * 2 * nb * nb; };
   auto inv_u = [&]( int k ) { return inv + (long)k * 2 * nb * nb + nb * nb; };
   auto diag = [&]( int k ) {
      ProfScope prof( kKDiag, 4.0 / 3.0 * nb * nb * nb, st );  // LU 2/3 nb^3 + two triangular inverses nb^3 / 3 each
      hipLaunchKernelGGL( k_lu_diag<T>, dim3( 1 ), dim3( 256 ), 0, st, 0, blk( k, k ), lda, inv_l( k ), check, T( g_growth ) );
   };
   auto urow = [&]( hipStream_t s, int k, int c0, int ncols ) {  // U row panel: L^-1 A over block columns c0..
      launch_gemm<T>( s, nb, ncols, nb, T( 1 ), inv_l( k ), nb, blk( k, c0 ), lda, T( 0 ), blk( k, c0 ), lda, nullptr );
   };
   // The panel below diagonal block k and the U row panel U(k, c0 .. c0 + ncols) in one launch
   // (both read only block k's inverses and what is already updated).
   auto panel_urow = [&]( int k, int c0, int ncols ) {
      const int m0 = npad - ( k + 1 ) * nb;
      const GemmProb<T> p0 { m0, nb, nb, T( 1 ), blk( k + 1, k ), lda, inv_u( k ), nb, T( 0 ), blk( k + 1, k ), lda, check,
                             T( g_growth ) };
      const GemmProb<T> p1 { nb, ncols, nb, T( 1 ), inv_l( k ), nb, blk( k, c0 ), lda, T( 0 ), blk( k, c0 ), lda, nullptr,
                             T( 0 ) };
      ProfScope prof( sizeof( T ) == 8 ? kKGemm64 : kKGemm32, 2.0 * ( (double)m0 + ncols ) * nb * nb, st );
      launch_gemm_pair<T>( st, p0, p1 );
   };
   auto order = [&]( hipEvent_t e, hipStream_t from, hipStream_t to ) {  // to waits for what from has so far
      BH_HIP( hipEventRecord( e, from ) );
      BH_HIP( hipStreamWaitEvent( to, e, 0 ) );
   };
   bool pending = false;    // an update on st2 (part A) that the next pair's U row panels must wait for
   bool pending_n = false;  // an update on st2 (part N) that the next pair's next-columns update must wait for
   bool used2 = false;      // st2 has work to join at the end
   bool deferred = false;   // the previous pair's part B is still to be applied (with this pair's)
   bool late_done = false;  // S.ev_late recorded
   for ( int k0 = 0; k0 < nblk; k0 += 2 ) {
      const int k1 = k0 + 1;
      if ( k1 >= nblk ) {  // a last single block (npad an odd multiple of 128)
         diag( k0 );
         break;
      }
      const bool last = k1 + 1 == nblk;
      const int c2 = k1 + 1, rest = npad - c2 * nb;
      const int nnext = std::min( 2 * nb, rest );
      if ( S.ev_late && !late_done && npad - k0 * nb < S.late_rest ) {
         BH_HIP( hipEventRecord( S.ev_late, st ) );
         late_done = true;
      }
      diag( k0 );
      if ( !last ) {  // U(k0, c2..) on st3, once part A of the previous update has reached those rows
         order( S.ev[ 0 ], st, st3 );
         if ( pending ) BH_HIP( hipStreamWaitEvent( st3, S.ev_g, 0 ) );
         pending = false;
         urow( st3, k0, c2, rest );
      }
      panel_urow( k0, k1, nb );  // the panel below k0 and U(k0, k1)
      if ( !last ) {
         // rows of block k1 right of the pair: A -= L(k1, k0) U(k0, c2..), on st3 once the panel is done
         order( S.ev[ 2 ], st, st3 );
         launch_gemm<T>( st3, nb, rest, nb, T( -1 ), blk( k1, k0 ), lda, blk( k0, c2 ), lda, T( 1 ), blk( k1, c2 ), lda,
                         nullptr );
         BH_HIP( hipEventRecord( S.ev[ 1 ], st3 ) );
      }
      launch_gemm<T>( st, npad - k1 * nb, nb, nb, T( -1 ), blk( k1, k0 ), lda, blk( k0, k1 ), lda, T( 1 ), blk( k1, k1 ), lda,
                      nullptr );
      diag( k1 );
      if ( last ) break;
      // the panel below k1 and U(k1, next pair), once block row k1 has its update by k0 (st3)
      BH_HIP( hipStreamWaitEvent( st, S.ev[ 1 ], 0 ) );
      panel_urow( k1, c2, nnext );
      // the rank-256 trailing update: L = the two panels below, U = the two row panels
      const bool more = rest > nnext;
      if ( more ) {
         // st2 reads this pair's panels and U row panels, not the next columns.  It forms U(k1, ..)
         // beyond the next pair itself first: st2 never waits for st3, which waits for st2 (part
         // A).  Two forked streams waiting for each other break the graph capture of this schedule
         // (the host process crashed in the capture when st2 waited for st3 here).
         order( S.ev_u, st, st2 );
         urow( st2, k1, c2 + nnext / nb, rest - nnext );
      }
      // the next pair's columns, after the previous update's part N has reached them
      if ( pending_n ) BH_HIP( hipStreamWaitEvent( st, S.ev[ 5 ], 0 ) );
      pending_n = false;
      launch_gemm<T>( st, rest, nnext, 2 * nb, T( -1 ), blk( c2, k0 ), lda, blk( k0, c2 ), lda, T( 1 ), blk( c2, c2 ), lda,
                      nullptr );
      if ( more ) {
         // The rest on st2, in three parts so that the next pair waits only for what it reads:
         // A, the next pair's rows (its U row panels, on st3); N, the columns of the pair after it
         // (its next-columns update, on st); B, the remainder, which overlaps the next pair.
         // While the trailing matrix is large, part B of every other pair is deferred and applied
         // with the next pair's: one rank-512 product, which runs the GEMM closer to its peak than
         // two rank-256 ones.  The deferring pair applies at once, besides A and N, what the next
         // pair's own parts A and N will cover (A2: the rows of the pair after next, right of it;
         // N2: the columns of the pair after that), so that the next pair's A and N stay rank-256
         // and nothing the next two pairs read waits for the deferred product.
         const int cr = c2 + nnext / nb, wr = rest - nnext, nn = std::min( 2 * nb, wr );
         const int cr2 = cr + nn / nb, nn2 = std::min( 2 * nb, wr - nn );
         const bool defer = !deferred && wr - nn - nn2 > 0 && rest >= kDeferMinRows;
         // A and N (and when deferring A2 and N2) write disjoint blocks: one launch
         const GemmProb<T> parts[ 4 ] = {
            { nnext, wr, 2 * nb, T( -1 ), blk( c2, k0 ), lda, blk( k0, cr ), lda, T( 1 ), blk( c2, cr ), lda, nullptr, T( 0 ) },
            { wr, nn, 2 * nb, T( -1 ), blk( cr, k0 ), lda, blk( k0, cr ), lda, T( 1 ), blk( cr, cr ), lda, nullptr, T( 0 ) },
            { nn, wr - nn, 2 * nb, T( -1 ), blk( cr, k0 ), lda, blk( k0, cr2 ), lda, T( 1 ), blk( cr, cr2 ), lda, nullptr, T( 0 ) },
            { wr - nn, nn2, 2 * nb, T( -1 ), blk( cr2, k0 ), lda, blk( k0, cr2 ), lda, T( 1 ), blk( cr2, cr2 ), lda, nullptr,
              T( 0 ) } };
         launch_gemm_multi<T>( st2, parts, defer ? 4 : 2 );
         BH_HIP( hipEventRecord( S.ev_g, st2 ) );
         pending = true;
         BH_HIP( hipEventRecord( S.ev[ 5 ], st2 ) );
         pending_n = true;
         if ( !defer && wr > nn ) {
            const int kf = deferred ? k0 - 2 : k0;  // first block column whose panels this product applies
            launch_gemm<T>( st2, wr, wr - nn, ( c2 - kf ) * nb, T( -1 ), blk( cr, kf ), lda, blk( kf, cr2 ), lda, T( 1 ),
                            blk( cr, cr2 ), lda, nullptr );
         }
         BH_HIP( hipEventRecord( S.ev[ 6 ], st2 ) );
         used2 = true;
         deferred = defer;
      }
   }
   // the caller's later work on st must see all of st2's updates (st3's work is joined into st
   // above)
   if ( used2 ) BH_HIP( hipStreamWaitEvent( st, S.ev[ 6 ], 0 ) );
   if ( S.ev_late && !late_done ) BH_HIP( hipEventRecord( S.ev_late, st ) );
}
template void lu_nopiv<double>( const LuStreams&, int, double*, long, double*, int* );
template void lu_nopiv<float>( const LuStreams&, int, float*, long, float*, int* );

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
      hipLaunchKernelGGL( k_lu_diag<T>, dim3( 1 ), dim3( 256 ), 0, st, 1, Akk, lda, invk, nullptr, T( 1 ) );
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

// The same product as one k_skinny_rows launch (M a multiple of 16, K of 64).
template <typename T>
void launch_skinny_rows( hipStream_t st, int M, int K, T alpha, const T* A, long lda, const T* X, T beta, T* Y ) {
   if ( M <= 0 ) return;
   if ( M % 16 != 0 || K % 64 != 0 ) throw std::runtime_error( "launch_skinny_rows: M must be a multiple of 16, K of 64" );
   ProfScope prof( kKTrsm, 2.0 * M * K * 11, st );
   hipLaunchKernelGGL( k_skinny_rows<T>, dim3( M / 16 ), dim3( 1024 ), 0, st, K, alpha, A, lda, X, beta, Y );
}

// Inverses of the S x S diagonal super-blocks of L (unit lower) and U, from the 128 x 128 block
// inverses by recursive doubling: with A, B the inverses of two adjacent s x s diagonal blocks and
// C the block between them, inv([A 0; C B]) has -B^-1 C A^-1 below (and inv of upper: -A^-1 C B^-1
// right of the diagonal).  Every level is two batched GEMMs per triangle over all pairs of all
// super-blocks; each product skips the K range where its inverse operand is zero (the
// triangles of A^-1 and B^-1).  tmp holds npad * S / 2 elements.
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
      launch_gemm_batched<T>( st, s, s, s, T( 1 ), LU + s, lda, XL, S, T( 0 ), tmp, s, nullptr, batch, zdiv, c_l, a_x, t_str, 1 );
      BatchStride x21 { SS, 2L * s * ( S + 1 ) };
      launch_gemm_batched<T>( st, s, s, s, T( -1 ), XL + (long)s * ( S + 1 ), S, tmp, s, T( 0 ), XL + s, S, nullptr, batch, zdiv,
                              a_x, t_str, x21, 2 );
      // U: T = C B^-1, X12 = -A^-1 T
      BatchStride c_u { (long)S + (long)S * lda, 2L * s + 2L * s * lda };
      launch_gemm_batched<T>( st, s, s, s, T( 1 ), LU + (long)s * lda, lda, XU + (long)s * ( S + 1 ), S, T( 0 ), tmp, s, nullptr,
                              batch, zdiv, c_u, a_x, t_str, 3 );
      launch_gemm_batched<T>( st, s, s, s, T( -1 ), XU, S, tmp, s, T( 0 ), XU + (long)s * S, S, nullptr, batch, zdiv, a_x, t_str,
                              x21, 4 );
   }
}
template void build_super_inverses<double>( hipStream_t, int, int, const double*, long, const double*, double*, double*,
                                            double* );
template void build_super_inverses<float>( hipStream_t, int, int, const float*, long, const float*, float*, float*, float* );

// Solves L U x = P b with the factors and the super-block inverses.  b, y, x are row-interleaved
// npad x 16 blocks: b is consumed (row-exchanged when pivoted, then overwritten by the forward
// updates), y is work space, the solution goes to x.  Two products per super-block and triangle:
// the inverse times the block (one k_skinny_rows launch), then the update of the rest of the
// right-hand side (split-K: a product and a reduce launch).
// piv: 0 no exchanges; 1 P b first (A = P^T L U); -1 the factors are those of A^T = U^T L^T P
// transposed in place (the packed factors and the super-block inverses transposed, XL and XU
// exchanged), so P^T is applied to the solution last.
template <typename T>
void lu_solve( hipStream_t st, int npad, int n3, int S, const T* LU, long lda, const T* XL, const T* XU, const int* ipiv,
               int piv, T* b, T* y, T* x, T* P ) {
   const int nsb = npad / S;
   if ( piv == 1 ) hipLaunchKernelGGL( k_apply_ipiv_il<T>, dim3( 1 ), dim3( 64 ), 0, st, n3, ipiv, b, 0 );
   for ( int q = 0; q < nsb; ++q ) {
      const long q0 = (long)q * S;
      launch_skinny_rows<T>( st, S, S, T( 1 ), XL + q * (long)S * S, S, b + q0 * 16, T( 0 ), y + q0 * 16 );
      launch_skinny<T>( st, npad - (int)q0 - S, S, T( -1 ), LU + q0 * lda + q0 + S, lda, y + q0 * 16, T( 1 ),
                        b + ( q0 + S ) * 16, P );
   }
   for ( int q = nsb - 1; q >= 0; --q ) {
      const long q0 = (long)q * S;
      launch_skinny_rows<T>( st, S, S, T( 1 ), XU + q * (long)S * S, S, y + q0 * 16, T( 0 ), x + q0 * 16 );
      launch_skinny<T>( st, (int)q0, S, T( -1 ), LU + q0 * lda, lda, x + q0 * 16, T( 1 ), y, P );
   }
   if ( piv == -1 ) hipLaunchKernelGGL( k_apply_ipiv_il<T>, dim3( 1 ), dim3( 64 ), 0, st, n3, ipiv, x, 1 );
}
template void lu_solve<double>( hipStream_t, int, int, int, const double*, long, const double*, const double*, const int*,
                                int, double*, double*, double*, double* );
template void lu_solve<float>( hipStream_t, int, int, int, const float*, long, const float*, const float*, const int*, int,
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
   constexpr int RM = kResidRows;
   dim3 grid( ( npad + 256 * RM - 1 ) / ( 256 * RM ), nchunks );
   ProfScope prof( kKResid, 2.0 * npad * npad * nrhs, st );
   hipLaunchKernelGGL( k_resid_part<RM>, grid, dim3( 256 ), chunk * 16 * sizeof( double ), st, npad, chunk, nrhs, A, lda, x,
                       ldx, part );
   hipLaunchKernelGGL( k_resid_final, dim3( ( npad + 255 ) / 256, nrhs ), dim3( 256 ), 0, st, npad, n3, nchunks, nrhs, part,
                       b, ldb, r, ldr, norms );
}

size_t residual_part_elems( int npad ) { return (size_t)( npad / resid_chunk( npad ) ) * 16 * npad; }

void ir_update( hipStream_t st, int n3, int nrhs, double* x, long ldx, const float* d, long ldd, double* norms ) {
   hipLaunchKernelGGL( k_ir_update, dim3( ( n3 + 255 ) / 256, nrhs ), dim3( 256 ), 0, st, n3, nrhs, x, ldx, d, ldd, norms );
}

void convert_d2f( hipStream_t st, long n, const double* a, float* b ) {
   hipLaunchKernelGGL( k_d2f, dim3( 2048 ), dim3( 256 ), 0, st, n, a, b );
}


}  // namespace bhg
