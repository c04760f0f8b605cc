// The H-matrix path on MI355X: the counterpart of the reference's hmatrix library
// (hmatrix/src/cluster.cpp, hmatrix.cpp, gmres.cpp, api.cpp) for meshes whose dense GR does not
// pay (or fit): 245,760 unknowns would need 483 GB in FP64.
//
//   cluster tree    hmatrix/src/cluster.cpp:13-108: median bisection of the incenters along the
//                   longest side of their box, all leaves at one depth (host)
//   block tree      hmatrix/src/hmatrix.cpp:27-54: admissible pairs min(diam) <= eta dist become
//                   low-rank leaves, pairs of leaf clusters dense leaves (host)
//   k_hm_aca        hmatrix/src/hmatrix.cpp:62-251: adaptive cross approximation with partial
//                   pivoting, one workgroup per block, a whole batch of blocks per launch
//   k_hm_trunc      hmatrix/src/mat.cpp:129-160: each approximation recompressed to eps (QR of
//                   both factors, SVD of the small core, singular values above eps s_max kept),
//                   one workgroup per block; k_hm_trunc_apply forms the new factors
//   k_hm_dense      dense leaves (hmatrix/src/kernel.cpp:251-275), one workgroup per block
//   k_hm_vtx, k_hm_rows   the product Y = H X (hmatrix/src/hmatrix.cpp:441-471): V^T X per
//                   low-rank block, then one workgroup per leaf row cluster sums its blocks in a
//                   fixed order, so the product is deterministic
//   GMRES           hmatrix/src/gmres.cpp:26-178: restarted, right-preconditioned, batched over
//                   the right-hand sides; Krylov vectors and products on the device, the small
//                   Hessenberg algebra on the host
//
// Differences from the reference, by design: the QR factors of the recompression come from
// Cholesky factors of the Gram matrices (one parallel pass over the factors; the kept ranks agree
// with the exact blocks' numerical ranks at eps); small admissible blocks and blocks whose
// approximation reaches kcap are stored dense (the reference compresses them by pivoted QR and
// SVD); the deflation is applied as its own
// rank-ndef term of the product instead of being appended to every block; the preconditioner is
// block Jacobi on the dense diagonal leaves (the reference factors a coarse H-LU).  GMRES stops
// at the same relative residual, so the solution agrees to that tolerance.
//
// Compiled with -ffp-contract=off: the kernel entries are the dense path's, bit for bit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <unordered_map>

#include "bh_common.h"
#include "hmatrix.h"
#include "pair_kernel.h"

namespace bhg {

namespace {

constexpr int kWg = 256;        // threads of the per-block kernels
constexpr int kRowsWg = 128;    // threads of the per-leaf-row kernels (a leaf has at most 3 * 32 rows)
constexpr int kMaxLeafRows = kRowsWg;  // rows of a leaf cluster (3 per triangle) at most

double now() {
   return std::chrono::duration<double>( std::chrono::steady_clock::now().time_since_epoch() ).count();
}

// ------------------------------------------------------------------ cluster and block trees (host)
//
// The rules are the reference's (hmatrix/src/cluster.cpp:13-117, hmatrix/src/hmatrix.cpp:27-54):
// a cluster is halved at the median incenter along the longest side of its incenters' box (the
// first half the floor(n/2) smallest by (coordinate, triangle)); every cluster of a level is
// halved, down to the first depth at which no cluster exceeds `leaf` triangles; a pair of clusters
// is admissible when the smaller diagonal of their vertex boxes is at most eta times the gap
// between the boxes.  The layout is this file's own: since every level is split whole, the tree
// is a complete binary tree stored in heap order (level l is clusters 2^l - 1 .. 2^(l+1) - 2, left
// to right; the sons of c are 2c + 1 and 2c + 2; the leaves are the last 2^depth, in tree order),
// one array per field, built one level at a time, and the block tree is enumerated level by level.

struct Tree {
   int n = 0, depth = 0;
   std::vector<int> beg, cnt;     // cluster c holds the triangles perm[beg[c] .. beg[c] + cnt[c])
   std::vector<double> box;       // cluster c: box[6c .. 6c + 2] lowest, [6c + 3 .. 6c + 5] highest vertex coordinates
   std::vector<int> perm, pos;    // tree position -> triangle, triangle -> tree position
   int clusters() const { return ( 2 << depth ) - 1; }
   int nleaves() const { return 1 << depth; }
   int first_leaf() const { return ( 1 << depth ) - 1; }
   bool is_leaf( int c ) const { return c >= first_leaf(); }
   static int level_of( int c ) {
      int l = 0;
      while ( ( 2 << l ) - 1 <= c ) ++l;
      return l;
   }
   // the leaves below c, as leaf indices lo .. lo + count - 1
   void leaf_range( int c, int& lo, int& count ) const {
      const int l = level_of( c );
      count = 1 << ( depth - l );
      lo = ( c - ( ( 1 << l ) - 1 ) ) * count;
   }
   int leaf_cluster( int i ) const { return first_leaf() + i; }
};

// v: v1, v2, v3 as n x 3 column-major arrays one after the other; x: incenters n x 3 column-major.
Tree build_tree( int n, const double* v, const double* x, int leaf ) {
   Tree T;
   T.n = n;
   while ( n > 0 && ( ( n - 1 ) >> T.depth ) + 1 > leaf ) ++T.depth;
   const int nc = T.clusters();
   T.beg.assign( nc, 0 );
   T.cnt.assign( nc, 0 );
   T.box.assign( 6 * (size_t)nc, 0.0 );
   T.perm.resize( n );
   std::iota( T.perm.begin(), T.perm.end(), 0 );
   T.cnt[ 0 ] = n;
   // one level at a time: each cluster's split axis, then one sort of all positions by (cluster,
   // coordinate on its axis, triangle), which orders every cluster's segment in place
   std::vector<int> owner( n, 0 ), axis( 1 << std::max( T.depth - 1, 0 ), 0 );
   for ( int l = 0; l < T.depth; ++l ) {
      const int c0 = ( 1 << l ) - 1, nl = 1 << l;
      for ( int k = 0; k < nl; ++k ) {
         const int c = c0 + k;
         double lo[ 3 ] = { HUGE_VAL, HUGE_VAL, HUGE_VAL }, hi[ 3 ] = { -HUGE_VAL, -HUGE_VAL, -HUGE_VAL };
         for ( int q = T.beg[ c ]; q < T.beg[ c ] + T.cnt[ c ]; ++q ) {
            owner[ q ] = k;
            for ( int d = 0; d < 3; ++d ) {
               const double w = x[ (size_t)d * n + T.perm[ q ] ];
               lo[ d ] = std::min( lo[ d ], w );
               hi[ d ] = std::max( hi[ d ], w );
            }
         }
         int a = 0;
         for ( int d = 1; d < 3; ++d )
            if ( hi[ d ] - lo[ d ] > hi[ a ] - lo[ a ] ) a = d;
         axis[ k ] = a;
      }
      std::vector<int> order( n );
      std::iota( order.begin(), order.end(), 0 );
      std::sort( order.begin(), order.end(), [&]( int p, int q ) {
         if ( owner[ p ] != owner[ q ] ) return owner[ p ] < owner[ q ];
         const int a = axis[ owner[ p ] ], tp = T.perm[ p ], tq = T.perm[ q ];
         const double xp = x[ (size_t)a * n + tp ], xq = x[ (size_t)a * n + tq ];
         return xp != xq ? xp < xq : tp < tq;
      } );
      std::vector<int> next( n );
      for ( int q = 0; q < n; ++q ) next[ q ] = T.perm[ order[ q ] ];
      T.perm.swap( next );
      for ( int k = 0; k < nl; ++k ) {
         const int c = c0 + k, s0 = 2 * c + 1, s1 = 2 * c + 2;
         T.beg[ s0 ] = T.beg[ c ];
         T.cnt[ s0 ] = T.cnt[ c ] / 2;
         T.beg[ s1 ] = T.beg[ c ] + T.cnt[ s0 ];
         T.cnt[ s1 ] = T.cnt[ c ] - T.cnt[ s0 ];
      }
   }
   // vertex boxes: the leaves from their triangles, every other cluster the union of its sons'
   for ( int c = nc - 1; c >= 0; --c ) {
      double* bx = &T.box[ 6 * (size_t)c ];
      for ( int d = 0; d < 3; ++d ) bx[ d ] = HUGE_VAL, bx[ 3 + d ] = -HUGE_VAL;
      if ( T.is_leaf( c ) ) {
         for ( int q = T.beg[ c ]; q < T.beg[ c ] + T.cnt[ c ]; ++q )
            for ( int j = 0; j < 3; ++j )
               for ( int d = 0; d < 3; ++d ) {
                  const double w = v[ (size_t)j * 3 * n + (size_t)d * n + T.perm[ q ] ];
                  bx[ d ] = std::min( bx[ d ], w );
                  bx[ 3 + d ] = std::max( bx[ 3 + d ], w );
               }
      } else {
         for ( int i = 1; i <= 2; ++i ) {
            const double* sb = &T.box[ 6 * (size_t)( 2 * c + i ) ];
            for ( int d = 0; d < 3; ++d ) bx[ d ] = std::min( bx[ d ], sb[ d ] ), bx[ 3 + d ] = std::max( bx[ 3 + d ], sb[ 3 + d ] );
         }
      }
   }
   T.pos.resize( n );
   for ( int q = 0; q < n; ++q ) T.pos[ T.perm[ q ] ] = q;
   return T;
}

// The smaller box diagonal at most eta times the distance between the boxes.
bool admissible_pair( const Tree& T, int t, int s, double eta ) {
   const double* a = &T.box[ 6 * (size_t)t ];
   const double* b = &T.box[ 6 * (size_t)s ];
   double da = 0.0, db = 0.0, gap = 0.0;
   for ( int d = 0; d < 3; ++d ) {
      da += ( a[ 3 + d ] - a[ d ] ) * ( a[ 3 + d ] - a[ d ] );
      db += ( b[ 3 + d ] - b[ d ] ) * ( b[ 3 + d ] - b[ d ] );
      const double g = std::max( 0.0, std::max( a[ d ] - b[ 3 + d ], b[ d ] - a[ 3 + d ] ) );
      gap += g * g;
   }
   return std::sqrt( std::min( da, db ) ) <= eta * std::sqrt( gap );
}

enum BlkType { kDense = 0, kLowRank = 1 };

struct Blk {
   int t, s, type;
   bool small = false;   // admissible but kept dense
   bool capped = false;  // the approximation reached kcap: kept dense
   int rank = 0;
   long nr, nc, r0, c0;  // rows and columns (unknowns, tree order)
};

// The leaf blocks of the block tree, coarsest level first: a pair of clusters of one level is a
// leaf when it is admissible (low-rank, or dense when small) or when both are leaf clusters
// (dense); otherwise its four pairs of sons go to the next level.
std::vector<Blk> make_blocks( const Tree& T, const HmOptions& opt ) {
   std::vector<Blk> out;
   std::vector<std::pair<int, int>> level { { 0, 0 } }, next;
   while ( !level.empty() ) {
      next.clear();
      for ( const auto& ts : level ) {
         const int t = ts.first, s = ts.second;
         Blk b { t, s, kDense, false, false, 0, 3L * T.cnt[ t ], 3L * T.cnt[ s ], 3L * T.beg[ t ], 3L * T.beg[ s ] };
         if ( t != s && admissible_pair( T, t, s, opt.eta ) ) {
            b.small = (long)T.cnt[ t ] * T.cnt[ s ] <= opt.dense_small;
            b.type = b.small ? kDense : kLowRank;
            out.push_back( b );
         } else if ( T.is_leaf( t ) ) {
            out.push_back( b );  // t and s are on one level: both leaves
         } else {
            for ( int i = 1; i <= 2; ++i )
               for ( int j = 1; j <= 2; ++j ) next.emplace_back( 2 * t + i, 2 * s + j );
         }
      }
      level.swap( next );
   }
   return out;
}

// ------------------------------------------------------------------ device helpers

// Wave and workgroup reductions in a fixed order (lane 0's result is broadcast, so every lane
// holds the same bits).
__device__ inline double wave_sum( double v ) {
#pragma unroll
   for ( int o = 32; o > 0; o >>= 1 ) v += __shfl_xor( v, o );
   return __shfl( v, 0 );
}

// Sum over the workgroup (blockDim.x a multiple of 64, at most 1024); red holds 16 doubles.
__device__ inline double wg_sum( double v, double* red ) {
   const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
   v = wave_sum( v );
   __syncthreads();
   if ( ( threadIdx.x & 63 ) == 0 ) red[ w ] = v;
   __syncthreads();
   double t = 0.0;
   for ( int k = 0; k < nw; ++k ) t += red[ k ];
   return t;
}

// Largest |value| over the workgroup, ties to the smallest index; idx < 0 for none.  red holds
// 16 doubles and ired 16 ints.
__device__ inline void wg_argmax( double& val, long& idx, double* red, long* ired ) {
   const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
   for ( int o = 32; o > 0; o >>= 1 ) {
      const double ov = __shfl_xor( val, o );
      const long oi = __shfl_xor( idx, o );
      if ( oi >= 0 && ( idx < 0 || ov > val || ( ov == val && oi < idx ) ) ) val = ov, idx = oi;
   }
   __syncthreads();
   if ( ( threadIdx.x & 63 ) == 0 ) red[ w ] = val, ired[ w ] = idx;
   __syncthreads();
   val = 0.0;
   idx = -1;
   for ( int k = 0; k < nw; ++k ) {
      const long oi = ired[ k ];
      const double ov = red[ k ];
      if ( oi >= 0 && ( idx < 0 || ov > val || ( ov == val && oi < idx ) ) ) val = ov, idx = oi;
   }
}

// The 3x3 block of GR (row a, column b at g[3 a + b]) for the triangles at tree positions pr, pc.
struct KernelData {
   const double* src;    // tree-ordered source records
   const double* xc;     // tree-ordered incenters, 3 x n (SoA)
   const double* selfb;  // tree-ordered self blocks
   int n;
   double cg;
};

__device__ inline void kblock( const KernelData& K, int pr, int pc, double g[ 9 ] ) {
   if ( pr == pc ) {
#pragma unroll
      for ( int k = 0; k < 9; ++k ) g[ k ] = K.selfb[ 9 * (size_t)pr + k ];
      return;
   }
   const double c[ 3 ] = { K.xc[ pr ], K.xc[ K.n + pr ], K.xc[ 2 * K.n + pr ] };
   double g6[ 6 ];
   pair_block( K.src + (size_t)kSrcStride * pc, c, K.cg, g6 );
#pragma unroll
   for ( int a = 0; a < 3; ++a )
#pragma unroll
      for ( int b = 0; b < 3; ++b ) g[ 3 * a + b ] = sym_at( g6, a, b );
}

// Tree-ordered copies of the per-triangle data.
__global__ void k_hm_permute( int n, const int* __restrict__ perm, const double* __restrict__ src,
                              const double* __restrict__ xc, const double* __restrict__ selfb, const double* __restrict__ xmu,
                              const int* __restrict__ comp, double* __restrict__ src_t, double* __restrict__ xc_t,
                              double* __restrict__ selfb_t, double* __restrict__ xmu_t, int* __restrict__ comp_t ) {
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   if ( p >= n ) return;
   const int t = perm[ p ];
   for ( int k = 0; k < kSrcStride; ++k ) src_t[ (size_t)kSrcStride * p + k ] = src[ (size_t)kSrcStride * t + k ];
   for ( int d = 0; d < 3; ++d ) xc_t[ (size_t)d * n + p ] = xc[ (size_t)d * n + t ];
   for ( int k = 0; k < 9; ++k ) selfb_t[ 9 * (size_t)p + k ] = selfb[ 9 * (size_t)t + k ];
   for ( int a = 0; a < 3; ++a ) xmu_t[ 3 * (size_t)p + a ] = xmu[ 3 * (size_t)t + a ];
   comp_t[ p ] = comp[ t ];
}

// Vectors between BEST's order and tree order: out(3 p + a, q) = in(3 perm[p] + a, q) (to_tree)
// or out(3 perm[p] + a, q) = in(3 p + a, q).
__global__ void k_hm_reorder( int n, int p, const int* __restrict__ perm, int to_tree, const double* __restrict__ in, long ldi,
                              double* __restrict__ out, long ldo ) {
   const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( i >= 3L * n * p ) return;
   const int q = (int)( i / ( 3L * n ) );
   const int r = (int)( i % ( 3L * n ) );
   const int pos = r / 3, a = r % 3;
   const long nat = 3L * perm[ pos ] + a;
   if ( to_tree )
      out[ q * ldo + r ] = in[ q * ldi + nat ];
   else
      out[ q * ldo + nat ] = in[ q * ldi + r ];
}

// ------------------------------------------------------------------ build kernels

struct DenseJob {
   int rb, m, cb, nc;  // row and column clusters (tree positions, triangle counts)
   void* d;            // nr x nc column-major (nr = 3 m), E
};

// Dense blocks: one workgroup per block, threads over (row triangle, column triangle) pairs.
template <typename E>
__global__ void __launch_bounds__( kWg ) k_hm_dense( const DenseJob* __restrict__ jobs, KernelData K,
                                                     unsigned long long* __restrict__ calls ) {
   const DenseJob J = jobs[ blockIdx.x ];
   E* D = static_cast<E*>( J.d );
   const long nr = 3L * J.m;
   const long total = (long)J.m * J.nc;
   for ( long e = threadIdx.x; e < total; e += kWg ) {
      const int it = (int)( e % J.m ), jt = (int)( e / J.m );
      double g[ 9 ];
      kblock( K, J.rb + it, J.cb + jt, g );
#pragma unroll
      for ( int b = 0; b < 3; ++b )
#pragma unroll
         for ( int a = 0; a < 3; ++a ) D[ ( 3L * jt + b ) * nr + 3L * it + a ] = (E)g[ 3 * a + b ];
   }
   if ( threadIdx.x == 0 ) atomicAdd( calls, (unsigned long long)total );
}

struct AcaJob {
   int rb, m, cb, nc;  // row and column clusters (tree positions, triangle counts)
   int kcap;
   double* U;          // workspace 3m x kcap, column-major
   double* V;          // workspace 3nc x kcap
   double* R;          // 3 x 3nc (the residual rows of the current row triangle)
   double* u;          // 3m
   double* v;          // 3nc
   int* used;          // m
};

// Adaptive cross approximation with partial pivoting (hmatrix/src/hmatrix.cpp:62-251): each row
// step evaluates the three rows of one row triangle, takes up to three rank-one terms from them,
// each with the column of its pivot; it stops after a row step whose every new term had
// |u| |v| <= eps |S|_F, S the approximation so far.  rank[job] = the rank, or -1 at kcap.
// One workgroup of NT threads per block (512 for the large blocks, whose products with the
// terms so far stream O(k (M + N)) entries per step); those products load 4 terms at a time.
template <int NT>
__global__ void __launch_bounds__( NT ) k_hm_aca( const AcaJob* __restrict__ jobs, KernelData K, double eps,
                                                  int* __restrict__ rank, unsigned long long* __restrict__ calls ) {
   constexpr int NW = NT / 64;
   __shared__ double red[ 16 ];
   __shared__ long ired[ 16 ];
   __shared__ double dots[ 2 * 256 ];
   const AcaJob J = jobs[ blockIdx.x ];
   const int tid = threadIdx.x;
   const int m = J.m, n = J.nc;
   const long M = 3L * m, N = 3L * n;
   double* U = J.U;
   double* V = J.V;
   double* R = J.R;
   double* u = J.u;
   double* v = J.v;
   for ( int it = tid; it < m; it += NT ) J.used[ it ] = 0;
   unsigned long long ncalls = 0;
   int k = 0, zero_rows = 0;
   double norm2 = 0.0, piv_max = 0.0;
   int i = m / 2;
   int result = -2;
   __syncthreads();
   for ( ;; ) {
      // the rows of row triangle i, less the terms so far
      for ( int jt = tid; jt < n; jt += NT ) {
         double g[ 9 ];
         kblock( K, J.rb + i, J.cb + jt, g );
#pragma unroll
         for ( int a = 0; a < 3; ++a )
#pragma unroll
            for ( int b = 0; b < 3; ++b ) R[ a * N + 3 * jt + b ] = g[ 3 * a + b ];
      }
      ncalls += n;
      __syncthreads();
      for ( long j = tid; j < N; j += NT ) {
         double r0 = R[ j ], r1 = R[ N + j ], r2 = R[ 2 * N + j ];
         int l = 0;
         for ( ; l + 4 <= k; l += 4 ) {
            double vv[ 4 ];
#pragma unroll
            for ( int d = 0; d < 4; ++d ) vv[ d ] = V[ ( l + d ) * N + j ];
#pragma unroll
            for ( int d = 0; d < 4; ++d ) {
               const double* ul = U + ( l + d ) * M + 3L * i;
               r0 -= ul[ 0 ] * vv[ d ];
               r1 -= ul[ 1 ] * vv[ d ];
               r2 -= ul[ 2 ] * vv[ d ];
            }
         }
         for ( ; l < k; ++l ) {
            const double vv = V[ l * N + j ];
            const double* ul = U + l * M + 3L * i;
            r0 -= ul[ 0 ] * vv;
            r1 -= ul[ 1 ] * vv;
            r2 -= ul[ 2 ] * vv;
         }
         R[ j ] = r0, R[ N + j ] = r1, R[ 2 * N + j ] = r2;
      }
      if ( tid == 0 ) J.used[ i ] = 1;
      __syncthreads();
      int added = 0;
      bool small = true;
      for ( int step = 0; step < 3 && k < J.kcap; ++step ) {
         // pivot: the largest |R| in row-major order over (a, j), first on ties
         double best = 0.0;
         long bi = -1;
         for ( int a = 0; a < 3; ++a )
            for ( long j = tid; j < N; j += NT ) {
               const double w = fabs( R[ a * N + j ] );
               if ( w > best ) best = w, bi = a * N + j;
            }
         if ( best == 0.0 ) bi = -1;
         wg_argmax( best, bi, red, ired );
         if ( bi < 0 || best <= 1e-15 * piv_max ) break;
         piv_max = fmax( piv_max, best );
         const int ab = (int)( bi / N );
         const long jb = bi % N;
         const int jt = (int)( jb / 3 ), bc = (int)( jb % 3 );
         const double piv = R[ ab * N + jb ];
         // the pivot column, less the terms so far
         for ( int it = tid; it < m; it += NT ) {
            double g[ 9 ];
            kblock( K, J.rb + it, J.cb + jt, g );
#pragma unroll
            for ( int a = 0; a < 3; ++a ) u[ 3L * it + a ] = g[ 3 * a + bc ];
         }
         ncalls += m;
         __syncthreads();
         for ( long r = tid; r < M; r += NT ) {
            double w = u[ r ];
            int l = 0;
            for ( ; l + 4 <= k; l += 4 ) {
               double uu[ 4 ];
#pragma unroll
               for ( int d = 0; d < 4; ++d ) uu[ d ] = U[ ( l + d ) * M + r ];
#pragma unroll
               for ( int d = 0; d < 4; ++d ) w -= V[ ( l + d ) * N + jb ] * uu[ d ];
            }
            for ( ; l < k; ++l ) w -= V[ l * N + jb ] * U[ l * M + r ];
            u[ r ] = w / piv;
         }
         for ( long j = tid; j < N; j += NT ) v[ j ] = R[ ab * N + j ];
         __syncthreads();
         const double ui0 = u[ 3L * i ], ui1 = u[ 3L * i + 1 ], ui2 = u[ 3L * i + 2 ];
         double nu = 0.0, nv = 0.0;
         for ( long j = tid; j < N; j += NT ) {
            const double vj = v[ j ];
            R[ j ] -= ui0 * vj;
            R[ N + j ] -= ui1 * vj;
            R[ 2 * N + j ] -= ui2 * vj;
            nv += vj * vj;
            V[ k * N + j ] = vj;
         }
         for ( long r = tid; r < M; r += NT ) {
            const double ur = u[ r ];
            nu += ur * ur;
            U[ k * M + r ] = ur;
         }
         nu = wg_sum( nu, red );
         nv = wg_sum( nv, red );
         // cross terms: sum_l (U_l . u)(V_l . v), one wave per l
         const int wave = tid >> 6, lane = tid & 63;
         for ( int l = wave; l < k; l += NW ) {
            const double* Ul = U + l * M;
            const double* Vl = V + l * N;
            double du[ 4 ] = { 0, 0, 0, 0 }, dv[ 4 ] = { 0, 0, 0, 0 };
            long r = lane;
            for ( ; r + 192 < M; r += 256 )
#pragma unroll
               for ( int d = 0; d < 4; ++d ) du[ d ] += Ul[ r + 64 * d ] * u[ r + 64 * d ];
            for ( ; r < M; r += 64 ) du[ 0 ] += Ul[ r ] * u[ r ];
            long j = lane;
            for ( ; j + 192 < N; j += 256 )
#pragma unroll
               for ( int d = 0; d < 4; ++d ) dv[ d ] += Vl[ j + 64 * d ] * v[ j + 64 * d ];
            for ( ; j < N; j += 64 ) dv[ 0 ] += Vl[ j ] * v[ j ];
            const double su = wave_sum( ( du[ 0 ] + du[ 1 ] ) + ( du[ 2 ] + du[ 3 ] ) );
            const double sv = wave_sum( ( dv[ 0 ] + dv[ 1 ] ) + ( dv[ 2 ] + dv[ 3 ] ) );
            if ( lane == 0 ) dots[ 2 * l ] = su, dots[ 2 * l + 1 ] = sv;
         }
         __syncthreads();
         double cross = 0.0;
         for ( int l = 0; l < k; ++l ) cross += dots[ 2 * l ] * dots[ 2 * l + 1 ];
         norm2 += 2.0 * cross + nu * nv;
         ++k;
         ++added;
         if ( sqrt( nu * nv ) > eps * sqrt( fmax( norm2, 0.0 ) ) ) small = false;
         __syncthreads();
      }
      zero_rows = added == 0 ? zero_rows + 1 : 0;
      if ( added > 0 && small ) break;
      if ( k >= J.kcap ) {
         result = -1;
         break;
      }
      if ( zero_rows >= 3 ) break;
      // next row triangle: the unused one where the last u is largest
      double bu = 0.0;
      long nx = -1;
      if ( k > 0 ) {
         for ( int it = tid; it < m; it += NT ) {
            if ( J.used[ it ] ) continue;
            const double* ul = U + ( k - 1 ) * M + 3L * it;
            const double w = fmax( fabs( ul[ 0 ] ), fmax( fabs( ul[ 1 ] ), fabs( ul[ 2 ] ) ) );
            if ( w > bu ) bu = w, nx = it;
         }
      }
      if ( bu == 0.0 ) nx = -1;
      wg_argmax( bu, nx, red, ired );
      if ( nx < 0 ) {
         double zero = 0.0;
         long first = -1;
         for ( int it = tid; it < m; it += NT )
            if ( !J.used[ it ] ) {
               first = it;
               break;
            }
         // the smallest unused index: an arg-max of a constant, ties to the smallest index
         wg_argmax( zero, first, red, ired );
         nx = first;
      }
      if ( nx < 0 ) break;
      i = (int)nx;
      __syncthreads();
   }
   if ( tid == 0 ) {
      rank[ blockIdx.x ] = result == -1 ? -1 : k;
      atomicAdd( calls, ncalls );
   }
}

// The first k columns of U (column-major) and of V (stored transposed, the rank index fastest)
// from the work space into the block's storage (E).
template <typename E>
__global__ void k_hm_compact( const AcaJob* __restrict__ jobs, const int* __restrict__ rank, void* const* __restrict__ uout,
                              void* const* __restrict__ vout ) {
   const AcaJob J = jobs[ blockIdx.x ];
   const int k = rank[ blockIdx.x ];
   if ( k <= 0 ) return;
   const long M = 3L * J.m, N = 3L * J.nc;
   E* Uo = static_cast<E*>( uout[ blockIdx.x ] );
   E* Vo = static_cast<E*>( vout[ blockIdx.x ] );
   for ( long e = threadIdx.x; e < M * k; e += blockDim.x ) Uo[ e ] = (E)J.U[ e ];
   // V with the rank index fastest: Vo[c k + l] = V(c, l)
   for ( long e = threadIdx.x; e < N * k; e += blockDim.x ) {
      const long c = e / k, l = e % k;
      Vo[ e ] = (E)J.V[ l * N + c ];
   }
}

// ------------------------------------------------------------------ recompression

// The reference recompresses each cross approximation U V^T (rank k, accurate to aca_factor *
// eps) to eps: U = Qu Ru and V = Qv Rv by Householder QR, the SVD of Ru Rv^T, and the singular
// values above eps * s_max kept (hmatrix/src/mat.cpp:129-160, kept_rank :22-32).  Here the QR
// factors come from the Gram matrices instead (Cholesky QR: one streaming pass over U and V, in
// parallel over the rows, where Householder QR makes k dependent passes): G = U^T U with the
// columns scaled to unit norm, G = R^T R, so U = Q R D with Q = U D^-1 R^-1.  U V^T = Qu M Qv^T
// holds for the computed factors whatever their rounding, M = (Ru Du)(Rv Dv)^T; one-sided Jacobi
// on M's columns gives M Z = B with orthogonal columns (the singular values are their norms), and
// the kept columns give U' = U (Du^-1 Ru^-1 B_r), V' = V (Dv^-1 Rv^-1 Z_r).  A Gram matrix whose
// scaled pivot falls below 1e-12 (columns dependent to 6 digits) is not factored: that block keeps
// its cross approximation unchanged.
//
// The blocks are recompressed in four classes by rank, each with a workgroup and an LDS size of
// its own (kp <= 32: 128 threads, 16 KiB; <= 64: 256, 64 KiB; <= 96: 512, 144 KiB; <= 128: 512,
// 128 KiB with the rotations Z in global memory), so that small blocks run several to a CU.
constexpr int kTrMaxK = 128;  // ranks up to this are recompressed
constexpr int kTrRows = 32;   // rows of U or V staged per step of the Gram products
constexpr int kTrApply = 256; // rows per workgroup of k_hm_trunc_apply

struct TruncJob {
   const double* U;  // M x k, leading dimension M (the cross approximation's work space)
   const double* V;  // N x k, leading dimension N
   int M, N, k, kp;  // kp = k rounded up to a multiple of 4
   double* G;        // 2 kp x kp: the Gram matrices of U and V, then their Cholesky factors
   double* Z;        // kp x kp: the right rotations of the SVD of M (used when they do not fit LDS)
   double* C;        // 2 kp x kp: Cu and Cv, U' = U Cu, V' = V Cv (r columns each)
   double* d;        // 2 kp: the column norms of U and V
};

// Gram matrix A^T A (A: R x k, column-major) into G (kp x kp, both halves), fixed summation order.
// Threads own 4 x 4 tiles of the upper triangle (at most Q per thread); with fewer tiles than
// threads, groups of threads take interleaved rows of each staged chunk and their sums are added
// in group order.
template <int NT, int Q>
__device__ void tr_gram( const double* __restrict__ A, int R, int k, int kp, double* __restrict__ G, double* S ) {
   const int tid = threadIdx.x;
   const int nt = kp / 4, npairs = nt * ( nt + 1 ) / 2;
   const int stride = npairs >= NT ? NT : npairs;
   const int ngrp = npairs >= NT ? 1 : NT / npairs;
   const int grp = tid / stride;
   const bool active = grp < ngrp;
   int ti[ Q ], tj[ Q ];
   bool has[ Q ];
#pragma unroll
   for ( int q = 0; q < Q; ++q ) {
      int p = tid % stride + stride * q;
      has[ q ] = active && p < npairs && ( q == 0 || npairs >= NT );
      int a = 0;
      while ( has[ q ] && p >= nt - a ) p -= nt - a, ++a;
      ti[ q ] = a, tj[ q ] = a + p;
   }
   double acc[ Q ][ 16 ];
#pragma unroll
   for ( int q = 0; q < Q; ++q )
#pragma unroll
      for ( int e = 0; e < 16; ++e ) acc[ q ][ e ] = 0.0;
   for ( int r0 = 0; r0 < R; r0 += kTrRows ) {
      __syncthreads();
      for ( int e = tid; e < kTrRows * kp; e += NT ) {
         const int r = e % kTrRows, c = e / kTrRows;
         S[ r * kp + c ] = ( c < k && r0 + r < R ) ? A[ (long)c * R + r0 + r ] : 0.0;
      }
      __syncthreads();
      if ( !active ) continue;
      for ( int r = grp; r < kTrRows; r += ngrp ) {
         const double* row = S + r * kp;
#pragma unroll
         for ( int q = 0; q < Q; ++q ) {
            if ( !has[ q ] ) continue;
            double a[ 4 ], b[ 4 ];
#pragma unroll
            for ( int x = 0; x < 4; ++x ) a[ x ] = row[ 4 * ti[ q ] + x ], b[ x ] = row[ 4 * tj[ q ] + x ];
#pragma unroll
            for ( int x = 0; x < 4; ++x )
#pragma unroll
               for ( int y = 0; y < 4; ++y ) acc[ q ][ 4 * x + y ] += a[ x ] * b[ y ];
         }
      }
   }
   __syncthreads();
   if ( ngrp > 1 ) {
      // per-group sums through LDS, added in group order
      if ( active )
#pragma unroll
         for ( int e = 0; e < 16; ++e ) S[ ( grp * npairs + tid % stride ) * 16 + e ] = acc[ 0 ][ e ];
      __syncthreads();
      for ( int o = tid; o < npairs * 16; o += NT ) {
         const int p0 = o / 16, e = o % 16;
         double s = 0.0;
         for ( int g = 0; g < ngrp; ++g ) s += S[ ( g * npairs + p0 ) * 16 + e ];
         int p = p0, a = 0;
         while ( p >= nt - a ) p -= nt - a, ++a;
         const int i = 4 * a + e / 4, j = 4 * ( a + p ) + e % 4;
         G[ (long)j * kp + i ] = s;
         G[ (long)i * kp + j ] = s;
      }
   } else {
#pragma unroll
      for ( int q = 0; q < Q; ++q ) {
         if ( !has[ q ] ) continue;
#pragma unroll
         for ( int e = 0; e < 16; ++e ) {
            const int i = 4 * ti[ q ] + e / 4, j = 4 * tj[ q ] + e % 4;
            G[ (long)j * kp + i ] = acc[ q ][ e ];
            G[ (long)i * kp + j ] = acc[ q ][ e ];
         }
      }
   }
   __syncthreads();
}

// Cholesky factor of the column-scaled Gram matrix G (kp x kp, in place: the upper triangle
// becomes R, the rest zero) and the column norms d.  Returns false if a scaled pivot is below
// 1e-12 or a column is zero (the result is then unused).
template <int NT>
__device__ bool tr_chol( double* __restrict__ G, double* __restrict__ d, int k, int kp, double* S, double* flag ) {
   const int tid = threadIdx.x;
   for ( int e = tid; e < kp * kp; e += NT ) S[ e ] = G[ e ];
   __syncthreads();
   for ( int l = tid; l < kp; l += NT ) d[ l ] = l < k ? sqrt( S[ l * kp + l ] ) : 1.0;
   if ( tid == 0 ) flag[ 0 ] = 0.0;
   __syncthreads();
   for ( int e = tid; e < k * k; e += NT ) {
      const int i = e % k, c = e / k;
      const double di = d[ i ], dc = d[ c ];
      if ( !( di > 0.0 ) || !( dc > 0.0 ) || !isfinite( di ) || !isfinite( dc ) ) flag[ 0 ] = 1.0;
      else S[ c * kp + i ] = S[ c * kp + i ] / ( di * dc );
   }
   __syncthreads();
   if ( flag[ 0 ] != 0.0 ) return false;
   for ( int j = 0; j < k; ++j ) {
      const double piv = S[ j * kp + j ];
      if ( !( piv >= 1e-12 ) ) return false;  // uniform: every thread reads the same LDS value
      const double rjj = sqrt( piv );
      __syncthreads();
      for ( int c = j + 1 + tid; c < k; c += NT ) S[ c * kp + j ] = S[ c * kp + j ] / rjj;
      if ( tid == 0 ) S[ j * kp + j ] = rjj;
      __syncthreads();
      const int w = k - j - 1;
      for ( int e = tid; e < w * w; e += NT ) {
         const int i = j + 1 + e % w, c = j + 1 + e / w;
         if ( i <= c ) S[ c * kp + i ] -= S[ i * kp + j ] * S[ c * kp + j ];
      }
      __syncthreads();
   }
   for ( int e = tid; e < kp * kp; e += NT ) {
      const int i = e % kp, c = e / kp;
      G[ e ] = ( i <= c && c < k ) ? S[ e ] : 0.0;
   }
   __syncthreads();
   return true;
}

// One-sided Jacobi on the k columns of M (LDS, column-major, leading dimension kp), the rotations
// accumulated in Z (LDS or global memory): cyclic round-robin pairs, all pairs of a round at once,
// each on a group of L lanes (a power of two, L <= 64, pairs x L <= the workgroup; the group's
// sums are its first lane's, so every lane applies the same rotation), at most 30 sweeps, a
// rotation wherever |m_a . m_b| > 1e-13 |m_a| |m_b|.
template <int NT>
__device__ void tr_jacobi( double* M, double* Z, int k, int kp, double* flag ) {
   const int tid = threadIdx.x;
   const int nk = k + ( k & 1 ), np = nk / 2;
   int L = 64;
   while ( L > 1 && np * L > NT ) L >>= 1;
   const int grp = tid / L, gl = tid % L, nslot = NT / L;
   for ( int sweep = 0; sweep < 30 && nk > 1; ++sweep ) {
      bool rotated = false;
      for ( int rd = 0; rd < nk - 1; ++rd ) {
         for ( int s = grp; s < np; s += nslot ) {
            int a, b;
            if ( s == 0 ) a = rd, b = nk - 1;
            else a = ( rd + s ) % ( nk - 1 ), b = ( rd - s + nk - 1 ) % ( nk - 1 );
            if ( a > b ) {
               const int t = a;
               a = b, b = t;
            }
            if ( b >= k ) continue;  // the whole group: its lanes share the pair
            double* Ma = M + a * kp;
            double* Mb = M + b * kp;
            double al = 0.0, be = 0.0, ga = 0.0;
            for ( int i = gl; i < k; i += L ) {
               const double x = Ma[ i ], y = Mb[ i ];
               al += x * x, be += y * y, ga += x * y;
            }
            for ( int o = L >> 1; o > 0; o >>= 1 ) {
               al += __shfl_xor( al, o );
               be += __shfl_xor( be, o );
               ga += __shfl_xor( ga, o );
            }
            al = __shfl( al, 0, L ), be = __shfl( be, 0, L ), ga = __shfl( ga, 0, L );
            if ( al == 0.0 || be == 0.0 || !( fabs( ga ) > 1e-13 * sqrt( al * be ) ) ) continue;
            const double zeta = ( be - al ) / ( 2.0 * ga );
            const double t = ( zeta >= 0.0 ? 1.0 : -1.0 ) / ( fabs( zeta ) + sqrt( 1.0 + zeta * zeta ) );
            const double c = 1.0 / sqrt( 1.0 + t * t ), sn = c * t;
            double* Za = Z + (long)a * kp;
            double* Zb = Z + (long)b * kp;
            for ( int i = gl; i < k; i += L ) {
               const double x = Ma[ i ], y = Mb[ i ];
               Ma[ i ] = c * x - sn * y;
               Mb[ i ] = sn * x + c * y;
               const double zx = Za[ i ], zy = Zb[ i ];
               Za[ i ] = c * zx - sn * zy;
               Zb[ i ] = sn * zx + c * zy;
            }
            rotated = true;
         }
         __syncthreads();
      }
      if ( tid == 0 ) flag[ 0 ] = 0.0;
      __syncthreads();
      if ( rotated ) flag[ 0 ] = 1.0;
      __syncthreads();
      const bool more = flag[ 0 ] != 0.0;
      __syncthreads();
      if ( !more ) break;
   }
}

// One workgroup per block (ranks with kp <= KP): Gram matrices, Cholesky factors, the SVD of M by
// one-sided Jacobi, the kept rank, and the coefficient matrices Cu, Cv.  rank[j] = the new rank
// (the old one if the block is not recompressed, with Cu = Cv = I), failed[j] = 1 in that case.
template <int NT, int KP>
__global__ void __launch_bounds__( NT ) k_hm_trunc( const TruncJob* __restrict__ jobs, double eps, int* __restrict__ rank,
                                                    int* __restrict__ failed ) {
   constexpr bool kZLds = KP <= 96;
   constexpr int kLds = ( kZLds ? 2 : 1 ) * KP * KP;
   constexpr int kTiles = ( KP / 4 ) * ( KP / 4 + 1 ) / 2;
   static_assert( kTrRows * KP <= kLds && 16 * NT <= kLds, "LDS too small for the Gram products" );
   __shared__ __attribute__( ( aligned( 16 ) ) ) double S[ kLds ];
   __shared__ double sig[ KP ];
   __shared__ int ord[ KP ];
   __shared__ double flag[ 4 ];
   const TruncJob J = jobs[ blockIdx.x ];
   const int tid = threadIdx.x, k = J.k, kp = J.kp;
   double* Gu = J.G;
   double* Gv = J.G + (long)kp * kp;
   double* Cu = J.C;
   double* Cv = J.C + (long)kp * kp;
   tr_gram<NT, ( kTiles + NT - 1 ) / NT>( J.U, J.M, k, kp, Gu, S );
   tr_gram<NT, ( kTiles + NT - 1 ) / NT>( J.V, J.N, k, kp, Gv, S );
   bool ok = tr_chol<NT>( Gu, J.d, k, kp, S, flag );
   if ( ok ) ok = tr_chol<NT>( Gv, J.d + kp, k, kp, S, flag );
   if ( !ok ) {
      for ( int e = tid; e < kp * kp; e += NT ) {
         const double v = ( e % kp == e / kp ) ? 1.0 : 0.0;
         Cu[ e ] = v, Cv[ e ] = v;
      }
      if ( tid == 0 ) rank[ blockIdx.x ] = k, failed[ blockIdx.x ] = 1;
      return;
   }
   const double* du = J.d;
   const double* dv = J.d + kp;
   double* M = S;
   double* Z = kZLds ? S + kp * kp : J.Z;
   // M = (Ru Du)(Rv Dv)^T (column-major); Z = I
   for ( int e = tid; e < kp * kp; e += NT ) {
      const int i = e % kp, j = e / kp;
      double s = 0.0;
      if ( i < k && j < k )
         for ( int l = max( i, j ); l < k; ++l ) s += ( Gu[ (long)l * kp + i ] * du[ l ] ) * ( Gv[ (long)l * kp + j ] * dv[ l ] );
      M[ e ] = s;
      Z[ e ] = ( i == j ) ? 1.0 : 0.0;
   }
   __syncthreads();
   tr_jacobi<NT>( M, Z, k, kp, flag );
   // singular values, their order (decreasing, ties by index) and the kept rank
   for ( int j = tid; j < k; j += NT ) {
      double s = 0.0;
      for ( int i = 0; i < k; ++i ) s += M[ j * kp + i ] * M[ j * kp + i ];
      sig[ j ] = sqrt( s );
   }
   __syncthreads();
   for ( int j = tid; j < k; j += NT ) {
      int pos = 0;
      for ( int i = 0; i < k; ++i ) pos += ( sig[ i ] > sig[ j ] || ( sig[ i ] == sig[ j ] && i < j ) ) ? 1 : 0;
      ord[ pos ] = j;
   }
   __syncthreads();
   const double smax = sig[ ord[ 0 ] ];
   int r = 0;
   if ( smax > 0.0 )
      for ( int j = 0; j < k; ++j ) r += sig[ j ] > eps * smax ? 1 : 0;
   // Cu(:, c) = Du^-1 Ru^-1 M(:, ord[c]) and Cv(:, c) = Dv^-1 Rv^-1 Z(:, ord[c]), c < r: back
   // substitution in place on the kept columns, row i of all 2 r columns at a time
   for ( int i = k - 1; i >= 0; --i ) {
      for ( int c = tid; c < 2 * r; c += NT ) {
         const bool isv = c >= r;
         double* x = ( isv ? Z : M ) + (long)ord[ isv ? c - r : c ] * kp;
         x[ i ] = x[ i ] / ( isv ? Gv : Gu )[ (long)i * kp + i ];
      }
      __syncthreads();
      for ( int e = tid; e < 2 * r * i; e += NT ) {
         const int l = e % i, c = e / i;
         const bool isv = c >= r;
         double* x = ( isv ? Z : M ) + (long)ord[ isv ? c - r : c ] * kp;
         x[ l ] -= ( isv ? Gv : Gu )[ (long)i * kp + l ] * x[ i ];
      }
      __syncthreads();
   }
   for ( int e = tid; e < 2 * r * k; e += NT ) {
      const int i = e % k, c = e / k;
      const bool isv = c >= r;
      const int cc = isv ? c - r : c;
      const double* x = ( isv ? Z : M ) + (long)ord[ cc ] * kp;
      ( isv ? Cv : Cu )[ (long)cc * kp + i ] = x[ i ] / ( isv ? dv : du )[ i ];
   }
   if ( tid == 0 ) rank[ blockIdx.x ] = r, failed[ blockIdx.x ] = 0;
}

// U' = U Cu (column-major, 3m x r) and V' = V Cv (stored with the rank index fastest, or with
// kVcol column-major like U) into the block's storage (E).  One workgroup per (job, side, chunk of
// kTrApply rows), one row per thread, 16 columns of the result per pass with that slice of C
// staged in LDS.
template <typename E, bool kVcol = false>
__global__ void __launch_bounds__( kTrApply ) k_hm_trunc_apply( const int4* __restrict__ tasks, const TruncJob* __restrict__ jobs,
                                                             const int* __restrict__ rank, void* const* __restrict__ uout,
                                                             void* const* __restrict__ vout ) {
   __shared__ double cs[ kTrMaxK * 16 ];
   const int4 T = tasks[ blockIdx.x ];
   const TruncJob J = jobs[ T.x ];
   const int side = T.y, r = rank[ T.x ], k = J.k, kp = J.kp;
   const int R = side ? J.N : J.M;
   const double* A = side ? J.V : J.U;
   const double* C = J.C + ( side ? (long)kp * kp : 0L );
   E* out = static_cast<E*>( side ? vout[ T.x ] : uout[ T.x ] );
   const long row = (long)T.z * kTrApply + threadIdx.x;
   for ( int c0 = 0; c0 < r; c0 += 16 ) {
      const int nc = min( 16, r - c0 );
      __syncthreads();
      for ( int e = threadIdx.x; e < k * 16; e += kTrApply ) {
         const int l = e / 16, c = e % 16;
         cs[ e ] = c < nc ? C[ (long)( c0 + c ) * kp + l ] : 0.0;
      }
      __syncthreads();
      if ( row >= R ) continue;
      double acc[ 16 ];
#pragma unroll
      for ( int c = 0; c < 16; ++c ) acc[ c ] = 0.0;
      for ( int l = 0; l < k; ++l ) {
         const double u = A[ (long)l * R + row ];
#pragma unroll
         for ( int c = 0; c < 16; ++c ) acc[ c ] = fma( u, cs[ l * 16 + c ], acc[ c ] );
      }
      if ( side == 0 || kVcol ) {
         for ( int c = 0; c < nc; ++c ) out[ (long)( c0 + c ) * R + row ] = (E)acc[ c ];
      } else {
         for ( int c = 0; c < nc; ++c ) out[ row * r + c0 + c ] = (E)acc[ c ];
      }
   }
}

// ------------------------------------------------------------------ the product

struct BlkDev {
   int type, rank;
   int nr, nc;
   long c0;
   const void* a;  // dense: D (nr x nc); low-rank: U (nr x rank)
   const void* b;  // low-rank: V (nc x rank), stored with the rank index fastest
   long w;         // low-rank: offset of V^T X in the work array (rank x p)
};

constexpr int kSeg = 1024;  // columns of V per task of k_hm_vtx
constexpr long kAcaBig = 6144;  // blocks with at least this many rows + columns: 512-thread approximation
constexpr long kAcaSmall = 768;  // blocks with fewer rows + columns: one wave each (a 256-thread
                                 // workgroup would leave most lanes idle in their row evaluations)
constexpr int kDotChunkDefl = 8192;  // rows per partial sum of the deflation product
constexpr int kBatch = 8;   // loads in flight per lane in the product kernels

// V_b^T X for every low-rank block.  V is stored with the rank index fastest (V[c * rank + l]),
// so a wave whose lanes are consecutive l (and, for ranks below 64, consecutive columns c) reads
// one contiguous run per step, and the entries of X it needs are staged in LDS.  One wave per
// task (block, l0, segment of kSeg columns); each lane keeps its P sums in registers and writes
// them for k_hm_wsum.
template <typename E, int P>
__global__ void __launch_bounds__( 64 ) k_hm_vtx( const int4* __restrict__ tasks, const BlkDev* __restrict__ blks,
                                                  const double* __restrict__ X, long ldx, double* __restrict__ part ) {
   __shared__ double xs[ 256 * P ];
   const int t = blockIdx.x, lane = threadIdx.x;
   const int4 T = tasks[ t ];
   const BlkDev B = blks[ T.x ];
   // rl rank indices from T.y on; a rank below 64 leaves lanes over: then lane = (column offset
   // cg, rank index lg) with rpad = rl rounded up to a power of two, G = 64 / rpad consecutive
   // columns per step (V is stored column after column, the rank index fastest, so the G columns'
   // entries are one contiguous run)
   const int rl = min( 64, B.rank - T.y );
   int rpad = 1;
   while ( rpad < rl ) rpad <<= 1;
   const int G = 64 / rpad, lg = lane % rpad, cg = lane / rpad;
   const int nj = min( kBatch, 256 / G );  // columns per lane per step
   const bool act = lg < rl;
   const E* v = static_cast<const E*>( B.b ) + T.y + lg;
   const double* x = X + B.c0;
   const int lo = T.z * kSeg, hi = min( B.nc, lo + kSeg );
   double acc[ P ];
#pragma unroll
   for ( int q = 0; q < P; ++q ) acc[ q ] = 0.0;
   for ( int cb = lo; cb < hi; cb += 256 ) {
      const int cn = min( 256, hi - cb ), step = G * nj, cnr = ( cn + step - 1 ) / step * step;
      __syncthreads();
      for ( int i = lane; i < cnr * P; i += 64 ) {
         const int c = i % cnr, q = i / cnr;
         xs[ c * P + q ] = c < cn ? x[ q * ldx + cb + c ] : 0.0;
      }
      __syncthreads();
      for ( int c0 = 0; c0 < cnr; c0 += step ) {
         double vv[ kBatch ];
#pragma unroll
         for ( int j = 0; j < kBatch; ++j ) {
            const int c = c0 + j * G + cg;
            vv[ j ] = ( act && j < nj && c < cn ) ? (double)v[ (long)( cb + c ) * B.rank ] : 0.0;
         }
#pragma unroll
         for ( int j = 0; j < kBatch; ++j ) {
            if ( j >= nj ) break;
            const int c = c0 + j * G + cg;
#pragma unroll
            for ( int q = 0; q < P; ++q ) acc[ q ] = fma( vv[ j ], xs[ c * P + q ], acc[ q ] );
         }
      }
   }
   // the G column groups of each rank index, added in a fixed order
   for ( int o = rpad; o < 64; o <<= 1 )
#pragma unroll
      for ( int q = 0; q < P; ++q ) acc[ q ] += __shfl_xor( acc[ q ], o );
   if ( cg == 0 )
#pragma unroll
      for ( int q = 0; q < P; ++q ) part[ ( (long)t * 64 + lg ) * kMaxRhs + q ] = acc[ q ];
}

// W_b(l, q) = sum over the segments, in order, of the partial sums; one thread per (block, l).
// wt = (block, l, task of the first segment, segments); the segments' tasks are ceil(rank / 64)
// apart.
__global__ void k_hm_wsum( const int4* __restrict__ wt, int nwt, const BlkDev* __restrict__ blks, int p,
                           const double* __restrict__ part, double* __restrict__ W ) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if ( t >= nwt ) return;
   const int4 T = wt[ t ];
   const BlkDev B = blks[ T.x ];
   const int nl0 = ( B.rank + 63 ) / 64, lane = T.y & 63;
   for ( int q = 0; q < p; ++q ) {
      double a = 0.0;
      for ( int g = 0; g < T.w; ++g ) a += part[ ( ( (long)T.z + (long)g * nl0 ) * 64 + lane ) * kMaxRhs + q ];
      W[ B.w + (long)T.y * p + q ] = a;
   }
}

constexpr int kSplit = 8;  // waves per leaf cluster in k_hm_rows, each summing a part of its blocks
constexpr int kRowChunk = 128;  // columns of X (or W) staged at a time in k_hm_rows
constexpr int kRowBatch = 16;   // loads in flight per lane in k_hm_rows

// Partial Y(rows of leaf L) over part s of the blocks covering those rows (the blocks in a fixed
// order, cut into kSplit consecutive parts): sums of D x or U w.  One wave per (leaf, part);
// lane r owns rows r and r + 64 for every right-hand side; kBatch columns of D or U are loaded
// before they are used, and the entries of X or W a block needs are staged in LDS (the same for
// every lane).  k_hm_ysum adds the parts in order.
template <typename E, int P>
__global__ void __launch_bounds__( 64 ) k_hm_rows( const int* __restrict__ leaf_nr, const int* __restrict__ ptr,
                                                   const int* __restrict__ ent_blk, const int* __restrict__ ent_off,
                                                   const BlkDev* __restrict__ blks, const double* __restrict__ X, long ldx,
                                                   const double* __restrict__ W, double* __restrict__ ypart ) {
   __shared__ double st[ kRowChunk * P ];
   const int L = blockIdx.x, part = blockIdx.y, r = threadIdx.x;
   const int nr = leaf_nr[ L ];
   const bool act0 = r < nr, act1 = r + 64 < nr;
   const int e0 = ptr[ L ], cnt = ptr[ L + 1 ] - e0;
   const int elo = e0 + (int)( (long)cnt * part / kSplit ), ehi = e0 + (int)( (long)cnt * ( part + 1 ) / kSplit );
   double acc0[ P ], acc1[ P ];
#pragma unroll
   for ( int q = 0; q < P; ++q ) acc0[ q ] = acc1[ q ] = 0.0;
   for ( int e = elo; e < ehi; ++e ) {
      const BlkDev B = blks[ ent_blk[ e ] ];
      const E* A = static_cast<const E*>( B.a ) + ent_off[ e ] + r;
      const bool dense = B.type == kDense;
      const int ncol = dense ? B.nc : B.rank;
      for ( int cb = 0; cb < ncol; cb += kRowChunk ) {
         const int cn = min( kRowChunk, ncol - cb ), cn8 = ( cn + kRowBatch - 1 ) / kRowBatch * kRowBatch;
         __syncthreads();
         if ( dense ) {
            const double* x = X + B.c0 + cb;
            for ( int i = r; i < cn8 * P; i += 64 ) {
               const int c = i % cn8, q = i / cn8;
               st[ c * P + q ] = c < cn ? x[ q * ldx + c ] : 0.0;
            }
         } else {
            const double* w = W + B.w + (long)cb * P;
            for ( int i = r; i < cn8 * P; i += 64 ) st[ i ] = i < cn * P ? w[ i ] : 0.0;
         }
         __syncthreads();
         for ( int c0 = 0; c0 < cn8; c0 += kRowBatch ) {
            double a0[ kRowBatch ], a1[ kRowBatch ];
#pragma unroll
            for ( int j = 0; j < kRowBatch; ++j ) {
               const bool in = c0 + j < cn;
               const long o = (long)( cb + c0 + j ) * B.nr;
               a0[ j ] = ( act0 && in ) ? (double)A[ o ] : 0.0;
               a1[ j ] = ( act1 && in ) ? (double)A[ o + 64 ] : 0.0;
            }
#pragma unroll
            for ( int j = 0; j < kRowBatch; ++j )
#pragma unroll
               for ( int q = 0; q < P; ++q ) {
                  const double xv = st[ ( c0 + j ) * P + q ];
                  acc0[ q ] = fma( a0[ j ], xv, acc0[ q ] );
                  acc1[ q ] = fma( a1[ j ], xv, acc1[ q ] );
               }
         }
      }
   }
   double* out = ypart + ( (long)L * kSplit + part ) * kRowsWg * kMaxRhs;
#pragma unroll
   for ( int q = 0; q < P; ++q ) {
      if ( act0 ) out[ q * kRowsWg + r ] = acc0[ q ];
      if ( act1 ) out[ q * kRowsWg + r + 64 ] = acc1[ q ];
   }
}

// Y(row, q) = the kSplit parts in order; one thread per (leaf, row slot, q).
__global__ void k_hm_ysum( int nleaves, int p, const int* __restrict__ leaf_r0, const int* __restrict__ leaf_nr,
                           const double* __restrict__ ypart, double* __restrict__ Y, long ldy ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t >= (long)nleaves * p * kRowsWg ) return;
   const int r = (int)( t % kRowsWg ), q = (int)( ( t / kRowsWg ) % p ), L = (int)( t / ( (long)kRowsWg * p ) );
   if ( r >= leaf_nr[ L ] ) return;
   const double* in = ypart + (long)L * kSplit * kRowsWg * kMaxRhs + q * kRowsWg + r;
   double a = 0.0;
#pragma unroll
   for ( int s = 0; s < kSplit; ++s ) a += in[ (long)s * kRowsWg * kMaxRhs ];
   Y[ q * ldy + leaf_r0[ L ] + r ] = a;
}

__global__ void k_hm_defl_final( int count, int nchunk, const double* __restrict__ part, double* __restrict__ out ) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if ( t >= count ) return;
   double a = 0.0;
   for ( int c = 0; c < nchunk; ++c ) a += part[ (long)t * nchunk + c ];
   out[ t ] = a;
}

// Partial sums of xmu x(:, q) over the unknowns of surface k (all unknowns when ndef = 1), one
// workgroup per (chunk of rows, k, q); k_hm_dots_final adds the chunks in order into s[k p + q].
__global__ void __launch_bounds__( kWg ) k_hm_defl_part( int n3, int ndef, int p, const double* __restrict__ xmu,
                                                         const int* __restrict__ comp, const double* __restrict__ X,
                                                         long ldx, double* __restrict__ part ) {
   __shared__ double red[ 16 ];
   const int c = blockIdx.x, nchunk = gridDim.x, kq = blockIdx.y, kk = kq / p, q = kq % p;
   const long lo = (long)c * kDotChunkDefl, hi = min( (long)n3, lo + kDotChunkDefl );
   double a = 0.0;
   for ( long i = lo + threadIdx.x; i < hi; i += kWg )
      if ( ndef == 1 || comp[ i / 3 ] == kk + 1 ) a += xmu[ i ] * X[ q * ldx + i ];
   a = wg_sum( a, red );
   if ( threadIdx.x == 0 ) part[ (long)kq * nchunk + c ] = a;
}

__global__ void k_hm_defl_add( int n3, int ndef, int p, const double* __restrict__ xmu, const int* __restrict__ comp,
                               const double* __restrict__ gdef, const double* __restrict__ s, double* __restrict__ Y,
                               long ldy ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t >= (long)n3 * p ) return;
   const int i = (int)( t % n3 ), q = (int)( t / n3 );
   const int kk = ndef == 1 ? 0 : comp[ i / 3 ] - 1;
   Y[ q * ldy + i ] += gdef[ kk ] * xmu[ i ] * s[ kk * p + q ];
}

// ------------------------------------------------------------------ block Jacobi preconditioner

// The inverse of each diagonal leaf block (GR's entries plus the deflation), by Gauss-Jordan
// elimination with partial pivoting in LDS; one workgroup per leaf.
template <typename E>
__global__ void __launch_bounds__( kWg ) k_hm_diag_inv( const int* __restrict__ diag_blk, const BlkDev* __restrict__ blks,
                                                        const int* __restrict__ leaf_r0, const double* __restrict__ xmu,
                                                        const int* __restrict__ comp, const double* __restrict__ gdef,
                                                        int ndef, const long* __restrict__ pinv_off,
                                                        double* __restrict__ pinv, int* __restrict__ fail ) {
   constexpr int LD = kMaxLeafRows + 1;
   __shared__ double A[ kMaxLeafRows * LD ];
   __shared__ int perm[ kMaxLeafRows ];
   __shared__ double red[ 16 ];
   __shared__ long ired[ 16 ];
   const int L = blockIdx.x, tid = threadIdx.x;
   const BlkDev B = blks[ diag_blk[ L ] ];
   const int n = B.nr;
   const long r0 = leaf_r0[ L ];
   const E* D = static_cast<const E*>( B.a );
   for ( int e = tid; e < n * n; e += kWg ) {
      const int i = e % n, j = e / n;
      double a = (double)D[ (long)j * n + i ];
      const long gi = r0 + i, gj = r0 + j;
      const int ci = comp[ gi / 3 ], cj = comp[ gj / 3 ];
      if ( ndef == 1 || ci == cj ) a += gdef[ ndef == 1 ? 0 : ci - 1 ] * xmu[ gj ] * xmu[ gi ];
      A[ j * LD + i ] = a;
   }
   if ( tid < n ) perm[ tid ] = tid;
   __syncthreads();
   // in-place Gauss-Jordan: column k of the inverse replaces column k of A
   for ( int k = 0; k < n; ++k ) {
      double best = 0.0;
      long bi = -1;
      for ( int i = k + tid; i < n; i += kWg ) {
         const double w = fabs( A[ k * LD + i ] );
         if ( w > best ) best = w, bi = i;
      }
      if ( best == 0.0 ) bi = -1;
      wg_argmax( best, bi, red, ired );
      if ( bi < 0 ) {
         if ( tid == 0 ) atomicOr( fail, 1 );
         return;
      }
      const int pr = (int)bi;
      if ( pr != k ) {
         for ( int j = tid; j < n; j += kWg ) {
            const double t = A[ j * LD + k ];
            A[ j * LD + k ] = A[ j * LD + pr ];
            A[ j * LD + pr ] = t;
         }
         if ( tid == 0 ) {
            const int t = perm[ k ];
            perm[ k ] = perm[ pr ];
            perm[ pr ] = t;
         }
      }
      __syncthreads();
      const double d = 1.0 / A[ k * LD + k ];
      __syncthreads();
      // row k scaled; column k becomes the multipliers' negatives
      for ( int j = tid; j < n; j += kWg ) A[ j * LD + k ] = ( j == k ? 1.0 : A[ j * LD + k ] ) * d;
      __syncthreads();
      for ( int e = tid; e < n * n; e += kWg ) {
         const int i = e % n, j = e / n;
         if ( i == k ) continue;
         const double f = A[ k * LD + i ];  // read before column k of this row is rewritten below
         if ( j != k ) A[ j * LD + i ] -= f * A[ j * LD + k ];
      }
      __syncthreads();
      for ( int i = tid; i < n; i += kWg )
         if ( i != k ) A[ k * LD + i ] = -A[ k * LD + i ] * d;
      __syncthreads();
   }
   // A now holds the inverse of P A0 (rows exchanged as perm says): A0^-1 = (P A0)^-1 P, so
   // column j of A goes to column perm[j].
   double* out = pinv + pinv_off[ L ];
   for ( int e = tid; e < n * n; e += kWg ) {
      const int i = e % n, j = e / n;
      out[ (long)perm[ j ] * n + i ] = A[ j * LD + i ];
   }
}

// Z(rows of L) = Pinv_L V(rows of L); one workgroup per leaf.
__global__ void __launch_bounds__( kRowsWg ) k_hm_precond( const int* __restrict__ leaf_r0, const int* __restrict__ leaf_nr,
                                                           const long* __restrict__ pinv_off, const double* __restrict__ pinv,
                                                           int p, const double* __restrict__ Vin, long ldv,
                                                           double* __restrict__ Z, long ldz ) {
   __shared__ double st[ kMaxLeafRows * kMaxRhs ];
   const int L = blockIdx.x, r = threadIdx.x;
   const int n = leaf_nr[ L ];
   const long r0 = leaf_r0[ L ];
   for ( int t = r; t < n * p; t += kRowsWg ) st[ ( t % n ) * kMaxRhs + t / n ] = Vin[ ( t / n ) * ldv + r0 + t % n ];
   __syncthreads();
   if ( r >= n ) return;
   const double* P = pinv + pinv_off[ L ];
   double acc[ kMaxRhs ];
#pragma unroll
   for ( int q = 0; q < kMaxRhs; ++q ) acc[ q ] = 0.0;
   for ( int c = 0; c < n; ++c ) {
      const double a = P[ (long)c * n + r ];
#pragma unroll
      for ( int q = 0; q < kMaxRhs; ++q )
         if ( q < p ) acc[ q ] += a * st[ c * kMaxRhs + q ];
   }
#pragma unroll
   for ( int q = 0; q < kMaxRhs; ++q )
      if ( q < p ) Z[ q * ldz + r0 + r ] = acc[ q ];
}

// ------------------------------------------------------------------ Krylov vector kernels

constexpr int kDotChunk = 4096;  // rows per partial sum

// part[(i p + q) nchunk + c] = sum over chunk c of Vb_i(:, q) . W(:, q), for i < nv; Vb_i is the
// n3 x p block at Vb + i vstride.
__global__ void __launch_bounds__( kWg ) k_hm_dots( int n3, int p, int nv, const double* __restrict__ Vb, long vstride,
                                                    const double* __restrict__ W, double* __restrict__ part ) {
   __shared__ double red[ 16 ];
   const int c = blockIdx.x, nchunk = gridDim.x;
   const int iq = blockIdx.y, i = iq / p, q = iq % p;
   const long lo = (long)c * kDotChunk, hi = min( (long)n3, lo + kDotChunk );
   const double* v = Vb + i * vstride + (long)q * n3;
   const double* w = W + (long)q * n3;
   double a = 0.0;
   for ( long r = lo + threadIdx.x; r < hi; r += kWg ) a += v[ r ] * w[ r ];
   a = wg_sum( a, red );
   if ( threadIdx.x == 0 ) part[ (long)iq * nchunk + c ] = a;
}

__global__ void k_hm_dots_final( int count, int nchunk, const double* __restrict__ part, double* __restrict__ out ) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if ( t >= count ) return;
   double a = 0.0;
   for ( int c = 0; c < nchunk; ++c ) a += part[ (long)t * nchunk + c ];
   out[ t ] = a;
}

// W(:, q) += sign * sum_i h[i p + q] Vb_i(:, q)
__global__ void k_hm_combine( int n3, int p, int nv, const double* __restrict__ Vb, long vstride, const double* __restrict__ h,
                              double sign, double* __restrict__ W ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t >= (long)n3 * p ) return;
   const int q = (int)( t / n3 );
   double a = W[ t ];
   for ( int i = 0; i < nv; ++i ) a += sign * h[ i * p + q ] * Vb[ i * vstride + t ];
   W[ t ] = a;
}

// Out(:, q) = In(:, q) * s[q] (s[q] = 0 gives zeros)
__global__ void k_hm_scale( int n3, int p, const double* __restrict__ in, const double* __restrict__ s, double* __restrict__ out ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t >= (long)n3 * p ) return;
   out[ t ] = in[ t ] * s[ t / n3 ];
}

// Y = A + sign * B (n3 x p, contiguous)
__global__ void k_hm_axpy( long count, const double* __restrict__ a, double sign, const double* __restrict__ b,
                           double* __restrict__ y ) {
   const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
   if ( t < count ) y[ t ] = a[ t ] + sign * b[ t ];
}

unsigned nblocks( long count, int threads ) { return (unsigned)( ( count + threads - 1 ) / threads ); }

}  // namespace

// ------------------------------------------------------------------ the hierarchical preconditioner
//
// The reference preconditions GMRES with a coarse H-LU: the H-matrix copied at eps_pre with the
// near field compressed too, factored in place by the recursive H-LU (hmatrix/src/api.cpp:502-548,
// hmatrix/src/hmatrix.cpp:516-1038) and applied once per step (lu_solve, :1126-1130).  That
// recursion runs down the diagonal one leaf at a time (4,096 leaves in sequence at 245,760
// unknowns), which on a GPU is a chain of tens of thousands of small dependent launches.  Here
// the factorization is hierarchical in the other direction: every level of the cluster tree is
// one batched step.  With the diagonal leaf blocks D (GR plus the deflation, exact) and, for every
// cluster tau with sons a, b, the two blocks between the sons compressed at eps_pre,
//     A(a, b) ~ U_ab V_ab^T,  A(b, a) ~ U_ba V_ba^T   (plus the deflation's rank-ndef terms),
// the approximation M of A factors as M = D (I + W_1 Z_1^T) ... (I + W_{d} Z_{d}^T), one factor per
// level, and M^-1 b is: the leaf inverses, then from the finest level to the root, at every
// cluster of the level at once,
//     y(tau) -= W^(tau) K_tau^-1 Z_tau^T y(tau),   K_tau = I + Z_tau^T W^(tau)   (2r x 2r),
// W^ being the bases after the finer levels' inverses (Sherman-Morrison-Woodbury, applied to the
// bases once during the factorization).  In the compact layout used here, level l keeps one
// n3 x r_l column block of bases Wc (U_ab on the rows of a, U_ba on the rows of b) and one of Zc
// (V_ba on the rows of a, V_ab on the rows of b); then Z_tau^T y = (Zc(b)^T y(b), Zc(a)^T y(a)) and
// K_tau = [[I, Zc(b)^T W^c(b)], [Zc(a)^T W^c(a), I]].  The pieces are batched kernels: products
// Zc(c)^T B over the rows of every son c of a level (k_hd_segtn, k_hd_segsum, fixed summation
// order), the small systems K_tau (k_hd_kinv: through their r x r Schur complement, inverted by
// Gauss-Jordan with partial pivoting in LDS, its pivots monitored as the reference's getrf_nopiv
// monitors its own), u = K^-1 (..) (k_hd_node), and y(c) -= W^c(c) u (k_hd_update).

constexpr int kHdRows = 1024;  // rows per chunk of k_hd_segtn
constexpr int kHdCols = 32;    // columns per tile of k_hd_segtn and k_hd_update
constexpr int kHdMaxR = 128;   // columns of a level's bases at most (compressed rank + deflation terms)
constexpr int kHdUpd = 256;    // rows per workgroup of k_hd_update
constexpr int kHdTile = 256;   // columns per pass of a level's correction

// Partial products Zc(rows, 0:r)^T B(rows, j0 : j0 + kHdCols) over chunk k of kHdRows rows of son
// ci of a level (heap cluster first + ci): grid.x = nchild * nchunk, grid.y = column tiles.
// part[(x * q + j) * r + i] for j < q (q columns of B in all).  Rows are staged 32 at a time.
__global__ void __launch_bounds__( 256 ) k_hd_segtn( int first, int nchunk, const int* __restrict__ cl_r0,
                                                     const int* __restrict__ cl_nr, const double* __restrict__ Zc, long ldz,
                                                     int r, const double* __restrict__ B, long ldb, int q,
                                                     double* __restrict__ part ) {
   constexpr int kOut = ( kHdMaxR * kHdCols + 255 ) / 256;
   __shared__ double zs[ kHdMaxR * 32 ];
   __shared__ double bs[ kHdCols * 32 ];
   const int tid = threadIdx.x;
   const int ci = blockIdx.x / nchunk, k = blockIdx.x % nchunk;
   const int c = first + ci;
   const long r0 = (long)cl_r0[ c ] + (long)k * kHdRows;
   const int nrow = min( kHdRows, cl_nr[ c ] - k * kHdRows );
   const int j0 = blockIdx.y * kHdCols, qt = min( kHdCols, q - j0 );
   double acc[ kOut ];
#pragma unroll
   for ( int t = 0; t < kOut; ++t ) acc[ t ] = 0.0;
   for ( int rr = 0; rr < nrow; rr += 32 ) {
      const int nb = min( 32, nrow - rr );
      __syncthreads();
      for ( int e = tid; e < 32 * r; e += 256 ) {
         const int row = e & 31, i = e >> 5;
         zs[ i * 32 + row ] = row < nb ? Zc[ (long)i * ldz + r0 + rr + row ] : 0.0;
      }
      for ( int e = tid; e < 32 * kHdCols; e += 256 ) {
         const int row = e & 31, j = e >> 5;
         bs[ j * 32 + row ] = ( row < nb && j < qt ) ? B[ (long)( j0 + j ) * ldb + r0 + rr + row ] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for ( int t = 0; t < kOut; ++t ) {
         const int o = tid + 256 * t;
         if ( o >= r * qt ) break;
         const int i = o % r, j = o / r;
         double s = acc[ t ];
         for ( int row = 0; row < 32; ++row ) s += zs[ i * 32 + row ] * bs[ j * 32 + row ];
         acc[ t ] = s;
      }
   }
#pragma unroll
   for ( int t = 0; t < kOut; ++t ) {
      const int o = tid + 256 * t;
      if ( o >= r * qt ) break;
      const int i = o % r, j = o / r;
      part[ ( (long)blockIdx.x * q + j0 + j ) * r + i ] = acc[ t ];
   }
}

// S[ci] (r x q, column-major) = the son's chunk partials added in order.
__global__ void k_hd_segsum( int nchild, int nchunk, long rq, const double* __restrict__ part, double* __restrict__ S ) {
   const long total = (long)nchild * rq;
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x ) {
      const long ci = e / rq, o = e % rq;
      double s = 0.0;
      for ( int k = 0; k < nchunk; ++k ) s += part[ ( ci * nchunk + k ) * rq + o ];
      S[ e ] = s;
   }
}

// The cluster systems of a level: K = [[I, P_b], [P_a, I]] with P_a = S[2 nd], P_b = S[2 nd + 1]
// (r x r, from k_hd_segsum with B = the level's bases).  K u = t is solved through the Schur
// complement C = I - P_a P_b: u_2 = C^-1 (t_2 - P_a t_1), u_1 = t_1 - P_b u_2.  Per cluster nd
// this stores P_a, P_b and C^-1 (3 r^2 at Kst + 3 nd r^2), C inverted by Gauss-Jordan with partial
// pivoting in LDS; piv[2 nd], piv[2 nd + 1] = its smallest and largest |pivot|; a zero or
// non-finite pivot sets *fail.
__global__ void __launch_bounds__( 256 ) k_hd_kinv( int r, const double* __restrict__ S, double* __restrict__ Kst,
                                                    double* __restrict__ piv, int* __restrict__ fail ) {
   constexpr int LD = kHdMaxR + 1;
   __shared__ double A[ kHdMaxR * LD ];
   __shared__ int perm[ kHdMaxR ];
   __shared__ double red[ 16 ];
   __shared__ long ired[ 16 ];
   const int nd = blockIdx.x, tid = threadIdx.x, n = r;
   const long rr = (long)r * r;
   const double* Pa = S + 2 * nd * rr;
   const double* Pb = S + ( 2 * nd + 1 ) * rr;
   double* out = Kst + 3 * nd * rr;
   for ( long e = tid; e < rr; e += 256 ) out[ e ] = Pa[ e ], out[ rr + e ] = Pb[ e ];
   // C = I - P_a P_b
   for ( int e = tid; e < n * n; e += 256 ) {
      const int i = e % n, j = e / n;
      double s = 0.0;
      for ( int k = 0; k < r; ++k ) s += Pa[ (long)k * r + i ] * Pb[ (long)j * r + k ];
      A[ j * LD + i ] = ( i == j ? 1.0 : 0.0 ) - s;
   }
   if ( tid < n ) perm[ tid ] = tid;
   __syncthreads();
   double pmin = HUGE_VAL, pmax = 0.0;
   for ( int k = 0; k < n; ++k ) {
      double best = 0.0;
      long bi = -1;
      for ( int i = k + tid; i < n; i += 256 ) {
         const double w = fabs( A[ k * LD + i ] );
         if ( w > best ) best = w, bi = i;
      }
      if ( best == 0.0 ) bi = -1;
      wg_argmax( best, bi, red, ired );
      if ( bi < 0 || !isfinite( best ) ) {
         if ( tid == 0 ) atomicOr( fail, 1 );
         return;
      }
      pmin = fmin( pmin, best ), pmax = fmax( pmax, best );
      const int pr = (int)bi;
      if ( pr != k ) {
         for ( int j = tid; j < n; j += 256 ) {
            const double t = A[ j * LD + k ];
            A[ j * LD + k ] = A[ j * LD + pr ];
            A[ j * LD + pr ] = t;
         }
         if ( tid == 0 ) {
            const int t = perm[ k ];
            perm[ k ] = perm[ pr ];
            perm[ pr ] = t;
         }
      }
      __syncthreads();
      const double d = 1.0 / A[ k * LD + k ];
      __syncthreads();
      for ( int j = tid; j < n; j += 256 ) A[ j * LD + k ] = ( j == k ? 1.0 : A[ j * LD + k ] ) * d;
      __syncthreads();
      for ( int e = tid; e < n * n; e += 256 ) {
         const int i = e % n, j = e / n;
         if ( i == k || j == k ) continue;
         A[ j * LD + i ] -= A[ k * LD + i ] * A[ j * LD + k ];
      }
      __syncthreads();
      for ( int i = tid; i < n; i += 256 )
         if ( i != k ) A[ k * LD + i ] = -A[ k * LD + i ] * d;
      __syncthreads();
   }
   for ( int e = tid; e < n * n; e += 256 ) {
      const int i = e % n, j = e / n;
      out[ 2 * rr + (long)perm[ j ] * n + i ] = A[ j * LD + i ];
   }
   if ( tid == 0 ) piv[ 2 * nd ] = pmin, piv[ 2 * nd + 1 ] = pmax;
}

// u[nd] (2r x q) = K_nd^-1 (t_1; t_2) with t_1 = S_b = S[2 nd + 1], t_2 = S_a = S[2 nd] (r x q
// each): w = t_2 - P_a t_1, u_2 = C^-1 w, u_1 = t_1 - P_b u_2.  One workgroup per (cluster,
// column tile of kHdCols); t_1, w and u_2 in LDS.
__global__ void __launch_bounds__( 256 ) k_hd_node( int r, int q, const double* __restrict__ Kst, const double* __restrict__ S,
                                                    double* __restrict__ u ) {
   __shared__ double t1[ kHdMaxR * kHdCols ];
   __shared__ double w[ kHdMaxR * kHdCols ];
   __shared__ double u2[ kHdMaxR * kHdCols ];
   const int nd = blockIdx.x, tid = threadIdx.x;
   const int j0 = blockIdx.y * kHdCols, qt = min( kHdCols, q - j0 );
   const long rr = (long)r * r, n2 = 2L * r;
   const double* Pa = Kst + 3 * nd * rr;
   const double* Pb = Pa + rr;
   const double* Ci = Pa + 2 * rr;
   const double* Sb = S + ( 2 * nd + 1 ) * (long)r * q;
   const double* Sa = S + ( 2 * nd ) * (long)r * q;
   double* un = u + (long)nd * n2 * q;
   for ( int e = tid; e < r * qt; e += 256 ) {
      const int i = e % r, j = e / r;
      t1[ j * kHdMaxR + i ] = Sb[ (long)( j0 + j ) * r + i ];
   }
   __syncthreads();
   for ( int e = tid; e < r * qt; e += 256 ) {  // w = t_2 - P_a t_1
      const int i = e % r, j = e / r;
      double s = Sa[ (long)( j0 + j ) * r + i ];
      for ( int k = 0; k < r; ++k ) s -= Pa[ (long)k * r + i ] * t1[ j * kHdMaxR + k ];
      w[ j * kHdMaxR + i ] = s;
   }
   __syncthreads();
   for ( int e = tid; e < r * qt; e += 256 ) {  // u_2 = C^-1 w
      const int i = e % r, j = e / r;
      double s = 0.0;
      for ( int k = 0; k < r; ++k ) s += Ci[ (long)k * r + i ] * w[ j * kHdMaxR + k ];
      u2[ j * kHdMaxR + i ] = s;
      un[ (long)( j0 + j ) * n2 + r + i ] = s;
   }
   __syncthreads();
   for ( int e = tid; e < r * qt; e += 256 ) {  // u_1 = t_1 - P_b u_2
      const int i = e % r, j = e / r;
      double s = t1[ j * kHdMaxR + i ];
      for ( int k = 0; k < r; ++k ) s -= Pb[ (long)k * r + i ] * u2[ j * kHdMaxR + k ];
      un[ (long)( j0 + j ) * n2 + i ] = s;
   }
}

// Y(rows of son ci, j0 : j0 + kHdCols) -= Wc(rows, 0:r) u[ci / 2](slot r : slot r + r, ...), slot
// = ci & 1 (the first son takes the first r rows of u).  grid.x = nchild * nchunk (kHdUpd rows per
// workgroup, one per thread), grid.y = column tiles.
__global__ void __launch_bounds__( kHdUpd ) k_hd_update( int first, int nchunk, const int* __restrict__ cl_r0,
                                                         const int* __restrict__ cl_nr, const double* __restrict__ Wc, long ldw,
                                                         int r, const double* __restrict__ u, int q, double* __restrict__ Y,
                                                         long ldy ) {
   __shared__ double us[ kHdMaxR * kHdCols ];
   const int ci = blockIdx.x / nchunk, k = blockIdx.x % nchunk;
   const int c = first + ci, nd = ci >> 1, slot = ci & 1;
   const int j0 = blockIdx.y * kHdCols, qt = min( kHdCols, q - j0 );
   const long n2 = 2L * r;
   for ( int e = threadIdx.x; e < r * kHdCols; e += kHdUpd ) {
      const int kk = e % r, j = e / r;
      us[ kk * kHdCols + j ] = j < qt ? u[ (long)nd * n2 * q + (long)( j0 + j ) * n2 + slot * r + kk ] : 0.0;
   }
   __syncthreads();
   const int lr = k * kHdUpd + threadIdx.x;
   if ( lr >= cl_nr[ c ] ) return;
   const long row = (long)cl_r0[ c ] + lr;
   double acc[ kHdCols ];
#pragma unroll
   for ( int j = 0; j < kHdCols; ++j ) acc[ j ] = 0.0;
   for ( int kk = 0; kk < r; ++kk ) {
      const double w = Wc[ (long)kk * ldw + row ];
#pragma unroll
      for ( int j = 0; j < kHdCols; ++j ) acc[ j ] += w * us[ kk * kHdCols + j ];
   }
   for ( int j = 0; j < qt; ++j ) Y[ (long)( j0 + j ) * ldy + row ] -= acc[ j ];
}

// One compressed block's factors into a level's bases: Wc(rows of t, col0 + 0:rank) = U and
// Zc(rows of s, col0 + 0:rank) = V (both column-major, leading dimensions 3 cnt).
struct HdScatter {
   const double* U;
   const double* V;
   int rank, tr0, tnr, sr0, snr;
};
__global__ void k_hd_scatter( const HdScatter* __restrict__ jobs, double* __restrict__ Wc, double* __restrict__ Zc, long ld ) {
   const HdScatter J = jobs[ blockIdx.x ];
   for ( long e = threadIdx.x; e < (long)J.rank * J.tnr; e += blockDim.x ) {
      const long l = e / J.tnr, i = e % J.tnr;
      Wc[ l * ld + J.tr0 + i ] = J.U[ l * J.tnr + i ];
   }
   for ( long e = threadIdx.x; e < (long)J.rank * J.snr; e += blockDim.x ) {
      const long l = e / J.snr, i = e % J.snr;
      Zc[ l * ld + J.sr0 + i ] = J.V[ l * J.snr + i ];
   }
}

// The deflation's terms in a level's bases, columns col0 + q (q < ndef): on every row i (tree
// order) Wc = gdef_q xmu_i and Zc = xmu_i where row i is on surface q + 1 (all rows when ndef = 1),
// which makes the block between the sons of every cluster carry sum_q gdef_q xmu_q(a) xmu_q(b)^T.
__global__ void k_hd_deflation( int n3, int ndef, const double* __restrict__ xmu, const int* __restrict__ comp,
                                const double* __restrict__ gdef, double* __restrict__ Wc, double* __restrict__ Zc, long ld,
                                int col0 ) {
   const long total = (long)n3 * ndef;
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x ) {
      const int i = (int)( e % n3 ), q = (int)( e / n3 );
      const bool on = ndef == 1 || comp[ i / 3 ] == q + 1;
      Wc[ (long)( col0 + q ) * ld + i ] = on ? gdef[ q ] * xmu[ i ] : 0.0;
      Zc[ (long)( col0 + q ) * ld + i ] = on ? xmu[ i ] : 0.0;
   }
}

// The stacked factors of a block from its four sons' blocks (s00 s01 / s10 s11; the first
// son's rows come first): U = [U00 U01 0 0; 0 0 U10 U11], V = [V00 0 V10 0; 0 V01 0 V11], column
// blocks in the order 00, 01, 10, 11.  One workgroup per merged block.
struct HdMerge {
   const double* U[ 4 ];
   const double* V[ 4 ];
   int k[ 4 ];
   int m0, m1, n0, n1;  // rows of the two row sons and of the two column sons (unknowns)
   double* Uo;          // (m0 + m1) x sum k, column-major
   double* Vo;          // (n0 + n1) x sum k
};
__global__ void k_hd_stack( const HdMerge* __restrict__ jobs ) {
   const HdMerge J = jobs[ blockIdx.x ];
   const int M = J.m0 + J.m1, N = J.n0 + J.n1;
   int col = 0;
   for ( int b = 0; b < 4; ++b ) {
      const int ib = b >> 1, jb = b & 1;
      const int mr = ib ? J.m1 : J.m0, moff = ib ? J.m0 : 0;
      const int nr = jb ? J.n1 : J.n0, noff = jb ? J.n0 : 0;
      for ( long e = threadIdx.x; e < (long)J.k[ b ] * M; e += blockDim.x ) {
         const long l = e / M, i = e % M;
         const long li = i - moff;
         J.Uo[ ( col + l ) * M + i ] = ( li >= 0 && li < mr ) ? J.U[ b ][ l * mr + li ] : 0.0;
      }
      for ( long e = threadIdx.x; e < (long)J.k[ b ] * N; e += blockDim.x ) {
         const long l = e / N, i = e % N;
         const long li = i - noff;
         J.Vo[ ( col + l ) * N + i ] = ( li >= 0 && li < nr ) ? J.V[ b ][ l * nr + li ] : 0.0;
      }
      col += J.k[ b ];
   }
}

// ------------------------------------------------------------------ the H-matrix object

struct HMatrix {
   int n = 0, n3 = 0, ndef = 1, single = 0;
   HmOptions opt;
   Tree T;
   std::vector<Blk> blks;
   // tree-ordered per-triangle data and the deflation
   DevBuf<int> perm;
   DevBuf<double> src_t, xc_t, selfb_t, xmu_t, gdef;
   DevBuf<int> comp_t;
   // block storage
   std::deque<DevBuf<char>> pools;  // a deque: growing it never moves (and frees) an allocation
   DevBuf<BlkDev> blk_dev;
   DevBuf<int4> vtx_tasks, w_tasks;  // (block, l, segment) and (block, l, first task, segments)
   DevBuf<double> vtx_part;
   int n_vtx = 0, n_w = 0;
   int n_lr = 0;
   long w_elems = 0;              // rank * p summed (p = kMaxRhs)
   DevBuf<double> w;              // V^T X per low-rank block
   // leaf row lists
   DevBuf<int> leaf_r0, leaf_nr, ptr, ent_blk, ent_off;
   DevBuf<double> ypart;          // kSplit partial sums per leaf row
   int nleaves = 0;
   // preconditioner
   DevBuf<int> diag_blk;
   DevBuf<long> pinv_off;
   DevBuf<double> pinv;
   // deflation sums and Krylov work space
   DevBuf<double> defl_s, defl_part;
   DevBuf<double> Vb, Wk, Zk, Xk, Bk, Rk, part, hbuf;
   // the hierarchical preconditioner: level l (clusters with sons, 0 .. hd_depth - 1) has the
   // bases hd_W, hd_Z columns hd_off[l] .. hd_off[l] + hd_r[l] (hd_W holds W^ once factored) and
   // its clusters' K^-1 at hd_K + hd_koff[l]
   bool hd_on = false;
   int hd_lmin = 0;  // the coarsest level applied (the levels above are left out)
   KernelData K {};  // the kernel's tree-ordered data (for rebuilding the preconditioner)
   int hd_depth = 0;
   double hd_eps = 0.0, hd_piv_ratio = 0.0;
   std::vector<int> hd_r;
   std::vector<long> hd_off, hd_koff;
   long hd_cols = 0;
   DevBuf<double> hd_W, hd_Z, hd_K, hd_part, hd_S, hd_u;
   DevBuf<int> cl_r0, cl_nr;
   size_t hd_part_n = 0, hd_s_n = 0;

   void* alloc( size_t bytes ) {
      pools.emplace_back();
      return pools.back().ensure( std::max<size_t>( bytes, 16 ) );
   }

   template <typename E, int P>
   void product( hipStream_t st, const double* X, double* Y ) {
      if ( n_vtx > 0 ) {
         hipLaunchKernelGGL( ( k_hm_vtx<E, P> ), dim3( n_vtx ), dim3( 64 ), 0, st, vtx_tasks.p, blk_dev.p, X, (long)n3,
                             vtx_part.p );
         hipLaunchKernelGGL( k_hm_wsum, dim3( nblocks( n_w, 256 ) ), dim3( 256 ), 0, st, w_tasks.p, n_w, blk_dev.p, P, vtx_part.p,
                             w.p );
      }
      hipLaunchKernelGGL( ( k_hm_rows<E, P> ), dim3( nleaves, kSplit ), dim3( 64 ), 0, st, leaf_nr.p, ptr.p, ent_blk.p, ent_off.p,
                          blk_dev.p, X, (long)n3, w.p, ypart.p );
   }

   template <typename E>
   void product_p( hipStream_t st, int p, const double* X, double* Y ) {
      switch ( p ) {
#define BH_HM_CASE( P ) \
   case P: product<E, P>( st, X, Y ); break;
         BH_HM_CASE( 1 ) BH_HM_CASE( 2 ) BH_HM_CASE( 3 ) BH_HM_CASE( 4 ) BH_HM_CASE( 5 ) BH_HM_CASE( 6 ) BH_HM_CASE( 7 )
         BH_HM_CASE( 8 ) BH_HM_CASE( 9 ) BH_HM_CASE( 10 ) BH_HM_CASE( 11 ) BH_HM_CASE( 12 ) BH_HM_CASE( 13 )
         BH_HM_CASE( 14 ) BH_HM_CASE( 15 ) BH_HM_CASE( 16 )
#undef BH_HM_CASE
         default: throw std::runtime_error( "H-matrix product: 1 to 16 columns" );
      }
   }

   // Y = A X in tree order (n3 x p, leading dimension n3).
   void apply( hipStream_t st, int p, const double* X, double* Y ) {
      if ( single )
         product_p<float>( st, p, X, Y );
      else
         product_p<double>( st, p, X, Y );
      hipLaunchKernelGGL( k_hm_ysum, dim3( nblocks( (long)nleaves * p * kRowsWg, 256 ) ), dim3( 256 ), 0, st, nleaves, p,
                          leaf_r0.p, leaf_nr.p, ypart.p, Y, (long)n3 );
      const int nchunk = ( n3 + kDotChunkDefl - 1 ) / kDotChunkDefl;
      hipLaunchKernelGGL( k_hm_defl_part, dim3( nchunk, ndef * p ), dim3( kWg ), 0, st, n3, ndef, p, xmu_t.p, comp_t.p, X, (long)n3,
                          defl_part.p );
      hipLaunchKernelGGL( k_hm_defl_final, dim3( nblocks( ndef * p, 64 ) ), dim3( 64 ), 0, st, ndef * p, nchunk, defl_part.p,
                          defl_s.p );
      hipLaunchKernelGGL( k_hm_defl_add, dim3( nblocks( (long)n3 * p, 256 ) ), dim3( 256 ), 0, st, n3, ndef, p, xmu_t.p, comp_t.p,
                          gdef.p, defl_s.p, Y, (long)n3 );
   }

   void precond( hipStream_t st, int p, const double* V, double* Z ) {
      hipLaunchKernelGGL( k_hm_precond, dim3( nleaves ), dim3( kRowsWg ), 0, st, leaf_r0.p, leaf_nr.p, pinv_off.p, pinv.p, p, V,
                          (long)n3, Z, (long)n3 );
      if ( hd_on )
         for ( int l = hd_depth - 1; l >= hd_lmin; --l ) hd_level( st, l, Z, (long)n3, p );
   }

   // Rows of the sons of level l's clusters: heap clusters first .. first + 2^(l+1) - 1.
   int hd_first( int l ) const { return ( 2 << l ) - 1; }
   int hd_chunks( int l, int rows ) const {
      int mx = 0;
      for ( int c = hd_first( l ); c < hd_first( l ) + ( 2 << l ); ++c ) mx = std::max( mx, 3 * T.cnt[ c ] );
      return std::max( 1, ( mx + rows - 1 ) / rows );
   }

   // S = Zc_l(son)^T B(son) for every son of level l (r_l x q each, B's columns 0 .. q - 1).
   void hd_products( hipStream_t st, int l, const double* B, long ldb, int q ) {
      const int r = hd_r[ l ], nchild = 2 << l, nch = hd_chunks( l, kHdRows );
      const double* Zc = hd_Z.p + hd_off[ l ] * (long)n3;
      if ( (size_t)nchild * nch * r * q > hd_part_n || (size_t)nchild * r * q > hd_s_n )
         throw std::runtime_error( "hierarchical preconditioner: work space" );
      hipLaunchKernelGGL( k_hd_segtn, dim3( nchild * nch, ( q + kHdCols - 1 ) / kHdCols ), dim3( 256 ), 0, st, hd_first( l ), nch,
                          cl_r0.p, cl_nr.p, Zc, (long)n3, r, B, ldb, q, hd_part.p );
      const long rq = (long)r * q;
      hipLaunchKernelGGL( k_hd_segsum, dim3( nblocks( nchild * rq, 256 ) ), dim3( 256 ), 0, st, nchild, nch, rq, hd_part.p, hd_S.p );
   }

   // B(:, 0 .. q-1) <- the correction of level l applied: B(tau) -= W^(tau) K_tau^-1 Z_tau^T B(tau),
   // columns in tiles of kHdTile.
   void hd_level( hipStream_t st, int l, double* B, long ldb, int q ) {
      const int r = hd_r[ l ], nchild = 2 << l, nnode = 1 << l;
      if ( r == 0 ) return;
      const double* Wc = hd_W.p + hd_off[ l ] * (long)n3;
      const int nchu = hd_chunks( l, kHdUpd );
      for ( int j0 = 0; j0 < q; j0 += kHdTile ) {
         const int qt = std::min( kHdTile, q - j0 );
         hd_products( st, l, B + (long)j0 * ldb, ldb, qt );
         hipLaunchKernelGGL( k_hd_node, dim3( nnode, ( qt + kHdCols - 1 ) / kHdCols ), dim3( 256 ), 0, st, r, qt,
                             hd_K.p + hd_koff[ l ], hd_S.p, hd_u.p );
         hipLaunchKernelGGL( k_hd_update, dim3( nchild * nchu, ( qt + kHdCols - 1 ) / kHdCols ), dim3( kHdUpd ), 0, st, hd_first( l ),
                             nchu, cl_r0.p, cl_nr.p, Wc, (long)n3, r, hd_u.p, qt, B + (long)j0 * ldb, ldb );
      }
   }

   // out[i p + q] = Vb_i(:, q) . W(:, q) for i < nv, copied to the host.
   void dots( hipStream_t st, int p, int nv, const double* V, long vstride, const double* W, double* h_out ) {
      const int nchunk = ( n3 + kDotChunk - 1 ) / kDotChunk;
      hipLaunchKernelGGL( k_hm_dots, dim3( nchunk, nv * p ), dim3( kWg ), 0, st, n3, p, nv, V, vstride, W, part.p );
      hipLaunchKernelGGL( k_hm_dots_final, dim3( nblocks( nv * p, 64 ) ), dim3( 64 ), 0, st, nv * p, nchunk, part.p, hbuf.p );
      BH_HIP( hipMemcpyAsync( h_out, hbuf.p, sizeof( double ) * nv * p, hipMemcpyDeviceToHost, st ) );
      BH_HIP( hipStreamSynchronize( st ) );
   }
};

void HMatrixDeleter::operator()( HMatrix* h ) const { delete h; }

namespace {

constexpr int kHdDirect = 5120;  // clusters of at most this many triangles: their blocks are approximated directly

// One compressed block A(t, s) of the preconditioner: U (3 cnt(t) x rank) and V (3 cnt(s) x
// rank), column-major, in the build's pools.
struct HdRec {
   int rank = 0;
   const double* U = nullptr;
   const double* V = nullptr;
};

// Truncates the candidates (U, V) of tj at eps (the reference's truncate, hmatrix/src/mat.cpp:
// 129-160, by the recompression kernels), keeps at most cap[j] columns (the largest singular
// values first), and stores the factors column-major in a new pool entry; out[j] receives them.
void hd_truncate_store( hipStream_t st, std::vector<TruncJob>& tj, const std::vector<int>& cap, double eps,
                        std::deque<DevBuf<char>>& pools, std::vector<HdRec*>& out ) {
   const int nt = (int)tj.size();
   if ( nt == 0 ) return;
   // classes by kp (<= 32, 64, 96, 128), each with its own workgroup size and LDS
   std::vector<int> order;
   int ncls[ 4 ] = { 0, 0, 0, 0 };
   for ( int cls = 0; cls < 4; ++cls )
      for ( int j = 0; j < nt; ++j )
         if ( tj[ j ].k >= 1 && ( tj[ j ].kp + 31 ) / 32 - 1 == cls ) order.push_back( j ), ++ncls[ cls ];
   std::vector<TruncJob> jobs;
   size_t scratch = 0;
   for ( int j : order ) {
      jobs.push_back( tj[ j ] );
      jobs.back().G = reinterpret_cast<double*>( scratch );
      scratch += 5 * (size_t)tj[ j ].kp * tj[ j ].kp + 2 * (size_t)tj[ j ].kp;
   }
   const int nj = (int)jobs.size();
   std::vector<int> rank( nt, 0 );
   if ( nj > 0 ) {
      DevBuf<double> ws;
      double* base = ws.ensure( scratch );
      for ( TruncJob& t : jobs ) {
         const size_t off = reinterpret_cast<size_t>( t.G ), kk = (size_t)t.kp * t.kp;
         t.G = base + off;
         t.Z = t.G + 2 * kk;
         t.C = t.Z + kk;
         t.d = t.C + 2 * kk;
      }
      DevBuf<TruncJob> jd;
      DevBuf<int> rd, fd;
      jd.ensure( nj );
      rd.ensure( nj );
      fd.ensure( nj );
      BH_HIP( hipMemcpyAsync( jd.p, jobs.data(), sizeof( TruncJob ) * nj, hipMemcpyHostToDevice, st ) );
      const int o1 = ncls[ 0 ], o2 = o1 + ncls[ 1 ], o3 = o2 + ncls[ 2 ];
      if ( ncls[ 0 ] )
         hipLaunchKernelGGL( ( k_hm_trunc<128, 32> ), dim3( ncls[ 0 ] ), dim3( 128 ), 0, st, jd.p, eps, rd.p, fd.p );
      if ( ncls[ 1 ] )
         hipLaunchKernelGGL( ( k_hm_trunc<256, 64> ), dim3( ncls[ 1 ] ), dim3( 256 ), 0, st, jd.p + o1, eps, rd.p + o1, fd.p + o1 );
      if ( ncls[ 2 ] )
         hipLaunchKernelGGL( ( k_hm_trunc<512, 96> ), dim3( ncls[ 2 ] ), dim3( 512 ), 0, st, jd.p + o2, eps, rd.p + o2, fd.p + o2 );
      if ( ncls[ 3 ] )
         hipLaunchKernelGGL( ( k_hm_trunc<512, 128> ), dim3( ncls[ 3 ] ), dim3( 512 ), 0, st, jd.p + o3, eps, rd.p + o3, fd.p + o3 );
      std::vector<int> r( nj );
      BH_HIP( hipMemcpyAsync( r.data(), rd.p, sizeof( int ) * nj, hipMemcpyDeviceToHost, st ) );
      BH_HIP( hipStreamSynchronize( st ) );
      for ( int q = 0; q < nj; ++q ) r[ q ] = std::min( r[ q ], cap[ order[ q ] ] );
      size_t total = 0;
      for ( int q = 0; q < nj; ++q ) total += (size_t)( jobs[ q ].M + jobs[ q ].N ) * r[ q ];
      pools.emplace_back();
      double* store = reinterpret_cast<double*>( pools.back().ensure( std::max<size_t>( total, 1 ) * sizeof( double ) ) );
      std::vector<void*> uo( nj ), vo( nj );
      std::vector<int4> tasks;
      size_t at = 0;
      for ( int q = 0; q < nj; ++q ) {
         uo[ q ] = store + at;
         vo[ q ] = store + at + (size_t)jobs[ q ].M * r[ q ];
         at += (size_t)( jobs[ q ].M + jobs[ q ].N ) * r[ q ];
         HdRec* o = out[ order[ q ] ];
         o->rank = r[ q ];
         o->U = static_cast<const double*>( uo[ q ] );
         o->V = static_cast<const double*>( vo[ q ] );
         if ( r[ q ] <= 0 ) continue;
         for ( int side = 0; side < 2; ++side ) {
            const int R = side ? jobs[ q ].N : jobs[ q ].M;
            for ( int c = 0; c * kTrApply < R; ++c ) tasks.push_back( make_int4( q, side, c, 0 ) );
         }
      }
      BH_HIP( hipMemcpyAsync( rd.p, r.data(), sizeof( int ) * nj, hipMemcpyHostToDevice, st ) );
      if ( !tasks.empty() ) {
         DevBuf<void*> ud, vd;
         DevBuf<int4> td;
         ud.ensure( nj );
         vd.ensure( nj );
         td.ensure( tasks.size() );
         BH_HIP( hipMemcpyAsync( ud.p, uo.data(), sizeof( void* ) * nj, hipMemcpyHostToDevice, st ) );
         BH_HIP( hipMemcpyAsync( vd.p, vo.data(), sizeof( void* ) * nj, hipMemcpyHostToDevice, st ) );
         BH_HIP( hipMemcpyAsync( td.p, tasks.data(), sizeof( int4 ) * tasks.size(), hipMemcpyHostToDevice, st ) );
         hipLaunchKernelGGL( ( k_hm_trunc_apply<double, true> ), dim3( (unsigned)tasks.size() ), dim3( kTrApply ), 0, st, td.p, jd.p,
                             rd.p, ud.p, vd.p );
         BH_HIP( hipStreamSynchronize( st ) );
      }
      BH_HIP( hipStreamSynchronize( st ) );
   }
}

}  // namespace

// The hierarchical preconditioner of the built H-matrix (the section on it above): the blocks
// between the sons of every cluster compressed at eps, directly by cross approximation where the
// clusters have at most kHdDirect triangles, and above that by merging the four sons' blocks and
// truncating again; then every level's bases with the deflation's terms, and their
// factorization.  Returns false with a log line if a cluster's system is singular.
bool hd_build( HMatrix& h, hipStream_t st, const KernelData& K, const HmOptions& opt, double eps, std::vector<std::string>* log ) {
   auto say = [&]( const char* fmt, auto... a ) {
      char buf[ 512 ];
      std::snprintf( buf, sizeof buf, fmt, a... );
      if ( log ) log->emplace_back( buf );
   };
   const double t0 = now();
   const Tree& T = h.T;
   const int D = T.depth, n3 = h.n3, ndef = h.ndef;
   h.hd_on = false;
   if ( D < 1 ) return false;
   const int rfinal = kHdMaxR - ndef;  // a level's blocks: rank at most this (plus the deflation terms)
   const int rmerge = kTrMaxK / 4;     // a block merged into a coarser one
   int Ls = 0;
   for ( ; Ls < D; ++Ls ) {
      int mx = 0;
      for ( int c = ( 1 << Ls ) - 1; c < ( 2 << Ls ) - 1; ++c ) mx = std::max( mx, T.cnt[ c ] );
      if ( mx <= kHdDirect ) break;
   }
   const long NC = T.clusters();
   std::vector<HdRec> recs;
   std::vector<long> keys;
   auto key = [&]( int t, int s ) { return (long)t * NC + s; };
   auto sibling = [&]( int t, int s ) { return t != s && ( t - 1 ) / 2 == ( s - 1 ) / 2; };
   // every block the build makes, in a sorted table (its key -> its record)
   std::vector<std::pair<int, int>> pairs;
   if ( Ls >= 1 )
      for ( int t = ( 1 << Ls ) - 1; t < ( 2 << Ls ) - 1; ++t )
         for ( int s = ( 1 << Ls ) - 1; s < ( 2 << Ls ) - 1; ++s )
            if ( t != s ) pairs.emplace_back( t, s );
   for ( int m = std::max( Ls, 0 ) + 1; m <= D; ++m )
      for ( int c = ( 1 << ( m - 1 ) ) - 1; c < ( 1 << m ) - 1; ++c ) pairs.emplace_back( 2 * c + 1, 2 * c + 2 ), pairs.emplace_back( 2 * c + 2, 2 * c + 1 );
   const size_t ndirect = pairs.size();
   for ( int m = Ls - 1; m >= 1; --m )
      for ( int t = ( 1 << m ) - 1; t < ( 2 << m ) - 1; ++t )
         for ( int s = ( 1 << m ) - 1; s < ( 2 << m ) - 1; ++s )
            if ( t != s ) pairs.emplace_back( t, s );
   std::vector<size_t> idx( pairs.size() );
   std::iota( idx.begin(), idx.end(), 0 );
   std::sort( idx.begin(), idx.end(), [&]( size_t a, size_t b ) { return key( pairs[ a ].first, pairs[ a ].second ) < key( pairs[ b ].first, pairs[ b ].second ); } );
   keys.resize( pairs.size() );
   recs.resize( pairs.size() );
   std::vector<size_t> slot( pairs.size() );
   for ( size_t q = 0; q < idx.size(); ++q ) keys[ q ] = key( pairs[ idx[ q ] ].first, pairs[ idx[ q ] ].second ), slot[ idx[ q ] ] = q;
   auto rec = [&]( int t, int s ) -> HdRec& {
      const auto it = std::lower_bound( keys.begin(), keys.end(), key( t, s ) );
      if ( it == keys.end() || *it != key( t, s ) ) throw std::runtime_error( "hierarchical preconditioner: missing block" );
      return recs[ it - keys.begin() ];
   };
   std::deque<DevBuf<char>> pools;
   long n_capped = 0;
   // 1. cross approximation at aca_factor * eps of the direct blocks, largest first, in batches
   //    within the work-space budget; each batch truncated to eps and stored
   {
      std::vector<size_t> dj( ndirect );
      std::iota( dj.begin(), dj.end(), 0 );
      std::stable_sort( dj.begin(), dj.end(), [&]( size_t a, size_t b ) {
         return (long)T.cnt[ pairs[ a ].first ] + T.cnt[ pairs[ a ].second ] > (long)T.cnt[ pairs[ b ].first ] + T.cnt[ pairs[ b ].second ];
      } );
      auto kcap_of = [&]( size_t p ) {
         return std::max( 3, std::min( kTrMaxK, 3 * std::min( T.cnt[ pairs[ p ].first ], T.cnt[ pairs[ p ].second ] ) / 2 ) );
      };
      auto need_of = [&]( size_t p ) {
         const long m = T.cnt[ pairs[ p ].first ], nc = T.cnt[ pairs[ p ].second ];
         return ( sizeof( double ) * ( (size_t)( 3 * m + 3 * nc ) * kcap_of( p ) + 9 * nc + 3 * m + 3 * nc ) + sizeof( int ) * ( m + 2 ) +
                  255 ) / 256 * 256;
      };
      DevBuf<unsigned long long> calls;
      calls.ensure( 1 );
      BH_HIP( hipMemsetAsync( calls.p, 0, sizeof( unsigned long long ), st ) );
      size_t pos = 0;
      while ( pos < dj.size() ) {
         std::vector<size_t> batch;
         size_t bytes = 0;
         while ( pos < dj.size() && ( batch.empty() || bytes + need_of( dj[ pos ] ) <= (size_t)opt.ws_bytes ) )
            bytes += need_of( dj[ pos ] ), batch.push_back( dj[ pos++ ] );
         DevBuf<char> ws;
         char* base = ws.ensure( bytes );
         std::vector<AcaJob> jobs;
         size_t at = 0;
         for ( size_t p : batch ) {
            AcaJob j {};
            j.rb = T.beg[ pairs[ p ].first ];
            j.m = T.cnt[ pairs[ p ].first ];
            j.cb = T.beg[ pairs[ p ].second ];
            j.nc = T.cnt[ pairs[ p ].second ];
            j.kcap = kcap_of( p );
            double* d = reinterpret_cast<double*>( base + at );
            at += need_of( p );
            const long M = 3L * j.m, N = 3L * j.nc;
            j.U = d;
            j.V = d + M * j.kcap;
            j.R = j.V + N * j.kcap;
            j.u = j.R + 3 * N;
            j.v = j.u + M;
            j.used = reinterpret_cast<int*>( j.v + N );
            jobs.push_back( j );
         }
         const int nj = (int)jobs.size();
         DevBuf<AcaJob> jd;
         DevBuf<int> rd;
         jd.ensure( nj );
         rd.ensure( nj );
         BH_HIP( hipMemcpyAsync( jd.p, jobs.data(), sizeof( AcaJob ) * nj, hipMemcpyHostToDevice, st ) );
         int nbig = 0;
         while ( nbig < nj && 3L * ( jobs[ nbig ].m + jobs[ nbig ].nc ) >= kAcaBig ) ++nbig;
         int nsm = nbig;
         while ( nsm < nj && 3L * ( jobs[ nsm ].m + jobs[ nsm ].nc ) >= kAcaSmall ) ++nsm;
         const double eaca = opt.aca_factor * eps;
         if ( nbig ) hipLaunchKernelGGL( k_hm_aca<512>, dim3( nbig ), dim3( 512 ), 0, st, jd.p, K, eaca, rd.p, calls.p );
         if ( nsm > nbig )
            hipLaunchKernelGGL( k_hm_aca<kWg>, dim3( nsm - nbig ), dim3( kWg ), 0, st, jd.p + nbig, K, eaca, rd.p + nbig, calls.p );
         if ( nj > nsm ) hipLaunchKernelGGL( k_hm_aca<64>, dim3( nj - nsm ), dim3( 64 ), 0, st, jd.p + nsm, K, eaca, rd.p + nsm, calls.p );
         std::vector<int> rk( nj );
         BH_HIP( hipMemcpyAsync( rk.data(), rd.p, sizeof( int ) * nj, hipMemcpyDeviceToHost, st ) );
         BH_HIP( hipStreamSynchronize( st ) );
         std::vector<TruncJob> tj( nj );
         std::vector<int> cap( nj );
         std::vector<HdRec*> out( nj );
         for ( int q = 0; q < nj; ++q ) {
            const auto& pr = pairs[ batch[ q ] ];
            int k = rk[ q ];
            if ( k < 0 ) k = jobs[ q ].kcap, ++n_capped;  // the cap reached: its terms so far
            TruncJob t {};
            t.U = jobs[ q ].U;
            t.V = jobs[ q ].V;
            t.M = 3 * jobs[ q ].m;
            t.N = 3 * jobs[ q ].nc;
            t.k = k;
            t.kp = ( k + 3 ) / 4 * 4;
            tj[ q ] = t;
            cap[ q ] = sibling( pr.first, pr.second ) ? rfinal : rmerge;
            out[ q ] = &recs[ slot[ batch[ q ] ] ];
         }
         hd_truncate_store( st, tj, cap, eps, pools, out );
      }
   }
   const double t_aca = now() - t0;
   // 2. the levels above: each block from its four sons' blocks, stacked and truncated again
   for ( int m = Ls - 1; m >= 1; --m ) {
      std::vector<std::pair<int, int>> lv;
      for ( int t = ( 1 << m ) - 1; t < ( 2 << m ) - 1; ++t )
         for ( int s = ( 1 << m ) - 1; s < ( 2 << m ) - 1; ++s )
            if ( t != s ) lv.emplace_back( t, s );
      std::vector<HdMerge> mj;
      size_t total = 0;
      for ( const auto& ts : lv ) {
         HdMerge g {};
         const int t = ts.first, s = ts.second;
         for ( int b = 0; b < 4; ++b ) {
            const HdRec& r = rec( 2 * t + 1 + ( b >> 1 ), 2 * s + 1 + ( b & 1 ) );
            g.U[ b ] = r.U, g.V[ b ] = r.V, g.k[ b ] = r.rank;
         }
         g.m0 = 3 * T.cnt[ 2 * t + 1 ], g.m1 = 3 * T.cnt[ 2 * t + 2 ];
         g.n0 = 3 * T.cnt[ 2 * s + 1 ], g.n1 = 3 * T.cnt[ 2 * s + 2 ];
         const int kt = g.k[ 0 ] + g.k[ 1 ] + g.k[ 2 ] + g.k[ 3 ];
         total += (size_t)( g.m0 + g.m1 + g.n0 + g.n1 ) * kt;
         mj.push_back( g );
      }
      DevBuf<double> stk;
      double* base = stk.ensure( std::max<size_t>( total, 1 ) );
      size_t at = 0;
      std::vector<TruncJob> tj( mj.size() );
      std::vector<int> cap( mj.size() );
      std::vector<HdRec*> out( mj.size() );
      for ( size_t q = 0; q < mj.size(); ++q ) {
         HdMerge& g = mj[ q ];
         const int kt = g.k[ 0 ] + g.k[ 1 ] + g.k[ 2 ] + g.k[ 3 ];
         g.Uo = base + at;
         at += (size_t)( g.m0 + g.m1 ) * kt;
         g.Vo = base + at;
         at += (size_t)( g.n0 + g.n1 ) * kt;
         TruncJob t {};
         t.U = g.Uo;
         t.V = g.Vo;
         t.M = g.m0 + g.m1;
         t.N = g.n0 + g.n1;
         t.k = kt;
         t.kp = ( kt + 3 ) / 4 * 4;
         tj[ q ] = t;
         cap[ q ] = sibling( lv[ q ].first, lv[ q ].second ) ? rfinal : rmerge;
         out[ q ] = &rec( lv[ q ].first, lv[ q ].second );
      }
      DevBuf<HdMerge> md;
      md.ensure( mj.size() );
      BH_HIP( hipMemcpyAsync( md.p, mj.data(), sizeof( HdMerge ) * mj.size(), hipMemcpyHostToDevice, st ) );
      hipLaunchKernelGGL( k_hd_stack, dim3( (unsigned)mj.size() ), dim3( 256 ), 0, st, md.p );
      hd_truncate_store( st, tj, cap, eps, pools, out );
   }
   // 3. the bases of every level: the blocks between its clusters' sons, and the deflation
   h.hd_depth = D;
   h.hd_r.assign( D, 0 );
   h.hd_off.assign( D + 1, 0 );
   h.hd_koff.assign( D + 1, 0 );
   for ( int l = 0; l < D; ++l ) {
      int r = 0;
      for ( int c = ( 1 << l ) - 1; c < ( 2 << l ) - 1; ++c )
         r = std::max( { r, rec( 2 * c + 1, 2 * c + 2 ).rank, rec( 2 * c + 2, 2 * c + 1 ).rank } );
      h.hd_r[ l ] = r + ndef;
      h.hd_off[ l + 1 ] = h.hd_off[ l ] + h.hd_r[ l ];
      h.hd_koff[ l + 1 ] = h.hd_koff[ l ] + ( 1L << l ) * 3L * h.hd_r[ l ] * h.hd_r[ l ];
   }
   h.hd_cols = h.hd_off[ D ];
   h.hd_W.ensure( (size_t)n3 * h.hd_cols );
   h.hd_Z.ensure( (size_t)n3 * h.hd_cols );
   h.hd_K.ensure( std::max<long>( h.hd_koff[ D ], 1 ) );
   BH_HIP( hipMemsetAsync( h.hd_W.p, 0, sizeof( double ) * n3 * h.hd_cols, st ) );
   BH_HIP( hipMemsetAsync( h.hd_Z.p, 0, sizeof( double ) * n3 * h.hd_cols, st ) );
   {
      std::vector<int> r0( NC ), nr( NC );
      for ( long c = 0; c < NC; ++c ) r0[ c ] = 3 * T.beg[ c ], nr[ c ] = 3 * T.cnt[ c ];
      h.cl_r0.ensure( NC );
      h.cl_nr.ensure( NC );
      BH_HIP( hipMemcpyAsync( h.cl_r0.p, r0.data(), sizeof( int ) * NC, hipMemcpyHostToDevice, st ) );
      BH_HIP( hipMemcpyAsync( h.cl_nr.p, nr.data(), sizeof( int ) * NC, hipMemcpyHostToDevice, st ) );
      DevBuf<HdScatter> sd;
      for ( int l = 0; l < D; ++l ) {
         std::vector<HdScatter> sj;
         for ( int c = ( 1 << l ) - 1; c < ( 2 << l ) - 1; ++c )
            for ( int side = 0; side < 2; ++side ) {
               const int t = 2 * c + 1 + side, s = 2 * c + 2 - side;
               const HdRec& r = rec( t, s );
               if ( r.rank > 0 ) sj.push_back( HdScatter { r.U, r.V, r.rank, 3 * T.beg[ t ], 3 * T.cnt[ t ], 3 * T.beg[ s ], 3 * T.cnt[ s ] } );
            }
         double* Wc = h.hd_W.p + h.hd_off[ l ] * (long)n3;
         double* Zc = h.hd_Z.p + h.hd_off[ l ] * (long)n3;
         if ( !sj.empty() ) {
            sd.ensure( sj.size() );
            BH_HIP( hipMemcpyAsync( sd.p, sj.data(), sizeof( HdScatter ) * sj.size(), hipMemcpyHostToDevice, st ) );
            hipLaunchKernelGGL( k_hd_scatter, dim3( (unsigned)sj.size() ), dim3( 256 ), 0, st, sd.p, Wc, Zc, (long)n3 );
            BH_HIP( hipStreamSynchronize( st ) );  // sd is reused by the next level
         }
         hipLaunchKernelGGL( k_hd_deflation, dim3( nblocks( (long)n3 * ndef, 256 ) ), dim3( 256 ), 0, st, n3, ndef, h.xmu_t.p,
                             h.comp_t.p, h.gdef.p, Wc, Zc, (long)n3, h.hd_r[ l ] - ndef );
      }
   }
   BH_HIP( hipStreamSynchronize( st ) );
   pools.clear();  // the blocks are in the bases now
   // work space of the level corrections (kHdTile columns at a time)
   h.hd_part_n = h.hd_s_n = 1;
   for ( int l = 0; l < D; ++l ) {
      const size_t r = h.hd_r[ l ], nchild = (size_t)2 << l;
      h.hd_part_n = std::max( h.hd_part_n, nchild * h.hd_chunks( l, kHdRows ) * r * kHdTile );
      h.hd_s_n = std::max( h.hd_s_n, nchild * r * kHdTile );
   }
   h.hd_part.ensure( h.hd_part_n );
   h.hd_S.ensure( h.hd_s_n );
   h.hd_u.ensure( h.hd_s_n );
   // 4. the factorization: the leaf inverses on every basis, then from the finest level up, each
   //    level's cluster systems and its correction of the coarser levels' bases
   for ( long c0 = 0; c0 < h.hd_cols; c0 += kMaxRhs ) {
      double* B = h.hd_W.p + c0 * n3;
      hipLaunchKernelGGL( k_hm_precond, dim3( h.nleaves ), dim3( kRowsWg ), 0, st, h.leaf_r0.p, h.leaf_nr.p, h.pinv_off.p, h.pinv.p,
                          (int)std::min<long>( kMaxRhs, h.hd_cols - c0 ), B, (long)n3, B, (long)n3 );
   }
   DevBuf<double> piv;
   DevBuf<int> fail;
   piv.ensure( (size_t)2 << D );
   fail.ensure( 1 );
   BH_HIP( hipMemsetAsync( fail.p, 0, sizeof( int ), st ) );
   long piv_at = 0;
   for ( int l = D - 1; l >= 0; --l ) {
      const int r = h.hd_r[ l ];
      if ( r == 0 ) continue;
      h.hd_products( st, l, h.hd_W.p + h.hd_off[ l ] * (long)n3, (long)n3, r );
      hipLaunchKernelGGL( k_hd_kinv, dim3( 1 << l ), dim3( 256 ), 0, st, r, h.hd_S.p, h.hd_K.p + h.hd_koff[ l ], piv.p + piv_at,
                          fail.p );
      piv_at += 2L << l;
      if ( h.hd_off[ l ] > 0 ) h.hd_level( st, l, h.hd_W.p, (long)n3, (int)h.hd_off[ l ] );
   }
   int hf = 0;
   std::vector<double> pv( piv_at );
   BH_HIP( hipMemcpyAsync( &hf, fail.p, sizeof hf, hipMemcpyDeviceToHost, st ) );
   if ( piv_at ) BH_HIP( hipMemcpyAsync( pv.data(), piv.p, sizeof( double ) * piv_at, hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipStreamSynchronize( st ) );
   if ( hf ) {
      say( "preconditioner: the hierarchical factorization at eps %.0e broke down (a singular cluster system)", eps );
      return false;
   }
   double pmin = HUGE_VAL, pmax = 0.0;
   for ( long q = 0; q < piv_at; q += 2 ) pmin = std::min( pmin, pv[ q ] ), pmax = std::max( pmax, pv[ q + 1 ] );
   h.hd_piv_ratio = pmax > 0 ? pmin / pmax : 0.0;
   h.hd_eps = eps;
   h.hd_on = true;
   h.hd_lmin = std::min( D, opt.pre_skip );
   int rmax = 0;
   for ( int r : h.hd_r ) rmax = std::max( rmax, r );
   say( "preconditioner: hierarchical factorization at eps %.0e over %d levels in %.3f s (approximations %.3f s, %ld at "
        "their cap), level ranks up to %d (%ld basis columns, %.2f GB), smallest pivot ratio of the cluster systems %.2e",
        eps, D, now() - t0, t_aca, n_capped, rmax, h.hd_cols, 2.0 * 8.0 * n3 * h.hd_cols / 1e9, h.hd_piv_ratio );
   return true;
}

HmStats hm_plan( int n, const double* h_v, const double* h_xc, const HmOptions& opt ) {
   HmStats S;
   Tree T = build_tree( n, h_v, h_xc, opt.leaf );
   const std::vector<Blk> blks = make_blocks( T, opt );
   S.depth = T.depth;
   S.nleaves = T.nleaves();
   S.leaf_min = 1 << 30;
   for ( int l = 0; l < T.nleaves(); ++l ) {
      const int c = T.cnt[ T.leaf_cluster( l ) ];
      S.leaf_min = std::min( S.leaf_min, c ), S.leaf_max = std::max( S.leaf_max, c );
   }
   for ( const Blk& b : blks ) {
      if ( b.type == kLowRank ) ++S.n_lowrank;
      else if ( b.small ) ++S.n_small;
      else ++S.n_dense;
   }
   return S;
}

HMatrixPtr hm_build( hipStream_t st, hipStream_t st_b, hipStream_t st_c, const HmInput& in, const HmOptions& opt, HmStats* stats,
                     std::vector<std::string>* log ) {
   auto say = [&]( const char* fmt, auto... a ) {
      char buf[ 512 ];
      std::snprintf( buf, sizeof buf, fmt, a... );
      if ( log ) log->emplace_back( buf );
   };
   HmStats S;
   HMatrixPtr H( new HMatrix );
   HMatrix& h = *H;
   h.n = in.n;
   h.n3 = 3 * in.n;
   h.ndef = std::max( 1, in.ndef );
   h.single = opt.single;
   h.opt = opt;
   const int n = in.n;
   double t0 = now();
   h.T = build_tree( n, in.h_v, in.h_xc, opt.leaf );
   h.blks = make_blocks( h.T, opt );
   const Tree& T = h.T;
   S.depth = T.depth;
   S.nleaves = T.nleaves();
   S.leaf_min = 1 << 30;
   for ( int l = 0; l < T.nleaves(); ++l ) {
      const int c = T.cnt[ T.leaf_cluster( l ) ];
      S.leaf_min = std::min( S.leaf_min, c ), S.leaf_max = std::max( S.leaf_max, c );
   }
   if ( 3 * S.leaf_max > kMaxLeafRows ) throw std::runtime_error( "H-matrix: leaf clusters larger than 42 triangles" );
   S.t_tree = now() - t0;
   say( "cluster tree: depth %d, %d leaf clusters of %d to %d triangles, %.2f s", S.depth, S.nleaves, S.leaf_min, S.leaf_max,
        S.t_tree );

   // The device memory the build and the solves will need, estimated before anything is allocated
   // (the reference's memory plan and refusal, hmatrix/src/api.cpp:146-339, :585-630): the dense
   // blocks exactly; the low-rank blocks at (rows + columns) x a rank model of the cross
   // approximations (measured at 245,760 unknowns: mean rank 48 at eps 1e-6, about 8 per decade of
   // eps, capped at kcap), which bounds the recompressed storage too; the approximations' work space
   // (its budget); the Krylov basis; the hierarchical factorization's bases at its widest (when it
   // is the preconditioner from the start, or -hlu); the vectors.  The floor is 5% of the device's
   // memory, at least 1 GB.
   {
      const size_t esz = opt.single ? sizeof( float ) : sizeof( double );
      const double kmodel = std::min<double>( opt.kcap, std::max( 4.0, 8.0 * std::log10( 1.0 / opt.eps ) ) );
      double dense = 0.0, lowrank = 0.0, ws = 0.0;
      for ( const Blk& b : h.blks ) {
         if ( b.type == kLowRank ) {
            lowrank += (double)( b.nr + b.nc ) * kmodel * esz;
            ws += 8.0 * (double)( b.nr + b.nc ) * std::min<long>( opt.kcap, std::min( b.nr, b.nc ) / 2 );
         } else {
            dense += (double)b.nr * b.nc * esz;
         }
      }
      ws = std::min( ws, opt.ws_bytes );
      const double n3d = 3.0 * n;
      const int restart = std::max( opt.restart, 30 );
      const double krylov = 8.0 * n3d * opt.nrhs * ( restart + 8 );
      const double pre = ( opt.precond == 1 || opt.direct ) ? 2.0 * 8.0 * n3d * T.depth * kHdMaxR : 8.0 * n3d * 96;
      const double est = dense + lowrank + ws + krylov + pre + 8.0 * n3d * ( kSrcStride + 20 );
      size_t freeb = 0, total = 0;
      BH_HIP( hipMemGetInfo( &freeb, &total ) );
      const double floor = opt.mem_floor >= 0.0 ? opt.mem_floor : std::max( 1e9, 0.05 * (double)total );
      S.mem_estimate = est, S.mem_free = (double)freeb, S.mem_floor = floor;
      say( "memory: about %.1f GB needed (dense blocks %.1f, low-rank about %.1f, approximation work space %.1f, Krylov "
           "basis %.1f, preconditioner %.1f), %.1f GB free on the device",
           est / 1e9, dense / 1e9, lowrank / 1e9, ws / 1e9, krylov / 1e9, pre / 1e9, freeb / 1e9 );
      if ( opt.memcheck && est > (double)freeb - floor ) {
         char buf[ 400 ];
         std::snprintf( buf, sizeof buf,
                        "not enough device memory for the H-matrix solver at eps %.0e: about %.1f GB needed, %.1f GB free above "
                        "the floor of %.1f GB (BEST_HMATRIX_MEMFLOOR sets the floor in GB; set BEST_HMATRIX_MEMCHECK=0 to "
                        "try anyway)%s",
                        opt.eps, est / 1e9, std::max( 0.0, ( (double)freeb - floor ) / 1e9 ), floor / 1e9,
                        opt.single ? "" : "; -hs (single precision) needs about half" );
         throw std::runtime_error( buf );
      }
   }

   // tree-ordered data
   t0 = now();
   h.perm.ensure( n );
   BH_HIP( hipMemcpyAsync( h.perm.p, T.perm.data(), sizeof( int ) * n, hipMemcpyHostToDevice, st ) );
   h.src_t.ensure( (size_t)kSrcStride * n );
   h.xc_t.ensure( 3 * (size_t)n );
   h.selfb_t.ensure( 9 * (size_t)n );
   h.xmu_t.ensure( 3 * (size_t)n );
   h.comp_t.ensure( n );
   hipLaunchKernelGGL( k_hm_permute, dim3( nblocks( n, 256 ) ), dim3( 256 ), 0, st, n, h.perm.p, in.d_src, in.d_xc, in.d_selfb,
                       in.d_xmu, in.d_comp, h.src_t.p, h.xc_t.p, h.selfb_t.p, h.xmu_t.p, h.comp_t.p );
   h.gdef.ensure( h.ndef );
   BH_HIP( hipMemcpyAsync( h.gdef.p, in.d_gdef, sizeof( double ) * h.ndef, hipMemcpyDeviceToDevice, st ) );
   KernelData K { h.src_t.p, h.xc_t.p, h.selfb_t.p, n, in.cg };
   DevBuf<unsigned long long> calls;
   calls.ensure( 1 );
   BH_HIP( hipMemsetAsync( calls.p, 0, sizeof( unsigned long long ), st ) );
   const size_t esz = opt.single ? sizeof( float ) : sizeof( double );

   // low-rank blocks: cross approximation in batches, largest first.  The largest blocks (3 (m + n)
   // >= kAcaBig: a few hundred, each a 512-thread workgroup that runs for up to hundreds of ms)
   // form a batch of their own on a second stream, launched first and finished last, so that the
   // other batches and their recompression run beside them instead of waiting for them.
   std::vector<int> lr;
   for ( int i = 0; i < (int)h.blks.size(); ++i )
      if ( h.blks[ i ].type == kLowRank ) lr.push_back( i );
   std::stable_sort( lr.begin(), lr.end(), [&]( int a, int b ) {
      return (double)h.blks[ a ].nr * h.blks[ a ].nc > (double)h.blks[ b ].nr * h.blks[ b ].nc;
   } );
   std::vector<const void*> blk_a( h.blks.size(), nullptr ), blk_b( h.blks.size(), nullptr );
   long n_aca = 0;
   {
      struct Ev {
         hipEvent_t e = nullptr;
         ~Ev() {
            if ( e ) (void)hipEventDestroy( e );
         }
      } aux;
      BH_HIP( hipEventCreateWithFlags( &aux.e, hipEventDisableTiming ) );
      hipStream_t st_big = st_b, st_aux = st_c;
      hipEvent_t ev = aux.e;
      struct Batch {
         std::vector<int> idx;  // positions in lr
         std::vector<AcaJob> jobs;
         DevBuf<char> ws;
         DevBuf<AcaJob> jobs_d;
         DevBuf<int> rank_d;
         DevBuf<void*> uo_d, vo_d;
         DevBuf<double> tr_ws;  // recompression: per-job Gram, rotation and coefficient matrices
         DevBuf<TruncJob> tj_d;
         DevBuf<int> trank_d, tfail_d;
         DevBuf<void*> tuo_d, tvo_d;
         DevBuf<int4> ttask_d;
      };
      auto kcap_of = [&]( const Blk& b ) { return std::max( 3, (int)std::min<long>( opt.kcap, std::min( b.nr, b.nc ) / 2 ) ); };
      auto need_of = [&]( const Blk& b ) {
         return sizeof( double ) * ( (size_t)( b.nr + b.nc ) * kcap_of( b ) + 3 * b.nc + b.nr + b.nc ) +
                sizeof( int ) * ( b.nr / 3 + 2 ) + 256;
      };
      // Work space and jobs of the batch's blocks, then the approximations: the 512-thread prefix
      // (the batch is in decreasing size) on s2, the rest on s (s2 == s: all on s), in 256-thread
      // workgroups and, for the smallest blocks, one wave each.
      auto launch = [&]( Batch& B, hipStream_t s, hipStream_t s2 ) {
         B.jobs.clear();
         std::vector<size_t> offs;
         size_t bytes = 0;
         for ( int p : B.idx ) {
            const Blk& b = h.blks[ lr[ p ] ];
            AcaJob j {};
            j.rb = (int)( b.r0 / 3 );
            j.m = (int)( b.nr / 3 );
            j.cb = (int)( b.c0 / 3 );
            j.nc = (int)( b.nc / 3 );
            j.kcap = kcap_of( b );
            offs.push_back( bytes );
            bytes += ( need_of( b ) + 255 ) / 256 * 256;
            B.jobs.push_back( j );
         }
         char* base = B.ws.ensure( bytes );
         for ( size_t q = 0; q < B.jobs.size(); ++q ) {
            AcaJob& j = B.jobs[ q ];
            double* d = reinterpret_cast<double*>( base + offs[ q ] );
            const long M = 3L * j.m, N = 3L * j.nc;
            j.U = d;
            j.V = d + M * j.kcap;
            j.R = j.V + N * j.kcap;
            j.u = j.R + 3 * N;
            j.v = j.u + M;
            j.used = reinterpret_cast<int*>( j.v + N );
         }
         const int nj = (int)B.jobs.size();
         B.jobs_d.ensure( nj );
         B.rank_d.ensure( nj );
         BH_HIP( hipMemcpyAsync( B.jobs_d.p, B.jobs.data(), sizeof( AcaJob ) * nj, hipMemcpyHostToDevice, s ) );
         int nbig = 0;
         while ( nbig < nj && 3L * ( B.jobs[ nbig ].m + B.jobs[ nbig ].nc ) >= kAcaBig ) ++nbig;
         if ( nbig > 0 && s2 != s ) {
            BH_HIP( hipEventRecord( ev, s ) );
            BH_HIP( hipStreamWaitEvent( s2, ev, 0 ) );
         }
         if ( nbig > 0 )
            hipLaunchKernelGGL( k_hm_aca<512>, dim3( nbig ), dim3( 512 ), 0, s2, B.jobs_d.p, K, opt.aca_factor * opt.eps,
                                B.rank_d.p, calls.p );
         int nsm = nbig;  // the sizes decrease: the one-wave blocks are a suffix
         while ( nsm < nj && 3L * ( B.jobs[ nsm ].m + B.jobs[ nsm ].nc ) >= kAcaSmall ) ++nsm;
         if ( nsm > nbig )
            hipLaunchKernelGGL( k_hm_aca<kWg>, dim3( nsm - nbig ), dim3( kWg ), 0, s, B.jobs_d.p + nbig, K,
                                opt.aca_factor * opt.eps, B.rank_d.p + nbig, calls.p );
         if ( nj > nsm )
            hipLaunchKernelGGL( k_hm_aca<64>, dim3( nj - nsm ), dim3( 64 ), 0, s, B.jobs_d.p + nsm, K, opt.aca_factor * opt.eps,
                                B.rank_d.p + nsm, calls.p );
         if ( nbig > 0 && s2 != s ) {
            BH_HIP( hipEventRecord( ev, s2 ) );
            BH_HIP( hipStreamWaitEvent( s, ev, 0 ) );
         }
      };
      // Waits for the batch's approximations, recompresses them (the two classes of large ranks on
      // s2), then stores every block (the recompressed ones by k_hm_trunc_apply, the rest copied).
      auto finish = [&]( Batch& B, hipStream_t s, hipStream_t s2 ) {
         const int nj = (int)B.jobs.size();
         const std::vector<AcaJob>& jobs = B.jobs;
         std::vector<int> ranks( nj );
         BH_HIP( hipMemcpyAsync( ranks.data(), B.rank_d.p, sizeof( int ) * nj, hipMemcpyDeviceToHost, s ) );
         BH_HIP( hipStreamSynchronize( s ) );
         const double t_tr = now();
         std::vector<int> tq;  // batch positions of the recompressed jobs
         std::vector<TruncJob> tj;
         std::vector<int> aca_rank = ranks;
         if ( opt.trunc ) {
            size_t scratch = 0;
            // by class (kp <= 32, 64, 96, 128), in batch order (decreasing size) within a class
            std::vector<int> order;
            for ( int cls = 0; cls < 4; ++cls )
               for ( int q = 0; q < nj; ++q ) {
                  const int k = ranks[ q ];
                  if ( k >= 1 && k <= kTrMaxK && ( ( k + 3 ) / 4 * 4 + 31 ) / 32 - 1 == cls ) order.push_back( q );
               }
            int ncls[ 4 ] = { 0, 0, 0, 0 };
            for ( int q : order ) {
               const int k = ranks[ q ];
               const int kp = ( k + 3 ) / 4 * 4;
               ++ncls[ ( kp + 31 ) / 32 - 1 ];
               TruncJob t {};
               t.U = jobs[ q ].U;
               t.V = jobs[ q ].V;
               t.M = 3 * jobs[ q ].m;
               t.N = 3 * jobs[ q ].nc;
               t.k = k;
               t.kp = kp;
               t.G = reinterpret_cast<double*>( scratch );  // offsets until the scratch exists
               scratch += 5 * (size_t)kp * kp + 2 * (size_t)kp;
               tq.push_back( q );
               tj.push_back( t );
            }
            if ( !tj.empty() ) {
               double* base = B.tr_ws.ensure( scratch );
               for ( TruncJob& t : tj ) {
                  const size_t off = reinterpret_cast<size_t>( t.G );
                  const size_t kk = (size_t)t.kp * t.kp;
                  t.G = base + off;
                  t.Z = t.G + 2 * kk;
                  t.C = t.Z + kk;
                  t.d = t.C + 2 * kk;
               }
               const int nt = (int)tj.size();
               B.tj_d.ensure( nt );
               B.trank_d.ensure( nt );
               B.tfail_d.ensure( nt );
               BH_HIP( hipMemcpyAsync( B.tj_d.p, tj.data(), sizeof( TruncJob ) * nt, hipMemcpyHostToDevice, s ) );
               const int o1 = ncls[ 0 ], o2 = o1 + ncls[ 1 ], o3 = o2 + ncls[ 2 ];
               BH_HIP( hipEventRecord( ev, s ) );
               BH_HIP( hipStreamWaitEvent( s2, ev, 0 ) );
               if ( ncls[ 3 ] )
                  hipLaunchKernelGGL( ( k_hm_trunc<512, 128> ), dim3( ncls[ 3 ] ), dim3( 512 ), 0, s2, B.tj_d.p + o3, opt.eps,
                                      B.trank_d.p + o3, B.tfail_d.p + o3 );
               if ( ncls[ 2 ] )
                  hipLaunchKernelGGL( ( k_hm_trunc<512, 96> ), dim3( ncls[ 2 ] ), dim3( 512 ), 0, s2, B.tj_d.p + o2, opt.eps,
                                      B.trank_d.p + o2, B.tfail_d.p + o2 );
               if ( ncls[ 1 ] )
                  hipLaunchKernelGGL( ( k_hm_trunc<256, 64> ), dim3( ncls[ 1 ] ), dim3( 256 ), 0, s, B.tj_d.p + o1, opt.eps,
                                      B.trank_d.p + o1, B.tfail_d.p + o1 );
               if ( ncls[ 0 ] )
                  hipLaunchKernelGGL( ( k_hm_trunc<128, 32> ), dim3( ncls[ 0 ] ), dim3( 128 ), 0, s, B.tj_d.p, opt.eps,
                                      B.trank_d.p, B.tfail_d.p );
               BH_HIP( hipEventRecord( ev, s2 ) );
               BH_HIP( hipStreamWaitEvent( s, ev, 0 ) );
               std::vector<int> tr( nt ), tf( nt );
               BH_HIP( hipMemcpyAsync( tr.data(), B.trank_d.p, sizeof( int ) * nt, hipMemcpyDeviceToHost, s ) );
               BH_HIP( hipMemcpyAsync( tf.data(), B.tfail_d.p, sizeof( int ) * nt, hipMemcpyDeviceToHost, s ) );
               BH_HIP( hipStreamSynchronize( s ) );
               for ( int i = 0; i < nt; ++i ) {
                  ranks[ tq[ i ] ] = tr[ i ];
                  S.n_trunc_failed += tf[ i ];
               }
            }
         }
         for ( int q = 0; q < nj; ++q )
            if ( aca_rank[ q ] > 0 ) S.kmean_aca += aca_rank[ q ], ++n_aca;
         // storage for the converged blocks
         size_t total = 0;
         for ( int q = 0; q < nj; ++q )
            if ( ranks[ q ] > 0 ) total += ( (size_t)( 3L * jobs[ q ].m + 3L * jobs[ q ].nc ) * ranks[ q ] * esz + 255 ) / 256 * 256;
         char* store = static_cast<char*>( h.alloc( total ) );
         std::vector<void*> uo( nj, nullptr ), vo( nj, nullptr );
         size_t at = 0;
         for ( int q = 0; q < nj; ++q ) {
            const int bi = lr[ B.idx[ q ] ];
            Blk& b = h.blks[ bi ];
            if ( ranks[ q ] < 0 ) {
               b.type = kDense;  // reached kcap: kept dense
               b.capped = true;
               ++S.n_capped;
               continue;
            }
            b.rank = ranks[ q ];
            if ( ranks[ q ] == 0 ) continue;
            uo[ q ] = store + at;
            vo[ q ] = store + at + (size_t)b.nr * b.rank * esz;
            at += ( (size_t)( b.nr + b.nc ) * b.rank * esz + 255 ) / 256 * 256;
            blk_a[ bi ] = uo[ q ];
            blk_b[ bi ] = vo[ q ];
         }
         B.uo_d.ensure( nj );
         B.vo_d.ensure( nj );
         // the recompressed blocks are written by k_hm_trunc_apply; the rest are copied
         std::vector<int> crank( nj );
         for ( int q = 0; q < nj; ++q ) crank[ q ] = ranks[ q ];
         for ( int q : tq ) crank[ q ] = 0;
         BH_HIP( hipMemcpyAsync( B.uo_d.p, uo.data(), sizeof( void* ) * nj, hipMemcpyHostToDevice, s ) );
         BH_HIP( hipMemcpyAsync( B.vo_d.p, vo.data(), sizeof( void* ) * nj, hipMemcpyHostToDevice, s ) );
         BH_HIP( hipMemcpyAsync( B.rank_d.p, crank.data(), sizeof( int ) * nj, hipMemcpyHostToDevice, s ) );
         if ( opt.single )
            hipLaunchKernelGGL( k_hm_compact<float>, dim3( nj ), dim3( kWg ), 0, s, B.jobs_d.p, B.rank_d.p, B.uo_d.p, B.vo_d.p );
         else
            hipLaunchKernelGGL( k_hm_compact<double>, dim3( nj ), dim3( kWg ), 0, s, B.jobs_d.p, B.rank_d.p, B.uo_d.p, B.vo_d.p );
         if ( !tj.empty() ) {
            const int nt = (int)tj.size();
            std::vector<void*> tuo( nt ), tvo( nt );
            std::vector<int4> tasks;
            for ( int i = 0; i < nt; ++i ) {
               tuo[ i ] = uo[ tq[ i ] ];
               tvo[ i ] = vo[ tq[ i ] ];
               if ( ranks[ tq[ i ] ] <= 0 ) continue;
               for ( int side = 0; side < 2; ++side ) {
                  const int R = side ? tj[ i ].N : tj[ i ].M;
                  for ( int c = 0; c * kTrApply < R; ++c ) tasks.push_back( make_int4( i, side, c, 0 ) );
               }
            }
            B.tuo_d.ensure( nt );
            B.tvo_d.ensure( nt );
            BH_HIP( hipMemcpyAsync( B.tuo_d.p, tuo.data(), sizeof( void* ) * nt, hipMemcpyHostToDevice, s ) );
            BH_HIP( hipMemcpyAsync( B.tvo_d.p, tvo.data(), sizeof( void* ) * nt, hipMemcpyHostToDevice, s ) );
            if ( !tasks.empty() ) {
               B.ttask_d.ensure( tasks.size() );
               BH_HIP( hipMemcpyAsync( B.ttask_d.p, tasks.data(), sizeof( int4 ) * tasks.size(), hipMemcpyHostToDevice, s ) );
               if ( opt.single )
                  hipLaunchKernelGGL( k_hm_trunc_apply<float>, dim3( (unsigned)tasks.size() ), dim3( kTrApply ), 0, s, B.ttask_d.p,
                                      B.tj_d.p, B.trank_d.p, B.tuo_d.p, B.tvo_d.p );
               else
                  hipLaunchKernelGGL( k_hm_trunc_apply<double>, dim3( (unsigned)tasks.size() ), dim3( kTrApply ), 0, s, B.ttask_d.p,
                                      B.tj_d.p, B.trank_d.p, B.tuo_d.p, B.tvo_d.p );
            }
         }
         BH_HIP( hipStreamSynchronize( s ) );
         S.t_trunc += now() - t_tr;
      };
      // the largest blocks first, on their own stream (within the work-space budget)
      Batch big;
      size_t pos = 0, big_bytes = 0;
      while ( pos < lr.size() ) {
         const Blk& b = h.blks[ lr[ pos ] ];
         if ( b.nr + b.nc < kAcaBig || ( !big.idx.empty() && big_bytes + need_of( b ) > opt.ws_bytes ) ) break;
         big_bytes += ( need_of( b ) + 255 ) / 256 * 256;
         big.idx.push_back( (int)pos++ );
      }
      if ( !big.idx.empty() ) {
         BH_HIP( hipEventRecord( ev, st ) );  // the tree-ordered data and the counter are ready
         BH_HIP( hipStreamWaitEvent( st_big, ev, 0 ) );
         launch( big, st_big, st_big );
      }
      // the rest in batches within the budget, each finished before the next.  The batches are
      // cut first and the shared buffers sized for the largest, so that none of them is freed
      // and reallocated while the large blocks run (a free waits for the whole device).
      std::vector<std::vector<int>> cuts;
      size_t ws_max = 0, nj_max = 0, tasks_max = 0, scratch_max = 0;
      while ( pos < lr.size() ) {
         std::vector<int> idx;
         size_t bytes = 0, tasks = 0, scratch = 0;
         while ( pos < lr.size() ) {
            const Blk& b = h.blks[ lr[ pos ] ];
            const size_t need = need_of( b );
            if ( !idx.empty() && bytes + need > opt.ws_bytes ) break;
            bytes += ( need + 255 ) / 256 * 256;
            tasks += ( b.nr + kTrApply - 1 ) / kTrApply + ( b.nc + kTrApply - 1 ) / kTrApply;
            const size_t kp = ( std::min( kcap_of( b ), kTrMaxK ) + 3 ) / 4 * 4;
            scratch += 5 * kp * kp + 2 * kp;
            idx.push_back( (int)pos++ );
         }
         ws_max = std::max( ws_max, bytes );
         nj_max = std::max( nj_max, idx.size() );
         tasks_max = std::max( tasks_max, tasks );
         scratch_max = std::max( scratch_max, scratch );
         cuts.push_back( std::move( idx ) );
      }
      Batch B;
      if ( !cuts.empty() ) {
         B.ws.ensure( ws_max );
         B.jobs_d.ensure( nj_max );
         B.rank_d.ensure( nj_max );
         B.uo_d.ensure( nj_max );
         B.vo_d.ensure( nj_max );
         B.tj_d.ensure( nj_max );
         B.trank_d.ensure( nj_max );
         B.tfail_d.ensure( nj_max );
         B.tuo_d.ensure( nj_max );
         B.tvo_d.ensure( nj_max );
         B.ttask_d.ensure( std::max<size_t>( tasks_max, 1 ) );
         if ( opt.trunc ) B.tr_ws.ensure( scratch_max );
      }
      // The large blocks are finished once the last batch's approximations are queued: their
      // recompression then runs beside those approximations instead of after them.
      bool big_done = big.idx.empty();
      for ( size_t c = 0; c < cuts.size(); ++c ) {
         B.idx = std::move( cuts[ c ] );
         launch( B, st, st_aux );
         if ( c + 1 == cuts.size() && !big_done ) {
            finish( big, st_big, st_aux );
            big_done = true;
         }
         finish( B, st, st_aux );
      }
      if ( !big.idx.empty() ) {
         if ( !big_done ) finish( big, st_big, st_aux );
         BH_HIP( hipEventRecord( ev, st_big ) );
         BH_HIP( hipStreamWaitEvent( st, ev, 0 ) );
      }
   }

   S.kmean_aca = n_aca ? S.kmean_aca / n_aca : 0.0;
   if ( opt.trunc )
      say( "recompression to eps by QR and SVD: %.3f s, mean rank %.1f after the cross approximation, %ld blocks kept as "
           "approximated (dependent columns or rank above %d)",
           S.t_trunc, S.kmean_aca, S.n_trunc_failed, kTrMaxK );

   // dense blocks (inadmissible leaves, small admissible blocks, capped approximations)
   {
      std::vector<DenseJob> dj;
      std::vector<int> di;
      size_t total = 0;
      for ( int i = 0; i < (int)h.blks.size(); ++i )
         if ( h.blks[ i ].type == kDense ) {
            di.push_back( i );
            total += ( (size_t)h.blks[ i ].nr * h.blks[ i ].nc * esz + 255 ) / 256 * 256;
         }
      char* store = static_cast<char*>( h.alloc( total ) );
      size_t at = 0;
      for ( int i : di ) {
         const Blk& b = h.blks[ i ];
         DenseJob j { (int)( b.r0 / 3 ), (int)( b.nr / 3 ), (int)( b.c0 / 3 ), (int)( b.nc / 3 ), store + at };
         blk_a[ i ] = store + at;
         at += ( (size_t)b.nr * b.nc * esz + 255 ) / 256 * 256;
         dj.push_back( j );
      }
      DevBuf<DenseJob> dj_d;
      const int nd = (int)dj.size();
      dj_d.ensure( std::max( nd, 1 ) );
      BH_HIP( hipMemcpyAsync( dj_d.p, dj.data(), sizeof( DenseJob ) * nd, hipMemcpyHostToDevice, st ) );
      if ( nd > 0 ) {
         if ( opt.single )
            hipLaunchKernelGGL( k_hm_dense<float>, dim3( nd ), dim3( kWg ), 0, st, dj_d.p, K, calls.p );
         else
            hipLaunchKernelGGL( k_hm_dense<double>, dim3( nd ), dim3( kWg ), 0, st, dj_d.p, K, calls.p );
      }
      BH_HIP( hipStreamSynchronize( st ) );
   }

   // block descriptors, the low-rank list and the leaf row lists
   std::vector<BlkDev> bd( h.blks.size() );
   std::vector<int> lrl;
   long wofs = 0;
   long kmax = 0;
   double ksum = 0;
   for ( size_t i = 0; i < h.blks.size(); ++i ) {
      const Blk& b = h.blks[ i ];
      BlkDev& d = bd[ i ];
      d.type = b.type;
      d.rank = b.type == kLowRank ? b.rank : 0;
      d.nr = (int)b.nr;
      d.nc = (int)b.nc;
      d.c0 = b.c0;
      d.a = blk_a[ i ];
      d.b = blk_b[ i ];
      d.w = 0;
      if ( b.type == kLowRank ) {
         ++S.n_lowrank;
         S.bytes_lowrank += (double)( b.nr + b.nc ) * b.rank * esz;
         kmax = std::max<long>( kmax, b.rank );
         ksum += b.rank;
         if ( b.rank > 0 ) {
            d.w = wofs;
            wofs += (long)b.rank * kMaxRhs;
            lrl.push_back( (int)i );
         }
      } else {
         if ( b.small ) ++S.n_small;
         else if ( !b.capped ) ++S.n_dense;
         S.bytes_dense += (double)b.nr * b.nc * esz;
      }
   }
   S.kmax = (int)kmax;
   S.kmean = S.n_lowrank ? ksum / S.n_lowrank : 0.0;
   h.n_lr = (int)lrl.size();
   h.w_elems = wofs;
   h.w.ensure( std::max<long>( wofs, 1 ) );
   h.blk_dev.ensure( bd.size() );
   BH_HIP( hipMemcpyAsync( h.blk_dev.p, bd.data(), sizeof( BlkDev ) * bd.size(), hipMemcpyHostToDevice, st ) );
   {
      // tasks of the V^T X product, blocks in the order of the build (largest first)
      std::vector<int4> vt, wt;
      for ( int i : lr ) {
         const Blk& b = h.blks[ i ];
         if ( b.type != kLowRank || b.rank <= 0 ) continue;
         const int nseg = (int)( ( b.nc + kSeg - 1 ) / kSeg ), nl0 = ( b.rank + 63 ) / 64;
         const int first = (int)vt.size();
         for ( int g = 0; g < nseg; ++g )
            for ( int l0 = 0; l0 < nl0; ++l0 ) vt.push_back( make_int4( i, 64 * l0, g, 0 ) );
         for ( int l = 0; l < b.rank; ++l ) wt.push_back( make_int4( i, l, first + l / 64, nseg ) );
      }
      h.n_vtx = (int)vt.size();
      h.n_w = (int)wt.size();
      h.vtx_tasks.ensure( std::max<size_t>( vt.size(), 1 ) );
      h.w_tasks.ensure( std::max<size_t>( wt.size(), 1 ) );
      h.vtx_part.ensure( std::max<size_t>( vt.size(), 1 ) * 64 * kMaxRhs );
      BH_HIP( hipMemcpyAsync( h.vtx_tasks.p, vt.data(), sizeof( int4 ) * vt.size(), hipMemcpyHostToDevice, st ) );
      BH_HIP( hipMemcpyAsync( h.w_tasks.p, wt.data(), sizeof( int4 ) * wt.size(), hipMemcpyHostToDevice, st ) );
      BH_HIP( hipStreamSynchronize( st ) );
   }

   // per leaf cluster: the blocks covering its rows, in block order (fixed)
   h.nleaves = T.nleaves();
   std::vector<std::vector<std::pair<int, int>>> rows( h.nleaves );
   std::vector<int> diag( h.nleaves, -1 );
   for ( size_t i = 0; i < h.blks.size(); ++i ) {
      const Blk& b = h.blks[ i ];
      int l0 = 0, nlf = 0;
      T.leaf_range( b.t, l0, nlf );
      for ( int l = l0; l < l0 + nlf; ++l ) rows[ l ].push_back( { (int)i, 3 * ( T.beg[ T.leaf_cluster( l ) ] - T.beg[ b.t ] ) } );
      if ( b.t == b.s && T.is_leaf( b.t ) ) diag[ b.t - T.first_leaf() ] = (int)i;
   }
   std::vector<int> r0v( h.nleaves ), nrv( h.nleaves ), ptrv( h.nleaves + 1, 0 ), eb, eo;
   std::vector<long> poff( h.nleaves );
   long ptot = 0;
   for ( int l = 0; l < h.nleaves; ++l ) {
      r0v[ l ] = 3 * T.beg[ T.leaf_cluster( l ) ];
      nrv[ l ] = 3 * T.cnt[ T.leaf_cluster( l ) ];
      for ( auto& e : rows[ l ] ) eb.push_back( e.first ), eo.push_back( e.second );
      ptrv[ l + 1 ] = (int)eb.size();
      poff[ l ] = ptot;
      ptot += (long)nrv[ l ] * nrv[ l ];
      if ( diag[ l ] < 0 ) throw std::runtime_error( "H-matrix: a leaf has no dense diagonal block" );
   }
   h.leaf_r0.ensure( h.nleaves );
   h.leaf_nr.ensure( h.nleaves );
   h.ptr.ensure( h.nleaves + 1 );
   h.ent_blk.ensure( eb.size() );
   h.ent_off.ensure( eo.size() );
   h.diag_blk.ensure( h.nleaves );
   h.pinv_off.ensure( h.nleaves );
   h.pinv.ensure( ptot );
   BH_HIP( hipMemcpyAsync( h.leaf_r0.p, r0v.data(), sizeof( int ) * h.nleaves, hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.leaf_nr.p, nrv.data(), sizeof( int ) * h.nleaves, hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.ptr.p, ptrv.data(), sizeof( int ) * ( h.nleaves + 1 ), hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.ent_blk.p, eb.data(), sizeof( int ) * eb.size(), hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.ent_off.p, eo.data(), sizeof( int ) * eo.size(), hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.diag_blk.p, diag.data(), sizeof( int ) * h.nleaves, hipMemcpyHostToDevice, st ) );
   BH_HIP( hipMemcpyAsync( h.pinv_off.p, poff.data(), sizeof( long ) * h.nleaves, hipMemcpyHostToDevice, st ) );
   unsigned long long hcalls = 0;
   BH_HIP( hipMemcpyAsync( &hcalls, calls.p, sizeof hcalls, hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipStreamSynchronize( st ) );
   S.kernel_blocks = (double)hcalls;
   S.t_fill = now() - t0;
   const double dense_bytes = 72.0 * n * (double)n;
   say( "H-matrix in %.2f s: %.2f GB (%.0f KB per triangle, %.2f%% of dense), %ld dense and %ld low-rank blocks (%ld small "
        "admissible and %ld at the rank cap kept dense), rank max %d mean %.1f",
        S.t_fill, ( S.bytes_dense + S.bytes_lowrank ) / 1e9, ( S.bytes_dense + S.bytes_lowrank ) / n / 1e3,
        100.0 * ( S.bytes_dense + S.bytes_lowrank ) / dense_bytes, S.n_dense, S.n_lowrank, S.n_small, S.n_capped, S.kmax,
        S.kmean );
   say( "H-matrix: eps %.0e, eta %.1f, %.3g kernel blocks per N^2, %s precision storage", opt.eps, opt.eta,
        S.kernel_blocks / ( (double)n * n ), opt.single ? "single" : "double" );

   // the preconditioner: inverses of the diagonal leaf blocks
   t0 = now();
   DevBuf<int> fail;
   fail.ensure( 1 );
   BH_HIP( hipMemsetAsync( fail.p, 0, sizeof( int ), st ) );
   if ( opt.single )
      hipLaunchKernelGGL( k_hm_diag_inv<float>, dim3( h.nleaves ), dim3( kWg ), 0, st, h.diag_blk.p, h.blk_dev.p, h.leaf_r0.p,
                          h.xmu_t.p, h.comp_t.p, h.gdef.p, h.ndef, h.pinv_off.p, h.pinv.p, fail.p );
   else
      hipLaunchKernelGGL( k_hm_diag_inv<double>, dim3( h.nleaves ), dim3( kWg ), 0, st, h.diag_blk.p, h.blk_dev.p, h.leaf_r0.p,
                          h.xmu_t.p, h.comp_t.p, h.gdef.p, h.ndef, h.pinv_off.p, h.pinv.p, fail.p );
   int hf = 0;
   BH_HIP( hipMemcpyAsync( &hf, fail.p, sizeof hf, hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipStreamSynchronize( st ) );
   if ( hf ) throw std::runtime_error( "H-matrix: a diagonal leaf block is singular" );
   h.K = K;
   if ( opt.precond == 1 || opt.direct ) {
      const double e = opt.direct ? opt.eps : opt.eps_pre;
      if ( hd_build( h, st, K, opt, e, log ) ) {
         S.eps_pre = h.hd_eps;
         S.piv_ratio = h.hd_piv_ratio;
         S.pre_levels = h.hd_depth;
         for ( int r : h.hd_r ) S.pre_rank_max = std::max( S.pre_rank_max, r );
         S.pre_bytes = 2.0 * 8.0 * h.n3 * (double)h.hd_cols;
      } else if ( opt.direct ) {
         throw std::runtime_error( "H-LU: the factorization broke down; without -hlu BEST solves by GMRES" );
      } else {
         say( "%s", "preconditioner: block Jacobi on the diagonal leaf blocks alone" );
      }
   }
   S.t_pre = now() - t0;
   say( "preconditioner: %s, %.3f s", h.hd_on ? "the hierarchical factorization on the block-Jacobi leaf inverses"
                                               : "block Jacobi on the diagonal leaf blocks", S.t_pre );
   {
      size_t freeb = 0, total = 0;
      BH_HIP( hipMemGetInfo( &freeb, &total ) );
      S.mem_used = S.mem_free - (double)freeb;
   }
   h.defl_s.ensure( (size_t)h.ndef * kMaxRhs );
   h.defl_part.ensure( (size_t)( ( h.n3 + kDotChunkDefl - 1 ) / kDotChunkDefl ) * h.ndef * kMaxRhs );
   h.ypart.ensure( (size_t)h.nleaves * kSplit * kRowsWg * kMaxRhs );
   if ( stats ) *stats = S;
   return H;
}

long hm_blocks( const HMatrix& h, long* out, long cap, int* perm ) {
   const long nb = (long)h.blks.size();
   if ( out )
      for ( long i = 0; i < std::min( nb, cap ); ++i ) {
         const Blk& b = h.blks[ i ];
         long* o = out + 6 * i;
         o[ 0 ] = b.type == kLowRank ? 1 : 0;
         o[ 1 ] = b.type == kLowRank ? b.rank : 0;
         o[ 2 ] = b.r0, o[ 3 ] = b.nr, o[ 4 ] = b.c0, o[ 5 ] = b.nc;
      }
   if ( perm ) std::copy( h.T.perm.begin(), h.T.perm.end(), perm );
   return nb;
}

void hm_apply( HMatrix& h, hipStream_t st, int p, const double* d_x, long ldx, double* d_y, long ldy ) {
   if ( p < 1 || p > kMaxRhs ) throw std::runtime_error( "H-matrix product: 1 to 16 columns" );
   const size_t nn = (size_t)h.n3 * p;
   h.Xk.ensure( nn );
   h.Wk.ensure( nn );
   hipLaunchKernelGGL( k_hm_reorder, dim3( nblocks( nn, 256 ) ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 1, d_x, ldx, h.Xk.p,
                       (long)h.n3 );
   h.apply( st, p, h.Xk.p, h.Wk.p );
   hipLaunchKernelGGL( k_hm_reorder, dim3( nblocks( nn, 256 ) ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 0, h.Wk.p, (long)h.n3,
                       d_y, ldy );
   BH_HIP( hipStreamSynchronize( st ) );
}

double hm_time_product( HMatrix& h, hipStream_t st, int p, int reps, double* bytes ) {
   if ( p < 1 || p > kMaxRhs ) throw std::runtime_error( "H-matrix product: 1 to 16 columns" );
   const size_t nn = (size_t)h.n3 * p;
   h.Xk.ensure( nn );
   h.Wk.ensure( nn );
   std::vector<double> ones( nn, 1.0 );
   BH_HIP( hipMemcpyAsync( h.Xk.p, ones.data(), nn * sizeof( double ), hipMemcpyHostToDevice, st ) );
   h.apply( st, p, h.Xk.p, h.Wk.p );  // warm
   hipEvent_t e0, e1;
   BH_HIP( hipEventCreate( &e0 ) );
   BH_HIP( hipEventCreate( &e1 ) );
   BH_HIP( hipEventRecord( e0, st ) );
   for ( int r = 0; r < reps; ++r ) h.apply( st, p, h.Xk.p, h.Wk.p );
   BH_HIP( hipEventRecord( e1, st ) );
   BH_HIP( hipEventSynchronize( e1 ) );
   float ms = 0.f;
   BH_HIP( hipEventElapsedTime( &ms, e0, e1 ) );
   (void)hipEventDestroy( e0 );
   (void)hipEventDestroy( e1 );
   if ( bytes ) {
      const size_t esz = h.single ? sizeof( float ) : sizeof( double );
      double b = 0.0;
      for ( const Blk& k : h.blks ) b += k.type == kLowRank ? (double)( k.nr + k.nc ) * k.rank * esz : (double)k.nr * k.nc * esz;
      *bytes = b + 2.0 * nn * sizeof( double );
   }
   return ms / reps;
}

namespace {

// Restarted GMRES (hmatrix/src/gmres.cpp:26-178) on the built H-matrix with its preconditioner;
// returns false if it did not reach the tolerance within maxsteps (the solution so far is in d_x).
bool hm_gmres( HMatrix& h, hipStream_t st, int p, const double* d_b, long ldb, double* d_x, long ldx, HmStats* stats,
               std::vector<std::string>* log ) {
   const double t_start = now();
   const int n3 = h.n3, m = h.opt.restart;
   const double tol = h.opt.tol > 0 ? h.opt.tol : h.opt.aca_factor * h.opt.eps;
   const size_t nn = (size_t)n3 * p;
   const long vstride = (long)nn;
   h.Vb.ensure( nn * ( m + 1 ) );
   h.Wk.ensure( nn );
   h.Zk.ensure( nn );
   h.Xk.ensure( nn );
   h.Bk.ensure( nn );
   h.Rk.ensure( nn );
   h.part.ensure( (size_t)( ( n3 + kDotChunk - 1 ) / kDotChunk ) * ( m + 1 ) * kMaxRhs );
   h.hbuf.ensure( (size_t)( m + 1 ) * kMaxRhs );
   DevBuf<double> sc;
   sc.ensure( kMaxRhs );
   const unsigned g = nblocks( nn, 256 );
   hipLaunchKernelGGL( k_hm_reorder, dim3( g ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 1, d_b, ldb, h.Bk.p, (long)n3 );
   BH_HIP( hipMemsetAsync( h.Xk.p, 0, sizeof( double ) * nn, st ) );
   BH_HIP( hipMemcpyAsync( h.Rk.p, h.Bk.p, sizeof( double ) * nn, hipMemcpyDeviceToDevice, st ) );
   std::vector<double> bnorm( p ), relres( p, 1.0 ), tmp( ( m + 1 ) * p );
   std::vector<int> iters( p, 0 );
   std::vector<char> done( p, 0 );
   h.dots( st, p, 1, h.Bk.p, 0, h.Bk.p, tmp.data() );
   for ( int q = 0; q < p; ++q ) {
      bnorm[ q ] = std::sqrt( tmp[ q ] );
      done[ q ] = bnorm[ q ] == 0.0;
      relres[ q ] = done[ q ] ? 0.0 : 1.0;
   }
   int total = 0;
   std::vector<double> H( (size_t)( m + 1 ) * m * p ), gv( ( m + 1 ) * p ), cs( m * p ), sn( m * p ), scale( p );
   std::vector<int> jlast( p );
   std::vector<char> going( p );
   auto Hat = [&]( int i, int j, int q ) -> double& { return H[ ( (size_t)q * m + j ) * ( m + 1 ) + i ]; };
   while ( total < h.opt.maxsteps ) {
      bool any = false;
      for ( int q = 0; q < p; ++q ) any |= !done[ q ];
      if ( !any ) break;
      // V_0 = R / |R| for the active columns, 0 for the others
      h.dots( st, p, 1, h.Rk.p, 0, h.Rk.p, tmp.data() );
      for ( int q = 0; q < p; ++q ) {
         const double beta = std::sqrt( tmp[ q ] );
         going[ q ] = !done[ q ] && beta > 0.0;
         scale[ q ] = going[ q ] ? 1.0 / beta : 0.0;
         std::fill( gv.begin() + q * ( m + 1 ), gv.begin() + ( q + 1 ) * ( m + 1 ), 0.0 );
         gv[ q * ( m + 1 ) ] = going[ q ] ? beta : 0.0;
         jlast[ q ] = -1;
      }
      BH_HIP( hipMemcpyAsync( sc.p, scale.data(), sizeof( double ) * p, hipMemcpyHostToDevice, st ) );
      hipLaunchKernelGGL( k_hm_scale, dim3( g ), dim3( 256 ), 0, st, n3, p, h.Rk.p, sc.p, h.Vb.p );
      for ( int j = 0; j < m && total < h.opt.maxsteps; ++j ) {
         bool cyc = false;
         for ( int q = 0; q < p; ++q ) cyc |= going[ q ];
         if ( !cyc ) break;
         double* Vj = h.Vb.p + (size_t)j * vstride;
         h.precond( st, p, Vj, h.Zk.p );
         h.apply( st, p, h.Zk.p, h.Wk.p );
         ++total;
         // classical Gram-Schmidt, twice
         std::vector<double> hc( ( j + 1 ) * p, 0.0 );
         for ( int pass = 0; pass < 2; ++pass ) {
            h.dots( st, p, j + 1, h.Vb.p, vstride, h.Wk.p, tmp.data() );
            for ( int i = 0; i < ( j + 1 ) * p; ++i ) hc[ i ] += tmp[ i ];
            BH_HIP( hipMemcpyAsync( h.hbuf.p, tmp.data(), sizeof( double ) * ( j + 1 ) * p, hipMemcpyHostToDevice, st ) );
            hipLaunchKernelGGL( k_hm_combine, dim3( g ), dim3( 256 ), 0, st, n3, p, j + 1, h.Vb.p, vstride, h.hbuf.p, -1.0,
                                h.Wk.p );
         }
         h.dots( st, p, 1, h.Wk.p, 0, h.Wk.p, tmp.data() );
         for ( int q = 0; q < p; ++q ) {
            const double hn = std::sqrt( tmp[ q ] );
            scale[ q ] = going[ q ] && hn > 0.0 ? 1.0 / hn : 0.0;
            if ( !going[ q ] ) continue;
            for ( int i = 0; i <= j; ++i ) Hat( i, j, q ) = hc[ i * p + q ];
            Hat( j + 1, j, q ) = hn;
            double* c = cs.data() + q * m;
            double* s = sn.data() + q * m;
            for ( int i = 0; i < j; ++i ) {
               const double t = c[ i ] * Hat( i, j, q ) + s[ i ] * Hat( i + 1, j, q );
               Hat( i + 1, j, q ) = -s[ i ] * Hat( i, j, q ) + c[ i ] * Hat( i + 1, j, q );
               Hat( i, j, q ) = t;
            }
            const double rr = std::hypot( Hat( j, j, q ), Hat( j + 1, j, q ) );
            c[ j ] = rr > 0 ? Hat( j, j, q ) / rr : 1.0;
            s[ j ] = rr > 0 ? Hat( j + 1, j, q ) / rr : 0.0;
            Hat( j, j, q ) = rr;
            Hat( j + 1, j, q ) = 0.0;
            double* gq = gv.data() + q * ( m + 1 );
            gq[ j + 1 ] = -s[ j ] * gq[ j ];
            gq[ j ] = c[ j ] * gq[ j ];
            jlast[ q ] = j;
            ++iters[ q ];
            if ( std::fabs( gq[ j + 1 ] ) <= tol * bnorm[ q ] || hn == 0.0 ) going[ q ] = 0;
         }
         BH_HIP( hipMemcpyAsync( sc.p, scale.data(), sizeof( double ) * p, hipMemcpyHostToDevice, st ) );
         hipLaunchKernelGGL( k_hm_scale, dim3( g ), dim3( 256 ), 0, st, n3, p, h.Wk.p, sc.p, h.Vb.p + (size_t)( j + 1 ) * vstride );
      }
      // x += M^-1 V y for every column of this cycle
      int kmaxc = 0;
      for ( int q = 0; q < p; ++q ) kmaxc = std::max( kmaxc, jlast[ q ] + 1 );
      std::vector<double> yall( (size_t)std::max( kmaxc, 1 ) * p, 0.0 );
      for ( int q = 0; q < p; ++q ) {
         const int k = jlast[ q ] + 1;
         for ( int i = k - 1; i >= 0; --i ) {
            double s = gv[ q * ( m + 1 ) + i ];
            for ( int l = i + 1; l < k; ++l ) s -= Hat( i, l, q ) * yall[ l * p + q ];
            yall[ i * p + q ] = s / Hat( i, i, q );
         }
      }
      if ( kmaxc > 0 ) {
         BH_HIP( hipMemcpyAsync( h.hbuf.p, yall.data(), sizeof( double ) * kmaxc * p, hipMemcpyHostToDevice, st ) );
         BH_HIP( hipMemsetAsync( h.Wk.p, 0, sizeof( double ) * nn, st ) );
         hipLaunchKernelGGL( k_hm_combine, dim3( g ), dim3( 256 ), 0, st, n3, p, kmaxc, h.Vb.p, vstride, h.hbuf.p, 1.0, h.Wk.p );
         h.precond( st, p, h.Wk.p, h.Zk.p );
         hipLaunchKernelGGL( k_hm_axpy, dim3( g ), dim3( 256 ), 0, st, (long)nn, h.Xk.p, 1.0, h.Zk.p, h.Xk.p );
      }
      // the true residual
      h.apply( st, p, h.Xk.p, h.Wk.p );
      hipLaunchKernelGGL( k_hm_axpy, dim3( g ), dim3( 256 ), 0, st, (long)nn, h.Bk.p, -1.0, h.Wk.p, h.Rk.p );
      h.dots( st, p, 1, h.Rk.p, 0, h.Rk.p, tmp.data() );
      for ( int q = 0; q < p; ++q ) {
         if ( bnorm[ q ] == 0.0 ) continue;
         relres[ q ] = std::sqrt( tmp[ q ] ) / bnorm[ q ];
         done[ q ] = relres[ q ] <= tol;
      }
   }
   int imin = 1 << 30, imax = 0;
   double worst = 0.0;
   for ( int q = 0; q < p; ++q ) {
      imin = std::min( imin, iters[ q ] );
      imax = std::max( imax, iters[ q ] );
      if ( std::isnan( relres[ q ] ) || relres[ q ] > worst ) worst = relres[ q ];
   }
   hipLaunchKernelGGL( k_hm_reorder, dim3( g ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 0, h.Xk.p, (long)n3, d_x, ldx );
   BH_HIP( hipStreamSynchronize( st ) );
   const double el = now() - t_start;
   if ( log ) {
      char buf[ 256 ];
      std::snprintf( buf, sizeof buf,
                     "solve: %d right-hand sides by GMRES, %d to %d iterations, largest relative residual %.1e, %.3f s", p,
                     imin, imax, worst, el );
      log->emplace_back( buf );
   }
   if ( stats ) {
      stats->iters_min = imin;
      stats->iters_max = imax;
      stats->relres = worst;
      stats->steps = total;
      stats->t_solve = el;
   }
   if ( !( worst <= tol ) ) {
      if ( log ) {
         char buf[ 160 ];
         std::snprintf( buf, sizeof buf, "GMRES did not reach a relative residual of %.0e in %d steps (%.1e)", tol,
                        h.opt.maxsteps, worst );
         log->emplace_back( buf );
      }
      return false;
   }
   return true;
}

// -hlu: x = M^-1 b with the hierarchical factorization at eps, refined against the H-matrix
// (x += M^-1 (b - A x)) until the relative residual is below the tolerance; no Krylov method.
// The reference's direct mode (hmatrix/src/api.cpp:461-495) solves with its H-LU factors once.
constexpr int kDirectSteps = 40;
void hm_direct( HMatrix& h, hipStream_t st, int p, const double* d_b, long ldb, double* d_x, long ldx, HmStats* stats,
                std::vector<std::string>* log ) {
   const double t_start = now();
   const int n3 = h.n3;
   const double tol = h.opt.tol > 0 ? h.opt.tol : h.opt.aca_factor * h.opt.eps;
   const size_t nn = (size_t)n3 * p;
   h.Wk.ensure( nn );
   h.Zk.ensure( nn );
   h.Xk.ensure( nn );
   h.Bk.ensure( nn );
   h.Rk.ensure( nn );
   h.part.ensure( (size_t)( ( n3 + kDotChunk - 1 ) / kDotChunk ) * kMaxRhs );
   h.hbuf.ensure( kMaxRhs );
   const unsigned g = nblocks( nn, 256 );
   hipLaunchKernelGGL( k_hm_reorder, dim3( g ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 1, d_b, ldb, h.Bk.p, (long)n3 );
   std::vector<double> bn( p ), rn( p );
   h.dots( st, p, 1, h.Bk.p, 0, h.Bk.p, bn.data() );
   h.precond( st, p, h.Bk.p, h.Xk.p );
   int steps = 0;
   double worst = 0.0;
   for ( ;; ) {
      h.apply( st, p, h.Xk.p, h.Wk.p );
      hipLaunchKernelGGL( k_hm_axpy, dim3( g ), dim3( 256 ), 0, st, (long)nn, h.Bk.p, -1.0, h.Wk.p, h.Rk.p );
      h.dots( st, p, 1, h.Rk.p, 0, h.Rk.p, rn.data() );
      worst = 0.0;
      for ( int q = 0; q < p; ++q ) {
         const double r = bn[ q ] > 0 ? std::sqrt( rn[ q ] / bn[ q ] ) : 0.0;
         worst = ( std::isnan( r ) || r > worst ) ? r : worst;
      }
      if ( worst <= tol || steps >= kDirectSteps || !std::isfinite( worst ) ) break;
      h.precond( st, p, h.Rk.p, h.Zk.p );
      hipLaunchKernelGGL( k_hm_axpy, dim3( g ), dim3( 256 ), 0, st, (long)nn, h.Xk.p, 1.0, h.Zk.p, h.Xk.p );
      ++steps;
   }
   hipLaunchKernelGGL( k_hm_reorder, dim3( g ), dim3( 256 ), 0, st, h.n, p, h.perm.p, 0, h.Xk.p, (long)n3, d_x, ldx );
   BH_HIP( hipStreamSynchronize( st ) );
   const double el = now() - t_start;
   if ( log ) {
      char buf[ 256 ];
      std::snprintf( buf, sizeof buf,
                     "solve: %d right-hand sides with the hierarchical factorization at eps %.0e (-hlu), %d refinement steps, "
                     "largest relative residual %.1e, %.3f s",
                     p, h.hd_eps, steps, worst, el );
      log->emplace_back( buf );
   }
   if ( stats ) {
      stats->iters_min = stats->iters_max = steps;
      stats->relres = worst;
      stats->steps = steps;
      stats->t_solve = el;
   }
   if ( !( worst <= tol ) ) {
      char buf[ 200 ];
      std::snprintf( buf, sizeof buf,
                     "H-LU: the factorization at eps %.0e does not reach a relative residual of %.0e (%.1e after %d refinement "
                     "steps); without -hlu BEST solves by GMRES",
                     h.hd_eps, tol, worst, steps );
      throw std::runtime_error( buf );
   }
}

constexpr double kPreMin = 1e-3;  // the finest preconditioner the retries build (hmatrix/src/api.cpp:554-578)

}  // namespace

void hm_solve( HMatrix& h, hipStream_t st, int p, const double* d_b, long ldb, double* d_x, long ldx, HmStats* stats,
               std::vector<std::string>* log ) {
   if ( p < 1 || p > kMaxRhs ) throw std::runtime_error( "H-matrix solve: 1 to 16 right-hand sides per call" );
   if ( h.opt.direct ) {
      hm_direct( h, st, p, d_b, ldb, d_x, ldx, stats, log );
      return;
   }
   // GMRES; if it does not converge, a stronger preconditioner is built and GMRES runs again: the
   // hierarchical factorization at eps_pre after block Jacobi, then that factorization 10 times
   // finer (down to 1e-3), as the reference's finer() does (hmatrix/src/api.cpp:554-578, :744-763)
   for ( ;; ) {
      if ( hm_gmres( h, st, p, d_b, ldb, d_x, ldx, stats, log ) ) return;
      if ( !h.hd_on ) {
         if ( log ) log->emplace_back( "GMRES preconditioner: block Jacobi was not enough; building the hierarchical "
                                       "factorization" );
         if ( !hd_build( h, st, h.K, h.opt, h.opt.eps_pre, log ) ) throw std::runtime_error( "GMRES preconditioner: broke down" );
         if ( stats ) stats->eps_pre = h.hd_eps, stats->piv_ratio = h.hd_piv_ratio;
         continue;
      }
      if ( h.hd_eps <= kPreMin * 1.001 ) {
         char buf[ 160 ];
         std::snprintf( buf, sizeof buf, "GMRES did not reach its relative residual in %d steps (%.1e)", h.opt.maxsteps,
                        stats ? stats->relres : 0.0 );
         throw std::runtime_error( buf );
      }
      const double next = h.hd_eps / 10;
      if ( log ) {
         char buf[ 96 ];
         std::snprintf( buf, sizeof buf, "GMRES preconditioner: rebuilding it at eps %.0e", next );
         log->emplace_back( buf );
      }
      if ( !hd_build( h, st, h.K, h.opt, next, log ) ) throw std::runtime_error( "GMRES preconditioner: broke down" );
      if ( stats ) stats->eps_pre = h.hd_eps, stats->piv_ratio = h.hd_piv_ratio;
   }
}


}  // namespace bhg
