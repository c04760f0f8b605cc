// COMPON on the device: the connected surfaces of the triangulation and its open edges
// (best/compon.f90:19-115), from vertices resident in HBM.
//
//   k_vertex_ids   corners with the same three coordinates (compared as reals, so -0 equals 0;
//                  a corner with a NaN coordinate matches nothing, as in the reference) get one
//                  id: the corner that claimed its slot in an open-addressing hash table
//   k_unite        union-find over triangles: each corner's triangle is joined to the triangle
//                  of its vertex's representative corner, by lock-free hooking of the larger root
//                  under the smaller (compare-and-swap), so that every root is the smallest
//                  triangle of its surface
//   k_roots, k_label_scan, k_label   ICOMP in 1..NCOMP numbered in the order of the surfaces'
//                  first triangles (an in-order scan of the root flags)
//   k_edges, k_open_count   the sides (smaller id, larger id) counted in a second hash table;
//                  NOPEN = the distinct sides not shared by exactly two triangles
//
// The reference sorts the corners (MSORT3) and runs a sequential union-find; the ids here are
// other numbers, but only their equality is used, so ICOMP, NCOMP and NOPEN are the same.
#include <stdexcept>

#include "bh_common.h"
#include "compon.h"

namespace bhg {
namespace {

constexpr unsigned long long kEmptyEdge = ~0ull;
constexpr int kMaxProbe = 1 << 20;  // bound on every probe and retry loop (a full table is an error)

__device__ inline double canon( double x ) { return x == 0.0 ? 0.0 : x; }  // -0 -> +0

__device__ inline unsigned long long mix64( unsigned long long h ) {
   h ^= h >> 33;
   h *= 0xff51afd7ed558ccdull;
   h ^= h >> 33;
   h *= 0xc4ceb9fe1a85ec53ull;
   h ^= h >> 33;
   return h;
}

// corner c = 3 t + k is vertex k (v1, v2, v3) of triangle t; v holds v1, v2, v3 as NAR x 3
// column-major arrays one after the other
__device__ inline void corner( const double* __restrict__ v, int n, int c, double& x, double& y, double& z ) {
   const int t = c / 3, k = c % 3;
   const double* b = v + (size_t)k * 3 * n;
   x = canon( b[ t ] ), y = canon( b[ n + t ] ), z = canon( b[ 2 * (size_t)n + t ] );
}

__global__ void k_fill_int( int* p, long n, int val ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) p[ i ] = val;
}

__global__ void k_fill_edges( unsigned long long* key, int* cnt, long n ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) {
      key[ i ] = kEmptyEdge;
      cnt[ i ] = 0;
   }
}

__global__ void k_vertex_ids( int n, const double* __restrict__ v, int* __restrict__ table, int hmask, int* __restrict__ vid,
                              int* __restrict__ err ) {
   const int nv = 3 * n;
   for ( int c = blockIdx.x * blockDim.x + threadIdx.x; c < nv; c += gridDim.x * blockDim.x ) {
      double x, y, z;
      corner( v, n, c, x, y, z );
      if ( x != x || y != y || z != z ) {  // NaN: equal to nothing
         vid[ c ] = c;
         continue;
      }
      const unsigned long long bx = __double_as_longlong( x ), by = __double_as_longlong( y ), bz = __double_as_longlong( z );
      unsigned h = (unsigned)mix64( bx ^ mix64( by ^ mix64( bz ) ) ) & hmask;
      int id = -1;
      for ( int probe = 0; probe < kMaxProbe; ++probe, h = ( h + 1 ) & hmask ) {
         int cur = atomicCAS( &table[ h ], -1, c );
         if ( cur == -1 ) {
            id = c;
            break;
         }
         double ox, oy, oz;
         corner( v, n, cur, ox, oy, oz );
         if ( ox == x && oy == y && oz == z ) {
            id = cur;
            break;
         }
      }
      if ( id < 0 ) {
         atomicOr( err, 1 );
         id = c;
      }
      vid[ c ] = id;
   }
}

// parent[] is read and written only by device-scope atomics: workgroups on different XCDs (each
// with its own L2) then never work from a stale copy.
__device__ inline int ld_parent( int* parent, int t ) {
   return __hip_atomic_load( &parent[ t ], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
}

__device__ inline int find_root( int* parent, int t ) {
   int p = ld_parent( parent, t );
   for ( int guard = 0; guard < kMaxProbe && p != t; ++guard ) {
      const int q = ld_parent( parent, p );
      // path halving: t's parent becomes its grandparent (an ancestor; parents only decrease)
      if ( q != p ) __hip_atomic_store( &parent[ t ], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
      t = p;
      p = q;
   }
   return t;
}

__global__ void k_unite( int n, const int* __restrict__ vid, int* parent, int* __restrict__ err ) {
   const int nv = 3 * n;
   for ( int c = blockIdx.x * blockDim.x + threadIdx.x; c < nv; c += gridDim.x * blockDim.x ) {
      const int r = vid[ c ];
      if ( r == c ) continue;
      int a = c / 3, b = r / 3;
      bool done = false;
      for ( int guard = 0; guard < kMaxProbe; ++guard ) {
         a = find_root( parent, a );
         b = find_root( parent, b );
         if ( a == b ) {
            done = true;
            break;
         }
         if ( a > b ) {
            const int t = a;
            a = b;
            b = t;
         }
         // hook the larger root b under a, if b is still a root
         if ( atomicCAS( &parent[ b ], b, a ) == b ) {
            done = true;
            break;
         }
      }
      if ( !done ) atomicOr( err, 2 );
   }
}

__global__ void k_roots( int n, int* parent, int* __restrict__ root, int* __restrict__ flag ) {
   for ( int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x ) {
      const int r = find_root( parent, t );
      root[ t ] = r;
      flag[ t ] = r == t ? 1 : 0;
   }
}

// In-order exclusive scan of flag[0..n) into lab (one workgroup of 1024 threads, chunk by chunk);
// *total = the number of roots.
__global__ void __launch_bounds__( 1024 ) k_label_scan( int n, const int* __restrict__ flag, int* __restrict__ lab,
                                                        int* __restrict__ total ) {
   __shared__ int part[ 1024 ];
   __shared__ int carry;
   if ( threadIdx.x == 0 ) carry = 0;
   __syncthreads();
   for ( int base = 0; base < n; base += 1024 ) {
      const int t = base + threadIdx.x;
      const int f = t < n ? flag[ t ] : 0;
      part[ threadIdx.x ] = f;
      __syncthreads();
      for ( int off = 1; off < 1024; off <<= 1 ) {  // inclusive Hillis-Steele scan
         const int add = threadIdx.x >= off ? part[ threadIdx.x - off ] : 0;
         __syncthreads();
         part[ threadIdx.x ] += add;
         __syncthreads();
      }
      if ( t < n ) lab[ t ] = carry + part[ threadIdx.x ] - f;
      __syncthreads();
      if ( threadIdx.x == 0 ) carry += part[ 1023 ];
      __syncthreads();
   }
   if ( threadIdx.x == 0 ) *total = carry;
}

__global__ void k_label( int n, const int* __restrict__ root, const int* __restrict__ lab, int* __restrict__ icomp ) {
   for ( int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x ) icomp[ t ] = lab[ root[ t ] ] + 1;
}

__global__ void k_edges( int n, const int* __restrict__ vid, unsigned long long* key, int* cnt, int hmask, int* __restrict__ err ) {
   const int ns = 3 * n;
   for ( int e = blockIdx.x * blockDim.x + threadIdx.x; e < ns; e += gridDim.x * blockDim.x ) {
      const int t = e / 3, c = e % 3;
      const int a = vid[ 3 * t + c ], b = vid[ 3 * t + ( c + 1 ) % 3 ];
      const unsigned long long k = ( (unsigned long long)(unsigned)min( a, b ) << 32 ) | (unsigned)max( a, b );
      unsigned h = (unsigned)mix64( k ) & hmask;
      bool done = false;
      for ( int probe = 0; probe < kMaxProbe; ++probe, h = ( h + 1 ) & hmask ) {
         const unsigned long long cur = atomicCAS( &key[ h ], kEmptyEdge, k );
         if ( cur == kEmptyEdge || cur == k ) {
            atomicAdd( &cnt[ h ], 1 );
            done = true;
            break;
         }
      }
      if ( !done ) atomicOr( err, 4 );
   }
}

__global__ void k_open_count( const unsigned long long* __restrict__ key, const int* __restrict__ cnt, long hsize,
                              int* __restrict__ nopen ) {
   int mine = 0;
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < hsize; i += (long)gridDim.x * blockDim.x )
      mine += ( key[ i ] != kEmptyEdge && cnt[ i ] != 2 ) ? 1 : 0;
   if ( mine ) atomicAdd( nopen, mine );
}

__global__ void k_iota( int n, int* p ) {
   for ( int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x ) p[ t ] = t;
}

int pow2_at_least( long v ) {
   int p = 1;
   while ( p < v ) p <<= 1;
   return p;
}

}  // namespace

void ComponWork::ensure( int n ) {
   const long nv = 3L * n;
   if ( nv > ( 1L << 29 ) ) throw std::runtime_error( "COMPON on the device: too many triangles" );
   hsize = pow2_at_least( 2 * nv );
   vtable.ensure( hsize );
   vid.ensure( nv );
   parent.ensure( n );
   root.ensure( n );
   flag.ensure( n );
   lab.ensure( n );
   ekey.ensure( hsize );
   ecnt.ensure( hsize );
   out.ensure( 4 );
}

void device_compon( hipStream_t st, int n, const double* d_v, int* d_icomp, ComponWork& w, int* ncomp, int* nopen ) {
   w.ensure( n );
   const int nv = 3 * n, hmask = w.hsize - 1;
   const int g = 1024, b = 256;
   hipLaunchKernelGGL( k_fill_int, dim3( g ), dim3( b ), 0, st, w.vtable.p, (long)w.hsize, -1 );
   hipLaunchKernelGGL( k_fill_edges, dim3( g ), dim3( b ), 0, st, w.ekey.p, w.ecnt.p, (long)w.hsize );
   hipLaunchKernelGGL( k_fill_int, dim3( 1 ), dim3( 64 ), 0, st, w.out.p, 4L, 0 );
   hipLaunchKernelGGL( k_vertex_ids, dim3( ( nv + b - 1 ) / b ), dim3( b ), 0, st, n, d_v, w.vtable.p, hmask, w.vid.p,
                       w.out.p + 2 );
   // parent[t] = t
   hipLaunchKernelGGL( k_iota, dim3( ( n + b - 1 ) / b ), dim3( b ), 0, st, n, w.parent.p );
   hipLaunchKernelGGL( k_unite, dim3( ( nv + b - 1 ) / b ), dim3( b ), 0, st, n, w.vid.p, w.parent.p, w.out.p + 2 );
   hipLaunchKernelGGL( k_roots, dim3( ( n + b - 1 ) / b ), dim3( b ), 0, st, n, w.parent.p, w.root.p, w.flag.p );
   hipLaunchKernelGGL( k_label_scan, dim3( 1 ), dim3( 1024 ), 0, st, n, w.flag.p, w.lab.p, w.out.p );
   hipLaunchKernelGGL( k_label, dim3( ( n + b - 1 ) / b ), dim3( b ), 0, st, n, w.root.p, w.lab.p, d_icomp );
   hipLaunchKernelGGL( k_edges, dim3( ( nv + b - 1 ) / b ), dim3( b ), 0, st, n, w.vid.p, w.ekey.p, w.ecnt.p, hmask,
                       w.out.p + 2 );
   hipLaunchKernelGGL( k_open_count, dim3( g ), dim3( b ), 0, st, w.ekey.p, w.ecnt.p, (long)w.hsize, w.out.p + 1 );
   int h[ 4 ];
   BH_HIP( hipMemcpyAsync( h, w.out.p, sizeof h, hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipStreamSynchronize( st ) );
   if ( h[ 2 ] ) throw std::runtime_error( "COMPON on the device: a hash table or union-find loop ran past its bound" );
   *ncomp = h[ 0 ];
   *nopen = h[ 1 ];
}

}  // namespace bhg
