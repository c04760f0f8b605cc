// Shared definitions of the MI355X BEST solver library (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace bhg {

// LU block size: the diagonal block of the blocked LU lives in LDS (128 x 128 FP64 = 128 KiB of
// the 160 KiB per CU) and every GEMM tile edge is a multiple of it, so matrices are padded to a
// multiple of kNB with an identity block.
constexpr int kNB = 128;

// Right-hand sides per solve call at most.
constexpr int kMaxRhs = 16;
// Columns of K per workgroup of the triangular solves' split-K products (k_skinny_part), and the
// most slices one product has: its K is at most a super-block order of 1024.
constexpr int kSkinnySlice = 128;
constexpr int kSkinnyParts = 1024 / kSkinnySlice;

// Per-source-triangle record used by the pair kernel: the vertices, the corrected area and the
// source-only quantities of the quadrature (edge vectors, their dot products, sqrt(d2121)).
// 32 doubles so that one record is two 128-byte lines and loads as scalars (uniform per wave).
constexpr int kSrcStride = 32;
enum SrcField {
   kP1 = 0,    // p1 x,y,z
   kP2 = 3,    // p2
   kP3 = 6,    // p3
   kArea = 9,  // corrected area
   kV21 = 10,  // p2 - p1
   kV31 = 13,  // p3 - p1
   kV32 = 16,  // p3 - p2
   kD2121 = 19,
   kD2131 = 20,
   kD2132 = 21,
   kD3131 = 22,
   kD3232 = 23,
   kD3231 = 24,
   kS21 = 25,  // sqrt(d2121)
};

// The streams and events of the blocked factorization: st (the caller's; the critical path),
// st2 (the trailing update), st3 (the U row panels); ev_u / ev_g order st2 after st and back, ev
// the forks and joins of st3 (ev[0..4]) and the later parts of st2's update (ev[5], ev[6]).
constexpr int kLuEvents = 7;
struct LuStreams {
   hipStream_t st, st2, st3;
   hipEvent_t ev_u, ev_g;
   hipEvent_t ev[ kLuEvents ];
   // optional: recorded on st when the trailing matrix first has fewer than late_rest rows (the
   // factorization's latency-bound end), or at its end
   hipEvent_t ev_late = nullptr;
   int late_rest = 0;
};

// Kernel classes timed by the optional launch profiler.
// kKGemm32 counts the FP32 GEMMs on 64 x 128 and 128 x 64 tiles (the panels below the diagonal
// blocks and the U row panels, products written in place), kKGemm32S those on 64 x 64 tiles (the
// narrow parts of the trailing update and the column updates between the blocks of a pair),
// kKGemm32T those on 128 x 128 tiles (the wide trailing updates).
enum KernelClass { kKPairs = 0, kKGemm64, kKGemm32, kKDiag, kKTrsm, kKResid, kKOther, kKGemm32T, kKGemm32S, kKClasses };

// When a solver turns profiling on, every launch of a timed class is bracketed by two HIP events
// on the launching stream; the algorithmic flops of the launch are added to its class.  The
// events come from a pool that grows once and is reused.
struct LaunchProfiler {
   bool on = false;
   hipStream_t st = nullptr;
   std::vector<hipEvent_t> pool;
   size_t used = 0;
   std::vector<int> cls;
   std::vector<double> flops;
   double total_ms[ kKClasses ] = {};
   double total_flops[ kKClasses ] = {};
   long count[ kKClasses ] = {};
   void reset() {
      used = 0;
      cls.clear();
      flops.clear();
      for ( int k = 0; k < kKClasses; ++k ) total_ms[ k ] = total_flops[ k ] = 0, count[ k ] = 0;
   }
   hipEvent_t next() {
      if ( used == pool.size() ) {
         hipEvent_t e;
         if ( hipEventCreate( &e ) != hipSuccess ) throw std::runtime_error( "hipEventCreate" );
         pool.push_back( e );
      }
      return pool[ used++ ];
   }
   void begin( int k, double f, hipStream_t s ) {
      if ( !on ) return;
      (void)hipEventRecord( next(), s ? s : st );
      cls.push_back( k );
      flops.push_back( f );
   }
   void end( hipStream_t s ) {
      if ( !on ) return;
      (void)hipEventRecord( next(), s ? s : st );
   }
   // After the stream has been synchronized.
   void collect() {
      for ( size_t i = 0; i < cls.size(); ++i ) {
         float ms = 0.f;
         if ( hipEventElapsedTime( &ms, pool[ 2 * i ], pool[ 2 * i + 1 ] ) != hipSuccess ) continue;
         total_ms[ cls[ i ] ] += ms;
         total_flops[ cls[ i ] ] += flops[ i ];
         count[ cls[ i ] ] += 1;
      }
      used = 0;
      cls.clear();
      flops.clear();
   }
   ~LaunchProfiler() {
      for ( auto e : pool ) (void)hipEventDestroy( e );
   }
};

// The profiler of the solver whose work is being enqueued on this host thread (or null).
extern thread_local LaunchProfiler* g_prof;

struct ProfScope {
   hipStream_t s;
   ProfScope( int k, double f, hipStream_t stream = nullptr ) : s( stream ) {
      if ( g_prof ) g_prof->begin( k, f, s );
   }
   ~ProfScope() {
      if ( g_prof ) g_prof->end( s );
   }
};

inline void hip_check( hipError_t e, const char* what, const char* file, int line ) {
   if ( e != hipSuccess ) {
      char buf[ 512 ];
      std::snprintf( buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line, hipGetErrorString( e ) );
      throw std::runtime_error( buf );
   }
}

// A device allocation owned by the host object that holds it.
template <typename T>
struct DevBuf {
   T* p = nullptr;
   size_t n = 0;
   ~DevBuf() { release(); }
   void release() {
      if ( p ) (void)hipFree( p );
      p = nullptr;
      n = 0;
   }
   // Grows to at least count elements (contents are not kept).
   T* ensure( size_t count ) {
      if ( count > n ) {
         release();
         hip_check( hipMalloc( &p, std::max<size_t>( count, 1 ) * sizeof( T ) ), "hipMalloc", __FILE__, __LINE__ );
         n = count;
      }
      return p;
   }
};

}  // namespace bhg

#define BH_HIP( x ) ::bhg::hip_check( ( x ), #x, __FILE__, __LINE__ )
