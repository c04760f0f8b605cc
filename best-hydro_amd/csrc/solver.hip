// The solver object behind the C ABI: one mesh, its matrix and factors resident in HBM, the
// sequence of kernels for one mesh solve, and the C entry points (include/best_hip.h).
//
// One mesh solve (the -v flow of best/best.f:261-498, without file I/O):
//   k_prep -> k_self -> k_deflation_weights -> k_pairs (assembly + deflation) -> LU -> 11-RHS
//   solve (FP64 LU, or FP32 LU + FP64 iterative refinement) -> k_sums -> host 3x3 algebra.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/best_hip.h"
#include "bh_common.h"
#include "hmatrix.h"
#include "transport.h"

namespace bhg {

thread_local LaunchProfiler* g_prof = nullptr;

// ------------------------------------------------------------------ launchers (other files)
void launch_prep( hipStream_t, int, double*, double*, double*, int, double, double*, double* );
void launch_self( hipStream_t, int, const double*, double, const double*, const double*, const double*, const double*,
                  double*, double*, double*, double* );
void launch_deflation_weights( hipStream_t, int, int, const double*, const double*, const int*, double*, double*, int );
void launch_pairs( hipStream_t, int, int, double, const double*, const double*, const double*, const double*, const int*,
                   const double*, int, double*, long );
void launch_pad_identity( hipStream_t, int, int, double*, long );
void launch_entries( hipStream_t, int, double, const double*, const double*, const double*, int, int, int, int, double*,
                     long );
template <typename T>
void lu_nopiv( hipStream_t, hipStream_t, hipEvent_t, hipEvent_t, int, T*, long, T*, int* );
template <typename T>
void lu_pivoted( hipStream_t, int, T*, long, T*, int*, int*, double*, int* );
template <typename T>
void build_super_inverses( hipStream_t, int, int, const T*, long, const T*, T*, T*, T* );
template <typename T>
void lu_solve( hipStream_t, int, int, int, const T*, long, const T*, const T*, const int*, bool, T*, T*, T*, T* );
template <typename S, typename D>
void to_interleaved( hipStream_t, int, int, const S*, long, D* );
template <typename S, typename D>
void from_interleaved( hipStream_t, int, int, const S*, D*, long );
void residual( hipStream_t, int, int, int, const double*, long, const double*, long, const double*, long, double*, long,
               double*, double* );
size_t residual_part_elems( int );
void ir_update( hipStream_t, int, int, double*, long, const float*, long, double* );
void convert_d2f( hipStream_t, long, const double*, float* );
void convert_f2d( hipStream_t, long, const float*, double* );


// Order of the diagonal super-blocks whose inverses the triangular solves use; the matrix is
// padded to a multiple of it.
constexpr int kSuper = 1024;
constexpr int kPairStrip = 64;  // source triangles per workgroup of the pair kernel

struct PhaseTimer {
   hipEvent_t ev[ 8 ] = {};
   int used = 0;
   void init() {
      for ( auto& e : ev )
         if ( !e ) BH_HIP( hipEventCreate( &e ) );
   }
   void mark( hipStream_t st, int k ) {
      BH_HIP( hipEventRecord( ev[ k ], st ) );
      used = std::max( used, k + 1 );
   }
   double ms( int a, int b ) {
      float t = 0.f;
      if ( hipEventElapsedTime( &t, ev[ a ], ev[ b ] ) != hipSuccess ) return -1.0;
      return t;
   }
   ~PhaseTimer() {
      for ( auto& e : ev )
         if ( e ) (void)hipEventDestroy( e );
   }
};

}  // namespace bhg

struct bhg_solver {
   int device = 0;
   hipStream_t st = nullptr;
   hipStream_t st2 = nullptr;          // look-ahead stream of the factorization
   hipEvent_t ev_u = nullptr, ev_g = nullptr;
   std::string last_error;
   std::deque<std::string> log;
   // mesh
   int n = 0, n3 = 0, npad = 0, ndef = 1, ncomp = 1, nopen = 0;
   double cg = 1.0, truearea = 0.0, truevol = 0.0;
   bool mesh_raw = false;     // raw (unshifted) vertices are resident (bhg_set_mesh)
   bool mesh_ready = false;   // prepared vertices, areas, incenters and self blocks are resident
   bhg::DevBuf<double> raw_v, v, area, xc, selfb, xmu, src, gdef, scal;
   bhg::DevBuf<int> comp;
   // matrix and factors
   bhg::DevBuf<double> a64, inv64, xl64, xu64, tmp64;
   bhg::DevBuf<float> a32, inv32, xl32, xu32, tmp32;
   bhg::DevBuf<int> ipiv, flags;
   bhg::DevBuf<double> pval;
   bhg::DevBuf<int> pidx;
   int factored = 0;          // 0, 64 or 32; -1: the H-matrix is built (method 1)
   // H-matrix path (method 1)
   int method = 0;            // 0 dense LU, 1 H-matrix + GMRES
   bhg::HmOptions hm_opt;
   bhg::HmStats hm_stats;
   bhg::HMatrixPtr hm;
   bool pivoted = false;      // the slow path ran (rows were exchanged)
   bool deflated = true;      // the factored matrix carries the deflation
   // vectors
   bhg::DevBuf<double> b, x, r, y64, x64il, p64, part, norms, sums;
   bhg::DevBuf<float> d32, y32, x32, p32;
   std::vector<double> host_sums;
   bhg::PhaseTimer timer;
   bhg::LaunchProfiler prof;
   bhg_stats stats {};
   double ir_tol = 1e-12;
   int ir_maxit = 30;

   void say( const char* fmt, ... ) __attribute__( ( format( printf, 2, 3 ) ) ) {
      char buf[ 512 ];
      va_list ap;
      va_start( ap, fmt );
      std::vsnprintf( buf, sizeof buf, fmt, ap );
      va_end( ap );
      log.emplace_back( buf );
   }
};

namespace bhg {

namespace {

int round_up( int v, int m ) { return ( v + m - 1 ) / m * m; }

void set_device( bhg_solver* s ) { BH_HIP( hipSetDevice( s->device ) ); }

// Sizes the per-mesh buffers for n triangles.
void size_mesh( bhg_solver* s, int n ) {
   s->n = n;
   s->n3 = 3 * n;
   s->npad = round_up( s->n3, kSuper );
   s->v.ensure( 9 * (size_t)n );
   s->area.ensure( n );
   s->xc.ensure( 3 * (size_t)n );
   s->selfb.ensure( 9 * (size_t)n );
   s->xmu.ensure( 3 * (size_t)n );
   s->src.ensure( (size_t)kSrcStride * n );
   s->scal.ensure( 8 );
   s->comp.ensure( n );
   s->flags.ensure( 4 );
   s->pval.ensure( 64 );
   s->pidx.ensure( 64 );
}

double* V1( bhg_solver* s ) { return s->v.p; }
double* V2( bhg_solver* s ) { return s->v.p + 3 * (size_t)s->n; }
double* V3( bhg_solver* s ) { return s->v.p + 6 * (size_t)s->n; }

// Incenters, self blocks, XMU, source records (K1); the deflation weights unless given.
// exact_order: DIAGM and the weights as the reference's running sums (bit for bit, the value of
// bhm_diag_mean); otherwise fixed-order tree sums (equal to rounding, much faster).
void self_terms( bhg_solver* s, bool weights, bool exact_order ) {
   launch_self( s->st, s->n, s->scal.p, s->cg, V1( s ), V2( s ), V3( s ), s->area.p, s->xc.p, s->selfb.p, s->xmu.p, s->src.p );
   if ( weights ) {
      s->gdef.ensure( std::max( s->ndef, 1 ) );
      launch_deflation_weights( s->st, s->n, s->ndef, s->selfb.p, s->xmu.p, s->comp.p, s->gdef.p, s->scal.p + 5,
                                exact_order ? 1 : 0 );
   }
}

void assemble( bhg_solver* s, bool deflate ) {
   const size_t np = s->npad;
   s->a64.ensure( np * np );
   launch_pairs( s->st, s->n, kPairStrip, s->cg, s->src.p, s->xc.p, s->selfb.p, s->xmu.p, s->comp.p, s->gdef.p, deflate ? 1 : 0,
                 s->a64.p, (long)np );
   launch_pad_identity( s->st, s->n3, s->npad, s->a64.p, (long)np );
   s->deflated = deflate;
}

int read_flag( bhg_solver* s ) {
   int h = 0;
   BH_HIP( hipMemcpyAsync( &h, s->flags.p, sizeof( int ), hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   return h;
}

// Factors the assembled matrix: precision 64 in place on a64 (GR is consumed), 32 on an FP32
// copy (a64 is kept for the refinement residuals).  Partial pivoting falls back to the slow path
// when the exchange-free factorization is not GEPP's (see lu.hip).
void factor( bhg_solver* s, int precision ) {
   s->hm.reset();
   const size_t np = s->npad;
   const int nblk = s->npad / kNB;
   s->ipiv.ensure( np );
   BH_HIP( hipMemsetAsync( s->flags.p, 0, 4 * sizeof( int ), s->st ) );
   s->pivoted = false;
   if ( precision == 64 ) {
      s->inv64.ensure( (size_t)nblk * 2 * kNB * kNB );
      lu_nopiv<double>( s->st, s->st2, s->ev_u, s->ev_g, s->npad, s->a64.p, (long)np, s->inv64.p, s->flags.p );
      if ( read_flag( s ) ) {
         s->say( "LU: a multiplier exceeded 1; refactoring with row exchanges" );
         assemble( s, s->deflated );
         BH_HIP( hipMemsetAsync( s->flags.p, 0, 4 * sizeof( int ), s->st ) );
         lu_pivoted<double>( s->st, s->npad, s->a64.p, (long)np, s->inv64.p, s->ipiv.p, s->flags.p + 1, s->pval.p, s->pidx.p );
         s->pivoted = true;
      }
      s->xl64.ensure( np * kSuper );
      s->xu64.ensure( np * kSuper );
      s->tmp64.ensure( np * kSuper / 2 );
      build_super_inverses<double>( s->st, s->npad, kSuper, s->a64.p, (long)np, s->inv64.p, s->xl64.p, s->xu64.p, s->tmp64.p );
   } else {
      s->a32.ensure( np * np );
      s->inv32.ensure( (size_t)nblk * 2 * kNB * kNB );
      convert_d2f( s->st, (long)( np * np ), s->a64.p, s->a32.p );
      lu_nopiv<float>( s->st, s->st2, s->ev_u, s->ev_g, s->npad, s->a32.p, (long)np, s->inv32.p, s->flags.p );
      if ( read_flag( s ) ) {
         s->say( "LU (FP32): a multiplier exceeded 1; refactoring with row exchanges" );
         convert_d2f( s->st, (long)( np * np ), s->a64.p, s->a32.p );
         BH_HIP( hipMemsetAsync( s->flags.p, 0, 4 * sizeof( int ), s->st ) );
         lu_pivoted<float>( s->st, s->npad, s->a32.p, (long)np, s->inv32.p, s->ipiv.p, s->flags.p + 1, s->pval.p, s->pidx.p );
         s->pivoted = true;
      }
      s->xl32.ensure( np * kSuper );
      s->xu32.ensure( np * kSuper );
      s->tmp32.ensure( np * kSuper / 2 );
      build_super_inverses<float>( s->st, s->npad, kSuper, s->a32.p, (long)np, s->inv32.p, s->xl32.p, s->xu32.p, s->tmp32.p );
   }
   s->factored = precision;
}

// The H-matrix of the deflated GR and its preconditioner, from the resident mesh (prepared
// vertices, incenters, self blocks, XMU, surfaces and weights).
void hm_factor( bhg_solver* s ) {
   s->hm.reset();
   s->a64.release();
   s->a32.release();
   std::vector<double> hv( 9 * (size_t)s->n ), hx( 3 * (size_t)s->n );
   BH_HIP( hipMemcpyAsync( hv.data(), s->v.p, hv.size() * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipMemcpyAsync( hx.data(), s->xc.p, hx.size() * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   HmInput in;
   in.n = s->n;
   in.cg = s->cg;
   in.h_v = hv.data();
   in.h_xc = hx.data();
   in.d_src = s->src.p;
   in.d_xc = s->xc.p;
   in.d_selfb = s->selfb.p;
   in.d_xmu = s->xmu.p;
   in.d_comp = s->comp.p;
   in.d_gdef = s->gdef.p;
   in.ndef = s->ndef;
   std::vector<std::string> lines;
   s->hm = hm_build( s->st, in, s->hm_opt, &s->hm_stats, &lines );
   for ( auto& l : lines ) s->log.push_back( std::move( l ) );
   s->factored = -1;
   s->pivoted = false;
}

void size_vectors( bhg_solver* s ) {
   const size_t np = s->npad;
   const size_t need = np * kMaxRhs;
   const bool grow = s->b.n < need;
   s->b.ensure( need );
   s->x.ensure( need );
   s->r.ensure( need );
   s->y64.ensure( need );
   s->x64il.ensure( need );
   s->p64.ensure( need * 4 );
   s->d32.ensure( need );
   s->y32.ensure( need );
   s->x32.ensure( need );
   s->p32.ensure( need * 4 );
   s->part.ensure( residual_part_elems( s->npad ) );
   s->norms.ensure( 32 );
   s->sums.ensure( (size_t)kSumGroups * kSumSlots );

   if ( grow ) {
      BH_HIP( hipMemsetAsync( s->b.p, 0, need * sizeof( double ), s->st ) );
      BH_HIP( hipMemsetAsync( s->x.p, 0, need * sizeof( double ), s->st ) );
   }
}

// Solves for the nrhs columns of s->b (npad x nrhs, rows >= n3 zero) into s->x.
void solve_device( bhg_solver* s, int nrhs ) {
   const long np = s->npad;
   s->stats.ir_iterations = 0;
   s->stats.ir_correction = 0.0;
   if ( s->factored == -1 ) {
      std::vector<std::string> lines;
      hm_solve( *s->hm, s->st, nrhs, s->b.p, np, s->x.p, np, &s->hm_stats, &lines );
      for ( auto& l : lines ) s->log.push_back( std::move( l ) );
      return;
   }
   if ( s->factored == 64 ) {
      to_interleaved<double, double>( s->st, s->npad, nrhs, s->b.p, np, s->r.p );
      lu_solve<double>( s->st, s->npad, s->n3, kSuper, s->a64.p, np, s->xl64.p, s->xu64.p, s->ipiv.p, s->pivoted, s->r.p,
                        s->y64.p, s->x64il.p, s->p64.p );
      from_interleaved<double, double>( s->st, s->n3, nrhs, s->x64il.p, s->x.p, np );
      return;
   }
   if ( s->factored != 32 ) throw std::runtime_error( "solve before the factorization" );
   // FP64 iterative refinement with the FP32 factors: x0 = LU32^-1 b, then r = b - A x,
   // x += LU32^-1 r until the correction is below ir_tol relative to x in every column.
   double hn[ 32 ];
   double prev = 1e300;
   for ( int it = 0; it < s->ir_maxit; ++it ) {
      BH_HIP( hipMemsetAsync( s->norms.p, 0, 32 * sizeof( double ), s->st ) );
      if ( it == 0 ) {
         BH_HIP( hipMemsetAsync( s->x.p, 0, sizeof( double ) * np * nrhs, s->st ) );
         to_interleaved<double, float>( s->st, s->npad, nrhs, s->b.p, np, s->d32.p );
      } else {
         residual( s->st, s->npad, s->n3, nrhs, s->a64.p, np, s->x.p, np, s->b.p, np, s->r.p, np, s->part.p, s->norms.p );
         to_interleaved<double, float>( s->st, s->npad, nrhs, s->r.p, np, s->d32.p );
         BH_HIP( hipMemsetAsync( s->norms.p, 0, 32 * sizeof( double ), s->st ) );
      }
      lu_solve<float>( s->st, s->npad, s->n3, kSuper, s->a32.p, np, s->xl32.p, s->xu32.p, s->ipiv.p, s->pivoted, s->d32.p,
                       s->y32.p, s->x32.p, s->p32.p );
      ir_update( s->st, s->n3, nrhs, s->x.p, np, s->x32.p, np, s->norms.p );
      BH_HIP( hipMemcpyAsync( hn, s->norms.p, sizeof hn, hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
      double worst = 0.0;
      for ( int k = 0; k < nrhs; ++k ) {
         const double rel = hn[ 16 + k ] > 0 ? hn[ k ] / hn[ 16 + k ] : 0.0;
         if ( !std::isfinite( rel ) ) throw std::runtime_error( "iterative refinement produced a non-finite correction" );
         worst = std::max( worst, rel );
      }
      s->stats.ir_iterations = it + 1;
      s->stats.ir_correction = worst;
      if ( it > 0 && worst < s->ir_tol ) break;
      if ( it > 2 && worst > 0.5 * prev ) break;  // stagnation at the FP64 floor
      prev = worst;
   }
}

// Host-ordered sum of the per-workgroup partials.
void collect_sums( bhg_solver* s, double* out ) {
   s->host_sums.resize( (size_t)kSumGroups * kSumSlots );
   BH_HIP( hipMemcpyAsync( s->host_sums.data(), s->sums.p, s->host_sums.size() * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   for ( int k = 0; k < kSumSlots; ++k ) {
      double t = 0.0;
      for ( int g = 0; g < kSumGroups; ++g ) t = t + s->host_sums[ (size_t)g * kSumSlots + k ];
      out[ k ] = t;
   }
}

}  // namespace

// COMPON (best/compon.f90:19-115) on the host: surfaces by exact-coordinate vertex matching and
// union-find, numbered by first triangle; open edges = edges not shared by exactly two triangles.
void compon( int n, const double* v1, const double* v2, const double* v3, int* icomp, int* ncomp, int* nopen ) {
   const int nv = 3 * n;
   std::vector<std::array<double, 3>> key( nv );
   for ( int t = 0; t < n; ++t )
      for ( int c = 0; c < 3; ++c ) {
         key[ 3 * t ][ c ] = v1[ (size_t)c * n + t ];
         key[ 3 * t + 1 ][ c ] = v2[ (size_t)c * n + t ];
         key[ 3 * t + 2 ][ c ] = v3[ (size_t)c * n + t ];
      }
   std::vector<int> idx( nv );
   std::iota( idx.begin(), idx.end(), 0 );
   std::stable_sort( idx.begin(), idx.end(), [&]( int a, int b ) { return key[ a ] < key[ b ]; } );
   std::vector<int> vid( nv );
   int nid = 0;
   for ( int j = 0; j < nv; ++j ) {
      if ( j == 0 || key[ idx[ j ] ] != key[ idx[ j - 1 ] ] ) ++nid;
      vid[ idx[ j ] ] = nid;
   }
   std::vector<int> root( n );
   std::iota( root.begin(), root.end(), 0 );
   auto find = [&]( int i ) {
      int p = i;
      while ( root[ p ] != p ) p = root[ p ];
      while ( root[ i ] != p ) {
         const int t = root[ i ];
         root[ i ] = p;
         i = t;
      }
      return p;
   };
   for ( int j = 1; j < nv; ++j )
      if ( vid[ idx[ j ] ] == vid[ idx[ j - 1 ] ] ) {
         const int a = find( idx[ j ] / 3 ), b = find( idx[ j - 1 ] / 3 );
         if ( a != b ) root[ std::max( a, b ) ] = std::min( a, b );
      }
   std::vector<int> lab( n, 0 );
   int nc = 0;
   for ( int t = 0; t < n; ++t ) {
      const int r = find( t );
      if ( !lab[ r ] ) lab[ r ] = ++nc;
      icomp[ t ] = lab[ r ];
   }
   std::vector<std::pair<int, int>> edges( nv );
   for ( int t = 0; t < n; ++t )
      for ( int c = 0; c < 3; ++c ) {
         const int a = vid[ 3 * t + c ], b = vid[ 3 * t + ( c + 1 ) % 3 ];
         edges[ 3 * t + c ] = { std::min( a, b ), std::max( a, b ) };
      }
   std::sort( edges.begin(), edges.end() );
   int open = 0;
   for ( int j = 0; j < nv; ) {
      int k = j;
      while ( k < nv && edges[ k ] == edges[ j ] ) ++k;
      if ( k - j != 2 ) ++open;
      j = k;
   }
   *ncomp = nc;
   *nopen = open;
}

}  // namespace bhg

// ====================================================================== C ABI
using bhg::DevBuf;

namespace {

template <class F>
int guarded( bhg_solver* s, const char* what, F&& f ) {
   try {
      f();
      return 0;
   } catch ( const std::exception& e ) {
      if ( s ) {
         s->last_error = std::string( what ) + ": " + e.what();
         s->say( "error: %s", s->last_error.c_str() );
      }
      return 1;
   }
}

}  // namespace

extern "C" {

bhg_solver* bhg_create( int device ) {
   auto s = std::make_unique<bhg_solver>();
   s->device = device;
   try {
      BH_HIP( hipSetDevice( device ) );
      BH_HIP( hipStreamCreateWithFlags( &s->st, hipStreamNonBlocking ) );
      BH_HIP( hipStreamCreateWithFlags( &s->st2, hipStreamNonBlocking ) );
      BH_HIP( hipEventCreateWithFlags( &s->ev_u, hipEventDisableTiming ) );
      BH_HIP( hipEventCreateWithFlags( &s->ev_g, hipEventDisableTiming ) );
      s->timer.init();
      s->flags.ensure( 4 );
   } catch ( ... ) {
      return nullptr;
   }
   return s.release();
}

void bhg_destroy( bhg_solver* s ) {
   if ( !s ) return;
   (void)hipSetDevice( s->device );
   if ( s->st ) (void)hipStreamSynchronize( s->st );
   if ( s->st2 ) (void)hipStreamSynchronize( s->st2 );
   hipStream_t st = s->st, st2 = s->st2;
   hipEvent_t e1 = s->ev_u, e2 = s->ev_g;
   delete s;
   if ( e1 ) (void)hipEventDestroy( e1 );
   if ( e2 ) (void)hipEventDestroy( e2 );
   if ( st ) (void)hipStreamDestroy( st );
   if ( st2 ) (void)hipStreamDestroy( st2 );
}

const char* bhg_last_error( bhg_solver* s ) { return s ? s->last_error.c_str() : "no solver"; }

int bhg_log_line( bhg_solver* s, char* buf, int len ) {
   if ( !s || s->log.empty() ) return -1;
   const std::string m = std::move( s->log.front() );
   s->log.pop_front();
   const int k = std::min<int>( len, (int)m.size() );
   std::memcpy( buf, m.data(), k );
   return k;
}

int bhg_compon( int n, const double* v1, const double* v2, const double* v3, int* icomp, int* ncomp, int* nopen ) {
   try {
      bhg::compon( n, v1, v2, v3, icomp, ncomp, nopen );
      return 0;
   } catch ( ... ) {
      return 1;
   }
}

int bhg_set_mesh( bhg_solver* s, int n, const double* v1, const double* v2, const double* v3, double truearea,
                  double truevol, const int* icomp ) {
   return guarded( s, "bhg_set_mesh", [&] {
      if ( n < 2 ) throw std::runtime_error( "need at least two triangles" );
      bhg::set_device( s );
      bhg::size_mesh( s, n );
      s->raw_v.ensure( 9 * (size_t)n );
      const size_t bytes = 3 * (size_t)n * sizeof( double );
      BH_HIP( hipMemcpyAsync( s->raw_v.p, v1, bytes, hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipMemcpyAsync( s->raw_v.p + 3 * (size_t)n, v2, bytes, hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipMemcpyAsync( s->raw_v.p + 6 * (size_t)n, v3, bytes, hipMemcpyHostToDevice, s->st ) );
      std::vector<int> comp( n );
      int ncomp = 1, nopen = 0;
      if ( icomp ) {
         std::memcpy( comp.data(), icomp, n * sizeof( int ) );
         ncomp = *std::max_element( comp.begin(), comp.end() );
      } else {
         bhg::compon( n, v1, v2, v3, comp.data(), &ncomp, &nopen );
      }
      // best/best.f:324-337: one deflation term per closed surface, one for all if any is open.
      s->ncomp = ncomp;
      s->nopen = nopen;
      s->ndef = 1;
      if ( nopen > 0 ) std::fill( comp.begin(), comp.end(), 1 );
      else if ( ncomp > 1 ) s->ndef = ncomp;
      BH_HIP( hipMemcpyAsync( s->comp.p, comp.data(), n * sizeof( int ), hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
      s->truearea = truearea;
      s->truevol = truevol;
      s->mesh_raw = true;
      s->mesh_ready = false;
      s->factored = 0;
   } );
}

int bhg_solve_mesh( bhg_solver* s, int iarea, double cg, int ivisco, int icov, int precision, bhg_result* out ) {
   return guarded( s, "bhg_solve_mesh", [&] {
      if ( !s->mesh_raw ) throw std::runtime_error( "no mesh (bhg_set_mesh)" );
      if ( precision != 32 && precision != 64 ) throw std::runtime_error( "precision must be 32 or 64" );
      bhg::set_device( s );
      const int n = s->n;
      const long np = s->npad;
      s->cg = cg;
      struct ProfBind {
         explicit ProfBind( bhg::LaunchProfiler* p ) { bhg::g_prof = p->on ? p : nullptr; }
         ~ProfBind() { bhg::g_prof = nullptr; }
      } bind( &s->prof );
      s->prof.st = s->st;
      auto& T = s->timer;
      T.mark( s->st, 0 );
      BH_HIP( hipMemcpyAsync( s->v.p, s->raw_v.p, 9 * (size_t)n * sizeof( double ), hipMemcpyDeviceToDevice, s->st ) );
      bhg::launch_prep( s->st, n, bhg::V1( s ), bhg::V2( s ), bhg::V3( s ), iarea, s->truearea, s->area.p, s->scal.p );
      bhg::self_terms( s, true, false );
      s->mesh_ready = true;
      T.mark( s->st, 1 );
      if ( s->method == 1 ) {  // the H-matrix build is the factorization phase
         T.mark( s->st, 2 );
         bhg::hm_factor( s );
      } else {
         bhg::assemble( s, true );
         T.mark( s->st, 2 );
         bhg::factor( s, precision );
      }
      T.mark( s->st, 3 );
      bhg::size_vectors( s );
      const int nrhs = ivisco ? 11 : 6;
      bhg::launch_rhs( s->st, n, s->xc.p, s->b.p, np, ivisco ? 1 : 0 );
      bhg::solve_device( s, nrhs );
      T.mark( s->st, 4 );
      const double zero[ 3 ] = { 0.0, 0.0, 0.0 };
      bhg::launch_sums( s->st, n, s->xc.p, s->area.p, s->x.p, np, ivisco ? 5 : 0, zero, icov ? 1 : 0, s->sums.p );
      double sums[ bhg::kSumSlots ];
      bhg::collect_sums( s, sums );
      bhg::Transport tr {};
      bhg::difco_algebra( sums, cg, tr );
      if ( ivisco ) {
         if ( icov ) {
            bhg::visco_centre( sums, tr.dvisc );
            bhg::launch_sums( s->st, n, s->xc.p, s->area.p, s->x.p, np, 5, tr.dvisc, 0, s->sums.p );
            double sums2[ bhg::kSumSlots ];
            bhg::collect_sums( s, sums2 );
            tr.chiprime = bhg::visco_chiprime( sums2, cg );
         } else {
            tr.chiprime = bhg::visco_chiprime( sums, cg );
         }
      }
      bhg::output_values( tr );
      T.mark( s->st, 5 );
      double scal[ 8 ];
      BH_HIP( hipMemcpyAsync( scal, s->scal.p, sizeof scal, hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
      s->stats.t_prep = T.ms( 0, 1 );
      s->stats.t_assemble = T.ms( 1, 2 );
      s->stats.t_factor = T.ms( 2, 3 );
      s->stats.t_solve = T.ms( 3, 4 );
      s->stats.t_reduce = T.ms( 4, 5 );
      s->stats.t_total = T.ms( 0, 5 );
      s->stats.pivoted = s->pivoted ? 1 : 0;
      s->stats.npad = s->npad;
      if ( s->prof.on ) s->prof.collect();
      if ( out ) {
         std::memcpy( out->dtt, tr.dtt, sizeof tr.dtt );
         std::memcpy( out->dtr, tr.dtr, sizeof tr.dtr );
         std::memcpy( out->drr, tr.drr, sizeof tr.drr );
         std::memcpy( out->urc, tr.urc, sizeof tr.urc );
         std::memcpy( out->udr, tr.udr, sizeof tr.udr );
         std::memcpy( out->ktt, tr.ktt, sizeof tr.ktt );
         std::memcpy( out->krt, tr.krt, sizeof tr.krt );
         std::memcpy( out->krr, tr.krr, sizeof tr.krr );
         std::memcpy( out->drr_eig, tr.drr_eig, sizeof tr.drr_eig );
         std::memcpy( out->dtt_eig, tr.dtt_eig, sizeof tr.dtt_eig );
         out->dr = tr.dr;
         out->dtrans = tr.dtrans;
         out->chiprime = tr.chiprime;
         out->chi = ( iarea == 1 && s->truevol != 0.0 ) ? tr.chiprime / s->truevol : 0.0;
         std::memcpy( out->dvisc, tr.dvisc, sizeof tr.dvisc );
         out->afac = scal[ 0 ];
         out->totarea = scal[ 1 ];
         std::memcpy( out->centroid, scal + 2, 3 * sizeof( double ) );
         out->diagm = scal[ 5 ];
         out->ncomp = s->ncomp;
         out->nopen = s->nopen;
         out->ndef = s->ndef;
      }
   } );
}

int bhg_get_solution( bhg_solver* s, int nrhs, double* out, long ldo ) {
   return guarded( s, "bhg_get_solution", [&] {
      if ( nrhs > bhg::kMaxRhs ) throw std::runtime_error( "too many right-hand sides" );
      BH_HIP( hipMemcpy2DAsync( out, ldo * sizeof( double ), s->x.p, s->npad * sizeof( double ), s->n3 * sizeof( double ),
                                nrhs, hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

int bhg_get_stats( bhg_solver* s, bhg_stats* out ) {
   if ( !s || !out ) return 1;
   *out = s->stats;
   return 0;
}

int bhg_set_profiling( bhg_solver* s, int on ) {
   if ( !s ) return 1;
   s->prof.on = on != 0;
   s->prof.reset();
   return 0;
}

int bhg_get_profile( bhg_solver* s, double* ms, double* flops, long* count ) {
   if ( !s ) return 1;
   for ( int k = 0; k < bhg::kKClasses; ++k ) {
      ms[ k ] = s->prof.total_ms[ k ];
      flops[ k ] = s->prof.total_flops[ k ];
      count[ k ] = s->prof.count[ k ];
   }
   return 0;
}

int bhg_set_refinement( bhg_solver* s, double tol, int maxit ) {
   if ( !s ) return 1;
   s->ir_tol = tol;
   s->ir_maxit = maxit;
   return 0;
}

int bhg_set_method( bhg_solver* s, int method, double eps, int single, double tol ) {
   if ( !s || ( method != 0 && method != 1 ) ) return 1;
   s->method = method;
   s->hm_opt.eps = eps > 0.0 ? eps : 1e-6;
   s->hm_opt.single = single ? 1 : 0;
   s->hm_opt.tol = tol;
   return 0;
}

int bhg_set_hm_option( bhg_solver* s, const char* name, double value ) {
   if ( !s || !name ) return 1;
   const std::string k( name );
   bhg::HmOptions& o = s->hm_opt;
   if ( k == "eta" ) o.eta = value;
   else if ( k == "aca_factor" ) o.aca_factor = value;
   else if ( k == "leaf" ) o.leaf = (int)value;
   else if ( k == "dense_small" ) o.dense_small = (long)value;
   else if ( k == "kcap" ) o.kcap = (int)value;
   else if ( k == "restart" ) o.restart = (int)value;
   else if ( k == "maxsteps" ) o.maxsteps = (int)value;
   else if ( k == "ws_gb" ) o.ws_bytes = value * 1e9;
   else return 1;
   return 0;
}

int bhg_get_hm_stats( bhg_solver* s, bhg_hm_stats* out ) {
   if ( !s || !out ) return 1;
   const bhg::HmStats& h = s->hm_stats;
   out->depth = h.depth;
   out->nleaves = h.nleaves;
   out->leaf_min = h.leaf_min;
   out->leaf_max = h.leaf_max;
   out->n_dense = h.n_dense;
   out->n_lowrank = h.n_lowrank;
   out->n_small = h.n_small;
   out->n_capped = h.n_capped;
   out->bytes_dense = h.bytes_dense;
   out->bytes_lowrank = h.bytes_lowrank;
   out->kmax = h.kmax;
   out->kmean = h.kmean;
   out->kernel_blocks = h.kernel_blocks;
   out->t_tree = h.t_tree;
   out->t_fill = h.t_fill;
   out->t_pre = h.t_pre;
   out->iters_min = h.iters_min;
   out->iters_max = h.iters_max;
   out->steps = h.steps;
   out->relres = h.relres;
   out->t_solve = h.t_solve;
   return 0;
}

int bhg_hm_plan( int n, const double* v1, const double* v2, const double* v3, const double* x, int leaf, double eta,
                 long* out ) {
   try {
      std::vector<double> hv( 9 * (size_t)n );
      std::memcpy( hv.data(), v1, 3 * (size_t)n * sizeof( double ) );
      std::memcpy( hv.data() + 3 * (size_t)n, v2, 3 * (size_t)n * sizeof( double ) );
      std::memcpy( hv.data() + 6 * (size_t)n, v3, 3 * (size_t)n * sizeof( double ) );
      bhg::HmOptions opt;
      opt.leaf = leaf;
      opt.eta = eta;
      const bhg::HmStats st = bhg::hm_plan( n, hv.data(), x, opt );
      out[ 0 ] = st.depth;
      out[ 1 ] = st.nleaves;
      out[ 2 ] = st.leaf_min;
      out[ 3 ] = st.leaf_max;
      out[ 4 ] = st.n_dense;
      out[ 5 ] = st.n_lowrank;
      out[ 6 ] = st.n_small;
      return 0;
   } catch ( ... ) {
      return 1;
   }
}

int bhg_hm_time_product( bhg_solver* s, int nrhs, int reps, double* ms, double* bytes ) {
   return guarded( s, "bhg_hm_time_product", [&] {
      if ( !s->hm ) throw std::runtime_error( "no H-matrix" );
      bhg::set_device( s );
      *ms = bhg::hm_time_product( *s->hm, s->st, nrhs, std::max( reps, 1 ), bytes );
   } );
}

int bhg_hm_apply( bhg_solver* s, int nrhs, const double* x, long ldx, double* y, long ldy ) {
   return guarded( s, "bhg_hm_apply", [&] {
      if ( !s->hm ) throw std::runtime_error( "no H-matrix" );
      if ( nrhs < 1 || nrhs > bhg::kMaxRhs ) throw std::runtime_error( "between 1 and 16 columns per call" );
      bhg::set_device( s );
      const long np = s->npad;
      bhg::size_vectors( s );
      BH_HIP( hipMemcpy2DAsync( s->b.p, np * sizeof( double ), x, ldx * sizeof( double ), s->n3 * sizeof( double ), nrhs,
                                hipMemcpyHostToDevice, s->st ) );
      bhg::hm_apply( *s->hm, s->st, nrhs, s->b.p, np, s->x.p, np );
      BH_HIP( hipMemcpy2DAsync( y, ldy * sizeof( double ), s->x.p, np * sizeof( double ), s->n3 * sizeof( double ), nrhs,
                                hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

// The prepared mesh of bhm_create (vertices already shifted, areas corrected).
int bhg_load_prepared( bhg_solver* s, int n, const double* v1, const double* v2, const double* v3, const double* area,
                       double afac, double cg ) {
   return guarded( s, "bhg_load_prepared", [&] {
      if ( n < 2 ) throw std::runtime_error( "need at least two triangles" );
      bhg::set_device( s );
      bhg::size_mesh( s, n );
      const size_t bytes = 3 * (size_t)n * sizeof( double );
      BH_HIP( hipMemcpyAsync( bhg::V1( s ), v1, bytes, hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipMemcpyAsync( bhg::V2( s ), v2, bytes, hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipMemcpyAsync( bhg::V3( s ), v3, bytes, hipMemcpyHostToDevice, s->st ) );
      BH_HIP( hipMemcpyAsync( s->area.p, area, n * sizeof( double ), hipMemcpyHostToDevice, s->st ) );
      double sc[ 8 ] = { afac, 0, 0, 0, 0, 0, 0, 0 };
      BH_HIP( hipMemcpyAsync( s->scal.p, sc, sizeof sc, hipMemcpyHostToDevice, s->st ) );
      std::vector<int> ones( n, 1 );
      BH_HIP( hipMemcpyAsync( s->comp.p, ones.data(), n * sizeof( int ), hipMemcpyHostToDevice, s->st ) );
      s->cg = cg;
      s->ndef = 1;
      bhg::self_terms( s, true, true );
      BH_HIP( hipStreamSynchronize( s->st ) );
      s->mesh_raw = false;
      s->mesh_ready = true;
      s->factored = 0;
   } );
}

int bhg_incenters( bhg_solver* s, double* x ) {
   return guarded( s, "bhg_incenters", [&] {
      if ( !s->mesh_ready ) throw std::runtime_error( "no prepared mesh" );
      BH_HIP( hipMemcpyAsync( x, s->xc.p, 3 * (size_t)s->n * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

int bhg_entries( bhg_solver* s, int i0, int m, int j0, int nc, double* a, long lda ) {
   return guarded( s, "bhg_entries", [&] {
      if ( !s->mesh_ready ) throw std::runtime_error( "no prepared mesh" );
      if ( m <= 0 || nc <= 0 ) return;
      if ( i0 < 0 || j0 < 0 || i0 + m > s->n3 || j0 + nc > s->n3 || lda < m )
         throw std::runtime_error( "entries out of range" );
      DevBuf<double> tmp;
      tmp.ensure( (size_t)m * nc );
      bhg::launch_entries( s->st, s->n, s->cg, s->src.p, s->xc.p, s->selfb.p, i0, m, j0, nc, tmp.p, m );
      BH_HIP( hipMemcpy2DAsync( a, lda * sizeof( double ), tmp.p, m * sizeof( double ), m * sizeof( double ), nc,
                                hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

double bhg_diag_mean( bhg_solver* s ) {
   double d = NAN;
   guarded( s, "bhg_diag_mean", [&] {
      BH_HIP( hipMemcpyAsync( &d, s->scal.p + 5, sizeof d, hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
   return d;
}

// TGETGR's X and GR (best/tgetgr.f:1) for the prepared mesh, without deflation, into host
// memory (column-major, leading dimension ldgr >= 3n).
int bhg_tgetgr( bhg_solver* s, double* x, double* gr, long ldgr ) {
   return guarded( s, "bhg_tgetgr", [&] {
      if ( !s->mesh_ready ) throw std::runtime_error( "no prepared mesh" );
      bhg::assemble( s, false );
      BH_HIP( hipMemcpy2DAsync( gr, ldgr * sizeof( double ), s->a64.p, s->npad * sizeof( double ), s->n3 * sizeof( double ),
                                s->n3, hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipMemcpyAsync( x, s->xc.p, 3 * (size_t)s->n * sizeof( double ), hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
      s->factored = 0;
   } );
}

// Assembles GR with the deflation GDEF(K) XMU XMU^T from the caller (bhm_build's arguments, or
// the device's own weights when gdef == nullptr) and factors it.
int bhg_factor_deflated( bhg_solver* s, int ndef, const double* gdef, const int* comp, const double* xmu, int precision ) {
   return guarded( s, "bhg_factor_deflated", [&] {
      if ( !s->mesh_ready ) throw std::runtime_error( "no prepared mesh" );
      if ( precision != 32 && precision != 64 ) throw std::runtime_error( "precision must be 32 or 64" );
      s->ndef = std::max( ndef, 1 );
      if ( gdef ) {
         s->gdef.ensure( s->ndef );
         BH_HIP( hipMemcpyAsync( s->gdef.p, gdef, s->ndef * sizeof( double ), hipMemcpyHostToDevice, s->st ) );
      }
      if ( comp && s->ndef > 1 ) {
         BH_HIP( hipMemcpyAsync( s->comp.p, comp, s->n * sizeof( int ), hipMemcpyHostToDevice, s->st ) );
      } else {
         std::vector<int> ones( s->n, 1 );
         BH_HIP( hipMemcpyAsync( s->comp.p, ones.data(), s->n * sizeof( int ), hipMemcpyHostToDevice, s->st ) );
         BH_HIP( hipStreamSynchronize( s->st ) );
      }
      if ( xmu ) BH_HIP( hipMemcpyAsync( s->xmu.p, xmu, 3 * (size_t)s->n * sizeof( double ), hipMemcpyHostToDevice, s->st ) );
      if ( !gdef ) {
         s->gdef.ensure( s->ndef );
         bhg::launch_deflation_weights( s->st, s->n, s->ndef, s->selfb.p, s->xmu.p, s->comp.p, s->gdef.p, s->scal.p + 5, 1 );
      }
      bhg::assemble( s, true );
      bhg::factor( s, precision );
      bhg::size_vectors( s );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

// BESOLV (best/besolve.f90:95-134) with the resident factors: b (3n x nrhs, column-major,
// leading dimension ldb) holds the right-hand sides and receives the solutions.
int bhg_besolv( bhg_solver* s, int nrhs, double* bh, long ldb ) {
   return guarded( s, "bhg_besolv", [&] {
      if ( !s->factored ) throw std::runtime_error( "no factorization" );
      if ( nrhs < 1 || nrhs > bhg::kMaxRhs ) throw std::runtime_error( "between 1 and 16 right-hand sides per call" );
      const long np = s->npad;
      BH_HIP( hipMemsetAsync( s->b.p, 0, sizeof( double ) * np * nrhs, s->st ) );
      BH_HIP( hipMemcpy2DAsync( s->b.p, np * sizeof( double ), bh, ldb * sizeof( double ), s->n3 * sizeof( double ), nrhs,
                                hipMemcpyHostToDevice, s->st ) );
      bhg::solve_device( s, nrhs );
      BH_HIP( hipMemcpy2DAsync( bh, ldb * sizeof( double ), s->x.p, np * sizeof( double ), s->n3 * sizeof( double ), nrhs,
                                hipMemcpyDeviceToHost, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

// ---------------------------------------------------------------------- hmatrix/best_hmatrix.h

static bhg_solver* g_solver = nullptr;
static std::deque<std::string> g_log;
static int g_precision = 32;
// Above this many unknowns bhm_build takes the H-matrix path (the reference's NHMAT, 10,000
// triangles, best/besolve.f90).
static const int kDenseMaxUnknowns = 30000;

static void g_drain() {
   if ( !g_solver ) return;
   while ( !g_solver->log.empty() ) {
      g_log.push_back( std::move( g_solver->log.front() ) );
      g_solver->log.pop_front();
   }
}

int bhm_available( void ) { return 1; }

int bhm_create( int n, const double* v1, const double* v2, const double* v3, const double* area, double afac, double cg,
                int threads ) {
   (void)threads;
   if ( g_solver ) {
      bhg_destroy( g_solver );
      g_solver = nullptr;
   }
   const char* dev = std::getenv( "BEST_HIP_DEVICE" );
   g_solver = bhg_create( dev ? std::atoi( dev ) : 0 );
   if ( !g_solver ) {
      g_log.emplace_back( "error: no HIP device for the BEST solver" );
      return 1;
   }
   const char* prec = std::getenv( "BEST_HIP_LU" );
   g_precision = ( prec && std::string( prec ) == "fp64" ) ? 64 : 32;
   const int rc = bhg_load_prepared( g_solver, n, v1, v2, v3, area, afac, cg );
   if ( rc == 0 )
      g_solver->say( "MI355X dense solver: %d triangles, %d unknowns (padded to %d)", n, 3 * n, g_solver->npad );
   g_drain();
   return rc;
}

void bhm_incenters( double* x ) {
   if ( g_solver ) bhg_incenters( g_solver, x );
   g_drain();
}

void bhm_entries( int i0, int m, int j0, int nc, double* a, int lda ) {
   if ( g_solver ) bhg_entries( g_solver, i0, m, j0, nc, a, lda );
   g_drain();
}

double bhm_diag_mean( void ) {
   const double d = g_solver ? bhg_diag_mean( g_solver ) : NAN;
   g_drain();
   return d;
}

int bhm_build( double eps, int direct, double eps_pre, int ndef, const double* gdef, const int* comp, const double* xmu,
               int single, int nrhs ) {
   (void)eps_pre;
   (void)nrhs;
   if ( !g_solver ) {
      g_log.emplace_back( "error: bhm_build before bhm_create" );
      return 1;
   }
   // The H-matrix path above kDenseMaxUnknowns (BEST_HIP_SOLVER=dense or hmatrix overrides).
   const char* which = std::getenv( "BEST_HIP_SOLVER" );
   const bool use_hm = which ? std::string( which ) == "hmatrix" : g_solver->n3 > kDenseMaxUnknowns;
   if ( use_hm ) {
      bhg_solver* s = g_solver;
      const int rc = guarded( s, "bhm_build", [&] {
         const char* tol = std::getenv( "BEST_HMATRIX_GMRES_TOL" );
         s->method = 1;
         s->hm_opt.eps = eps > 0.0 ? eps : 1e-4;
         s->hm_opt.single = single ? 1 : 0;
         s->hm_opt.tol = tol ? std::atof( tol ) : 0.0;
         if ( direct ) s->say( "the direct H-LU (-hlu) is not on the GPU path: GMRES runs to the same tolerance instead" );
         s->ndef = std::max( ndef, 1 );
         s->gdef.ensure( s->ndef );
         BH_HIP( hipMemcpyAsync( s->gdef.p, gdef, s->ndef * sizeof( double ), hipMemcpyHostToDevice, s->st ) );
         std::vector<int> ones( s->n, 1 );
         BH_HIP( hipMemcpyAsync( s->comp.p, ( comp && s->ndef > 1 ) ? comp : ones.data(), s->n * sizeof( int ),
                                 hipMemcpyHostToDevice, s->st ) );
         BH_HIP( hipMemcpyAsync( s->xmu.p, xmu, 3 * (size_t)s->n * sizeof( double ), hipMemcpyHostToDevice, s->st ) );
         BH_HIP( hipStreamSynchronize( s->st ) );
         bhg::hm_factor( s );
         bhg::size_vectors( s );
      } );
      g_drain();
      return rc;
   }
   const int prec = single ? 32 : g_precision;
   const size_t np = g_solver->npad;
   size_t freeb = 0, total = 0;
   (void)hipMemGetInfo( &freeb, &total );
   const double need = (double)np * np * ( prec == 32 ? 12.0 : 8.0 );
   if ( need > 0.95 * (double)freeb ) {
      g_solver->say( "error: not enough device memory for the dense matrix (%.1f GB needed, %.1f GB free)", need / 1e9,
                     freeb / 1e9 );
      g_drain();
      return 1;
   }
   g_solver->say( "dense GPU LU (%s)%s; eps = %g requested, the dense solve is exact to rounding",
                  prec == 32 ? "FP32 on MFMA + FP64 iterative refinement" : "FP64 on MFMA", direct ? ", direct" : "", eps );
   const int rc = bhg_factor_deflated( g_solver, ndef, gdef, comp, xmu, prec );
   g_drain();
   return rc;
}

int bhm_solve( int nrhs, double* b ) {
   if ( !g_solver ) {
      g_log.emplace_back( "error: bhm_solve before bhm_build" );
      return 1;
   }
   int rc = 0;
   const int n3 = g_solver->n3;
   for ( int c0 = 0; c0 < nrhs && rc == 0; c0 += bhg::kMaxRhs ) {
      const int k = std::min( bhg::kMaxRhs, nrhs - c0 );
      rc = bhg_besolv( g_solver, k, b + (size_t)c0 * n3, n3 );
   }
   if ( rc == 0 && g_solver->factored == 32 )
      g_solver->say( "solve: %d right-hand sides, %d refinement steps, last correction %.2e", nrhs,
                     g_solver->stats.ir_iterations, g_solver->stats.ir_correction );
   g_drain();
   return rc;
}

int bhm_log_line( char* buf, int len ) {
   g_drain();
   if ( g_log.empty() ) return -1;
   const std::string m = std::move( g_log.front() );
   g_log.pop_front();
   const int k = std::min<int>( len, (int)m.size() );
   std::memcpy( buf, m.data(), k );
   return k;
}

void bhm_destroy( void ) {
   if ( g_solver ) bhg_destroy( g_solver );
   g_solver = nullptr;
   g_log.clear();
}

}  // extern "C"
