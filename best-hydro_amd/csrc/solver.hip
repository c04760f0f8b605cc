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
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/best_hip.h"
#include "adjoint.h"
#include "bh_common.h"
#include "compon.h"
#include "files.h"
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
                   const double*, int, double*, long, float* gr32 = nullptr );
void launch_pad_identity( hipStream_t, int, int, double*, long );
void launch_pad_identity32( hipStream_t, int, int, float*, long );
void launch_entries( hipStream_t, int, double, const double*, const double*, const double*, int, int, int, int, double*,
                     long );
template <typename T>
void lu_nopiv( const LuStreams&, int, T*, long, T*, int* );
template <typename T>
void lu_pivoted( hipStream_t, int, T*, long, T*, int*, int*, double*, int* );
template <typename T>
void build_super_inverses( hipStream_t, int, int, const T*, long, const T*, T*, T*, T* );
template <typename T>
void lu_solve( hipStream_t, int, int, int, const T*, long, const T*, const T*, const int*, int, T*, T*, T*, T* );
template <typename S, typename D>
void to_interleaved( hipStream_t, int, int, const S*, long, D* );
template <typename S, typename D>
void from_interleaved( hipStream_t, int, int, const S*, D*, long );
void residual( hipStream_t, int, int, int, const double*, long, const double*, long, const double*, long, double*, long,
               double*, double* );
size_t residual_part_elems( int );
void ir_update( hipStream_t, int, int, double*, long, const float*, long, double* );
void convert_d2f( hipStream_t, long, const double*, float* );
extern thread_local int g_gemm_bk_f32;
extern thread_local double g_growth;

// Largest multiplier the exchange-free FP32 factorization accepts: iterative refinement against
// the FP64 matrix removes the effect of the larger growth (the FP64 factorization keeps partial
// pivoting's own test, 1, so its factors are DGETRF's).
constexpr double kFp32Growth = 256.0;


// Order of the diagonal super-blocks whose inverses the triangular solves use; the matrix is
// padded to a multiple of it.
constexpr int kSuper = 1024;
static_assert( kSuper / kSkinnySlice <= kSkinnyParts, "the triangular solves' split-K partials of one super-block" );
constexpr int kPairStrip = 16;  // source triangles per workgroup of the pair kernel

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

// Work vectors of one solve (solve_device and the refinement): the solver's own buffers, or the
// static work space of bhg_besolv_noalloc.
struct SolveWork {
   double *b, *x, *r, *y64, *x64il, *p64, *part, *norms;
   float *d32, *y32, *x32, *p32;
};

}  // namespace bhg

struct bhg_solver {
   int device = 0;
   hipStream_t st = nullptr;
   hipStream_t st2 = nullptr;          // look-ahead stream of the factorization
   hipStream_t st3 = nullptr;          // its U row panels
   hipEvent_t ev_u = nullptr, ev_g = nullptr;
   hipEvent_t ev_lu[ bhg::kLuEvents ] = {};
   bhg::LuStreams lu_streams() const {
      bhg::LuStreams l { st, st2, st3, ev_u, ev_g, {} };
      std::copy( ev_lu, ev_lu + bhg::kLuEvents, l.ev );
      return l;
   }
   std::string last_error;
   std::deque<std::string> log;
   // mesh
   int n = 0, n3 = 0, npad = 0, ndef = 1, ncomp = 1, nopen = 0;
   double cg = 1.0, truearea = 0.0, truevol = 0.0;
   bool mesh_raw = false;     // raw (unshifted) vertices are resident (bhg_set_mesh)
   bool mesh_ready = false;   // prepared vertices, areas, incenters and self blocks are resident
   bhg::DevBuf<double> raw_v, v, area, xc, selfb, xmu, src, gdef, scal;
   bhg::DevBuf<int> comp;     // the deflation's surface of each triangle (all 1 when a surface is open)
   bhg::DevBuf<int> icomp;    // COMPON's ICOMP
   bhg::ComponWork compw;
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
   bool a32_fresh = false;    // a32 holds the FP32 rounding of the assembled a64 (the pair kernel stored both)
   bool transposed = false;   // the factors (and the FP64 matrix) are transposed in place: solves are with GR^T
   // vectors
   bhg::DevBuf<double> b, x, r, y64, x64il, p64, part, norms, sums, trdev;
   bhg::DevBuf<float> d32, y32, x32, p32;
   // adjoint (bhg_gradient)
   bhg::DevBuf<double> adj_F, adj_y, adj_lam, adj_ghat, adj_gsrc_part, adj_gcol_part, adj_gsrc, adj_gcol;
   bhg::DevBuf<bhg::AdjTri> adj_tri;
   std::vector<double> host_sums;
   bhg::PhaseTimer timer;
   bhg::LaunchProfiler prof;
   bhg_stats stats {};
   double ir_tol = 1e-12;
   int ir_maxit = 30;
   double fp32_growth = bhg::kFp32Growth;  // largest multiplier the exchange-free FP32 LU accepts
   int ir_steps_nopiv = 0;                 // refinement steps with those factors before the pivoted
                                           // refactorization (0: ir_maxit)
   int host_algebra = 0;                   // 1: the reductions summed and BEST's algebra done on the
                                           // host (a check of the device algebra)
   // The refinement steps the last FP32 dense solve of the resident mesh took, and the arguments
   // its matrix and right-hand sides came from: a solve with the same ones enqueues that many
   // steps before its first host check (0: none known; any setter that changes the matrix or the
   // refinement clears it).  ir_ahead is the count in force for the solve running now.
   int ir_hint = 0, hint_iarea = 0, hint_nrhs = 0;
   double hint_cg = 0.0;
   int ir_ahead = 0;
   bool flag_pending = false;  // the FP32 LU's multiplier check is read at the first host check
   const bhg::SolveWork* work_override = nullptr;  // solve_device's work vectors when not the buffers above

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

// Fills by kernel rather than hipMemsetAsync, so the solve can be captured in a HIP graph.
__global__ void k_fill_d( long n, double* __restrict__ p, double v ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) p[ i ] = v;
}
__global__ void k_fill_i( long n, int* __restrict__ p, int v ) {
   for ( long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x ) p[ i ] = v;
}
void fill( hipStream_t st, double* p, long n, double v ) {
   hipLaunchKernelGGL( k_fill_d, dim3( (unsigned)std::min<long>( 2048, ( n + 255 ) / 256 ) ), dim3( 256 ), 0, st, n, p, v );
}
void fill( hipStream_t st, int* p, long n, int v ) {
   hipLaunchKernelGGL( k_fill_i, dim3( (unsigned)std::min<long>( 2048, ( n + 255 ) / 256 ) ), dim3( 256 ), 0, st, n, p, v );
}

// The refinement's norms: the last step's 32 values move to norms[32..64) and norms[0..32) is
// cleared for the next step.
__global__ void k_ir_shift( double* norms ) {
   const int k = threadIdx.x;
   if ( k < 32 ) {
      norms[ 32 + k ] = norms[ k ];
      norms[ k ] = 0.0;
   }
}

// ------------------------------------------------------------------ the static work space of
// bhg_besolv_noalloc: every work vector of one solve (the FP64 right-hand sides, solutions and
// residuals, the split-K partial sums of the triangular solves in both precisions, the residual
// product's partial sums, the refinement's norms) for matrices of order up to kNoAllocMaxPad and
// up to kMaxRhs right-hand sides, in one file-scope device array (101 MiB on each GPU the library
// runs on, held until the process ends).  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (noalloc_scratch below).
constexpr long kNoAllocMaxPad = 16384;
constexpr long kNaVec = kNoAllocMaxPad * kMaxRhs;                      // one npad x 16 block
constexpr long kNaPart = ( kNoAllocMaxPad / 512 ) * 16 * kNoAllocMaxPad;  // residual partials at that order
constexpr size_t kNaDoubles = ( 5 + kSkinnyParts ) * (size_t)kNaVec + (size_t)kNaPart + 64;  // b x r y64 x64il, p64, part, norms
constexpr size_t kNaFloats = ( 3 + kSkinnyParts ) * (size_t)kNaVec;                          // d32 y32 x32, p32
constexpr size_t kNaBytes = ( ( kNaDoubles * sizeof( double ) + kNaFloats * sizeof( float ) ) + 255 ) / 256 * 256;
static __device__ __align__( 256 ) char scratch[ kNaBytes ];

// b (npad x nrhs, leading dimension npad) = the caller's n3 x nrhs block, zero below row n3.
__global__ void k_noalloc_copy_in( int n3, int npad, int nrhs, const double* __restrict__ src, long lds, double* __restrict__ b ) {
   const long total = (long)npad * nrhs;
   for ( long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x ) {
      const long i = e % npad, j = e / npad;
      b[ e ] = i < n3 ? src[ j * lds + i ] : 0.0;
   }
}

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

// also32: the pair kernel also stores the FP32 rounding of every entry into a32, the copy the FP32
// factorization consumes (so it needs no conversion pass).
void assemble( bhg_solver* s, bool deflate, bool also32 = false ) {
   const size_t np = s->npad;
   s->a64.ensure( np * np );
   if ( also32 ) s->a32.ensure( np * np );
   launch_pairs( s->st, s->n, kPairStrip, s->cg, s->src.p, s->xc.p, s->selfb.p, s->xmu.p, s->comp.p, s->gdef.p, deflate ? 1 : 0,
                 s->a64.p, (long)np, also32 ? s->a32.p : nullptr );
   launch_pad_identity( s->st, s->n3, s->npad, s->a64.p, (long)np );
   if ( also32 ) launch_pad_identity32( s->st, s->n3, s->npad, s->a32.p, (long)np );
   s->deflated = deflate;
   s->a32_fresh = also32;
}

int read_flag( bhg_solver* s ) {
   int h = 0;
   BH_HIP( hipMemcpyAsync( &h, s->flags.p, sizeof( int ), hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   return h;
}

// Factors the assembled matrix: precision 64 in place on a64 (GR is consumed), 32 on an FP32
// copy (a64 is kept for the refinement residuals).  Partial pivoting falls back to the slow path
// when the exchange-free factorization is not GEPP's (see lu.hip).  With defer_flag (FP32) the
// multiplier check is not read here: solve_device reads it at the refinement's first host check
// and refactors with row exchanges then.
void factor( bhg_solver* s, int precision, bool defer_flag = false ) {
   s->hm.reset();
   s->flag_pending = false;
   const size_t np = s->npad;
   const int nblk = s->npad / kNB;
   s->ipiv.ensure( np );
   fill( s->st, s->flags.p, 4, 0 );
   s->pivoted = false;
   s->transposed = false;
   const bool fresh32 = s->a32_fresh;
   s->a32_fresh = false;
   if ( precision == 64 ) {
      s->inv64.ensure( (size_t)nblk * 2 * kNB * kNB );
      lu_nopiv<double>( s->lu_streams(), s->npad, s->a64.p, (long)np, s->inv64.p, s->flags.p );
      if ( read_flag( s ) ) {
         s->say( "LU: a multiplier exceeded 1; refactoring with row exchanges" );
         assemble( s, s->deflated );
         fill( s->st, s->flags.p, 4, 0 );
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
      if ( !fresh32 ) convert_d2f( s->st, (long)( np * np ), s->a64.p, s->a32.p );
      g_growth = s->fp32_growth;
      lu_nopiv<float>( s->lu_streams(), s->npad, s->a32.p, (long)np, s->inv32.p, s->flags.p );
      g_growth = 1.0;
      if ( defer_flag ) {
         s->flag_pending = true;
      } else if ( read_flag( s ) ) {
         s->say( "LU (FP32): a multiplier exceeded %g; refactoring with row exchanges", s->fp32_growth );
         convert_d2f( s->st, (long)( np * np ), s->a64.p, s->a32.p );
         fill( s->st, s->flags.p, 4, 0 );
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
   s->hm = hm_build( s->st, s->st2, s->st3, in, s->hm_opt, &s->hm_stats, &lines );
   for ( auto& l : lines ) s->log.push_back( std::move( l ) );
   s->factored = -1;
   s->pivoted = false;
   s->transposed = false;
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
   s->p64.ensure( need * kSkinnyParts );
   s->d32.ensure( need );
   s->y32.ensure( need );
   s->x32.ensure( need );
   s->p32.ensure( need * kSkinnyParts );
   s->part.ensure( residual_part_elems( s->npad ) );
   s->norms.ensure( 64 );
   s->sums.ensure( (size_t)kSumGroups * kSumSlots );
   s->trdev.ensure( sizeof( Transport ) / sizeof( double ) );

   if ( grow ) {
      BH_HIP( hipMemsetAsync( s->b.p, 0, need * sizeof( double ), s->st ) );
      BH_HIP( hipMemsetAsync( s->x.p, 0, need * sizeof( double ), s->st ) );
   }
}

// The work vectors solve_device uses.
SolveWork work( bhg_solver* s ) {
   if ( s->work_override ) return *s->work_override;
   return SolveWork { s->b.p, s->x.p, s->r.p, s->y64.p, s->x64il.p, s->p64.p, s->part.p, s->norms.p,
                      s->d32.p, s->y32.p, s->x32.p, s->p32.p };
}

// How lu_solve applies the row exchanges of the resident factors.
int piv_mode( const bhg_solver* s ) { return s->pivoted ? ( s->transposed ? -1 : 1 ) : 0; }

// One refinement step of the FP32 path: x0 = LU32^-1 b (it == 0), else r = b - A x in FP64 and
// x += LU32^-1 r; norms[k] = max |dx|, norms[16 + k] = max |x| per column, norms[32 .. 64) the
// previous step's.  No host interaction.
void ir_step( bhg_solver* s, int nrhs, int it ) {
   const long np = s->npad;
   const SolveWork w = work( s );
   hipLaunchKernelGGL( k_ir_shift, dim3( 1 ), dim3( 64 ), 0, s->st, w.norms );
   if ( it == 0 ) {
      fill( s->st, w.x, np * nrhs, 0.0 );
      to_interleaved<double, float>( s->st, s->npad, nrhs, w.b, np, w.d32 );
   } else {
      residual( s->st, s->npad, s->n3, nrhs, s->a64.p, np, w.x, np, w.b, np, w.r, np, w.part, w.norms );
      to_interleaved<double, float>( s->st, s->npad, nrhs, w.r, np, w.d32 );
      fill( s->st, w.norms, 32, 0.0 );
   }
   lu_solve<float>( s->st, s->npad, s->n3, kSuper, s->a32.p, np, s->xl32.p, s->xu32.p, s->ipiv.p, piv_mode( s ), w.d32, w.y32,
                    w.x32, w.p32 );
   ir_update( s->st, s->n3, nrhs, w.x, np, w.x32, np, w.norms );
}

// The last refinement step, per column k: the correction c_k = max |dx| / max |x| and, from the
// step before (correction p_k), the error left in x estimated as c_k (c_k / p_k): refinement
// converges linearly, each step's correction being the error of the x before it, so c_k / p_k
// estimates the contraction.  Columns that do not contract by half or more get no estimate.
struct IrCorrection {
   double worst = 0.0;  // max_k c_k
   double est = 0.0;    // max_k c_k^2 / p_k (infinite without a contraction estimate)
};
IrCorrection ir_correction( bhg_solver* s, int nrhs ) {
   double hn[ 64 ];
   BH_HIP( hipMemcpyAsync( hn, work( s ).norms, sizeof hn, hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   IrCorrection r;
   for ( int k = 0; k < nrhs; ++k ) {
      const double c = hn[ 16 + k ] > 0 ? hn[ k ] / hn[ 16 + k ] : 0.0;
      const double p = hn[ 48 + k ] > 0 ? hn[ 32 + k ] / hn[ 48 + k ] : 0.0;
      if ( !std::isfinite( c ) ) throw std::runtime_error( "iterative refinement produced a non-finite correction" );
      r.worst = std::max( r.worst, c );
      const double e = c == 0.0 ? 0.0 : ( p > 0.0 && c < 0.5 * p ) ? c * ( c / p ) : HUGE_VAL;
      r.est = std::max( r.est, e );
   }
   return r;
}

// Converged after step it: the last correction, or the error it leaves estimated from the two
// last corrections, below tol (step 0's correction is 1, the whole of x0).
bool ir_converged( const IrCorrection& c, int it, double tol ) { return it > 0 && ( c.worst < tol || c.est < tol ); }

// FP64 iterative refinement with the FP32 factors from step it0 up to step maxit, until the
// correction, or the error estimated to remain, is below ir_tol relative to x in every column.
// A refinement that stagnates (the correction no longer halves) below kIrFloor is accepted: the
// FP64 residuals themselves then limit it; the log says so and bhg_stats.ir_above_tol is set.
// Returns false if it stopped above kIrFloor (the FP32 factors too poor for the refinement).
constexpr double kIrFloor = 1e-10;
enum IrOutcome { kIrConverged, kIrFailed, kIrFlagged };
IrOutcome ir_run( bhg_solver* s, int nrhs, int it0, int maxit ) {
   double prev = 1e300, worst = 1e300;
   // Host checks begin after step `first`: step 0's correction is the whole of x0 and decides
   // nothing, so step 1 follows it without a round trip (its norms stay on the device, where step
   // 1's estimate reads them); a solve that knows how many steps its matrix needs (ir_ahead)
   // enqueues them all first.
   const int first = std::min( maxit, it0 + ( it0 == 0 ? std::max( 2, s->ir_ahead ) : 1 ) ) - 1;
   for ( int it = it0; it < maxit; ++it ) {
      ir_step( s, nrhs, it );
      if ( it < first ) continue;
      if ( s->flag_pending ) {  // before any correction is read: flagged factors may give non-finite ones
         s->flag_pending = false;
         if ( read_flag( s ) ) return kIrFlagged;
      }
      const IrCorrection c = ir_correction( s, nrhs );
      worst = c.worst;
      s->stats.ir_iterations = it + 1;
      s->stats.ir_correction = worst;
      if ( ir_converged( c, it, s->ir_tol ) ) return kIrConverged;
      if ( it > 2 && worst > 0.5 * prev ) break;  // stagnation
      prev = worst;
   }
   if ( worst < kIrFloor ) {
      s->stats.ir_above_tol = 1;
      s->say( "refinement: stagnated at a correction of %.2e, above the tolerance %.0e; accepted (below %.0e)", worst,
              s->ir_tol, kIrFloor );
      return kIrConverged;
   }
   return kIrFailed;
}
bool ir_loop( bhg_solver* s, int nrhs, int it0, int maxit ) {
   const IrOutcome o = ir_run( s, nrhs, it0, maxit );
   if ( o == kIrFlagged ) throw std::runtime_error( "refinement: the multiplier check was still pending" );
   return o == kIrConverged;
}
bool ir_loop( bhg_solver* s, int nrhs, int it0 ) { return ir_loop( s, nrhs, it0, s->ir_maxit ); }

// The FP32 factors of the resident (deflated) matrix with row exchanges, after the exchange-free
// ones did not let the refinement converge.
void refactor_pivoted32( bhg_solver* s, bool flagged = false ) {
   const size_t np = s->npad;
   if ( flagged )
      s->say( "LU (FP32): a multiplier exceeded %g; refactoring with row exchanges", s->fp32_growth );
   else
      s->say( "LU (FP32): refinement did not converge with the exchange-free factors; refactoring with row exchanges" );
   convert_d2f( s->st, (long)( np * np ), s->a64.p, s->a32.p );
   fill( s->st, s->flags.p, 4, 0 );
   lu_pivoted<float>( s->st, s->npad, s->a32.p, (long)np, s->inv32.p, s->ipiv.p, s->flags.p + 1, s->pval.p, s->pidx.p );
   s->pivoted = true;
   build_super_inverses<float>( s->st, s->npad, kSuper, s->a32.p, (long)np, s->inv32.p, s->xl32.p, s->xu32.p, s->tmp32.p );
}

// Solves for the nrhs columns of b (npad x nrhs, rows >= n3 zero) into x (the work vectors:
// s->b and s->x unless overridden).
void solve_device( bhg_solver* s, int nrhs ) {
   const long np = s->npad;
   const SolveWork w = work( s );
   s->stats.ir_iterations = 0;
   s->stats.ir_correction = 0.0;
   s->stats.ir_above_tol = 0;
   if ( s->factored == -1 ) {
      if ( s->work_override ) throw std::runtime_error( "the H-matrix solve needs its own work space" );
      std::vector<std::string> lines;
      hm_solve( *s->hm, s->st, nrhs, s->b.p, np, s->x.p, np, &s->hm_stats, &lines );
      for ( auto& l : lines ) s->log.push_back( std::move( l ) );
      return;
   }
   if ( s->factored == 64 ) {
      to_interleaved<double, double>( s->st, s->npad, nrhs, w.b, np, w.r );
      lu_solve<double>( s->st, s->npad, s->n3, kSuper, s->a64.p, np, s->xl64.p, s->xu64.p, s->ipiv.p, piv_mode( s ), w.r,
                        w.y64, w.x64il, w.p64 );
      from_interleaved<double, double>( s->st, s->n3, nrhs, w.x64il, w.x, np );
      return;
   }
   if ( s->factored != 32 ) throw std::runtime_error( "solve before the factorization" );
   const int first = s->ir_steps_nopiv > 0 && !s->pivoted ? std::min( s->ir_steps_nopiv, s->ir_maxit ) : s->ir_maxit;
   const IrOutcome o = ir_run( s, nrhs, 0, first );
   if ( o == kIrConverged ) return;
   if ( o == kIrFailed && ( s->pivoted || s->transposed ) ) throw std::runtime_error( "iterative refinement did not converge" );
   // the exchange-free factors had a multiplier above the threshold (read only now), or did not
   // let the refinement converge: the same matrix with row exchanges
   refactor_pivoted32( s, o == kIrFlagged );
   s->ir_ahead = 0;
   if ( !ir_loop( s, nrhs, 0 ) ) throw std::runtime_error( "iterative refinement did not converge" );
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

// out from the transport values and the prep's scalars (afac, total area, centroid, diagm).
void fill_result( const bhg_solver* s, int iarea, const Transport& tr, const double* scal, bhg_result* out ) {
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

// Copies the transport values the device computed (launch_transport into s->trdev) and the prep's
// scalars to the host, into out.  Synchronizes.
void read_results( bhg_solver* s, int iarea, bhg_result* out ) {
   Transport tr {};
   double scal[ 8 ];
   BH_HIP( hipMemcpyAsync( &tr, s->trdev.p, sizeof tr, hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipMemcpyAsync( scal, s->scal.p, sizeof scal, hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   if ( out ) fill_result( s, iarea, tr, scal, out );
}

// The DIFCO/VISCO reductions of the solutions in s->x and BEST's algebra on them (best/difco.f,
// best/visco.f, best/output.f) into out, on the device; with host_algebra the partial sums are
// added and the algebra done on the host instead.  Synchronizes.
void compute_results( bhg_solver* s, int iarea, double cg, int ivisco, int icov, bhg_result* out ) {
   const int n = s->n;
   const long np = s->npad;
   if ( !s->host_algebra ) {
      launch_transport( s->st, n, s->xc.p, s->area.p, s->x.p, np, ivisco, icov, cg, s->sums.p,
                        reinterpret_cast<Transport*>( s->trdev.p ) );
      read_results( s, iarea, out );
      return;
   }
   const double zero[ 3 ] = { 0.0, 0.0, 0.0 };
   launch_sums( s->st, n, s->xc.p, s->area.p, s->x.p, np, ivisco ? 5 : 0, zero, icov ? 1 : 0, s->sums.p );
   double sums[ kSumSlots ];
   collect_sums( s, sums );
   Transport tr {};
   difco_algebra( sums, cg, tr );
   if ( ivisco ) {
      if ( icov ) {
         visco_centre( sums, tr.dvisc );
         launch_sums( s->st, n, s->xc.p, s->area.p, s->x.p, np, 5, tr.dvisc, 0, s->sums.p );
         double sums2[ kSumSlots ];
         collect_sums( s, sums2 );
         tr.chiprime = visco_chiprime( sums2, cg );
      } else {
         tr.chiprime = visco_chiprime( sums, cg );
      }
   }
   output_values( tr );
   double scal[ 8 ];
   BH_HIP( hipMemcpyAsync( scal, s->scal.p, sizeof scal, hipMemcpyDeviceToHost, s->st ) );
   BH_HIP( hipStreamSynchronize( s->st ) );
   if ( out ) fill_result( s, iarea, tr, scal, out );
}

// Tile choices of one mesh of a series (picked by the autotuner).
struct MeshTune {
   int strip = kPairStrip;  // source triangles per workgroup of the pair kernel
   int gemm_bk = 0;         // K slice of the FP32 GEMM (0: 16 on the 128 x 128 tiles, 32 on the others)
   int ir_iters = 4;        // refinement steps in the captured schedule
};

// Every buffer of the FP32 path for the resident mesh, so that nothing is allocated once a
// schedule is captured.
void prepare_buffers( bhg_solver* s ) {
   const size_t np = s->npad;
   const int nblk = s->npad / kNB;
   s->hm.reset();
   s->a64.ensure( np * np );
   s->a32.ensure( np * np );
   s->inv32.ensure( (size_t)nblk * 2 * kNB * kNB );
   s->xl32.ensure( np * kSuper );
   s->xu32.ensure( np * kSuper );
   s->tmp32.ensure( np * kSuper / 2 );
   s->ipiv.ensure( np );
   s->gdef.ensure( std::max( s->ndef, 1 ) );
   size_vectors( s );
}

// The whole FP32-LU + FP64-refinement solve of the resident mesh, enqueued on s->st (and s->st2)
// with no host interaction: nothing here reads a device value, allocates or sets memory, so the
// sequence can be captured in a HIP graph.  The pivot check and the refinement's convergence are
// read afterwards (finish_enqueued).
void enqueue_solve( bhg_solver* s, int iarea, double cg, int ivisco, int icov, const MeshTune& t,
                    hipEvent_t late = nullptr ) {
   const int n = s->n;
   const long np = s->npad;
   g_gemm_bk_f32 = t.gemm_bk;
   s->cg = cg;
   BH_HIP( hipMemcpyAsync( s->v.p, s->raw_v.p, 9 * (size_t)n * sizeof( double ), hipMemcpyDeviceToDevice, s->st ) );
   launch_prep( s->st, n, V1( s ), V2( s ), V3( s ), iarea, s->truearea, s->area.p, s->scal.p );
   self_terms( s, true, false );
   launch_pairs( s->st, n, t.strip, s->cg, s->src.p, s->xc.p, s->selfb.p, s->xmu.p, s->comp.p, s->gdef.p, 1, s->a64.p, np,
                 s->a32.p );
   launch_pad_identity( s->st, s->n3, s->npad, s->a64.p, np );
   launch_pad_identity32( s->st, s->n3, s->npad, s->a32.p, np );
   s->deflated = true;
   s->a32_fresh = false;
   fill( s->st, s->flags.p, 4, 0 );
   g_growth = s->fp32_growth;
   LuStreams ls = s->lu_streams();
   ls.ev_late = late;  // the staggered series starts the next mesh here (half of the rows left)
   ls.late_rest = s->npad / 2;
   lu_nopiv<float>( ls, s->npad, s->a32.p, np, s->inv32.p, s->flags.p );
   g_growth = 1.0;
   build_super_inverses<float>( s->st, s->npad, kSuper, s->a32.p, np, s->inv32.p, s->xl32.p, s->xu32.p, s->tmp32.p );
   launch_rhs( s->st, n, s->xc.p, s->b.p, np, ivisco ? 1 : 0 );
   const int nrhs = ivisco ? 11 : 6;
   for ( int it = 0; it < t.ir_iters; ++it ) ir_step( s, nrhs, it );
   // the reductions and the algebra on the solutions as they stand (used when no fallback runs)
   launch_transport( s->st, n, s->xc.p, s->area.p, s->x.p, np, ivisco, icov, cg, s->sums.p,
                     reinterpret_cast<Transport*>( s->trdev.p ) );
   g_gemm_bk_f32 = 0;
}

// After an enqueued solve has run: if partial pivoting would have exchanged rows, the mesh is
// solved again on the pivoted path; if the refinement has not converged it continues, and the
// reductions and BEST's algebra are done again on its solutions.  Otherwise the enqueued ones
// stand.
void finish_enqueued( bhg_solver* s, int iarea, double cg, int ivisco, int icov, const MeshTune& t, bhg_result* out ) {
   const int nrhs = ivisco ? 11 : 6;
   s->factored = 32;
   s->pivoted = false;
   s->transposed = false;
   s->mesh_ready = true;
   s->stats.ir_above_tol = 0;
   bool redo = true;
   if ( read_flag( s ) ) {
      s->say( "series: a multiplier exceeded %g; the mesh is solved again with row exchanges", s->fp32_growth );
      assemble( s, true, true );
      factor( s, 32 );
      launch_rhs( s->st, s->n, s->xc.p, s->b.p, s->npad, ivisco ? 1 : 0 );
      solve_device( s, nrhs );
   } else {
      const IrCorrection c = ir_correction( s, nrhs );
      s->stats.ir_iterations = t.ir_iters;
      s->stats.ir_correction = c.worst;
      if ( ir_converged( c, t.ir_iters - 1, s->ir_tol ) ) {
         redo = s->host_algebra != 0;
      } else if ( !ir_loop( s, nrhs, t.ir_iters ) ) {
         refactor_pivoted32( s );
         if ( !ir_loop( s, nrhs, 0 ) ) throw std::runtime_error( "iterative refinement did not converge" );
      }
   }
   s->stats.pivoted = s->pivoted ? 1 : 0;
   s->stats.npad = s->npad;
   if ( redo ) compute_results( s, iarea, cg, ivisco, icov, out );
   else read_results( s, iarea, out );
}

// One dense (or H-matrix) solve of the resident mesh: prep, self terms, assembly, factorization,
// the 6 or 11 right-hand sides, reductions and algebra.
void solve_mesh( bhg_solver* s, int iarea, double cg, int ivisco, int icov, int precision, bhg_result* out ) {
   if ( !s->mesh_raw ) throw std::runtime_error( "no mesh (bhg_set_mesh)" );
   if ( precision != 32 && precision != 64 ) throw std::runtime_error( "precision must be 32 or 64" );
   set_device( s );
   const int n = s->n;
   const long np = s->npad;
   s->cg = cg;
   s->flag_pending = false;
   struct ProfBind {
      explicit ProfBind( LaunchProfiler* p ) { g_prof = p->on ? p : nullptr; }
      ~ProfBind() { g_prof = nullptr; }
   } bind( &s->prof );
   s->prof.st = s->st;
   auto& T = s->timer;
   T.mark( s->st, 0 );
   BH_HIP( hipMemcpyAsync( s->v.p, s->raw_v.p, 9 * (size_t)n * sizeof( double ), hipMemcpyDeviceToDevice, s->st ) );
   launch_prep( s->st, n, V1( s ), V2( s ), V3( s ), iarea, s->truearea, s->area.p, s->scal.p );
   self_terms( s, true, false );
   s->mesh_ready = true;
   T.mark( s->st, 1 );
   if ( s->method == 1 ) {  // the H-matrix build is the factorization phase
      T.mark( s->st, 2 );
      hm_factor( s );
   } else {
      assemble( s, true, precision == 32 );
      T.mark( s->st, 2 );
      factor( s, precision, precision == 32 );
   }
   T.mark( s->st, 3 );
   size_vectors( s );
   const int nrhs = ivisco ? 11 : 6;
   const bool dense32 = s->method == 0 && precision == 32;
   s->ir_ahead = dense32 && s->hint_iarea == iarea && s->hint_cg == cg && s->hint_nrhs == nrhs ? s->ir_hint : 0;
   launch_rhs( s->st, n, s->xc.p, s->b.p, np, ivisco ? 1 : 0 );
   solve_device( s, nrhs );
   s->ir_ahead = 0;
   if ( dense32 ) {
      s->ir_hint = s->pivoted ? 0 : s->stats.ir_iterations;
      s->hint_iarea = iarea;
      s->hint_cg = cg;
      s->hint_nrhs = nrhs;
   }
   T.mark( s->st, 4 );
   compute_results( s, iarea, cg, ivisco, icov, out );
   T.mark( s->st, 5 );
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
}


// ------------------------------------------------------------------ adjoint

// ghat[o][q] = d(output o)/d(slot q), slots 0..44 (DIFCO's forces and torques, the CHI moments;
// no centre-of-viscosity move): exact derivatives of the algebra (algebra_gradient, forward mode).
void adjoint_weights( const double* sums, double cg, double* ghat ) { algebra_gradient( sums, cg, kAdjSlots, ghat ); }

template <typename T>
void swap_bufs( DevBuf<T>& a, DevBuf<T>& b ) {
   std::swap( a.p, b.p );
   std::swap( a.n, b.n );
}

// Transposes the resident factorization in place: P GR = L U becomes the factors of
// GR^T = U^T L^T P (the packed factors transposed, the super-block inverses of L and U
// transposed and exchanged); on the FP32 path the FP64 matrix of the refinement residuals too.
// Applied twice it restores the factors.
void transpose_factors( bhg_solver* s ) {
   const int np = s->npad, nsb = s->npad / kSuper;
   const long SS = (long)kSuper * kSuper;
   if ( s->factored == 64 ) {
      launch_transpose<double>( s->st, np, s->a64.p, np, 1, 0 );
      launch_transpose<double>( s->st, kSuper, s->xl64.p, kSuper, nsb, SS );
      launch_transpose<double>( s->st, kSuper, s->xu64.p, kSuper, nsb, SS );
      swap_bufs( s->xl64, s->xu64 );
   } else if ( s->factored == 32 ) {
      launch_transpose<double>( s->st, np, s->a64.p, np, 1, 0 );
      launch_transpose<float>( s->st, np, s->a32.p, np, 1, 0 );
      launch_transpose<float>( s->st, kSuper, s->xl32.p, kSuper, nsb, SS );
      launch_transpose<float>( s->st, kSuper, s->xu32.p, kSuper, nsb, SS );
      swap_bufs( s->xl32, s->xu32 );
   } else {
      throw std::runtime_error( "the gradient needs the dense factorization" );
   }
   s->transposed = !s->transposed;
}

// The gradient of D_t, D_r and CHI * V with respect to the raw vertex coordinates, after a dense
// solve of the resident mesh with 11 right-hand sides (solutions in s->x):
//   grad[o * 9 n + v * 3 n + c * n + i] = d(output o) / d(vertex v of triangle i, coordinate c).
void gradient( bhg_solver* s, int iarea, double cg, double* grad ) {
   const int n = s->n, n3 = s->n3;
   const long np = s->npad;
   constexpr int K = 11, O = kAdjOutputs;
   hipStream_t st = s->st;
   // solutions and the output weights of the reductions
   s->adj_F.ensure( np * K );
   BH_HIP( hipMemcpyAsync( s->adj_F.p, s->x.p, np * K * sizeof( double ), hipMemcpyDeviceToDevice, st ) );
   const double zero[ 3 ] = { 0.0, 0.0, 0.0 };
   launch_sums( st, n, s->xc.p, s->area.p, s->x.p, np, 5, zero, 0, s->sums.p );
   double sums[ kSumSlots ];
   collect_sums( s, sums );
   double ghat[ O * kAdjSlots ];
   adjoint_weights( sums, cg, ghat );
   s->adj_ghat.ensure( O * kAdjSlots );
   BH_HIP( hipMemcpyAsync( s->adj_ghat.p, ghat, sizeof ghat, hipMemcpyHostToDevice, st ) );
   // adjoint right-hand sides and solves with the transposed factors
   s->adj_y.ensure( np * K * O );
   s->adj_lam.ensure( np * K * O );
   auto& T = s->timer;
   T.mark( st, 0 );
   fill( st, s->adj_y.p, np * K * O, 0.0 );
   launch_adj_rhs( st, n, s->xc.p, s->area.p, s->adj_ghat.p, s->adj_y.p, np );
   transpose_factors( s );
   try {
      for ( int o = 0; o < O; ++o ) {
         BH_HIP( hipMemcpyAsync( s->b.p, s->adj_y.p + o * K * np, np * K * sizeof( double ), hipMemcpyDeviceToDevice, st ) );
         solve_device( s, K );
         BH_HIP( hipMemcpyAsync( s->adj_lam.p + o * K * np, s->x.p, np * K * sizeof( double ), hipMemcpyDeviceToDevice, st ) );
      }
   } catch ( ... ) {
      transpose_factors( s );
      throw;
   }
   transpose_factors( s );
   BH_HIP( hipMemcpyAsync( s->x.p, s->adj_F.p, np * K * sizeof( double ), hipMemcpyDeviceToDevice, st ) );
   T.mark( st, 1 );
   // pair terms (source strips sized for at least 16,384 workgroups, so that the last round of
   // the launch is short: 27.5 ms at 5,120 triangles against 28.3 with 1,024) and per-triangle terms
   const int nb = ( n + 255 ) / 256;
   const int strip = std::max( 8, std::min( 128, (int)( ( (long)n * nb + 16383 ) / 16384 ) ) );
   const int ns = ( n + strip - 1 ) / strip;
   s->adj_gsrc_part.ensure( (size_t)nb * n * O * 10 );
   s->adj_gcol_part.ensure( (size_t)ns * n * O * 3 );
   s->adj_gsrc.ensure( (size_t)n * O * 10 );
   s->adj_gcol.ensure( (size_t)n * O * 3 );
   s->adj_tri.ensure( n );
   launch_adj_pairs( st, n, strip, s->cg, s->src.p, s->xc.p, s->adj_lam.p, s->adj_F.p, np, s->adj_gsrc_part.p, s->adj_gcol_part.p,
                     s->adj_gsrc.p, s->adj_gcol.p );
   T.mark( st, 2 );
   launch_adj_tri( st, n, V1( s ), V2( s ), V3( s ), s->scal.p, s->cg, s->adj_lam.p, s->adj_F.p, np, s->adj_ghat.p,
                   s->adj_tri.p );
   T.mark( st, 3 );
   // to the host: the chain rule through the deflation, the area factor and the centroid shift
   std::vector<double> gsrc( (size_t)n * O * 10 ), gcol( (size_t)n * O * 3 ), lam( (size_t)n3 * K * O ), F( (size_t)n3 * K );
   std::vector<AdjTri> tri( n );
   std::vector<int> comp( n );
   const int ndef = std::max( s->ndef, 1 );
   std::vector<double> gdef( ndef );
   double scal[ 8 ];
   BH_HIP( hipMemcpyAsync( gsrc.data(), s->adj_gsrc.p, gsrc.size() * sizeof( double ), hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpyAsync( gcol.data(), s->adj_gcol.p, gcol.size() * sizeof( double ), hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpyAsync( tri.data(), s->adj_tri.p, n * sizeof( AdjTri ), hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpy2DAsync( lam.data(), n3 * sizeof( double ), s->adj_lam.p, np * sizeof( double ), n3 * sizeof( double ), K * O,
                             hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpy2DAsync( F.data(), n3 * sizeof( double ), s->adj_F.p, np * sizeof( double ), n3 * sizeof( double ), K,
                             hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpyAsync( comp.data(), s->comp.p, n * sizeof( int ), hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpyAsync( gdef.data(), s->gdef.p, ndef * sizeof( double ), hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipMemcpyAsync( scal, s->scal.p, sizeof scal, hipMemcpyDeviceToHost, st ) );
   BH_HIP( hipStreamSynchronize( st ) );
   s->stats.t_adjoint_solve = T.ms( 0, 1 );
   s->stats.t_adjoint_pairs = T.ms( 1, 2 );
   s->stats.t_adjoint_tri = T.ms( 2, 3 );
   const double afac = scal[ 0 ], tot = scal[ 1 ], diagm = scal[ 5 ];
   auto Lm = [&]( int o, int k, int i, int c ) { return lam[ ( (size_t)o * K + k ) * n3 + 3 * (size_t)i + c ]; };
   auto Fm = [&]( int k, int i, int c ) { return F[ (size_t)k * n3 + 3 * (size_t)i + c ]; };
   auto surf = [&]( int i ) { return ndef > 1 ? comp[ i ] - 1 : 0; };
   // deflation: per surface, X = sum |xmu|^2, alpha = lambda . xmu, beta = F . xmu, P = alpha . beta
   std::vector<double> X( ndef, 0.0 ), alpha( (size_t)O * K * ndef, 0.0 ), beta( (size_t)K * ndef, 0.0 );
   for ( int i = 0; i < n; ++i ) {
      const int S = surf( i );
      const double* xm = tri[ i ].xmu;
      X[ S ] += xm[ 0 ] * xm[ 0 ] + xm[ 1 ] * xm[ 1 ] + xm[ 2 ] * xm[ 2 ];
      for ( int k = 0; k < K; ++k ) {
         beta[ k * ndef + S ] += Fm( k, i, 0 ) * xm[ 0 ] + Fm( k, i, 1 ) * xm[ 1 ] + Fm( k, i, 2 ) * xm[ 2 ];
         for ( int o = 0; o < O; ++o )
            alpha[ ( o * K + k ) * ndef + S ] += Lm( o, k, i, 0 ) * xm[ 0 ] + Lm( o, k, i, 1 ) * xm[ 1 ] + Lm( o, k, i, 2 ) * xm[ 2 ];
      }
   }
   std::vector<double> P( (size_t)O * ndef, 0.0 );
   double Pt[ O ] = {};
   for ( int o = 0; o < O; ++o )
      for ( int S = 0; S < ndef; ++S ) {
         for ( int k = 0; k < K; ++k ) P[ o * ndef + S ] += alpha[ ( o * K + k ) * ndef + S ] * beta[ k * ndef + S ];
         Pt[ o ] += gdef[ S ] * P[ o * ndef + S ];
      }
   const double ddiag = 1.0 / ( 3.0 * n * diagm );
   double dafac[ O ] = {};
   const size_t n9 = 9 * (size_t)n;
   for ( int i = 0; i < n; ++i ) {
      const AdjTri& R = tri[ i ];
      const int S = surf( i );
      const double g = gdef[ S ];
      for ( int o = 0; o < O; ++o ) {
         double w[ 3 ];
         for ( int c = 0; c < 3; ++c ) {
            double a = 0.0;
            for ( int k = 0; k < K; ++k ) a += Lm( o, k, i, c ) * beta[ k * ndef + S ] + alpha[ ( o * K + k ) * ndef + S ] * Fm( k, i, c );
            w[ c ] = a;
         }
         const double* gs = &gsrc[ ( (size_t)i * O + o ) * 10 ];
         const double* gc = &gcol[ ( (size_t)i * O + o ) * 3 ];
         for ( int d = 0; d < 10; ++d ) {
            double xdx = 0.0, wdx = 0.0;
            for ( int c = 0; c < 3; ++c ) xdx += R.xmu[ c ] * R.dxmu[ c ][ d ], wdx += w[ c ] * R.dxmu[ c ][ d ];
            const double defl = Pt[ o ] * R.dtr[ d ] * ddiag - g * P[ o * ndef + S ] * 2.0 * xdx / X[ S ] + g * wdx;
            if ( d == 9 ) {
               dafac[ o ] += ( R.gsa[ o ] - gs[ 9 ] ) * R.A - R.gself[ o ][ 9 ] - defl;
               continue;
            }
            const double dar = afac * R.dA[ d ];
            double v = ( R.gsa[ o ] - gs[ 9 ] ) * dar - gs[ d ] - R.gself[ o ][ d ] - defl;
            for ( int c = 0; c < 3; ++c ) v += ( R.gsx[ o ][ c ] + R.grhs[ o ][ c ] - gc[ c ] ) * R.dx[ c ][ d ];
            grad[ o * n9 + ( d / 3 ) * 3 * (size_t)n + ( d % 3 ) * (size_t)n + i ] = v;
         }
      }
   }
   for ( int o = 0; o < O; ++o ) {
      double* G = grad + o * n9;
      // afac = truearea / (sum of the geometric areas) when the areas are corrected
      if ( iarea == 1 )
         for ( int i = 0; i < n; ++i )
            for ( int d = 0; d < 9; ++d ) G[ ( d / 3 ) * 3 * (size_t)n + ( d % 3 ) * (size_t)n + i ] -= dafac[ o ] * afac / tot * tri[ i ].dA[ d ];
      // the vertices are shifted by the mean of the triangle centroids first
      for ( int c = 0; c < 3; ++c ) {
         double m = 0.0;
         for ( int v = 0; v < 3; ++v )
            for ( int i = 0; i < n; ++i ) m += G[ v * 3 * (size_t)n + c * (size_t)n + i ];
         m /= 3.0 * n;
         for ( int v = 0; v < 3; ++v )
            for ( int i = 0; i < n; ++i ) G[ v * 3 * (size_t)n + c * (size_t)n + i ] -= m;
      }
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
      BH_HIP( hipStreamCreateWithFlags( &s->st3, hipStreamNonBlocking ) );
      BH_HIP( hipEventCreateWithFlags( &s->ev_u, hipEventDisableTiming ) );
      BH_HIP( hipEventCreateWithFlags( &s->ev_g, hipEventDisableTiming ) );
      for ( auto& e : s->ev_lu ) BH_HIP( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
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
   if ( s->st3 ) (void)hipStreamSynchronize( s->st3 );
   hipStream_t st = s->st, st2 = s->st2, st3 = s->st3;
   hipEvent_t e1 = s->ev_u, e2 = s->ev_g;
   hipEvent_t el[ bhg::kLuEvents ];
   std::copy( s->ev_lu, s->ev_lu + bhg::kLuEvents, el );
   delete s;
   if ( e1 ) (void)hipEventDestroy( e1 );
   if ( e2 ) (void)hipEventDestroy( e2 );
   for ( hipEvent_t e : el )
      if ( e ) (void)hipEventDestroy( e );
   if ( st ) (void)hipStreamDestroy( st );
   if ( st2 ) (void)hipStreamDestroy( st2 );
   if ( st3 ) (void)hipStreamDestroy( st3 );
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
      s->icomp.ensure( n );
      int ncomp = 1, nopen = 0;
      if ( icomp ) {
         ncomp = *std::max_element( icomp, icomp + n );
         BH_HIP( hipMemcpyAsync( s->icomp.p, icomp, n * sizeof( int ), hipMemcpyHostToDevice, s->st ) );
      } else {
         // COMPON on the device, on the vertices after the centroid shift, where best/best.f:323
         // calls it (the shift is the prep's, bit for bit; each solve repeats it from raw_v)
         BH_HIP( hipMemcpyAsync( s->v.p, s->raw_v.p, 9 * (size_t)n * sizeof( double ), hipMemcpyDeviceToDevice, s->st ) );
         bhg::launch_prep( s->st, n, bhg::V1( s ), bhg::V2( s ), bhg::V3( s ), 0, 0.0, s->area.p, s->scal.p );
         bhg::device_compon( s->st, n, s->v.p, s->icomp.p, s->compw, &ncomp, &nopen );
      }
      // best/best.f:324-337: one deflation term per closed surface, one for all if any is open.
      s->ncomp = ncomp;
      s->nopen = nopen;
      s->ndef = 1;
      if ( nopen > 0 ) bhg::fill( s->st, s->comp.p, n, 1 );
      else BH_HIP( hipMemcpyAsync( s->comp.p, s->icomp.p, n * sizeof( int ), hipMemcpyDeviceToDevice, s->st ) );
      if ( nopen == 0 && ncomp > 1 ) s->ndef = ncomp;
      BH_HIP( hipStreamSynchronize( s->st ) );
      s->truearea = truearea;
      s->truevol = truevol;
      s->mesh_raw = true;
      s->mesh_ready = false;
      s->factored = 0;
      s->ir_hint = 0;
   } );
}

int bhg_set_mesh_file( bhg_solver* s, const char* path, int iarea ) {
   if ( !s ) return 1;
   bhg::BestMesh m;
   try {
      if ( !path ) throw std::runtime_error( "no path" );
      m = bhg::read_best( path, iarea );
   } catch ( const std::exception& e ) {
      s->last_error = std::string( "bhg_set_mesh_file: " ) + e.what();
      return 1;
   }
   return bhg_set_mesh( s, m.nar, m.v1.data(), m.v2.data(), m.v3.data(), m.truearea, m.truevol, nullptr );
}

int bhg_get_components( bhg_solver* s, int* nar, int* icomp, int* ncomp, int* nopen ) {
   return guarded( s, "bhg_get_components", [&] {
      if ( !s->mesh_raw ) throw std::runtime_error( "no mesh (bhg_set_mesh)" );
      bhg::set_device( s );
      if ( nar ) *nar = s->n;
      if ( icomp ) {
         BH_HIP( hipMemcpyAsync( icomp, s->icomp.p, s->n * sizeof( int ), hipMemcpyDeviceToHost, s->st ) );
         BH_HIP( hipStreamSynchronize( s->st ) );
      }
      if ( ncomp ) *ncomp = s->ncomp;
      if ( nopen ) *nopen = s->nopen;
   } );
}

int bhg_solve_mesh( bhg_solver* s, int iarea, double cg, int ivisco, int icov, int precision, bhg_result* out ) {
   return guarded( s, "bhg_solve_mesh", [&] { bhg::solve_mesh( s, iarea, cg, ivisco, icov, precision, out ); } );
}

int bhg_gradient( bhg_solver* s, int iarea, double cg, int precision, bhg_result* out, double* grad ) {
   return guarded( s, "bhg_gradient", [&] {
      if ( s->method != 0 ) throw std::runtime_error( "the gradient needs the dense path (bhg_set_method 0)" );
      if ( !grad ) throw std::runtime_error( "no gradient buffer" );
      bhg::solve_mesh( s, iarea, cg, 1, 0, precision, out );
      bhg::gradient( s, iarea, cg, grad );
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
   s->ir_hint = 0;
   return 0;
}

int bhg_set_option( bhg_solver* s, const char* name, double value ) {
   if ( !s || !name ) return 1;
   const std::string k( name );
   s->ir_hint = 0;
   if ( k == "fp32_growth" && value >= 1.0 ) {
      s->fp32_growth = value;
   } else if ( k == "ir_steps_before_pivoting" && value >= 0.0 ) {
      s->ir_steps_nopiv = (int)value;
   } else if ( k == "host_algebra" && ( value == 0.0 || value == 1.0 ) ) {
      s->host_algebra = (int)value;
   } else {
      s->last_error = "bhg_set_option: unknown option or value out of range: " + k;
      return 1;
   }
   return 0;
}

int bhg_set_method( bhg_solver* s, int method, double eps, int single, double tol ) {
   if ( !s || ( method != 0 && method != 1 ) ) return 1;
   s->method = method;
   s->ir_hint = 0;
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
   else if ( k == "trunc" ) o.trunc = value != 0.0 ? 1 : 0;
   else if ( k == "precond" ) o.precond = value != 0.0 ? 1 : 0;
   else if ( k == "eps_pre" && value > 0.0 ) o.eps_pre = value;
   else if ( k == "direct" ) o.direct = value != 0.0 ? 1 : 0;
   else if ( k == "pre_skip" ) o.pre_skip = std::max( 0, (int)value );
   else if ( k == "memcheck" ) o.memcheck = value != 0.0 ? 1 : 0;
   else if ( k == "mem_floor_gb" ) o.mem_floor = value * 1e9;
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
   out->n_trunc_failed = h.n_trunc_failed;
   out->kmean_aca = h.kmean_aca;
   out->t_trunc = h.t_trunc;
   out->eps_pre = h.eps_pre;
   out->piv_ratio = h.piv_ratio;
   out->pre_levels = h.pre_levels;
   out->pre_rank_max = h.pre_rank_max;
   out->pre_bytes = h.pre_bytes;
   out->mem_estimate = h.mem_estimate;
   out->mem_free = h.mem_free;
   out->mem_floor = h.mem_floor;
   out->mem_used = h.mem_used;
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

int bhg_hm_blocks( bhg_solver* s, long* out, long cap, int* perm, long* count ) {
   return guarded( s, "bhg_hm_blocks", [&] {
      if ( !s->hm ) throw std::runtime_error( "no H-matrix" );
      *count = bhg::hm_blocks( *s->hm, out, cap, perm );
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
      bhg::assemble( s, true, precision == 32 );
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

}  // extern "C"

namespace bhg {
namespace {

// The device address of `scratch` on the current device, looked up once per device.
char* noalloc_scratch( int device ) {
   static std::mutex m;
   static std::array<char*, 64> ptr {};
   if ( device < 0 || device >= (int)ptr.size() ) throw std::runtime_error( "device index out of range" );
   std::lock_guard<std::mutex> lk( m );
   if ( !ptr[ device ] ) {
      void* p = nullptr;
      BH_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) ) );
      if ( !p ) throw std::runtime_error( "hipGetSymbolAddress returned no address for the work space" );
      ptr[ device ] = static_cast<char*>( p );
   }
   return ptr[ device ];
}

}  // namespace
}  // namespace bhg

extern "C" {

int bhg_besolv_noalloc( bhg_solver* s, int nrhs, const double* d_b, long ldb, double* d_x, long ldx ) {
   return guarded( s, "bhg_besolv_noalloc", [&] {
      if ( s->factored != 32 && s->factored != 64 )
         throw std::runtime_error( s->factored == -1 ? "the H-matrix solve needs its own work space (use bhg_besolv)"
                                                     : "no factorization" );
      if ( nrhs < 1 || nrhs > bhg::kMaxRhs ) throw std::runtime_error( "between 1 and 16 right-hand sides per call" );
      if ( s->npad > bhg::kNoAllocMaxPad ) {
         char buf[ 160 ];
         std::snprintf( buf, sizeof buf, "matrix order %d (padded) exceeds the %ld the static work space holds (use bhg_besolv)",
                        s->npad, bhg::kNoAllocMaxPad );
         throw std::runtime_error( buf );
      }
      if ( ldb < s->n3 || ldx < s->n3 ) throw std::runtime_error( "leading dimensions below 3 n" );
      if ( !d_b || !d_x ) throw std::runtime_error( "null vector" );
      bhg::set_device( s );
      char* base = bhg::noalloc_scratch( s->device );
      const long np = s->npad;
      double* d = reinterpret_cast<double*>( base );
      bhg::SolveWork w {};
      w.b = d, w.x = d + bhg::kNaVec, w.r = d + 2 * bhg::kNaVec, w.y64 = d + 3 * bhg::kNaVec, w.x64il = d + 4 * bhg::kNaVec;
      w.p64 = d + 5 * bhg::kNaVec;
      w.part = d + ( 5 + bhg::kSkinnyParts ) * bhg::kNaVec;
      w.norms = w.part + bhg::kNaPart;
      float* f = reinterpret_cast<float*>( base + bhg::kNaDoubles * sizeof( double ) );
      w.d32 = f, w.y32 = f + bhg::kNaVec, w.x32 = f + 2 * bhg::kNaVec, w.p32 = f + 3 * bhg::kNaVec;
      if ( bhg::residual_part_elems( s->npad ) > (size_t)bhg::kNaPart ) throw std::runtime_error( "residual work space" );
      hipLaunchKernelGGL( bhg::k_noalloc_copy_in, dim3( 1024 ), dim3( 256 ), 0, s->st, s->n3, s->npad, nrhs, d_b, ldb, w.b );
      bhg::fill( s->st, w.norms, 64, 0.0 );
      struct Restore {
         bhg_solver* s;
         ~Restore() { s->work_override = nullptr; }
      } restore { s };
      s->work_override = &w;
      bhg::solve_device( s, nrhs );
      BH_HIP( hipMemcpy2DAsync( d_x, ldx * sizeof( double ), w.x, np * sizeof( double ), s->n3 * sizeof( double ), nrhs,
                                hipMemcpyDeviceToDevice, s->st ) );
      BH_HIP( hipStreamSynchronize( s->st ) );
   } );
}

// ---------------------------------------------------------------------- hmatrix/best_hmatrix.h

// ------------------------------------------------------------------ series of meshes

struct bhg_series {
   int device = 0;
   std::vector<bhg_solver*> m;
   std::vector<bhg::MeshTune> tune;
   hipStream_t so = nullptr;
   std::vector<hipEvent_t> ev;       // fork and join per mesh
   std::vector<hipEvent_t> ev_late;  // per mesh: its factorization's latency-bound end begins
   hipEvent_t t0 = nullptr, t1 = nullptr;
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   int parallel = 1;  // meshes side by side (1), one after another (0), or staggered (2): each mesh
                      // starts when the one before it has half of its rows left to factor
   int key_iarea = -1, key_ivisco = -1, key_icov = -1;
   double key_cg = -1.0;
   std::string last_error;
   void drop_graph() {
      if ( exec ) (void)hipGraphExecDestroy( exec );
      if ( graph ) (void)hipGraphDestroy( graph );
      exec = nullptr;
      graph = nullptr;
   }
   ~bhg_series() {
      drop_graph();
      for ( auto e : ev ) (void)hipEventDestroy( e );
      for ( auto e : ev_late ) (void)hipEventDestroy( e );
      if ( t0 ) (void)hipEventDestroy( t0 );
      if ( t1 ) (void)hipEventDestroy( t1 );
      if ( so ) (void)hipStreamDestroy( so );
      for ( auto p : m ) bhg_destroy( p );
   }
};

}  // extern "C"

namespace bhg {
namespace {

// Captures the solves of every mesh (enqueue_solve) as one HIP graph: forked from the series
// stream to each mesh's streams and joined back, side by side, chained one after another, or
// staggered: mesh k starts when mesh k - 1's factorization reaches its latency-bound end (half
// of its rows left), so that the next mesh's assembly and early, GEMM-bound factorization fill
// what the previous one's chain of small launches leaves idle.
void series_capture( bhg_series* S, int iarea, double cg, int ivisco, int icov, int parallel ) {
   BH_HIP( hipSetDevice( S->device ) );
   S->drop_graph();
   for ( auto* m : S->m ) prepare_buffers( m );
   BH_HIP( hipDeviceSynchronize() );
   const int nm = (int)S->m.size();
   BH_HIP( hipStreamBeginCapture( S->so, hipStreamCaptureModeRelaxed ) );
   try {
      hipStream_t prev = S->so;
      for ( int k = 0; k < nm; ++k ) {
         bhg_solver* m = S->m[ k ];
         if ( parallel == 2 && k > 0 ) {
            BH_HIP( hipStreamWaitEvent( m->st, S->ev_late[ k - 1 ], 0 ) );
         } else {
            BH_HIP( hipEventRecord( S->ev[ 2 * k ], parallel ? S->so : prev ) );
            BH_HIP( hipStreamWaitEvent( m->st, S->ev[ 2 * k ], 0 ) );
         }
         enqueue_solve( m, iarea, cg, ivisco, icov, S->tune[ k ], parallel == 2 ? S->ev_late[ k ] : nullptr );
         BH_HIP( hipEventRecord( S->ev[ 2 * k + 1 ], m->st ) );
         prev = m->st;
      }
      if ( parallel )
         for ( int k = 0; k < nm; ++k ) BH_HIP( hipStreamWaitEvent( S->so, S->ev[ 2 * k + 1 ], 0 ) );
      else
         BH_HIP( hipStreamWaitEvent( S->so, S->ev[ 2 * nm - 1 ], 0 ) );
   } catch ( ... ) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture( S->so, &g );
      if ( g ) (void)hipGraphDestroy( g );
      throw;
   }
   BH_HIP( hipStreamEndCapture( S->so, &S->graph ) );
   BH_HIP( hipGraphInstantiate( &S->exec, S->graph, nullptr, nullptr, 0 ) );
   S->parallel = parallel;
   S->key_iarea = iarea;
   S->key_cg = cg;
   S->key_ivisco = ivisco;
   S->key_icov = icov;
}

// One launch of the captured schedule; the device time in ms.
double series_replay( bhg_series* S ) {
   BH_HIP( hipEventRecord( S->t0, S->so ) );
   BH_HIP( hipGraphLaunch( S->exec, S->so ) );
   BH_HIP( hipEventRecord( S->t1, S->so ) );
   BH_HIP( hipEventSynchronize( S->t1 ) );
   float ms = 0.f;
   BH_HIP( hipEventElapsedTime( &ms, S->t0, S->t1 ) );
   return ms;
}

template <class F>
int series_guarded( bhg_series* S, const char* what, F&& f ) {
   if ( !S ) return 1;
   try {
      f();
      return 0;
   } catch ( const std::exception& e ) {
      S->last_error = std::string( what ) + ": " + e.what();
      return 1;
   }
}

}  // namespace
}  // namespace bhg

extern "C" {

bhg_series* bhg_series_create( int device, int nmesh ) {
   if ( nmesh < 1 ) return nullptr;
   auto S = std::make_unique<bhg_series>();
   S->device = device;
   try {
      BH_HIP( hipSetDevice( device ) );
      BH_HIP( hipStreamCreateWithFlags( &S->so, hipStreamNonBlocking ) );
      BH_HIP( hipEventCreate( &S->t0 ) );
      BH_HIP( hipEventCreate( &S->t1 ) );
      S->ev.resize( 2 * nmesh, nullptr );
      for ( auto& e : S->ev ) BH_HIP( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
      S->ev_late.resize( nmesh, nullptr );
      for ( auto& e : S->ev_late ) BH_HIP( hipEventCreateWithFlags( &e, hipEventDisableTiming ) );
      for ( int k = 0; k < nmesh; ++k ) {
         bhg_solver* m = bhg_create( device );
         if ( !m ) throw std::runtime_error( "bhg_create" );
         S->m.push_back( m );
      }
      S->tune.assign( nmesh, bhg::MeshTune {} );
   } catch ( ... ) {
      return nullptr;
   }
   return S.release();
}

void bhg_series_destroy( bhg_series* S ) { delete S; }

const char* bhg_series_last_error( bhg_series* S ) { return S ? S->last_error.c_str() : "no series"; }

int bhg_series_set_mesh_file( bhg_series* S, int k, const char* path, int iarea ) {
   return bhg::series_guarded( S, "bhg_series_set_mesh_file", [&] {
      if ( k < 0 || k >= (int)S->m.size() ) throw std::runtime_error( "mesh index out of range" );
      S->drop_graph();
      if ( bhg_set_mesh_file( S->m[ k ], path, iarea ) != 0 ) throw std::runtime_error( S->m[ k ]->last_error );
   } );
}

int bhg_series_set_mesh( bhg_series* S, int k, int n, const double* v1, const double* v2, const double* v3, double truearea,
                         double truevol ) {
   return bhg::series_guarded( S, "bhg_series_set_mesh", [&] {
      if ( k < 0 || k >= (int)S->m.size() ) throw std::runtime_error( "mesh index out of range" );
      S->drop_graph();
      if ( bhg_set_mesh( S->m[ k ], n, v1, v2, v3, truearea, truevol, nullptr ) != 0 )
         throw std::runtime_error( S->m[ k ]->last_error );
   } );
}

int bhg_series_set_tuning( bhg_series* S, int k, int strip, int gemm_bk, int ir_iters ) {
   return bhg::series_guarded( S, "bhg_series_set_tuning", [&] {
      if ( k < 0 || k >= (int)S->m.size() ) throw std::runtime_error( "mesh index out of range" );
      if ( strip < 1 || ( gemm_bk != 0 && gemm_bk != 16 && gemm_bk != 32 && gemm_bk != 64 ) || ir_iters < 1 )
         throw std::runtime_error( "strip >= 1, gemm_bk 0, 16, 32 or 64, ir_iters >= 1" );
      // the eager continuation after the replay needs steps left below the refinement limit
      if ( ir_iters >= S->m[ k ]->ir_maxit ) throw std::runtime_error( "ir_iters must be below the refinement step limit" );
      S->tune[ k ] = bhg::MeshTune { strip, gemm_bk, ir_iters };
      S->drop_graph();
   } );
}

int bhg_series_set_option( bhg_series* S, const char* name, double value ) {
   return bhg::series_guarded( S, "bhg_series_set_option", [&] {
      for ( auto* m : S->m )
         if ( bhg_set_option( m, name, value ) != 0 ) throw std::runtime_error( m->last_error );
      S->drop_graph();
   } );
}

int bhg_series_get_tuning( bhg_series* S, int k, int* out ) {
   if ( !S || k < 0 || k >= (int)S->m.size() ) return 1;
   out[ 0 ] = S->tune[ k ].strip;
   out[ 1 ] = S->tune[ k ].gemm_bk;
   out[ 2 ] = S->tune[ k ].ir_iters;
   out[ 3 ] = S->parallel;
   return 0;
}

int bhg_series_autotune( bhg_series* S, int iarea, double cg, int ivisco ) {
   return bhg::series_guarded( S, "bhg_series_autotune", [&] {
      BH_HIP( hipSetDevice( S->device ) );
      hipEvent_t e0, e1;
      BH_HIP( hipEventCreate( &e0 ) );
      BH_HIP( hipEventCreate( &e1 ) );
      struct Ev {
         hipEvent_t a, b;
         ~Ev() {
            (void)hipEventDestroy( a );
            (void)hipEventDestroy( b );
         }
      } guard { e0, e1 };
      for ( size_t k = 0; k < S->m.size(); ++k ) {
         bhg_solver* m = S->m[ k ];
         bhg::MeshTune t;
         // an eager solve: allocations, and the refinement steps this mesh needs
         if ( bhg_solve_mesh( m, iarea, cg, ivisco, 0, 32, nullptr ) != 0 ) throw std::runtime_error( m->last_error );
         t.ir_iters = std::min( std::max( 2, m->stats.ir_iterations ), m->ir_maxit - 1 );
         bhg::prepare_buffers( m );
         const long np = m->npad;
         auto time = [&]( auto&& f ) {
            double best = 1e300;
            for ( int rep = 0; rep < 2; ++rep ) {
               BH_HIP( hipEventRecord( e0, m->st ) );
               f();
               BH_HIP( hipEventRecord( e1, m->st ) );
               BH_HIP( hipEventSynchronize( e1 ) );
               float ms = 0.f;
               BH_HIP( hipEventElapsedTime( &ms, e0, e1 ) );
               best = std::min( best, (double)ms );
            }
            return best;
         };
         double tb = 1e300;
         for ( int strip : { 16, 32, 64, 128 } ) {
            const double ms = time( [&] {
               bhg::launch_pairs( m->st, m->n, strip, m->cg, m->src.p, m->xc.p, m->selfb.p, m->xmu.p, m->comp.p, m->gdef.p, 1,
                                  m->a64.p, np, m->a32.p );
            } );
            if ( ms < tb ) tb = ms, t.strip = strip;
         }
         tb = 1e300;
         // one K slice for every tile shape: the per-shape default (0) times fastest for a mesh
         // alone but slowed the side-by-side schedule of the series (352 -> 370 ms)
         for ( int bk : { 16, 32, 64 } ) {
            const double ms = time( [&] {
               bhg::g_gemm_bk_f32 = bk;
               bhg::fill( m->st, m->flags.p, 4, 0 );
               bhg::convert_d2f( m->st, np * np, m->a64.p, m->a32.p );
               bhg::g_growth = m->fp32_growth;
               bhg::lu_nopiv<float>( m->lu_streams(), m->npad, m->a32.p, np, m->inv32.p, m->flags.p );
               bhg::g_growth = 1.0;
               bhg::build_super_inverses<float>( m->st, m->npad, bhg::kSuper, m->a32.p, np, m->inv32.p, m->xl32.p, m->xu32.p,
                                                 m->tmp32.p );
               bhg::g_gemm_bk_f32 = 0;
            } );
            if ( ms < tb ) tb = ms, t.gemm_bk = bk;
         }
         S->tune[ k ] = t;
      }
      // The schedule, timed as a whole: side by side or one after another, each with the K slices
      // chosen per mesh or with 32 for every mesh (a slice that is fastest for a mesh alone can be
      // slower when the meshes share the GPU).
      const std::vector<bhg::MeshTune> tuned = S->tune;
      std::vector<bhg::MeshTune> k32 = tuned;
      for ( auto& t : k32 ) t.gemm_bk = 32;
      double best = 1e300;
      int choice = 1, choice_k = 0;
      for ( int kset : { 0, 1 } ) {
         S->tune = kset ? k32 : tuned;
         for ( int par : { 1, 0, 2 } ) {
            bhg::series_capture( S, iarea, cg, ivisco, 0, par );
            bhg::series_replay( S );
            const double ms = bhg::series_replay( S );
            if ( ms < best ) best = ms, choice = par, choice_k = kset;
         }
      }
      S->tune = choice_k ? k32 : tuned;
      bhg::series_capture( S, iarea, cg, ivisco, 0, choice );
   } );
}

int bhg_series_solve( bhg_series* S, int iarea, double cg, int ivisco, int icov, bhg_result* out, double* graph_ms ) {
   return bhg::series_guarded( S, "bhg_series_solve", [&] {
      for ( auto* m : S->m )
         if ( !m->mesh_raw ) throw std::runtime_error( "a mesh of the series is not set" );
      if ( !S->exec || S->key_iarea != iarea || S->key_cg != cg || S->key_ivisco != ivisco || S->key_icov != icov )
         bhg::series_capture( S, iarea, cg, ivisco, icov, S->parallel );
      const double ms = bhg::series_replay( S );
      if ( graph_ms ) *graph_ms = ms;
      for ( size_t k = 0; k < S->m.size(); ++k )
         bhg::finish_enqueued( S->m[ k ], iarea, cg, ivisco, icov, S->tune[ k ], out ? out + k : nullptr );
   } );
}

int bhg_series_get_stats( bhg_series* S, int k, bhg_stats* out ) {
   if ( !S || k < 0 || k >= (int)S->m.size() ) return 1;
   return bhg_get_stats( S->m[ k ], out );
}

int bhg_extrapolate_1n( int nmesh, const int* n, const double* y, double* intercept, double* slope ) {
   // least squares y = a + b / N
   if ( nmesh < 2 ) return 1;
   double sx = 0, sy = 0, sxx = 0, sxy = 0;
   for ( int k = 0; k < nmesh; ++k ) {
      const double x = 1.0 / n[ k ];
      sx += x;
      sy += y[ k ];
      sxx += x * x;
      sxy += x * y[ k ];
   }
   const double d = nmesh * sxx - sx * sx;
   if ( d == 0.0 ) return 1;
   *slope = ( nmesh * sxy - sx * sy ) / d;
   *intercept = ( sy - *slope * sx ) / nmesh;
   return 0;
}

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
         // BEST_HMATRIX_PRE_EPS overrides the preconditioner's accuracy (hmatrix/best_hmatrix.h:48-60)
         const char* pre = std::getenv( "BEST_HMATRIX_PRE_EPS" );
         s->hm_opt.eps_pre = pre ? std::atof( pre ) : ( eps_pre > 0.0 ? eps_pre : 0.1 );
         s->hm_opt.direct = direct ? 1 : 0;
         // BEST_HIP_PRECOND=hierarchical: GMRES preconditioned by the hierarchical factorization at
         // eps_pre from the start (the reference's choice, hmatrix/best_hmatrix.h:48-60); by default
         // block Jacobi, which is faster here, and the factorization only if GMRES stalls
         const char* pc = std::getenv( "BEST_HIP_PRECOND" );
         if ( pc ) s->hm_opt.precond = std::string( pc ) == "hierarchical" ? 1 : 0;
         // the memory contract of hmatrix/best_hmatrix.h:48-60, for device memory
         const char* mc = std::getenv( "BEST_HMATRIX_MEMCHECK" );
         s->hm_opt.memcheck = mc ? ( std::atof( mc ) != 0.0 ) : 1;
         const char* mf = std::getenv( "BEST_HMATRIX_MEMFLOOR" );
         s->hm_opt.mem_floor = mf ? 1e9 * std::atof( mf ) : -1.0;
         s->hm_opt.nrhs = std::min( std::max( nrhs, 1 ), bhg::kMaxRhs );
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
