// The hierarchical-matrix path of the solver: the counterpart of hmatrix/src (cluster tree, block
// tree, low-rank blocks by adaptive cross approximation, the H-matrix product, GMRES) with the
// blocks, the products and the Krylov vectors in device memory.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "bh_common.h"

namespace bhg {

struct HmOptions {
   double eps = 1e-6;          // relative accuracy of the low-rank blocks
   double eta = 2.0;           // admissibility: min(diam t, diam s) <= eta dist(t, s)
   double aca_factor = 0.1;    // the cross approximation runs to aca_factor * eps
   int leaf = 32;              // triangles per leaf cluster at most
   long dense_small = 1024;    // admissible blocks of at most this many triangle pairs are kept dense
   int kcap = 192;             // rank at which a cross approximation gives up (the block is then kept dense)
   int single = 0;             // 1: blocks stored in FP32 (arithmetic in FP64)
   int trunc = 1;              // 1: approximations recompressed to eps by QR and SVD, as the reference does
   double tol = 0.0;           // GMRES relative residual; 0: aca_factor * eps
   int restart = 120;          // GMRES restart length (the reference's 30 with its H-LU; 120 with block Jacobi)
   int maxsteps = 500;         // GMRES steps (batched products) per solve
   double ws_bytes = 24e9;     // device work space of one batch of cross approximations
   int precond = 0;            // GMRES preconditioner: 0 block Jacobi (then the hierarchical factorization at
                               //   eps_pre if GMRES stalls), 1 the hierarchical factorization at eps_pre
   double eps_pre = 0.1;       // its accuracy (BEST's EPSPRE, best/besolve.f90:34)
   int pre_skip = 0;           // coarsest levels of the hierarchical factorization left out (diagnostic)
   int nrhs = 16;              // right-hand sides per solve at most (for the memory estimate)
   int memcheck = 1;           // refuse a build the estimate says does not fit above the floor
   double mem_floor = -1.0;    // device bytes to keep free (< 0: 5% of the device's memory, at least 1 GB)
   int direct = 0;             // 1 (-hlu): the hierarchical factorization at eps solves directly, refined
                               //   against the H-matrix, no GMRES
};

struct HmStats {
   int depth = 0, nleaves = 0, leaf_min = 0, leaf_max = 0;
   long n_dense = 0;           // inadmissible leaf blocks (leaf cluster pairs)
   long n_lowrank = 0;         // admissible blocks stored as U V^T
   long n_small = 0;           // admissible blocks small enough to be kept dense
   long n_capped = 0;          // admissible blocks whose approximation reached kcap (kept dense)
   double bytes_dense = 0, bytes_lowrank = 0;
   int kmax = 0;
   double kmean = 0;
   double kernel_blocks = 0;   // 3x3 kernel blocks evaluated by the build
   double t_tree = 0, t_fill = 0, t_pre = 0;   // seconds
   int iters_min = 0, iters_max = 0, steps = 0;
   double relres = 0, t_solve = 0;
   long n_trunc_failed = 0;    // approximations left as they were (dependent columns, or rank above the limit)
   double kmean_aca = 0;       // mean rank before the recompression
   double t_trunc = 0;         // seconds in the recompression (part of t_fill)
   double eps_pre = 0;         // accuracy of the hierarchical preconditioner (0: block Jacobi)
   double piv_ratio = 0;       // its smallest pivot ratio (cluster systems)
   int pre_levels = 0, pre_rank_max = 0;
   double pre_bytes = 0;
   double mem_estimate = 0, mem_free = 0, mem_floor = 0;  // bytes: the build's estimate, free before it, the floor
   double mem_used = 0;        // bytes the build took (free memory before it less free memory after it)
};

// The geometry: vertices (v1, v2, v3 as BEST's NAR x 3 column-major arrays, one after the other)
// and incenters (NAR x 3, column-major) on the host; the per-triangle records, incenters, self
// blocks and the deflation (XMU, the surface of each triangle, GDEF per surface) on the device.
struct HmInput {
   int n = 0;
   double cg = 1.0;
   const double* h_v = nullptr;
   const double* h_xc = nullptr;
   const double* d_src = nullptr;
   const double* d_xc = nullptr;
   const double* d_selfb = nullptr;
   const double* d_xmu = nullptr;
   const int* d_comp = nullptr;
   const double* d_gdef = nullptr;
   int ndef = 1;
};

// Cluster and block trees alone (host): the counts the build would log.
HmStats hm_plan( int n, const double* h_v, const double* h_xc, const HmOptions& opt );

struct HMatrix;
struct HMatrixDeleter {
   void operator()( HMatrix* h ) const;
};
using HMatrixPtr = std::unique_ptr<HMatrix, HMatrixDeleter>;

// Builds the H-matrix of GR + sum_k gdef[k] xmu_k xmu_k^T and the GMRES preconditioner, on st
// with two more streams beside it (the largest blocks' approximations, the recompression of large
// ranks); the three should sit on different hardware queues.
HMatrixPtr hm_build( hipStream_t st, hipStream_t st_b, hipStream_t st_c, const HmInput& in, const HmOptions& opt, HmStats* stats,
                     std::vector<std::string>* log );

// Solves for nrhs (<= 16) right-hand sides: b and x are device arrays in BEST's unknown order
// (3 n x nrhs, column-major, leading dimensions ldb and ldx).  Throws if GMRES does not converge.
void hm_solve( HMatrix& h, hipStream_t st, int nrhs, const double* d_b, long ldb, double* d_x, long ldx, HmStats* stats,
               std::vector<std::string>* log );

// Mean duration (ms, HIP events) of the product with p columns of device-resident vectors over
// reps repetitions; *bytes = the blocks' storage plus X and Y (the algorithmic HBM bytes).
double hm_time_product( HMatrix& h, hipStream_t st, int p, int reps, double* bytes );

// The leaf blocks, 6 longs each (1 if low-rank else 0, rank, r0, nr, c0, nc; unknowns in tree
// order), at most cap of them into out; perm (n ints) = the triangle at each tree position.
// Returns the number of blocks.
long hm_blocks( const HMatrix& h, long* out, long cap, int* perm );

// y = (GR + deflation) x on the device, BEST's order (for checks).
void hm_apply( HMatrix& h, hipStream_t st, int nrhs, const double* d_x, long ldx, double* d_y, long ldy );

}  // namespace bhg
