/* best_hip.h -- BEST's boundary-element Stokes solve on one MI355X GPU.
 *
 * Part 1 is the C ABI of the reference's H-matrix library, hmatrix/best_hmatrix.h:26-73, with
 * the same nine names, arguments, layouts and error contract, so that BEST links this library
 * where it linked $(HM_OBJ) (Makefile:54-62) or best/hmstub.f90, and best/besolve.f90:48-89
 * binds it unchanged.  Behind it the whole dense path runs on the GPU: the GR entries
 * (best/tgetgr.f, best/triangleint.f), the deflation, an LU on the matrix cores and the solves.
 *
 * Part 2 is the handle-based API the Python layer, the tests and the benchmark use: one solver
 * per handle, a mesh and its factors resident in device memory, and the whole mesh solve of
 * best/best.f:261-498 (without file I/O) in one call.
 *
 * Layouts follow BEST: vertex and incenter arrays are NAR x 3 column-major (x[c*n + i]);
 * matrices and right-hand-side blocks are column-major; unknown 3*i + a is component a on
 * triangle i (0-based).  All pointers are host pointers.
 */
#ifndef BEST_HIP_H
#define BEST_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ Part 1: hmatrix ABI */

/* hmatrix/best_hmatrix.h:27 (stand-in best/hmstub.f90:7-11).  Returns 1. */
int bhm_available( void );

/* hmatrix/best_hmatrix.h:35-36.  The mesh as HMPREP passes it (best/besolve.f90:164): vertices
 * after the centroid shift, areas after the area correction, afac, cg.  Computes the incenters,
 * self blocks and the mean diagonal on the GPU (device BEST_HIP_DEVICE, default 0).  threads is
 * ignored.  Returns 0, or nonzero after queuing a message. */
int bhm_create( int n, const double* v1, const double* v2, const double* v3, const double* area, double afac,
                double cg, int threads );

/* hmatrix/best_hmatrix.h:39: TGETGR's X, n x 3 column-major. */
void bhm_incenters( double* x );

/* hmatrix/best_hmatrix.h:43: GR(i0 : i0+m-1, j0 : j0+nc-1) before deflation, column-major. */
void bhm_entries( int i0, int m, int j0, int nc, double* a, int lda );

/* hmatrix/best_hmatrix.h:46: the mean diagonal entry of GR in BEST's summation order. */
double bhm_diag_mean( void );

/* hmatrix/best_hmatrix.h:62-63.  Up to 30,000 unknowns (or with BEST_HIP_SOLVER=dense): assembles
 * GR + sum_k gdef[k] xmu_k xmu_k^T densely on the GPU and factors it: single = 1 (or BEST_HIP_LU
 * unset) FP32 LU on MFMA with FP64 iterative refinement in bhm_solve; BEST_HIP_LU=fp64 and single
 * = 0: FP64 LU; eps, direct and eps_pre do not apply (the dense solve is exact to rounding).  Above
 * that (or with BEST_HIP_SOLVER=hmatrix): the H-matrix at eps (FP32 storage with single = 1), and
 * with direct = 0 GMRES preconditioned by block Jacobi, then by the hierarchical factorization at
 * eps_pre if GMRES stalls (BEST_HIP_PRECOND=hierarchical: from the start); with direct = 1 that
 * factorization at eps solves directly.  Returns nonzero (with a message) when the problem does not
 * fit in device memory. */
int bhm_build( double eps, int direct, double eps_pre, int ndef, const double* gdef, const int* comp, const double* xmu,
               int single, int nrhs );

/* hmatrix/best_hmatrix.h:67: b (3n x nrhs, column-major) in, solutions out. */
int bhm_solve( int nrhs, double* b );

/* hmatrix/best_hmatrix.h:71: the next queued message (no NUL), or -1. */
int bhm_log_line( char* buf, int len );

/* hmatrix/best_hmatrix.h:74: frees everything. */
void bhm_destroy( void );

/* ------------------------------------------------------------------ Part 2: solver handles */

typedef struct bhg_solver bhg_solver;

/* The values BEST writes to its .be file (best/output.f) and the tensors behind them; 3x3
 * tensors column-major as in best/difco.f. */
typedef struct bhg_result {
   double dtt[ 9 ], dtr[ 9 ], drr[ 9 ]; /* diffusion tensors at the centre of diffusion (1/A, 1/A^2, 1/A^3) */
   double urc[ 3 ], udr[ 3 ];           /* centre of resistance - centroid, centre of diffusion - resistance */
   double ktt[ 9 ], krt[ 9 ], krr[ 9 ]; /* friction tensors at the centre of resistance */
   double drr_eig[ 3 ], dtt_eig[ 3 ];   /* JACOBI eigenvalues */
   double dr;                           /* 1/3 Trace DRR(EXACT), 1/s */
   double dtrans;                       /* 1/3 Trace DTT(EXACT), cm^2/s */
   double chiprime;                     /* CHI * VOLUME, A^3 (with ivisco) */
   double chi;                          /* CHIPRIME / TRUEVOL (with the area correction) */
   double dvisc[ 3 ];                   /* centre of viscosity - centroid (icov), else 0 */
   double afac, totarea, centroid[ 3 ], diagm;
   int ncomp, nopen, ndef;
} bhg_result;

/* Timings (ms, HIP events on the solver's stream) and diagnostics of the last call. */
typedef struct bhg_stats {
   double t_prep, t_assemble, t_factor, t_solve, t_reduce, t_total;
   int ir_iterations;    /* refinement steps of the FP32 path (0 with FP64 LU) */
   double ir_correction; /* last max_j |dx_j| / |x_j| */
   int pivoted;          /* 1 if the factorization needed row exchanges */
   int npad;             /* padded order of the matrix */
   double t_adjoint_solve, t_adjoint_pairs, t_adjoint_tri; /* bhg_gradient: the transposed solves
                                                              (3 x 11 RHS), the pair term, the rest */
   int ir_above_tol;     /* 1 if the refinement stagnated below 1e-10 but above its tolerance and the
                            solution was accepted there (the log says so) */
} bhg_stats;

bhg_solver* bhg_create( int device );
void bhg_destroy( bhg_solver* s );
const char* bhg_last_error( bhg_solver* s );
int bhg_log_line( bhg_solver* s, char* buf, int len );

/* COMPON (best/compon.f90:19-115) on the host: icomp[n] in 1..ncomp, and the open-edge count. */
int bhg_compon( int n, const double* v1, const double* v2, const double* v3, int* icomp, int* ncomp, int* nopen );

/* A mesh as BEST reads it (best/best.f:189-220, before the centroid shift), made resident.  icomp
 * may be NULL: COMPON then runs on the device, on the copied vertices after the centroid shift
 * (where best/best.f:323 calls it). */
int bhg_set_mesh( bhg_solver* s, int n, const double* v1, const double* v2, const double* v3, double truearea,
                  double truevol, const int* icomp );
/* The mesh file BEST reads (best/best.f:189-220: OPEN(10), READ(10,*) NAR[, TRUEAREA, TRUEVOL]
 * with iarea 1, then one READ of 9 values per triangle), read on the host and made resident as
 * bhg_set_mesh does, COMPON on the device. */
int bhg_set_mesh_file( bhg_solver* s, const char* path, int iarea );
/* The resident mesh's NAR, and COMPON's result for it: icomp[NAR] (may be NULL), NCOMP, NOPEN. */
int bhg_get_components( bhg_solver* s, int* nar, int* icomp, int* ncomp, int* nopen );

/* One mesh solve on the resident mesh: best/best.f:261-498 without file I/O (prep, TGETGR,
 * deflation, DGETRF, DIFCO, VISCO, the .be values).  iarea = 1 applies the area correction
 * (BEST's default, -na gives 0); cg = 1, or 0 with -p; ivisco = 1 solves the strain flows (-v);
 * icov = 1 moves to the centre of viscosity (-vc); precision = 64 (FP64 LU) or 32 (FP32 LU on
 * MFMA + FP64 refinement). */
int bhg_solve_mesh( bhg_solver* s, int iarea, double cg, int ivisco, int icov, int precision, bhg_result* out );

/* The adjoint of bhg_solve_mesh (ivisco = 1, icov = 0, dense path): one solve, then the gradient
 * of D_t (out->dtrans), D_r (out->dr) and CHI * V (out->chiprime; [eta] = 0.602214 CHI V / M)
 * with respect to the vertex coordinates as bhg_set_mesh took them.  grad holds 3 x 9 n doubles:
 * grad[o * 9 n + v * 3 n + c * n + i] = d(output o) / d(coordinate c of vertex v of triangle i),
 * o = 0 D_t, 1 D_r, 2 CHI V; the layout of v1, v2, v3 one after the other.  The adjoint system
 * GR^T lambda = dQ/dF is solved with the factors of the forward solve, transposed in place.  The
 * surfaces (icomp) are held fixed.  No counterpart in the reference, which has no derivatives:
 * it differentiates the path of best/best.f:261-498 (assembly best/tgetgr.f, best/triangleint.f,
 * deflation best/best.f:384-426, reductions best/difco.f, best/visco.f). */
int bhg_gradient( bhg_solver* s, int iarea, double cg, int precision, bhg_result* out, double* grad );

/* The last solutions (3n x nrhs, columns: 3 translations, 3 rotations, 5 strain flows). */
int bhg_get_solution( bhg_solver* s, int nrhs, double* out, long ldo );
int bhg_get_stats( bhg_solver* s, bhg_stats* out );
int bhg_set_refinement( bhg_solver* s, double tol, int maxit );
/* Settings of the dense path that BEST does not have (it calls LAPACK's partial pivoting):
 * "fp32_growth" (256): the largest multiplier the exchange-free FP32 LU accepts before it
 * refactors with row exchanges (the FP64 LU keeps partial pivoting's own test, 1);
 * "ir_steps_before_pivoting" (0 = the refinement step limit): refinement steps allowed with the
 * exchange-free FP32 factors before the matrix is refactored with row exchanges.  Both exist so
 * that the fallbacks can be exercised.  Returns 1 for an unknown name or a value out of range. */
int bhg_set_option( bhg_solver* s, const char* name, double value );

/* Launch profiling of bhg_solve_mesh: with on = 1 every launch of the timed kernel classes is
 * bracketed by HIP events on the solver's stream.  bhg_get_profile returns, per class (0 pair
 * assembly, 1 FP64 GEMM, 2 FP32 GEMM on 64 x 128 and 128 x 64 tiles, 3 diagonal blocks,
 * 4 triangular solves, 5 refinement residuals, 6 other, 7 FP32 GEMM on 128 x 128 tiles, 8 FP32
 * GEMM on 64 x 64 tiles: 9 entries), the summed duration (ms), algorithmic flops and launch count
 * since profiling was turned on. */
int bhg_set_profiling( bhg_solver* s, int on );
int bhg_get_profile( bhg_solver* s, double* ms, double* flops, long* count );

/* The prepared mesh of bhm_create (shifted vertices, corrected areas). */
int bhg_load_prepared( bhg_solver* s, int n, const double* v1, const double* v2, const double* v3, const double* area,
                       double afac, double cg );
int bhg_incenters( bhg_solver* s, double* x );
int bhg_entries( bhg_solver* s, int i0, int m, int j0, int nc, double* a, long lda );
double bhg_diag_mean( bhg_solver* s );

/* TGETGR twin (best/tgetgr.f:1): X and GR (no deflation) of the prepared mesh into host memory. */
int bhg_tgetgr( bhg_solver* s, double* x, double* gr, long ldgr );

/* best/best.f:384-468 on the GPU: GR + deflation assembled and factored (DGETRF twin). gdef NULL:
 * the weights are computed on the device in BEST's order. */
int bhg_factor_deflated( bhg_solver* s, int ndef, const double* gdef, const int* comp, const double* xmu, int precision );

/* BESOLV twin (best/besolve.f90:95-134): b (3n x nrhs, leading dimension ldb) in place. */
int bhg_besolv( bhg_solver* s, int nrhs, double* b, long ldb );

/* bhg_besolv for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  b and x are
 * DEVICE pointers (3n x nrhs, leading dimensions ldb and ldx); the solve's work vectors live in
 * one static device array of the library (101 MiB on each GPU it runs on, held until the process
 * ends), so nothing is allocated or freed: the resident factors of bhg_solve_mesh or
 * bhg_factor_deflated (FP64 LU, or FP32 LU + FP64 refinement) are used as they are.  Limits: padded
 * matrix order up to 16,384 (5,461 triangles) and up to 16 right-hand sides, else nonzero with a
 * message; not the H-matrix path (its Krylov basis needs its own memory).  Calls that overlap on
 * one GPU share that array and race; the solver's own solution (bhg_get_solution) is not updated.
 * Synchronizes its stream before returning. */
int bhg_besolv_noalloc( bhg_solver* s, int nrhs, const double* b, long ldb, double* x, long ldx );

/* ------------------------------------------------------------------ the H-matrix path */

/* Diagnostics of the last H-matrix build and solve (the figures hmatrix/src/api.cpp:380-386 and
 * :722-735 log). */
typedef struct bhg_hm_stats {
   int depth, nleaves, leaf_min, leaf_max; /* cluster tree */
   long n_dense, n_lowrank, n_small, n_capped; /* leaf blocks: inadmissible, U V^T, small admissible
                                                 (dense), approximations that reached the rank cap */
   double bytes_dense, bytes_lowrank;
   int kmax;
   double kmean, kernel_blocks; /* ranks; 3x3 kernel blocks evaluated by the build */
   double t_tree, t_fill, t_pre; /* seconds */
   int iters_min, iters_max, steps; /* GMRES iterations per right-hand side, batched products */
   double relres, t_solve;
   long n_trunc_failed; /* approximations not recompressed (dependent columns or rank above 128) */
   double kmean_aca, t_trunc; /* mean rank before the recompression; its time (s, part of t_fill) */
   double eps_pre;            /* accuracy of the hierarchical preconditioner (0: block Jacobi only) */
   double piv_ratio;          /* its smallest pivot ratio (the cluster systems) */
   int pre_levels, pre_rank_max; /* its levels and their largest basis width */
   double pre_bytes;          /* device bytes of its bases */
   double mem_estimate, mem_free, mem_floor, mem_used; /* device bytes: the estimate made before the build
                                                          (hmatrix/src/api.cpp:146-339), free then, the
                                                          floor kept free, what the build took */
} bhg_hm_stats;

/* method 0: dense LU (default); 1: H-matrix + GMRES (hmatrix/src/api.cpp:585-735) for the next
 * bhg_solve_mesh: low-rank blocks by cross approximation at eps (BEST's -he), stored in FP32 with
 * single = 1 (-hs), GMRES to the relative residual tol (0: eps / 10, as BEST_HMATRIX_GMRES_TOL's
 * default, hmatrix/src/api.cpp:840).  bhm_build takes this path above 30,000 unknowns, or always
 * with BEST_HIP_SOLVER=hmatrix. */
int bhg_set_method( bhg_solver* s, int method, double eps, int single, double tol );
int bhg_get_hm_stats( bhg_solver* s, bhg_hm_stats* out );

/* The H-matrix settings that are not BEST options (hmatrix/src/api.cpp:37-41, hmatrix.h:60-66):
 * "eta" (admissibility, 2), "aca_factor" (0.1), "leaf" (triangles per leaf cluster, 32),
 * "dense_small" (admissible blocks of at most this many triangle pairs stay dense, 1024), "kcap"
 * (rank at which an approximation is abandoned for a dense block, 192), "precond" (0: block
 * Jacobi, then the hierarchical factorization at eps_pre if GMRES stalls; 1: the factorization from
 * the start, hmatrix/src/api.cpp:502-578), "eps_pre" (its accuracy, 0.1), "direct" (1: -hlu, the
 * factorization at eps refined against the H-matrix, hmatrix/src/api.cpp:461-495), "restart" (GMRES restart
 * length, 120: the reference restarts at 30, hmatrix/src/api.cpp:39, with its H-LU preconditioner;
 * with block Jacobi the longer basis saves iterations, 109 -> 78 on 245,760 unknowns),
 * "maxsteps" (500), "ws_gb" (device work space of the cross approximations, GB), "trunc" (1:
 * recompress each approximation to eps by QR and SVD, hmatrix/src/hmatrix.cpp:270-273 and
 * hmatrix/src/mat.cpp:129-160; 0: keep the approximation at aca_factor * eps).  Returns 1 for an
 * unknown name. */
int bhg_set_hm_option( bhg_solver* s, const char* name, double value );

/* The cluster and block trees alone, on the host (hmatrix/src/cluster.cpp:91-112,
 * hmatrix/src/hmatrix.cpp:27-54): out[0..6] = depth, leaf clusters, smallest and largest leaf,
 * inadmissible leaf blocks, admissible blocks above 1,024 triangle pairs, admissible blocks of at
 * most 1,024.  v1, v2, v3, x as BEST's NAR x 3 column-major arrays. */
int bhg_hm_plan( int n, const double* v1, const double* v2, const double* v3, const double* x, int leaf, double eta,
                 long* out );

/* The H-matrix product on device-resident vectors, timed with HIP events: *ms the mean over reps,
 * *bytes the algorithmic HBM bytes of one product (the blocks' storage plus X and Y).  No
 * counterpart in the reference: it measures the kernel hmatrix/src/hmatrix.cpp:441-471 runs. */
int bhg_hm_time_product( bhg_solver* s, int nrhs, int reps, double* ms, double* bytes );

/* The leaf blocks of the built H-matrix, 6 longs each: 1 if low-rank (else 0, dense), rank, first
 * row, rows, first column, columns (unknowns in tree order); at most cap blocks into out (may be
 * null), *count = their number.  perm (may be null; n ints) = BEST's triangle (from 0) at each
 * tree position.  A diagnostic with no counterpart in the reference's C ABI (its HStats,
 * hmatrix/src/hmatrix.h:58-70, are the aggregates of these). */
int bhg_hm_blocks( bhg_solver* s, long* out, long cap, int* perm, long* count );

/* y = (GR + deflation) x with the built H-matrix (hmatrix/src/hmatrix.cpp:441-471), BEST's order. */
int bhg_hm_apply( bhg_solver* s, int nrhs, const double* x, long ldx, double* y, long ldy );

/* ------------------------------------------------------------------ a series of meshes */

/* The series of triangulations of one body at different N that rcoal makes so that BEST's results
 * can be extrapolated linearly in 1/N (README.md:11-15): every mesh resident, the whole FP32-LU +
 * FP64-refinement solve of all of them (assembly, factorization, refinement) captured as one HIP
 * graph and launched at once.  No counterpart in the reference's call surface, which runs BEST
 * once per mesh. */
typedef struct bhg_series bhg_series;

bhg_series* bhg_series_create( int device, int nmesh );
void bhg_series_destroy( bhg_series* s );
const char* bhg_series_last_error( bhg_series* s );

/* Mesh k as BEST reads it (as bhg_set_mesh). */
int bhg_series_set_mesh( bhg_series* s, int k, int n, const double* v1, const double* v2, const double* v3,
                         double truearea, double truevol );
/* Mesh k from BEST's mesh file (as bhg_set_mesh_file). */
int bhg_series_set_mesh_file( bhg_series* s, int k, const char* path, int iarea );

/* The autotuner: for each mesh, the pair kernel's strip of source triangles per workgroup (16, 32,
 * 64, 128) and the FP32 GEMM's K slice (16, 32, 64) that time fastest, the refinement steps an
 * eager solve needed; then, by timing six graphs, whether the meshes run side by side, one after
 * another, or staggered (each starts when the one before it has half of its rows left to factor),
 * and whether they keep those K slices or all take 32 (what is fastest for a mesh alone can be
 * slower when the meshes share the GPU).  Captures the schedule. */
int bhg_series_autotune( bhg_series* s, int iarea, double cg, int ivisco );
/* gemm_bk: 16, 32 or 64 on every tile, or 0 (a single solve's default): 16 on the 128 x 128 tiles
 * of the trailing updates and 32 on the smaller ones. */
int bhg_series_set_tuning( bhg_series* s, int k, int strip, int gemm_bk, int ir_iters );
/* bhg_set_option for every mesh of the series (the schedule is captured again). */
int bhg_series_set_option( bhg_series* s, const char* name, double value );
/* out[0..3] = strip, gemm_bk, ir_iters, side by side (1), chained (0) or staggered (2). */
int bhg_series_get_tuning( bhg_series* s, int k, int* out );

/* One launch of the captured schedule (captured first if the options changed), then per mesh: the
 * pivot check (a mesh that would need row exchanges is solved again on the pivoted path), more
 * refinement steps if the correction is not yet below the tolerance, and the reductions; out[k]
 * as bhg_solve_mesh's.  *graph_ms: the device time of the launch. */
int bhg_series_solve( bhg_series* s, int iarea, double cg, int ivisco, int icov, bhg_result* out, double* graph_ms );
int bhg_series_get_stats( bhg_series* s, int k, bhg_stats* out );

/* ---- the .be file ---------------------------------------------------------------------------
 * BEST's OUTPUT (best/output.f:1-176, called at best/best.f:498) writes one mesh's results to the
 * .be file; these two write the same records, byte for byte, from a bhg_result.  What OUTPUT takes
 * besides the result (best/output.f:1-3): */
typedef struct bhg_be_params {
   const char* ver;           /* VER, padded or cut to 16 characters (BEST: "4.0 SFSU/ARAGON", best/best.f:180) */
   int iarea;                 /* IAREA: 1 prints True Area, Areafactor, True Volume and CHI */
   double truearea, truevol;  /* TRUEAREA, TRUEVOL (A^2, A^3) */
   int mw;                    /* MW: 1 prints the molecular weight and ETA */
   double wm;                 /* WM (g/mol) */
   int nar;                   /* NAR: triangles */
   int ivisco, icov;          /* IVISCO, ICOV: the viscosity block, at the centre of viscosity */
   double cg;                 /* CG: 0 prints the pre-averaging line */
} bhg_be_params;
/* The .be text into buf (len bytes with the terminating NUL; NULL or 0 to size it): returns the
 * length of the whole text, or -1 with the message in err when the tensors or the viscosity values
 * are not finite (best/best.f:483-496: BEST then stops and the .be file stays empty). */
long bhg_format_be( const bhg_result* r, const bhg_be_params* p, char* buf, long len, char* err, int errlen );
/* The same text written to path (replaced); 0 or 1 with the message in err. */
int bhg_write_be( const char* path, const bhg_result* r, const bhg_be_params* p, char* err, int errlen );

/* ---- the mesh file ---------------------------------------------------------------------------
 * The triangulation as BEST reads it (best/best.f:189-220) and rcoal writes it
 * (rcoal/scoalesce.f:891-899: the list-directed header NAR TRUEAREA TRUEVOL, then format(9e17.8)
 * per triangle).  v1, v2, v3 are NAR x 3 column-major, as BEST's V1, V2, V3.
 * bhg_read_best: with v1, v2 or v3 NULL only *nar (and the header's areas) are returned, to size
 * the arrays; cap = the triangles the arrays hold.  iarea 1 reads TRUEAREA and TRUEVOL from the
 * header, 0 reads NAR alone as BEST does with -na.  Nonzero with the message in err on a missing,
 * short or malformed file (BEST stops there with the Fortran runtime's error). */
int bhg_read_best( const char* path, int iarea, int cap, int* nar, double* truearea, double* truevol, double* v1, double* v2,
                   double* v3, char* err, int errlen );
int bhg_write_best( const char* path, int nar, double truearea, double truevol, const double* v1, const double* v2,
                    const double* v3, char* err, int errlen );

/* Least-squares line y = intercept + slope / N through nmesh points: the 1/N extrapolation. */
int bhg_extrapolate_1n( int nmesh, const int* n, const double* y, double* intercept, double* slope );

#ifdef __cplusplus
}
#endif

#endif
