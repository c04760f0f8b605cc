// Device parts of the adjoint (gradient) of the dense solve: declarations shared by adjoint.hip
// and the host orchestration in solver.hip.
#pragma once

#include "bh_common.h"

namespace bhg {

// Outputs differentiated: 0 D_t (1/3 trace DTT), 1 D_r (1/3 trace DRR), 2 CHI * V.
constexpr int kAdjOutputs = 3;
// Reductions they depend on (k_sums slots 0..44: DIFCO's forces and torques, the CHI moments).
constexpr int kAdjSlots = 45;

// Per triangle: values and derivatives along its own 9 vertex coordinates (v1 x y z, v2, v3)
// and the area factor (direction 9), and the contractions with the adjoint solutions.
struct AdjTri {
   double x[ 3 ], xmu[ 3 ], A;                    // incenter, XMU, geometric area
   double dx[ 3 ][ 10 ], dxmu[ 3 ][ 10 ], dA[ 10 ];
   double dtr[ 10 ];                               // trace of the self block
   double gself[ kAdjOutputs ][ 10 ];              // sum_k lambda_k,i^T dS_i F_k,i
   double grhs[ kAdjOutputs ][ 3 ];                // sum_k lambda_k,i^T dU_k,i / dx_i
   double gsx[ kAdjOutputs ][ 3 ], gsa[ kAdjOutputs ];  // d(outputs)/dx_i, d/d(corrected area) at fixed F
};

// y(o * 11 * ld + k * ld + 3 i + c) = d(output o)/dF(3 i + c, k) from ghat = d(output)/d(slot).
void launch_adj_rhs( hipStream_t st, int n, const double* xc, const double* area, const double* ghat, double* y, long ld );

// In-place transpose of nb square blocks of order m (a multiple of 32), block b at a + b * stride.
template <typename T>
void launch_transpose( hipStream_t st, int m, T* a, long ld, int nb, long stride );

// The pair terms: gsrc (n x kAdjOutputs x 10: source coordinates, then the corrected area) and
// gcol (n x kAdjOutputs x 3: the collocation point); the partials are work space of
// ceil(n / 256) * n * 30 and ceil(n / strip) * n * 9 doubles.
void launch_adj_pairs( hipStream_t st, int n, int strip, double cg, const double* src, const double* xc, const double* lam,
                       const double* F, long ld, double* gsrc_part, double* gcol_part, double* gsrc, double* gcol );

// The per-triangle terms (vertices as NAR x 3 column-major arrays, afac on the device).
void launch_adj_tri( hipStream_t st, int n, const double* v1, const double* v2, const double* v3, const double* afac, double cg,
                     const double* lam, const double* F, long ld, const double* ghat, AdjTri* out );

}  // namespace bhg
