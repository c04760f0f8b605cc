// Transport properties from the solutions of BEST's 11 flows.
#pragma once

#include <hip/hip_runtime.h>

namespace bhg {

constexpr int kSumSlots = 54;   // see k_sums in transport.hip
constexpr int kSumGroups = 64;  // workgroups of the reduction (partials summed on the host in order)

// The tensors of best/difco.f and the printed values of best/output.f (3x3 column-major).
struct Transport {
   double dtt[ 9 ], dtr[ 9 ], drr[ 9 ];
   double urc[ 3 ], udr[ 3 ];
   double ktt[ 9 ], ktr[ 9 ], krt[ 9 ], krr[ 9 ];
   double drr_eig[ 3 ], dtt_eig[ 3 ];
   double dr, dtrans;          // 1/3 Trace DRR(EXACT) (1/s), 1/3 Trace DTT(EXACT) (cm^2/s)
   double chiprime;            // CHI * VOLUME (A^3)
   double dvisc[ 3 ];          // centre of viscosity - centroid (-vc), else 0
};

void launch_rhs( hipStream_t st, int n, const double* xc, double* b, long ldb, int with_strain );
void launch_sums( hipStream_t st, int n, const double* xc, const double* area, const double* b, long ldb, int nstrain,
                  const double d[ 3 ], int want_bb, double* partial );
__host__ __device__ void difco_algebra( const double* sums, double cg, Transport& T );
__host__ __device__ void visco_centre( const double* sums, double d[ 3 ] );
__host__ __device__ double visco_chiprime( const double* sums, double cg );
__host__ __device__ void output_values( Transport& T );
void jacobi3( const double* a, double d[ 3 ] );  // host
// d(D_t, D_r, CHI * V) / d(sum q), q < nslots, exactly (forward mode through the algebra; host):
// ghat[o * nslots + q].
void algebra_gradient( const double* sums, double cg, int nslots, double* ghat );
// The reductions and the whole algebra on the device, into d_out (device memory): partial holds
// kSumGroups x kSumSlots doubles.
void launch_transport( hipStream_t st, int n, const double* xc, const double* area, const double* b, long ldb, int ivisco,
                       int icov, double cg, double* partial, Transport* d_out );

}  // namespace bhg
