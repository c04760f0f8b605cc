// COMPON on the device (compon.hip).
#pragma once

#include "bh_common.h"

namespace bhg {

// Work space of device_compon (two hash tables of 2 x 3 n slots at least, per-corner ids and
// per-triangle union-find arrays); kept by the caller between meshes.
struct ComponWork {
   int hsize = 0;
   DevBuf<int> vtable, vid, parent, root, flag, lab, ecnt, out;
   DevBuf<unsigned long long> ekey;
   void ensure( int n );
};

// ICOMP (device, n ints, 1..NCOMP in the order of the surfaces' first triangles), NCOMP and NOPEN
// of the n triangles whose raw vertices d_v holds (v1, v2, v3 as NAR x 3 column-major arrays, one
// after the other), as best/compon.f90:19-115 computes them.  Synchronizes st.
void device_compon( hipStream_t st, int n, const double* d_v, int* d_icomp, ComponWork& w, int* ncomp, int* nopen );

}  // namespace bhg
