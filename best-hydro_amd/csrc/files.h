// BEST's mesh file (files.hip), for the solver's file entry points.
#pragma once

#include <string>
#include <vector>

namespace bhg {

// A mesh as BEST reads it: V1, V2, V3 as NAR x 3 column-major arrays.
struct BestMesh {
   int nar = 0;
   double truearea = 0.0, truevol = 0.0;
   std::vector<double> v1, v2, v3;
};

// best/best.f:189-220 (iarea 1 reads TRUEAREA and TRUEVOL from the header); throws with the file
// name and the line on a malformed or short file.
BestMesh read_best( const std::string& path, int iarea );

}  // namespace bhg
