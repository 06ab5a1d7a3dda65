// nsp_build.hpp -- mesh basis and tree build entry (host side).
#pragma once

#include "nsp_internal.hpp"

namespace nsp {

// FP64 basis build, then the 2^N-tree by per-level binning (max_depth and
// min_objects as in src/painter/shapes/mesh.cpp:33-58).
void build_mesh_device(MeshDev& m, int max_depth, int min_objects, hipStream_t s);

}  // namespace nsp
