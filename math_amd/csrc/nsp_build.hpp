// nsp_build.hpp -- mesh basis and tree build entry (host side).
#pragma once

#include "nsp_internal.hpp"

namespace nsp {

// FP64 basis build, then the 2^N-tree by per-level binning (max_depth and
// min_objects as in src/painter/shapes/mesh.cpp:33-58).
void build_mesh_device(MeshDev& m, int max_depth, int min_objects, hipStream_t s);
// The leaf-ordered SoA copy of the FP32 records (record layout 0), built once,
// when a caller first selects that layout.
void build_soa_device(MeshDev& m, hipStream_t s);

}  // namespace nsp
