// nsp_raster.hpp -- primary visibility of a parallel projection by rasterising
// the simplices: every primary ray of a ParallelProjector has the same
// direction, so the simplices a sample's ray can hit are those whose projected
// bounding box contains the sample.  Each (simplex, sample) pair gets the same
// FP32 test (with the FP64 re-test inline) the tree walk applies to a leaf
// candidate, and a 64-bit atomicMin on (t, simplex index) keeps the nearest
// hit with the walk's tie rule (smaller t, then smaller index).  The result is
// the walk's result: the nearest of all accepted candidates.
#pragma once

#include <vector>

#include "nsp_internal.hpp"
#include "nsp_render.hpp"

namespace nsp {

struct RasterPrep {
    int n = 0, S = 0;
    int64_t F = 0;
    int64_t ppu = 0;      // pixels of a simplex's box per work unit
    int hmax = 0;         // largest box extent - 1 along the last screen axis
    int last = 0;         // screen extent of the last screen axis
    int64_t plane = 0;    // pixels per slice of the last screen axis
    int64_t units = 0;
    bool small = false;   // every box volume and unit item count below 2^22 (float-quotient index split)
    int screen[NSP_MAX_DIM - 1] = {0};
    DevBuf box;           // int32 [F][2(n-1)]: first and last pixel per screen axis
    DevBuf planes;        // float [F][n][n]: per vertex k, the unit normal and offset of the projected face opposite k
    DevBuf unit;          // int2 (simplex, chunk of its box) in the order of the box's low corner, last axis major
    DevBuf ustart;        // int64 [last + 1]: first unit whose box starts at last-axis pixel v
};

// Boxes and work units of the mesh's simplices under the scene's projector
// (computed once per scene and sample count).
void raster_prepare(RasterPrep& rp, const MeshDev& m, const SceneK<double>& kd, int S, hipStream_t s);

// Depth-0 closest hits of the batch in flight (desc) into the ray queue q:
// clears the per-path keys, rasterises, writes q.hit / q.hit_t.  dir: the
// direction raygen gave every primary ray (the projector's, in FP32).
void raster_launch(const RasterPrep& rp, const MeshDev& m, const TreeArgs& A, const BatchDesc* desc,
                   RayQueue<float> q, int64_t cap, unsigned long long* keys, const float* dir, hipStream_t s);

}  // namespace nsp
