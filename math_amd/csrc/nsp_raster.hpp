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

// Shadow rays towards one point light, rasterised as the light sees the mesh:
// every shadow ray of the depth ends at the light, so, projected from the
// light, it is a single point, and the simplices it can hit are those whose
// projection covers that point.  The rays are binned by the cell of their hit
// point's projection (a count, a scan and a scatter into cell order); then
// each simplex's footprint cells are culled by its projected face planes and
// every ray of the kept cells gets the any-hit test the shadow walk applies to
// a leaf candidate.  The occlusion of each ray is therefore the walk's.  Used
// when every vertex lies well in front of the light along the axis from the
// light to the mesh centre (then a simplex's footprint is the simplex of its
// projected vertices).
struct SRGrid {
    double light[NSP_MAX_DIM];
    double axis[NSP_MAX_DIM];
    double basis[NSP_MAX_DIM - 1][NSP_MAX_DIM];
    double smin[NSP_MAX_DIM - 1];
    double inv_cell[NSP_MAX_DIM - 1];
    int grid[NSP_MAX_DIM - 1];
};

struct ShadowRasterPrep {
    bool ok = false;      // the light sees the mesh in front of it (else the walks take the shadow rays)
    int n = 0;
    int64_t F = 0;
    double light[NSP_MAX_DIM] = {0};
    SRGrid grid{};
    int64_t cells = 0;
    int64_t units = 0;
    DevBuf box;           // int32 [F][2(n-1)]: first and last cell per axis of each footprint
    DevBuf planes;        // float [F][n][n]: the footprint's face planes in cell coordinates
    DevBuf unit;          // int2 (simplex, chunk of 64 cells of its box)
    DevBuf count, start, tiles;  // u32 per cell: rays, first ray in cell order; scan tiles
};

// Projection, grid, footprints and work units for the light at `light` (once
// per mesh and light position); false when the light does not see the whole
// mesh in front of it.
bool shadow_raster_prepare(ShadowRasterPrep& sp, const MeshDev& m, const double* light, hipStream_t s);
// Per-ray buffers of the binned shadow rays (scap rays).
void shadow_raster_buffers(RenderBuffers& b, int n, int64_t scap);
// Occlusion of the shadow rays of one depth (queue sq, counters dc); the
// contributions of the rays that reach the light are added to Ls (and to the
// direct-light AOV when direct is set), as the shadow walks add them.
// rays: when not null, the number of rays is added to it.  The queue holds
// each ray's hit point alone (shade's spoint mode); the ray is rebuilt from it
// towards the light at light32 (the FP32 position shading used) with the ray
// offset shading used.
void shadow_raster_launch(const ShadowRasterPrep& sp, const MeshDev& m, const TreeArgs& A, ShadowQueue<float> sq,
                          int64_t scap, const unsigned long long* dc, RenderBuffers& b, float* Ls, int64_t cap,
                          float* direct, const float* light32, float offset, unsigned long long* rays,
                          hipStream_t s);

}  // namespace nsp
