// nsp_batch.hpp -- many scenes of uneven size and dimension rendered by one
// sequence of wavefront launches (the batched persistent launch).
#pragma once

#include "nsp_render.hpp"

namespace nsp {

// Scene descriptor in device memory: everything a stage kernel needs to
// process a ray of this scene.
struct SceneSlot {
    int n;
    int S;                 // samples per pixel
    int64_t path_begin;    // first path of the scene in the batch's global path space
    int64_t pixels;
    float* pix_sum;        // 3 * pixels, component-major
    int32_t* pix_cnt;
    SceneK<float> sc;
    TreeArgs A;
    MeshShadeArgs M;
};

struct MultiBuffers {
    int64_t cap = 0;
    int depth_slots = 0;
    DevBuf qa[7], qb[7], sq[5];  // ray and shadow queues with NMAX components
    DevBuf qa_scene, qb_scene, sq_scene;
    DevBuf L, counters, desc, batch_list, batch_idx, totals;
};

struct MultiLaunch {
    const SceneSlot* slots;       // device
    const int64_t* path_begins;   // device, nslots + 1 entries
    int nslots;
    MultiBuffers* buf;
    int max_depth;
};

void alloc_multi_buffers(MultiBuffers& b, int64_t cap, int max_depth);
void multi_pass_reset(MultiBuffers& b, hipStream_t s);
void render_batch_multi(const MultiLaunch& ml, hipStream_t s);

}  // namespace nsp
