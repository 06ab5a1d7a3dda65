// nsp_batch.hip -- the batched persistent launch: one sequence of wavefront
// launches renders many scenes of uneven size and dimension.
//
// All scenes' paths form one global path space (scene by scene, pixel by
// pixel, sample by sample).  A batch is a range of it; every ray carries its
// scene slot, and each stage kernel dispatches on the slot's dimension n per
// ray (scenes are contiguous in the path space, so a wave rarely mixes n).
// The stage bodies are the single-scene ones (nsp_stages.hpp), so a scene's
// image equals the image of a single-scene render with the same batch rule.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nsp_batch.hpp"
#include "nsp_device.hpp"
#include "nsp_stages.hpp"
#include "nsp_trace.hpp"

namespace nsp {

#define NSP_DISPATCH_N(n, F) \
    switch (n) {             \
    case 3: F(3); break;     \
    case 4: F(4); break;     \
    case 5: F(5); break;     \
    case 6: F(6); break;     \
    default: break;          \
    }

__device__ __forceinline__ int find_slot(const int64_t* __restrict__ begins, int nslots, int64_t gp)
{
    int lo = 0, hi = nslots - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= gp)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_multi_reset(unsigned long long* cnt, int ncnt, const BatchDesc* __restrict__ list,
                              const long long* __restrict__ batch_idx, BatchDesc* __restrict__ cur)
{
    int i = threadIdx.x;
    for (int k = i; k < ncnt; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    if (i == 0) {
        BatchDesc d = list[*batch_idx];
        *cur = d;
        cnt[0] = (unsigned long long)d.paths;
    }
}

__global__ void k_multi_totals(const unsigned long long* __restrict__ cnt, int depth_slots,
                               unsigned long long* __restrict__ totals, long long* __restrict__ batch_idx)
{
    if (threadIdx.x != 0) return;
    unsigned long long c = 0, s = 0, p = 0;
    for (int d = 0; d < depth_slots; ++d) {
        c += cnt[d * CNT_STRIDE + 0];
        s += cnt[d * CNT_STRIDE + 1];
        p += cnt[d * CNT_STRIDE + 2];
    }
    totals[0] += c;
    totals[1] += s;
    totals[2] += p;
    *batch_idx += 1;
}

__global__ void k_multi_pass_reset(unsigned long long* totals, long long* batch_idx)
{
    if (threadIdx.x < 3) totals[threadIdx.x] = 0;
    if (threadIdx.x == 0) *batch_idx = 0;
}

template <int N>
__device__ __forceinline__ void mraygen_one(const SceneSlot& sl, int64_t local, RayQueue<float>& q, int64_t cap,
                                            int64_t p)
{
    float o[N], d[N];
    primary_ray<N, float>(sl.sc, local / sl.S, (int)(local % sl.S), o, d);
#pragma unroll
    for (int j = 0; j < N; ++j) {
        q.org[(int64_t)j * cap + p] = o[j];
        q.dir[(int64_t)j * cap + p] = d[j];
    }
}

__global__ void __launch_bounds__(TRACE_BLOCK)
    k_mraygen(const SceneSlot* __restrict__ slots, const int64_t* __restrict__ begins, int nslots,
              const BatchDesc* __restrict__ desc, RayQueue<float> q, int32_t* __restrict__ qscene, int64_t cap,
              float* __restrict__ L)
{
    const int64_t paths = desc->paths, path0 = desc->pixel0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x; p < paths; p += stride) {
        const int64_t gp = path0 + p;
        const int s = find_slot(begins, nslots, gp);
        const SceneSlot& sl = slots[s];
        const int64_t local = gp - sl.path_begin;
#define F(NN) mraygen_one<NN>(sl, local, q, cap, p)
        NSP_DISPATCH_N(sl.n, F)
#undef F
        qscene[p] = s;
        q.path[p] = (int32_t)p;
        q.beta[p] = 1.0f;
        q.beta[cap + p] = 1.0f;
        q.beta[2 * cap + p] = 1.0f;
        q.level[p] = 1.0f;
        L[p] = 0.0f;
        L[cap + p] = 0.0f;
        L[2 * cap + p] = 0.0f;
    }
}

template <int N, bool ANY>
__device__ __forceinline__ HitResult mtrace_one(const TreeArgs& A, const float* org, const float* dir, int64_t cap,
                                                int64_t i, float tmax, int* st)
{
    float o[N], d[N];
    double o64[N], d64[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        o[k] = org[(int64_t)k * cap + i];
        d[k] = dir[(int64_t)k * cap + i];
        o64[k] = o[k];
        d64[k] = d[k];
    }
    return traverse<N, float, ANY>(A, o, d, tmax, st, o64, d64);
}

__global__ void __launch_bounds__(TRACE_BLOCK)
    k_mclosest(const SceneSlot* __restrict__ slots, RayQueue<float> q, const int32_t* __restrict__ qscene, int64_t cap,
               const unsigned long long* __restrict__ count_in)
{
    __shared__ int stack[MAX_LEVELS * TRACE_BLOCK];
    int* st = stack + threadIdx.x;
    const int64_t count = (int64_t)*count_in;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x; i < count; i += stride) {
        const SceneSlot& sl = slots[qscene[i]];
        HitResult h{-1, 0.0, false};
#define F(NN) h = mtrace_one<NN, false>(sl.A, q.org, q.dir, cap, i, INFINITY, st)
        NSP_DISPATCH_N(sl.n, F)
#undef F
        q.hit[i] = h.idx;
        q.hit_t[i] = (float)h.t;
    }
}

struct UniformOut {
    bool next, shadow;
    float norg[NMAX], ndir[NMAX], nbeta[3], nlevel;
    float sorg[NMAX], sdir[NMAX], stmax, contrib[3];
};

template <int N>
__device__ __forceinline__ void mshade_one(const SceneSlot& sl, int depth, int64_t local, const RayQueue<float>& qin,
                                           int64_t cap, int64_t i, int* st, unsigned long long* probe,
                                           float* add, UniformOut& u)
{
    float o[N], d[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        o[k] = qin.org[(int64_t)k * cap + i];
        d[k] = qin.dir[(int64_t)k * cap + i];
    }
    float beta[3] = {qin.beta[i], qin.beta[cap + i], qin.beta[2 * cap + i]};
    ShadeOut<N, float> out;
    shade_one<N, float>(sl.sc, sl.A, sl.M, depth, local / sl.S, (int)(local % sl.S), o, d, qin.hit[i], qin.hit_t[i],
                        beta, qin.level[i], st, probe, nullptr, out);
    add[0] = out.add[0];
    add[1] = out.add[1];
    add[2] = out.add[2];
    u.next = out.next;
    u.shadow = out.shadow;
    if (out.next) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            u.norg[k] = out.norg[k];
            u.ndir[k] = out.ndir[k];
        }
        u.nbeta[0] = out.nbeta[0];
        u.nbeta[1] = out.nbeta[1];
        u.nbeta[2] = out.nbeta[2];
        u.nlevel = out.nlevel;
    }
    if (out.shadow) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            u.sorg[k] = out.sorg[k];
            u.sdir[k] = out.sdir[k];
        }
        u.stmax = out.stmax;
        u.contrib[0] = out.contrib[0];
        u.contrib[1] = out.contrib[1];
        u.contrib[2] = out.contrib[2];
    }
}

__global__ void __launch_bounds__(TRACE_BLOCK)
    k_mshade(const SceneSlot* __restrict__ slots, const BatchDesc* __restrict__ desc, int depth, RayQueue<float> qin,
             const int32_t* __restrict__ qin_scene, RayQueue<float> qout, int32_t* __restrict__ qout_scene,
             ShadowQueue<float> sq, int32_t* __restrict__ sq_scene, int64_t cap, unsigned long long* __restrict__ cnt,
             float* __restrict__ L)
{
    __shared__ int stack[MAX_LEVELS * TRACE_BLOCK];
    int* st = stack + threadIdx.x;
    const int64_t count = (int64_t)cnt[depth * CNT_STRIDE + 0];
    unsigned long long* next_count = &cnt[(depth + 1) * CNT_STRIDE + 0];
    unsigned long long* shadow_count = &cnt[depth * CNT_STRIDE + 1];
    unsigned long long* probe_count = &cnt[depth * CNT_STRIDE + 2];
    const int64_t path0 = desc->pixel0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t base = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
    const int64_t iters = (count + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i = base + it * stride;
        UniformOut u;
        u.next = false;
        u.shadow = false;
        int32_t path = 0, s = 0, n = 0;
        if (i < count) {
            path = qin.path[i];
            s = qin_scene[i];
            const SceneSlot& sl = slots[s];
            n = sl.n;
            const int64_t local = path0 + path - sl.path_begin;
            float add[3] = {0, 0, 0};
#define F(NN) mshade_one<NN>(sl, depth, local, qin, cap, i, st, probe_count, add, u)
            NSP_DISPATCH_N(sl.n, F)
#undef F
#pragma unroll
            for (int c = 0; c < 3; ++c) L[c * cap + path] += add[c];
        }
        int64_t slot = wave_append(u.next, next_count);
        if (slot >= 0) {
            for (int k = 0; k < n; ++k) {
                qout.org[(int64_t)k * cap + slot] = u.norg[k];
                qout.dir[(int64_t)k * cap + slot] = u.ndir[k];
            }
            qout.path[slot] = path;
            qout_scene[slot] = s;
            qout.beta[slot] = u.nbeta[0];
            qout.beta[cap + slot] = u.nbeta[1];
            qout.beta[2 * cap + slot] = u.nbeta[2];
            qout.level[slot] = u.nlevel;
        }
        int64_t ss = wave_append(u.shadow, shadow_count);
        if (ss >= 0) {
            for (int k = 0; k < n; ++k) {
                sq.org[(int64_t)k * cap + ss] = u.sorg[k];
                sq.dir[(int64_t)k * cap + ss] = u.sdir[k];
            }
            sq.tmax[ss] = u.stmax;
            sq.path[ss] = path;
            sq_scene[ss] = s;
            sq.contrib[ss] = u.contrib[0];
            sq.contrib[cap + ss] = u.contrib[1];
            sq.contrib[2 * cap + ss] = u.contrib[2];
        }
    }
}

__global__ void __launch_bounds__(TRACE_BLOCK)
    k_mshadow(const SceneSlot* __restrict__ slots, ShadowQueue<float> sq, const int32_t* __restrict__ sq_scene,
              int64_t cap, const unsigned long long* __restrict__ count_in, float* __restrict__ L)
{
    __shared__ int stack[MAX_LEVELS * TRACE_BLOCK];
    int* st = stack + threadIdx.x;
    const int64_t count = (int64_t)*count_in;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x; i < count; i += stride) {
        const SceneSlot& sl = slots[sq_scene[i]];
        HitResult h{-1, 0.0, false};
        const float tmax = sq.tmax[i];
#define F(NN) h = mtrace_one<NN, true>(sl.A, sq.org, sq.dir, cap, i, tmax, st)
        NSP_DISPATCH_N(sl.n, F)
#undef F
        if (h.idx < 0) {
            int32_t path = sq.path[i];
            L[path] += sq.contrib[i];
            L[cap + path] += sq.contrib[cap + i];
            L[2 * cap + path] += sq.contrib[2 * cap + i];
        }
    }
}

// A pixel's samples are consecutive paths; the thread of sample 0 sums them in
// order into the scene's image.
__global__ void __launch_bounds__(256)
    k_maccumulate(const SceneSlot* __restrict__ slots, const int64_t* __restrict__ begins, int nslots,
                  const BatchDesc* __restrict__ desc, const float* __restrict__ L, int64_t cap)
{
    const int64_t paths = desc->paths, path0 = desc->pixel0;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < paths; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t gp = path0 + p;
        const int s = find_slot(begins, nslots, gp);
        const SceneSlot& sl = slots[s];
        const int64_t local = gp - sl.path_begin;
        if (local % sl.S != 0) continue;
        float s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < sl.S; ++k) {
            s0 += L[p + k];
            s1 += L[cap + p + k];
            s2 += L[2 * cap + p + k];
        }
        const int64_t px = local / sl.S;
        sl.pix_sum[px] += s0;
        sl.pix_sum[sl.pixels + px] += s1;
        sl.pix_sum[2 * sl.pixels + px] += s2;
        sl.pix_cnt[px] += sl.S;
    }
}

static unsigned mgrid(int64_t n, int block = TRACE_BLOCK, int64_t max_blocks = 8192)
{
    int64_t b = (n + block - 1) / block;
    b = std::max<int64_t>(1, std::min<int64_t>(b, max_blocks));
    return (unsigned)((b + 7) / 8 * 8);
}

void alloc_multi_buffers(MultiBuffers& b, int64_t cap, int max_depth)
{
    if (b.cap >= cap && b.depth_slots >= max_depth + 2) return;
    b.cap = cap;
    b.depth_slots = max_depth + 2;
    for (DevBuf* q : {b.qa, b.qb}) {
        q[0].alloc(sizeof(float) * NMAX * cap);
        q[1].alloc(sizeof(float) * NMAX * cap);
        q[2].alloc(sizeof(int32_t) * cap);
        q[3].alloc(sizeof(float) * 3 * cap);
        q[4].alloc(sizeof(float) * cap);
        q[5].alloc(sizeof(int32_t) * cap);
        q[6].alloc(sizeof(float) * cap);
    }
    b.sq[0].alloc(sizeof(float) * NMAX * cap);
    b.sq[1].alloc(sizeof(float) * NMAX * cap);
    b.sq[2].alloc(sizeof(float) * cap);
    b.sq[3].alloc(sizeof(int32_t) * cap);
    b.sq[4].alloc(sizeof(float) * 3 * cap);
    b.qa_scene.alloc(sizeof(int32_t) * cap);
    b.qb_scene.alloc(sizeof(int32_t) * cap);
    b.sq_scene.alloc(sizeof(int32_t) * cap);
    b.L.alloc(sizeof(float) * 3 * cap);
    b.counters.alloc(sizeof(unsigned long long) * CNT_STRIDE * (max_depth + 2));
    b.desc.alloc(sizeof(BatchDesc));
    b.batch_idx.alloc(sizeof(long long));
    b.totals.alloc(sizeof(unsigned long long) * 3);
}

void multi_pass_reset(MultiBuffers& b, hipStream_t s)
{
    hipLaunchKernelGGL(k_multi_pass_reset, dim3(1), dim3(64), 0, s, b.totals.as<unsigned long long>(),
                       b.batch_idx.as<long long>());
    NSP_HIP(hipGetLastError());
}

void render_batch_multi(const MultiLaunch& ml, hipStream_t s)
{
    MultiBuffers& b = *ml.buf;
    const int64_t cap = b.cap;
    RayQueue<float> qa{b.qa[0].as<float>(), b.qa[1].as<float>(), b.qa[2].as<int32_t>(), b.qa[3].as<float>(),
                       b.qa[4].as<float>(), b.qa[5].as<int32_t>(), b.qa[6].as<float>()};
    RayQueue<float> qb{b.qb[0].as<float>(), b.qb[1].as<float>(), b.qb[2].as<int32_t>(), b.qb[3].as<float>(),
                       b.qb[4].as<float>(), b.qb[5].as<int32_t>(), b.qb[6].as<float>()};
    ShadowQueue<float> sq{b.sq[0].as<float>(), b.sq[1].as<float>(), b.sq[2].as<float>(), b.sq[3].as<int32_t>(),
                          b.sq[4].as<float>()};
    int32_t* sa = b.qa_scene.as<int32_t>();
    int32_t* sb = b.qb_scene.as<int32_t>();
    unsigned long long* cnt = b.counters.as<unsigned long long>();
    const BatchDesc* desc = b.desc.as<BatchDesc>();
    float* L = b.L.as<float>();
    const unsigned g = mgrid(cap);
    hipLaunchKernelGGL(k_multi_reset, dim3(1), dim3(256), 0, s, cnt, CNT_STRIDE * (ml.max_depth + 2),
                       b.batch_list.as<BatchDesc>(), b.batch_idx.as<long long>(), b.desc.as<BatchDesc>());
    hipLaunchKernelGGL(k_mraygen, dim3(g), dim3(TRACE_BLOCK), 0, s, ml.slots, ml.path_begins, ml.nslots, desc, qa, sa,
                       cap, L);
    NSP_HIP(hipGetLastError());
    for (int depth = 0; depth <= ml.max_depth; ++depth) {
        RayQueue<float>& qin = (depth & 1) ? qb : qa;
        RayQueue<float>& qout = (depth & 1) ? qa : qb;
        int32_t* sin = (depth & 1) ? sb : sa;
        int32_t* sout = (depth & 1) ? sa : sb;
        hipLaunchKernelGGL(k_mclosest, dim3(g), dim3(TRACE_BLOCK), 0, s, ml.slots, qin, sin, cap,
                           cnt + depth * CNT_STRIDE);
        hipLaunchKernelGGL(k_mshade, dim3(g), dim3(TRACE_BLOCK), 0, s, ml.slots, desc, depth, qin, sin, qout, sout, sq,
                           b.sq_scene.as<int32_t>(), cap, cnt, L);
        hipLaunchKernelGGL(k_mshadow, dim3(g), dim3(TRACE_BLOCK), 0, s, ml.slots, sq, b.sq_scene.as<int32_t>(), cap,
                           cnt + depth * CNT_STRIDE + 1, L);
        NSP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_maccumulate, dim3(mgrid(cap, 256, 4096)), dim3(256), 0, s, ml.slots, ml.path_begins,
                       ml.nslots, desc, L, cap);
    hipLaunchKernelGGL(k_multi_totals, dim3(1), dim3(64), 0, s, cnt, ml.max_depth + 1,
                       b.totals.as<unsigned long long>(), b.batch_idx.as<long long>());
    NSP_HIP(hipGetLastError());
}

}  // namespace nsp
