// nsp_render.hip -- first-hit kernel and the wavefront path-tracing stages
// (gfx950).  One FP32 (or FP64) ray state per lane; stages:
//   raygen    paint_pixels' sample + ParallelProjector::ray (src/painter/painter.cpp:419-442,
//             src/painter/sampling/engine.h:86-156, src/painter/projectors/projector.h:194-201)
//   closest   Mesh::intersect_precise (src/painter/shapes/mesh.cpp:143-166) by tree traversal
//   shade     trace_path's shading, direct_diffuse_lighting's shadow-ray set-up and
//             diffuse_lighting's bounce (painter.cpp:197-397); shadow rays and bounce
//             rays are appended to compacted queues with one atomic per wave
//   shadow    light_source_is_visible (painter.cpp:142-186) as an any-hit traversal
//   accumulate Pixels::add_color_and_samples (painter.cpp:57-88), samples summed in a
//             fixed order so every image is bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "nsp_device.hpp"
#include "nsp_raster.hpp"
#include "nsp_render.hpp"
#include "nsp_stages.hpp"
#include "nsp_trace.hpp"

namespace nsp {

static unsigned grid_for(int64_t n, int block = TRACE_BLOCK, int64_t max_blocks = 8192)
{
    int64_t b = (n + block - 1) / block;
    b = std::max<int64_t>(1, std::min<int64_t>(b, max_blocks));
    // multiple of 8 so the XCD remap is a permutation
    b = (b + 7) / 8 * 8;
    return (unsigned)b;
}

// Share (per cent) of a full persistent grid given to the shadow walks when they
// run on the second stream.  Measured with one batch per pass: n = 4 (C1) 1312
// Mrays/s at 75 %, 1270 at 100 %, 1157 at 50 %; n = 5 (C2) 276 at 100 %, 258 at
// 75 %, 229 at 50 %.  -DNSP_SIDE_GRID_PCT=P sets one share for every dimension.
template <int N>
constexpr int side_grid_pct()
{
#ifdef NSP_SIDE_GRID_PCT
    return NSP_SIDE_GRID_PCT;
#else
    return N >= 5 ? 100 : 75;
#endif
}

// Blocks of a persistent launch: every CU filled to the kernel's occupancy
// (computed once per kernel), rounded to a multiple of 8 for the XCD remap.
template <class K>
static unsigned persistent_grid(K kernel)
{
    static std::mutex mu;
    static std::map<const void*, unsigned> cache;
    std::lock_guard<std::mutex> lock(mu);
    unsigned& blocks = cache[reinterpret_cast<const void*>(kernel)];
    if (!blocks) {
        int dev = 0, cus = 0, per_cu = 0;
        NSP_HIP(hipGetDevice(&dev));
        NSP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        NSP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TRACE_BLOCK, 0));
        unsigned b = (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus);
        blocks = (b + 7) / 8 * 8;
    }
    return blocks;
}

TreeArgs make_tree_args(const MeshDev& m)
{
    const TreeDev& t = m.tree;
    TreeArgs a{};
    a.nodes = t.nodes.as<int2>();
    a.entries = t.entries.as<int32_t>();
    a.soa = t.soa.as<float>();
    a.pad = t.entry_pad;
    a.rec32 = m.rec32.as<float>();
    a.layout = m.layout;
    a.rec64 = m.rec64.as<double>();
    a.lmax = t.levels - 1;
    if (t.levels > stack_levels(m.n)) raise(NSP_E_INVALID, "tree deeper than the walk's node stack");
    if ((t.levels - 1) * m.n > 32) raise(NSP_E_INVALID, "tree cell coordinates exceed the walk's packed word");
    double ext = 0;
    for (int i = 0; i < m.n; ++i) {
        a.lo[i] = (float)t.root_lo[i];
        a.size[i] = (float)t.root_size[i];
        a.inv_fine[i] = (float)((double)(1 << a.lmax) / t.root_size[i]);
        a.lo64[i] = t.root_lo[i];
        a.size64[i] = t.root_size[i];
        a.inv_fine64[i] = (double)(1 << a.lmax) / t.root_size[i];
        ext = std::max(ext, std::fabs(t.root_lo[i]) + t.root_size[i]);
    }
    a.tadv = (float)(ext * 2.0e-6);
    a.tadv64 = ext * 1.0e-12;
    a.counters = nullptr;
    return a;
}

// ---------------------------------------------------------------------------
// First hit

template <int N>
__global__ void __launch_bounds__(TRACE_BLOCK) k_first_hit(TreeArgs A, int64_t count, const float* __restrict__ org,
                                                           const float* __restrict__ dir, int32_t* __restrict__ idx,
                                                           double* __restrict__ tout, uint8_t* __restrict__ near)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x; i < count; i += stride) {
        float o[N], d[N];
        double o64[N], d64[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o[k] = org[i * N + k];
            d[k] = dir[i * N + k];
            o64[k] = o[k];
            d64[k] = d[k];
        }
        normalize<N, float>(d);
        normalize<N, double>(d64);  // Ray normalises its direction, src/com/ray.h:34-36
        HitResult h = traverse<N, float, false>(A, o, d, INFINITY, st, o64, d64);
        idx[i] = h.idx;
        tout[i] = h.idx >= 0 ? h.t : 0.0;
        if (near) near[i] = h.near ? 1 : 0;
    }
}

void launch_first_hit(const MeshDev& m, const float* org, const float* dir, int64_t count, int32_t* idx, double* t,
                      uint8_t* near, unsigned long long* counters, hipStream_t s)
{
    TreeArgs a = make_tree_args(m);
    a.counters = counters;
    unsigned g = grid_for(count);
    switch (m.n) {
    case 3: hipLaunchKernelGGL(k_first_hit<3>, dim3(g), dim3(TRACE_BLOCK), 0, s, a, count, org, dir, idx, t, near); break;
    case 4: hipLaunchKernelGGL(k_first_hit<4>, dim3(g), dim3(TRACE_BLOCK), 0, s, a, count, org, dir, idx, t, near); break;
    case 5: hipLaunchKernelGGL(k_first_hit<5>, dim3(g), dim3(TRACE_BLOCK), 0, s, a, count, org, dir, idx, t, near); break;
    case 6: hipLaunchKernelGGL(k_first_hit<6>, dim3(g), dim3(TRACE_BLOCK), 0, s, a, count, org, dir, idx, t, near); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
    NSP_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Render stages

// Selects the current batch from the pass's batch list (device memory, so a
// captured graph replays unchanged for every batch), clears its counters and
// writes the queue segment layout: the primary rays lie contiguous in QSEG
// segments of ceil(paths / QSEG); the queues of later depths have segments of
// segcap rays, the shadow queues of ssegcap.
__global__ void k_batch_reset(unsigned long long* cnt, int depth_slots, const BatchDesc* __restrict__ list,
                              const long long* __restrict__ batch_idx, BatchDesc* __restrict__ cur, int64_t segcap,
                              int64_t ssegcap)
{
    const int i = threadIdx.x;
    const int ncnt = CNT_STRIDE * depth_slots;
    for (int k = i; k < ncnt; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const BatchDesc d = list[*batch_idx];
    if (i < depth_slots) {
        cnt[(int64_t)i * CNT_STRIDE + CNT_SEGCAP] = (unsigned long long)segcap;
        cnt[(int64_t)i * CNT_STRIDE + CNT_SSEGCAP] = (unsigned long long)ssegcap;
    }
    __syncthreads();
    if (i == 0) {
        *cur = d;
        const int64_t per = (d.paths + QSEG - 1) / QSEG;
        cnt[CNT_RAYS] = (unsigned long long)d.paths;
        cnt[CNT_SEGCAP] = (unsigned long long)per;
        for (int k = 0; k < QSEG; ++k) {
            const int64_t left = d.paths - k * per;
            cnt[cnt_rayseg(k)] = (unsigned long long)(left < 0 ? 0 : left < per ? left : per);
        }
    }
}

// Adds the batch's ray counts to the pass totals and advances the batch index.
__global__ void k_batch_totals(const unsigned long long* __restrict__ cnt, int depth_slots,
                               unsigned long long* __restrict__ totals, long long* __restrict__ batch_idx)
{
    if (threadIdx.x != 0) return;
    unsigned long long c = 0, s = 0, p = 0;
    for (int d = 0; d < depth_slots; ++d) {
        const unsigned long long* dc = cnt + (int64_t)d * CNT_STRIDE;
        for (int k = 0; k < QSEG; ++k) {
            c += dc[cnt_rayseg(k)];
            s += dc[cnt_shadowseg(k)];
        }
        s += dc[CNT_EXTRA];
        p += dc[CNT_PROBE];
    }
    totals[0] += c;
    totals[1] += s;
    totals[2] += p;
    *batch_idx += 1;
}

// Clears the pass totals and the batch index (a kernel, not a memset, so it
// may sit inside captured work).
__global__ void k_pass_reset(unsigned long long* totals, long long* batch_idx, unsigned long long* trav, int ntrav,
                             int* pass_dev, int pass)
{
    int i = threadIdx.x;
    if (i < 3) totals[i] = 0;
    if (i == 0) *batch_idx = 0;
    if (i == 0 && pass_dev) *pass_dev = pass;
    if (trav && i < ntrav) trav[i] = 0;
}

void pass_reset_state(unsigned long long* totals, long long* batch_idx, int* pass_dev, int pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_pass_reset, dim3(1), dim3(64), 0, s, totals, batch_idx, (unsigned long long*)nullptr, 0,
                       pass_dev, pass);
    NSP_HIP(hipGetLastError());
}

void pass_reset(RenderBuffers& b, bool counters, int* pass_dev, int pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_pass_reset, dim3(1), dim3(64), 0, s, b.totals.as<unsigned long long>(),
                       b.batch_idx.as<long long>(), counters ? b.trav_counters.as<unsigned long long>() : nullptr, 9,
                       pass_dev, pass);
    NSP_HIP(hipGetLastError());
}

// Primary rays of the batch.  slim (primary visibility rasterised, parallel
// projection): only the origins, which the rasteriser reads, and the cleared
// shadow sums; shade of depth 0 then takes the path index from the queue
// position, the direction from the projector, beta and level 1, and stores L.
template <int N, typename T>
__global__ void __launch_bounds__(TRACE_BLOCK) k_raygen(SceneK<T> sc, const BatchDesc* __restrict__ desc,
                                                        RayQueue<T> q, int64_t cap, float* __restrict__ L,
                                                        float* __restrict__ Ls, int slim)
{
    const int64_t paths = desc->paths;
    const int64_t pixel0 = desc->pixel0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)xcd_block() * blockDim.x + threadIdx.x; p < paths; p += stride) {
        T o[N], d[N];
        primary_ray<N, T>(sc, pixel0 + p / sc.S, (int)(p % sc.S), o, d);
#pragma unroll
        for (int j = 0; j < N; ++j) q.org[(int64_t)j * cap + p] = o[j];
        Ls[p] = 0.0f;
        Ls[cap + p] = 0.0f;
        Ls[2 * cap + p] = 0.0f;
        if (slim) continue;
#pragma unroll
        for (int j = 0; j < N; ++j) q.dir[(int64_t)j * cap + p] = d[j];
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

// Waves per SIMD asked of the compiler for k_shade: 5 for the plain FP32
// variant at n <= 4 (C1: 96 VGPRs with 72 bytes of spills per lane instead of
// 122 VGPRs at 4 waves, shade 6.0 against 6.4 ms per pass; 6 and 8 waves spill
// 140 and 176 bytes and are slower), otherwise what the register use allows
// (the smooth-normal and several-light variants would spill 480 and 256 bytes).
template <int N, typename T, bool PLAIN>
constexpr int shade_min_waves()
{
    return PLAIN && N <= 4 && sizeof(T) == 4 ? 5 : 1;
}

template <int N, typename T, bool SMOOTH, bool SHAPES, bool MULTI>
__global__ void __launch_bounds__(TRACE_BLOCK, (shade_min_waves<N, T, !SMOOTH && !SHAPES && !MULTI>()))
    k_shade(SceneK<T> sc, TreeArgs A, MeshShadeArgs M, const BatchDesc* __restrict__ desc, int depth,
            RayQueue<T> qin, RayQueue<T> qout, ShadowQueue<T> sq, int64_t cap, unsigned long long* __restrict__ cnt,
            float* __restrict__ L, float* __restrict__ aov, int prim, int spoint)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    // input segment = output segments = blockIdx.x % QSEG (the grid is a
    // multiple of QSEG blocks); the blocks of a segment stride over its rays
    const unsigned long long* dc = cnt + (int64_t)depth * CNT_STRIDE;
    const int seg = (int)(blockIdx.x % QSEG);
    const int64_t count = (int64_t)dc[cnt_rayseg(seg)];
    const int64_t in0 = (int64_t)seg * (int64_t)dc[CNT_SEGCAP];
    unsigned long long* next_count = &cnt[(int64_t)(depth + 1) * CNT_STRIDE + cnt_rayseg(seg)];
    const int64_t out0 = (int64_t)seg * (int64_t)cnt[(int64_t)(depth + 1) * CNT_STRIDE + CNT_SEGCAP];
    unsigned long long* shadow_count = &cnt[(int64_t)depth * CNT_STRIDE + cnt_shadowseg(seg)];
    const int64_t sout0 = (int64_t)seg * (int64_t)dc[CNT_SSEGCAP];
    unsigned long long* probe_count = &cnt[(int64_t)depth * CNT_STRIDE + CNT_PROBE];
    const int64_t pixel0 = desc->pixel0;
    const int64_t stride = (int64_t)(gridDim.x / QSEG) * blockDim.x;
    const int64_t base = (int64_t)(blockIdx.x / QSEG) * blockDim.x + threadIdx.x;
    const int64_t scap = cap * (sc.light_count > 1 ? sc.light_count : 1);
    // every lane of a wave iterates the same number of times (wave-level appends)
    const int64_t iters = (count + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t j = base + it * stride;
        const bool valid = j < count;
        const int64_t i = in0 + j;
        ShadeOut<N, T> out;
        out.next = false;
        out.shadow = false;
        out.spoint = spoint != 0;
        int32_t path = 0;
        if (valid) {
            const int hidx = qin.hit[i];
            float beta[3], level;
            T o[N], d[N];
            if (prim) {
                // the primary queue of a slim raygen: the queue position is the
                // path index; a miss without shapes reads no ray
                path = (int32_t)i;
                beta[0] = beta[1] = beta[2] = 1.0f;
                level = 1.0f;
                const bool ray = SHAPES || hidx >= 0;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    o[k] = ray ? qin.org[(int64_t)k * cap + i] : (T)0;
                    d[k] = sc.cam_dir[k];
                }
            } else {
                path = qin.path[i];
                beta[0] = qin.beta[i];
                beta[1] = qin.beta[cap + i];
                beta[2] = qin.beta[2 * cap + i];
                level = qin.level[i];
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    o[k] = qin.org[(int64_t)k * cap + i];
                    d[k] = qin.dir[(int64_t)k * cap + i];
                }
            }
            const int64_t gp = pixel0 * sc.S + path;
            shade_one<N, T, SMOOTH, SHAPES, MULTI>(sc, A, M, depth, gp / sc.S, (int)(gp % sc.S), o, d, hidx,
                                            qin.hit_t[i], beta, level, st, probe_count, aov && depth == 0, out);
            if (prim) {
#pragma unroll
                for (int c = 0; c < 3; ++c) L[c * cap + path] = out.add[c];
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) L[c * cap + path] += out.add[c];
            }
            if (aov && depth == 0) {
#pragma unroll
                for (int c = 0; c < 3 + N; ++c) aov[c * cap + path] = out.av[c];
                // direct channels start with what the primary ray sees itself
                // (background or emitter); first-hit shadow rays add to them
#pragma unroll
                for (int c = 0; c < 3; ++c) aov[(3 + N + c) * cap + path] = out.add[c];
            }
        }
        // compaction: one atomic per wave per queue
        int64_t slot = wave_append(out.next, next_count);
        if (slot >= 0) {
            slot += out0;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                qout.org[(int64_t)k * cap + slot] = out.norg[k];
                qout.dir[(int64_t)k * cap + slot] = out.ndir[k];
            }
            qout.path[slot] = path;
            qout.beta[slot] = out.nbeta[0];
            qout.beta[cap + slot] = out.nbeta[1];
            qout.beta[2 * cap + slot] = out.nbeta[2];
            qout.level[slot] = out.nlevel;
        }
        // one shadow ray per light (a single pass when the shade body set it up)
        const int lights = MULTI ? sc.light_count : 1;
        for (int l = 0; l < lights; ++l) {
            if (MULTI) {
                out.shadow = false;
                if (valid && out.direct)
                    light_shadow_ray<N, T, SMOOTH, SHAPES>(sc, A, l, out.sp, out.sgn, out.ssn, out.mesh, out.sbeta,
                                                           out.sscol, st, probe_count, out);
            }
            int64_t ss = wave_append(out.shadow, shadow_count);
            if (ss >= 0) {
                ss += sout0;
                // spoint: the hit point alone (the shadow raster rebuilds the ray)
#pragma unroll
                for (int k = 0; k < N; ++k) sq.org[(int64_t)k * scap + ss] = out.sorg[k];
                if (!spoint) {
#pragma unroll
                    for (int k = 0; k < N; ++k) sq.dir[(int64_t)k * scap + ss] = out.sdir[k];
                    sq.tmax[ss] = out.stmax;
                }
                // a first-hit shadow ray is marked by a negative path index while
                // the direct-light AOV is collected
                sq.path[ss] = (aov && depth == 0) ? -(path + 1) : path;
                sq.contrib[ss] = out.contrib[0];
                sq.contrib[scap + ss] = out.contrib[1];
                sq.contrib[2 * scap + ss] = out.contrib[2];
                if (sq.shape_count > 0) sq.shape[ss] = out.sshape;
            }
        }
    }
}

// Entries of the FP64-decision lists: ray index, then PENDING candidates (-1: none).
constexpr int RES_STRIDE = 1 + PENDING;

__device__ __forceinline__ void push_pending(unsigned long long* res_count, int32_t* res, int64_t ray,
                                             const int* pend, int npend)
{
    const int64_t k = (int64_t)atomicAdd(res_count, 1ull);
    res[RES_STRIDE * k] = (int32_t)ray;
#pragma unroll
    for (int c = 0; c < PENDING; ++c) res[RES_STRIDE * k + 1 + c] = c < npend ? pend[c] : -1;
}

// Persistent traversal (shadow stage): each lane holds one ray walk (RayWalk)
// across the iterations of the loop and does one leaf step per iteration; when
// at least refill_lanes<N>() lanes of a wave have finished, the wave takes that
// many new rays with one atomic.  The queue's QSEG segments are the slices of
// the 8 XCDs (workgroups are dispatched round robin over them, so block b runs
// on XCD b % 8, the XCD whose shade blocks filled segment b % 8): a wave takes
// rays from its own XCD's segment, so neighbouring rays share that XCD's L2,
// and moves on to the next segments once its own is empty.  Every wave leaves
// the loop after all segments are empty and its own rays are done.
// Measured with the shadow rays of the primary hits rasterised (C1, Mrays/s,
// two runs each): 8 lanes 3183, 3194; 16 lanes 3517, 3516; 24 lanes 3646,
// 3640; 32 lanes 3646, 3647; 40 lanes 3651, 3628.  (When the depth-0 launch
// also walked those shadow rays, 16 lanes were best: 2483 against 2452 at 32.)
// -DNSP_REFILL_LANES=L sets another threshold.
template <int N>
constexpr int refill_lanes()
{
#ifdef NSP_REFILL_LANES
    return NSP_REFILL_LANES;
#else
    return 32;
#endif
}

// Minimum waves per SIMD asked of the compiler for the persistent walks
// (-DNSP_POOL_MIN_WAVES=W); by default their register use alone sets it.
#ifdef NSP_POOL_MIN_WAVES
#define NSP_POOL_BOUNDS __launch_bounds__(TRACE_BLOCK, NSP_POOL_MIN_WAVES)
#else
#define NSP_POOL_BOUNDS __launch_bounds__(TRACE_BLOCK)
#endif

template <int REFILL_LANES>
struct RayPool {
    unsigned long long* work;         // rays taken from segment k: work[k * CNT_LINE]
    const unsigned long long* segn;   // rays in segment k: segn[k * CNT_LINE]
    int64_t segcap;                   // segment k starts at k * segcap
    int64_t lo, hi;                   // the current segment's rays
    int slice, visited;
    bool drained;

    __device__ __forceinline__ RayPool(unsigned long long* w, const unsigned long long* n, int64_t cap)
        : work(w), segn(n), segcap(cap), slice((int)(blockIdx.x % QSEG)), visited(0), drained(false)
    {
        load();
    }

    __device__ __forceinline__ void load()
    {
        lo = (int64_t)slice * segcap;
        hi = lo + (int64_t)segn[slice * CNT_LINE];
    }

    // Ray index for each lane with want set (or -1); wave-uniform call.
    __device__ __forceinline__ int64_t take(bool want)
    {
        const unsigned long long idle = __ballot(want);
        const int nidle = __popcll(idle);
        if (nidle < REFILL_LANES) return -1;
        const int64_t l0 = lo, h0 = hi;
        const int lane = (int)(threadIdx.x & 63);
        const int leader = __ffsll((unsigned long long)idle) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&work[slice * CNT_LINE], (unsigned long long)nidle);
        const unsigned blo = __shfl((unsigned)base, leader), bhi = __shfl((unsigned)(base >> 32), leader);
        base = ((unsigned long long)bhi << 32) | blo;
        if (l0 + (int64_t)base + nidle >= h0) {
            slice = (slice + 1) % QSEG;
            if (++visited == QSEG)
                drained = true;
            else
                load();
        }
        if (!want) return -1;
        const int64_t i = l0 + (int64_t)base + __popcll(idle & ((1ull << lane) - 1ull));
        return i < h0 ? i : -1;
    }
};

// Closest hit of primary rays (when they are walked, not rasterised): lane =
// ray, grid-stride over the queue, the whole walk with the FP64 re-test
// inline.  Consecutive rays are samples of one pixel, so the lanes of a wave
// walk nearly the same nodes and leaves; this static assignment keeps them
// together, which measured faster for them than the persistent refill loop
// (k_closest_pool) that the incoherent bounce rays use.
template <int N, typename T, int LAY>
__global__ void __launch_bounds__(TRACE_BLOCK) k_closest(TreeArgs A, RayQueue<T> q, int64_t cap,
                                                         const unsigned long long* __restrict__ cnt)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    const int64_t count = (int64_t)cnt[CNT_RAYS];
    for (int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        RayWalk<N, T, false, LAY, false> w;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            w.o[k] = q.org[(int64_t)k * cap + i];
            w.d[k] = q.dir[(int64_t)k * cap + i];
        }
        HitResult h{-1, 0.0, false};
        if (w.start(A, (T)INFINITY)) {
            while (!w.step(A, st, nullptr, nullptr)) {
            }
            w.flush_counters(A);
            h = w.result();
        }
        q.hit[i] = h.idx;
        q.hit_t[i] = (T)h.t;
    }
}

// Closest hit of bounce rays with the persistent loop of the shadow stage
// (FP64 walks and FP32 walks over the leaf-ordered SoA records; FP32 walks over
// simplex-ordered records run in k_walk): these rays are incoherent, so
// refilling finished lanes keeps waves full.  Candidates within the FP64
// re-test band are decided afterwards by k_closest_resolve; walks with more
// than PENDING of them are walked again by k_closest_exact.
template <int N, typename T, int LAY>
__global__ void NSP_POOL_BOUNDS k_closest_pool(TreeArgs A, RayQueue<T> q, int64_t cap,
                                                              unsigned long long* __restrict__ cnt,
                                                              unsigned long long* __restrict__ redo_count,
                                                              int32_t* __restrict__ redo,
                                                              unsigned long long* __restrict__ res_count,
                                                              int32_t* __restrict__ res)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    RayPool<refill_lanes<N>()> pool(cnt + cnt_cwork(0), cnt + cnt_rayseg(0), (int64_t)cnt[CNT_SEGCAP]);
    RayWalk<N, T, false, LAY, true> w;
    int64_t ray = -1;
    for (;;) {
        if (!pool.drained) {
            const int64_t i = pool.take(ray < 0);
            if (i >= 0) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    w.o[k] = q.org[(int64_t)k * cap + i];
                    w.d[k] = q.dir[(int64_t)k * cap + i];
                }
                if (w.start(A, (T)INFINITY)) {
                    ray = i;
                } else {
                    q.hit[i] = -1;
                    q.hit_t[i] = (T)0;
                }
            }
        }
        const bool ended = ray >= 0 && w.step(A, st, nullptr, nullptr);
        if (ended) {
            if (w.deferred) {
                redo[atomicAdd(redo_count, 1ull)] = (int32_t)ray;
            } else {
                const HitResult h = w.result();
                q.hit[ray] = h.idx;
                q.hit_t[ray] = (T)h.t;
                if (w.npend > 0) push_pending(res_count, res, ray, w.pend, w.npend);
            }
            w.flush_counters(A);
            ray = -1;
        }
        if (pool.drained && __ballot(ray >= 0) == 0) break;
    }
}

template <int N>
__global__ void __launch_bounds__(256) k_closest_resolve(TreeArgs A, RayQueue<float> q, int64_t cap,
                                                         const unsigned long long* __restrict__ res_count,
                                                         const int32_t* __restrict__ res)
{
    const int64_t count = (int64_t)*res_count;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = res[RES_STRIDE * j];
        float o[N], d[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o[k] = q.org[(int64_t)k * cap + i];
            d[k] = q.dir[(int64_t)k * cap + i];
        }
        int best = q.hit[i];
        double best_t = best >= 0 ? (double)q.hit_t[i] : (double)INFINITY;
        if (pending_resolve<N, false>(A, o, d, res + RES_STRIDE * j + 1, INFINITY, &best, &best_t)) {
            q.hit[i] = best;
            q.hit_t[i] = (float)best_t;
        }
    }
}

template <int N, typename T>
__global__ void __launch_bounds__(TRACE_BLOCK) k_closest_exact(TreeArgs A, RayQueue<T> q, int64_t cap,
                                                               const unsigned long long* __restrict__ redo_count,
                                                               const int32_t* __restrict__ redo)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    const int64_t count = (int64_t)*redo_count;
    for (int64_t j = (int64_t)xcd_block() * blockDim.x + threadIdx.x; j < count;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = redo[j];
        T o[N], d[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o[k] = q.org[(int64_t)k * cap + i];
            d[k] = q.dir[(int64_t)k * cap + i];
        }
        const HitResult h = traverse<N, T, false>(A, o, d, (T)INFINITY, st, nullptr, nullptr);
        q.hit[i] = h.idx;
        q.hit_t[i] = (T)h.t;
    }
}

template <int N, typename T, int LAY>
__global__ void NSP_POOL_BOUNDS k_shadow(TreeArgs A, ShadowQueue<T> sq, int64_t scap,
                                                        unsigned long long* __restrict__ cnt,
                                                        float* __restrict__ L, int64_t cap,
                                                        float* __restrict__ direct,
                                                        unsigned long long* __restrict__ redo_count,
                                                        int32_t* __restrict__ redo,
                                                        unsigned long long* __restrict__ res_count,
                                                        int32_t* __restrict__ res)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    RayPool<refill_lanes<N>()> pool(cnt + cnt_swork(0), cnt + cnt_shadowseg(0), (int64_t)cnt[CNT_SSEGCAP]);
    RayWalk<N, T, true, LAY, true> w;
    int64_t ray = -1;
    for (;;) {
        if (!pool.drained) {
            const int64_t i = pool.take(ray < 0);
            if (i >= 0) {
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    w.o[k] = sq.org[(int64_t)k * scap + i];
                    w.d[k] = sq.dir[(int64_t)k * scap + i];
                }
                if (w.start(A, sq.tmax[i]))
                    ray = i;
                else
                    shadow_unoccluded<N, T>(sq, scap, i, L, cap, direct);
            }
        }
        const bool ended = ray >= 0 && w.step(A, st, nullptr, nullptr);
        if (ended) {
            if (w.deferred) {
                redo[atomicAdd(redo_count, 1ull)] = (int32_t)ray;
            } else if (w.best < 0) {
                if (w.npend > 0) {
                    push_pending(res_count, res, ray, w.pend, w.npend);
                } else {
                    shadow_unoccluded<N, T>(sq, scap, ray, L, cap, direct);
                }
            }
            w.flush_counters(A);
            ray = -1;
        }
        if (pool.drained && __ballot(ray >= 0) == 0) break;
    }
}

template <int N, typename T>
__global__ void __launch_bounds__(TRACE_BLOCK) k_shadow_exact(TreeArgs A, ShadowQueue<T> sq, int64_t scap,
                                                              float* __restrict__ L, int64_t cap,
                                                              float* __restrict__ direct,
                                                              const unsigned long long* __restrict__ redo_count,
                                                              const int32_t* __restrict__ redo)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    const int64_t count = (int64_t)*redo_count;
    for (int64_t j = (int64_t)xcd_block() * blockDim.x + threadIdx.x; j < count;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = redo[j];
        T o[N], d[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o[k] = sq.org[(int64_t)k * scap + i];
            d[k] = sq.dir[(int64_t)k * scap + i];
        }
        if (traverse<N, T, true>(A, o, d, sq.tmax[i], st, nullptr, nullptr).idx < 0)
            shadow_unoccluded<N, T>(sq, scap, i, L, cap, direct);
    }
}

// FP64 decisions of the pending candidates of FP32 shadow walks that found
// no FP32 hit: the light is visible when none of them is a hit before tmax.
template <int N>
__global__ void __launch_bounds__(256) k_shadow_resolve(TreeArgs A, ShadowQueue<float> sq, int64_t scap,
                                                        float* __restrict__ L, int64_t cap, float* __restrict__ direct,
                                                        const unsigned long long* __restrict__ res_count,
                                                        const int32_t* __restrict__ res)
{
    const int64_t count = (int64_t)*res_count;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = res[RES_STRIDE * j];
        float o[N], d[N];
#pragma unroll
        for (int k = 0; k < N; ++k) {
            o[k] = sq.org[(int64_t)k * scap + i];
            d[k] = sq.dir[(int64_t)k * scap + i];
        }
        int best = -1;
        double best_t = (double)sq.tmax[i];
        if (!pending_resolve<N, true>(A, o, d, res + RES_STRIDE * j + 1, sq.tmax[i], &best, &best_t))
            shadow_unoccluded<N, float>(sq, scap, i, L, cap, direct);
    }
}

// The rays of one walk launch: the closest-hit rays of depth d + 1 (counters
// cc) and the shadow rays of depth d (counters sc), both made by shade(d).
// Segment k of the combined ray space holds segment k's closest-hit rays, then
// its shadow rays; a wave takes rays from its own XCD's segment first (as
// RayPool) and then from the others, with one atomic per refill on the
// segment's counter sc[cnt_swork(k)].
template <int REFILL_LANES>
struct WalkPool {
    const unsigned long long* cc;
    unsigned long long* sc;
    int64_t csegcap, ssegcap;
    int64_t nc, total;  // the current segment's closest-hit rays and all its rays
    int slice, visited;
    bool drained;
    bool shadows;  // false: the shadow rays are rasterised from the light, the pool holds closest-hit rays only

    __device__ __forceinline__ WalkPool(const unsigned long long* c, unsigned long long* s, bool with_shadows)
        : cc(c), sc(s), csegcap((int64_t)c[CNT_SEGCAP]), ssegcap((int64_t)s[CNT_SSEGCAP]),
          slice((int)(blockIdx.x % QSEG)), visited(0), drained(false), shadows(with_shadows)
    {
        load();
    }

    __device__ __forceinline__ void load()
    {
        nc = (int64_t)cc[cnt_rayseg(slice)];
        total = nc + (shadows ? (int64_t)sc[cnt_shadowseg(slice)] : 0);
    }

    // For each lane with want set: the index of its ray in the closest-hit
    // queue (*shadow false) or in the shadow queue (*shadow true), or -1.
    // Wave-uniform call.
    __device__ __forceinline__ int64_t take(bool want, bool* shadow)
    {
        const unsigned long long idle = __ballot(want);
        const int nidle = __popcll(idle);
        if (nidle < REFILL_LANES) return -1;
        const int s0 = slice;
        const int64_t nc0 = nc, tot0 = total;
        const int lane = (int)(threadIdx.x & 63);
        const int leader = __ffsll((unsigned long long)idle) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&sc[cnt_swork(s0)], (unsigned long long)nidle);
        const unsigned blo = __shfl((unsigned)base, leader), bhi = __shfl((unsigned)(base >> 32), leader);
        base = ((unsigned long long)bhi << 32) | blo;
        if ((int64_t)base + nidle >= tot0) {
            slice = (slice + 1) % QSEG;
            if (++visited == QSEG)
                drained = true;
            else
                load();
        }
        if (!want) return -1;
        const int64_t x = (int64_t)base + __popcll(idle & ((1ull << lane) - 1ull));
        if (x >= tot0) return -1;
        *shadow = x >= nc0;
        return x < nc0 ? s0 * csegcap + x : s0 * ssegcap + (x - nc0);
    }
};

// One persistent walk launch per depth d for FP32 walks over simplex-ordered
// records: the closest-hit walks of depth d + 1's bounce rays and the shadow
// walks of depth d share the launch, so a single grid fills the machine and
// no second persistent kernel competes with it for registers, LDS and
// dispatch (two such kernels on two streams held back the launch stream's
// small kernels until the shadow walks ended).  Each lane holds one walk of
// either kind; the wave tests its lanes' leaf candidates together
// (coop_test with a per-lane mode).  Finished walks end as k_closest_pool's
// and k_shadow's do: hits written, FP64 decisions and re-walks listed, an
// unoccluded shadow ray's contribution added to Ls.  COUNT: the traversal
// counters are kept (a separate instance, so that the timed kernel carries
// no counter registers).
// Waves per SIMD asked of the compiler for k_walk: 5 for n <= 4 (C1: 2665
// Mrays/s at 5 waves with 96 VGPRs and 2 spilled, 2516 at the 4 waves its 105
// VGPRs allow; 5 blocks of 30.8 KB of LDS fill a CU); 4 for n = 5 and 3 for
// n = 6, where 5 waves would spill 33 and 55 VGPRs (and 4 waves 6 at n = 6).
template <int N>
constexpr int walk_min_waves()
{
    return N <= 4 ? 5 : N == 5 ? 4 : 3;
}

template <int N, bool COUNT>
__global__ void __launch_bounds__(TRACE_BLOCK, walk_min_waves<N>()) k_walk(TreeArgs A, unsigned long long* __restrict__ shadow_counters,
                                       RayQueue<float> q, int64_t cap, ShadowQueue<float> sq, int64_t scap,
                                       unsigned long long* __restrict__ cc, unsigned long long* __restrict__ sc,
                                       float* __restrict__ Ls, float* __restrict__ direct,
                                       int32_t* __restrict__ redo_c, int32_t* __restrict__ res_c,
                                       int32_t* __restrict__ redo_s, int32_t* __restrict__ res_s, int shadows)
{
    __shared__ int2 stack[stack_levels(N) * TRACE_BLOCK];
    int2* st = stack + threadIdx.x;
    __shared__ CoopWave coop[TRACE_BLOCK / 64];
    CoopWave& L = coop[threadIdx.x >> 6];
    const int lane = (int)(threadIdx.x & 63);
    WalkPool<refill_lanes<N>()> pool(cc, sc, shadows != 0);
    RayWalk<N, float, false, 1, true> w;
    int32_t ray = -1;  // ray indices fit 32 bits (at most 2^27 paths times 8 lights)
    bool any = false;  // the lane's walk is a shadow (any-hit) walk
    // COUNT: the lane's traversal counters over its walks (closest-hit, then
    // shadow), added to the global counters once per wave at the end
    unsigned long long acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (;;) {
        if (!pool.drained) {
            bool shadow = false;
            const int64_t i = pool.take(ray < 0, &shadow);
            if (i >= 0) {
                float tmax;
                if (shadow) {
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        w.o[k] = sq.org[(int64_t)k * scap + i];
                        w.d[k] = sq.dir[(int64_t)k * scap + i];
                    }
                    tmax = sq.tmax[i];
                } else {
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        w.o[k] = q.org[(int64_t)k * cap + i];
                        w.d[k] = q.dir[(int64_t)k * cap + i];
                    }
                    tmax = INFINITY;
                }
                any = shadow;
                L.npend[lane] = 0;  // the walk's pending list (coop_test keeps it in L)
                if (w.start(A, tmax)) {
                    ray = (int32_t)i;
                } else if (shadow) {
                    shadow_unoccluded<N, float>(sq, scap, i, Ls, cap, direct);
                } else {
                    q.hit[i] = -1;
                    q.hit_t[i] = 0.0f;
                }
            }
        }
        int e0 = 0, n = 0;
        bool ended = ray >= 0 && w.locate(A, st, &e0, &n);
        const bool live = ray >= 0 && !ended;
        coop_test<N>(A, L, w, live, e0, n, any);
        // a hit ends an any-hit walk (coop_test then leaves deferred unset)
        if (live) ended = (any && w.best >= 0) || w.deferred || w.advance(A);
        if (ended) {
            if (any) {
                if (w.deferred) {
                    redo_s[atomicAdd(sc + CNT_REDO_S, 1ull)] = (int32_t)ray;
                } else if (w.best < 0) {
                    const int np = L.npend[lane];
                    if (np > 0)
                        push_pending(sc + CNT_RES_S, res_s, ray, &L.pend[lane * PENDING], np);
                    else
                        shadow_unoccluded<N, float>(sq, scap, ray, Ls, cap, direct);
                }
                if constexpr (COUNT) w.add_to(acc + 4);
            } else {
                if (w.deferred) {
                    redo_c[atomicAdd(cc + CNT_REDO_C, 1ull)] = (int32_t)ray;
                } else {
                    const HitResult h = w.result();
                    q.hit[ray] = h.idx;
                    q.hit_t[ray] = (float)h.t;
                    const int np = L.npend[lane];
                    if (np > 0) push_pending(cc + CNT_RES_C, res_c, ray, &L.pend[lane * PENDING], np);
                }
                if constexpr (COUNT) w.add_to(acc);
            }
            ray = -1;
        }
        if (pool.drained && __ballot(ray >= 0) == 0) break;
    }
    if constexpr (COUNT) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            unsigned long long v = acc[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            unsigned long long* dst = k < 4 ? A.counters : shadow_counters;
            if (lane == 0 && dst) atomicAdd(&dst[k & 3], v);
        }
    }
}

// Sum of the samples of each pixel in sample order; Pixels::add_color_and_samples.
// A pixel's S samples are consecutive in each channel, so a block stages the
// samples of up to ACC_PIX pixels through a 32 KiB LDS tile with coalesced loads,
// and each thread then sums one pixel's row from LDS. Rows are padded to an odd
// word stride, so the 64 lanes of a wave read 64 different banks.
constexpr int ACC_PIX = 256;
constexpr int ACC_WORDS = 8192;
constexpr int ACC_MAX_S = 63;  // larger S: k_accumulate_wide

__host__ __device__ constexpr int acc_pixels_per_block(int S)
{
    return ACC_WORDS / (S | 1) < ACC_PIX ? ACC_WORDS / (S | 1) : ACC_PIX;
}

__device__ __forceinline__ void acc_channel(float* tile, const float* __restrict__ a,
                                            const float* __restrict__ b, int64_t first, int S, int R,
                                            int P, float inv_s, float* out)
{
    // index k of the block's P * S <= ACC_WORDS samples -> row k / S, column k % S.
    // The float quotient is within 2^-22 (k + 0.5) / S < 0.002 / S of (k + 0.5) / S,
    // which is at least 0.5 / S from an integer, so the truncation is exact.
    const int T = P * S;
    __syncthreads();
    for (int k = threadIdx.x; k < T; k += blockDim.x) {
        const int q = (int)(((float)k + 0.5f) * inv_s);
        const int r = k - q * S;
        float v = a[first + k];
        if (b) v += b[first + k];
        tile[q * R + r] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < P) {
        float s = 0;
        const float* row = tile + threadIdx.x * R;
        for (int j = 0; j < S; ++j) s += row[j];
        *out = s;
    }
}

__global__ void __launch_bounds__(ACC_PIX) k_accumulate(const BatchDesc* __restrict__ desc, int S,
                                                        const float* __restrict__ L,
                                                        const float* __restrict__ Ls, int64_t cap,
                                                        float* __restrict__ pix_sum,
                                                        int32_t* __restrict__ pix_cnt, int64_t pixels,
                                                        const float* __restrict__ aov,
                                                        float* __restrict__ pix_aov, int aovc)
{
    __shared__ float tile[ACC_WORDS];
    const int64_t npix = desc->paths / S;
    const int64_t pixel0 = desc->pixel0;
    const int R = S | 1;
    const int PB = acc_pixels_per_block(S);
    const float inv_s = 1.0f / (float)S;
    for (int64_t j0 = (int64_t)blockIdx.x * PB; j0 < npix; j0 += (int64_t)gridDim.x * PB) {
        const int P = (int)std::min<int64_t>(PB, npix - j0);
        const int64_t first = j0 * S;
        const int64_t px = pixel0 + j0 + threadIdx.x;
        float s = 0;
        for (int c = 0; c < 3; ++c) {
            acc_channel(tile, L + c * cap, Ls + c * cap, first, S, R, P, inv_s, &s);
            if ((int)threadIdx.x < P) pix_sum[c * pixels + px] += s;
        }
        if ((int)threadIdx.x < P) pix_cnt[px] += S;
        for (int c = 0; aov && c < aovc; ++c) {
            acc_channel(tile, aov + c * cap, nullptr, first, S, R, P, inv_s, &s);
            if ((int)threadIdx.x < P) pix_aov[c * pixels + px] += s;
        }
        if (aov && (int)threadIdx.x < P) {
            // the radiance of the pixel's first sample: a one-sample image per pass
            const int64_t f0 = first + (int64_t)threadIdx.x * S;
            for (int c = 0; c < 3; ++c) pix_aov[(aovc + c) * pixels + px] += L[c * cap + f0] + Ls[c * cap + f0];
        }
    }
}

// The same sum for S > ACC_MAX_S: one wave per pixel, lanes strided over the
// samples, then a butterfly sum across the wave.
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__global__ void __launch_bounds__(256) k_accumulate_wide(const BatchDesc* __restrict__ desc, int S,
                                                         const float* __restrict__ L,
                                                         const float* __restrict__ Ls, int64_t cap,
                                                         float* __restrict__ pix_sum,
                                                         int32_t* __restrict__ pix_cnt, int64_t pixels,
                                                         const float* __restrict__ aov,
                                                         float* __restrict__ pix_aov, int aovc)
{
    const int64_t npix = desc->paths / S;
    const int64_t pixel0 = desc->pixel0;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t j = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); j < npix; j += waves) {
        const int64_t first = j * S;
        const int64_t px = pixel0 + j;
        for (int c = 0; c < 3; ++c) {
            float v = 0;
            for (int s = lane; s < S; s += 64) v += L[c * cap + first + s] + Ls[c * cap + first + s];
            v = wave_sum(v);
            if (lane == 0) pix_sum[c * pixels + px] += v;
        }
        if (lane == 0) pix_cnt[px] += S;
        for (int c = 0; aov && c < aovc; ++c) {
            float v = 0;
            for (int s = lane; s < S; s += 64) v += aov[c * cap + first + s];
            v = wave_sum(v);
            if (lane == 0) pix_aov[c * pixels + px] += v;
        }
        if (aov && lane == 0)
            for (int c = 0; c < 3; ++c) pix_aov[(aovc + c) * pixels + px] += L[c * cap + first] + Ls[c * cap + first];
    }
}

// ---------------------------------------------------------------------------

void alloc_render_buffers(RenderBuffers& b, int n, int precision, int64_t cap, int max_depth, int lights)
{
    // DevBuf::alloc keeps a buffer that is already large enough, so scenes of
    // different dimension and precision share one workspace that only grows
    size_t ts = precision ? sizeof(double) : sizeof(float);
    b.n = n;
    b.precision = precision;
    cap = (cap + QSEG - 1) / QSEG * QSEG;  // whole queue segments
    b.cap = cap;
    b.depth_slots = max_depth + 2;
    b.lights = std::max(1, lights);
    for (DevBuf* q : {b.qa, b.qb}) {
        q[0].alloc(ts * n * cap);
        q[1].alloc(ts * n * cap);
        q[2].alloc(sizeof(int32_t) * cap);
        q[3].alloc(sizeof(float) * 3 * cap);
        q[4].alloc(sizeof(float) * cap);
        q[5].alloc(sizeof(int32_t) * cap);
        q[6].alloc(ts * cap);
    }
    int64_t scap = cap * b.lights;
    for (DevBuf* q : {b.sq[0], b.sq[1]}) {
        q[0].alloc(ts * n * scap);
        q[1].alloc(ts * n * scap);
        q[2].alloc(ts * scap);
        q[3].alloc(sizeof(int32_t) * scap);
        q[4].alloc(sizeof(float) * 3 * scap);
        q[5].alloc(sizeof(int32_t) * scap);
    }
    b.L.alloc(sizeof(float) * 3 * cap);
    b.Ls.alloc(sizeof(float) * 3 * cap);
    for (hipEvent_t* e : {&b.fork_ev, &b.join_ev, &b.done_ev[0], &b.done_ev[1]})
        if (!*e) NSP_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    b.redo.alloc(sizeof(int32_t) * scap);
    b.redo_c.alloc(sizeof(int32_t) * cap);
    b.res_c.alloc(sizeof(int32_t) * RES_STRIDE * cap);
    b.res.alloc(sizeof(int32_t) * RES_STRIDE * scap);
    b.counters.alloc(sizeof(unsigned long long) * CNT_STRIDE * (max_depth + 2));
    b.desc.alloc(sizeof(BatchDesc));
    b.totals.alloc(sizeof(unsigned long long) * 3);
    b.batch_idx.alloc(sizeof(long long));
    b.trav_counters.alloc(sizeof(unsigned long long) * 9);
}

int64_t render_bytes_per_path(int n, int precision, int lights, bool aov, bool raster, bool shadow_raster)
{
    // the per-path sizes alloc_render_buffers (and the raster keys, the AOV
    // planes, the binned shadow rays) allocate
    const int64_t ts = precision ? sizeof(double) : sizeof(float);
    const int64_t L = std::max(1, lights);
    const int64_t queue = 2 * ts * n + 4 + 12 + 4 + 4 + ts;
    const int64_t shadow = 2 * ts * n + ts + 4 + 12 + 4;
    int64_t b = 2 * queue + 2 * shadow * L + 24 + 4 * L + 4 + 4 * RES_STRIDE * (1 + L);
    if (raster) b += 8;
    if (shadow_raster) b += (4 + 4 + 4 * n + 4 + 1) * L;
    if (aov) b += 4 * aov_channels(n);
    return b;
}

RenderBuffers::~RenderBuffers()
{
    for (hipEvent_t e : {fork_ev, join_ev, done_ev[0], done_ev[1]})
        if (e) (void)hipEventDestroy(e);
}

template <typename T>
static RayQueue<T> queue_of(DevBuf* q)
{
    return RayQueue<T>{q[0].as<T>(), q[1].as<T>(), q[2].as<int32_t>(), q[3].as<float>(), q[4].as<float>(),
                       q[5].as<int32_t>(), q[6].as<T>()};
}

struct StageEv {
    hipEvent_t* ev;
    int stage;
};

template <int N, typename T>
static void render_batch_t(const RenderLaunch& rl, hipStream_t s)
{
    RenderBuffers& b = *rl.buf;
    const SceneK<T>& sc = *static_cast<const SceneK<T>*>(rl.scene);
    const int64_t cap = b.cap;
    if (std::max(1, sc.light_count) > b.lights) raise(NSP_E_INVALID, "render buffers hold fewer lights than the scene");
    const int64_t scap = cap * std::max(1, sc.light_count);
    TreeArgs A = make_tree_args(*rl.mesh);
    TreeArgs As = A;
    if (rl.counters) {
        A.counters = b.trav_counters.as<unsigned long long>();
        As.counters = A.counters + 4;
    }
    MeshShadeArgs M{rl.mesh->normal32.as<float>(), rl.mesh->rec64.as<double>(),
                    rl.mesh->has_normals ? rl.mesh->vn64.as<double>() : nullptr,
                    rl.mesh->has_normals ? rl.mesh->fnrm.as<int32_t>() : nullptr, rl.mesh->ntype.as<int32_t>()};
    RayQueue<T> qa = queue_of<T>(b.qa), qb = queue_of<T>(b.qb);
    ShadowQueue<T> sqs[2];
    for (int k = 0; k < 2; ++k)
        sqs[k] = ShadowQueue<T>{b.sq[k][0].as<T>(), b.sq[k][1].as<T>(), b.sq[k][2].as<T>(), b.sq[k][3].as<int32_t>(),
                                b.sq[k][4].as<float>(), b.sq[k][5].as<int32_t>(), sc.shape_count, nullptr,
                                sc.light_count > 1};
    unsigned long long* cnt = b.counters.as<unsigned long long>();
    const BatchDesc* desc = b.desc.as<BatchDesc>();
    const BatchDesc* list = rl.batch_list ? rl.batch_list : b.batch_list.as<BatchDesc>();
    long long* bidx = rl.batch_idx ? rl.batch_idx : b.batch_idx.as<long long>();
    unsigned long long* totals = rl.totals ? rl.totals : b.totals.as<unsigned long long>();
    float* L = b.L.as<float>();
    float* Ls = b.Ls.as<float>();
    hipStream_t ss = rl.side ? rl.side : s;  // the shadow stages' stream
    float* aov = nullptr;
    if (rl.pix_aov) {
        // allocated by the caller before any graph capture
        if (b.aov.bytes < sizeof(float) * aov_channels(N) * cap) raise(NSP_E_INVALID, "AOV buffer not allocated");
        aov = b.aov.as<float>();
    }
    unsigned g = grid_for(cap);
    // FP32 walks read the record layout chosen at compile time; FP64 walks read the FP64 records
    const bool aos = sizeof(T) == 4 && A.layout == 1;
    auto kc = aos ? k_closest<N, T, 1> : k_closest<N, T, 0>;
    // (the persistent two-kernel walks serve FP64 walks and the leaf-ordered SoA records)
    auto kcp = k_closest_pool<N, T, 0>;
    const unsigned gcp = persistent_grid(k_closest_pool<N, T, 0>);

    // shade variants: several lights (with shapes compiled in), shapes, or the plain mesh scene
    auto kshade = sc.light_count > 1
                      ? (sc.smooth ? k_shade<N, T, true, true, true> : k_shade<N, T, false, true, true>)
                  : sc.shape_count > 0
                      ? (sc.smooth ? k_shade<N, T, true, true, false> : k_shade<N, T, false, true, false>)
                      : (sc.smooth ? k_shade<N, T, true, false, false> : k_shade<N, T, false, false, false>);
    auto ks = k_shadow<N, T, 0>;
    const unsigned gs_full = persistent_grid(k_shadow<N, T, 0>);
    // beside the closest-hit walks, the persistent shadow walk leaves part of the
    // machine to the launch stream's kernels; the last depth's shadow stage has
    // no closest-hit walk beside it and takes the whole grid
    const unsigned gs_side = rl.side ? std::max(8u, (gs_full * side_grid_pct<N>() / 100 + 7) / 8 * 8) : gs_full;
    const unsigned gx = 8 * 32;  // FP64 decisions and re-walks of shadow rays: a few per cent or fewer
    // FP32 walks over the simplex-ordered records: the bounce rays' closest-hit
    // walks of depth d + 1 and the shadow walks of depth d share one launch
    // (k_walk) after shade(d), on the launch stream; otherwise the two walk
    // kernels of each depth run as before, the shadow stage on the side stream
    const bool merged = sizeof(T) == 4 && aos;
    unsigned gw = 0;
    if constexpr (sizeof(T) == 4) gw = rl.counters ? persistent_grid(k_walk<N, true>) : persistent_grid(k_walk<N, false>);
    int32_t* redo_s = b.redo.as<int32_t>();
    int32_t* res_s = b.res.as<int32_t>();
    int32_t* redo_c = b.redo_c.as<int32_t>();
    int32_t* res_c = b.res_c.as<int32_t>();
    auto timed_on = [&](hipStream_t st, int stage, auto&& launch) {
        if (!rl.marks) {
            launch();
            NSP_HIP(hipGetLastError());
            return;
        }
        hipEvent_t e0 = rl.marks->get(), e1 = rl.marks->get();
        NSP_HIP(hipEventRecord(e0, st));
        launch();
        NSP_HIP(hipGetLastError());
        NSP_HIP(hipEventRecord(e1, st));
        rl.marks->marks.emplace_back(stage, e0, e1);
    };
    auto timed = [&](int stage, auto&& launch) { timed_on(s, stage, launch); };
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(256), 0, s, cnt, rl.max_depth + 2, list, bidx,
                       b.desc.as<BatchDesc>(), cap / QSEG, cap / QSEG * std::max(1, sc.light_count));
    NSP_HIP(hipGetLastError());
    // rasterised primaries: raygen writes only what the rasteriser and shade(0) read
    const bool slim = sizeof(T) == 4 && rl.raster != nullptr;
    timed(0, [&] {
        hipLaunchKernelGGL((k_raygen<N, T>), dim3(g), dim3(TRACE_BLOCK), 0, s, sc, desc, qa, cap, L, Ls, (int)slim);
    });
    bool sr_pending = false;  // the shadow raster of depth 0 runs on the side stream
    for (int depth = 0; depth <= rl.max_depth; ++depth) {
        RayQueue<T>& qin = (depth & 1) ? qb : qa;
        RayQueue<T>& qout = (depth & 1) ? qa : qb;
        ShadowQueue<T> sq = sqs[depth & 1];
        sq.extra = cnt + depth * CNT_STRIDE + CNT_EXTRA;
        unsigned long long* dc = cnt + depth * CNT_STRIDE;
        const bool rasterise = sizeof(T) == 4 && depth == 0 && rl.raster;
        // with merged walks, the closest hits of depth > 0 were found by the walk launch of depth - 1
        if (depth == 0 || !merged) timed(rasterise ? 5 : 1, [&] {
            if constexpr (sizeof(T) == 4) {
                if (rasterise) {
                    raster_launch(*rl.raster, *rl.mesh, A, desc, qin, cap, b.keys.as<unsigned long long>(),
                                  sc.cam_dir, s);
                    return;
                }
            }
            if (depth > 0) {
                hipLaunchKernelGGL(kcp, dim3(gcp), dim3(TRACE_BLOCK), 0, s, A, qin, cap, dc, dc + CNT_REDO_C, redo_c,
                                   dc + CNT_RES_C, res_c);
                NSP_HIP(hipGetLastError());
                if constexpr (sizeof(T) == 4) {
                    hipLaunchKernelGGL((k_closest_resolve<N>), dim3(gx), dim3(256), 0, s, A, qin, cap, dc + CNT_RES_C,
                                       res_c);
                    NSP_HIP(hipGetLastError());
                }
                hipLaunchKernelGGL((k_closest_exact<N, T>), dim3(gx), dim3(TRACE_BLOCK), 0, s, A, qin, cap,
                                   dc + CNT_REDO_C, redo_c);
                return;
            }
            hipLaunchKernelGGL(kc, dim3(g), dim3(TRACE_BLOCK), 0, s, A, qin, cap, dc);
        });
        // shade(depth) refills the shadow queue that the shadow stage of depth - 2 read
        if (!merged && rl.side && depth >= 2) NSP_HIP(hipStreamWaitEvent(s, b.done_ev[depth & 1], 0));
        timed(2, [&] {
            hipLaunchKernelGGL(kshade, dim3(g), dim3(TRACE_BLOCK), 0, s, sc, A, M, desc, depth, qin, qout, sq, cap,
                               cnt, L, aov, (int)(slim && depth == 0), (int)(rl.sraster != nullptr && depth == 0));
        });
        if constexpr (sizeof(T) == 4) {
            if (merged) {
                // one walk launch: the bounce rays of depth + 1 (in qout) and the shadow rays of depth
                unsigned long long* dc1 = cnt + (depth + 1) * CNT_STRIDE;
                float* direct = (aov && depth == 0) ? aov + (int64_t)(3 + N) * cap : nullptr;
                // the shadow rays of the primary hits are rasterised from the light;
                // the fewer of later depths are walked (a rasterising launch costs
                // about 0.5 ms for C1's 663,552 footprints even with few rays)
                const bool sr = rl.sraster != nullptr && depth == 0;
                // a walk launch that takes shadow rays adds to Ls: it waits for the
                // shadow raster of depth 0, which ran on the side stream beside the
                // depth-0 walk and shade(1) (neither reads or writes Ls)
                if (sr_pending && !sr) {
                    NSP_HIP(hipStreamWaitEvent(s, b.join_ev, 0));
                    sr_pending = false;
                }
                if (sr) {
                    if (rl.side) {
                        NSP_HIP(hipEventRecord(b.fork_ev, s));
                        NSP_HIP(hipStreamWaitEvent(ss, b.fork_ev, 0));
                    }
                    timed_on(ss, 3, [&] {
                        shadow_raster_launch(*rl.sraster, *rl.mesh, A, sq, scap, dc, b, Ls, cap, direct,
                                             sc.light_pos[0], sc.ray_offset,
                                             rl.counters ? b.trav_counters.as<unsigned long long>() + 8 : nullptr,
                                             ss);
                    });
                    if (rl.side) {
                        NSP_HIP(hipEventRecord(b.join_ev, ss));
                        sr_pending = true;
                    }
                }
                timed(6, [&] {
                    if (rl.counters)
                        hipLaunchKernelGGL((k_walk<N, true>), dim3(gw), dim3(TRACE_BLOCK), 0, s, A, As.counters, qout,
                                           cap, sq, scap, dc1, dc, Ls, direct, redo_c, res_c, redo_s, res_s, sr ? 0 : 1);
                    else
                        hipLaunchKernelGGL((k_walk<N, false>), dim3(gw), dim3(TRACE_BLOCK), 0, s, A, As.counters, qout,
                                           cap, sq, scap, dc1, dc, Ls, direct, redo_c, res_c, redo_s, res_s, sr ? 0 : 1);
                    NSP_HIP(hipGetLastError());
                    hipLaunchKernelGGL((k_closest_resolve<N>), dim3(gx), dim3(256), 0, s, A, qout, cap,
                                       dc1 + CNT_RES_C, res_c);
                    NSP_HIP(hipGetLastError());
                    hipLaunchKernelGGL((k_closest_exact<N, T>), dim3(gx), dim3(TRACE_BLOCK), 0, s, A, qout, cap,
                                       dc1 + CNT_REDO_C, redo_c);
                    NSP_HIP(hipGetLastError());
                    if (sr) return;
                    hipLaunchKernelGGL((k_shadow_resolve<N>), dim3(gx), dim3(256), 0, s, As, sq, scap, Ls, cap,
                                       direct, dc + CNT_RES_S, res_s);
                    NSP_HIP(hipGetLastError());
                    hipLaunchKernelGGL((k_shadow_exact<N, T>), dim3(gx), dim3(TRACE_BLOCK), 0, s, As, sq, scap, Ls,
                                       cap, direct, dc + CNT_REDO_S, redo_s);
                });
                continue;
            }
        }
        if (rl.side) {
            NSP_HIP(hipEventRecord(b.fork_ev, s));
            NSP_HIP(hipStreamWaitEvent(ss, b.fork_ev, 0));
        }
        timed_on(ss, 3, [&] {
            float* direct = (aov && depth == 0) ? aov + (int64_t)(3 + N) * cap : nullptr;
            const unsigned gs = depth < rl.max_depth ? gs_side : gs_full;
            hipLaunchKernelGGL(ks, dim3(gs), dim3(TRACE_BLOCK), 0, ss, As, sq, scap, dc, Ls, cap, direct,
                               dc + CNT_REDO_S, redo_s, dc + CNT_RES_S, res_s);
            NSP_HIP(hipGetLastError());
            if constexpr (sizeof(T) == 4) {
                hipLaunchKernelGGL((k_shadow_resolve<N>), dim3(gx), dim3(256), 0, ss, As, sq, scap, Ls, cap, direct,
                                   dc + CNT_RES_S, res_s);
                NSP_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL((k_shadow_exact<N, T>), dim3(gx), dim3(TRACE_BLOCK), 0, ss, As, sq, scap, Ls, cap,
                               direct, dc + CNT_REDO_S, redo_s);
        });
        if (rl.side) NSP_HIP(hipEventRecord(b.done_ev[depth & 1], ss));
    }
    if (rl.side && !merged) {
        NSP_HIP(hipEventRecord(b.join_ev, ss));
        NSP_HIP(hipStreamWaitEvent(s, b.join_ev, 0));
    }
    if (sr_pending) NSP_HIP(hipStreamWaitEvent(s, b.join_ev, 0));
    timed(4, [&] {
        const int64_t max_pix = cap / std::max(1, sc.S);
        if (sc.S <= ACC_MAX_S)
            hipLaunchKernelGGL(k_accumulate, dim3(grid_for(max_pix, acc_pixels_per_block(sc.S), 8192)),
                               dim3(ACC_PIX), 0, s, desc, sc.S, L, Ls, cap, rl.pix_sum, rl.pix_cnt, rl.pixels,
                               aov, rl.pix_aov, aov_channels(N));
        else
            hipLaunchKernelGGL(k_accumulate_wide, dim3(grid_for(max_pix, 4, 8192)), dim3(256), 0, s, desc, sc.S,
                               L, Ls, cap, rl.pix_sum, rl.pix_cnt, rl.pixels, aov, rl.pix_aov, aov_channels(N));
    });
    hipLaunchKernelGGL(k_batch_totals, dim3(1), dim3(64), 0, s, cnt, rl.max_depth + 1, totals, bidx);
    NSP_HIP(hipGetLastError());
}

hipEvent_t StageMarks::get()
{
    if (used == pool.size()) {
        hipEvent_t e;
        NSP_HIP(hipEventCreate(&e));
        pool.push_back(e);
    }
    return pool[used++];
}

StageMarks::~StageMarks()
{
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void render_batch(const RenderLaunch& rl, hipStream_t s)
{
    const int n = rl.mesh->n;
    const bool dbl = rl.buf->precision != 0;
#define NSP_CASE(NN)                                 \
    case NN:                                         \
        if (dbl)                                     \
            render_batch_t<NN, double>(rl, s);       \
        else                                         \
            render_batch_t<NN, float>(rl, s);        \
        break;
    switch (n) {
        NSP_CASE(3)
        NSP_CASE(4)
        NSP_CASE(5)
        NSP_CASE(6)
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
#undef NSP_CASE
}

}  // namespace nsp
