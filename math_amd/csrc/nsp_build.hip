// nsp_build.hip -- FP64 simplex basis build and 2^N-tree construction by
// per-level binning, for gfx950.
//
// Compiled with -ffp-contract=off: the FP64 records must equal the CPU
// checker's bit for bit, so no multiply-add may be fused except the explicit
// fma() of dot() (src/com/vec.h:258-267).
//
// Tree build (replaces SpatialSubdivisionTree::decompose / extend,
// src/painter/space/tree.h:327-387, :466-525, whose CPU form pops boxes from
// a spin-locked LIFO and tests child x object with an LP):
//   level l holds A active parents and P (parent, simplex) pairs sorted by
//   (parent, simplex).  Per level:
//     1. k_child_mask   one thread per pair: overlap bitmask over the 2^N children
//     2. k_child_count  per-child counts (wave-aggregated atomics)
//     3. k_classify     leaf / active per child (reference leaf rule: depth or
//                       <= min objects, src/painter/shapes/mesh.cpp:33-58)
//     4. scan of active flags -> next-level parent ranks
//     5. one flat exclusive scan over (child slot c, pair p) in c-major order
//        of the packed flag {active, leaf}; its down-sweep scatters each pair
//        into the next level's pair list or into the leaf entry list.
//   c-major order with children numbered (c, a) keeps every child's list
//   contiguous and in ascending simplex order, the order in which the
//   reference's leaf scan breaks ties (src/painter/space/ray_intersection.h:141-150).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "nsp_build.hpp"
#include "nsp_device.hpp"
#include "nsp_scan.hpp"

namespace nsp {

// ---------------------------------------------------------------------------
// Scans

void scan_u64(const uint64_t* d_in, uint64_t* d_out, int64_t n, uint64_t* total_host, hipStream_t s)
{
    if (n <= 0) {
        if (total_host) *total_host = 0;
        return;
    }
    int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    DevBuf sums, offs;
    sums.alloc(sizeof(uint64_t) * tiles);
    offs.alloc(sizeof(uint64_t) * tiles);
    ArrayScanOp op{d_in, d_out};
    hipLaunchKernelGGL(k_tile_reduce<ArrayScanOp>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, n,
                       sums.as<uint64_t>());
    NSP_HIP(hipGetLastError());
    uint64_t total = 0;
    if (tiles == 1) {
        NSP_HIP(hipMemsetAsync(offs.p, 0, sizeof(uint64_t), s));
        NSP_HIP(hipMemcpyAsync(&total, sums.p, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    } else {
        scan_u64(sums.as<uint64_t>(), offs.as<uint64_t>(), tiles, &total, s);
    }
    hipLaunchKernelGGL(k_tile_downsweep<ArrayScanOp>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, n,
                       offs.as<uint64_t>());
    NSP_HIP(hipGetLastError());
    NSP_HIP(hipStreamSynchronize(s));
    if (total_host) *total_host = total;
}

// ---------------------------------------------------------------------------
// FP64 basis build: MeshHyperplaneSimplex ctor (src/painter/shapes/mesh_hyperplane_simplex.cpp:59-143),
// HyperplaneSimplexGeometry::set_data (src/painter/space/hyperplane_geometry.h:83-115),
// HyperplaneSimplexGeometry::constraints (hyperplane_geometry.h:120-160).

template <int N>
__global__ void __launch_bounds__(256) k_basis(int64_t F, int64_t V, const double* __restrict__ vtx,
                                               const int32_t* __restrict__ fac, const double* __restrict__ vn,
                                               const int32_t* __restrict__ fnrm, int has_normals,
                                               double* __restrict__ rec, double* __restrict__ brec,
                                               float* __restrict__ normal32, int32_t* __restrict__ ntype,
                                               int* __restrict__ status)
{
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    int32_t vi[N];
    for (int k = 0; k < N; ++k) {
        vi[k] = fac[f * N + k];
        if (vi[k] < 0 || vi[k] >= V) {
            atomicMin(status, 0);  // status 0 = bad index (min over failures)
            return;
        }
    }
    double verts[N * N];
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < N; ++q) verts[k * N + q] = vtx[(int64_t)vi[k] * N + q];
    // normal = normalize(ortho_nn(points, indices)), linear_algebra.h:315-328
    double vec[(N - 1) * N];
    for (int i = 0; i < N - 1; ++i)
        for (int q = 0; q < N; ++q) vec[i * N + q] = verts[(i + 1) * N + q] - verts[q];
    double nrm[N];
    ortho_nn<N, double>(vec, nrm);
    normalize<N, double>(nrm);
    bool finite = true;
    for (int q = 0; q < N; ++q) finite = finite && isfinite(nrm[q]);
    if (!finite) {
        atomicMin(status, 1);  // status 1 = non-finite normal
        return;
    }
    // set_data: planes through each facet opposite vertex i (i < N-1)
    double pn[(N - 1) * N], pd[N - 1];
    {
        double vectors[(N - 1) * N];
        for (int i = 0; i < N - 1; ++i)
            for (int q = 0; q < N; ++q) vectors[i * N + q] = verts[i * N + q] - verts[(N - 1) * N + q];
        for (int i = 0; i < N - 1; ++i) {
            double saved[N];
            for (int q = 0; q < N; ++q) {
                saved[q] = vectors[i * N + q];
                vectors[i * N + q] = nrm[q];
            }
            double p[N];
            ortho_nn<N, double>(vectors, p);
            for (int q = 0; q < N; ++q) vectors[i * N + q] = saved[q];
            double d = dot<N, double>(verts + (N - 1) * N, p);
            double distance = dot<N, double>(verts + i * N, p) - d;
            for (int q = 0; q < N; ++q) pn[i * N + q] = p[q] / distance;
            pd[i] = d / distance;
        }
    }
    // vertex-normal classification (None / Use / Reverse, LIMIT_COSINE 0.7)
    int type = 0, revmask = 0;
    if (has_normals) {
        bool ok = true;
        int32_t ni[N];
        for (int k = 0; k < N; ++k) {
            ni[k] = fnrm[f * N + k];
            ok = ok && ni[k] >= 0;
        }
        if (ok) {
            double dots[N];
            bool small = false, pos = true, neg = true;
            for (int k = 0; k < N; ++k) {
                dots[k] = dot<N, double>(vn + (int64_t)ni[k] * N, nrm);
                small = small || fabs(dots[k]) < 0.7;
                pos = pos && dots[k] > 0;
                neg = neg && dots[k] < 0;
            }
            if (!small) {
                if (pos) {
                    type = 1;
                } else if (neg) {
                    type = 1;
                    for (int q = 0; q < N; ++q) nrm[q] = -nrm[q];
                } else {
                    type = 2;
                    for (int k = 0; k < N; ++k)
                        if (dots[k] < 0) revmask |= 1 << k;
                }
            }
        }
    }
    double* r = rec + f * rec64_stride(N);
    for (int q = 0; q < N; ++q) {
        r[q] = nrm[q];
        r[N + q] = verts[q];
    }
    for (int k = 0; k < (N - 1) * N; ++k) r[2 * N + k] = pn[k];
    for (int i = 0; i < N - 1; ++i) r[2 * N + (N - 1) * N + i] = pd[i];
    for (int q = 0; q < N; ++q) normal32[f * N + q] = (float)nrm[q];
    ntype[f] = type | (revmask << 8);

    // overlap-test data: wrapper vertex min/max and constraints
    double* b = brec + f * build64_stride(N);
    for (int q = 0; q < N; ++q) {
        double lo = verts[q], hi = verts[q];
        for (int k = 1; k < N; ++k) {
            lo = fmin(lo, verts[k * N + q]);
            hi = fmax(hi, verts[k * N + q]);
        }
        b[q] = lo;
        b[N + q] = hi;
    }
    double* c = b + 2 * N;
    for (int i = 0; i < N - 1; ++i) {
        double len = sqrt(dot<N, double>(pn + i * N, pn + i * N));
        for (int q = 0; q < N; ++q) c[i * (N + 1) + q] = pn[i * N + q] / len;
        c[i * (N + 1) + N] = -pd[i] / len;
    }
    {
        double vectors[(N - 1) * N];
        for (int i = 0; i < N - 2; ++i)
            for (int q = 0; q < N; ++q) vectors[i * N + q] = verts[(i + 1) * N + q] - verts[q];
        for (int q = 0; q < N; ++q) vectors[(N - 2) * N + q] = nrm[q];
        double nn[N];
        ortho_nn<N, double>(vectors, nn);
        normalize<N, double>(nn);
        double d = dot<N, double>(verts, nn);
        bool to_vertex = dot<N, double>(verts + (N - 1) * N, nn) - d >= 0;
        for (int q = 0; q < N; ++q) c[(N - 1) * (N + 1) + q] = to_vertex ? nn[q] : -nn[q];
        c[(N - 1) * (N + 1) + N] = to_vertex ? -d : d;
        double* e = c + N * (N + 1);
        double d0 = dot<N, double>(verts, nrm);
        for (int q = 0; q < N; ++q) e[q] = nrm[q];
        e[N] = -d0;
    }
}

// ---------------------------------------------------------------------------
// Box / simplex overlap (conservative): the separating planes of
// shapes_not_intersect_by_planes (src/painter/space/shape_intersection.h:150-189)
// -- box faces, the simplex's N facet constraints and its hyperplane --
// evaluated on the box centre and half sizes with guard delta.  A superset of
// the reference's LP-exact overlap (shape_intersection.h:191-278).

struct RootBox {
    double lo[NSP_MAX_DIM];
    double size[NSP_MAX_DIM];
};

template <int N>
__device__ __forceinline__ bool overlap(const double* __restrict__ b, const double* org, const double* size,
                                        double delta)
{
    double c[N], h[N];
    for (int i = 0; i < N; ++i) {
        h[i] = size[i] * 0.5;
        c[i] = org[i] + h[i];
    }
    for (int i = 0; i < N; ++i) {
        if (b[i] > org[i] + size[i] + delta) return false;
        if (b[N + i] < org[i] - delta) return false;
    }
    const double* cons = b + 2 * N;
    for (int k = 0; k < N; ++k) {
        const double* a = cons + k * (N + 1);
        double reach = fabs(a[0]) * h[0];
        for (int i = 1; i < N; ++i) reach = fma(fabs(a[i]), h[i], reach);
        double v = dot<N, double>(a, c) + a[N];
        if (v + reach < -delta) return false;
    }
    const double* a = cons + N * (N + 1);
    double reach = fabs(a[0]) * h[0];
    for (int i = 1; i < N; ++i) reach = fma(fabs(a[i]), h[i], reach);
    double v = dot<N, double>(a, c) + a[N];
    if (fabs(v) > reach + delta) return false;
    return true;
}

template <int N>
__global__ void __launch_bounds__(256) k_child_mask(int64_t P, int level, RootBox root, double delta,
                                                    const int32_t* __restrict__ pair_a,
                                                    const int32_t* __restrict__ pair_s,
                                                    const int32_t* __restrict__ parent_q,
                                                    const double* __restrict__ brec, uint64_t* __restrict__ mask)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int a = pair_a[p];
    int64_t s = pair_s[p];
    const double* b = brec + s * build64_stride(N);
    double brl[build64_stride(N)];
    for (int k = 0; k < build64_stride(N); ++k) brl[k] = b[k];
    int cl = level + 1;
    double cell[N];
    for (int i = 0; i < N; ++i) cell[i] = ldexp(root.size[i], -cl);
    uint64_t m = 0;
    for (int c = 0; c < (1 << N); ++c) {
        double org[N];
        for (int i = 0; i < N; ++i) {
            int64_t q = 2 * (int64_t)parent_q[a * N + i] + ((c >> i) & 1);
            org[i] = root.lo[i] + (double)q * cell[i];
        }
        if (overlap<N>(brl, org, cell, delta)) m |= 1ull << c;
    }
    mask[p] = m;
}

template <int N>
__global__ void __launch_bounds__(256) k_child_count(int64_t P, int A, const int32_t* __restrict__ pair_a,
                                                     const uint64_t* __restrict__ mask, uint32_t* __restrict__ cnt)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = p < P;
    int a = valid ? pair_a[p] : -1;
    uint64_t m = valid ? mask[p] : 0;
    int a0 = __shfl(a, 0, 64);
    // uniform parent across the wave (first lane's parent shared by all valid lanes)
    bool uniform = __all(!valid || a == a0);
    for (int c = 0; c < (1 << N); ++c) {
        bool bit = (m >> c) & 1;
        if (uniform) {
            uint64_t bal = __ballot(bit);
            if (__lane_id() == 0 && bal) atomicAdd(&cnt[(int64_t)c * A + a0], (uint32_t)__popcll(bal));
        } else if (bit) {
            atomicAdd(&cnt[(int64_t)c * A + a], 1u);
        }
    }
}

__global__ void k_classify(int64_t C, const uint32_t* __restrict__ cnt, int min_objects, int child_can_split,
                           uint64_t* __restrict__ act)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= C) return;
    act[j] = (child_can_split && cnt[j] > (uint32_t)min_objects) ? 1 : 0;
}

__global__ void k_set_parents(int A, const int32_t* __restrict__ parent_node, int64_t child_base, int2* nodes)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A) return;
    nodes[parent_node[a]] = make_int2((int)(child_base + a), -A);
}

template <int N>
__global__ void k_child_meta(int A, int level, int64_t child_base, const int32_t* __restrict__ parent_q,
                             int8_t* __restrict__ node_level, int32_t* __restrict__ node_q)
{
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ((int64_t)A << N)) return;
    int c = (int)(j / A), a = (int)(j % A);
    node_level[child_base + j] = (int8_t)(level + 1);
    for (int i = 0; i < N; ++i) node_q[(child_base + j) * N + i] = 2 * parent_q[a * N + i] + ((c >> i) & 1);
}

// Flat (c, p) flag: bit c of mask[p]; packed {leaf count in low 32, active count in high 32}.
template <int N>
struct ScatterOp {
    int64_t P;
    int A;
    int level;
    const uint64_t* mask;
    const int32_t* pair_a;
    const int32_t* pair_s;
    const int64_t* pair_begin;
    const uint64_t* act;
    const uint64_t* act_rank;
    const uint32_t* cnt;
    const int32_t* parent_q;
    int64_t child_base;
    int64_t entry_base;
    int2* nodes;
    int32_t* entries;
    int32_t* next_pair_a;
    int32_t* next_pair_s;
    int64_t* next_pair_begin;
    int32_t* next_parent_node;
    int32_t* next_parent_q;

    __device__ uint64_t value(int64_t i) const
    {
        int64_t c = i / P, p = i - c * P;
        if (!((mask[p] >> c) & 1)) return 0;
        int64_t j = c * A + pair_a[p];
        return act[j] ? (1ull << 32) : 1ull;
    }
    __device__ void emit(int64_t i, uint64_t ex, uint64_t v) const
    {
        int64_t c = i / P, p = i - c * P;
        int a = pair_a[p];
        int64_t j = c * A + a;
        bool active = act[j] != 0;
        if (v) {
            if (active) {
                int64_t pos = (int64_t)(ex >> 32);
                next_pair_a[pos] = (int32_t)act_rank[j];
                next_pair_s[pos] = pair_s[p];
            } else {
                entries[entry_base + (int64_t)(ex & 0xffffffffull)] = pair_s[p];
            }
        }
        if (p == pair_begin[a]) {
            if (active) {
                int64_t r = (int64_t)act_rank[j];
                next_parent_node[r] = (int32_t)(child_base + j);
                next_pair_begin[r] = (int64_t)(ex >> 32);
                for (int q = 0; q < N; ++q) next_parent_q[r * N + q] = 2 * parent_q[a * N + q] + (int)((c >> q) & 1);
            } else {
                nodes[child_base + j] = make_int2((int)(entry_base + (int64_t)(ex & 0xffffffffull)), (int)cnt[j]);
            }
        }
    }
};

// Leaf-ordered FP32 SoA of the simplex records (normal, d = normal.v0, then the
// N-1 barycentric planes), so that a leaf's candidates are contiguous and a
// wave reading one leaf reads each component array coalesced.
// The FP32 test components of one simplex, in the order both layouts share.
template <int N>
__device__ inline void fp32_components(const double* r, float* out)
{
    int comp = 0;
    for (int q = 0; q < N; ++q) out[comp++] = (float)r[q];
    out[comp++] = (float)dot<N, double>(r, r + N);
    for (int i = 0; i < N - 1; ++i) {
        for (int q = 0; q < N; ++q) out[comp++] = (float)r[2 * N + i * N + q];
        out[comp++] = (float)r[2 * N + (N - 1) * N + i];
    }
    for (; comp < rec32_stride(N); ++comp) out[comp] = 0.0f;
}

template <int N>
__global__ void __launch_bounds__(256) k_pack_soa(int64_t E, int64_t pad, const int32_t* __restrict__ entries,
                                                  const double* __restrict__ rec, float* __restrict__ soa)
{
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int64_t s = entries[e];
    float c[rec32_stride(N)];
    fp32_components<N>(rec + s * rec64_stride(N), c);
    for (int k = 0; k < soa_components(N); ++k) soa[(int64_t)k * pad + e] = c[k];
}

// Simplex-ordered AoS copy of the same FP32 components, padded to 16-byte
// records: a candidate test reads its record with rec32_stride/4 vector loads,
// and the whole table (F * rec32_stride * 4 bytes) is small enough to stay in
// the last-level cache where the leaf-ordered SoA is not.
template <int N>
__global__ void __launch_bounds__(256) k_pack_rec32(int64_t F, const double* __restrict__ rec,
                                                    float* __restrict__ rec32)
{
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= F) return;
    float c[rec32_stride(N)];
    fp32_components<N>(rec + s * rec64_stride(N), c);
    for (int k = 0; k < rec32_stride(N); ++k) rec32[s * rec32_stride(N) + k] = c[k];
}

__global__ void k_iota(int64_t n, int32_t* out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)i;
}

__global__ void k_fill_i32(int64_t n, int32_t* out, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = v;
}

static unsigned blocks_for(int64_t n, int b = 256) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// ---------------------------------------------------------------------------

template <int N>
void build_basis_t(MeshDev& m, hipStream_t s)
{
    DevBuf status;
    status.alloc(sizeof(int));
    int init = 0x7fffffff;
    NSP_HIP(hipMemcpyAsync(status.p, &init, sizeof(int), hipMemcpyHostToDevice, s));
    m.rec64.alloc(sizeof(double) * m.F * rec64_stride(N));
    m.build64.alloc(sizeof(double) * m.F * build64_stride(N));
    m.normal32.alloc(sizeof(float) * m.F * N);
    m.ntype.alloc(sizeof(int32_t) * m.F);
    hipLaunchKernelGGL(k_basis<N>, dim3(blocks_for(m.F)), dim3(256), 0, s, m.F, m.V, m.vtx64.as<double>(),
                       m.fac.as<int32_t>(), m.has_normals ? m.vn64.as<double>() : nullptr,
                       m.has_normals ? m.fnrm.as<int32_t>() : nullptr, (int)m.has_normals, m.rec64.as<double>(),
                       m.build64.as<double>(), m.normal32.as<float>(), m.ntype.as<int32_t>(), status.as<int>());
    NSP_HIP(hipGetLastError());
    m.rec32.alloc(sizeof(float) * m.F * rec32_stride(N));
    hipLaunchKernelGGL(k_pack_rec32<N>, dim3(blocks_for(m.F)), dim3(256), 0, s, m.F, m.rec64.as<double>(),
                       m.rec32.as<float>());
    NSP_HIP(hipGetLastError());
    int st = 0;
    NSP_HIP(hipMemcpyAsync(&st, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NSP_HIP(hipStreamSynchronize(s));
    if (st == 0) raise(NSP_E_INVALID, "facet vertex index out of range");
    if (st == 1) raise(NSP_E_INVALID, "Simplex normal is not finite");
}

template <int N>
void build_tree_t(MeshDev& m, int max_depth, int min_objects, hipStream_t s)
{
    TreeDev& t = m.tree;
    t.n = N;
    t.max_depth = max_depth;
    // min_max_and_distance, src/painter/space/tree.h:167-226
    const double eps = 2.220446049250313e-16;
    double mn[N], mx[N], all_max = -INFINITY;
    for (int i = 0; i < N; ++i) {
        mn[i] = m.mn[i];
        mx[i] = m.mx[i];
        if (!(mn[i] < mx[i])) raise(NSP_E_INVALID, "Objects for (2^N)-tree don't form N-dimensional object");
    }
    for (int i = 0; i < N; ++i) {
        double abs_max = std::max(std::fabs(mn[i]), std::fabs(mx[i]));
        double guard = abs_max * (20 * eps);
        mn[i] -= guard;
        mx[i] += guard;
        all_max = std::max(abs_max, all_max);
    }
    double dist = all_max * (20 * eps);
    int max_div = 1 << (max_depth - 1);
    for (int i = 0; i < N; ++i)
        if (dist >= (mx[i] - mn[i]) / max_div / 2)
            raise(NSP_E_INVALID, "The minimal distance from facets is greater than one half of the minimum box size");
    RootBox root;
    for (int i = 0; i < N; ++i) {
        root.lo[i] = t.root_lo[i] = mn[i];
        root.size[i] = t.root_size[i] = mx[i] - mn[i];
    }
    t.dist_from_facet = dist;

    // node arrays grow per level; keep host-side vectors of device chunks
    // and concatenate at the end.
    std::vector<int2> h_nodes;  // assembled on device per level, copied back
    DevBuf d_nodes, d_level, d_q;
    int64_t node_cap = 1;
    auto grow_nodes = [&](int64_t need) {
        if (need <= node_cap && d_nodes.p) return;
        int64_t cap = std::max<int64_t>(need, node_cap * 2);
        DevBuf nn, nl, nq;
        nn.alloc(sizeof(int2) * cap);
        nl.alloc(sizeof(int8_t) * cap);
        nq.alloc(sizeof(int32_t) * cap * N);
        if (d_nodes.p) {
            NSP_HIP(hipMemcpyAsync(nn.p, d_nodes.p, sizeof(int2) * node_cap, hipMemcpyDeviceToDevice, s));
            NSP_HIP(hipMemcpyAsync(nl.p, d_level.p, sizeof(int8_t) * node_cap, hipMemcpyDeviceToDevice, s));
            NSP_HIP(hipMemcpyAsync(nq.p, d_q.p, sizeof(int32_t) * node_cap * N, hipMemcpyDeviceToDevice, s));
            NSP_HIP(hipStreamSynchronize(s));
        }
        std::swap(d_nodes.p, nn.p);
        std::swap(d_nodes.bytes, nn.bytes);
        std::swap(d_level.p, nl.p);
        std::swap(d_level.bytes, nl.bytes);
        std::swap(d_q.p, nq.p);
        std::swap(d_q.bytes, nq.bytes);
        node_cap = cap;
    };
    grow_nodes(1);
    {
        int8_t l0 = 0;
        int32_t q0[N] = {0};
        NSP_HIP(hipMemcpyAsync(d_level.p, &l0, 1, hipMemcpyHostToDevice, s));
        NSP_HIP(hipMemcpyAsync(d_q.p, q0, sizeof q0, hipMemcpyHostToDevice, s));
    }
    int64_t node_count = 1;

    DevBuf entries;
    int64_t entry_count = 0, entry_cap = 0;
    auto grow_entries = [&](int64_t need) {
        if (need <= entry_cap) return;
        int64_t cap = std::max<int64_t>(need, entry_cap * 2);
        DevBuf ne;
        ne.alloc(sizeof(int32_t) * cap);
        if (entries.p && entry_count)
            NSP_HIP(hipMemcpyAsync(ne.p, entries.p, sizeof(int32_t) * entry_count, hipMemcpyDeviceToDevice, s));
        NSP_HIP(hipStreamSynchronize(s));
        std::swap(entries.p, ne.p);
        std::swap(entries.bytes, ne.bytes);
        entry_cap = cap;
    };

    int levels = 1;
    if (m.F <= min_objects || max_depth <= 1) {
        grow_entries(m.F);
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(m.F)), dim3(256), 0, s, m.F, entries.as<int32_t>());
        int2 leaf = make_int2(0, (int)m.F);
        NSP_HIP(hipMemcpyAsync(d_nodes.p, &leaf, sizeof leaf, hipMemcpyHostToDevice, s));
        entry_count = m.F;
    } else {
        // level state
        int A = 1;
        int64_t P = m.F;
        DevBuf parent_node, parent_q, pair_a, pair_s, pair_begin;
        parent_node.alloc(sizeof(int32_t));
        parent_q.alloc(sizeof(int32_t) * N);
        pair_a.alloc(sizeof(int32_t) * P);
        pair_s.alloc(sizeof(int32_t) * P);
        pair_begin.alloc(sizeof(int64_t));
        NSP_HIP(hipMemsetAsync(parent_node.p, 0, sizeof(int32_t), s));
        NSP_HIP(hipMemsetAsync(parent_q.p, 0, sizeof(int32_t) * N, s));
        NSP_HIP(hipMemsetAsync(pair_a.p, 0, sizeof(int32_t) * P, s));
        NSP_HIP(hipMemsetAsync(pair_begin.p, 0, sizeof(int64_t), s));
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(P)), dim3(256), 0, s, P, pair_s.as<int32_t>());
        int level = 0;
        while (A > 0) {
            const int64_t C = (int64_t)A << N;
            const int64_t child_base = node_count;
            grow_nodes(node_count + C);
            hipLaunchKernelGGL(k_set_parents, dim3(blocks_for(A)), dim3(256), 0, s, A, parent_node.as<int32_t>(),
                               child_base, d_nodes.as<int2>());
            hipLaunchKernelGGL(k_child_meta<N>, dim3(blocks_for(C)), dim3(256), 0, s, A, level, child_base,
                               parent_q.as<int32_t>(), d_level.as<int8_t>(), d_q.as<int32_t>());
            DevBuf mask, cnt, act, act_rank;
            mask.alloc(sizeof(uint64_t) * P);
            cnt.alloc(sizeof(uint32_t) * C);
            act.alloc(sizeof(uint64_t) * C);
            act_rank.alloc(sizeof(uint64_t) * C);
            NSP_HIP(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t) * C, s));
            hipLaunchKernelGGL(k_child_mask<N>, dim3(blocks_for(P)), dim3(256), 0, s, P, level, root, dist,
                               pair_a.as<int32_t>(), pair_s.as<int32_t>(), parent_q.as<int32_t>(),
                               m.build64.as<double>(), mask.as<uint64_t>());
            NSP_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_child_count<N>, dim3(blocks_for(P)), dim3(256), 0, s, P, A, pair_a.as<int32_t>(),
                               mask.as<uint64_t>(), cnt.as<uint32_t>());
            NSP_HIP(hipGetLastError());
            // child at level+1 has reference depth level+2; split if depth < max_depth
            int can_split = (level + 2) < max_depth ? 1 : 0;
            hipLaunchKernelGGL(k_classify, dim3(blocks_for(C)), dim3(256), 0, s, C, cnt.as<uint32_t>(), min_objects,
                               can_split, act.as<uint64_t>());
            uint64_t A_next = 0;
            scan_u64(act.as<uint64_t>(), act_rank.as<uint64_t>(), C, &A_next, s);

            // totals of the flat scan first (to size outputs): reduce pass
            ScatterOp<N> op;
            op.P = P;
            op.A = A;
            op.level = level;
            op.mask = mask.as<uint64_t>();
            op.pair_a = pair_a.as<int32_t>();
            op.pair_s = pair_s.as<int32_t>();
            op.pair_begin = pair_begin.as<int64_t>();
            op.act = act.as<uint64_t>();
            op.act_rank = act_rank.as<uint64_t>();
            op.cnt = cnt.as<uint32_t>();
            op.parent_q = parent_q.as<int32_t>();
            op.child_base = child_base;
            op.entry_base = entry_count;
            op.nodes = d_nodes.as<int2>();
            const int64_t M = C == 0 ? 0 : ((int64_t)1 << N) * P;
            int64_t tiles = (M + SCAN_TILE - 1) / SCAN_TILE;
            DevBuf sums, offs;
            sums.alloc(sizeof(uint64_t) * std::max<int64_t>(tiles, 1));
            offs.alloc(sizeof(uint64_t) * std::max<int64_t>(tiles, 1));
            hipLaunchKernelGGL(k_tile_reduce<ScatterOp<N>>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, M,
                               sums.as<uint64_t>());
            NSP_HIP(hipGetLastError());
            uint64_t total = 0;
            scan_u64(sums.as<uint64_t>(), offs.as<uint64_t>(), tiles, &total, s);
            int64_t n_leaf_entries = (int64_t)(total & 0xffffffffull);
            int64_t P_next = (int64_t)(total >> 32);
            grow_entries(entry_count + n_leaf_entries);
            DevBuf npa, nps, npb, npn, npq;
            npa.alloc(sizeof(int32_t) * std::max<int64_t>(P_next, 1));
            nps.alloc(sizeof(int32_t) * std::max<int64_t>(P_next, 1));
            npb.alloc(sizeof(int64_t) * std::max<uint64_t>(A_next, 1));
            npn.alloc(sizeof(int32_t) * std::max<uint64_t>(A_next, 1));
            npq.alloc(sizeof(int32_t) * std::max<uint64_t>(A_next, 1) * N);
            op.entries = entries.as<int32_t>();
            op.nodes = d_nodes.as<int2>();
            op.next_pair_a = npa.as<int32_t>();
            op.next_pair_s = nps.as<int32_t>();
            op.next_pair_begin = npb.as<int64_t>();
            op.next_parent_node = npn.as<int32_t>();
            op.next_parent_q = npq.as<int32_t>();
            hipLaunchKernelGGL(k_tile_downsweep<ScatterOp<N>>, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, s, op, M,
                               offs.as<uint64_t>());
            NSP_HIP(hipGetLastError());
            NSP_HIP(hipStreamSynchronize(s));
            entry_count += n_leaf_entries;
            node_count += C;
            ++levels;
            ++level;
            A = (int)A_next;
            P = P_next;
            std::swap(pair_a.p, npa.p);
            std::swap(pair_a.bytes, npa.bytes);
            std::swap(pair_s.p, nps.p);
            std::swap(pair_s.bytes, nps.bytes);
            std::swap(pair_begin.p, npb.p);
            std::swap(pair_begin.bytes, npb.bytes);
            std::swap(parent_node.p, npn.p);
            std::swap(parent_node.bytes, npn.bytes);
            std::swap(parent_q.p, npq.p);
            std::swap(parent_q.bytes, npq.bytes);
            if ((double)node_count > 2.0e9) raise(NSP_E_INVALID, "tree too large");
        }
    }
    NSP_HIP(hipStreamSynchronize(s));
    t.levels = levels;
    t.node_count = node_count;
    t.entry_count = entry_count;
    // move into the tree
    std::swap(t.nodes.p, d_nodes.p);
    std::swap(t.nodes.bytes, d_nodes.bytes);
    std::swap(t.node_level.p, d_level.p);
    std::swap(t.node_level.bytes, d_level.bytes);
    std::swap(t.node_q.p, d_q.p);
    std::swap(t.node_q.bytes, d_q.bytes);
    std::swap(t.entries.p, entries.p);
    std::swap(t.entries.bytes, entries.bytes);
    // leaf count (host pass over node records)
    std::vector<int2> hn(node_count);
    NSP_HIP(hipMemcpy(hn.data(), t.nodes.p, sizeof(int2) * node_count, hipMemcpyDeviceToHost));
    int64_t leaves = 0;
    for (const int2& nd : hn)
        if (nd.y >= 0) ++leaves;
    t.leaf_count = leaves;
    // the leaf-ordered SoA copy of the records is built on demand (build_soa_device)
    t.entry_pad = (entry_count + 63) / 64 * 64;
    if (t.entry_pad == 0) t.entry_pad = 64;
    NSP_HIP(hipStreamSynchronize(s));
}

template <int N>
static void build_soa_t(MeshDev& m, hipStream_t s)
{
    TreeDev& t = m.tree;
    t.soa.alloc(sizeof(float) * t.entry_pad * soa_components(N));
    if (t.entry_count > 0) {
        hipLaunchKernelGGL(k_pack_soa<N>, dim3(blocks_for(t.entry_count)), dim3(256), 0, s, t.entry_count,
                           t.entry_pad, t.entries.as<int32_t>(), m.rec64.as<double>(), t.soa.as<float>());
        NSP_HIP(hipGetLastError());
    }
    NSP_HIP(hipStreamSynchronize(s));
}

void build_soa_device(MeshDev& m, hipStream_t s)
{
    if (m.tree.soa.p) return;
    switch (m.n) {
    case 3: build_soa_t<3>(m, s); break;
    case 4: build_soa_t<4>(m, s); break;
    case 5: build_soa_t<5>(m, s); break;
    case 6: build_soa_t<6>(m, s); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
}

template <int N>
static void build_t(MeshDev& m, int max_depth, int min_objects, hipStream_t s)
{
    auto t0 = std::chrono::steady_clock::now();
    build_basis_t<N>(m, s);
    auto t1 = std::chrono::steady_clock::now();
    build_tree_t<N>(m, max_depth, min_objects, s);
    auto t2 = std::chrono::steady_clock::now();
    m.build64.release();
    m.basis_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    m.build_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
}

void build_mesh_device(MeshDev& m, int max_depth, int min_objects, hipStream_t s)
{
    switch (m.n) {
    case 3: build_t<3>(m, max_depth, min_objects, s); break;
    case 4: build_t<4>(m, max_depth, min_objects, s); break;
    case 5: build_t<5>(m, max_depth, min_objects, s); break;
    case 6: build_t<6>(m, max_depth, min_objects, s); break;
    default: raise(NSP_E_INVALID, "dimension not supported");
    }
}

}  // namespace nsp
