// Centerpoint linking and the centerline graph (runNewCenterlineAlg,
// parallelCenterlineExtraction.cpp:912-1160, 1270-1316; kernels
// linkLengths, compact, linkCenterpoints, graphComponentLabeling,
// calculateTreeLength, removeSmallTrees, removeDuplicateEdges,
// kernels_no_3d_write.cl:727-968, and the 2D pyramid compaction of the edge
// image, 349-390 and 624-725).
//
// The reference materialises a count x count image of pair lengths, a
// compacted count x count image of (length, neighbour) per vertex, and a
// count x count edge image that a 2D histogram pyramid turns into the edge
// list.  Here:
//   k_link_cand    per vertex, the neighbours at 0 < |xb - xa| < maxDistance
//                  in ascending index (one row of count ints per vertex)
//   k_link_pairs   per vertex, the reference's search for the best pair of
//                  neighbours (angle, mean TDF along both segments)
//   k_link_keys    the pairs as edges, with the duplicate rule, each keyed by
//                  its position in the 2D pyramid's traversal: the descent
//                  visits the children of a 2x2 cell in the order (0,0),
//                  (0,1), (1,0), (1,1) (cubeOffsets2D), so edge (x, y) comes
//                  at key sum_b (2 x_b + y_b) 4^b
//   k_link_rank    each edge's rank among the keys (the list order)
// A vertex links at most two neighbours, so there are at most 2 * count
// edges and no count x count edge image is needed.
//   k_components   one workgroup: labels by minimum propagation along the
//                  edges with pointer jumping until no edge joins two labels
//                  (the label of a component is then its smallest vertex
//                  index, as the reference's iteration converges to), then
//                  the component sizes
//   k_rasterize    removeSmallTrees: the voxels of the edges of large trees
#include "tsf_common.hpp"

namespace tsf {

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 vert(const int32_t* v, int i) {
    return F3{(float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]};
}
// OpenCL distance(a, b) = length(a - b): sqrt of the sum of squares in
// x, y, z order (exact for the integer positions of a volume below 2^12).
__device__ __forceinline__ float dist3(F3 a, F3 b) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// CLK_FILTER_LINEAR with unnormalised coordinates and clamp-to-edge on a
// float volume: the sample at (u, v, w) - 0.5, the eight terms in the OpenCL
// specification's order (as the TDF's sampler).
__device__ __forceinline__ float linear_sample(const float* __restrict__ T, const Dims& d, float u, float v, float w) {
    const float uu = u - 0.5f, vv = v - 0.5f, ww = w - 0.5f;
    const float fu = floorf(uu), fv = floorf(vv), fw = floorf(ww);
    const float a = uu - fu, b = vv - fv, c = ww - fw;
    int i0 = (int)fu, j0 = (int)fv, k0 = (int)fw;
    const int i1 = clampi(i0 + 1, 0, d.X - 1), j1 = clampi(j0 + 1, 0, d.Y - 1), k1 = clampi(k0 + 1, 0, d.Z - 1);
    i0 = clampi(i0, 0, d.X - 1);
    j0 = clampi(j0, 0, d.Y - 1);
    k0 = clampi(k0, 0, d.Z - 1);
    const float t0 = T[d.at(i0, j0, k0)], t1 = T[d.at(i1, j0, k0)], t2 = T[d.at(i0, j1, k0)];
    const float t3 = T[d.at(i1, j1, k0)], t4 = T[d.at(i0, j0, k1)], t5 = T[d.at(i1, j0, k1)];
    const float t6 = T[d.at(i0, j1, k1)], t7 = T[d.at(i1, j1, k1)];
    float s = (1.0f - a) * (1.0f - b) * (1.0f - c) * t0;
    s = s + a * (1.0f - b) * (1.0f - c) * t1;
    s = s + (1.0f - a) * b * (1.0f - c) * t2;
    s = s + a * b * (1.0f - c) * t3;
    s = s + (1.0f - a) * (1.0f - b) * c * t4;
    s = s + a * (1.0f - b) * c * t5;
    s = s + (1.0f - a) * b * c * t6;
    s = s + a * b * c * t7;
    return s;
}

// Mean of the TDF at xa + ab * k / n, k = 0..n (linkCenterpoints).
__device__ __forceinline__ float segment_mean_tdf(const float* __restrict__ T, const Dims& d, F3 xa, F3 ab, int n) {
    float s = 0.0f;
    for (int k = 0; k <= n; k++) {
        const float alpha = (float)k / (float)n;
        s += linear_sample(T, d, xa.x + ab.x * alpha, xa.y + ab.y * alpha, xa.z + ab.z * alpha);
    }
    return s / (float)(n + 1);
}

// linkLengths + compact: neighbours b of a with 0 < |xb - xa| < maxDistance,
// in ascending b (the reference's order is that of its atomic increments).
__global__ void k_link_cand(const int32_t* __restrict__ v, int count, float maxd, int32_t* __restrict__ cand,
                            int32_t* __restrict__ ncand) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    const F3 xa = vert(v, a);
    int32_t* row = cand + (i64)a * count;
    int n = 0;
    for (int b = 0; b < count; b++) {
        const float l = dist3(xa, vert(v, b));
        if (l < maxd && l > 0.0f) row[n++] = b;
    }
    ncand[a] = n;
}

// linkCenterpoints (kernels_no_3d_write.cl:753-882): among the pairs (b, c)
// of candidates, c listed before b, the one of smallest round|ab| +
// round|ac| (the first found on ties) whose angle at a exceeds 2 rad and
// whose segments ab and ac both have a mean TDF >= minAvgTDF.
__global__ void k_link_pairs(const int32_t* __restrict__ v, int count, const int32_t* __restrict__ cand,
                             const int32_t* __restrict__ ncand, const float* __restrict__ T, Dims d, float maxd,
                             float min_avg, int32_t* __restrict__ pair) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= count) return;
    const F3 xa = vert(v, a);
    const int32_t* row = cand + (i64)a * count;
    const int n = ncand[a];
    float shortest = maxd * 2.0f;
    int bx = -1, by = -1;
    for (int i = 0; i < n; i++) {
        const int b = row[i];
        const F3 xb = vert(v, b);
        const int db = (int)roundf(dist3(xa, xb));
        if ((float)db >= shortest) continue;
        for (int j = 0; j < i; j++) {
            const int c = row[j];
            if (c == b) continue;
            const F3 xc = vert(v, c);
            const int dc = (int)roundf(dist3(xa, xc));
            if (!((float)(db + dc) < shortest)) continue;
            const F3 ab{xb.x - xa.x, xb.y - xa.y, xb.z - xa.z}, ac{xc.x - xa.x, xc.y - xa.y, xc.z - xa.z};
            const float lb = sqrtf(ab.x * ab.x + ab.y * ab.y + ab.z * ab.z);
            const float lc = sqrtf(ac.x * ac.x + ac.y * ac.y + ac.z * ac.z);
            const float dot = (ab.x / lb) * (ac.x / lc) + (ab.y / lb) * (ac.y / lc) + (ab.z / lb) * (ac.z / lc);
            if (acosf(dot) < 2.0f) continue;
            if (segment_mean_tdf(T, d, xa, ab, db) < min_avg) continue;
            if (segment_mean_tdf(T, d, xa, ac, dc) < min_avg) continue;
            bx = b;
            by = c;
            shortest = (float)(db + dc);
        }
    }
    pair[2 * a] = bx;
    pair[2 * a + 1] = by;
}

// 2D pyramid traversal key of cell (x, y) < 2^16: sum_b (2 x_b + y_b) 4^b.
__host__ __device__ __forceinline__ uint64_t key2d(uint32_t x, uint32_t y) {
    uint64_t k = 0;
    for (int b = 0; b < 16; b++) k |= (uint64_t)(((x >> b) & 1u) * 2u + ((y >> b) & 1u)) << (2 * b);
    return k;
}
constexpr uint64_t NO_EDGE = ~(uint64_t)0;

// Edge slots 2a, 2a+1: (a, pair[2a]) and (a, pair[2a+1]); removeDuplicateEdges
// drops (x, y), x > y, when (y, x) is an edge too.
__global__ void k_link_keys(const int32_t* __restrict__ pair, int count, uint64_t* __restrict__ keys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= 2 * count) return;
    const int x = s >> 1, y = pair[s];
    uint64_t k = NO_EDGE;
    if (y >= 0) {
        const bool opposite = pair[2 * y] == x || pair[2 * y + 1] == x;
        if (!(x > y && opposite)) k = key2d((uint32_t)x, (uint32_t)y);
    }
    keys[s] = k;
}

// Number of edges (one block).
__global__ void __launch_bounds__(1024) k_link_count(const uint64_t* __restrict__ keys, int n, int32_t* out) {
    __shared__ int part[16];
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 1024) c += keys[i] != NO_EDGE;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; w++) t += part[w];
        *out = t;
    }
}

// The list position of every edge: the number of smaller keys (keys are
// distinct), through LDS tiles of 1024 keys.
__global__ void __launch_bounds__(256) k_link_rank(const uint64_t* __restrict__ keys, const int32_t* __restrict__ pair,
                                                   int n, int32_t* __restrict__ edges, int capacity) {
    __shared__ uint64_t tile[1024];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t k = s < n ? keys[s] : NO_EDGE;
    int rank = 0;
    for (int t0 = 0; t0 < n; t0 += 1024) {
        __syncthreads();
        for (int i = threadIdx.x; i < 1024; i += 256) tile[i] = t0 + i < n ? keys[t0 + i] : NO_EDGE;
        __syncthreads();
        const int m = min(1024, n - t0);
        for (int i = 0; i < m; i++) rank += tile[i] < k;
    }
    if (k != NO_EDGE && rank < capacity) {
        edges[2 * rank] = s >> 1;
        edges[2 * rank + 1] = pair[s];
    }
}

// graphComponentLabeling + calculateTreeLength in one workgroup.  Labels
// only decrease and stay inside their component, so the iteration ends
// (at most `count` passes: the smallest index of a component spreads one
// edge per pass at least) with every vertex labelled by its component's
// smallest index; *status = 0, or 1 if the pass bound was reached first
// (not expected).
__global__ void __launch_bounds__(1024) k_components(const int32_t* __restrict__ edges, int n_edges, int count,
                                                     int32_t* __restrict__ L, int32_t* __restrict__ S,
                                                     int32_t* __restrict__ status) {
    __shared__ int changed;
    for (int i = threadIdx.x; i < count; i += 1024) {
        L[i] = i;
        S[i] = 0;
    }
    __syncthreads();
    int pass = 0;
    for (;; pass++) {
        if (threadIdx.x == 0) changed = 0;
        __syncthreads();
        int any = 0;
        for (int e = threadIdx.x; e < n_edges; e += 1024) {
            const int x = edges[2 * e], y = edges[2 * e + 1];
            const int a = __hip_atomic_load(&L[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int b = __hip_atomic_load(&L[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (a != b) {
                const int m = min(a, b);
                atomicMin(&L[x], m);
                atomicMin(&L[y], m);
                any = 1;
            }
        }
        if (any) changed = 1;
        __syncthreads();
        // pointer jumping: L[v] = L[L[v]] (an index of the same component, not larger)
        for (int i = threadIdx.x; i < count; i += 1024) {
            const int l = __hip_atomic_load(&L[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int ll = __hip_atomic_load(&L[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (ll < l) atomicMin(&L[i], ll);
        }
        __syncthreads();
        const int c = changed;
        __syncthreads();
        if (!c || pass >= count) break;
    }
    for (int i = threadIdx.x; i < count; i += 1024) atomicAdd(&S[L[i]], 1);
    if (threadIdx.x == 0) *status = changed ? 1 : 0;
}

// removeSmallTrees (kernels_no_3d_write.cl:920-945): for each edge whose
// first vertex's tree has at least minTreeLength vertices, the voxels
// round(xa + (xb - xa) i / l), i = 0..l-1, l = round|xb - xa|, are set.
__global__ void k_rasterize(const int32_t* __restrict__ edges, int n_edges, const int32_t* __restrict__ v,
                            const int32_t* __restrict__ L, const int32_t* __restrict__ S, int min_tree,
                            uint8_t* __restrict__ out, Dims d) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges) return;
    const int x = edges[2 * e], y = edges[2 * e + 1];
    if (S[L[x]] < min_tree) return;
    const F3 xa = vert(v, x), xb = vert(v, y);
    const F3 ab{xb.x - xa.x, xb.y - xa.y, xb.z - xa.z};
    const int l = (int)roundf(sqrtf(ab.x * ab.x + ab.y * ab.y + ab.z * ab.z));
    for (int i = 0; i < l; i++) {
        const float alpha = (float)i / (float)l;
        const int px = (int)roundf(xa.x + ab.x * alpha), py = (int)roundf(xa.y + ab.y * alpha);
        const int pz = (int)roundf(xa.z + ab.z * alpha);
        if (px >= 0 && py >= 0 && pz >= 0 && px < d.X && py < d.Y && pz < d.Z) out[d.at(px, py, pz)] = 1;
    }
}

static unsigned blocks(i64 n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_link_centerpoints_workspace_size(int32_t count) {
    if (count < 1) return 0;
    const size_t c = (size_t)count;
    size_t a = (c * c * sizeof(int32_t) + 255) & ~(size_t)255;  // candidate rows
    a += (c * sizeof(int32_t) + 255) & ~(size_t)255;            // candidate counts
    a += (2 * c * sizeof(int32_t) + 255) & ~(size_t)255;        // pairs
    a += (2 * c * sizeof(uint64_t) + 255) & ~(size_t)255;       // edge keys
    return a + 256;                                             // edge count
}

tsf_status tsf_link_centerpoints(const int32_t* vertices, int32_t count, const float* tdf, tsf_dims dims,
                                 float max_distance, float min_mean_tdf, int32_t* edges, int32_t capacity,
                                 int32_t* n_edges, void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_link_centerpoints";
    if (!vertices || !tdf || !n_edges) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (count < 1 || count > 65536) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: count must be in 1..65536", W);
    if (capacity < 0 || (capacity > 0 && !edges)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: bad edge list", W);
    if (!(max_distance > 0.0f) || !__builtin_isfinite(max_distance))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: max_distance must be finite and > 0", W);
    TSF_TRY(check_dims(dims, 1, W));
    Dims d = D(dims);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_link_centerpoints_workspace_size(count), W));
    const size_t c = (size_t)count;
    char* p = (char*)sc.ptr;
    int32_t* cand = (int32_t*)p;
    p += (c * c * sizeof(int32_t) + 255) & ~(size_t)255;
    int32_t* ncand = (int32_t*)p;
    p += (c * sizeof(int32_t) + 255) & ~(size_t)255;
    int32_t* pair = (int32_t*)p;
    p += (2 * c * sizeof(int32_t) + 255) & ~(size_t)255;
    uint64_t* keys = (uint64_t*)p;
    p += (2 * c * sizeof(uint64_t) + 255) & ~(size_t)255;
    int32_t* dcount = (int32_t*)p;
    hipLaunchKernelGGL(k_link_cand, blocks(count, 64), 64, 0, stream, vertices, count, max_distance, cand, ncand);
    hipLaunchKernelGGL(k_link_pairs, blocks(count, 64), 64, 0, stream, vertices, count, cand, ncand, tdf, d,
                       max_distance, min_mean_tdf, pair);
    hipLaunchKernelGGL(k_link_keys, blocks(2 * (i64)count, 256), 256, 0, stream, pair, count, keys);
    hipLaunchKernelGGL(k_link_count, 1, 1024, 0, stream, keys, 2 * count, dcount);
    if (capacity > 0)
        hipLaunchKernelGGL(k_link_rank, blocks(2 * (i64)count, 256), 256, 0, stream, keys, pair, 2 * count, edges,
                           capacity);
    TSF_TRY(check_launch(W));
    int32_t total = 0;
    hipError_t e = hipMemcpyAsync(&total, dcount, sizeof total, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    *n_edges = total;
    if (total > capacity && edges != nullptr)
        return set_error(TSF_ERR_CAPACITY, "%s: %d edges, capacity %d", W, total, capacity);
    return TSF_OK;
}

tsf_status tsf_graph_components(const int32_t* edges, int32_t n_edges, int32_t count, int32_t* labels,
                                int32_t* sizes, hipStream_t stream) {
    const char* W = "tsf_graph_components";
    if (!labels || !sizes || (n_edges > 0 && !edges)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (count < 1 || n_edges < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: bad count", W);
    Scratch sc;
    TSF_TRY(sc.get(nullptr, 0, 256, W));
    int32_t* status = (int32_t*)sc.ptr;
    hipLaunchKernelGGL(k_components, 1, 1024, 0, stream, edges, n_edges, count, labels, sizes, status);
    TSF_TRY(check_launch(W));
    int32_t h = 0;
    hipError_t e = hipMemcpyAsync(&h, status, sizeof h, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    if (h != 0) return set_error(TSF_ERR_RANGE, "%s: labels did not settle within the pass bound", W);
    return TSF_OK;
}

tsf_status tsf_rasterize_tree_edges(const int32_t* edges, int32_t n_edges, const int32_t* vertices,
                                    const int32_t* labels, const int32_t* sizes, int32_t min_tree_length,
                                    uint8_t* centerlines, tsf_dims dims, hipStream_t stream) {
    const char* W = "tsf_rasterize_tree_edges";
    if (!vertices || !labels || !sizes || !centerlines || (n_edges > 0 && !edges))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (n_edges < 0) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: bad edge count", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (n_edges == 0) return TSF_OK;
    hipLaunchKernelGGL(k_rasterize, blocks(n_edges, 256), 256, 0, stream, edges, n_edges, vertices, labels, sizes,
                       min_tree_length, centerlines, D(dims));
    return check_launch(W);
}

}  // extern "C"
