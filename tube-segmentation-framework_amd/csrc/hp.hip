// Histogram-pyramid compaction of a voxel mask.
//
// The reference builds a full pyramid (log2(HP_SIZE) levels of 2x2x2 sums)
// and lets every output thread descend it (kernels_no_3d_write.cl:246-698).
// The descent visits children in the order k = x + 2z + 4y at every level,
// so the output list is the set voxels sorted by the key
//     key(p) = sum_b (x_b + 2 z_b + 4 y_b) * 8^b.
// An aligned 16^3 cube is a contiguous key range of 4096 keys, and the cubes
// themselves are ordered by the same key of their cube coordinates.  This
// file therefore keeps only the top of the pyramid explicitly (one count per
// 16^3 cube, in cube-key order, then an exclusive scan = the pyramid's sums
// read in traversal order) and resolves the bottom four levels inside each
// cube with a 256-thread wavefront scan over the cube's keys.  Three
// launches: count (one byte read per voxel), scan, scatter (one byte read per
// voxel of non-empty cubes + 12 bytes per output entry).
#include "tsf_common.hpp"

namespace tsf {

constexpr int HB = 16;           // cube edge
constexpr int HB3 = HB * HB * HB;
constexpr int HT = 256;          // threads per cube
constexpr int KPT = HB3 / HT;    // keys per thread (16)

__device__ __forceinline__ void decode_key(uint64_t k, int levels, int& x, int& y, int& z) {
    x = y = z = 0;
    for (int b = 0; b < levels; b++) {
        int dgt = (int)((k >> (3 * b)) & 7);
        x |= (dgt & 1) << b;
        z |= ((dgt >> 1) & 1) << b;
        y |= ((dgt >> 2) & 1) << b;
    }
}

// inclusive wavefront scan (64 lanes)
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// exclusive scan over the 256 threads of a block; returns the block total in *tot
__device__ __forceinline__ int block_excl_scan(int v, int* tot) {
    __shared__ int wsum[HT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < HT / 64; w++) {
        if (w < wid) pre += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *tot = all;
    return pre + inc - v;
}

__global__ void __launch_bounds__(HT) k_hp_count(const uint8_t* __restrict__ mask, Dims d, int cube_levels,
                                                 int* __restrict__ counts) {
    const uint64_t bkey = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
    int bx, by, bz;
    decode_key(bkey, cube_levels, bx, by, bz);
    const int x0 = bx * HB, y0 = by * HB, z0 = bz * HB;
    int c = 0;
    if (x0 < d.X && y0 < d.Y && z0 < d.Z) {
        const int ly = threadIdx.x & 15, lz = threadIdx.x >> 4;
        const int y = y0 + ly, z = z0 + lz;
        if (y < d.Y && z < d.Z) {
            const uint8_t* row = mask + d.at(x0, y, z);
            const int nx = min(HB, d.X - x0);
            for (int i = 0; i < nx; i++) c += row[i] != 0;
        }
    }
    int tot;
    (void)block_excl_scan(c, &tot);
    if (threadIdx.x == 0) counts[bkey] = tot;
}

// Exclusive scan of the cube counts (one workgroup), total in offsets[n].
__global__ void __launch_bounds__(1024) k_hp_scan(const int* __restrict__ counts, long long* __restrict__ offsets,
                                                  long long n) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const long long per = (n + 1023) / 1024;
    const long long b = t * per, e = min(n, b + per);
    long long s = 0;
    for (long long i = b; i < e; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    // Hillis-Steele over 1024 partial sums
    for (int o = 1; o < 1024; o <<= 1) {
        long long v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - s;
    for (long long i = b; i < e; i++) {
        offsets[i] = run;
        run += counts[i];
    }
    if (t == 1023) offsets[n] = part[1023];
}

__global__ void __launch_bounds__(HT) k_hp_scatter(const uint8_t* __restrict__ mask, Dims d, int cube_levels,
                                                   const int* __restrict__ counts,
                                                   const long long* __restrict__ offsets, int32_t* __restrict__ pos,
                                                   long long capacity) {
    const uint64_t bkey = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
    if (counts[bkey] == 0) return;  // uniform per block
    __shared__ uint8_t cube[HB3];
    int bx, by, bz;
    decode_key(bkey, cube_levels, bx, by, bz);
    const int x0 = bx * HB, y0 = by * HB, z0 = bz * HB;
    {
        const int ly = threadIdx.x & 15, lz = threadIdx.x >> 4;
        const int y = y0 + ly, z = z0 + lz;
        const int nx = (y < d.Y && z < d.Z) ? min(HB, d.X - x0) : 0;
        const uint8_t* row = mask + (nx > 0 ? d.at(x0, y, z) : 0);
        for (int i = 0; i < HB; i++) cube[(lz * HB + ly) * HB + i] = (i < nx && row[i] != 0) ? 1 : 0;
    }
    __syncthreads();
    // thread t owns local keys [KPT*t, KPT*t + KPT)
    int lx[KPT], ly[KPT], lz[KPT];
    unsigned bits = 0;
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        decode_key((uint64_t)(KPT * threadIdx.x + j), 4, lx[j], ly[j], lz[j]);
        if (cube[(lz[j] * HB + ly[j]) * HB + lx[j]]) bits |= 1u << j;
    }
    int tot;
    const int pre = block_excl_scan(__builtin_popcount(bits), &tot);
    long long o = offsets[bkey] + pre;
#pragma unroll
    for (int j = 0; j < KPT; j++) {
        if (bits & (1u << j)) {
            if (o < capacity) {
                pos[3 * o] = x0 + lx[j];
                pos[3 * o + 1] = y0 + ly[j];
                pos[3 * o + 2] = z0 + lz[j];
            }
            o++;
        }
    }
}

struct HPPlan {
    int cube_levels;   // levels of the cube-key (HP_SIZE / 16 = 2^cube_levels)
    long long ncubes;  // 8^cube_levels
};

static HPPlan plan(Dims d) {
    int m = d.X > d.Y ? d.X : d.Y;
    m = m > d.Z ? m : d.Z;
    int size = HB, L = 0;
    while (size < m) {
        size *= 2;
        L++;
    }
    return HPPlan{L, 1LL << (3 * L)};
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_hp_compact_workspace_size(tsf_dims dims) {
    HPPlan p = plan(D(dims));
    size_t a = (size_t)p.ncubes * sizeof(int);
    a = (a + 255) & ~(size_t)255;
    return a + (size_t)(p.ncubes + 1) * sizeof(long long);
}

tsf_status tsf_hp_compact(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity, int64_t* count,
                          void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_hp_compact";
    if (!mask || !count) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (capacity < 0 || (capacity > 0 && !positions))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: bad output list", W);
    TSF_TRY(check_dims(dims, 1, W));
    Dims d = D(dims);
    HPPlan p = plan(d);
    if (p.ncubes > (1LL << 30)) return set_error(TSF_ERR_DIMENSIONS, "%s: pyramid too large", W);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_hp_compact_workspace_size(dims), W));
    int* counts = (int*)sc.ptr;
    size_t a = ((size_t)p.ncubes * sizeof(int) + 255) & ~(size_t)255;
    long long* offsets = (long long*)((char*)sc.ptr + a);
    // cubes in key order; split over a 2D grid so gridDim.x stays small
    unsigned gx = (unsigned)(p.ncubes < 65536 ? p.ncubes : 65536);
    unsigned gy = (unsigned)(p.ncubes / gx);
    hipLaunchKernelGGL(k_hp_count, dim3(gx, gy), HT, 0, stream, mask, d, p.cube_levels, counts);
    TSF_TRY(check_launch(W));
    hipLaunchKernelGGL(k_hp_scan, 1, 1024, 0, stream, counts, offsets, p.ncubes);
    TSF_TRY(check_launch(W));
    if (capacity > 0) {
        hipLaunchKernelGGL(k_hp_scatter, dim3(gx, gy), HT, 0, stream, mask, d, p.cube_levels, counts, offsets,
                           positions, (long long)capacity);
        TSF_TRY(check_launch(W));
    }
    long long total = 0;
    hipError_t e = hipMemcpyAsync(&total, offsets + p.ncubes, sizeof total, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    *count = total;
    if (total > capacity && positions != nullptr)  // positions == NULL: count-only query
        return set_error(TSF_ERR_CAPACITY, "%s: %lld set voxels, capacity %lld", W, total, (long long)capacity);
    return TSF_OK;
}

}  // extern "C"
