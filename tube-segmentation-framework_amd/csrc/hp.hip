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
// launches: count (one byte read per voxel, each workgroup reading a strip
// of cubes along x so that the reads are whole rows), scan, scatter (one
// byte read per voxel of non-empty cubes + 12 bytes per output entry).
#include "hp.hpp"

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

__device__ __forceinline__ uint64_t encode_key(int x, int y, int z, int levels) {
    uint64_t k = 0;
    for (int b = 0; b < levels; b++)
        k |= (uint64_t)(((x >> b) & 1) + 2 * ((z >> b) & 1) + 4 * ((y >> b) & 1)) << (3 * b);
    return k;
}

// Number of nonzero bytes of a 32-bit word: bit 7 of each byte of
// ((b & 0x7f) + 0x7f) | b is set exactly when b != 0 (no carry crosses bytes).
__device__ __forceinline__ int nz_bytes(unsigned w) {
    return __builtin_popcount((((w & 0x7f7f7f7fu) + 0x7f7f7f7fu) | w) & 0x80808080u);
}

// Cube counts for one strip of cubes (by, bz) = (blockIdx.x, blockIdx.y):
// the strip's 16 x 16 rows are read along x in 16-byte items, one item per
// cube and row, so consecutive threads read consecutive 16-byte segments of
// a row.  A thread keeps one cube column (bx) across the rows it visits and
// adds its count to the strip's LDS counter once.  Every cube of the
// pyramid's domain is written (zero outside the volume), in key order.
template <bool ALIGNED>
__global__ void __launch_bounds__(HT) k_hp_count_strip(const uint8_t* __restrict__ mask, Dims d, int cube_levels,
                                                       int* __restrict__ counts) {
    __shared__ int cnt[1024];  // ncx = 2^cube_levels <= 1024 (the plan caps the pyramid at 2^30 cubes)
    const int ncx = 1 << cube_levels;
    const int by = blockIdx.x, bz = blockIdx.y;
    const int y0 = by * HB, z0 = bz * HB;
    const int tid = threadIdx.x;
    for (int i = tid; i < ncx; i += HT) cnt[i] = 0;
    __syncthreads();
    const bool strip_in = y0 < d.Y && z0 < d.Z;
    const int nbx = (d.X + HB - 1) / HB;  // cubes along x inside the volume
    auto item = [&](int r, int bx) -> int {
        const int y = y0 + (r & 15), z = z0 + (r >> 4);
        if (y >= d.Y || z >= d.Z) return 0;
        const uint8_t* p = mask + d.at(bx * HB, y, z);
        if (ALIGNED) {
            const uint4 w = *reinterpret_cast<const uint4*>(p);
            return nz_bytes(w.x) + nz_bytes(w.y) + nz_bytes(w.z) + nz_bytes(w.w);
        }
        const int nx = min(HB, d.X - bx * HB);
        int c = 0;
        for (int i = 0; i < nx; i++) c += p[i] != 0;
        return c;
    };
    if (strip_in) {
        if (nbx <= HT) {
            const int rpp = HT / nbx;  // rows per pass
            const int r0 = tid / nbx, bx = tid - r0 * nbx;
            if (r0 < rpp) {
                int c = 0;
                for (int r = r0; r < HB * HB; r += rpp) c += item(r, bx);
                if (c) atomicAdd(&cnt[bx], c);
            }
        } else {
            for (int i = tid; i < HB * HB * nbx; i += HT) {
                const int r = i / nbx, bx = i - r * nbx;
                const int c = item(r, bx);
                if (c) atomicAdd(&cnt[bx], c);
            }
        }
    }
    __syncthreads();
    for (int bx = tid; bx < ncx; bx += HT) counts[encode_key(bx, by, bz, cube_levels)] = cnt[bx];
}

// Exclusive scan of the cube counts (one workgroup), total in offsets[n]:
// tiles of 4096 counts, four consecutive ones per thread (coalesced), a
// wavefront scan of the thread sums and a scan of the 16 wave totals; the
// running offset is 64-bit (a tile's total is below 2^24).
__global__ void __launch_bounds__(1024) k_hp_scan(const int* __restrict__ counts, long long* __restrict__ offsets,
                                                  long long n) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    long long carry = 0;
    for (long long base = 0; base < n; base += 4096) {
        const long long i0 = base + 4 * t;
        int v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = i0 + k < n ? counts[i0 + k] : 0;
        const int s = v[0] + v[1] + v[2] + v[3];
        const int inc = wave_incl_scan(s);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        int pre = 0, tile = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int c = wsum[k];
            pre += k < w ? c : 0;
            tile += c;
        }
        long long run = carry + pre + inc - s;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (i0 + k < n) offsets[i0 + k] = run;
            run += v[k];
        }
        carry += tile;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (t == 0) offsets[n] = carry;
}

template <bool ALIGNED>
__global__ void __launch_bounds__(HT) k_hp_scatter(const uint8_t* __restrict__ mask, Dims d, int cube_levels,
                                                   const int* __restrict__ counts,
                                                   const long long* __restrict__ offsets, int32_t* __restrict__ pos,
                                                   long long capacity) {
    const uint64_t bkey = blockIdx.x + (uint64_t)blockIdx.y * gridDim.x;
    if (counts[bkey] == 0) return;  // uniform per block
    __shared__ __align__(16) uint8_t cube[HB3];
    int bx, by, bz;
    decode_key(bkey, cube_levels, bx, by, bz);
    const int x0 = bx * HB, y0 = by * HB, z0 = bz * HB;
    {
        const int ly = threadIdx.x & 15, lz = threadIdx.x >> 4;
        const int y = y0 + ly, z = z0 + lz;
        const int nx = (y < d.Y && z < d.Z) ? min(HB, d.X - x0) : 0;
        const uint8_t* row = mask + (nx > 0 ? d.at(x0, y, z) : 0);
        if (ALIGNED) {  // X % 16 == 0: whole 16-byte rows
            *reinterpret_cast<uint4*>(&cube[(lz * HB + ly) * HB]) =
                nx > 0 ? *reinterpret_cast<const uint4*>(row) : make_uint4(0, 0, 0, 0);
        } else {
            for (int i = 0; i < HB; i++) cube[(lz * HB + ly) * HB + i] = (i < nx && row[i] != 0) ? 1 : 0;
        }
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

HPPlan hp_plan(Dims d) {
    int m = d.X > d.Y ? d.X : d.Y;
    m = m > d.Z ? m : d.Z;
    int size = HB, L = 0;
    while (size < m) {
        size *= 2;
        L++;
    }
    return HPPlan{L, 1LL << (3 * L)};
}

// The three launches on counts / offsets laid out as the workspace of
// tsf_hp_compact_workspace_size, then the total read back to the host.
tsf_status hp_run(const uint8_t* mask, Dims d, const HPPlan& p, void* ws, int32_t* positions, int64_t capacity,
                  int64_t* count, hipStream_t stream, const char* W) {
    int* counts = (int*)ws;
    size_t a = ((size_t)p.ncubes * sizeof(int) + 255) & ~(size_t)255;
    long long* offsets = (long long*)((char*)ws + a);
    const int ncx = 1 << p.cube_levels;
    const bool aligned = d.X % HB == 0 && (uintptr_t)mask % 16 == 0;
    if (aligned)
        hipLaunchKernelGGL(k_hp_count_strip<true>, dim3(ncx, ncx), HT, 0, stream, mask, d, p.cube_levels, counts);
    else
        hipLaunchKernelGGL(k_hp_count_strip<false>, dim3(ncx, ncx), HT, 0, stream, mask, d, p.cube_levels, counts);
    TSF_TRY(check_launch(W));
    hipLaunchKernelGGL(k_hp_scan, 1, 1024, 0, stream, counts, offsets, p.ncubes);
    TSF_TRY(check_launch(W));
    if (capacity > 0) {
        // cubes in key order; split over a 2D grid so gridDim.x stays small
        unsigned gx = (unsigned)(p.ncubes < 65536 ? p.ncubes : 65536);
        unsigned gy = (unsigned)(p.ncubes / gx);
        if (aligned)
            hipLaunchKernelGGL(k_hp_scatter<true>, dim3(gx, gy), HT, 0, stream, mask, d, p.cube_levels, counts,
                               offsets, positions, (long long)capacity);
        else
            hipLaunchKernelGGL(k_hp_scatter<false>, dim3(gx, gy), HT, 0, stream, mask, d, p.cube_levels, counts,
                               offsets, positions, (long long)capacity);
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

tsf_status hp_check(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity, int64_t* count,
                    const char* W) {
    if (!mask || !count) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (capacity < 0 || (capacity > 0 && !positions))
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: bad output list", W);
    TSF_TRY(check_dims(dims, 1, W));
    if (hp_plan(D(dims)).ncubes > (1LL << 30)) return set_error(TSF_ERR_DIMENSIONS, "%s: pyramid too large", W);
    return TSF_OK;
}

}  // namespace tsf

using namespace tsf;

extern "C" {

size_t tsf_hp_compact_workspace_size(tsf_dims dims) {
    HPPlan p = hp_plan(D(dims));
    size_t a = (size_t)p.ncubes * sizeof(int);
    a = (a + 255) & ~(size_t)255;
    return a + (size_t)(p.ncubes + 1) * sizeof(long long);
}

tsf_status tsf_hp_compact(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity, int64_t* count,
                          void* workspace, size_t workspace_bytes, hipStream_t stream) {
    const char* W = "tsf_hp_compact";
    TSF_TRY(hp_check(mask, dims, positions, capacity, count, W));
    Dims d = D(dims);
    Scratch sc;
    TSF_TRY(sc.get(workspace, workspace_bytes, tsf_hp_compact_workspace_size(dims), W));
    return hp_run(mask, d, hp_plan(d), sc.ptr, positions, capacity, count, stream, W);
}

}  // extern "C"
