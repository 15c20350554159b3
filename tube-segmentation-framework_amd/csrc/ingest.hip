// Volume ingest (readDatasetAndTransfer, tube-segmentation.cpp:1399-1905):
// the intensity limits (getMinimum / getMaximum, 1356-1374), the scan-line
// counts of threshold and lung cropping (cropDatasetThreshold /
// cropDatasetLung, kernels_no_3d_write.cl:1236-1368), the crop region from
// those counts (tube-segmentation.cpp:1612-1745, host code) and toFloat
// fused with the copy of the cropped region.
//
// The reference reads the raw file on the host, finds the limits there,
// uploads the whole volume, counts scan lines with one work-item per slice
// (each walking a whole slice: X, Y or Z work-items in all), copies the
// region to a new image and converts it.  Here the raw volume is read by
// many waves: one kernel walks every z-line (the counts of the x and y
// slices come from the same lines, and for threshold cropping also the
// x-line verdicts, by wave ballots); lung cropping walks every x-line in a
// second kernel through an LDS transpose; the conversion reads the region
// straight from the uploaded volume.
#include <cfloat>
#include <cmath>

#include "tsf_common.hpp"

namespace tsf {

namespace {

// ---------------------------------------------------------------- limits
// std::min / std::max over the values in index order, from the initial
// value numeric_limits<T>::max() / ::min() (for float, ::min() is the
// smallest positive normal, FLT_MIN): the result is the initial value unless
// some element compares below (above) it, else the first element, in index
// order, of the least (greatest) value.  NaN never compares, so never wins;
// of two equal values (+0 and -0) the first stays.  Each candidate carries
// its index so that the parallel reduction keeps the same one.
struct Cand {
    float v;
    long long i;  // -1: none
};
__device__ __forceinline__ Cand pick_min(Cand a, Cand b) {
    if (b.i < 0) return a;
    if (a.i < 0) return b;
    if (b.v < a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}
__device__ __forceinline__ Cand pick_max(Cand a, Cand b) {
    if (b.i < 0) return a;
    if (a.i < 0) return b;
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
    return a;
}
__device__ __forceinline__ Cand shfl_xor_cand(Cand c, int m) {
    return Cand{__shfl_xor(c.v, m), __shfl_xor(c.i, m)};
}

struct Limits {
    Cand mn, mx;
};

constexpr int LIM_THREADS = 256;

__device__ __forceinline__ void block_reduce(Cand& mn, Cand& mx, Limits* out) {
    for (int m = 32; m >= 1; m >>= 1) {
        mn = pick_min(mn, shfl_xor_cand(mn, m));
        mx = pick_max(mx, shfl_xor_cand(mx, m));
    }
    __shared__ Limits part[LIM_THREADS / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) part[wave] = Limits{mn, mx};
    __syncthreads();
    if (threadIdx.x == 0) {
        Limits r = part[0];
        for (int w = 1; w < LIM_THREADS / 64; w++) {
            r.mn = pick_min(r.mn, part[w].mn);
            r.mx = pick_max(r.mx, part[w].mx);
        }
        *out = r;
    }
}

// Each thread visits its elements in increasing index order (a grid-stride
// loop, coalesced across the wave), so a strict comparison keeps its first of
// equal values; the reduction then prefers the lower index on ties.
template <class T>
__global__ void __launch_bounds__(LIM_THREADS) k_limits(const T* __restrict__ in, i64 n, Limits* __restrict__ part) {
    const i64 stride = (i64)gridDim.x * LIM_THREADS;
    Cand mn{0.0f, -1}, mx{0.0f, -1};
    for (i64 i = (i64)blockIdx.x * LIM_THREADS + threadIdx.x; i < n; i += stride) {
        const float v = (float)in[i];
        if (v != v) continue;
        if (mn.i < 0 || v < mn.v) mn = Cand{v, i};
        if (mx.i < 0 || v > mx.v) mx = Cand{v, i};
    }
    block_reduce(mn, mx, part + blockIdx.x);
}

// 16 bytes per thread and iteration (n a multiple of the vector length,
// `in` 16-byte aligned): the elements of each vector in index order.
template <class T>
__global__ void __launch_bounds__(LIM_THREADS) k_limits16(const uint4* __restrict__ in, i64 nvec,
                                                          Limits* __restrict__ part) {
    constexpr int E = 16 / sizeof(T);
    const i64 stride = (i64)gridDim.x * LIM_THREADS;
    Cand mn{0.0f, -1}, mx{0.0f, -1};
    for (i64 c = (i64)blockIdx.x * LIM_THREADS + threadIdx.x; c < nvec; c += stride) {
        const uint4 u = in[c];
        T e[E];
        __builtin_memcpy(e, &u, 16);
#pragma unroll
        for (int k = 0; k < E; k++) {
            const float v = (float)e[k];
            if (v != v) continue;
            const i64 i = c * E + k;
            if (mn.i < 0 || v < mn.v) mn = Cand{v, i};
            if (mx.i < 0 || v > mx.v) mx = Cand{v, i};
        }
    }
    block_reduce(mn, mx, part + blockIdx.x);
}

template <class T>
void launch_limits(const void* in, i64 n, unsigned g, Limits* part, hipStream_t s) {
    constexpr int E = 16 / sizeof(T);
    if (n % E == 0 && (uintptr_t)in % 16 == 0)
        hipLaunchKernelGGL(k_limits16<T>, g, LIM_THREADS, 0, s, (const uint4*)in, n / E, part);
    else
        hipLaunchKernelGGL(k_limits<T>, g, LIM_THREADS, 0, s, (const T*)in, n, part);
}

__global__ void __launch_bounds__(LIM_THREADS) k_limits_final(const Limits* __restrict__ part, int n,
                                                              Limits* __restrict__ out) {
    Cand mn{0.0f, -1}, mx{0.0f, -1};
    for (int i = threadIdx.x; i < n; i += LIM_THREADS) {
        mn = pick_min(mn, part[i].mn);
        mx = pick_max(mx, part[i].mx);
    }
    block_reduce(mn, mx, out);
}

// ---------------------------------------------------------------- scan lines
// One scan line's verdict, fed the line's values in order.
//   threshold: some value > threshold (cropDatasetThreshold: integer reads
//              compared as float, exact for 8- and 16-bit values);
//   lung:      the run-length automaton of cropDatasetLung on HU = (short)
//              value against -150 (+1024 for unsigned types): a white
//              (black) area is counted when a run of 30 white (black) values
//              continues, and the line counts with two white areas and one
//              black, or more than one of each.
struct ThresholdLine {
    bool found = false;
    __device__ __forceinline__ void step(float v, float thr, int) { found = found || v > thr; }
    __device__ __forceinline__ bool inside() const { return found; }
};
struct LungLine {
    int wc = 0, bc = 0, dw = 0, db = 0;
    __device__ __forceinline__ void step(float v, float, int hu_limit) {
        const short hu = (short)(int)v;
        if (hu > hu_limit) {
            if (wc == 30) {
                dw++;
                bc = 0;
            }
            wc++;
        } else {
            if (bc == 30) {
                db++;
                wc = 0;
            }
            bc++;
        }
    }
    __device__ __forceinline__ bool inside() const { return (dw == 2 && db == 1) || (db > 1 && dw > 1); }
};

__global__ void k_zero_i32(int32_t* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// z-lines: one thread per (x, y) walks z, 16 planes of loads in flight.
// lines_x[x] counts the y with an inside z-line (slices of constant x,
// scan lines along y, elements along z); lines_y[y] counts the x (slices of
// constant y, scan lines along x).  With ROWS (threshold cropping, whose
// x-line verdict is an OR along x) each wave also marks row (y, z) in
// `rows` when one of its 64 values passes, so the x-lines need no second
// read of the volume.
constexpr int ZB = 16;
template <class T, class L, bool ROWS>
__global__ void __launch_bounds__(256) k_scan_z(const T* __restrict__ in, Dims d, float thr, int hu_limit,
                                                int32_t* __restrict__ lines_x, int32_t* __restrict__ lines_y,
                                                uint8_t* __restrict__ rows) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    const bool on = x < d.X && y < d.Y;
    L line;
    const i64 plane = (i64)d.X * d.Y;
    const T* p = in + (on ? (i64)y * d.X + x : 0);
    for (int z0 = 0; z0 < d.Z; z0 += ZB) {
        float v[ZB];
#pragma unroll
        for (int j = 0; j < ZB; j++) v[j] = on && z0 + j < d.Z ? (float)p[(i64)(z0 + j) * plane] : 0.0f;
#pragma unroll
        for (int j = 0; j < ZB; j++) {
            if (z0 + j >= d.Z) continue;  // uniform: the planes past the volume's end
            if (on) line.step(v[j], thr, hu_limit);
            if (ROWS) {
                const unsigned long long hit = __ballot(on && v[j] > thr);
                if (threadIdx.x == 0 && hit) rows[(i64)(z0 + j) * d.Y + y] = 1;
            }
        }
    }
    const bool ins = on && line.inside();
    const unsigned long long b = __ballot(ins);
    if (threadIdx.x == 0 && y < d.Y && b) atomicAdd(lines_y + y, (int)__popcll(b));
    __shared__ int cnt[4][64];
    cnt[threadIdx.y][threadIdx.x] = ins ? 1 : 0;
    __syncthreads();
    if (threadIdx.y == 0 && x < d.X) {
        const int s = cnt[0][threadIdx.x] + cnt[1][threadIdx.x] + cnt[2][threadIdx.x] + cnt[3][threadIdx.x];
        if (s) atomicAdd(lines_x + x, s);
    }
}

__global__ void k_zero_u8(uint8_t* __restrict__ a, i64 n) {
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) a[i] = 0;
}

// lines_z[z] = number of marked rows of plane z (one workgroup per plane)
__global__ void __launch_bounds__(256) k_count_rows(const uint8_t* __restrict__ rows, Dims d,
                                                    int32_t* __restrict__ lines_z) {
    const int z = blockIdx.x;
    int c = 0;
    for (int y = threadIdx.x; y < d.Y; y += 256) c += rows[(i64)z * d.Y + y];
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) lines_z[z] = part[0] + part[1] + part[2] + part[3];
}

// x-lines: one wave per 64 rows of one z plane; each 64-column chunk of the
// rows is loaded row by row (coalesced) into an LDS tile, then every lane
// walks its own row through the tile.  lines_z[z] counts the y with an
// inside x-line (slices of constant z, scan lines along y, elements along x).
template <class T, class L>
__global__ void __launch_bounds__(64) k_scan_x(const T* __restrict__ in, Dims d, float thr, int hu_limit,
                                               int32_t* __restrict__ lines_z) {
    __shared__ float tile[64][65];
    const int lane = threadIdx.x, y0 = blockIdx.x * 64, z = blockIdx.y;
    const T* plane = in + (i64)z * d.X * d.Y;
    L line;
    for (int x0 = 0; x0 < d.X; x0 += 64) {
        const int x = x0 + lane;
#pragma unroll 8
        for (int r = 0; r < 64; r++)
            tile[r][lane] = (y0 + r < d.Y && x < d.X) ? (float)plane[(i64)(y0 + r) * d.X + x] : 0.0f;
        __syncthreads();
        const int w = d.X - x0 < 64 ? d.X - x0 : 64;
        for (int c = 0; c < w; c++) line.step(tile[lane][c], thr, hu_limit);
        __syncthreads();
    }
    const bool ins = y0 + lane < d.Y && line.inside();
    const unsigned long long b = __ballot(ins);
    if (lane == 0 && b) atomicAdd(lines_z + z, (int)__popcll(b));
}

template <class T, class L>
void launch_scan(const void* in, Dims d, float thr, int hu, int32_t* lx, int32_t* ly, int32_t* lz, uint8_t* rows,
                 hipStream_t s) {
    const dim3 gz((d.X + 63) / 64, (d.Y + 3) / 4), bz(64, 4);
    if (rows) {
        // threshold: the x-line verdicts come from the z-line pass
        const i64 nr = (i64)d.Y * d.Z;
        hipLaunchKernelGGL(k_zero_u8, (unsigned)((nr + 255) / 256 < 65536 ? (nr + 255) / 256 : 65536), 256, 0, s,
                           rows, nr);
        hipLaunchKernelGGL((k_scan_z<T, L, true>), gz, bz, 0, s, (const T*)in, d, thr, hu, lx, ly, rows);
        hipLaunchKernelGGL(k_count_rows, d.Z, 256, 0, s, (const uint8_t*)rows, d, lz);
        return;
    }
    hipLaunchKernelGGL((k_scan_z<T, L, false>), gz, bz, 0, s, (const T*)in, d, thr, hu, lx, ly, nullptr);
    hipLaunchKernelGGL((k_scan_x<T, L>), dim3((d.Y + 63) / 64, d.Z), dim3(64), 0, s, (const T*)in, d, thr, hu, lz);
}

// ---------------------------------------------------------------- toFloat of a region
template <class T>
__global__ void __launch_bounds__(256) k_to_float_region(const T* __restrict__ in, Dims din, int ox, int oy, int oz,
                                                         float* __restrict__ out, Dims dout, float mn, float mx) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= dout.X) return;
    float v = (float)in[din.at(x + ox, y + oy, z + oz)];
    v = v > mx ? mx : v;
    v = v < mn ? mn : v;
    out[dout.at(x, y, z)] = (v - mn) / (mx - mn);
}

// Four voxels per thread (one 4-element load, one float4 store) when the
// region's x offset, both row lengths and both pointers are 4-element
// aligned.
template <class T> struct Vec4;
template <> struct Vec4<uint8_t> { typedef uchar4 V; };
template <> struct Vec4<int8_t> { typedef char4 V; };
template <> struct Vec4<uint16_t> { typedef ushort4 V; };
template <> struct Vec4<int16_t> { typedef short4 V; };
template <> struct Vec4<float> { typedef float4 V; };

__device__ __forceinline__ float to_unit(float v, float mn, float mx) {
    v = v > mx ? mx : v;
    v = v < mn ? mn : v;
    return (v - mn) / (mx - mn);
}

template <class T>
__global__ void __launch_bounds__(256) k_to_float_region4(const T* __restrict__ in, Dims din, int ox, int oy, int oz,
                                                          float* __restrict__ out, Dims dout, float mn, float mx) {
    const int x = 4 * (blockIdx.x * 256 + threadIdx.x), y = blockIdx.y, z = blockIdx.z;
    if (x >= dout.X) return;
    const typename Vec4<T>::V v = *reinterpret_cast<const typename Vec4<T>::V*>(in + din.at(x + ox, y + oy, z + oz));
    *reinterpret_cast<float4*>(out + dout.at(x, y, z)) =
        make_float4(to_unit((float)v.x, mn, mx), to_unit((float)v.y, mn, mx), to_unit((float)v.z, mn, mx),
                    to_unit((float)v.w, mn, mx));
}

template <class T>
void launch_to_float_region(const void* in, Dims di, tsf_dims off, float* out, Dims dout, float mn, float mx,
                            hipStream_t s) {
    const bool vec = off.x % 4 == 0 && di.X % 4 == 0 && dout.X % 4 == 0 &&
                     (uintptr_t)in % (4 * sizeof(T)) == 0 && (uintptr_t)out % 16 == 0;
    if (vec) {
        const dim3 g((dout.X / 4 + 255) / 256, dout.Y, dout.Z);
        hipLaunchKernelGGL(k_to_float_region4<T>, g, 256, 0, s, (const T*)in, di, off.x, off.y, off.z, out, dout, mn,
                           mx);
    } else {
        const dim3 g((dout.X + 255) / 256, dout.Y, dout.Z);
        hipLaunchKernelGGL(k_to_float_region<T>, g, 256, 0, s, (const T*)in, di, off.x, off.y, off.z, out, dout, mn,
                           mx);
    }
}

bool valid_type(int t) { return t >= 0 && t <= 4; }

// first index from `from` stepping by `dir` (while `i > stop` going down,
// `i < stop` going up) whose count satisfies a*lines > a*min
int first_slice(const int32_t* lines, int from, int stop, int dir, int a, int min_lines, int fallback) {
    if (dir > 0) {
        for (int i = from; i < stop; i++)
            if (a * lines[i] > a * min_lines) return i;
    } else {
        for (int i = from; i > stop; i--)
            if (a * lines[i] > a * min_lines) return i;
    }
    return fallback;
}

// grows [lo, hi) by one voxel at a time, alternately at the low end (when
// `lower` and lo > 0) and the high end, until its length is a multiple of 4
// or the whole extent; then trims the length down to a multiple of 4.  The
// alternation flag carries over from one axis to the next.
int make_divisible(int& lo, int& hi, int extent, bool& lower) {
    int size = hi - lo;
    while (size % 4 != 0 && size < extent) {
        if (lower && lo > 0) lo--;
        else if (hi < extent) hi++;
        lower = !lower;
        size = hi - lo;
    }
    while (size % 4 != 0) size--;
    return size;
}

}  // namespace

}  // namespace tsf

using namespace tsf;

extern "C" {

tsf_status tsf_volume_limits(const void* in, int32_t in_type, tsf_dims dims, float* minimum, float* maximum,
                             hipStream_t stream) {
    const char* W = "tsf_volume_limits";
    if (!in || !minimum || !maximum) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (!valid_type(in_type)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown input type %d", W, in_type);
    TSF_TRY(check_dims(dims, 1, W));
    const i64 n = D(dims).n();
    // about 64 elements per thread, at most 4096 workgroups
    const i64 threads_wanted = (n + 63) / 64;
    i64 blocks = (threads_wanted + LIM_THREADS - 1) / LIM_THREADS;
    if (blocks > 4096) blocks = 4096;
    Scratch sc;
    TSF_TRY(sc.get(nullptr, 0, sizeof(Limits) * (blocks + 1), W));
    Limits* part = (Limits*)sc.ptr;
    const unsigned g = (unsigned)blocks;
    switch (in_type) {
        case 0: launch_limits<uint8_t>(in, n, g, part, stream); break;
        case 1: launch_limits<int8_t>(in, n, g, part, stream); break;
        case 2: launch_limits<uint16_t>(in, n, g, part, stream); break;
        case 3: launch_limits<int16_t>(in, n, g, part, stream); break;
        default: launch_limits<float>(in, n, g, part, stream); break;
    }
    TSF_TRY(check_launch(W));
    hipLaunchKernelGGL(k_limits_final, 1, LIM_THREADS, 0, stream, part, (int)blocks, part + blocks);
    TSF_TRY(check_launch(W));
    Limits r;
    hipError_t e = hipMemcpyAsync(&r, part + blocks, sizeof r, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
    static const float init_min[5] = {255.0f, 127.0f, 65535.0f, 32767.0f, FLT_MAX};
    static const float init_max[5] = {0.0f, -128.0f, 0.0f, -32768.0f, FLT_MIN};
    const float mn0 = init_min[in_type], mx0 = init_max[in_type];
    *minimum = (r.mn.i >= 0 && r.mn.v < mn0) ? r.mn.v : mn0;
    *maximum = (r.mx.i >= 0 && r.mx.v > mx0) ? r.mx.v : mx0;
    return TSF_OK;
}

tsf_status tsf_crop_scan_lines(const void* in, int32_t in_type, tsf_dims dims, int32_t method, float threshold,
                               int32_t* lines_x, int32_t* lines_y, int32_t* lines_z, hipStream_t stream) {
    const char* W = "tsf_crop_scan_lines";
    if (!in || !lines_x || !lines_y || !lines_z) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    if (!valid_type(in_type)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown input type %d", W, in_type);
    if (method != TSF_CROP_THRESHOLD && method != TSF_CROP_LUNG)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown cropping method %d", W, method);
    if (method == TSF_CROP_LUNG && in_type == 4)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: lung cropping reads integer volumes only", W);
    TSF_TRY(check_dims(dims, 1, W));
    const Dims d = D(dims);
    hipLaunchKernelGGL(k_zero_i32, (d.X + 255) / 256, 256, 0, stream, lines_x, d.X);
    hipLaunchKernelGGL(k_zero_i32, (d.Y + 255) / 256, 256, 0, stream, lines_y, d.Y);
    hipLaunchKernelGGL(k_zero_i32, (d.Z + 255) / 256, 256, 0, stream, lines_z, d.Z);
    // HU limit of the lung automaton: -150, +1024 for unsigned volumes
    const int hu = (in_type == 0 || in_type == 2) ? -150 + 1024 : -150;
    if (method == TSF_CROP_THRESHOLD) {
        // one byte per (y, z) row for the x-line verdicts
        Scratch sc;
        TSF_TRY(sc.get(nullptr, 0, (size_t)d.Y * d.Z, W));
        uint8_t* rows = (uint8_t*)sc.ptr;
        switch (in_type) {
            case 0: launch_scan<uint8_t, ThresholdLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, rows, stream); break;
            case 1: launch_scan<int8_t, ThresholdLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, rows, stream); break;
            case 2: launch_scan<uint16_t, ThresholdLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, rows, stream); break;
            case 3: launch_scan<int16_t, ThresholdLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, rows, stream); break;
            default: launch_scan<float, ThresholdLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, rows, stream); break;
        }
        TSF_TRY(check_launch(W));
        const hipError_t e = hipStreamSynchronize(stream);  // the rows scratch is freed on return
        if (e != hipSuccess) return set_error(TSF_ERR_HIP, "%s: %s", W, hipGetErrorString(e));
        return TSF_OK;
    }
    switch (in_type) {
        case 0: launch_scan<uint8_t, LungLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, nullptr, stream); break;
        case 1: launch_scan<int8_t, LungLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, nullptr, stream); break;
        case 2: launch_scan<uint16_t, LungLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, nullptr, stream); break;
        default: launch_scan<int16_t, LungLine>(in, d, threshold, hu, lines_x, lines_y, lines_z, nullptr, stream); break;
    }
    return check_launch(W);
}

tsf_status tsf_crop_region(const int32_t* lines_x, const int32_t* lines_y, const int32_t* lines_z, tsf_dims dims,
                           int32_t min_scan_lines, int32_t start_z_middle, tsf_dims* offset, tsf_dims* size) {
    const char* W = "tsf_crop_region";
    if (!lines_x || !lines_y || !lines_z || !offset || !size)
        return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL argument", W);
    TSF_TRY(check_dims(dims, 1, W));
    const int X = dims.x, Y = dims.y, Z = dims.z, m = min_scan_lines;
    int x1 = first_slice(lines_x, 0, X, 1, 1, m, 0), x2 = first_slice(lines_x, X - 1, 0, -1, 1, m, X);
    int y1 = first_slice(lines_y, 0, Y, 1, 1, m, 0), y2 = first_slice(lines_y, Y - 1, 0, -1, 1, m, Y);
    // z: from the middle outwards to the first slice at or below the minimum
    // ("middle"), or from both ends inwards to the first slice above it
    const int start = start_z_middle ? Z / 2 : 0, a = start_z_middle ? -1 : 1;
    int z2 = first_slice(lines_z, start, Z, 1, a, m, Z);
    int z1 = first_slice(lines_z, Z - start - 1, 0, -1, a, m, 0);
    if (!start_z_middle) {
        const int t = z1;
        z1 = z2;
        z2 = t;
    }
    if (x2 - x1 <= 0 || y2 - y1 <= 0 || z2 - z1 <= 0)
        return set_error(TSF_ERR_RANGE, "Invalid cropping to new size %d, %d, %d", x2 - x1, y2 - y1, z2 - z1);
    bool lower = false;
    const int sx = make_divisible(x1, x2, X, lower);
    const int sy = make_divisible(y1, y2, Y, lower);
    const int sz = make_divisible(z1, z2, Z, lower);
    if (sx <= 0 || sy <= 0 || sz <= 0)
        return set_error(TSF_ERR_RANGE, "Invalid cropping to new size %d, %d, %d", sx, sy, sz);
    *offset = tsf_dims{x1, y1, z1};
    *size = tsf_dims{sx, sy, sz};
    return TSF_OK;
}

tsf_status tsf_to_float_region(const void* in, int32_t in_type, tsf_dims in_dims, tsf_dims offset, float* out,
                               tsf_dims out_dims, float minimum, float maximum, hipStream_t stream) {
    const char* W = "tsf_to_float_region";
    if (!in || !out) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: NULL volume", W);
    if (!valid_type(in_type)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: unknown input type %d", W, in_type);
    TSF_TRY(check_dims(in_dims, 1, W));
    TSF_TRY(check_dims(out_dims, 1, W));
    if (offset.x < 0 || offset.y < 0 || offset.z < 0 || (i64)offset.x + out_dims.x > in_dims.x ||
        (i64)offset.y + out_dims.y > in_dims.y || (i64)offset.z + out_dims.z > in_dims.z)
        return set_error(TSF_ERR_DIMENSIONS, "%s: region outside the volume", W);
    if (!(maximum != minimum)) return set_error(TSF_ERR_INVALID_ARGUMENT, "%s: maximum == minimum", W);
    const Dims di = D(in_dims), dout = D(out_dims);
    switch (in_type) {
        case 0: launch_to_float_region<uint8_t>(in, di, offset, out, dout, minimum, maximum, stream); break;
        case 1: launch_to_float_region<int8_t>(in, di, offset, out, dout, minimum, maximum, stream); break;
        case 2: launch_to_float_region<uint16_t>(in, di, offset, out, dout, minimum, maximum, stream); break;
        case 3: launch_to_float_region<int16_t>(in, di, offset, out, dout, minimum, maximum, stream); break;
        default: launch_to_float_region<float>(in, di, offset, out, dout, minimum, maximum, stream); break;
    }
    return check_launch(W);
}

}  // extern "C"
