/*
 * tsf.h -- C ABI of the MI355X (gfx950) volume-filter library (libtsf_hip.so).
 *
 * The library is the device side of the Tube Segmentation Framework's
 * volume-filter path: gradient vector field creation, gradient vector flow
 * (Euler iterations and the FMG multigrid solver), Hessian eigenanalysis,
 * circle-fitting TDF, and histogram-pyramid compaction of centreline
 * candidates.  Each entry point below names the reference function or kernel
 * it replaces (paths relative to the reference source tree).
 *
 * Conventions
 *  - Volumes are dense X*Y*Z arrays in device memory, linear index
 *    x + X*(y + Y*z) (the reference's LPOS macro, kernels_no_3d_write.cl:6).
 *  - A vector field is a tsf_field: four separate planes x, y, z, w (SoA).
 *    Plane w holds the per-voxel magnitude written by the GVF finish step
 *    (|v|, or 1 where |v| == 0) and is read by the normalised Hessian.  The
 *    reference stores the same four channels interleaved as one RGBA texel.
 *  - All calls are asynchronous on the given stream and return a tsf_status;
 *    nothing throws across the ABI.  tsf_last_error() returns a message for
 *    the last failure on the calling thread.
 *  - Entry points that need temporary device memory take (workspace,
 *    workspace_bytes).  Pass NULL/0 to let the library allocate it (the call
 *    then synchronises the stream before freeing); query the size with the
 *    matching *_workspace_size function to supply it yourself.  A supplied
 *    workspace must be 256-byte aligned (tsf_malloc / hipMalloc memory is).
 */
#ifndef TSF_H
#define TSF_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tsf_status {
    TSF_OK = 0,
    TSF_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad enum, bad parameter value */
    TSF_ERR_DIMENSIONS = 2,       /* volume too small for the operator */
    TSF_ERR_HIP = 3,              /* a HIP runtime call or kernel launch failed */
    TSF_ERR_OUT_OF_MEMORY = 4,    /* workspace allocation failed or too small */
    TSF_ERR_CAPACITY = 5,         /* output list capacity exceeded (count still reported) */
    TSF_ERR_RANGE = 6             /* result outside the range the caller accepts */
} tsf_status;

typedef enum tsf_storage {
    TSF_STORAGE_F32 = 0,  /* float planes */
    TSF_STORAGE_F16 = 1,  /* IEEE half planes holding value * scale */
    TSF_STORAGE_BF16 = 2, /* bfloat16 planes holding value * scale */
    TSF_STORAGE_SN16 = 3  /* int16 SNORM planes, the reference's 16-bit vector format
                             (kernels_no_3d_write.cl:9-21): stored as
                             convert_short_sat_rte(v * 32767), read as
                             max(-1, s / 32767); scale must be 1 */
} tsf_storage;

typedef struct tsf_dims {
    int32_t x, y, z;
} tsf_dims;

typedef struct tsf_field {
    void* x;         /* plane of component x (X*Y*Z elements of `storage`) */
    void* y;
    void* z;
    void* w;         /* magnitude plane; may be NULL where an entry point says so */
    int32_t storage; /* tsf_storage */
    float scale;     /* stored = value * scale; 1 for F32; a power of two */
} tsf_field;

/* ------------------------------------------------------------ housekeeping */
const char* tsf_last_error(void);
const char* tsf_version(void);
/* Device memory and stream helpers for hosts without their own allocator. */
tsf_status tsf_malloc(void** ptr, size_t bytes);
tsf_status tsf_free(void* ptr);
tsf_status tsf_memcpy_htod(void* dst, const void* src, size_t bytes, hipStream_t stream);
tsf_status tsf_memcpy_dtoh(void* dst, const void* src, size_t bytes, hipStream_t stream);
tsf_status tsf_memcpy_dtod(void* dst, const void* src, size_t bytes, hipStream_t stream);
tsf_status tsf_memset(void* dst, int value, size_t bytes, hipStream_t stream);
tsf_status tsf_stream_create(hipStream_t* stream);
tsf_status tsf_stream_destroy(hipStream_t stream);
tsf_status tsf_stream_synchronize(hipStream_t stream);
/* Event timing on a stream: ms between two recorded events. */
tsf_status tsf_event_create(hipEvent_t* ev);
tsf_status tsf_event_destroy(hipEvent_t ev);
tsf_status tsf_event_record(hipEvent_t ev, hipStream_t stream);
tsf_status tsf_event_elapsed_ms(hipEvent_t start, hipEvent_t stop, float* ms);
/* Work queued on `stream` after this call waits for `ev` (recorded on another
 * stream): independent stages of the path can run concurrently. */
tsf_status tsf_stream_wait_event(hipStream_t stream, hipEvent_t ev);
tsf_status tsf_device_count(int* count);
/* Free and total device memory of the current device (hipMemGetInfo). */
tsf_status tsf_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------ volume prep */

/* Normalised Gaussian mask of createBlurMask (tube-segmentation.cpp:188-211):
 * maskSize m = clamp(ceil(sigma/0.5), 1, 5), (2m+1)^3 weights
 * exp(-(a^2+b^2+c^2)/(2 sigma^2)) / sum, index a + b*M + c*M*M.  Host only;
 * `mask` must hold 1331 floats. */
tsf_status tsf_create_blur_mask(float sigma, float* mask, int32_t* mask_size);

/* blurVolumeWithGaussian (kernels_no_3d_write.cl:1442-1465) with the mask of
 * tsf_create_blur_mask(sigma): clamp-to-edge (2m+1)^3 correlation, summed in
 * the reference's loop order.  Replaces the two blur call sites of
 * runCircleFittingMethod (tube-segmentation.cpp:247-291, 536-582).
 * m <= 4 (sigma <= 2): the weights travel in the kernel arguments, so the
 * call is asynchronous and allocates nothing.  m = 5: the 1331 weights are
 * copied to a device buffer the call allocates, and it synchronises the
 * stream before freeing it. */
tsf_status tsf_blur_gaussian(const float* in, float* out, tsf_dims dims, float sigma, hipStream_t stream);

/* toFloat (kernels_no_3d_write.cl:1414-1440): clamp to [minimum, maximum],
 * then (v - minimum) / (maximum - minimum).  in_type: 0 = uint8, 1 = int8,
 * 2 = uint16, 3 = int16, 4 = float (the ElementTypes readDatasetAndTransfer
 * accepts, tube-segmentation.cpp:1490-1536). */
tsf_status tsf_to_float(const void* in, int32_t in_type, float* out, tsf_dims dims, float minimum,
                        float maximum, hipStream_t stream);

/* ------------------------------------------------------------ ingest */

/* getLimits with "minimum"/"maximum" off (getMinimum / getMaximum,
 * tube-segmentation.cpp:1356-1374, over the raw volume as read from the file):
 * std::min / std::max in index order from numeric_limits<T>::max() / ::min()
 * (for float ::min() = FLT_MIN), NaN ignored, of equal values the first.
 * in_type as tsf_to_float.  The results are host floats; the stream is
 * synchronised. */
tsf_status tsf_volume_limits(const void* in, int32_t in_type, tsf_dims dims, float* minimum, float* maximum,
                             hipStream_t stream);

enum { TSF_CROP_THRESHOLD = 0, TSF_CROP_LUNG = 1 };

/* The scan-line counts of cropping (cropDatasetThreshold / cropDatasetLung,
 * kernels_no_3d_write.cl:1236-1368, enqueued at tube-segmentation.cpp:
 * 1577-1610), on the raw volume (in_type as tsf_to_float):
 *   lines_x[x] = number of y whose z-line (x, y, .) is inside,
 *   lines_y[y] = number of x whose z-line (x, y, .) is inside,
 *   lines_z[z] = number of y whose x-line (., y, z) is inside.
 * TSF_CROP_THRESHOLD: a line is inside when some value > threshold.
 * TSF_CROP_LUNG: the reference's run-length test on HU = (short)value against
 * -150 (+1024 for unsigned types); integer volumes only.  Counts are int32
 * (the reference's are short, equal below 32768). */
tsf_status tsf_crop_scan_lines(const void* in, int32_t in_type, tsf_dims dims, int32_t method, float threshold,
                               int32_t* lines_x, int32_t* lines_y, int32_t* lines_z, hipStream_t stream);

/* Host only: the crop region from the scan-line counts (host arrays), as
 * readDatasetAndTransfer computes it (tube-segmentation.cpp:1612-1745): the
 * first slices with more than min_scan_lines lines from either end (z: from
 * the middle outwards to the first slice with fewer, when start_z_middle,
 * as lung cropping and "cropping-start-z middle" do), then each extent grown
 * to a multiple of 4 alternately at its two ends (or trimmed once it spans
 * the volume).  TSF_ERR_RANGE when an extent is empty or negative (the
 * reference throws "Invalid cropping" for an empty one). */
tsf_status tsf_crop_region(const int32_t* lines_x, const int32_t* lines_y, const int32_t* lines_z, tsf_dims dims,
                           int32_t min_scan_lines, int32_t start_z_middle, tsf_dims* offset, tsf_dims* size);

/* The copy of a region (enqueueCopyImage, tube-segmentation.cpp:1747-1764)
 * fused with toFloat (as tsf_to_float): out(x, y, z) = toFloat(in(x + offset.x,
 * y + offset.y, z + offset.z)). */
tsf_status tsf_to_float_region(const void* in, int32_t in_type, tsf_dims in_dims, tsf_dims offset, float* out,
                               tsf_dims out_dims, float minimum, float maximum, hipStream_t stream);

/* createVectorField (kernels_no_3d_write.cl:1470-1498): F = sign * central
 * difference of I (clamp-to-edge); F /= Fmax if |F| < Fmax else F /= |F|;
 * w = 1.  `out` must have all four planes.  sign = -1 for mode "black"
 * (tube-segmentation.cpp:220). */
tsf_status tsf_create_vector_field(const float* volume, const tsf_field* out, tsf_dims dims, float fmax,
                                   int32_t sign, hipStream_t stream);

/* ------------------------------------------------------------ GVF */

/* runFastGVF (gradientVectorFlow.cpp:1007-1110) with GVF3DIteration
 * (kernels_no_3d_write.cl:1746-1787) and GVF3DFinish (1798-1807):
 * v0 = f; `iterations` Jacobi steps
 *   v(w) <- v(q) + mu*lap(v)(q) - (v(q) - f(q))*|f(q)|^2,  q = mirror(w)
 * with the reference's boundary rule (0 -> 2, N-1 -> N-3); then
 * w = |v| (or 1 where |v| == 0).  f needs planes x, y, z; out needs all four.
 * f and out use the same storage type.  `fuse`: iterations per launch, 1, 2
 * or 4 (>= 4); a remainder runs first as one two-step and one single launch.
 * The result does not depend on it for F32 and SN16 storage (bit for bit);
 * for F16/BF16 the intermediate iterates of a launch stay in float.  Two per
 * launch is the fastest for 16-bit storage, four for F32.  Note: the reference returns iterate
 * 2*floor(iterations/2) (its finish kernel reads the first ping-pong buffer);
 * this entry point returns iterate `iterations`, the C++ runGVF adapter
 * reproduces the reference's rounding.  Dimensions must be >= 4. */
tsf_status tsf_gvf_euler(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                         int32_t fuse, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_gvf_euler_workspace_size(tsf_dims dims, int32_t storage);

/* runLowMemoryGVF (gradientVectorFlow.cpp:1173-1445; GVF3DInit_one_component,
 * GVF3DIteration_one_component, GVF3DFinish_one_component,
 * kernels_no_3d_write.cl:1669-1744), which the reference selects when a whole
 * field exceeds one allocation (tube-segmentation.cpp:743-762): each
 * component iterates alone with f's component and |f|^2, so besides f and
 * out the device holds one |f|^2 plane and one ping-pong plane.  |f|^2 is
 * stored in the field's format, as in the reference: F32 results equal
 * tsf_gvf_euler's bit for bit; SNORM16 ones differ from them where |f|^2
 * rounds.  One step per launch.  f, out: F32 or SNORM16, the same storage;
 * returns iterate `iterations` (the C++ runGVF adapter rounds to the
 * reference's 2*floor(N/2)).  Workspace:
 * tsf_gvf_euler_low_memory_workspace_size (two planes). */
tsf_status tsf_gvf_euler_low_memory(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu,
                                    int32_t iterations, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_gvf_euler_low_memory_workspace_size(tsf_dims dims, int32_t storage);

/* 1 when tsf_gvf_euler_sr runs on volumes of these dimensions (X even and
 * >= 64, Y >= 32, Z >= 4, and a 38-plane chunk of 16-bit values under
 * 2 GiB), else 0: its precondition, for callers that pick another path
 * instead.  No counterpart in the reference (the precision form is this
 * library's). */
int32_t tsf_gvf_euler_sr_supported(tsf_dims dims);

/* The precision form of tsf_gvf_euler (GVF3DIteration, kernels_no_3d_write.cl:
 * 1746-1787, the recurrence of runFastGVF, gradientVectorFlow.cpp:1007-1172):
 * the iterate streams through device memory as 16-bit planes and every step
 * computes in float, two Jacobi steps per launch.  Each launch rounds its
 * result to 16 bits stochastically: the magnitude rounds up with
 * probability equal to the discarded fraction (to 2^-10), the random bits a
 * hash of the voxel index and seed ^ (launch * 0x9e3779b9), so results are
 * reproducible.  Round to nearest would drop every update smaller than half
 * a 16-bit ulp and stall the iteration short of its fixed point.  The last
 * launch writes `out` as F32 planes with w = |v| (1 where 0): the final
 * iterate at full precision, as the Hessian of the TDF needs it (see
 * tsf_tdf_circle_mixed).  An odd count starts with one step rounded to
 * nearest.  The update is the reference's v + mu lap(v) - (v - f)|f|^2
 * arranged as v (1 - 6 mu - |f|^2) + f |f|^2 + mu sum(neighbours), with the
 * reference's mirror rule at the faces.  The last f32_tail iterations (if
 * > 0) run with the iterate in F32 planes, in the reference's operation
 * order (as tsf_gvf_euler on F32 fields), on f's stored values: they damp
 * the rounding noise of the last 16-bit stores, which the Hessian's central
 * differences would otherwise amplify.
 * f: F16 (value * scale) or BF16 planes -- the iterate then has f's storage
 * and scale -- or SNORM16 planes, the reference's 16-bit vector format
 * (the integer s read as max(s, -32767) * (unit / 32767) in units of the
 * iterate, F16 at unit TSF_SR_SN16_UNIT).  SNORM16 holds a field of
 * components in [-1, 1] to 1.5e-5 where F16 holds it to 2^-12 relative, and
 * f's rounding is the larger part of the F16 form's error.  x, y, z planes,
 * 4-byte aligned; out: F32 with w.  Needs X even and >= 64, Y >= 32, Z >= 4.
 * Workspace: tsf_gvf_euler_sr_workspace_size (two 16-bit iterates, and F32
 * planes for the tail). */
#define TSF_SR_SN16_UNIT 16384.0f
tsf_status tsf_gvf_euler_sr(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                            int32_t f32_tail, uint32_t seed, void* workspace, size_t workspace_bytes,
                            hipStream_t stream);
size_t tsf_gvf_euler_sr_workspace_size(tsf_dims dims, int32_t f32_tail);

/* runFMGGVF (gradientVectorFlow.cpp:729-1005): per component, `iterations`
 * defect-correction steps r = fmgResidual(f); e = fullMultigrid(r); f += e
 * with red-black Gauss-Seidel smoothing (GVFgaussSeidel/GVFgaussSeidel2,
 * kernels_no_3d_write.cl:2071-2150), 2x2x2 restriction, injection
 * prolongation, v0 V-cycles of v1 pre- and v2 post-sweeps over
 * levels+1 grids (reference: v0=1, v1=2, v2=2, levels = l_max = 2); then
 * w = |v| (no zero guard, MGGVFFinish 2202-2216).  F32 storage only.
 * `residual_norms` (optional, host array of 3*iterations doubles) receives
 * the L2 norm of each component's residual before every correction, reduced
 * in double on the device. */
tsf_status tsf_gvf_fmg(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu, int32_t iterations,
                       int32_t levels, int32_t v0, int32_t v1, int32_t v2, double* residual_norms,
                       void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_gvf_fmg_workspace_size(tsf_dims dims, int32_t levels);

/* ------------------------------------------------------------ Hessian / eigen */

/* Per-voxel Jacobian "Hessian" of the field (gradient / gradientNormalized,
 * kernels_no_3d_write.cl:90-200; matrix layout 1522-1526) followed by the
 * EISPACK tred2/tql2 eigensolver (1824-2068): eigenvalues sorted by ascending
 * |lambda|.  normalize != 0 divides every sample by its voxel's w plane.
 * lambda: 3 planes (float, X*Y*Z each, contiguous).  vectors (optional, may
 * be NULL): 9 planes, plane 3*j + k = component k of eigenvector j (e1 = the
 * tube direction).  Replaces the device-side use of eigen_decomposition and
 * the CPU doEigen batch (eigenanalysisOfHessian.cpp:343-389). */
tsf_status tsf_hessian_eigen(const tsf_field* field, tsf_dims dims, int32_t normalize, float* lambda,
                             float* vectors, hipStream_t stream);

/* doEigen / getTubeDirection (eigenanalysisOfHessian.cpp:343-389,
 * eigenanalysisOfHessian.hpp:7-9) for a batch of points: Jacobian of the
 * field where every sample is divided by sqrt(Fx^2+Fy^2+Fz^2) of its voxel,
 * then tred2/tql2.  positions: device int32[3*count] (x,y,z), each at least
 * one voxel from every border.  out: device float[12*count] per point
 * (lambda x,y,z, e1 xyz, e2 xyz, e3 xyz). */
tsf_status tsf_do_eigen_points(const float* fx, const float* fy, const float* fz, tsf_dims dims,
                               const int32_t* positions, int64_t count, float* out, hipStream_t stream);

/* ------------------------------------------------------------ TDF */

/* circleFittingTDF (kernels_no_3d_write.cl:1503-1572) via runCircleFittingTDF
 * (tubeDetectionFilters.cpp:57-73): Hessian (normalised when rmax >= 4) ->
 * eigenvectors e2, e3 -> for r = rmin; r <= rmax; r += rstep (float
 * accumulation) the mean over 32 directions V_j = cos(j) e3 + sin(j) e2
 * (j in integer radians) of dot(-F_lin(p + r V_j), V_j), F_lin = trilinear
 * sampling with the OpenCL linear-filter convention (-0.5 shift,
 * clamp-to-edge); keep the first maximum, stop at the first r that does not
 * increase it.  tdf = max(0, best), radius = its r (0 if none).  The field
 * needs plane w when rmax >= 4.  Workspace: the field repacked as one texel
 * (x, y, z, w) per voxel for the trilinear gathers; an F32 field with
 * max(|rmin|, |rmax|) < 5.5 is read from its planes and needs none (the
 * workspace, if given, is unused; wider radii gather the packed texels from
 * device memory). */
tsf_status tsf_tdf_circle(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep, float* tdf,
                          float* radius, void* workspace, size_t workspace_bytes, hipStream_t stream);

/* splineTDF (kernels_no_3d_write.cl:1574-1667) as runSplineTDF enqueues it
 * (tubeDetectionFilters.cpp:4-55: 12 arms; the caller passes
 * max(1, radiusMin) and min_average_mag 0.1), used for the large-radius TDF
 * when "use-spline-tdf" is set (tube-segmentation.cpp:785-786).  Per voxel:
 * normalised Hessian (w required), eigenvectors e1..e3; along 12 arms
 * cos(a) e3 + sin(a) e2, a = 2 pi j / 12, the first radius from rmin + rstep
 * (step rstep, up to rmax) where the trilinear field magnitude, after rising,
 * falls with a two-sample mean above the voxel's own |F| by min_average_mag.
 * radius = sum of the borders found / 12; tdf = 0 when an arm finds none or
 * its border field points outward, else min(1, mean(1 - |n . e1|) *
 * symmetry + 0.2).  Workspace: tsf_tdf_spline_workspace_size (packed
 * texels). */
size_t tsf_tdf_spline_workspace_size(tsf_dims dims, int32_t storage);
tsf_status tsf_tdf_spline(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep,
                          float min_average_mag, float* tdf, float* radius, void* workspace, size_t workspace_bytes,
                          hipStream_t stream);
size_t tsf_tdf_circle_workspace_size(tsf_dims dims, int32_t storage);

/* The circleFittingTDF of tsf_tdf_circle in the precision form's mixed mode:
 * the Hessian and its eigenvectors from the F32 `field` itself, the 32 x n_r
 * trilinear samples from a copy of its x, y, z planes rounded to 16-bit
 * texels (sample_storage F16 or BF16, holding value * sample_scale, a power
 * of two; 2^14 suits |v| <= 1 fields).  The eigenvectors are the part of the
 * TDF that amplifies storage rounding (the Hessian is a central difference);
 * the samples are averages and are not sensitive to it, and the 8-byte
 * texels keep the sampling at the 16-bit path's speed.  field: F32, with w
 * when rmax >= 4.  Workspace: tsf_tdf_circle_mixed_workspace_size (8 bytes
 * per voxel).  Replaces runCircleFittingTDF (tubeDetectionFilters.cpp:57-73)
 * where the field comes from the precision form's GVF. */
tsf_status tsf_tdf_circle_mixed(const tsf_field* field, int32_t sample_storage, float sample_scale, tsf_dims dims,
                                float rmin, float rmax, float rstep, float* tdf, float* radius, void* workspace,
                                size_t workspace_bytes, hipStream_t stream);
size_t tsf_tdf_circle_mixed_workspace_size(tsf_dims dims);

/* combine (kernels_no_3d_write.cl:1087-1098): where tdf_large < tdf_small,
 * take the small-scale TDF and radius.  In place on the large buffers. */
tsf_status tsf_combine(const float* tdf_small, const float* radius_small, float* tdf_large, float* radius_large,
                       int64_t count, hipStream_t stream);

/* The reference's 16-bit TDF buffers (TDF_TYPE ushort with 16bit-vectors,
 * kernels_no_3d_write.cl:19-21, written at 1570 and 1663): rounds each value
 * in place to what that format holds, convert_ushort_sat_rte(t * 65535) / 65535
 * (NaN -> 0).  Applied after tsf_tdf_circle on an SN16 field it gives the
 * reference's 16-bit-mode TDF values; tsf_combine and tsf_candidates then
 * compare them as the reference compares the codes. */
tsf_status tsf_tdf_round_unorm16(float* tdf, int64_t count, hipStream_t stream);

/* ------------------------------------------------------------ candidates */

/* findCandidateCenterpoints (kernels_no_3d_write.cl:1006-1017):
 * mask = tdf < threshold ? 0 : 1. */
tsf_status tsf_candidates(const float* tdf, uint8_t* mask, int64_t count, float threshold, hipStream_t stream);

/* Histogram-pyramid compaction (oul::HistogramPyramid3DBuffer create /
 * getSum / createPositionBuffer as used at parallelCenterlineExtraction.cpp:
 * 727-754 and 839-892; kernels constructHPLevel*, scanHPLevel*,
 * traverseHP3DBuffer, createPositions3DBuffer, kernels_no_3d_write.cl:
 * 246-698).  Writes the (x,y,z) of every nonzero mask voxel into `positions`
 * (device int32[3*capacity]) in the pyramid's traversal order: ascending key
 * sum_b (x_b + 2 z_b + 4 y_b) 8^b.  *count (host) receives the number of
 * set voxels (the pyramid's sum); the call synchronises the stream to read
 * it.  Returns TSF_ERR_CAPACITY (with *count set and the first `capacity`
 * entries written) when the list does not fit.  positions == NULL with
 * capacity == 0 only counts (the pyramid's getSum). */
tsf_status tsf_hp_compact(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity,
                          int64_t* count, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_hp_compact_workspace_size(tsf_dims dims);

/* tsf_hp_compact without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The cube counts and their scan use a static device array of
 * 3 MiB + 8 B that each GPU holds until the process ends; volumes whose
 * largest extent exceeds 1024 are rejected with TSF_ERR_DIMENSIONS.  Same
 * outputs, count read-back and errors as tsf_hp_compact.  Calls that overlap
 * on one GPU (two streams or two host threads) share the array and race. */
tsf_status tsf_hp_compact_noalloc(const uint8_t* mask, tsf_dims dims, int32_t* positions, int64_t capacity,
                                  int64_t* count, hipStream_t stream);

/* findCandidateCenterpoints2 (kernels_no_3d_write.cl:1019-1084; the image
 * path kernels.cl:1309-1390 differs only in the cap of maxD): for each listed
 * candidate, maxD = clamp(round(radius), 1, maxd_cap); e1 from the normalised
 * Hessian; the candidate is rejected if a voxel n = p + r, 0 < |r| < maxD,
 * within 0.5 rad of the plane normal to e1 has w(n) < w(p).  Writes 1 (kept)
 * or 0 (rejected) into `out` at each candidate; other voxels are untouched.
 * maxd_cap: 5 (buffer path) or 8 (image path). */
tsf_status tsf_candidate_nms(const float* tdf, const float* radius, const tsf_field* field, tsf_dims dims,
                             const int32_t* positions, int64_t count, int32_t maxd_cap, uint8_t* out,
                             hipStream_t stream);

/* dd (kernels_no_3d_write.cl:970-1004): for every cube x cube x cube cell,
 * mark in `out` the candidate (mask == 1) with the strictly greatest TDF
 * (first in x-outer, z-inner order; TDF must be > 0).  `out` is not cleared. */
tsf_status tsf_cube_max(const float* tdf, const uint8_t* candidates, tsf_dims dims, int32_t cube, uint8_t* out,
                        hipStream_t stream);

/* ------------------------------------------------------------ centerline graph */

/* Centerpoint linking: linkLengths, compact, linkCenterpoints and
 * removeDuplicateEdges (kernels_no_3d_write.cl:727-882, 947-968) and the 2D
 * pyramid compaction of the edge image (createPositions2D/traverseHP2D,
 * 349-390, 624-725), as runNewCenterlineAlg runs them
 * (parallelCenterlineExtraction.cpp:912-1060).  vertices: int32 (x, y, z)
 * triples, `count` of them (< 65536).  For each vertex a, among its
 * candidates b (0 < |xb - xa| < max_distance, in ascending index) and pairs
 * (b, c) with c listed before b, the first of smallest round|ab| + round|ac|
 * with an angle at a of at least 2 rad (acos of the normalised dot product)
 * and a mean TDF (linearly interpolated, clamp-to-edge, at xa + ab k/n,
 * k = 0..n, n = round|ab|) of at least min_mean_tdf on both segments; then
 * edges (a, b) and (a, c), where (x, y), x > y, is dropped when (y, x) is an
 * edge.  The edges (int32 pairs) are written in the order of the pyramid's
 * traversal, key sum_k (2 x_k + y_k) 4^k of the bits of x and y.  *n_edges
 * = their number (the stream is synchronised); TSF_ERR_CAPACITY if it
 * exceeds `capacity` (edges == NULL, capacity 0: count only).  Workspace:
 * a row of `count` candidates per vertex (count^2 int32, as the reference's
 * count x count images). */
tsf_status tsf_link_centerpoints(const int32_t* vertices, int32_t count, const float* tdf, tsf_dims dims,
                                 float max_distance, float min_mean_tdf, int32_t* edges, int32_t capacity,
                                 int32_t* n_edges, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_link_centerpoints_workspace_size(int32_t count);

/* graphComponentLabeling + calculateTreeLength (kernels_no_3d_write.cl:
 * 884-918; parallelCenterlineExtraction.cpp:1062-1160): labels[v] = the
 * smallest vertex index of v's connected component (the fixed point the
 * reference's minimum propagation reaches), sizes[l] = the number of vertices
 * labelled l (0 for indices that are no label).  labels, sizes: `count`
 * int32 each.  Synchronises the stream. */
tsf_status tsf_graph_components(const int32_t* edges, int32_t n_edges, int32_t count, int32_t* labels,
                                int32_t* sizes, hipStream_t stream);

/* removeSmallTrees (kernels_no_3d_write.cl:920-945), the rasterisation used
 * when loop removal is off: for every edge (a, b) with
 * sizes[labels[a]] >= min_tree_length, sets centerlines (one byte per voxel)
 * to 1 at round(xa + (xb - xa) i / l), i = 0..l-1, l = round|xb - xa|.
 * Other voxels are untouched. */
tsf_status tsf_rasterize_tree_edges(const int32_t* edges, int32_t n_edges, const int32_t* vertices,
                                    const int32_t* labels, const int32_t* sizes, int32_t min_tree_length,
                                    uint8_t* centerlines, tsf_dims dims, hipStream_t stream);

/* ------------------------------------------------------------ segmentation */

/* runInverseGradientSegmentation (segmentation.cpp:5-221; kernels
 * initGrowing, grow, dilate, erode, kernels_no_3d_write.cl:1100-1207,
 * 1370-1412).  centerlines: one byte per voxel (1 = centerline); radius:
 * float per voxel (the averaged radius along the centerline); field: the
 * GVF field with w = |v|.  Seeds: a cube of half size
 * clamp(round(radius/2), 1, 4) around each centerline voxel.  A pass
 * accepts a candidate X when a neighbour Y not yet accepted, with
 * w(Y) > w(X), has its normalised field pointing most nearly at X among
 * Y's 26 neighbour directions (the first of equal ones in x, y, z loop
 * order); Y becomes a candidate; candidates without such a Y are dropped.
 * Where a candidate's verdict and a nomination of it meet, accepted >
 * nominated > dropped (the reference leaves this to the write order).
 * At least two passes, then until a pass accepts nothing; then a 3x3x3
 * dilation and erosion.  segmentation: one byte per voxel (0/1).
 * *iterations (if not NULL): the number of passes.  Synchronises once per
 * pass (the reference polls its stop flag likewise). */
tsf_status tsf_segment_inverse_gradient(const uint8_t* centerlines, const float* radius, const tsf_field* field,
                                        tsf_dims dims, uint8_t* segmentation, int32_t* iterations, void* workspace,
                                        size_t workspace_bytes, hipStream_t stream);
size_t tsf_segment_inverse_gradient_workspace_size(tsf_dims dims);

/* runSphereSegmentation / sphereSegmentation (segmentation.cpp:223-307;
 * kernels_no_3d_write.cl:1209-1234): 1 within |offset| < radius of every
 * centerline voxel, 0 elsewhere. */
tsf_status tsf_segment_spheres(const uint8_t* centerlines, const float* radius, tsf_dims dims, uint8_t* segmentation,
                               hipStream_t stream);

/* ------------------------------------------------------------ adjoints
 * Reverse-mode derivatives of the chain
 *   B = blur_sigma(I), F = createVectorField(B), V = GVF_N(F; mu), W = |V|,
 *   T = circleFittingTDF(V, W),   L = sum_p gT(p) T(p).
 * No reference counterpart: the reference has no gradients.  Discrete
 * choices are held fixed (blur mask size, the Fmax branch of the vector
 * field, the eigenvalue order, the TDF radius r*).  Float arithmetic; the
 * scalar gradients (d/dsigma, d/dmu) are accumulated in double on the
 * device.  Entry points that return a host double synchronise the stream. */

/* dI = B^T dB for the blur of tsf_blur_gaussian(sigma) (clamp-to-edge
 * transpose); dsigma (optional, needs `in`) = sum_p dB(p) sum_k dw_k/dsigma
 * in(clamp(p+k)) with the mask size held fixed. */
tsf_status tsf_blur_adjoint(const float* dB, const float* in, float* dI, tsf_dims dims, float sigma, double* dsigma,
                            void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_blur_adjoint_workspace_size(void);

/* dvolume = (d createVectorField / d volume)^T dF (planes x, y, z of dF). */
tsf_status tsf_vector_field_adjoint(const float* volume, const tsf_field* dF, tsf_dims dims, float fmax, int32_t sign,
                                    float* dvolume, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_vector_field_adjoint_workspace_size(tsf_dims dims);

/* Forward GVF (F32) that stores v^0, v^k, v^2k, ... (k = checkpoint_every)
 * into `checkpoints` (tsf_gvf_checkpoint_count(...) states of 3 float
 * planes, contiguous) and the result (with w) into out.  Same values as
 * tsf_gvf_euler. */
int64_t tsf_gvf_checkpoint_count(int32_t iterations, int32_t checkpoint_every);
tsf_status tsf_gvf_euler_checkpointed(const tsf_field* f, const tsf_field* out, tsf_dims dims, float mu,
                                      int32_t iterations, int32_t checkpoint_every, float* checkpoints,
                                      size_t checkpoint_bytes, hipStream_t stream);

/* Adjoint of the GVF iterations: from dV (3 float planes, dL/dv^N) returns
 * dF (3 planes: through v^0 = f and the data term of every step; mirror
 * boundary transposed) and dmu (host).  Each checkpoint segment is
 * recomputed once; memory (checkpoint_every + 2) states + the checkpoints. */
tsf_status tsf_gvf_euler_adjoint(const tsf_field* f, tsf_dims dims, float mu, int32_t iterations,
                                 int32_t checkpoint_every, const float* checkpoints, const float* dV, float* dF,
                                 double* dmu, void* workspace, size_t workspace_bytes, hipStream_t stream);
size_t tsf_gvf_euler_adjoint_workspace_size(tsf_dims dims, int32_t checkpoint_every);

/* Adjoint of the finish step w = |v|: dv.xyz += dv.w * v / |v| (|v| > 0). */
tsf_status tsf_gvf_finish_adjoint(const tsf_field* v, const tsf_field* dv, tsf_dims dims, hipStream_t stream);

/* Adjoint of circleFittingTDF at the forward's radius r*: accumulates
 * (+=) dL/dV into dfield.xyz and, when rmax >= 4 (normalised Hessian),
 * dL/dW into dfield.w.  Includes the sampled values, the sample positions
 * and the eigenvector directions (first-order perturbation of e2, e3).
 * Voxels with T = 0 carry no gradient.  radius: the forward's radius output
 * (tsf_tdf_circle on the same field and radius range), from which r* is
 * read and the response at r* recomputed (the forward's own code on the same
 * packed texels, so bit-identical); NULL repeats the forward's radius search
 * instead.  Contributions are summed in an LDS tile per 16x4x4 brick and
 * then added with float atomics, so the summation order (last bits) is not
 * reproducible between runs.  Workspace: the packed field (16 B/voxel). */
tsf_status tsf_tdf_adjoint(const tsf_field* field, tsf_dims dims, float rmin, float rmax, float rstep,
                           const float* tdf_grad, const float* radius, const tsf_field* dfield, void* workspace,
                           size_t workspace_bytes, hipStream_t stream);
size_t tsf_tdf_adjoint_workspace_size(tsf_dims dims);

/* Mean squared error of t against target over n floats: *loss (host) =
 * (1/n) sum (t - target)^2 and grad = dL/dt = 2 (t - target) / n, the
 * gradient that starts a backward of the chain when fitting mu and sigma to
 * a target response.  Synchronises the stream. */
tsf_status tsf_mse(const float* t, const float* target, int64_t n, float* grad, double* loss, void* workspace,
                   size_t workspace_bytes, hipStream_t stream);
size_t tsf_mse_workspace_size(void);

#ifdef __cplusplus
}
#endif

#endif /* TSF_H */
