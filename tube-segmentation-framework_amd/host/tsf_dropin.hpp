// C++ drop-in layer over the C ABI (include/tsf.h).
//
// Mirrors the reference's host interface for the volume-filter path -- same
// function names, argument order and meaning, ownership and error behaviour
// -- with the OpenCL handle types replaced by HIP-backed ones of the same
// role:
//   OpenCL  (commons.hpp:8-15)        -> tsf::OpenCL: one in-order HIP stream
//   cl::Image3D / cl::Buffer           -> tsf::Image3D / tsf::Buffer: device
//                                         volumes / linear buffers (shared
//                                         ownership, freed with the last copy)
//   SIPL::int3 / SIPL::float3          -> tsf::int3 / tsf::float3
//   paramList (parameters.hpp:78-91)   -> tsf::paramList with the same accessors
//   cl::Error / SIPL::SIPLException    -> tsf::Error (carries the tsf_status)
// Nothing crosses the C ABI as a C++ type; every call below is a thin
// sequence of tsf_* calls on the context's stream.
#pragma once

#include <map>
#include <memory>
#include <stack>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tsf.h"

namespace tsf {

struct int3 {
    int x = 0, y = 0, z = 0;
    int3() = default;
    int3(int a, int b, int c) : x(a), y(b), z(c) {}
};
struct float3 {
    float x = 0, y = 0, z = 0;
    float3() = default;
    float3(float a, float b, float c) : x(a), y(b), z(c) {}
};
struct int2 {
    int x = 0, y = 0;
    int2() = default;
    int2(int a, int b) : x(a), y(b) {}
};

class Error : public std::runtime_error {
   public:
    Error(tsf_status s, const std::string& what) : std::runtime_error(what), status(s) {}
    tsf_status status;
};

// SIPL::IOException (a file that cannot be opened or read).
class IOException : public Error {
   public:
    explicit IOException(const std::string& filename);
    std::string filename;
};

// Throws tsf::Error for a failed tsf_* call (the reference's calls throw
// cl::Error from the OpenCL C++ bindings).
void check(tsf_status st);

// ----------------------------------------------------------------- parameters
// parameters.hpp: typed parameters with ranges and allowed values.
class BoolParameter {
   public:
    BoolParameter() = default;
    BoolParameter(bool v, std::string description, std::string group);
    bool get() const { return value; }
    void set(bool v) { value = v; }
    std::string getDescription() const { return description; }
    std::string getGroup() const { return group; }

   private:
    bool value = false;
    std::string description, group;
};
class NumericParameter {
   public:
    NumericParameter() = default;
    NumericParameter(float v, float min, float max, float step, std::string description, std::string group);
    float get() const { return value; }
    void set(float v);  // throws tsf::Error outside [min, max]
    bool validate(float v) const { return v >= min && v <= max; }
    float getMin() const { return min; }
    float getMax() const { return max; }
    float getStep() const { return step; }
    std::string getDescription() const { return description; }

   private:
    float value = 0, min = 0, max = 0, step = 0;
    std::string description, group;
};
class StringParameter {
   public:
    StringParameter() = default;
    StringParameter(std::string v, std::vector<std::string> possibilities, std::string description,
                    std::string group);
    std::string get() const { return value; }
    void set(const std::string& v);  // throws tsf::Error if not one of the possibilities
    void setWithoutValidation(const std::string& v) { value = v; }
    bool validate(const std::string& v) const;
    std::vector<std::string> getPossibilities() const { return possibilities; }

   private:
    std::string value;
    std::vector<std::string> possibilities;
    std::string description, group;
};
struct paramList {
    std::map<std::string, BoolParameter> bools;
    std::map<std::string, NumericParameter> numerics;
    std::map<std::string, StringParameter> strings;
};

// Defaults of every parameter the volume-filter path reads (the reference's
// parameters/parameters file), plus "3d_write" (set by run() from the device
// in the reference) and "vector-storage" (f32 | f16 | bf16 | sn16, new: the
// storage of the GVF field; "auto" follows 16bit-vectors and 32bit-vectors,
// choosing sn16, the reference's 16-bit format, or f32).
paramList initParameters();
// Reads a parameters file in the reference's format (name type default ...).
paramList initParameters(const std::string& parameter_dir);
// Applies parameter_dir/centerline-<method>/<preset> ("name value" lines) when
// the "parameters" parameter is not "none" (parameters.cpp:67-106).
void loadParameterPreset(paramList& parameters, const std::string& parameter_dir);
void setParameter(paramList& parameters, const std::string& name, const std::string& value);
float getParam(const paramList& parameters, const std::string& name);
bool getParamBool(const paramList& parameters, const std::string& name);
std::string getParamStr(const paramList& parameters, const std::string& name);

// ----------------------------------------------------------------- context and memory
struct OpenCL {
    hipStream_t queue = nullptr;
    OpenCL();
    ~OpenCL();
    OpenCL(const OpenCL&) = delete;
    OpenCL& operator=(const OpenCL&) = delete;
    void finish();
};

class Buffer {
   public:
    Buffer() = default;
    explicit Buffer(size_t bytes);
    void* get() const { return mem ? mem.get() : nullptr; }
    size_t size() const { return bytes; }
    void upload(OpenCL& ocl, const void* host, size_t n);
    void download(OpenCL& ocl, void* host, size_t n) const;

   private:
    std::shared_ptr<void> mem;
    size_t bytes = 0;
};

// A device volume: `channels` planes (1 = scalar volume, 4 = vector field
// x, y, z, w) of one storage type, each size.x*size.y*size.z elements.
class Image3D {
   public:
    Image3D() = default;
    Image3D(int3 size, int channels, tsf_storage storage = TSF_STORAGE_F32, float scale = 1.0f);
    int3 size() const { return sz; }
    int channels() const { return nch; }
    tsf_storage storage() const { return st; }
    float scale() const { return sc; }
    void* plane(int c) const { return planes.at(c).get(); }
    bool valid() const { return !planes.empty(); }
    tsf_field field() const;  // channels == 4
    tsf_dims dims() const { return tsf_dims{sz.x, sz.y, sz.z}; }
    // host transfer of plane c as float values (scaled/rounded as stored)
    void uploadPlane(OpenCL& ocl, int c, const float* host);
    std::vector<float> downloadPlane(OpenCL& ocl, int c) const;

   private:
    int3 sz;
    int nch = 0;
    tsf_storage st = TSF_STORAGE_F32;
    float sc = 1.0f;
    std::vector<Buffer> planes;
};

// Storage the path uses for the vector field under these parameters
// (vector-storage precision: SNORM16 for the GVF's f).
tsf_storage vectorStorage(const paramList& parameters);
// vector-storage precision: the GVF iterate in F16 with stochastic rounding
// (tsf_gvf_euler_sr, the last gvf-f32-tail iterations in F32, the result F32)
// and both circle TDFs in the mixed mode (tsf_tdf_circle_mixed).
bool precisionForm(const paramList& parameters);
float vectorScale(tsf_storage s);

// ----------------------------------------------------------------- the volume-filter path
// createBlurMask (tube-segmentation.cpp:188-211): caller frees with delete[].
float* createBlurMask(float sigma, int* maskSizePointer);

// gradientVectorFlow.hpp:8-10.  Consume *vectorField (it is released, as the
// reference deletes it) and return the new field (channels 4, w = |v|).
// runGVF returns iterate 2*floor(gvf-iterations/2): the reference's finish
// step reads the first ping-pong buffer (gradientVectorFlow.cpp:1071, 1161).
// useLessMemory selects the reference's one-component variant
// (runLowMemoryGVF, tsf_gvf_euler_low_memory) for F32 and SNORM16 fields:
// equal to the full-field result in F32; in the 16-bit mode |f|^2 is stored
// as SNORM16 as the reference does.  F16 / BF16 fields run the full-field
// path (a note on stderr).  vector-storage precision runs tsf_gvf_euler_sr.
Image3D runGVF(OpenCL& ocl, Image3D* vectorField, paramList& parameters, int3& size, bool useLessMemory);
Image3D runFMGGVF(OpenCL& ocl, Image3D* vectorField, paramList& parameters, int3& size);

// tubeDetectionFilters.hpp:15.  TDF and radius are caller-allocated device
// buffers of size.x*size.y*size.z floats.
void runCircleFittingTDF(OpenCL& ocl, int3& size, Image3D* vectorField, Buffer* TDF, Buffer* radius,
                         float radiusMin, float radiusMax, float radiusStep);

// tubeDetectionFilters.hpp: the spline TDF (12 arms, rMin raised to 1),
// same buffers as runCircleFittingTDF.
void runSplineTDF(OpenCL& ocl, int3& size, Image3D* vectorField, Buffer* TDF, Buffer* radius, float radiusMin,
                  float radiusMax, float radiusStep);

// tube-segmentation.cpp:212-939: blur -> vector field -> (small TDF) -> blur
// -> vector field -> GVF (Euler or FMG) -> large TDF (circle fitting, or the
// spline TDF with "use-spline-tdf") -> combine.  dataset is a
// float scalar volume in [0, 1] (toFloat output); it is consumed.  Outputs
// are set by reference as in the reference.
void runCircleFittingMethod(OpenCL& ocl, Image3D* dataset, int3 size, paramList& parameters, Image3D& vectorField,
                            Image3D& TDF, Image3D& radiusImage);

// The candidate / compaction part of runNewCenterlineAlg
// (parallelCenterlineExtraction.cpp:689-899): threshold -> pyramid compaction
// -> cross-section non-maximum suppression -> cube maximum -> pyramid
// compaction of the vertices.  Throws tsf::Error with TSF_ERR_RANGE when the
// candidate count is <= 0 or > half the volume, or the vertex count is < 8 or
// >= 16384, as the reference does.
struct CenterpointVertices {
    Buffer centerpoints;   // uint8[size.x*size.y*size.z]: candidates kept by the suppression
    Buffer vertices;       // int32[3*count] in pyramid order
    int count = 0;
    int candidates = 0;
};
CenterpointVertices runNewCenterlineAlgCandidates(OpenCL& ocl, int3 size, paramList& parameters,
                                                  Image3D& vectorField, Image3D& TDF, Image3D& radius);

// runNewCenterlineAlg (parallelCenterlineExtraction.hpp:7;
// parallelCenterlineExtraction.cpp:667-1332): the candidates and vertices
// above, then centerpoint linking, component labelling and the centerline
// voxels.  Returns the centerline mask, one byte per voxel (the reference's
// CL_SIGNED_INT8 image), as a device Buffer; with "centerpoints-only" the
// suppressed candidates instead.  With "loop-removal" (the default) or a
// "centerline-vtk-file" the large trees go through removeLoops and
// createCenterlineVoxels on the host, which also rewrites `radius` along the
// centerline (and the VTK file is written); otherwise the device rasterises
// the large trees' edges.  Throws tsf::Error (TSF_ERR_RANGE) where the
// reference throws: no edges, more than 10 million edges, an empty graph.
Buffer runNewCenterlineAlg(OpenCL& ocl, int3 size, paramList& parameters, Image3D& vectorField, Image3D& TDF,
                           Image3D& radius);
// The host steps it uses (file-local in the reference,
// parallelCenterlineExtraction.cpp:129-289, 291-322; inputOutput.cpp:48-64):
// loop removal by a spanning tree over the graph with its degree-two nodes
// fused; the voxels (and averaged radius) along each edge, ceil|ab| samples;
// the VTK polydata file.  createCenterlineVoxels returns new[]'d memory.
void removeLoops(std::vector<int3>& vertices, std::vector<int2>& edges, int3& size);
char* createCenterlineVoxels(std::vector<int3>& vertices, std::vector<int2>& edges, float* radius, int3& size);
void writeToVtkFile(paramList& parameters, std::vector<int3> vertices, std::vector<int2> edges);

// segmentation.hpp:8-10: the segmentation from the centerline mask (one byte
// per voxel, as runNewCenterlineAlg returns it), the GVF field (with w) and
// the radius volume; returns the mask (one byte per voxel, 0/1) as a device
// Buffer.  Inverse-gradient region growing with a 3x3x3 closing (the
// default), or balls of the radius around every centerline voxel
// ("sphere-segmentation").
Buffer runInverseGradientSegmentation(OpenCL& ocl, Buffer& centerline, Image3D& vectorField, Image3D& radius,
                                      int3 size, paramList parameters);
Buffer runSphereSegmentation(OpenCL& ocl, Buffer& centerline, Image3D& radius, int3 size, paramList parameters);

// ----------------------------------------------------------------- entry point, ingest and output
// TSFOutput (inputOutput.hpp:10-52): the results of run().  It owns its
// context (one stream), the size, and every volume set on it; the device
// volumes are copied to the host on the first get*() (new[]'d arrays that
// the output frees).  Centerline and segmentation are one byte per voxel.
class TSFOutput {
   public:
    explicit TSFOutput(int3* size, bool TDFis16bit = false);
    ~TSFOutput();
    TSFOutput(const TSFOutput&) = delete;
    TSFOutput& operator=(const TSFOutput&) = delete;
    bool hasSegmentation() const { return oclSegmentation || segmentation; }
    bool hasCenterlineVoxels() const { return oclCenterlineVoxels || centerlineVoxels; }
    bool hasTDF() const { return oclTDF || TDF; }
    void setTDF(Image3D* image);
    void setSegmentation(Buffer* mask);
    void setCenterlineVoxels(Buffer* mask);
    void setTDF(float* data);
    void setSegmentation(char* data);
    void setCenterlineVoxels(char* data);
    void setSize(int3* size);
    char* getSegmentation();
    char* getCenterlineVoxels();
    float* getTDF();
    int3* getSize() { return size; }
    int3 getShiftVector() const { return shiftVector; }
    void setShiftVector(int3 v) { shiftVector = v; }
    float3 getSpacing() const { return spacing; }
    void setSpacing(float3 v) { spacing = v; }
    OpenCL& getContext() { return ocl; }
    bool isTDF16bit() const { return TDFis16bit; }

   private:
    OpenCL ocl;
    int3* size;
    bool TDFis16bit;
    int3 shiftVector;
    float3 spacing{1.0f, 1.0f, 1.0f};
    Image3D* oclTDF = nullptr;
    Buffer* oclSegmentation = nullptr;
    Buffer* oclCenterlineVoxels = nullptr;
    float* TDF = nullptr;
    char* segmentation = nullptr;
    char* centerlineVoxels = nullptr;
};

// tube-segmentation.hpp:47 (tube-segmentation.cpp:53-163): reads the .mhd
// volume, runs the centerline method ("gpu": circle fitting, the centerline
// graph and the segmentation; "ridge": circle fitting, the ridge traversal
// and the segmentation) and returns the results.  kernel_dir is accepted for
// the reference's signature; the kernels are in the library.  Throws
// IOException for a missing file, tsf::Error otherwise (the experimental
// "test" method is not provided).
TSFOutput* run(std::string filename, paramList& parameters, std::string kernel_dir);

// tube-segmentation.cpp:1399-1905: parses the .mhd header (ElementType,
// ElementDataFile, DimSize, ElementSpacing), uploads the raw volume, finds
// the intensity limits ("minimum" / "maximum", or the volume's own), crops
// ("cropping" lung / threshold by scan lines, the AAA preset's 15 % margin,
// else each extent shrunk to a multiple of 4) and converts the region to a
// float volume in [0, 1].  *size becomes the cropped size; the shift vector
// and spacing are set on output (when not NULL).
Image3D readDatasetAndTransfer(OpenCL& ocl, std::string filename, paramList& parameters, int3* size,
                               TSFOutput* output);

// tube-segmentation.cpp:943-986: runCircleFittingMethod, runNewCenterlineAlg
// and the segmentation, results set on output; storage-dir writes them.
void runCircleFittingAndNewCenterlineAlg(OpenCL& ocl, Image3D* dataset, int3* size, paramList& parameters,
                                         TSFOutput* output);

// inputOutput.cpp:14-47: <dir><name>.centerline.mhd/.raw and
// <dir><name>.segmentation.mhd/.raw (MET_CHAR) for the results present.
void writeDataToDisk(TSFOutput* output, std::string storageDirectory, std::string name);

// eigenanalysisOfHessian.hpp:7-9 on the host arrays of TubeSegmentation
// (tube-segmentation.hpp:22-30).  The arrays are mirrored to the device on
// first use; pos must lie one voxel inside every border.
struct TubeSegmentation {
    float *Fx = nullptr, *Fy = nullptr, *Fz = nullptr;
    float *FxSmall = nullptr, *FySmall = nullptr, *FzSmall = nullptr;
    float* TDF = nullptr;
    float* radius = nullptr;
    char* centerline = nullptr;
    char* segmentation = nullptr;
    float* intensity = nullptr;
};
// eigenanalysisOfHessian.hpp:7-9.  Host code, as in the reference: each call
// reads T's current host arrays (no device copy to go stale).
void doEigen(TubeSegmentation& T, int3 pos, int3 size, float3* lambda, float3* e1, float3* e2, float3* e3);
float3 getTubeDirection(TubeSegmentation& T, int3 pos, int3 size);
// Many points on the host (12 floats each: lambda, e1, e2, e3).
void doEigenBatch(TubeSegmentation& T, int3 size, const std::vector<int3>& pos, std::vector<float>& out12);
// A device copy of T's arrays, owned by the caller, for large batches
// (tsf_do_eigen_points): the copy is taken at construction and by
// refresh(), which a caller that changes the host arrays calls; batch()
// reuses its point buffers, grown as needed.  Not thread-safe.
class EigenField {
   public:
    EigenField(OpenCL& ocl, const TubeSegmentation& T, int3 size);
    void refresh(const TubeSegmentation& T);
    void batch(const std::vector<int3>& pos, std::vector<float>& out12);

   private:
    OpenCL& ocl_;
    int3 size_;
    Buffer fx_, fy_, fz_, points_, results_;
    size_t cap_ = 0;
};

// ridgeTraversalCenterlineExtraction.hpp:9-15: the ridge-traversal method.
// Returns the centerline mask (new[]'d, one byte per voxel): the longest
// traversed tree and every tree longer than min-tree-length.  It sets voxel
// 0 of T.TDF / T.Fx / T.Fy / T.Fz to (0, 1, 0, 0), as the reference does,
// and reads the Hessian through doEigen.  centerlineStack is taken by value,
// as in the reference.  Throws tsf::Error (TSF_ERR_RANGE) when no start point
// exists.
struct CenterlinePoint {
    int3 pos;
    bool large = false;
    CenterlinePoint* next = nullptr;
};
char* runRidgeTraversal(TubeSegmentation& T, int3 size, paramList& parameters,
                        std::stack<CenterlinePoint> centerlineStack);
// tube-segmentation.cpp:1260-1350: runCircleFittingMethod, the traversal on
// the host copies of the field, TDF and radius, then the segmentation.
void runCircleFittingAndRidgeTraversal(OpenCL& ocl, Image3D* dataset, int3* size, paramList& parameters,
                                       TSFOutput* output);

}  // namespace tsf
