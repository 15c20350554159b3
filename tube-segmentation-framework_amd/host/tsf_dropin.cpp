// C++ drop-in layer over the C ABI: see tsf_dropin.hpp.
#include "tsf_dropin.hpp"

#include "../csrc/eigen3.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <locale>
#include <sstream>

namespace tsf {

void check(tsf_status st) {
    if (st != TSF_OK) throw Error(st, tsf_last_error());
}

static void fail(tsf_status st, const std::string& what) { throw Error(st, what); }

// ----------------------------------------------------------------- parameters
BoolParameter::BoolParameter(bool v, std::string d, std::string g) : value(v), description(d), group(g) {}
NumericParameter::NumericParameter(float v, float mn, float mx, float st, std::string d, std::string g)
    : value(v), min(mn), max(mx), step(st), description(d), group(g) {}
void NumericParameter::set(float v) {
    if (!validate(v)) fail(TSF_ERR_INVALID_ARGUMENT, "numeric parameter value out of range");
    value = v;
}
StringParameter::StringParameter(std::string v, std::vector<std::string> p, std::string d, std::string g)
    : value(v), possibilities(p), description(d), group(g) {}
bool StringParameter::validate(const std::string& v) const {
    return possibilities.empty() || std::find(possibilities.begin(), possibilities.end(), v) != possibilities.end();
}
void StringParameter::set(const std::string& v) {
    if (!validate(v)) fail(TSF_ERR_INVALID_ARGUMENT, "string parameter value not allowed: " + v);
    value = v;
}

static float to_float_c(const std::string& s) {
    float v = 0.0f;
    std::istringstream is(s);
    is.imbue(std::locale("C"));
    is >> v;
    return v;
}

paramList initParameters() {
    // name, default, range or allowed values, group: the values of the
    // reference's parameters/parameters file for every key it defines.
    paramList p;
    auto B = [&](const char* n, bool v, const char* g) { p.bools[n] = BoolParameter(v, "", g); };
    auto N = [&](const char* n, float v, float mn, float mx, float st, const char* g) {
        p.numerics[n] = NumericParameter(v, mn, mx, st, "", g);
    };
    auto S = [&](const char* n, const char* v, std::vector<std::string> ok, const char* g) {
        p.strings[n] = StringParameter(v, ok, "", g);
    };
    S("device", "gpu", {"gpu", "cpu"}, "general");
    S("mode", "white", {"black", "white"}, "general");
    B("display", false, "advanced");
    S("centerline-method", "gpu", {"test", "gpu", "ridge"}, "general");
    N("gvf-iterations", 250, 0, 10000, 50, "gradient-vector-flow");
    N("radius-min", 0.5f, 0.5f, 50.0f, 0.5f, "tube-detection-filter");
    N("radius-max", 15.0f, 2.0f, 300.0f, 1.0f, "tube-detection-filter");
    N("radius-step", 1.0f, 0.0f, 5.0f, 0.5f, "tube-detection-filter");
    N("fmax", 0.2f, 0.01f, 0.9f, 0.01f, "general");
    N("gvf-mu", 0.05f, 0.0f, 0.5f, 0.01f, "gradient-vector-flow");
    N("small-blur", 0.0f, 0.0f, 5.0f, 0.5f, "general");
    N("large-blur", 1.0f, 0.0f, 15.0f, 0.5f, "general");
    N("tdf-high", 0.5f, 0.1f, 1.0f, 0.1f, "centerline-general");
    N("min-distance", 5, 0, 100, 1, "centerline-ridge");
    N("m-low", 0.05f, 0.0f, 0.2f, 0.01f, "centerline-ridge");
    N("tdf-low", 0.5f, 0.0f, 1.0f, 0.1f, "centerline-ridge");
    N("max-below-tdf-low", 0, 0, 5, 1, "centerline-ridge");
    N("min-mean-tdf", 0.5f, 0.0f, 1.0f, 0.01f, "centerline-general");
    N("min-tree-length", 5, 0, 1000, 5, "centerline-general");
    B("timing", false, "advanced");
    N("cube-size", 4, 0, 10, 1, "centerline-gpu");
    N("max-distance", 25.0f, 0.0f, 50.0f, 1.0f, "centerline-gpu");
    B("centerpoints-only", false, "centerline-gpu");
    B("tdf-only", false, "tube-detection-filter");
    B("no-segmentation", false, "general");
    S("centerline-vtk-file", "off", {}, "storage");
    B("sphere-segmentation", false, "general");
    S("minimum", "off", {}, "general");
    S("maximum", "off", {}, "general");
    S("cropping", "no", {"no", "lung", "threshold"}, "cropping");
    N("min-scan-lines-threshold", 10, 0, 1024, 1, "cropping");
    N("min-scan-lines-lung", 200, 0, 1024, 1, "cropping");
    N("cropping-threshold", 0, 0, 3000, 10, "cropping");
    S("cropping-start-z", "end", {"end", "middle"}, "cropping");
    B("buffers-only", false, "advanced");
    S("storage-dir", "off", {}, "storage");
    S("storage-name", "unnamed", {}, "storage");
    B("32bit-vectors", false, "advanced");
    B("16bit-vectors", true, "advanced");
    S("parameters", "none", {}, "preset");
    B("loop-removal", true, "advanced");
    B("timer-total", false, "advanced");
    N("max-edge-distance", 3, 2, 30, 1, "centerline-gpu");
    B("use-spline-tdf", false, "tube-detection-filter");
    B("use-fmg-gvf", false, "gradient-vector-flow");
    B("3d_write", true, "advanced");
    S("vector-storage", "auto", {"auto", "f32", "f16", "bf16", "sn16", "precision"}, "advanced");
    N("gvf-fuse", 2, 1, 16, 1, "gradient-vector-flow");
    N("gvf-f32-tail", 32, 0, 10000, 1, "gradient-vector-flow");
    return p;
}

paramList initParameters(const std::string& dir) {
    paramList p = initParameters();
    std::ifstream file(dir + "/parameters");
    if (!file.is_open()) fail(TSF_ERR_INVALID_ARGUMENT, "cannot open " + dir + "/parameters");
    std::string line;
    std::getline(file, line);  // comment line
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        std::istringstream is(line);
        std::string name, type, def;
        is >> name >> type >> def;
        if (type == "bool") {
            p.bools[name].set(def == "true");
        } else if (type == "num") {
            float mn, mx, st;
            std::string a, b, c;
            is >> a >> b >> c;
            mn = to_float_c(a), mx = to_float_c(b), st = to_float_c(c);
            p.numerics[name] = NumericParameter(to_float_c(def), mn, mx, st, "", "");
        } else if (type == "str") {
            std::vector<std::string> ok;
            std::string tok;
            while (is >> tok && tok[0] != '"') ok.push_back(tok);
            p.strings[name] = StringParameter(def, ok, "", "");
        } else {
            fail(TSF_ERR_INVALID_ARGUMENT, "Could not parse parameter of type: " + type);
        }
    }
    return p;
}

void setParameter(paramList& p, const std::string& name, const std::string& value) {
    if (p.bools.count(name)) {
        p.bools[name].set(value == "true");
    } else if (p.numerics.count(name)) {
        p.numerics[name].set(to_float_c(value));
    } else if (p.strings.count(name)) {
        if (name == "parameters") p.strings[name].setWithoutValidation(value);
        else p.strings[name].set(value);
    } else {
        fail(TSF_ERR_INVALID_ARGUMENT, "Can not set value for parameter with name: " + name);
    }
}

void loadParameterPreset(paramList& p, const std::string& dir) {
    if (getParamStr(p, "parameters") == "none") return;
    const std::string method = getParamStr(p, "centerline-method");
    const std::string fn = dir + "/centerline-" + method + "/" + getParamStr(p, "parameters");
    std::ifstream file(fn);
    if (!file.is_open()) fail(TSF_ERR_INVALID_ARGUMENT, "cannot open preset " + fn);
    std::string line;
    while (std::getline(file, line)) {
        if (line.empty()) continue;
        const size_t sp = line.find(' ');
        if (sp != std::string::npos) setParameter(p, line.substr(0, sp), line.substr(sp + 1));
        else setParameter(p, line, "true");
    }
}

float getParam(const paramList& p, const std::string& n) {
    auto it = p.numerics.find(n);
    if (it == p.numerics.end()) fail(TSF_ERR_INVALID_ARGUMENT, "numeric parameter not found: " + n);
    return it->second.get();
}
bool getParamBool(const paramList& p, const std::string& n) {
    auto it = p.bools.find(n);
    if (it == p.bools.end()) fail(TSF_ERR_INVALID_ARGUMENT, "bool parameter not found: " + n);
    return it->second.get();
}
std::string getParamStr(const paramList& p, const std::string& n) {
    auto it = p.strings.find(n);
    if (it == p.strings.end()) fail(TSF_ERR_INVALID_ARGUMENT, "string parameter not found: " + n);
    return it->second.get();
}

// ----------------------------------------------------------------- context and memory
OpenCL::OpenCL() { check(tsf_stream_create(&queue)); }
OpenCL::~OpenCL() {
    if (queue) (void)tsf_stream_destroy(queue);
}
void OpenCL::finish() { check(tsf_stream_synchronize(queue)); }

Buffer::Buffer(size_t n) : bytes(n) {
    void* p = nullptr;
    check(tsf_malloc(&p, n ? n : 1));
    mem = std::shared_ptr<void>(p, [](void* q) { (void)tsf_free(q); });
}
void Buffer::upload(OpenCL& ocl, const void* host, size_t n) {
    if (n > bytes) fail(TSF_ERR_INVALID_ARGUMENT, "Buffer::upload larger than the buffer");
    check(tsf_memcpy_htod(get(), host, n, ocl.queue));
}
void Buffer::download(OpenCL& ocl, void* host, size_t n) const {
    if (n > bytes) fail(TSF_ERR_INVALID_ARGUMENT, "Buffer::download larger than the buffer");
    check(tsf_memcpy_dtoh(host, get(), n, ocl.queue));
}

static size_t elt(tsf_storage s) { return s == TSF_STORAGE_F32 ? 4 : 2; }
static size_t voxels(int3 s) { return (size_t)s.x * s.y * s.z; }

Image3D::Image3D(int3 size, int channels, tsf_storage storage, float scale)
    : sz(size), nch(channels), st(storage), sc(scale) {
    if (size.x <= 0 || size.y <= 0 || size.z <= 0 || channels < 1)
        fail(TSF_ERR_INVALID_ARGUMENT, "Image3D: bad size or channel count");
    for (int c = 0; c < channels; c++) planes.emplace_back(voxels(size) * elt(storage));
}
tsf_field Image3D::field() const {
    if (nch != 4) fail(TSF_ERR_INVALID_ARGUMENT, "Image3D is not a vector field");
    return tsf_field{plane(0), plane(1), plane(2), plane(3), (int32_t)st, sc};
}

// The reference's 16-bit vector format (kernels_no_3d_write.cl:9-21).
static int16_t f32_to_sn16(float v) {
    const float x = v * 32767.0f;
    if (x != x) return 0;
    const float r = std::nearbyint(x);
    return (int16_t)(r > 32767.0f ? 32767 : (r < -32768.0f ? -32768 : (int)r));
}
static float sn16_to_f32(int16_t s) { return std::max(-1.0f, (float)s / 32767.0f); }

static uint16_t f32_to_f16(float f) {
    // IEEE half, round to nearest even (host side conversion for uploads)
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    int32_t e = (int32_t)((u >> 23) & 0xff) - 127 + 15;
    uint32_t m = u & 0x7fffffu;
    if (((u >> 23) & 0xff) == 0xff) return (uint16_t)(sign | 0x7c00u | (m ? 0x200u : 0));
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        m |= 0x800000u;
        const int shift = 14 - e;
        uint32_t r = m >> shift, rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (r & 1))) r++;
        return (uint16_t)(sign | r);
    }
    uint32_t r = ((uint32_t)e << 10) | (m >> 13), rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1))) r++;
    return (uint16_t)(sign | r);
}
static float f16_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    int e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ffu, u;
    if (e == 0) {
        if (m == 0) {
            u = sign;
        } else {
            e = 1;
            while (!(m & 0x400u)) {
                m <<= 1;
                e--;
            }
            m &= 0x3ffu;
            u = sign | ((uint32_t)(e - 15 + 127) << 23) | (m << 13);
        }
    } else if (e == 31) {
        u = sign | 0x7f800000u | (m << 13);
    } else {
        u = sign | ((uint32_t)(e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
static uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static float bf16_to_f32(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void Image3D::uploadPlane(OpenCL& ocl, int c, const float* host) {
    const size_t n = voxels(sz);
    if (st == TSF_STORAGE_F32) {
        planes.at(c).upload(ocl, host, n * 4);
        return;
    }
    std::vector<uint16_t> t(n);
    for (size_t i = 0; i < n; i++)
        t[i] = st == TSF_STORAGE_F16    ? f32_to_f16(host[i] * sc)
               : st == TSF_STORAGE_SN16 ? (uint16_t)f32_to_sn16(host[i])
                                        : f32_to_bf16(host[i] * sc);
    planes.at(c).upload(ocl, t.data(), n * 2);
}
std::vector<float> Image3D::downloadPlane(OpenCL& ocl, int c) const {
    const size_t n = voxels(sz);
    std::vector<float> out(n);
    if (st == TSF_STORAGE_F32) {
        planes.at(c).download(ocl, out.data(), n * 4);
        return out;
    }
    std::vector<uint16_t> t(n);
    planes.at(c).download(ocl, t.data(), n * 2);
    for (size_t i = 0; i < n; i++)
        out[i] = st == TSF_STORAGE_SN16 ? sn16_to_f32((int16_t)t[i])
                                        : (st == TSF_STORAGE_F16 ? f16_to_f32(t[i]) : bf16_to_f32(t[i])) / sc;
    return out;
}

bool precisionForm(const paramList& p) { return getParamStr(p, "vector-storage") == "precision"; }

tsf_storage vectorStorage(const paramList& p) {
    const std::string s = getParamStr(p, "vector-storage");
    if (s == "precision") return TSF_STORAGE_SN16;  // the large branch's f; the small branch's field is F32
    if (s == "f32") return TSF_STORAGE_F32;
    if (s == "f16") return TSF_STORAGE_F16;
    if (s == "bf16") return TSF_STORAGE_BF16;
    if (s == "sn16") return TSF_STORAGE_SN16;
    // "auto": the reference's own 16-bit format (SNORM16) unless 32-bit
    // vectors are forced
    return getParamBool(p, "16bit-vectors") && !getParamBool(p, "32bit-vectors") ? TSF_STORAGE_SN16 : TSF_STORAGE_F32;
}
float vectorScale(tsf_storage s) { return s == TSF_STORAGE_F16 ? 16384.0f : 1.0f; }

// ----------------------------------------------------------------- the path
float* createBlurMask(float sigma, int* maskSizePointer) {
    float* mask = new float[11 * 11 * 11];
    int32_t m = 1;
    tsf_status st = tsf_create_blur_mask(sigma, mask, &m);
    if (st != TSF_OK) {
        delete[] mask;
        check(st);
    }
    *maskSizePointer = m;
    return mask;
}

Image3D runGVF(OpenCL& ocl, Image3D* vectorField, paramList& parameters, int3& size, bool useLessMemory) {
    if (!vectorField || !vectorField->valid()) fail(TSF_ERR_INVALID_ARGUMENT, "runGVF: no input field");
    const int iters = (int)getParam(parameters, "gvf-iterations");
    const float mu = getParam(parameters, "gvf-mu");
    Image3D in = *vectorField;
    tsf_field f = in.field();
    Image3D out;
    if (precisionForm(parameters) && in.storage() == TSF_STORAGE_SN16) {
        // the precision form: 16-bit iterate with stochastic rounding, the
        // last gvf-f32-tail iterations in float, the result F32
        out = Image3D(size, 4);
        tsf_field o = out.field();
        const int tail = (int)getParam(parameters, "gvf-f32-tail");
        Buffer ws(tsf_gvf_euler_sr_workspace_size(out.dims(), tail));
        check(tsf_gvf_euler_sr(&f, &o, out.dims(), mu, 2 * (iters / 2), tail, 0u, ws.get(), ws.size(), ocl.queue));
    } else if (useLessMemory && (in.storage() == TSF_STORAGE_F32 || in.storage() == TSF_STORAGE_SN16)) {
        // runLowMemoryGVF: one component at a time (gradientVectorFlow.cpp:1173-1445)
        out = Image3D(size, 4, in.storage(), in.scale());
        tsf_field o = out.field();
        Buffer ws(tsf_gvf_euler_low_memory_workspace_size(out.dims(), out.storage()));
        check(tsf_gvf_euler_low_memory(&f, &o, out.dims(), mu, 2 * (iters / 2), ws.get(), ws.size(), ocl.queue));
    } else {
        if (useLessMemory)
            std::fprintf(stderr, "NOTE: the one-component GVF takes F32 or SNORM16 fields; running the full-field "
                                 "GVF on this %s field\n", in.storage() == TSF_STORAGE_F16 ? "F16" : "BF16");
        out = Image3D(size, 4, in.storage(), in.scale());
        tsf_field o = out.field();
        Buffer ws(tsf_gvf_euler_workspace_size(out.dims(), out.storage()));
        check(tsf_gvf_euler(&f, &o, out.dims(), mu, 2 * (iters / 2), (int)getParam(parameters, "gvf-fuse"), ws.get(),
                            ws.size(), ocl.queue));
    }
    ocl.finish();
    *vectorField = Image3D();  // the reference deletes the input field
    return out;
}

// The reference runs the one-component GVF when a whole 4-channel field
// exceeds the device's largest allocation (tube-segmentation.cpp:743-762).
// HIP allocations are bounded by free memory instead: the one-component
// path is taken when the full-field one (the output field and three
// workspace planes) does not fit in it.
static bool gvfNeedsLowMemory(const Image3D& F, const paramList& p) {
    if (precisionForm(p)) return false;
    size_t freeB = 0, total = 0;
    check(tsf_mem_info(&freeB, &total));
    const size_t e = F.storage() == TSF_STORAGE_F32 ? 4 : 2;
    return 7 * e * voxels(F.size()) > freeB;
}

Image3D runFMGGVF(OpenCL& ocl, Image3D* vectorField, paramList& parameters, int3& size) {
    if (!vectorField || !vectorField->valid()) fail(TSF_ERR_INVALID_ARGUMENT, "runFMGGVF: no input field");
    const int iters = (int)getParam(parameters, "gvf-iterations");
    const float mu = getParam(parameters, "gvf-mu");
    Image3D in = *vectorField;
    if (in.storage() != TSF_STORAGE_F32) {
        // the multigrid solver runs in float planes
        Image3D f32(size, 4);
        for (int c = 0; c < 4; c++) f32.uploadPlane(ocl, c, in.downloadPlane(ocl, c).data());
        in = f32;
    }
    Image3D out(size, 4);
    tsf_field f = in.field(), o = out.field();
    Buffer ws(tsf_gvf_fmg_workspace_size(out.dims(), 2));
    // v0 = 1, v1 = v2 = 2, l_max = 2 (gradientVectorFlow.cpp:747-751)
    check(tsf_gvf_fmg(&f, &o, out.dims(), mu, iters, 2, 1, 2, 2, nullptr, ws.get(), ws.size(), ocl.queue));
    ocl.finish();
    *vectorField = Image3D();
    return out;
}

// The TDF of one field.  With the reference's 16-bit vectors the TDF buffers
// are 16-bit too (TDF_TYPE ushort, kernels_no_3d_write.cl:19-21): an SN16
// field gets its TDF rounded to that format.  Synchronises (the workspace is
// released on return).
static void circleTDF(OpenCL& ocl, const Image3D& field, float rmin, float rmax, float rstep, float* T, float* R,
                      bool mixed = false) {
    tsf_field f = field.field();
    if (mixed && field.storage() == TSF_STORAGE_F32) {
        // the precision form's TDF: Hessian from the F32 planes, samples from F16 texels
        Buffer ws(tsf_tdf_circle_mixed_workspace_size(field.dims()));
        check(tsf_tdf_circle_mixed(&f, TSF_STORAGE_F16, 16384.0f, field.dims(), rmin, rmax, rstep, T, R, ws.get(),
                                   ws.size(), ocl.queue));
        ocl.finish();
        return;
    }
    Buffer ws(tsf_tdf_circle_workspace_size(field.dims(), field.storage()));
    check(tsf_tdf_circle(&f, field.dims(), rmin, rmax, rstep, T, R, ws.get(), ws.size(), ocl.queue));
    if (field.storage() == TSF_STORAGE_SN16) check(tsf_tdf_round_unorm16(T, (int64_t)voxels(field.size()), ocl.queue));
    ocl.finish();
}

void runCircleFittingTDF(OpenCL& ocl, int3& size, Image3D* vectorField, Buffer* TDF, Buffer* radius, float radiusMin,
                         float radiusMax, float radiusStep) {
    if (!vectorField || !TDF || !radius) fail(TSF_ERR_INVALID_ARGUMENT, "runCircleFittingTDF: NULL argument");
    const size_t n = voxels(size);
    if (TDF->size() < n * 4 || radius->size() < n * 4)
        fail(TSF_ERR_INVALID_ARGUMENT, "runCircleFittingTDF: output buffers too small");
    circleTDF(ocl, *vectorField, radiusMin, radiusMax, radiusStep, (float*)TDF->get(), (float*)radius->get());
}

// runSplineTDF (tubeDetectionFilters.cpp:4-55): rMin raised to 1, 12 arms,
// minimum average magnitude 0.1; a 16-bit run stores the TDF as UNORM16.
static void splineTDF(OpenCL& ocl, const Image3D& field, float rmin, float rmax, float rstep, float* T, float* R) {
    tsf_field f = field.field();
    Buffer ws(tsf_tdf_spline_workspace_size(field.dims(), field.storage()));
    check(tsf_tdf_spline(&f, field.dims(), std::max(1.0f, rmin), rmax, rstep, 0.1f, T, R, ws.get(), ws.size(),
                         ocl.queue));
    if (field.storage() == TSF_STORAGE_SN16) check(tsf_tdf_round_unorm16(T, (int64_t)voxels(field.size()), ocl.queue));
    ocl.finish();
}

void runSplineTDF(OpenCL& ocl, int3& size, Image3D* vectorField, Buffer* TDF, Buffer* radius, float radiusMin,
                  float radiusMax, float radiusStep) {
    if (!vectorField || !TDF || !radius) fail(TSF_ERR_INVALID_ARGUMENT, "runSplineTDF: NULL argument");
    const size_t n = voxels(size);
    if (TDF->size() < n * 4 || radius->size() < n * 4)
        fail(TSF_ERR_INVALID_ARGUMENT, "runSplineTDF: output buffers too small");
    splineTDF(ocl, *vectorField, radiusMin, radiusMax, radiusStep, (float*)TDF->get(), (float*)radius->get());
}

static Image3D blurOrSame(OpenCL& ocl, Image3D& data, float sigma) {
    if (!(sigma > 0)) return data;
    Image3D out(data.size(), 1);
    check(tsf_blur_gaussian((const float*)data.plane(0), (float*)out.plane(0), data.dims(), sigma, ocl.queue));
    return out;
}

static Image3D vectorFieldOf(OpenCL& ocl, Image3D& vol, paramList& p, bool large) {
    // the precision form keeps the small branch's field in F32 (its TDF takes
    // the Hessian from it) and stores the large branch's, the GVF's f, as
    // SNORM16; on volumes the 16-bit GVF does not take (odd or small X, ...)
    // the large branch stays F32 and runs the FP32 GVF
    tsf_storage s = precisionForm(p) && !large ? TSF_STORAGE_F32 : vectorStorage(p);
    if (precisionForm(p) && large && !tsf_gvf_euler_sr_supported(vol.dims())) {
        std::fprintf(stderr, "NOTE: the precision form's 16-bit GVF needs X even and >= 64, Y >= 32, Z >= 4; "
                             "running the FP32 GVF on this %dx%dx%d volume\n",
                     vol.size().x, vol.size().y, vol.size().z);
        s = TSF_STORAGE_F32;
    }
    Image3D F(vol.size(), 4, s, vectorScale(s));
    tsf_field f = F.field();
    const int sign = getParamStr(p, "mode") == "black" ? -1 : 1;
    check(tsf_create_vector_field((const float*)vol.plane(0), &f, vol.dims(), getParam(p, "fmax"), sign, ocl.queue));
    return F;
}

void runCircleFittingMethod(OpenCL& ocl, Image3D* dataset, int3 size, paramList& p, Image3D& vectorField,
                            Image3D& TDF, Image3D& radiusImage) {
    if (!dataset || !dataset->valid() || dataset->channels() != 1 || dataset->storage() != TSF_STORAGE_F32)
        fail(TSF_ERR_INVALID_ARGUMENT, "runCircleFittingMethod: dataset must be a float scalar volume");
    const float rMin = getParam(p, "radius-min"), rMax = getParam(p, "radius-max"), rStep = getParam(p, "radius-step");
    const size_t n = voxels(size);
    Image3D tdfSmall, radSmall;
    if (rMin < 2.5f) {
        Image3D blurred = blurOrSame(ocl, *dataset, getParam(p, "small-blur"));
        Image3D Fs = vectorFieldOf(ocl, blurred, p, false);
        tdfSmall = Image3D(size, 1);
        radSmall = Image3D(size, 1);
        circleTDF(ocl, Fs, rMin, 3.0f, 0.5f, (float*)tdfSmall.plane(0), (float*)radSmall.plane(0), precisionForm(p));
        if (rMax < 2.5f) {
            TDF = tdfSmall;
            radiusImage = radSmall;
            vectorField = Fs;
            *dataset = Image3D();
            return;
        }
    }
    Image3D blurred = blurOrSame(ocl, *dataset, getParam(p, "large-blur"));
    Image3D F = vectorFieldOf(ocl, blurred, p, true);
    ocl.finish();
    *dataset = Image3D();
    blurred = Image3D();
    if (getParamBool(p, "use-fmg-gvf")) vectorField = runFMGGVF(ocl, &F, p, size);
    else vectorField = runGVF(ocl, &F, p, size, gvfNeedsLowMemory(F, p));
    TDF = Image3D(size, 1);
    radiusImage = Image3D(size, 1);
    // the large-radius TDF: spline (from 1.5) or circle fitting (from 2.5),
    // tube-segmentation.cpp:785-789
    if (getParamBool(p, "use-spline-tdf"))
        splineTDF(ocl, vectorField, std::max(1.5f, rMin), rMax, rStep, (float*)TDF.plane(0),
                  (float*)radiusImage.plane(0));
    else
        circleTDF(ocl, vectorField, std::max(2.5f, rMin), rMax, rStep, (float*)TDF.plane(0),
                  (float*)radiusImage.plane(0), precisionForm(p));
    if (rMin < 2.5f)
        check(tsf_combine((const float*)tdfSmall.plane(0), (const float*)radSmall.plane(0), (float*)TDF.plane(0),
                          (float*)radiusImage.plane(0), (int64_t)n, ocl.queue));
    ocl.finish();
}

CenterpointVertices runNewCenterlineAlgCandidates(OpenCL& ocl, int3 size, paramList& p, Image3D& vectorField,
                                                  Image3D& TDF, Image3D& radius) {
    const size_t n = voxels(size);
    const tsf_dims d{size.x, size.y, size.z};
    CenterpointVertices R;
    Buffer mask(n);
    check(tsf_candidates((const float*)TDF.plane(0), (uint8_t*)mask.get(), (int64_t)n, getParam(p, "tdf-high"),
                         ocl.queue));
    Buffer ws(tsf_hp_compact_workspace_size(d));
    int64_t sum = 0;
    check(tsf_hp_compact((const uint8_t*)mask.get(), d, nullptr, 0, &sum, ws.get(), ws.size(), ocl.queue));
    if (sum <= 0 || sum > 0.5 * (double)n)
        fail(TSF_ERR_RANGE, "The number of candidate voxels is too low or too high. Something went wrong... Wrong "
                            "parameters? Out of memory?");
    R.candidates = (int)sum;
    Buffer pos((size_t)sum * 12);
    check(tsf_hp_compact((const uint8_t*)mask.get(), d, (int32_t*)pos.get(), sum, &sum, ws.get(), ws.size(),
                         ocl.queue));
    R.centerpoints = Buffer(n);
    check(tsf_memset(R.centerpoints.get(), 0, n, ocl.queue));
    tsf_field f = vectorField.field();
    const int cap = getParamBool(p, "3d_write") ? 8 : 5;  // kernels.cl:1334 / kernels_no_3d_write.cl:1043
    check(tsf_candidate_nms((const float*)TDF.plane(0), (const float*)radius.plane(0), &f, d, (const int32_t*)pos.get(),
                            sum, cap, (uint8_t*)R.centerpoints.get(), ocl.queue));
    if (getParamBool(p, "centerpoints-only")) {
        ocl.finish();
        return R;
    }
    Buffer cells(n);
    check(tsf_memset(cells.get(), 0, n, ocl.queue));
    check(tsf_cube_max((const float*)TDF.plane(0), (const uint8_t*)R.centerpoints.get(), d,
                       (int)getParam(p, "cube-size"), (uint8_t*)cells.get(), ocl.queue));
    int64_t nv = 0;
    check(tsf_hp_compact((const uint8_t*)cells.get(), d, nullptr, 0, &nv, ws.get(), ws.size(), ocl.queue));
    R.vertices = Buffer((size_t)std::max<int64_t>(nv, 1) * 12);
    check(tsf_hp_compact((const uint8_t*)cells.get(), d, (int32_t*)R.vertices.get(), nv, &nv, ws.get(), ws.size(),
                         ocl.queue));
    R.count = (int)nv;
    if (nv < 8) fail(TSF_ERR_RANGE, "Too few centerpoints detected. Revise parameters.");
    if (nv >= 16384) fail(TSF_ERR_RANGE, "Too many centerpoints detected. More cropping of dataset is probably needed.");
    return R;
}

// ----------------------------------------------------------------- doEigen
// doEigen (eigenanalysisOfHessian.cpp:264-389) is host code in the reference:
// it reads T's host arrays at every call.  So is this one -- the 6 samples
// divided by sqrt(Fx^2 + Fy^2 + Fz^2) of their voxel, central differences,
// the register eigensolver of the device kernels (eigen3.hpp, the same
// operation order, no contraction) -- so a caller that changes the arrays in
// place sees the change at the next call.  Positions whose neighbours leave
// the volume (the reference reads outside its arrays there) give NaN.
static void eigen_at(const float* Fx, const float* Fy, const float* Fz, int3 size, int3 p, float o[12]) {
    if (p.x < 1 || p.y < 1 || p.z < 1 || p.x > size.x - 2 || p.y > size.y - 2 || p.z > size.z - 2) {
        for (int k = 0; k < 12; k++) o[k] = std::nanf("");
        return;
    }
    auto nv = [&](const float* c, int x, int y, int z) -> float {
        const size_t i = (size_t)x + (size_t)size.x * ((size_t)y + (size_t)size.y * (size_t)z);
        return c[i] / std::sqrt(Fx[i] * Fx[i] + Fy[i] * Fy[i] + Fz[i] * Fz[i]);
    };
    const int x = p.x, y = p.y, z = p.z;
    const float fxx = 0.5f * (nv(Fx, x + 1, y, z) - nv(Fx, x - 1, y, z));
    const float fyx = 0.5f * (nv(Fy, x + 1, y, z) - nv(Fy, x - 1, y, z));
    const float fyy = 0.5f * (nv(Fy, x, y + 1, z) - nv(Fy, x, y - 1, z));
    const float fzx = 0.5f * (nv(Fz, x + 1, y, z) - nv(Fz, x - 1, y, z));
    const float fzy = 0.5f * (nv(Fz, x, y + 1, z) - nv(Fz, x, y - 1, z));
    const float fzz = 0.5f * (nv(Fz, x, y, z + 1) - nv(Fz, x, y, z - 1));
    const float H[3][3] = {{fxx, fyx, fzx}, {fyx, fyy, fzy}, {fzx, fzy, fzz}};
    float V[3][3], ev[3];
    eigen3<float>(H, V, ev);
    for (int k = 0; k < 3; k++) {
        o[k] = ev[k];
        o[3 + k] = V[k][0];
        o[6 + k] = V[k][1];
        o[9 + k] = V[k][2];
    }
}

static void requireArrays(const TubeSegmentation& T) {
    if (!T.Fx || !T.Fy || !T.Fz) fail(TSF_ERR_INVALID_ARGUMENT, "doEigen: TubeSegmentation has no field arrays");
}

void doEigenBatch(TubeSegmentation& T, int3 size, const std::vector<int3>& pos, std::vector<float>& out12) {
    requireArrays(T);
    out12.assign(12 * pos.size(), 0.0f);
    for (size_t i = 0; i < pos.size(); i++) eigen_at(T.Fx, T.Fy, T.Fz, size, pos[i], out12.data() + 12 * i);
}

void doEigen(TubeSegmentation& T, int3 pos, int3 size, float3* lambda, float3* e1, float3* e2, float3* e3) {
    requireArrays(T);
    float o[12];
    eigen_at(T.Fx, T.Fy, T.Fz, size, pos, o);
    if (lambda) *lambda = float3(o[0], o[1], o[2]);
    if (e1) *e1 = float3(o[3], o[4], o[5]);
    if (e2) *e2 = float3(o[6], o[7], o[8]);
    if (e3) *e3 = float3(o[9], o[10], o[11]);
}

float3 getTubeDirection(TubeSegmentation& T, int3 pos, int3 size) {
    float3 e1;
    doEigen(T, pos, size, nullptr, &e1, nullptr, nullptr);
    return e1;
}

EigenField::EigenField(OpenCL& ocl, const TubeSegmentation& T, int3 size) : ocl_(ocl), size_(size) {
    requireArrays(T);
    const size_t n = voxels(size);
    fx_ = Buffer(n * 4);
    fy_ = Buffer(n * 4);
    fz_ = Buffer(n * 4);
    refresh(T);
}

void EigenField::refresh(const TubeSegmentation& T) {
    requireArrays(T);
    const size_t n = voxels(size_);
    fx_.upload(ocl_, T.Fx, n * 4);
    fy_.upload(ocl_, T.Fy, n * 4);
    fz_.upload(ocl_, T.Fz, n * 4);
}

void EigenField::batch(const std::vector<int3>& pos, std::vector<float>& out12) {
    const size_t k = pos.size();
    out12.assign(12 * k, 0.0f);
    if (!k) return;
    if (k > cap_) {  // grow the point and result buffers, never shrink them
        cap_ = std::max(k, 2 * cap_);
        points_ = Buffer(cap_ * 12);
        results_ = Buffer(cap_ * 48);
    }
    std::vector<int32_t> p(3 * k);
    for (size_t i = 0; i < k; i++) {
        p[3 * i] = pos[i].x;
        p[3 * i + 1] = pos[i].y;
        p[3 * i + 2] = pos[i].z;
    }
    points_.upload(ocl_, p.data(), p.size() * 4);
    check(tsf_do_eigen_points((const float*)fx_.get(), (const float*)fy_.get(), (const float*)fz_.get(),
                              tsf_dims{size_.x, size_.y, size_.z}, (const int32_t*)points_.get(), (int64_t)k,
                              (float*)results_.get(), ocl_.queue));
    results_.download(ocl_, out12.data(), out12.size() * 4);
}

}  // namespace tsf
