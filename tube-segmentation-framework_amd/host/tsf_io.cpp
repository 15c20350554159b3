// The reference's entry point and its ingest and output around the path:
// run() (tube-segmentation.cpp:53-163), readDatasetAndTransfer (1399-1905),
// runCircleFittingAndNewCenterlineAlg (943-986), TSFOutput and
// writeDataToDisk (inputOutput.cpp:7-239).
//
// The volume is read from the .mhd/.raw pair on the host and uploaded once;
// the limits, the scan-line counts of cropping and toFloat of the cropped
// region run on the device (tsf_volume_limits, tsf_crop_scan_lines,
// tsf_to_float_region), the crop region on the host (tsf_crop_region).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

#include "tsf_dropin.hpp"

namespace tsf {

namespace {

struct MhdHeader {
    std::string type, raw;
    int3 size;
    float3 spacing{1.0f, 1.0f, 1.0f};
};

std::string after(const std::string& line, size_t key_length) {
    return line.size() > key_length + 3 ? line.substr(key_length + 3) : std::string();
}

// three space-separated fields, as the reference splits DimSize and
// ElementSpacing
void three_fields(std::string s, std::string out[3]) {
    for (int k = 0; k < 3; k++) {
        const size_t sp = s.find(' ');
        out[k] = s.substr(0, sp);
        s = sp == std::string::npos ? std::string() : s.substr(sp + 1);
    }
}

// The header lines the reference reads (tube-segmentation.cpp:1407-1470):
// "ElementType = ", "ElementDataFile = " (cut at its first space, relative
// to the .mhd's directory), "DimSize = ", "ElementSpacing = ".
MhdHeader read_mhd(const std::string& filename) {
    std::ifstream file(filename);
    if (!file) throw IOException(filename);
    MhdHeader h;
    bool type_found = false, size_found = false, raw_found = false;
    std::string line;
    while (std::getline(file, line)) {
        if (line.substr(0, 11) == "ElementType") {
            h.type = after(line, 11);
            type_found = true;
        } else if (line.substr(0, 15) == "ElementDataFile") {
            h.raw = after(line, 15);
            raw_found = true;
            const size_t sp = h.raw.find(' ');
            if (sp != std::string::npos && sp > 0) h.raw = h.raw.substr(0, sp);
            const size_t slash = filename.rfind('/');
            if (slash != std::string::npos && slash > 0) h.raw = filename.substr(0, slash + 1) + h.raw;
        } else if (line.substr(0, 7) == "DimSize") {
            std::string f[3];
            three_fields(after(line, 7), f);
            h.size = int3(std::atoi(f[0].c_str()), std::atoi(f[1].c_str()), std::atoi(f[2].c_str()));
            size_found = true;
        } else if (line.substr(0, 14) == "ElementSpacing") {
            std::string f[3];
            three_fields(after(line, 14), f);
            h.spacing = float3((float)std::atof(f[0].c_str()), (float)std::atof(f[1].c_str()),
                               (float)std::atof(f[2].c_str()));
        }
    }
    const size_t sp = h.type.find(' ');
    if (sp != std::string::npos && sp > 0) h.type = h.type.substr(0, sp);
    if (!type_found || !size_found || !raw_found)
        throw Error(TSF_ERR_INVALID_ARGUMENT, "Error reading mhd file. Type, filename or size not found");
    return h;
}

// ElementType -> (tsf_to_float type code, bytes per voxel, signed class of
// the reference's cropping kernels)
struct RawType {
    int code;
    size_t bytes;
};
RawType raw_type(const std::string& t) {
    if (t == "MET_SHORT") return {3, 2};
    if (t == "MET_USHORT") return {2, 2};
    if (t == "MET_CHAR") return {1, 1};
    if (t == "MET_UCHAR") return {0, 1};
    if (t == "MET_FLOAT") return {4, 4};
    throw Error(TSF_ERR_INVALID_ARGUMENT, "unsupported data type " + t);
}

void set_number_string(paramList& p, const char* name, float v) {
    char str[64];
    std::snprintf(str, sizeof str, "%f", v);
    setParameter(p, name, str);
}

}  // namespace

IOException::IOException(const std::string& filename)
    : Error(TSF_ERR_INVALID_ARGUMENT, "Failed to read the file " + filename), filename(filename) {}

// ----------------------------------------------------------------- TSFOutput
TSFOutput::TSFOutput(int3* size, bool TDFis16bit) : size(size), TDFis16bit(TDFis16bit) {}

TSFOutput::~TSFOutput() {
    delete[] TDF;
    delete[] segmentation;
    delete[] centerlineVoxels;
    delete oclTDF;
    delete oclSegmentation;
    delete oclCenterlineVoxels;
    delete size;
}

void TSFOutput::setTDF(Image3D* image) { oclTDF = image; }
void TSFOutput::setSegmentation(Buffer* mask) { oclSegmentation = mask; }
void TSFOutput::setCenterlineVoxels(Buffer* mask) { oclCenterlineVoxels = mask; }
void TSFOutput::setTDF(float* data) { TDF = data; }
void TSFOutput::setSegmentation(char* data) { segmentation = data; }
void TSFOutput::setCenterlineVoxels(char* data) { centerlineVoxels = data; }
void TSFOutput::setSize(int3* s) { size = s; }

float* TSFOutput::getTDF() {
    if (TDF) return TDF;
    if (!oclTDF) throw Error(TSF_ERR_INVALID_ARGUMENT, "Trying to fetch non existing data from TSFOutput");
    // a 16-bit TDF already holds the UNORM16 values t/65535 as floats
    std::vector<float> t = oclTDF->downloadPlane(ocl, 0);
    TDF = new float[t.size()];
    std::memcpy(TDF, t.data(), t.size() * sizeof(float));
    return TDF;
}

char* TSFOutput::getSegmentation() {
    if (segmentation) return segmentation;
    if (!oclSegmentation) throw Error(TSF_ERR_INVALID_ARGUMENT, "Trying to fetch non existing data from TSFOutput");
    const size_t n = (size_t)size->x * size->y * size->z;
    segmentation = new char[n];
    oclSegmentation->download(ocl, segmentation, n);
    return segmentation;
}

char* TSFOutput::getCenterlineVoxels() {
    if (centerlineVoxels) return centerlineVoxels;
    if (!oclCenterlineVoxels)
        throw Error(TSF_ERR_INVALID_ARGUMENT, "Trying to fetch non existing data from TSFOutput");
    const size_t n = (size_t)size->x * size->y * size->z;
    centerlineVoxels = new char[n];
    oclCenterlineVoxels->download(ocl, centerlineVoxels, n);
    return centerlineVoxels;
}

// ----------------------------------------------------------------- ingest
Image3D readDatasetAndTransfer(OpenCL& ocl, std::string filename, paramList& parameters, int3* size,
                               TSFOutput* output) {
    const MhdHeader h = read_mhd(filename);
    const RawType rt = raw_type(h.type);
    *size = h.size;
    if (size->x < 1 || size->y < 1 || size->z < 1)
        throw Error(TSF_ERR_DIMENSIONS, "Error reading mhd file. Invalid DimSize");
    const size_t n = (size_t)size->x * size->y * size->z, bytes = n * rt.bytes;
    std::vector<char> host(bytes);
    {
        std::ifstream raw(h.raw, std::ios::binary);
        if (!raw || !raw.read(host.data(), (std::streamsize)bytes)) throw IOException(h.raw);
    }
    Buffer dataset(bytes);
    dataset.upload(ocl, host.data(), bytes);
    host.clear();
    host.shrink_to_fit();
    const tsf_dims d{size->x, size->y, size->z};

    // getLimits (tube-segmentation.cpp:1376-1397): a parameter, else the
    // volume's own minimum / maximum
    float minimum = 0.0f, maximum = 1.0f;
    const bool auto_min = getParamStr(parameters, "minimum") == "off";
    const bool auto_max = getParamStr(parameters, "maximum") == "off";
    if (auto_min || auto_max) check(tsf_volume_limits(dataset.get(), rt.code, d, &minimum, &maximum, ocl.queue));
    if (!auto_min) minimum = (float)std::atof(getParamStr(parameters, "minimum").c_str());
    if (!auto_max) maximum = (float)std::atof(getParamStr(parameters, "maximum").c_str());
    // unsigned CT volumes under the airway and AAA presets: HU + 1024
    // (tube-segmentation.cpp:1500-1512)
    const std::string preset = getParamStr(parameters, "parameters");
    if (h.type == "MET_USHORT" && (preset == "Lung-Airways-CT" || preset == "AAA-Vessels-CT")) {
        minimum = (float)std::atof(getParamStr(parameters, "minimum").c_str()) + 1024.0f;
        set_number_string(parameters, "minimum", minimum);
        maximum = (float)std::atof(getParamStr(parameters, "maximum").c_str()) + 1024.0f;
        set_number_string(parameters, "maximum", maximum);
    }

    // cropping (tube-segmentation.cpp:1556-1853)
    tsf_dims offset{0, 0, 0}, crop = d;
    int3 shift;
    const std::string cropping = getParamStr(parameters, "cropping");
    if (cropping == "lung" || cropping == "threshold") {
        const bool lung = cropping == "lung";
        const int min_lines = (int)getParam(parameters, lung ? "min-scan-lines-lung" : "min-scan-lines-threshold");
        const bool middle = lung || getParamStr(parameters, "cropping-start-z") == "middle";
        Buffer lx(sizeof(int32_t) * size->x), ly(sizeof(int32_t) * size->y), lz(sizeof(int32_t) * size->z);
        check(tsf_crop_scan_lines(dataset.get(), rt.code, d, lung ? TSF_CROP_LUNG : TSF_CROP_THRESHOLD,
                                  lung ? 0.0f : getParam(parameters, "cropping-threshold"), (int32_t*)lx.get(),
                                  (int32_t*)ly.get(), (int32_t*)lz.get(), ocl.queue));
        std::vector<int32_t> hx(size->x), hy(size->y), hz(size->z);
        lx.download(ocl, hx.data(), hx.size() * sizeof(int32_t));
        ly.download(ocl, hy.data(), hy.size() * sizeof(int32_t));
        lz.download(ocl, hz.data(), hz.size() * sizeof(int32_t));
        check(tsf_crop_region(hx.data(), hy.data(), hz.data(), d, min_lines, middle ? 1 : 0, &offset, &crop));
        shift = int3(offset.x, offset.y, offset.z);
    } else if (preset == "AAA-Vessels-CT") {
        // 15 % off each side in x and y (tube-segmentation.cpp:1782-1806)
        const float percent = 0.15f;
        offset.x = (int)std::round(size->x * percent);
        offset.y = (int)std::round(size->y * percent);
        crop.x = size->x - 2 * offset.x;
        crop.y = size->y - 2 * offset.y;
        while (crop.x % 4 != 0) crop.x--;
        while (crop.y % 4 != 0) crop.y--;
        while (crop.z % 4 != 0) crop.z--;
    } else {
        // each extent shrunk to a multiple of 4 (tube-segmentation.cpp:1807-1835)
        while (crop.x % 4 != 0) crop.x--;
        while (crop.y % 4 != 0) crop.y--;
        while (crop.z % 4 != 0) crop.z--;
    }
    if (crop.x < 1 || crop.y < 1 || crop.z < 1)
        throw Error(TSF_ERR_DIMENSIONS, "volume smaller than 4 voxels in some direction");
    *size = int3(crop.x, crop.y, crop.z);
    if (output) {
        output->setShiftVector(shift);
        output->setSpacing(h.spacing);
    }
    Image3D converted(*size, 1);
    check(tsf_to_float_region(dataset.get(), rt.code, d, offset, (float*)converted.plane(0), converted.dims(),
                              minimum, maximum, ocl.queue));
    return converted;
}

// ----------------------------------------------------------------- the method
void runCircleFittingAndNewCenterlineAlg(OpenCL& ocl, Image3D* dataset, int3* size, paramList& parameters,
                                         TSFOutput* output) {
    Image3D vectorField, radius;
    Image3D* TDF = new Image3D;
    runCircleFittingMethod(ocl, dataset, *size, parameters, vectorField, *TDF, radius);
    output->setTDF(TDF);
    if (getParamBool(parameters, "tdf-only")) return;
    Buffer* centerline = new Buffer(runNewCenterlineAlg(ocl, *size, parameters, vectorField, *TDF, radius));
    output->setCenterlineVoxels(centerline);
    if (!getParamBool(parameters, "no-segmentation")) {
        Buffer* segmentation = new Buffer(
            getParamBool(parameters, "sphere-segmentation")
                ? runSphereSegmentation(ocl, *centerline, radius, *size, parameters)
                : runInverseGradientSegmentation(ocl, *centerline, vectorField, radius, *size, parameters));
        output->setSegmentation(segmentation);
    }
    if (getParamStr(parameters, "storage-dir") != "off")
        writeDataToDisk(output, getParamStr(parameters, "storage-dir"), getParamStr(parameters, "storage-name"));
}

TSFOutput* run(std::string filename, paramList& parameters, std::string kernel_dir) {
    (void)kernel_dir;  // the kernels are compiled into the library
    if (getParamStr(parameters, "device") != "gpu") setParameter(parameters, "16bit-vectors", "false");
    // kernels.cl (3D image writes) unless buffers-only: the candidate cap
    // of the suppression follows it
    setParameter(parameters, "3d_write", getParamBool(parameters, "buffers-only") ? "false" : "true");
    TSFOutput* output = new TSFOutput(new int3(), getParamBool(parameters, "16bit-vectors"));
    try {
        Image3D dataset = readDatasetAndTransfer(output->getContext(), filename, parameters, output->getSize(),
                                                 output);
        const std::string method = getParamStr(parameters, "centerline-method");
        if (method == "gpu")
            runCircleFittingAndNewCenterlineAlg(output->getContext(), &dataset, output->getSize(), parameters,
                                                output);
        else if (method == "ridge")
            runCircleFittingAndRidgeTraversal(output->getContext(), &dataset, output->getSize(), parameters, output);
        else
            throw Error(TSF_ERR_INVALID_ARGUMENT, "centerline-method " + method + " is not provided by this library");
        output->getContext().finish();
    } catch (...) {
        delete output;
        throw;
    }
    return output;
}

// ----------------------------------------------------------------- output
void writeDataToDisk(TSFOutput* output, std::string storageDirectory, std::string name) {
    const int3* size = output->getSize();
    const size_t n = (size_t)size->x * size->y * size->z;
    auto write = [&](const char* kind, char* data) {
        const std::string base = storageDirectory + name + "." + kind;
        std::ofstream mhd(base + ".mhd");
        mhd << "ObjectType = Image\n";
        mhd << "NDims = 3\n";
        mhd << "DimSize = " << size->x << " " << size->y << " " << size->z << "\n";
        mhd << "ElementSpacing = " << output->getSpacing().x << " " << output->getSpacing().y << " "
            << output->getSpacing().z << "\n";
        mhd << "ElementType = MET_CHAR\n";
        mhd << "ElementDataFile = " << name << "." << kind << ".raw\n";
        mhd.close();
        std::ofstream raw(base + ".raw", std::ios::binary);
        raw.write(data, (std::streamsize)n);
        if (!mhd || !raw) throw IOException(base);
    };
    if (output->hasCenterlineVoxels()) write("centerline", output->getCenterlineVoxels());
    if (output->hasSegmentation()) write("segmentation", output->getSegmentation());
}

}  // namespace tsf
