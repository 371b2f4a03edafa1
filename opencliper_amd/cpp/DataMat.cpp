// File constructors of KData / XData, matlabSave and KData::saveCFLData (KData.cpp:246-435,688-710,837-936,
// XData.cpp:500-556,898-930 of the reference).  MAT v5 layout, as the reference reads and writes it:
//   Dims            uint32 1 x n: [nSpatial, allSizesEqual, nCoils, nTemporal, temporal..., spatial...], spatial sizes
//                   in OpenCLIPER order (width first) (MatVarDimsData.cpp:39-58,106-127)
//   KData           single complex [H, W, (D...), nCoils, temporal...]; a file may leave the coil dimension out when
//                   nCoils is 1 (KData.cpp:437-462); the reference always writes it
//   SensitivityMaps single complex [H, W, (D...), nCoils] (coil dimension absent for one coil) (SensitivityMapsData.cpp:147-200)
//   SamplingMasks   uint8 pixel masks [H, W, temporal...] or uint32 row masks [H, temporal...] with elements 0 / 1
//                   (SamplingMasksData.cpp:170-219,389-441)
//   XData           single complex [H, W, (D...), temporal...], nCoils = 0 in Dims (XData.cpp:500-556)
// MATLAB stores column-major with the row (height) index fastest; an NDArray here is row-major with the width index
// fastest, so each H x W plane is transposed on the way in and out (NDArray.cpp:74-124, MatVarInfo.cpp:143-190).
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/MatFile.hpp>
#include <OpenCLIPER/XData.hpp>

#include <fstream>
#include <sstream>

namespace OpenCLIPER {

namespace {

struct FileDims {
    std::vector<dimIndexType> spatial;   // OpenCLIPER order: width, height, depth...
    std::vector<dimIndexType> temporal;
    dimIndexType allSizesEqual = 1;
    numCoilsType nCoils = 0;
};

// Dims variable (Data.cpp:685-715, MatVarDimsData.cpp:106-127)
FileDims readDims(const std::map<std::string, mat::Var>& vars) {
    auto it = vars.find("Dims");
    if (it == vars.end()) throw std::invalid_argument("Error: no Dims variable in matlab file");
    const mat::Var& v = it->second;
    if (v.complex || v.mclass != mat::mxUINT32)
        throw std::invalid_argument("Matlab dims variable should be an array of 32 bits unsigned integers");
    const std::vector<uint32_t> d = v.asUint32();
    if (d.size() < 4 || d.size() != 4 + (size_t)d[0] + d[3])
        throw std::invalid_argument("Matlab dims variable: length does not match its spatial / temporal counts");
    FileDims fd;
    fd.allSizesEqual = d[1];
    fd.nCoils = d[2];
    fd.temporal.assign(d.begin() + 4, d.begin() + 4 + d[3]);
    fd.spatial.assign(d.begin() + 4 + d[3], d.end());
    if (fd.spatial.empty()) throw std::invalid_argument("Matlab dims variable: no spatial dimensions");
    return fd;
}

mat::Var dimsVar(const std::vector<dimIndexType>& spatial, bool allEqual, numCoilsType nCoils,
                 const std::vector<dimIndexType>& temporal) {
    std::vector<uint32_t> d = {(uint32_t)spatial.size(), allEqual ? 1u : 0u, nCoils, (uint32_t)temporal.size()};
    d.insert(d.end(), temporal.begin(), temporal.end());
    d.insert(d.end(), spatial.begin(), spatial.end());
    return mat::makeUint32("Dims", {1, (uint64_t)d.size()}, d.data());
}

// MATLAB dims of data with the given spatial dims (OpenCLIPER order) followed by extra dims: rows (height) first, at
// least two dims (MatVarInfo.cpp:239-264)
std::vector<uint64_t> matlabDims(const std::vector<dimIndexType>& spatial, const std::vector<dimIndexType>& extra) {
    std::vector<uint64_t> d;
    if (spatial.size() == 1) {
        d = {1, spatial[0]};
    } else {
        d = {spatial[1], spatial[0]};
        for (size_t i = 2; i < spatial.size(); ++i) d.push_back(spatial[i]);
    }
    d.insert(d.end(), extra.begin(), extra.end());
    return d;
}

std::string dimsString(const std::vector<uint64_t>& d) {
    std::ostringstream o;
    for (size_t i = 0; i < d.size(); ++i) o << (i ? " x " : "") << d[i];
    return o.str();
}

// A variable's dims must equal the expected ones up to trailing singleton dimensions (MATLAB and most writers drop
// them, libmatio keeps them)
void checkDims(const mat::Var& v, const std::vector<uint64_t>& want, const char* what) {
    std::vector<uint64_t> a = v.dims, b = want;
    while (a.size() > 2 && a.back() == 1) a.pop_back();
    while (b.size() > 2 && b.back() == 1) b.pop_back();
    if (a != b)
        throw std::out_of_range(std::string("Invalid ") + what + " dimensions (" + dimsString(v.dims) + "), should be " +
                                dimsString(want));
}

// Splits a column-major MATLAB array of n stacked planes into n row-major NDArrays of the given spatial dims
// (NDArray.cpp:74-124)
template <class T, class Get>
std::vector<NDArray*>* splitPlanes(const std::vector<dimIndexType>& spatial, size_t n, Get get) {
    const size_t W = spatial[0], H = spatial.size() > 1 ? spatial[1] : 1;
    size_t S = 1;
    for (size_t i = 2; i < spatial.size(); ++i) S *= spatial[i];
    const size_t plane = W * H * S;
    auto* arrays = new std::vector<NDArray*>();
    for (size_t a = 0; a < n; ++a) {
        auto* dims = new std::vector<dimIndexType>(spatial);
        auto* data = new std::vector<T>(plane);
        for (size_t s = 0; s < S; ++s)
            for (size_t h = 0; h < H; ++h)
                for (size_t w = 0; w < W; ++w) (*data)[w + W * (h + H * s)] = get(a * plane + h + H * (w + W * s));
        arrays->push_back(new ConcreteNDArray<T>(dims, data));
    }
    return arrays;
}

// Joins row-major NDArrays into one column-major MATLAB array (MatVarInfo.cpp:143-190)
template <class T> std::vector<T> joinPlanes(const Data& d) {
    const std::vector<dimIndexType>& spatial = *d.getNDArray(0)->getDims();
    const size_t W = spatial[0], H = spatial.size() > 1 ? spatial[1] : 1;
    size_t S = 1;
    for (size_t i = 2; i < spatial.size(); ++i) S *= spatial[i];
    const size_t plane = W * H * S;
    std::vector<T> out(plane * d.getNumNDArrays());
    for (dimIndexType a = 0; a < d.getNumNDArrays(); ++a) {
        const T* src = (const T*)d.getHostBuffer(a);
        for (size_t s = 0; s < S; ++s)
            for (size_t h = 0; h < H; ++h)
                for (size_t w = 0; w < W; ++w) out[a * plane + h + H * (w + W * s)] = src[w + W * (h + H * s)];
    }
    return out;
}

std::vector<NDArray*>* complexPlanes(const mat::Var& v, const std::vector<dimIndexType>& spatial, size_t n,
                                     const char* what) {
    if (v.mclass != mat::mxSINGLE)
        throw std::invalid_argument(std::string("Unsupported element data type in matlab data (") + what +
                                    " must be single precision)");
    return splitPlanes<complexType>(spatial, n, [&](size_t i) {
        return complexType((float)v.re[i], v.complex ? (float)v.im[i] : 0.f);
    });
}

dimIndexType product(const std::vector<dimIndexType>& d) {
    dimIndexType n = 1;
    for (auto x : d) n *= x;
    return n;
}

mat::Var complexVar(const std::string& name, const std::vector<uint64_t>& dims, const Data& d) {
    const std::vector<complexType> all = joinPlanes<complexType>(d);
    return mat::makeComplexSingle(name, dims, all.data());
}

}  // namespace

KData::KData(const std::shared_ptr<CLapp>& app, const std::string& fileName) : Data(app, TYPEID_COMPLEX) {
    if (mat::isMatFile(fileName)) loadMat(fileName);
    else loadCFL(fileName);
}

void KData::loadMat(const std::string& fileName) {
    const std::shared_ptr<CLapp>& app = app_;
    const std::map<std::string, mat::Var> vars = mat::read(fileName);
    const FileDims fd = readDims(vars);
    if (fd.nCoils < 1) throw std::invalid_argument("KData: Dims variable holds no coils");
    nCoils_ = fd.nCoils;
    dynDims_ = fd.temporal;
    auto it = vars.find("KData");
    if (it == vars.end()) throw std::invalid_argument("Error: no KData variable in matlab file");
    std::vector<dimIndexType> extra;
    if (nCoils_ > 1 || it->second.dims.size() > fd.spatial.size() + fd.temporal.size() ||
        (fd.spatial.size() == 1 && it->second.dims.size() > 2 + fd.temporal.size()))
        extra.push_back(nCoils_);
    extra.insert(extra.end(), fd.temporal.begin(), fd.temporal.end());
    checkDims(it->second, matlabDims(fd.spatial, extra), "KData");
    const size_t frames = fd.temporal.empty() ? 1 : product(fd.temporal);
    if (dynDims_.empty()) dynDims_ = {1};
    std::vector<NDArray*>* arrays = complexPlanes(it->second, fd.spatial, (size_t)nCoils_ * frames, "KData");
    setNDArrays(arrays, true);

    auto sit = vars.find("SensitivityMaps");
    if (sit != vars.end()) {
        std::vector<dimIndexType> cdim;
        if (nCoils_ > 1) cdim.push_back(nCoils_);
        checkDims(sit->second, matlabDims(fd.spatial, cdim), "SensitivityMapsData");
        std::vector<NDArray*>* maps = complexPlanes(sit->second, fd.spatial, nCoils_, "SensitivityMaps");
        sens_ = std::make_shared<SensitivityMapsData>(app, maps, nCoils_);
    }

    auto mit = vars.find("SamplingMasks");
    if (mit != vars.end()) {
        const mat::Var& m = mit->second;
        const std::vector<dimIndexType> sp2(fd.spatial.begin(), fd.spatial.begin() + std::min<size_t>(2, fd.spatial.size()));
        std::vector<NDArray*>* masks = nullptr;
        if (m.mclass == mat::mxUINT32 && !m.complex) {
            // ROWMASK: one 0 / 1 entry per k-space line and frame, expanded to whole rows of the k-space width
            const dimIndexType H = fd.spatial.size() > 1 ? fd.spatial[1] : 1;
            std::vector<uint64_t> want = {H};
            want.insert(want.end(), fd.temporal.begin(), fd.temporal.end());
            if (want.size() < 2) want.push_back(1);
            checkDims(m, want, "SamplingMasks");
            for (double x : m.re)
                if (x != 0.0 && x != 1.0)
                    throw std::invalid_argument("sampling maks data element must be 0 or 1 (it is " +
                                                std::to_string((long long)x) + ")");
            masks = new std::vector<NDArray*>();
            for (size_t f = 0; f < frames; ++f) {
                auto* dims = new std::vector<dimIndexType>({H});
                auto* data = new std::vector<uint8_t>(H);
                for (dimIndexType h = 0; h < H; ++h) (*data)[h] = m.re[f * H + h] != 0.0 ? 1 : 0;
                masks->push_back(new ConcreteNDArray<uint8_t>(dims, data));
            }
        } else if (m.mclass == mat::mxUINT8 && !m.complex) {
            checkDims(m, matlabDims(sp2, fd.temporal), "SamplingMasks");
            masks = splitPlanes<uint8_t>(sp2, frames, [&](size_t i) { return (uint8_t)m.re[i]; });
        } else {
            throw std::invalid_argument("Invalid mask format (supported: ROWMASK and PIXELMASK");
        }
        auto* mdyn = new std::vector<dimIndexType>(dynDims_);
        masks_ = std::make_shared<SamplingMasksData>(app, masks, mdyn, fd.spatial[0]);
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// CFL file sets (KData.cpp:246-345,688-710, Data.cpp:1071-1178): raw little-endian NDArrays back to back in the device
// order (width fastest, coil fastest across NDArrays), and a text header "<width> <height> <nCoils> <frames> 1" whose
// first non-comment line is read.
// ---------------------------------------------------------------------------------------------------------------------
namespace {

std::string baseName(const std::string& fileName) {
    const size_t dot = fileName.find_last_of('.');
    return dot == std::string::npos ? fileName : fileName.substr(0, dot);
}

bool fileExists(const std::string& name) { return std::ifstream(name).good(); }

template <class T>
std::vector<NDArray*>* readRawArrays(const std::string& name, const std::vector<dimIndexType>& dims, size_t n) {
    std::ifstream f(name, std::ios::binary);
    if (!f) throw std::invalid_argument(name + " cannot be read");
    auto* arrays = new std::vector<NDArray*>();
    try {
        for (size_t i = 0; i < n; ++i) {
            auto* d = new std::vector<dimIndexType>(dims);
            auto* a = new ConcreteNDArray<T>(d);
            arrays->push_back(a);
            const std::streamsize bytes = (std::streamsize)(a->size() * sizeof(T));
            if (!f.read((char*)a->getHostDataAsVoidPtr(), bytes))
                throw std::invalid_argument(name + ": file too short for " + std::to_string(n) + " arrays");
        }
    } catch (...) {
        for (NDArray* a : *arrays) delete a;
        delete arrays;
        throw;
    }
    return arrays;
}

void writeRaw(std::ofstream& f, const Data& d, size_t elemBytes) {
    for (dimIndexType i = 0; i < d.getNumNDArrays(); ++i)
        f.write((const char*)d.getHostBuffer(i), (std::streamsize)(d.getNDArray(i)->size() * elemBytes));
}

}  // namespace

void KData::loadCFL(const std::string& fileName) {
    const std::string base = baseName(fileName);
    const std::string hdr = base + ".hdr";
    std::ifstream h(hdr);
    if (!h) throw std::invalid_argument(hdr + " cannot be read (neither a MAT v5 file nor a CFL set)");
    dimIndexType W = 0, H = 0, C = 0, F = 0;
    std::string line;
    bool found = false;
    while (std::getline(h, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        if (!(iss >> W >> H >> C >> F))
            throw std::invalid_argument("Invalid line format: " + line +
                                        " (should be <width> <depth> <numCoils> <numFrames>)");
        found = true;
        break;
    }
    if (!found || W == 0 || H == 0 || C == 0 || F == 0)
        throw std::invalid_argument(hdr + ": no valid <width> <height> <numCoils> <numFrames> line");
    nCoils_ = C;
    dynDims_ = {F};
    const std::vector<dimIndexType> spatial = {W, H};
    std::vector<NDArray*>* arrays = readRawArrays<complexType>(fileName, spatial, (size_t)C * F);
    setNDArrays(arrays, true);

    const std::string sensName = base + "_SensitivityMaps.cfl";
    if (fileExists(sensName)) {
        std::vector<NDArray*>* maps = readRawArrays<complexType>(sensName, spatial, C);
        sens_ = std::make_shared<SensitivityMapsData>(app_, maps, nCoils_);
    }
    const std::string rowName = base + "_SamplingMasks.u32", pixName = base + "_SamplingMasks.u8";
    std::vector<NDArray*>* masks = nullptr;
    if (fileExists(rowName)) {
        // ROWMASK: H uint32 entries (0 / 1) per frame, expanded to whole rows (SamplingMasksData.cpp:389-441)
        std::vector<NDArray*>* rows = readRawArrays<uint32_t>(rowName, {H}, F);
        masks = new std::vector<NDArray*>();
        for (NDArray* a : *rows) {
            const uint32_t* r = (const uint32_t*)a->getHostDataAsVoidPtr();
            auto* d = new std::vector<dimIndexType>({H});
            auto* m = new std::vector<uint8_t>(H);
            for (dimIndexType i = 0; i < H; ++i) {
                if (r[i] > 1) {
                    const uint32_t bad = r[i];
                    for (NDArray* b : *rows) delete b;
                    delete rows;
                    for (NDArray* b : *masks) delete b;
                    delete masks;
                    delete d;
                    delete m;
                    throw std::invalid_argument("sampling maks data element must be 0 or 1 (it is " +
                                                std::to_string(bad) + ")");
                }
                (*m)[i] = (uint8_t)r[i];
            }
            masks->push_back(new ConcreteNDArray<uint8_t>(d, m));
        }
        for (NDArray* a : *rows) delete a;
        delete rows;
    } else if (fileExists(pixName)) {
        masks = readRawArrays<uint8_t>(pixName, spatial, F);
    }
    if (masks) {
        auto* mdyn = new std::vector<dimIndexType>(dynDims_);
        masks_ = std::make_shared<SamplingMasksData>(app_, masks, mdyn, W);
    }
}

void KData::saveCFLData(const std::string& baseFileName) {
    const std::vector<dimIndexType>& spatial = *getNDArray(0)->getDims();
    if (!getAllSizesEqual() || spatial.size() != 2 || dynDims_.size() != 1)
        throw std::invalid_argument("KData::saveCFLData: equal-size 2-D NDArrays with one temporal dimension only");
    device2Host();
    {
        std::ofstream f(baseFileName + ".cfl", std::ios::binary | std::ios::trunc);
        if (!f) throw std::invalid_argument(baseFileName + ".cfl cannot be created");
        writeRaw(f, *this, sizeof(complexType));
    }
    if (sens_) {
        sens_->device2Host();
        std::ofstream f(baseFileName + "_SensitivityMaps.cfl", std::ios::binary | std::ios::trunc);
        writeRaw(f, *sens_, sizeof(complexType));
    }
    if (masks_) {
        masks_->device2Host();
        std::ofstream f(baseFileName + "_SamplingMasks.u8", std::ios::binary | std::ios::trunc);
        writeRaw(f, *masks_, 1);
    }
    std::ofstream h(baseFileName + ".hdr", std::ios::trunc);
    if (!h) throw std::invalid_argument(baseFileName + ".hdr cannot be created");
    for (dimIndexType d : spatial) h << d << " ";
    h << nCoils_ << " " << dynDims_[0] << " 1\n";
}

void KData::matlabSave(const std::string& fileName, bool saveSamplingMasks) {
    if (getNDArray(0)->getDims()->at(0) == 0) throw std::invalid_argument("Invalid data size (width is 0)");
    if (!getAllSizesEqual())
        throw std::invalid_argument("KData::matlabSave: only NDArrays of equal spatial dimensions can be saved");
    device2Host();
    const std::vector<dimIndexType>& spatial = *getNDArray(0)->getDims();
    std::vector<mat::Var> vars;
    vars.push_back(dimsVar(spatial, true, nCoils_, dynDims_));
    std::vector<dimIndexType> extra = {nCoils_};
    extra.insert(extra.end(), dynDims_.begin(), dynDims_.end());
    vars.push_back(complexVar("KData", matlabDims(spatial, extra), *this));
    if (sens_) {
        sens_->device2Host();
        vars.push_back(complexVar("SensitivityMaps", matlabDims(*sens_->getNDArray(0)->getDims(), {sens_->getNCoils()}),
                                  *sens_));
    }
    if (saveSamplingMasks && masks_) {
        masks_->device2Host();
        const std::vector<uint8_t> all = joinPlanes<uint8_t>(*masks_);
        vars.push_back(mat::makeUint8("SamplingMasks", matlabDims(*masks_->getNDArray(0)->getDims(), dynDims_),
                                      all.data()));
    }
    mat::write(fileName, vars, true);
}

XData::XData(const std::shared_ptr<CLapp>& app, const std::string& fileName) : Data(app, TYPEID_COMPLEX) {
    const std::map<std::string, mat::Var> vars = mat::read(fileName);
    const FileDims fd = readDims(vars);
    if (fd.nCoils > 1) throw std::out_of_range("Invalid XData dimensions: Dims variable holds more than one coil");
    auto it = vars.find("XData");
    if (it == vars.end()) throw std::invalid_argument("Error: no XData variable in matlab file");
    checkDims(it->second, matlabDims(fd.spatial, fd.temporal), "XData");
    dynDims_ = fd.temporal.empty() ? std::vector<dimIndexType>{1} : fd.temporal;
    std::vector<NDArray*>* arrays = complexPlanes(it->second, fd.spatial, product(dynDims_), "XData");
    setNDArrays(arrays, true);
}

void XData::matlabSave(const std::string& fileName) {
    if (getNDArray(0)->getDims()->at(0) == 0) throw std::invalid_argument("Invalid data size (width is 0)");
    if (!getAllSizesEqual())
        throw std::invalid_argument("XData::matlabSave: only NDArrays of equal spatial dimensions can be saved");
    device2Host();
    const std::vector<dimIndexType>& spatial = *getNDArray(0)->getDims();
    std::vector<mat::Var> vars;
    vars.push_back(dimsVar(spatial, true, 0, dynDims_));
    vars.push_back(complexVar("XData", matlabDims(spatial, dynDims_), *this));
    mat::write(fileName, vars, true);
}

}  // namespace OpenCLIPER
