// CLapp and the Data classes of the OpenCLIPER-compatible C++ API, on top of the C ABI (include/ocl_amd.h).
#include <OpenCLIPER/CLapp.hpp>
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/XData.hpp>

#include <cstring>

#include "ocl_amd.h"

namespace OpenCLIPER {

void checkOcl(int rc, const char* where) {
    if (rc != OCL_OK) throw CLError(rc, std::string(where) + ": " + ocl_last_error());
}

// ---------------------------------------------------------------------------------------------------------------------
// CLapp (src/CLapp.cpp:139-569 of the reference: choose a device, create its queue)
// ---------------------------------------------------------------------------------------------------------------------
std::shared_ptr<CLapp> CLapp::create() { return create(PlatformTraits(), DeviceTraits()); }

std::shared_ptr<CLapp> CLapp::create(const PlatformTraits& pt, const DeviceTraits& dt) {
    std::shared_ptr<CLapp> app(new CLapp());
    app->init(pt, dt);
    return app;
}

void CLapp::init(const PlatformTraits&, const DeviceTraits& dt) {
    if (ctx_) return;
    if (!dt.useHIP) throw CLError(OCL_E_INVALID, "CLapp::init: this build runs on HIP devices only");
    checkOcl(ocl_ctx_create(dt.deviceNumber, &ctx_), "CLapp::init");
    // the reference never seeds drand48, so the myNormest restarts of every NestaUp::launch of a process continue one
    // stream (NestaUp.cpp:375); the context carries that stream the same way
    checkOcl(ocl_ctx_set_normest_stream(ctx_, 1), "CLapp::init");
}

CLapp::~CLapp() {
    if (ctx_) ocl_ctx_destroy(ctx_);
}

void* CLapp::getHIPStream() const { return ocl_ctx_stream(ctx_); }

std::string CLapp::getDeviceName() const {
    char name[256] = {0};
    checkOcl(ocl_device_info(ctx_, name, sizeof(name), nullptr, nullptr), "CLapp::getDeviceName");
    return name;
}

size_t CLapp::getDeviceGlobalMemory() const {
    size_t m = 0;
    checkOcl(ocl_device_info(ctx_, nullptr, 0, &m, nullptr), "CLapp::getDeviceGlobalMemory");
    return m;
}

void CLapp::finish() const { checkOcl(ocl_ctx_sync(ctx_), "CLapp::finish"); }

// ---------------------------------------------------------------------------------------------------------------------
// Data
// ---------------------------------------------------------------------------------------------------------------------
Data::~Data() {
    if (dev_ && app_ && app_->getOclContext()) ocl_free(app_->getOclContext(), dev_);
}

dimIndexType Data::getDynDimsTotalSize() const {
    dimIndexType n = 1;
    for (auto d : dynDims_) n *= d;
    return dynDims_.empty() ? 0 : n;
}

bool Data::getAllSizesEqual() const {
    for (size_t i = 1; i < view_.size(); ++i)
        if (*view_[i]->getDims() != *view_[0]->getDims()) return false;
    return true;
}

void Data::setNDArrays(std::vector<NDArray*>*& arrays, bool upload) {
    if (!arrays || arrays->empty()) throw std::invalid_argument("Data: no NDArrays given");
    std::unique_ptr<std::vector<NDArray*>> own(arrays);
    arrays = nullptr;
    bytes_ = 0;
    for (NDArray* a : *own) {
        if (!a) throw std::invalid_argument("Data: null NDArray");
        if (a->getElementDataType() != type_) throw std::invalid_argument("Data: NDArray element type mismatch");
        arrays_.emplace_back(a);
        view_.push_back(a);
        bytes_ += (size_t)a->size() * NDArray::getElementSize(type_);
    }
    checkOcl(ocl_malloc(app_->getOclContext(), bytes_, &dev_), "Data::allocate");
    if (upload) host2Device();
}

void Data::cloneShape(const Data& src, bool copyData) {
    auto* arrays = new std::vector<NDArray*>();
    for (const NDArray* a : src.view_) {
        auto* dims = new std::vector<dimIndexType>(*a->getDims());
        if (type_ == TYPEID_COMPLEX) arrays->push_back(new ConcreteNDArray<complexType>(dims));
        else if (type_ == TYPEID_REAL) arrays->push_back(new ConcreteNDArray<realType>(dims));
        else if (type_ == TYPEID_UINT32) arrays->push_back(new ConcreteNDArray<uint32_t>(dims));
        else if (type_ == TYPEID_INT32) arrays->push_back(new ConcreteNDArray<int32_t>(dims));
        else arrays->push_back(new ConcreteNDArray<uint8_t>(dims));
    }
    dynDims_ = src.dynDims_;
    setNDArrays(arrays, false);
    if (copyData) {
        checkOcl(ocl_memcpy_d2d(app_->getOclContext(), dev_, src.dev_, bytes_), "Data copy");
    } else {
        checkOcl(ocl_memset_zero(app_->getOclContext(), dev_, bytes_), "Data zero-init");
    }
}

void Data::host2Device() {
    size_t off = 0;
    for (auto& a : arrays_) {
        const size_t b = (size_t)a->size() * NDArray::getElementSize(type_);
        checkOcl(ocl_memcpy_h2d(app_->getOclContext(), (char*)dev_ + off, a->getHostDataAsVoidPtr(), b),
                 "Data::host2Device");
        off += b;
    }
}

void Data::device2Host() {
    size_t off = 0;
    for (auto& a : arrays_) {
        const size_t b = (size_t)a->size() * NDArray::getElementSize(type_);
        checkOcl(ocl_memcpy_d2h(app_->getOclContext(), a->getHostDataAsVoidPtr(), (char*)dev_ + off, b),
                 "Data::device2Host");
        off += b;
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// XData / KData / SensitivityMapsData / SamplingMasksData
// ---------------------------------------------------------------------------------------------------------------------
XData::XData(const std::shared_ptr<CLapp>& app, std::vector<NDArray*>*& arrays, std::vector<dimIndexType>*& dynDims)
    : Data(app, TYPEID_COMPLEX) {
    if (dynDims) {
        dynDims_ = *dynDims;
        delete dynDims;
        dynDims = nullptr;
    } else {
        dynDims_ = {(dimIndexType)(arrays ? arrays->size() : 0)};
    }
    setNDArrays(arrays, true);
    if (getDynDimsTotalSize() != getNumNDArrays())
        throw std::invalid_argument("XData: temporal dimensions do not match the number of NDArrays");
}

namespace {
bool supportedElementType(ElementDataType t) {
    return t == TYPEID_COMPLEX || t == TYPEID_REAL || t == TYPEID_INDEX || t == TYPEID_UINT32 || t == TYPEID_INT32;
}
}  // namespace

XData::XData(const std::shared_ptr<CLapp>& app, NDArray*& nd, ElementDataType t) : Data(app, t) {
    if (!nd) throw std::invalid_argument("XData: null NDArray");
    if (!supportedElementType(t)) throw std::invalid_argument("XData: unsupported element type");
    auto* arrays = new std::vector<NDArray*>({nd});
    nd = nullptr;
    dynDims_ = {};
    setNDArrays(arrays, true);
}

XData::XData(const std::shared_ptr<CLapp>& app, std::vector<dimIndexType>*& dims, ElementDataType t) : Data(app, t) {
    if (!dims) throw std::invalid_argument("XData: null dims");
    if (!supportedElementType(t)) throw std::invalid_argument("XData: unsupported element type");
    NDArray* nd;
    if (t == TYPEID_COMPLEX) nd = new ConcreteNDArray<complexType>(dims);
    else if (t == TYPEID_REAL) nd = new ConcreteNDArray<realType>(dims);
    else if (t == TYPEID_UINT32) nd = new ConcreteNDArray<uint32_t>(dims);
    else if (t == TYPEID_INT32) nd = new ConcreteNDArray<int32_t>(dims);
    else nd = new ConcreteNDArray<uint8_t>(dims);
    auto* arrays = new std::vector<NDArray*>({nd});
    dynDims_ = {};
    setNDArrays(arrays, false);
    checkOcl(ocl_memset_zero(app_->getOclContext(), dev_, bytes_), "XData zero-init");
}

XData::XData(const std::shared_ptr<CLapp>& app, const std::shared_ptr<XData>& src, bool copyData, bool)
    : Data(app, TYPEID_COMPLEX) {
    if (!src) throw std::invalid_argument("XData: null source");
    cloneShape(*src, copyData);
}

XData::XData(const std::shared_ptr<CLapp>& app, const std::shared_ptr<KData>& src) : Data(app, TYPEID_COMPLEX) {
    if (!src) throw std::invalid_argument("XData: null source");
    if (!src->getAllSizesEqual())
        throw std::invalid_argument(
            "XData copy constructor with KData source not valid if KData NDArrays do not have the same spatial dimensions");
    auto* arrays = new std::vector<NDArray*>();
    for (dimIndexType f = 0; f < src->getDynDimsTotalSize(); ++f) {
        auto* dims = new std::vector<dimIndexType>(*src->getNDArray(0)->getDims());
        arrays->push_back(new ConcreteNDArray<complexType>(dims));
    }
    dynDims_ = *src->getDynDims();
    setNDArrays(arrays, false);
    checkOcl(ocl_memset_zero(app_->getOclContext(), dev_, bytes_), "XData zero-init");
}

KData::KData(const std::shared_ptr<CLapp>& app, std::vector<NDArray*>*& arrays, numCoilsType nCoils,
             std::vector<dimIndexType>*& dynDims)
    : Data(app, TYPEID_COMPLEX), nCoils_(nCoils) {
    if (nCoils < 1) throw std::invalid_argument("KData: number of coils must be positive");
    if (dynDims) {
        dynDims_ = *dynDims;
        delete dynDims;
        dynDims = nullptr;
    } else {
        dynDims_ = {(dimIndexType)((arrays ? arrays->size() : 0) / nCoils)};
    }
    setNDArrays(arrays, true);
    if ((size_t)getDynDimsTotalSize() * nCoils_ != getNumNDArrays())
        throw std::invalid_argument("KData: frames x coils does not match the number of NDArrays");
}

KData::KData(const std::shared_ptr<CLapp>& app, std::vector<std::vector<dimIndexType>*>*& arraysDims,
             numCoilsType nCoils, std::vector<dimIndexType>*& dynDims)
    : Data(app, TYPEID_COMPLEX), nCoils_(nCoils) {
    if (!arraysDims) throw std::invalid_argument("KData: null dims");
    auto* arrays = new std::vector<NDArray*>();
    for (auto* d : *arraysDims) arrays->push_back(new ConcreteNDArray<complexType>(d));
    delete arraysDims;
    arraysDims = nullptr;
    dynDims_ = *dynDims;
    delete dynDims;
    dynDims = nullptr;
    setNDArrays(arrays, false);
    checkOcl(ocl_memset_zero(app_->getOclContext(), dev_, bytes_), "KData zero-init");
}

KData::KData(const std::shared_ptr<CLapp>& app, const std::shared_ptr<KData>& src, bool copyData, bool copySens,
             bool copyMasks)
    : Data(app, TYPEID_COMPLEX) {
    if (!src) throw std::invalid_argument("KData: null source");
    nCoils_ = src->nCoils_;
    cloneShape(*src, copyData);
    if (copySens) sens_ = src->sens_;
    if (copyMasks) masks_ = src->masks_;
}

SensitivityMapsData::SensitivityMapsData(const std::shared_ptr<CLapp>& app, std::vector<NDArray*>*& data,
                                         numCoilsType nCoils)
    : Data(app, TYPEID_COMPLEX) {
    if (!data || data->size() != nCoils)
        throw std::invalid_argument("SensitivityMapsData: one NDArray per coil is required");
    dynDims_ = {};
    setNDArrays(data, true);
}

SamplingMasksData::SamplingMasksData(const std::shared_ptr<CLapp>& app, std::vector<NDArray*>*& data,
                                     std::vector<dimIndexType>*& dynDims, dimIndexType kDataNumCols,
                                     ElementDataType t)
    : Data(app, TYPEID_INDEX) {
    if (t != TYPEID_INDEX) throw std::invalid_argument("SamplingMasksData: masks are uint8 in this library");
    if (!data || data->empty()) throw std::invalid_argument("SamplingMasksData: no masks given");
    // ROWMASK input ({H} per frame) is expanded to a pixel mask of kDataNumCols columns (255 where sampled)
    auto* arrays = new std::vector<NDArray*>();
    for (NDArray* a : *data) {
        if (a->getDims()->size() == 1) {
            const dimIndexType H = a->getDims()->at(0);
            auto* dims = new std::vector<dimIndexType>({kDataNumCols, H});
            auto* pix = new ConcreteNDArray<uint8_t>(dims);
            const uint8_t* row = (const uint8_t*)a->getHostDataAsVoidPtr();
            uint8_t* dst = (uint8_t*)pix->getHostDataAsVoidPtr();
            for (dimIndexType h = 0; h < H; ++h) std::memset(dst + (size_t)h * kDataNumCols, row[h] ? 255 : 0, kDataNumCols);
            arrays->push_back(pix);
            delete a;
        } else {
            arrays->push_back(a);
        }
    }
    delete data;
    data = nullptr;
    if (dynDims) {
        dynDims_ = *dynDims;
        delete dynDims;
        dynDims = nullptr;
    } else {
        dynDims_ = {(dimIndexType)arrays->size()};
    }
    setNDArrays(arrays, true);
}

}  // namespace OpenCLIPER
