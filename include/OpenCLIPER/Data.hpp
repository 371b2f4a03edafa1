// Data: a group of same-typed NDArrays with one HBM-resident device copy (include/OpenCLIPER/Data.hpp:47-348).
// Device layout: the NDArrays back to back in index order, no padding and no in-band header (the reference prefixes
// a dims/strides header, src/DeviceDataProperties.cpp:156-230; here shapes stay on the host).
#pragma once
#include <OpenCLIPER/CLapp.hpp>
#include <OpenCLIPER/NDArray.hpp>

namespace OpenCLIPER {

class Data : public std::enable_shared_from_this<Data> {
  public:
    virtual ~Data();
    Data(const Data&) = delete;
    Data& operator=(const Data&) = delete;

    const std::vector<const NDArray*>* getData() const { return &view_; }
    const NDArray* getNDArray(dimIndexType index) const { return view_.at(index); }
    dimIndexType getNumNDArrays() const { return (dimIndexType)view_.size(); }
    const void* getHostBuffer(dimIndexType index = 0) const { return arrays_.at(index)->getHostDataAsVoidPtr(); }
    void* getHostBuffer(dimIndexType index = 0) { return arrays_.at(index)->getHostDataAsVoidPtr(); }
    /// HBM address of NDArray 0 (Data::getHIPDeviceBuffer in the reference, Data.hpp:132)
    void* getHIPDeviceBuffer() { return dev_; }
    void* getDeviceBuffer() { return dev_; }
    size_t getDeviceBufferBytes() const { return bytes_; }
    ElementDataType getElementDataType() const { return type_; }
    const std::vector<dimIndexType>* getDynDims() const { return &dynDims_; }
    dimIndexType getDynDimsTotalSize() const;
    bool getAllSizesEqual() const;
    const std::shared_ptr<CLapp>& getApp() const { return app_; }

    /// Copies every NDArray to the device buffer (synchronous).
    void host2Device();
    /// Copies the device buffer back into the NDArrays (synchronous; the reference's blocking read, Data.cpp:1493).
    void device2Host();

  protected:
    Data(const std::shared_ptr<CLapp>& app, ElementDataType t) : app_(app), type_(t) {}
    /// Takes ownership of the NDArrays (the caller's pointer is set to nullptr), allocates the device copy and, if
    /// upload is true, fills it.
    void setNDArrays(std::vector<NDArray*>*& arrays, bool upload);
    /// Same shapes as src, device memory allocated; copyData copies src's device contents.
    void cloneShape(const Data& src, bool copyData);

    std::shared_ptr<CLapp> app_;
    ElementDataType type_;
    std::vector<std::unique_ptr<NDArray>> arrays_;
    std::vector<const NDArray*> view_;
    std::vector<dimIndexType> dynDims_;
    void* dev_ = nullptr;
    size_t bytes_ = 0;
};

}  // namespace OpenCLIPER
