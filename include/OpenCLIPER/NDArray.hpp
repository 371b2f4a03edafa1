// Host-side n-dimensional arrays (include/OpenCLIPER/NDArray.hpp, ConcreteNDArray.hpp of the reference).
// dims[0] is the width (fastest), dims[1] the height, dims[2] the depth.
#pragma once
#include <OpenCLIPER/defs.hpp>

#define NDARRAYWIDTH(a) ((a)->getDims()->at(0))
#define NDARRAYHEIGHT(a) ((a)->getDims()->size() > 1 ? (a)->getDims()->at(1) : 1)
#define NDARRAYDEPTH(a) ((a)->getDims()->size() > 2 ? (a)->getDims()->at(2) : 0)

namespace OpenCLIPER {

class NDArray {
  public:
    virtual ~NDArray() {}
    const std::vector<dimIndexType>* getDims() const { return dims_.get(); }
    /// number of elements
    dimIndexType size() const {
        dimIndexType n = 1;
        for (auto d : *dims_) n *= d;
        return n;
    }
    virtual ElementDataType getElementDataType() const = 0;
    virtual const void* getHostDataAsVoidPtr() const = 0;
    virtual void* getHostDataAsVoidPtr() = 0;
    static dimIndexType getElementSize(ElementDataType t) {
        if (t == TYPEID_COMPLEX) return sizeof(complexType);
        if (t == TYPEID_REAL) return sizeof(realType);
        if (t == TYPEID_UINT32 || t == TYPEID_INT32) return 4;
        return 1;
    }

  protected:
    std::unique_ptr<std::vector<dimIndexType>> dims_;
};

template <class T> class ConcreteNDArray : public NDArray {
  public:
    /// Takes ownership of both vectors (the caller's pointers are set to nullptr, as in the reference).
    ConcreteNDArray(std::vector<dimIndexType>*& pDims, std::vector<T>*& pData) {
        dims_.reset(pDims);
        data_.reset(pData);
        pDims = nullptr;
        pData = nullptr;
        if (data_->size() != size()) throw std::invalid_argument("ConcreteNDArray: data size does not match dims");
    }
    /// Zero-filled array of the given dims.
    explicit ConcreteNDArray(std::vector<dimIndexType>*& pDims) {
        dims_.reset(pDims);
        pDims = nullptr;
        data_.reset(new std::vector<T>(size()));
    }
    ElementDataType getElementDataType() const override { return std::type_index(typeid(T)); }
    const void* getHostDataAsVoidPtr() const override { return data_->data(); }
    void* getHostDataAsVoidPtr() override { return data_->data(); }
    std::vector<T>* getHostData() { return data_.get(); }
    const std::vector<T>* getHostData() const { return data_.get(); }

  private:
    std::unique_ptr<std::vector<T>> data_;
};

}  // namespace OpenCLIPER
