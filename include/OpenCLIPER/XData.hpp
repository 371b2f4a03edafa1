// XData: image-space data, one NDArray per frame (include/OpenCLIPER/XData.hpp:56-196).
#pragma once
#include <OpenCLIPER/Data.hpp>

namespace OpenCLIPER {

class KData;

class XData : public Data {
  public:
    /// From a group of NDArrays (one per frame) and the temporal dimensions (ownership is taken).
    XData(const std::shared_ptr<CLapp>& pCLapp, std::vector<NDArray*>*& pNDArrays, std::vector<dimIndexType>*& pDynDims);
    /// Same shape as sourceData; copyData copies its contents (XData.cpp:403-406).
    XData(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<XData>& sourceData, bool copyData,
          bool copyDataToDevice = true);
    /// Uninitialised image with one frame per KData frame (XData.cpp:425-440).
    XData(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<KData>& sourceData);
    /// One NDArray of any supported element type (complex, real, uint8, uint32, int32) and no temporal dimension:
    /// the form the reference's registration uses for meshes and deformations (XData.cpp, XData(pCLapp, NDArray*&,
    /// ElementDataType)).  Ownership of the NDArray is taken.
    XData(const std::shared_ptr<CLapp>& pCLapp, NDArray*& pNDArray, ElementDataType elementDataType);
    /// A zero-filled single NDArray of the given dims and element type.
    XData(const std::shared_ptr<CLapp>& pCLapp, std::vector<dimIndexType>*& pDims, ElementDataType elementDataType);
    /// From a MAT v5 file holding Dims and XData (XData.cpp:500-556).
    XData(const std::shared_ptr<CLapp>& pCLapp, const std::string& fileName);

    /// Saves Dims and XData to a zlib-compressed MAT v5 file (XData.cpp:898-930).
    void matlabSave(const std::string& fileName);
};

}  // namespace OpenCLIPER
