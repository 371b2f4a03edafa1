// KData: multi-coil k-space, NDArrays ordered frame-major with the coil index fastest (index = coil + C * frame, as
// hostKernelFunctions.cl:175-183 of the reference lays them out), plus the attached sensitivity maps and sampling
// masks (include/OpenCLIPER/KData.hpp:66-351).
#pragma once
#include <OpenCLIPER/SamplingMasksData.hpp>
#include <OpenCLIPER/SensitivityMapsData.hpp>

namespace OpenCLIPER {

class KData : public Data {
  public:
    /// From a group of NDArrays (F*C of them, coil fastest), the number of coils and the temporal dimensions.
    KData(const std::shared_ptr<CLapp>& pCLapp, std::vector<NDArray*>*& pNDArrays, numCoilsType nCoils,
          std::vector<dimIndexType>*& pDynDims);
    /// Uninitialised k-space of the given per-NDArray dims (KData.hpp:101).
    KData(const std::shared_ptr<CLapp>& pCLapp, std::vector<std::vector<dimIndexType>*>*& pArraysDims,
          numCoilsType nCoils, std::vector<dimIndexType>*& pDynDims);
    /// Same shape as sourceData (KData.cpp:145-213): copyData copies the samples, copySensitivityMaps /
    /// copySamplingMasks share the source's maps / masks.
    KData(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<KData>& sourceData, bool copyData = false,
          bool copySensitivityMaps = false, bool copySamplingMasks = false);
    /// From a MAT v5 file holding Dims, KData and optionally SensitivityMaps and SamplingMasks (KData.cpp:292-435), or,
    /// when the file is not a MAT file, from the CFL set <base>.cfl + <base>.hdr with optional
    /// <base>_SensitivityMaps.cfl and <base>_SamplingMasks.u32 (row masks) or .u8 (pixel masks) (KData.cpp:246-345).
    KData(const std::shared_ptr<CLapp>& pCLapp, const std::string& fileName);

    /// Writes <base>.cfl, <base>.hdr and, when attached, <base>_SensitivityMaps.cfl and <base>_SamplingMasks.u8
    /// (KData.cpp:688-710).
    void saveCFLData(const std::string& baseFileName);

    /// Saves Dims, KData and (when attached) SensitivityMaps to a zlib-compressed MAT v5 file (KData.cpp:837-936);
    /// saveSamplingMasks also writes the masks as a uint8 pixel mask [H, W, temporal...], which the reference leaves out.
    void matlabSave(const std::string& fileName, bool saveSamplingMasks = false);

    numCoilsType getNCoils() const { return nCoils_; }
    std::shared_ptr<SensitivityMapsData> getSensitivityMapsData() const { return sens_; }
    std::shared_ptr<SamplingMasksData> getSamplingMasksData() const { return masks_; }
    void setSensitivityMapsData(SensitivityMapsData*& p) { sens_.reset(p); p = nullptr; }
    void setSensitivityMapsData(const std::shared_ptr<SensitivityMapsData>& p) { sens_ = p; }
    void setSamplingMasksData(SamplingMasksData*& p) { masks_.reset(p); p = nullptr; }
    void setSamplingMasksData(const std::shared_ptr<SamplingMasksData>& p) { masks_ = p; }

  private:
    void loadMat(const std::string& fileName);
    void loadCFL(const std::string& fileName);

    numCoilsType nCoils_ = 1;
    std::shared_ptr<SensitivityMapsData> sens_;
    std::shared_ptr<SamplingMasksData> masks_;
};

}  // namespace OpenCLIPER
