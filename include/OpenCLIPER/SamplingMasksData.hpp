// Sampling masks, one NDArray per frame (include/OpenCLIPER/SamplingMasksData.hpp).  Stored on the device as
// per-pixel bytes [F][H][W], non-zero = sampled (the reference's PIXELMASK format; ROWMASK input with one entry per
// k-space line is expanded across kDataNumCols columns, as SamplingMasksData.cpp:389-441 does).
#pragma once
#include <OpenCLIPER/Data.hpp>

namespace OpenCLIPER {

class SamplingMasksData : public Data {
  public:
    enum MasksFormat { ROWMASK, ROWSTOBLANK, PIXELMASK };
    /// pData: per frame either {W, H} uint8 pixel masks or {H} uint8 row masks (ROWMASK).
    SamplingMasksData(const std::shared_ptr<CLapp>& pCLapp, std::vector<NDArray*>*& pData,
                      std::vector<dimIndexType>*& pDynDims, dimIndexType kDataNumCols,
                      ElementDataType elementDataType = TYPEID_INDEX);
};

}  // namespace OpenCLIPER
