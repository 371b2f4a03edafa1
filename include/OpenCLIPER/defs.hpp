// Basic types of the OpenCLIPER-compatible C++ API (include/OpenCLIPER/defs.hpp and kernels/defs.h of the reference
// define the same names).  Single precision only; complex data is std::complex<float>.
#pragma once
#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <vector>

namespace OpenCLIPER {
typedef uint32_t dimIndexType;
typedef uint32_t numCoilsType;
typedef float realType;
typedef std::complex<float> complexType;
typedef std::type_index ElementDataType;
}  // namespace OpenCLIPER

#define TYPEID_COMPLEX std::type_index(typeid(OpenCLIPER::complexType))
#define TYPEID_REAL std::type_index(typeid(OpenCLIPER::realType))
#define TYPEID_INDEX std::type_index(typeid(uint8_t))   // sampling masks are stored as bytes (non-zero = sampled)
#define TYPEID_UINT32 std::type_index(typeid(uint32_t)) // meshes and index arrays (the reference's dimIndexType)
#define TYPEID_INT32 std::type_index(typeid(int32_t))
