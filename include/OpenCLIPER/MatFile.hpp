// MATLAB Level 5 MAT-file reading and writing for the Data classes (the role libmatio plays for the reference's
// Data::readMatlabVariablesFromFile / matlabSave, src/Data.cpp:615-676,1477-1483).  Numeric arrays only (real or
// complex; double, single, int8..uint64), uncompressed or zlib-compressed (miCOMPRESSED) elements, little-endian.
// Host-side only: nothing here touches the device.
#pragma once
#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace OpenCLIPER {
namespace mat {

/// MATLAB array classes (mxClassID) of the numeric types handled here
enum MatClass : int {
    mxDOUBLE = 6, mxSINGLE = 7, mxINT8 = 8, mxUINT8 = 9, mxINT16 = 10, mxUINT16 = 11, mxINT32 = 12, mxUINT32 = 13,
    mxINT64 = 14, mxUINT64 = 15
};

/// One numeric variable.  dims are in MATLAB order (dims[0] varies fastest); re / im hold the elements converted to
/// double (im is empty for a real array), which is exact for every class except 64-bit integers beyond 2^53.
struct Var {
    std::string name;
    int mclass = mxDOUBLE;
    bool complex = false;
    std::vector<uint64_t> dims;
    std::vector<double> re, im;

    uint64_t numel() const;
    std::vector<std::complex<float>> asComplexFloat() const;
    std::vector<uint32_t> asUint32() const;
    std::vector<uint8_t> asUint8() const;
};

/// True when the file starts with a MAT v5 header (the reference tells MAT input from CFL input by whether libmatio
/// opens it, KData.cpp:300-302).
bool isMatFile(const std::string& fileName);

/// Reads every numeric variable of a MAT v5 file (other variable kinds are skipped).  Throws std::invalid_argument
/// when the file is not a little-endian MAT v5 file.
std::map<std::string, Var> read(const std::string& fileName);

/// Writes the variables, in order, as a MAT v5 file; compress = true stores each as a zlib miCOMPRESSED element (what
/// the reference's matlabSave asks libmatio for, MAT_COMPRESSION_ZLIB).  Elements are stored in their class's type.
void write(const std::string& fileName, const std::vector<Var>& vars, bool compress = true);

/// Helpers building variables from host buffers (dims in MATLAB order)
Var makeComplexSingle(const std::string& name, const std::vector<uint64_t>& dims, const std::complex<float>* data);
Var makeUint32(const std::string& name, const std::vector<uint64_t>& dims, const uint32_t* data);
Var makeUint8(const std::string& name, const std::vector<uint64_t>& dims, const uint8_t* data);

}  // namespace mat
}  // namespace OpenCLIPER
