// MATLAB Level 5 MAT-file reader / writer (numeric arrays), written from the published MAT-file format: a 128-byte
// header (text, subsystem offset, version 0x0100, endian indicator "IM"), then data elements, each an 8-byte tag
// (type, byte count; or the 4-byte "small element" form when the count is at most 4) followed by the data padded to
// 8 bytes.  A variable is an miMATRIX element holding array flags (class, complex bit), dimensions, name, and the
// real and imaginary parts; miCOMPRESSED elements hold one zlib-deflated element.
#include <OpenCLIPER/MatFile.hpp>

#include <zlib.h>

#include <cstring>
#include <ctime>
#include <fstream>
#include <stdexcept>

namespace OpenCLIPER {
namespace mat {

namespace {

enum MiType : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6, miSINGLE = 7, miDOUBLE = 9,
    miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15
};
const uint32_t kComplexFlag = 0x0800;

size_t typeSize(uint32_t t) {
    switch (t) {
        case miINT8: case miUINT8: return 1;
        case miINT16: case miUINT16: return 2;
        case miINT32: case miUINT32: case miSINGLE: return 4;
        case miDOUBLE: case miINT64: case miUINT64: return 8;
        default: return 0;
    }
}

bool numericClass(int c) { return c >= mxDOUBLE && c <= mxUINT64; }

uint32_t storageType(int mclass) {
    switch (mclass) {
        case mxDOUBLE: return miDOUBLE;
        case mxSINGLE: return miSINGLE;
        case mxINT8: return miINT8;
        case mxUINT8: return miUINT8;
        case mxINT16: return miINT16;
        case mxUINT16: return miUINT16;
        case mxINT32: return miINT32;
        case mxUINT32: return miUINT32;
        case mxINT64: return miINT64;
        default: return miUINT64;
    }
}

template <class T> T rd(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

double elementAsDouble(uint32_t t, const uint8_t* p) {
    switch (t) {
        case miINT8: return rd<int8_t>(p);
        case miUINT8: return rd<uint8_t>(p);
        case miINT16: return rd<int16_t>(p);
        case miUINT16: return rd<uint16_t>(p);
        case miINT32: return rd<int32_t>(p);
        case miUINT32: return rd<uint32_t>(p);
        case miSINGLE: return rd<float>(p);
        case miDOUBLE: return rd<double>(p);
        case miINT64: return (double)rd<int64_t>(p);
        case miUINT64: return (double)rd<uint64_t>(p);
        default: throw std::invalid_argument("MAT-file: unsupported element type " + std::to_string(t));
    }
}

// One data element inside [p, end): returns its type, data pointer and byte count, and advances p past it
struct Element {
    uint32_t type;
    const uint8_t* data;
    uint32_t bytes;
};

Element nextElement(const uint8_t*& p, const uint8_t* end, bool pad) {
    if (end - p < 8) throw std::invalid_argument("MAT-file: truncated data element");
    const uint32_t w0 = rd<uint32_t>(p);
    Element e;
    if (w0 >> 16) {   // small data element: byte count in the upper half, data in the second word
        e.type = w0 & 0xFFFF;
        e.bytes = w0 >> 16;
        if (e.bytes > 4) throw std::invalid_argument("MAT-file: small data element longer than 4 bytes");
        e.data = p + 4;
        p += 8;
        return e;
    }
    e.type = w0;
    e.bytes = rd<uint32_t>(p + 4);
    e.data = p + 8;
    if ((size_t)(end - e.data) < e.bytes) throw std::invalid_argument("MAT-file: truncated data element");
    p = e.data + (pad ? (e.bytes + 7) / 8 * 8 : e.bytes);
    if (p > end) p = end;
    return e;
}

std::vector<double> convertPart(const Element& e) {
    const size_t s = typeSize(e.type);
    if (!s) throw std::invalid_argument("MAT-file: unsupported numeric storage type " + std::to_string(e.type));
    std::vector<double> v(e.bytes / s);
    for (size_t i = 0; i < v.size(); ++i) v[i] = elementAsDouble(e.type, e.data + i * s);
    return v;
}

// Parses the contents of one miMATRIX element; returns false for non-numeric classes
bool parseMatrix(const uint8_t* p, const uint8_t* end, Var& v) {
    const Element flags = nextElement(p, end, true);
    if (flags.type != miUINT32 || flags.bytes < 8) throw std::invalid_argument("MAT-file: bad array flags");
    const uint32_t f0 = rd<uint32_t>(flags.data);
    v.mclass = (int)(f0 & 0xFF);
    v.complex = (f0 & kComplexFlag) != 0;
    if (!numericClass(v.mclass)) return false;
    const Element dims = nextElement(p, end, true);
    if (dims.type != miINT32) throw std::invalid_argument("MAT-file: bad dimensions");
    v.dims.clear();
    for (uint32_t i = 0; i < dims.bytes / 4; ++i) v.dims.push_back((uint64_t)rd<int32_t>(dims.data + 4 * i));
    const Element name = nextElement(p, end, true);
    v.name.assign((const char*)name.data, name.bytes);
    v.re = convertPart(nextElement(p, end, true));
    if (v.complex) v.im = convertPart(nextElement(p, end, true));
    if (v.re.size() != v.numel() || (v.complex && v.im.size() != v.numel()))
        throw std::invalid_argument("MAT-file: variable " + v.name + " has a data size that does not match its dims");
    return true;
}

std::vector<uint8_t> inflateAll(const uint8_t* src, size_t n) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) throw std::runtime_error("MAT-file: zlib init failed");
    std::vector<uint8_t> out(std::max<size_t>(4 * n, 1024));
    zs.next_in = const_cast<Bytef*>(src);
    zs.avail_in = (uInt)n;
    int rc = Z_OK;
    while (rc != Z_STREAM_END) {
        if (zs.total_out == out.size()) out.resize(out.size() * 2);
        zs.next_out = out.data() + zs.total_out;
        zs.avail_out = (uInt)(out.size() - zs.total_out);
        rc = inflate(&zs, Z_NO_FLUSH);
        if (rc == Z_BUF_ERROR && zs.avail_in == 0) break;   // input ended without an end marker
        if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR) {
            inflateEnd(&zs);
            throw std::invalid_argument("MAT-file: corrupt compressed element");
        }
    }
    out.resize(zs.total_out);
    inflateEnd(&zs);
    return out;
}

void put32(std::vector<uint8_t>& b, uint32_t v) {
    const size_t o = b.size();
    b.resize(o + 4);
    std::memcpy(b.data() + o, &v, 4);
}

void putElement(std::vector<uint8_t>& b, uint32_t type, const void* data, size_t bytes) {
    put32(b, type);
    put32(b, (uint32_t)bytes);
    const size_t o = b.size();
    b.resize(o + (bytes + 7) / 8 * 8, 0);
    if (bytes) std::memcpy(b.data() + o, data, bytes);
}

template <class T> std::vector<uint8_t> packAs(const std::vector<double>& v) {
    std::vector<uint8_t> out(v.size() * sizeof(T));
    for (size_t i = 0; i < v.size(); ++i) {
        const T x = (T)v[i];
        std::memcpy(out.data() + i * sizeof(T), &x, sizeof(T));
    }
    return out;
}

std::vector<uint8_t> packPart(int mclass, const std::vector<double>& v) {
    switch (mclass) {
        case mxDOUBLE: return packAs<double>(v);
        case mxSINGLE: return packAs<float>(v);
        case mxINT8: return packAs<int8_t>(v);
        case mxUINT8: return packAs<uint8_t>(v);
        case mxINT16: return packAs<int16_t>(v);
        case mxUINT16: return packAs<uint16_t>(v);
        case mxINT32: return packAs<int32_t>(v);
        case mxUINT32: return packAs<uint32_t>(v);
        case mxINT64: return packAs<int64_t>(v);
        default: return packAs<uint64_t>(v);
    }
}

std::vector<uint8_t> matrixElement(const Var& v) {
    std::vector<uint8_t> body;
    const uint32_t flags[2] = {(uint32_t)v.mclass | (v.complex ? kComplexFlag : 0u), 0u};
    putElement(body, miUINT32, flags, 8);
    std::vector<int32_t> d;
    for (uint64_t x : v.dims) d.push_back((int32_t)x);
    putElement(body, miINT32, d.data(), 4 * d.size());
    putElement(body, miINT8, v.name.data(), v.name.size());
    const std::vector<uint8_t> re = packPart(v.mclass, v.re);
    putElement(body, storageType(v.mclass), re.data(), re.size());
    if (v.complex) {
        const std::vector<uint8_t> im = packPart(v.mclass, v.im);
        putElement(body, storageType(v.mclass), im.data(), im.size());
    }
    std::vector<uint8_t> el;
    put32(el, miMATRIX);
    put32(el, (uint32_t)body.size());
    el.insert(el.end(), body.begin(), body.end());
    return el;
}

}  // namespace

uint64_t Var::numel() const {
    uint64_t n = 1;
    for (uint64_t d : dims) n *= d;
    return dims.empty() ? 0 : n;
}

std::vector<std::complex<float>> Var::asComplexFloat() const {
    std::vector<std::complex<float>> out(re.size());
    for (size_t i = 0; i < re.size(); ++i) out[i] = {(float)re[i], complex ? (float)im[i] : 0.f};
    return out;
}

std::vector<uint32_t> Var::asUint32() const {
    std::vector<uint32_t> out(re.size());
    for (size_t i = 0; i < re.size(); ++i) out[i] = (uint32_t)re[i];
    return out;
}

std::vector<uint8_t> Var::asUint8() const {
    std::vector<uint8_t> out(re.size());
    for (size_t i = 0; i < re.size(); ++i) out[i] = (uint8_t)re[i];
    return out;
}

bool isMatFile(const std::string& fileName) {
    std::ifstream f(fileName, std::ios::binary);
    uint8_t h[128];
    if (!f.read((char*)h, sizeof(h))) return false;
    return h[124] == 0x00 && h[125] == 0x01 && h[126] == 'I' && h[127] == 'M';
}

std::map<std::string, Var> read(const std::string& fileName) {
    std::ifstream f(fileName, std::ios::binary | std::ios::ate);
    if (!f) throw std::invalid_argument("MAT-file: cannot open " + fileName);
    const std::streamsize size = f.tellg();
    std::vector<uint8_t> buf(size > 0 ? (size_t)size : 0);
    f.seekg(0);
    if (size > 0 && !f.read((char*)buf.data(), size)) throw std::invalid_argument("MAT-file: cannot read " + fileName);
    if (buf.size() < 128 || buf[126] != 'I' || buf[127] != 'M')
        throw std::invalid_argument("MAT-file: " + fileName + " is not a little-endian MAT v5 file");
    if (buf[124] != 0x00 || buf[125] != 0x01)
        throw std::invalid_argument("MAT-file: " + fileName + " is not a MAT v5 file (v7.3 HDF5 files are not read here)");
    std::map<std::string, Var> vars;
    const uint8_t* p = buf.data() + 128;
    const uint8_t* end = buf.data() + buf.size();
    while (end - p >= 8) {
        const Element e = nextElement(p, end, true);
        if (e.type == miCOMPRESSED) {
            // compressed elements are not padded
            p = e.data + e.bytes;
            const std::vector<uint8_t> raw = inflateAll(e.data, e.bytes);
            const uint8_t* q = raw.data();
            const Element m = nextElement(q, raw.data() + raw.size(), true);
            Var v;
            if (m.type == miMATRIX && parseMatrix(m.data, m.data + m.bytes, v)) vars[v.name] = std::move(v);
        } else if (e.type == miMATRIX) {
            Var v;
            if (parseMatrix(e.data, e.data + e.bytes, v)) vars[v.name] = std::move(v);
        }
    }
    return vars;
}

void write(const std::string& fileName, const std::vector<Var>& vars, bool compress) {
    std::vector<uint8_t> out(128, ' ');
    char text[117];
    std::time_t now = std::time(nullptr);
    char date[64];
    std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y", std::gmtime(&now));
    std::snprintf(text, sizeof(text), "MATLAB 5.0 MAT-file, Platform: GLNXA64, Created on: %s", date);
    std::memcpy(out.data(), text, std::strlen(text));
    std::memset(out.data() + 116, 0, 8);   // no subsystem data
    out[124] = 0x00;
    out[125] = 0x01;                       // version 0x0100
    out[126] = 'I';
    out[127] = 'M';
    for (const Var& v : vars) {
        if (!numericClass(v.mclass)) throw std::invalid_argument("MAT-file: only numeric variables can be written");
        if (v.re.size() != v.numel() || (v.complex && v.im.size() != v.numel()))
            throw std::invalid_argument("MAT-file: variable " + v.name + " has a data size that does not match its dims");
        const std::vector<uint8_t> el = matrixElement(v);
        if (compress) {
            uLongf n = compressBound(el.size());
            std::vector<uint8_t> z(n);
            if (compress2(z.data(), &n, el.data(), el.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
                throw std::runtime_error("MAT-file: zlib compression failed");
            put32(out, miCOMPRESSED);
            put32(out, (uint32_t)n);
            out.insert(out.end(), z.begin(), z.begin() + n);
        } else {
            out.insert(out.end(), el.begin(), el.end());
        }
    }
    std::ofstream f(fileName, std::ios::binary);
    if (!f) throw std::invalid_argument("MAT-file: cannot create " + fileName);
    f.write((const char*)out.data(), (std::streamsize)out.size());
    if (!f) throw std::runtime_error("MAT-file: write failed for " + fileName);
}

Var makeComplexSingle(const std::string& name, const std::vector<uint64_t>& dims, const std::complex<float>* data) {
    Var v;
    v.name = name;
    v.mclass = mxSINGLE;
    v.complex = true;
    v.dims = dims;
    const uint64_t n = v.numel();
    v.re.resize(n);
    v.im.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        v.re[i] = data[i].real();
        v.im[i] = data[i].imag();
    }
    return v;
}

Var makeUint32(const std::string& name, const std::vector<uint64_t>& dims, const uint32_t* data) {
    Var v;
    v.name = name;
    v.mclass = mxUINT32;
    v.dims = dims;
    v.re.assign(data, data + v.numel());
    return v;
}

Var makeUint8(const std::string& name, const std::vector<uint64_t>& dims, const uint8_t* data) {
    Var v;
    v.name = name;
    v.mclass = mxUINT8;
    v.dims = dims;
    v.re.assign(data, data + v.numel());
    return v;
}

}  // namespace mat
}  // namespace OpenCLIPER
