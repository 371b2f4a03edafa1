// CLapp: the device context of the OpenCLIPER API (include/OpenCLIPER/CLapp.hpp:114-359 of the reference).
// Here it owns one HIP device and one in-order stream through the C ABI of libocl_amd.so; kernels are compiled ahead
// of time for gfx950, so loadKernels()/addKernelFile() have nothing left to do.
#pragma once
#include <OpenCLIPER/defs.hpp>

struct ocl_ctx;

namespace OpenCLIPER {

/// Error raised by the API (the reference's CLError keeps an error code and a message, CLapp.hpp:114-144).
class CLError : public std::exception {
  public:
    CLError(int err, const std::string& msg) : err_(err), msg_(msg) {}
    int err() const { return err_; }
    const char* what() const noexcept override { return msg_.c_str(); }

  private:
    int err_;
    std::string msg_;
};

/// Throws CLError with the C ABI's last message when rc != 0.
void checkOcl(int rc, const char* where);

class CLapp {
  public:
    /// Accepted for source compatibility; a HIP build has a single platform.
    struct PlatformTraits {
        std::string vendor;
    };
    /// Device selection (CLapp::DeviceTraits of the reference); useHIP is always true.
    struct DeviceTraits {
        int deviceNumber = 0;
        bool useHIP = true;
    };

    static std::shared_ptr<CLapp> create();
    static std::shared_ptr<CLapp> create(const PlatformTraits& platformTraits, const DeviceTraits& deviceTraits);
    ~CLapp();

    void init(const PlatformTraits& platformTraits, const DeviceTraits& deviceTraits);
    void loadKernels() {}
    void addKernelFile(const std::string&) {}

    /// The in-order stream every Process of this CLapp enqueues on (getCommandQueue() in the reference).
    void* getHIPStream() const;
    /// The C ABI context handle.
    ocl_ctx* getOclContext() const { return ctx_; }
    std::string getDeviceName() const;
    size_t getDeviceGlobalMemory() const;
    void finish() const;

  private:
    CLapp() = default;
    ocl_ctx* ctx_ = nullptr;
};

}  // namespace OpenCLIPER
