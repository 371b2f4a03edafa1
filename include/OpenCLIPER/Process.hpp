// Process: creation through Process::create<T> (include/OpenCLIPER/Process.hpp:35-110).
#pragma once
#include <OpenCLIPER/ProcessCore.hpp>
#include <type_traits>

namespace OpenCLIPER {

class Process : public ProcessCore {
  public:
    virtual ~Process() {}
    const std::shared_ptr<CLapp> getApp() const { return app_; }
    void setApp(const std::shared_ptr<CLapp>& pCLapp) { app_ = pCLapp; }
    /// The reference names the OpenCL source file a Process compiles at run time.  Here every kernel is compiled ahead
    /// of time for gfx950 into libocl_amd.so, so no Process has a source file to name: this returns "" for all of them.
    virtual const std::string getKernelFile() const { return ""; }

    template <typename T>
    static std::shared_ptr<T> create(const std::shared_ptr<CLapp>& pCLapp,
                                     const std::shared_ptr<ProfileParameters>& pPP = nullptr) {
        static_assert(std::is_base_of<Process, T>::value, "Process::create: class does not derive from Process");
        auto p = std::shared_ptr<T>(new T(pCLapp, pPP));
        pCLapp->addKernelFile(p->getKernelFile());
        return p;
    }
    template <typename T>
    static std::shared_ptr<T> create(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<Data> pIn,
                                     const std::shared_ptr<Data> pOut,
                                     const std::shared_ptr<ProfileParameters>& pPP = nullptr) {
        static_assert(std::is_base_of<Process, T>::value, "Process::create: class does not derive from Process");
        auto p = std::shared_ptr<T>(new T(pCLapp, pPP));
        p->setInput(pIn);
        p->setOutput(pOut);
        pCLapp->addKernelFile(p->getKernelFile());
        return p;
    }

  protected:
    Process(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<ProfileParameters>& pPP = nullptr)
        : ProcessCore(pPP), app_(pCLapp) {
        if (!pCLapp) throw std::invalid_argument("Invalid CLapp pointer");
    }
    void startProfiling();
    void stopProfiling();

  private:
    std::shared_ptr<CLapp> app_;
    void* ev_[2] = {nullptr, nullptr};
};

#define OCL_PROCESS_BOILERPLATE(Cls)                                                              \
  protected:                                                                                      \
    Cls(const std::shared_ptr<CLapp>& pCLapp, const std::shared_ptr<ProfileParameters>& pPP = nullptr) \
        : Process(pCLapp, pPP) {}                                                                 \
    friend class Process;

}  // namespace OpenCLIPER
