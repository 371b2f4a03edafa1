// ProcessCore: input/output binding and parameter structs of every Process (include/OpenCLIPER/ProcessCore.hpp:35-272).
#pragma once
#include <OpenCLIPER/Data.hpp>

namespace OpenCLIPER {

class ProcessCore {
  public:
    struct InitParameters {
        virtual ~InitParameters() {}
    };
    struct LaunchParameters {
        virtual ~LaunchParameters() {}
    };
    struct ProfileParameters {
        bool enable;
        unsigned long loops;
        ProfileParameters() : enable(false), loops(1) {}
        ProfileParameters(bool e, unsigned long l) : enable(e), loops(l) {}
        virtual ~ProfileParameters() {}
    };

    explicit ProcessCore(const std::shared_ptr<ProfileParameters>& pPP = nullptr)
        : pProfileParameters(pPP ? pPP : std::make_shared<ProfileParameters>()) {}
    virtual ~ProcessCore() {}

    void setInput(std::shared_ptr<Data> pInputData) { input_ = pInputData; }
    void setOutput(std::shared_ptr<Data> pOutputData) { output_ = pOutputData; }
    virtual void setInitParameters(const std::shared_ptr<InitParameters>& p) { pInitParameters = p; }
    virtual void setLaunchParameters(const std::shared_ptr<LaunchParameters>& p) { pLaunchParameters = p; }
    virtual const std::shared_ptr<InitParameters> getInitParameters() const { return pInitParameters; }
    virtual const std::shared_ptr<LaunchParameters> getLaunchParameters() const { return pLaunchParameters; }
    virtual void init() {}
    virtual void launch() = 0;
    /// Device time of the last launch in milliseconds when profiling is enabled (ProcessCore.cpp:184-241 keeps
    /// per-kernel event times; here one pair of HIP events brackets the launch).
    double getLastLaunchMs() const { return lastMs_; }

  protected:
    std::shared_ptr<Data> getInput() { return input_; }
    std::shared_ptr<Data> getOutput() { return output_; }

    std::shared_ptr<InitParameters> pInitParameters;
    std::shared_ptr<LaunchParameters> pLaunchParameters;
    std::shared_ptr<ProfileParameters> pProfileParameters;
    double lastMs_ = 0.0;

  private:
    std::shared_ptr<Data> input_, output_;
};

}  // namespace OpenCLIPER
