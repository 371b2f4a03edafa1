// NESTA with temporal-TV regularisation and mu continuation (include/OpenCLIPER/processes/nesta/NestaUp.hpp).
// Input: KData with attached SensitivityMapsData and SamplingMasksData; output: XData (one image per frame).
#pragma once
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/XData.hpp>
#include <OpenCLIPER/Process.hpp>
#include <OpenCLIPER/processes/GroupwiseRegistration.hpp>

namespace OpenCLIPER {

class NestaUp : public Process {
  public:
    /// Same fields and constructor argument order as the reference (NestaUp.hpp:67-83).  Like the reference,
    /// launch() overwrites tolVar (NestaUp.cpp:583,609).  With argsMC the sparsity operator is tTV o MC and the
    /// output's current contents are the starting estimate (NestaUp.cpp:201-278,483-491); that case runs the fused
    /// form (SPLIT falls back to it) when the mesh is GroupwiseRegistration's original one, else the PLAIN form.
    struct LaunchParameters : Process::LaunchParameters {
        float lambda_i;
        float mu_f;
        float La;
        uint32_t maxIntIter;
        float tolVar;
        int verbose;
        unsigned int maxIter;
        uint32_t stoptest;
        uint32_t miniter;
        ArgumentsMotionCompensation* argsMC;
        bool showProgress;
        LaunchParameters(float lambda_i, float mu_f, float La, uint32_t maxIntIter, float tolVar, int verbose, int maxIter,
                         uint32_t stoptest, uint32_t miniter, ArgumentsMotionCompensation* args, bool sp = false)
            : lambda_i(lambda_i), mu_f(mu_f), La(La), maxIntIter(maxIntIter), tolVar(tolVar), verbose(verbose),
              maxIter(maxIter), stoptest(stoptest), miniter(miniter), argsMC(args), showProgress(sp) {}
    };
    /// Execution form: the fused iteration (default), the reference's op sequence on the GPU, or the fused iteration
    /// with split-precision BF16-MFMA DFTs (H, W in {64, 128, 192, 256}).
    enum Form { FUSED = 1, PLAIN = 0, SPLIT = 2 };
    void setForm(Form f) { form_ = f; }

    /// What the last launch computed: normU, lambda, mu0, iterations per stage, f per inner iteration.
    struct Report {
        double normU, maxXref, maxUXref, lambda, mu0;
        std::vector<int> itersPerStage;
        std::vector<double> f;
    };
    const Report& getReport() const { return report_; }

    void init() override;
    void launch() override;
    OCL_PROCESS_BOILERPLATE(NestaUp)

  private:
    Form form_ = FUSED;
    Report report_{};
};

}  // namespace OpenCLIPER
