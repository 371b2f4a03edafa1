// The Processes of the NESTA-TV path in the OpenCLIPER-compatible C++ API, each one call into the C ABI.
// Reference Processes mirrored: TemporalTV, VectorNormalization, FFT, ComplexElementProd, ApplyMask, XImageSum,
// ComplexAbs, ComplexAbsPow2, CopyDataGPU, RSoS, SimpleMRIRecon, SimpleMRIReconSOS, NestaUp (src/processes/**).
#include <OpenCLIPER/processes/ApplyMask.hpp>
#include <OpenCLIPER/processes/ComplexAbs.hpp>
#include <OpenCLIPER/processes/ComplexAbsPow2.hpp>
#include <OpenCLIPER/processes/ComplexElementProd.hpp>
#include <OpenCLIPER/processes/FFT.hpp>
#include <OpenCLIPER/processes/XImageSum.hpp>
#include <OpenCLIPER/processes/examples/SimpleMRIRecon.hpp>
#include <OpenCLIPER/processes/examples/SimpleMRIReconSOS.hpp>
#include <OpenCLIPER/processes/RSoS.hpp>
#include <OpenCLIPER/processes/groupwiseRegistration/initialize/CopyDataGPU.hpp>
#include <OpenCLIPER/processes/nesta/NestaUp.hpp>
#include <OpenCLIPER/processes/nesta/TemporalTV.hpp>
#include <OpenCLIPER/processes/nesta/VectorNormalization.hpp>

#include "ocl_amd.h"

namespace OpenCLIPER {

namespace {
ocl_ctx* ctxOf(const Process& p) { return p.getApp()->getOclContext(); }
long pixels(const Data& d) { return (long)d.getNDArray(0)->size(); }
void needIO(const std::shared_ptr<Data>& in, const std::shared_ptr<Data>& out, const char* who) {
    if (!in) throw std::invalid_argument(std::string(who) + ": launch() called before setInput()");
    if (!out) throw std::invalid_argument(std::string(who) + ": launch() called before setOutput()");
}
}  // namespace

void Process::startProfiling() {
    if (pProfileParameters && pProfileParameters->enable)
        checkOcl(ocl_timer_begin(app_->getOclContext(), (ocl_timer**)&ev_[0]), "Process::startProfiling");
}

void Process::stopProfiling() {
    if (pProfileParameters && pProfileParameters->enable && ev_[0]) {
        checkOcl(ocl_timer_end(app_->getOclContext(), (ocl_timer*)ev_[0], &lastMs_), "Process::stopProfiling");
        ev_[0] = nullptr;
    }
}

// ---- TemporalTV (src/processes/nesta/TemporalTV.cpp:49-86) ----
void TemporalTV::init() {
    if (!std::dynamic_pointer_cast<InitParameters>(pInitParameters))
        throw std::invalid_argument("TemporalTV::init: InitParameters with a direction are required");
}
void TemporalTV::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "TemporalTV");
    auto ip = std::dynamic_pointer_cast<InitParameters>(pInitParameters);
    if (!ip) throw std::invalid_argument("TemporalTV::launch: init() not called");
    startProfiling();
    checkOcl(ocl_ttv(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(), (int)in->getNumNDArrays(),
                     pixels(*in), ip->dir == FORWARD ? 1 : 0),
             "TemporalTV::launch");
    stopProfiling();
}

// ---- VectorNormalization (src/processes/nesta/VectorNormalization.cpp:50-87) ----
void VectorNormalization::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "VectorNormalization");
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp) throw std::invalid_argument("VectorNormalization::launch: LaunchParameters(mu) required");
    startProfiling();
    checkOcl(ocl_vecnorm(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(),
                         (long)in->getNumNDArrays() * pixels(*in), lp->mu),
             "VectorNormalization::launch");
    stopProfiling();
}

// ---- FFT (src/processes/FFT.cpp:95-146,497-525, dim = -1) ----
void FFT::init() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "FFT::init");
    auto ip = std::dynamic_pointer_cast<InitParameters>(pInitParameters);
    if (ip && (ip->dim != -1 || ip->samplingMask))
        throw std::invalid_argument("FFT::init: only full 2-D transforms (dim = -1, no mask) are provided");
    if (in->getData()->size() != out->getData()->size())
        throw std::invalid_argument("OpenCLIPER::FFT::launch(): inputData and outputData must have the same number of images");
    if (!in->getAllSizesEqual())
        throw std::invalid_argument("FFT for variable-size data objects is not implemented at this time");
}
void FFT::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "FFT");
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    const int dir = (lp && lp->dir == BACKWARD) ? 1 : -1;
    const NDArray* a = in->getNDArray(0);
    startProfiling();
    checkOcl(ocl_fft2d(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(), (long)in->getNumNDArrays(),
                       (int)NDARRAYHEIGHT(a), (int)NDARRAYWIDTH(a), dir),
             "FFT::launch");
    stopProfiling();
}

// ---- ComplexElementProd (src/processes/ComplexElementProd.cpp:60-131) ----
void ComplexElementProd::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "ComplexElementProd");
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp || !lp->sensitivityMapsData)
        throw std::invalid_argument("ComplexElementProd::launch: non-existing SensitivityMaps");
    auto kout = std::dynamic_pointer_cast<KData>(out);
    if (!kout) throw std::invalid_argument("ComplexElementProd::launch: outputData should be of type KData");
    const int C = (int)lp->sensitivityMapsData->getNCoils();
    const int F = (int)kout->getDynDimsTotalSize();
    const int in_coils = std::dynamic_pointer_cast<KData>(in) ? 1 : 0;
    startProfiling();
    checkOcl(ocl_cep(ctxOf(*this), in->getDeviceBuffer(), lp->sensitivityMapsData->getDeviceBuffer(),
                     out->getDeviceBuffer(), F, C, pixels(*in), in_coils,
                     lp->conjugateSensMap == conjugate ? 1 : 0),
             "ComplexElementProd::launch");
    stopProfiling();
}

// ---- ApplyMask (src/processes/ApplyMask.cpp:21-51) ----
void ApplyMask::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "ApplyMask");
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp || !lp->samplingMasksData) throw std::invalid_argument("ApplyMask::launch: null sampling masks");
    auto k = std::dynamic_pointer_cast<KData>(in);
    if (!k) throw std::invalid_argument("ApplyMask::launch: input must be KData");
    if (in != out) checkOcl(ocl_copy(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(),
                                     (long)in->getNumNDArrays() * pixels(*in)), "ApplyMask copy");
    startProfiling();
    checkOcl(ocl_apply_mask(ctxOf(*this), out->getDeviceBuffer(),
                            (const uint8_t*)lp->samplingMasksData->getDeviceBuffer(), (int)k->getDynDimsTotalSize(),
                            (int)k->getNCoils(), pixels(*in)),
             "ApplyMask::launch");
    stopProfiling();
}

// ---- XImageSum (src/processes/XImageSum.cpp:66-96) ----
void XImageSum::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "XImageSum");
    auto k = std::dynamic_pointer_cast<KData>(in);
    if (!k) throw std::invalid_argument("XImageSum::launch: input must be KData");
    startProfiling();
    checkOcl(ocl_ximage_sum(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(), (int)k->getDynDimsTotalSize(),
                            (int)k->getNCoils(), pixels(*in)),
             "XImageSum::launch");
    stopProfiling();
}

// ---- RSoS (src/processes/RSoS.cpp, src/kernels/rss.cl) ----
void RSoS::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "RSoS");
    auto k = std::dynamic_pointer_cast<KData>(in);
    if (!k) throw std::invalid_argument("RSoS::launch: input must be KData");
    startProfiling();
    checkOcl(ocl_rsos(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(), (int)k->getDynDimsTotalSize(),
                      (int)k->getNCoils(), pixels(*in)),
             "RSoS::launch");
    stopProfiling();
}

// ---- ComplexAbs / ComplexAbsPow2 (src/processes/ComplexAbs.cpp:48-78, ComplexAbsPow2.cpp:48-78) ----
void ComplexAbs::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "ComplexAbs");
    checkOcl(ocl_complex_abs(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(),
                             (long)in->getNumNDArrays() * pixels(*in), 0),
             "ComplexAbs::launch");
}
void ComplexAbsPow2::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "ComplexAbsPow2");
    checkOcl(ocl_complex_abs(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(),
                             (long)in->getNumNDArrays() * pixels(*in), 1),
             "ComplexAbsPow2::launch");
}

// ---- CopyDataGPU (src/processes/groupwiseRegistration/initialize/CopyDataGPU.cpp:53-74) ----
void CopyDataGPU::launch() {
    auto in = getInput(), out = getOutput();
    needIO(in, out, "CopyDataGPU");
    checkOcl(ocl_copy(ctxOf(*this), in->getDeviceBuffer(), out->getDeviceBuffer(),
                      (long)in->getNumNDArrays() * pixels(*in)),
             "CopyDataGPU::launch");
}

// ---- SimpleMRIRecon (src/processes/examples/SimpleMRIRecon.cpp:59-101) ----
void SimpleMRIRecon::init() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    if (!k || !getOutput()) throw std::invalid_argument("SimpleMRIRecon::init: KData input and XData output required");
    if (!k->getSensitivityMapsData()) throw std::invalid_argument("SimpleMRIRecon::init: non-existing SensitivityMaps");
}
void SimpleMRIRecon::launch() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    auto out = getOutput();
    needIO(k, out, "SimpleMRIRecon");
    const NDArray* a = k->getNDArray(0);
    startProfiling();
    checkOcl(ocl_simple_mri_recon(ctxOf(*this), k->getDeviceBuffer(), k->getSensitivityMapsData()->getDeviceBuffer(),
                                  out->getDeviceBuffer(), (int)k->getDynDimsTotalSize(), (int)k->getNCoils(),
                                  (int)NDARRAYHEIGHT(a), (int)NDARRAYWIDTH(a)),
             "SimpleMRIRecon::launch");
    stopProfiling();
}

// ---- SimpleMRIReconSOS (src/processes/examples/SimpleMRIReconSOS.cpp) ----
void SimpleMRIReconSOS::init() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    if (!k || !getOutput()) throw std::invalid_argument("SimpleMRIReconSOS::init: KData input and XData output required");
}
void SimpleMRIReconSOS::launch() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    auto out = getOutput();
    needIO(k, out, "SimpleMRIReconSOS");
    const NDArray* a = k->getNDArray(0);
    startProfiling();
    checkOcl(ocl_simple_mri_recon_sos(ctxOf(*this), k->getDeviceBuffer(), out->getDeviceBuffer(),
                                      (int)k->getDynDimsTotalSize(), (int)k->getNCoils(), (int)NDARRAYHEIGHT(a),
                                      (int)NDARRAYWIDTH(a)),
             "SimpleMRIReconSOS::launch");
    stopProfiling();
}

// ---- NestaUp (src/processes/nesta/NestaUp.cpp:81-106,442-910) ----
void NestaUp::init() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    if (!k) throw std::invalid_argument("NestaUp::init: input must be KData");
    if (!std::dynamic_pointer_cast<XData>(getOutput())) throw std::invalid_argument("NestaUp::init: output must be XData");
}
void NestaUp::launch() {
    auto k = std::dynamic_pointer_cast<KData>(getInput());
    auto x = std::dynamic_pointer_cast<XData>(getOutput());
    if (!k || !x) throw std::invalid_argument("NestaUp::launch: KData input and XData output required");
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp) throw std::invalid_argument("NestaUp::launch: LaunchParameters required");
    if (lp->argsMC) throw std::invalid_argument("NestaUp::launch: the motion-compensated operator is not provided");
    auto S = k->getSensitivityMapsData();
    auto M = k->getSamplingMasksData();
    if (!S) throw std::invalid_argument("NestaUp::launch: non-existing SensitivityMaps");
    if (!M) throw std::invalid_argument("NestaUp::launch: non-existing SamplingMasks");
    const NDArray* a = k->getNDArray(0);
    ocl_nesta_params p{lp->lambda_i, lp->mu_f, lp->La, (int)lp->maxIntIter, lp->tolVar, lp->verbose,
                       (int)lp->maxIter, (int)lp->stoptest, (int)lp->miniter};
    auto rep = std::unique_ptr<ocl_nesta_report>(new ocl_nesta_report());
    startProfiling();
    checkOcl(ocl_nesta_run(ctxOf(*this), k->getDeviceBuffer(), S->getDeviceBuffer(),
                           (const uint8_t*)M->getDeviceBuffer(), x->getDeviceBuffer(), (int)k->getDynDimsTotalSize(),
                           (int)k->getNCoils(), (int)NDARRAYHEIGHT(a), (int)NDARRAYWIDTH(a), &p,
                           form_ == FUSED ? OCL_FORM_FUSED : form_ == SPLIT ? OCL_FORM_SPLIT : OCL_FORM_PLAIN, rep.get()),
             "NestaUp::launch");
    stopProfiling();
    lp->tolVar = (float)rep->final_tolVar;   // the reference overwrites the caller's tolVar
    report_.normU = rep->normU;
    report_.maxXref = rep->maxXref;
    report_.maxUXref = rep->maxUXref;
    report_.lambda = rep->lambda;
    report_.mu0 = rep->mu0;
    report_.itersPerStage.assign(rep->iters, rep->iters + rep->stages);
    report_.f.assign(rep->trace_f, rep->trace_f + rep->ntrace);
}

}  // namespace OpenCLIPER
