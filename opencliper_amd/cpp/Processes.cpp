// The Processes of the NESTA-TV path in the OpenCLIPER-compatible C++ API, each one call into the C ABI.
// Reference Processes mirrored: TemporalTV, VectorNormalization, FFT, ComplexElementProd, ApplyMask, XImageSum,
// ComplexAbs, ComplexAbsPow2, CopyDataGPU, RSoS, SimpleMRIRecon, SimpleMRIReconSOS, MotionCompensation,
// AdjointMotionCompensation, GroupwiseRegistration, NestaUp (src/processes/**).
#include <OpenCLIPER/processes/AdjointMotionCompensation.hpp>
#include <OpenCLIPER/processes/ApplyMask.hpp>
#include <OpenCLIPER/processes/MotionCompensation.hpp>
#include <OpenCLIPER/processes/ComplexAbs.hpp>
#include <OpenCLIPER/processes/ComplexAbsPow2.hpp>
#include <OpenCLIPER/processes/ComplexElementProd.hpp>
#include <OpenCLIPER/processes/FFT.hpp>
#include <OpenCLIPER/processes/GroupwiseRegistration.hpp>
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

// ---- MotionCompensation / AdjointMotionCompensation (src/processes/MotionCompensation.cpp,
//      AdjointMotionCompensation.cpp) ----
namespace {
// Device view of ArgumentsMotionCompensation for F frames of H x W, with the sizes of its Data checked
ocl_mc_args mcArgs(const ArgumentsMotionCompensation* a, int F, int H, int W, const char* who) {
    if (!a) throw std::invalid_argument(std::string(who) + ": null argsMC");
    if (!a->xnData || !a->xData || !a->bound_box || a->bound_box->size() != 4)
        throw std::invalid_argument(std::string(who) + ": argsMC needs xnData, xData and a 4-value bound_box");
    const size_t N = (size_t)H * W;
    if (a->xnData->getElementDataType() != TYPEID_REAL || a->xnData->getDeviceBufferBytes() < 2 * (size_t)F * N * 4)
        throw std::invalid_argument(std::string(who) + ": xnData must hold 2*F*H*W floats");
    const ElementDataType xt = a->xData->getElementDataType();
    if ((xt != TYPEID_UINT32 && xt != TYPEID_INT32) || a->xData->getDeviceBufferBytes() < 2 * N * 4)
        throw std::invalid_argument(std::string(who) + ": xData must hold 2*H*W 32-bit integers");
    ocl_mc_args m;
    m.xn = (const float*)a->xnData->getDeviceBuffer();
    m.x = (const int*)a->xData->getDeviceBuffer();
    for (int i = 0; i < 4; ++i) m.bound_box[i] = a->bound_box->at(i);
    return m;
}

void launchMC(Process& p, const std::shared_ptr<Data>& in, const std::shared_ptr<Data>& out,
              ArgumentsMotionCompensation* args, bool adjoint, const char* who) {
    needIO(in, out, who);
    const NDArray* a = in->getNDArray(0);
    const int F = (int)in->getNumNDArrays(), H = (int)NDARRAYHEIGHT(a), W = (int)NDARRAYWIDTH(a);
    const ocl_mc_args m = mcArgs(args, F, H, W, who);
    checkOcl(adjoint ? ocl_adjoint_motion_compensation(ctxOf(p), in->getDeviceBuffer(), out->getDeviceBuffer(), F, H,
                                                       W, &m)
                     : ocl_motion_compensation(ctxOf(p), in->getDeviceBuffer(), out->getDeviceBuffer(), F, H, W, &m),
             who);
}
}  // namespace

void MotionCompensation::launch() {
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp) throw std::invalid_argument("MotionCompensation::launch: LaunchParameters required");
    startProfiling();
    launchMC(*this, getInput(), getOutput(), lp->argsMC, false, "MotionCompensation::launch");
    stopProfiling();
}

void AdjointMotionCompensation::launch() {
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp) throw std::invalid_argument("AdjointMotionCompensation::launch: LaunchParameters required");
    startProfiling();
    launchMC(*this, getInput(), getOutput(), lp->argsMC, true, "AdjointMotionCompensation::launch");
    stopProfiling();
}

// ---- GroupwiseRegistration (src/processes/GroupwiseRegistration.cpp:93-615, groupwiseRegistration/*.cpp) ----
GroupwiseRegistration::~GroupwiseRegistration() {
    if (plan_) ocl_gwr_destroy(plan_);
    delete bound_box_;
}

void GroupwiseRegistration::init() {
    auto ip = std::dynamic_pointer_cast<InitParameters>(pInitParameters);
    if (!ip) throw std::invalid_argument("GroupwiseRegistration::init: InitParameters required");
    auto in = getInput();
    if (!in) throw CLError(OCL_E_INVALID, "GroupwiseRegistration::init: init() called before setInput()");
    if (in->getElementDataType() != TYPEID_COMPLEX)
        throw std::invalid_argument("GroupwiseRegistration::init: the input must be complex image frames");
    if (!ip->Dp) throw std::invalid_argument("GroupwiseRegistration::init: null Dp");
    if (ip->lambda.size() < 4) throw std::invalid_argument("GroupwiseRegistration::init: lambda needs four weights");
    const NDArray* a = in->getNDArray(0);
    const int F = (int)in->getNumNDArrays(), H = (int)NDARRAYHEIGHT(a), W = (int)NDARRAYWIDTH(a);
    ocl_gwr_params prm;
    prm.W = ip->W;
    prm.flagW = ip->flagW ? 1 : 0;
    prm.radius = ip->radius;
    prm.E = ip->E;
    prm.Dp[0] = ip->Dp[0];
    prm.Dp[1] = ip->Dp[1];
    prm.nmax = ip->nmax;
    prm.et = ip->et;
    prm.eh = ip->eh;
    for (int k = 0; k < 4; ++k) prm.lambda[k] = ip->lambda[k];
    if (plan_) ocl_gwr_destroy(plan_);
    plan_ = nullptr;
    checkOcl(ocl_gwr_create(ctxOf(*this), F, H, W, &prm, &plan_), "GroupwiseRegistration::init");
    ocl_gwr_info_t info;
    checkOcl(ocl_gwr_info(plan_, &info), "GroupwiseRegistration::init");
    F_ = F, N_ = H, Cp_ = info.Cp;
    // the mesh x (GroupwiseRegistration.cpp:150-164): int32 {H, W, 2}, element [d][c][r]
    auto* dims = new std::vector<dimIndexType>({(dimIndexType)H, (dimIndexType)W, 2});
    xData_ = std::make_shared<XData>(getApp(), dims, TYPEID_INT32);
    checkOcl(ocl_memcpy_d2d(ctxOf(*this), xData_->getDeviceBuffer(), info.mesh, sizeof(int32_t) * 2 * (size_t)H * W),
             "GroupwiseRegistration::init");
    checkOcl(ocl_ctx_sync(ctxOf(*this)), "GroupwiseRegistration::init");
    delete bound_box_;
    bound_box_ = new std::vector<dimIndexType>(info.bound_box, info.bound_box + 4);
}

void GroupwiseRegistration::launch() {
    auto lp = std::dynamic_pointer_cast<LaunchParameters>(pLaunchParameters);
    if (!lp || !lp->argsMC) throw std::invalid_argument("GroupwiseRegistration::launch: LaunchParameters with argsMC "
                                                        "required");
    if (!plan_) throw std::invalid_argument("GroupwiseRegistration::launch: init() not called");
    auto in = getInput();
    if (!in || (int)in->getNumNDArrays() != F_ || (int)NDARRAYHEIGHT(in->getNDArray(0)) != N_)
        throw std::invalid_argument("GroupwiseRegistration::launch: the input changed shape since init()");
    const size_t nT = 2 * (size_t)F_ * Cp_ * Cp_;
    // coefficients: the caller's TData (the previous motion iteration's estimate, optimizer.cpp:73-82) or zero
    std::shared_ptr<Data> T = lp->TData;
    if (T && (T->getElementDataType() != TYPEID_REAL || T->getDeviceBufferBytes() < nT * sizeof(float)))
        throw std::invalid_argument("GroupwiseRegistration::launch: TData must hold 2*F*Cp*Cp floats");
    const bool T_in = (bool)T;
    if (!T) {
        auto* dims = new std::vector<dimIndexType>({(dimIndexType)Cp_, (dimIndexType)Cp_, (dimIndexType)F_, 2});
        T = std::make_shared<XData>(getApp(), dims, TYPEID_REAL);
    }
    auto* xdims = new std::vector<dimIndexType>({(dimIndexType)N_, (dimIndexType)N_, (dimIndexType)F_, 2});
    auto xn = std::make_shared<XData>(getApp(), xdims, TYPEID_REAL);
    ocl_gwr_report rep;
    startProfiling();
    checkOcl(ocl_gwr_run(plan_, in->getDeviceBuffer(), (float*)T->getDeviceBuffer(), T_in ? 1 : 0,
                         (float*)xn->getDeviceBuffer(), nullptr, &rep),
             "GroupwiseRegistration::launch");
    stopProfiling();
    report_.iters = rep.iters;
    report_.oldmax = rep.oldmax;
    report_.H.assign(rep.H, rep.H + rep.ntrace + 1);
    report_.Tdif.assign(rep.Tdif, rep.Tdif + rep.ntrace);
    report_.Hdif.assign(rep.Hdif, rep.Hdif + rep.ntrace);
    report_.Wn.assign(rep.Wn_trace, rep.Wn_trace + rep.ntrace);
    xnData_ = xn;
    TData_ = T;
    // optimizer.cpp:362-365
    lp->argsMC->xnData = xnData_;
    lp->argsMC->TData = TData_;
    lp->argsMC->xData = xData_;
    lp->argsMC->bound_box = bound_box_;
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
    auto S = k->getSensitivityMapsData();
    auto M = k->getSamplingMasksData();
    if (!S) throw std::invalid_argument("NestaUp::launch: non-existing SensitivityMaps");
    if (!M) throw std::invalid_argument("NestaUp::launch: non-existing SamplingMasks");
    const NDArray* a = k->getNDArray(0);
    ocl_nesta_params p{lp->lambda_i, lp->mu_f, lp->La, (int)lp->maxIntIter, lp->tolVar, lp->verbose,
                       (int)lp->maxIter, (int)lp->stoptest, (int)lp->miniter};
    auto rep = std::unique_ptr<ocl_nesta_report>(new ocl_nesta_report());
    const int F = (int)k->getDynDimsTotalSize(), H = (int)NDARRAYHEIGHT(a), W = (int)NDARRAYWIDTH(a);
    startProfiling();
    if (lp->argsMC) {
        // argsMC != nullptr: tTV o MC, starting from the output's contents (NestaUp.cpp:201-278,483-491)
        const ocl_mc_args m = mcArgs(lp->argsMC, F, H, W, "NestaUp::launch");
        // the fused engine needs GroupwiseRegistration's identity mesh; other meshes and PLAIN take the plain form
        int form = form_ == PLAIN ? OCL_FORM_PLAIN : OCL_FORM_FUSED;
        int rc = ocl_nesta_run_mc(ctxOf(*this), k->getDeviceBuffer(), S->getDeviceBuffer(),
                                  (const uint8_t*)M->getDeviceBuffer(), x->getDeviceBuffer(), F, (int)k->getNCoils(), H,
                                  W, &p, form, &m, rep.get());
        if (rc == OCL_E_UNSUPPORTED && form == OCL_FORM_FUSED)
            rc = ocl_nesta_run_mc(ctxOf(*this), k->getDeviceBuffer(), S->getDeviceBuffer(),
                                  (const uint8_t*)M->getDeviceBuffer(), x->getDeviceBuffer(), F, (int)k->getNCoils(), H,
                                  W, &p, OCL_FORM_PLAIN, &m, rep.get());
        checkOcl(rc, "NestaUp::launch");
    } else {
        // sizes the fused or split form does not take (H > 512, W > 1024, a tile beyond LDS, split sizes other than
        // 64..256) run in the plain form, which takes every size the reference's FFT does
        const int form = form_ == FUSED ? OCL_FORM_FUSED : form_ == SPLIT ? OCL_FORM_SPLIT : OCL_FORM_PLAIN;
        int rc = ocl_nesta_run(ctxOf(*this), k->getDeviceBuffer(), S->getDeviceBuffer(),
                               (const uint8_t*)M->getDeviceBuffer(), x->getDeviceBuffer(), F, (int)k->getNCoils(), H, W,
                               &p, form, rep.get());
        if (rc == OCL_E_UNSUPPORTED && form != OCL_FORM_PLAIN)
            rc = ocl_nesta_run(ctxOf(*this), k->getDeviceBuffer(), S->getDeviceBuffer(),
                               (const uint8_t*)M->getDeviceBuffer(), x->getDeviceBuffer(), F, (int)k->getNCoils(), H, W,
                               &p, OCL_FORM_PLAIN, rep.get());
        checkOcl(rc, "NestaUp::launch");
    }
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
