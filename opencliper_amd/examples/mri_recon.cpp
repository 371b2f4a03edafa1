// NESTA-TV driver written against the OpenCLIPER-compatible API, shaped like the reference's tests/MRIRecon.cpp:
// create a CLapp, build KData (with maps and masks) and an XData from it, set up NestaUp with the driver's launch
// parameters and launch it.  Input and output are raw little-endian files so the Python tests can drive it:
//   mri_recon <in.raw> <out.raw> F C H W maxIntIter maxIter stoptest miniter form [profile]
// in.raw holds y [F][C][H][W] complex64, S [C][H][W] complex64, M [F][H][W] uint8; out.raw receives x [F][H][W].
// With files in the reference's formats (KData with SensitivityMaps and SamplingMasks in, from a MAT v5 file or a CFL
// set; XData out as a MAT v5 file), as the reference's KData / XData file constructors and matlabSave lay them out:
//   mri_recon <in.mat|in.cfl> <out.mat> maxIntIter maxIter stoptest miniter form [profile]
// form: 0 plain, 1 fused, 2 split precision.
#include <OpenCLIPER/processes/nesta/NestaUp.hpp>
#include <OpenCLIPER/processes/nesta/TemporalTV.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

using namespace OpenCLIPER;

// k-space, maps and masks from the raw input layout
static std::shared_ptr<KData> readRaw(const std::shared_ptr<CLapp>& pCLapp, char* argv[]) {
    const int F = atoi(argv[3]), C = atoi(argv[4]), H = atoi(argv[5]), W = atoi(argv[6]);
    std::ifstream in(argv[1], std::ios::binary);
    if (!in) throw std::runtime_error("cannot open input");
    auto readC = [&](size_t n) {
        auto* v = new std::vector<complexType>(n);
        in.read((char*)v->data(), sizeof(complexType) * n);
        return v;
    };
    // k-space NDArrays: frame-major, coil fastest
    auto* karr = new std::vector<NDArray*>();
    for (int i = 0; i < F * C; ++i) {
        auto* dims = new std::vector<dimIndexType>({(dimIndexType)W, (dimIndexType)H});
        auto* data = readC((size_t)H * W);
        karr->push_back(new ConcreteNDArray<complexType>(dims, data));
    }
    auto* sarr = new std::vector<NDArray*>();
    for (int c = 0; c < C; ++c) {
        auto* dims = new std::vector<dimIndexType>({(dimIndexType)W, (dimIndexType)H});
        auto* data = readC((size_t)H * W);
        sarr->push_back(new ConcreteNDArray<complexType>(dims, data));
    }
    auto* marr = new std::vector<NDArray*>();
    for (int f = 0; f < F; ++f) {
        auto* dims = new std::vector<dimIndexType>({(dimIndexType)W, (dimIndexType)H});
        auto* data = new std::vector<uint8_t>((size_t)H * W);
        in.read((char*)data->data(), (std::streamsize)data->size());
        marr->push_back(new ConcreteNDArray<uint8_t>(dims, data));
    }
    if (!in) throw std::runtime_error("input file too short");
    auto* dyn = new std::vector<dimIndexType>({(dimIndexType)F});
    auto pKData = std::make_shared<KData>(pCLapp, karr, (numCoilsType)C, dyn);
    SensitivityMapsData* pS = new SensitivityMapsData(pCLapp, sarr, (numCoilsType)C);
    auto* dynm = new std::vector<dimIndexType>({(dimIndexType)F});
    SamplingMasksData* pM = new SamplingMasksData(pCLapp, marr, dynm, (dimIndexType)W);
    pKData->setSensitivityMapsData(pS);
    pKData->setSamplingMasksData(pM);
    return pKData;
}

static bool isDataFile(const char* name) {
    const size_t n = std::strlen(name);
    return n > 4 && (!std::strcmp(name + n - 4, ".mat") || !std::strcmp(name + n - 4, ".cfl"));
}

int main(int argc, char* argv[]) {
    const bool mat = argc > 1 && isDataFile(argv[1]);
    if (argc < (mat ? 8 : 12)) {
        std::cerr << "usage: mri_recon in.raw out.raw F C H W maxIntIter maxIter stoptest miniter form [profile]\n"
                     "       mri_recon in.mat|in.cfl out.mat maxIntIter maxIter stoptest miniter form [profile]\n";
        return 2;
    }
    const int a0 = mat ? 3 : 7;   // first launch-parameter argument
    try {
        auto pCLapp = CLapp::create();
        std::shared_ptr<KData> pKData;
        if (mat) {
            pKData = std::make_shared<KData>(pCLapp, std::string(argv[1]));
            if (!pKData->getSensitivityMapsData() || !pKData->getSamplingMasksData())
                throw std::invalid_argument("the input file needs SensitivityMaps and SamplingMasks");
        } else {
            pKData = readRaw(pCLapp, argv);
        }
        const int H = (int)NDARRAYHEIGHT(pKData->getNDArray(0)), W = (int)NDARRAYWIDTH(pKData->getNDArray(0));
        auto pOutputXData = std::make_shared<XData>(pCLapp, pKData);

        // as in tests/MRIRecon.cpp:94,195-197,232-235 of the reference
        auto pp = std::make_shared<ProcessCore::ProfileParameters>(argc > a0 + 5, 1);
        auto nesta = Process::create<NestaUp>(pCLapp, pp);
        const int form = atoi(argv[a0 + 4]);
        if (form < 0 || form > 2) throw std::invalid_argument("form must be 0 (plain), 1 (fused) or 2 (split)");
        nesta->setForm((NestaUp::Form)form);
        nesta->setInput(pKData);
        nesta->setOutput(pOutputXData);
        nesta->init();
        const float lambda_i = 1e-2f, mu_f = 3e-4f, La = 1.f, tolVar = 3e-4f;
        auto params = std::make_shared<NestaUp::LaunchParameters>(lambda_i, mu_f, La, (uint32_t)atoi(argv[a0]), tolVar, 1,
                                                                  atoi(argv[a0 + 1]), (uint32_t)atoi(argv[a0 + 2]),
                                                                  (uint32_t)atoi(argv[a0 + 3]), nullptr);
        nesta->setLaunchParameters(params);
        nesta->launch();
        pOutputXData->device2Host();
        if (mat) {
            pOutputXData->matlabSave(argv[2]);
        } else {
            std::ofstream out(argv[2], std::ios::binary);
            for (dimIndexType f = 0; f < pOutputXData->getNumNDArrays(); ++f)
                out.write((const char*)pOutputXData->getHostBuffer(f), sizeof(complexType) * (size_t)H * W);
        }
        const auto& r = nesta->getReport();
        std::cout << "normU " << r.normU << " lambda " << r.lambda << " iters";
        for (int n : r.itersPerStage) std::cout << " " << n;
        std::cout << " ms " << nesta->getLastLaunchMs() << "\n";

        // TemporalTV round trip through the same API: ADJOINT(FORWARD(x)) on the reconstruction
        auto u = std::make_shared<XData>(pCLapp, pOutputXData, false);
        auto v = std::make_shared<XData>(pCLapp, pOutputXData, false);
        auto tv = Process::create<TemporalTV>(pCLapp, pOutputXData, u);
        tv->setInitParameters(std::make_shared<TemporalTV::InitParameters>(TemporalTV::FORWARD));
        tv->init();
        tv->launch();
        auto tva = Process::create<TemporalTV>(pCLapp, u, v);
        tva->setInitParameters(std::make_shared<TemporalTV::InitParameters>(TemporalTV::ADJOINT));
        tva->init();
        tva->launch();
        v->device2Host();
        std::ofstream out2(std::string(argv[2]) + ".ttv", std::ios::binary);
        for (dimIndexType f = 0; f < v->getNumNDArrays(); ++f)
            out2.write((const char*)v->getHostBuffer(f), sizeof(complexType) * (size_t)H * W);
    } catch (CLError& e) {
        std::cerr << "CLError " << e.err() << ": " << e.what() << "\n";
        return 1;
    } catch (std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
