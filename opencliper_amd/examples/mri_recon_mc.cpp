// Motion-compensated NESTA-TV through the OpenCLIPER-compatible API, chained as the reference's tests/MRIRecon.cpp
// does it (:209-282): a first NESTA without motion compensation, then NESTA with argsMC starting from that result.
//   mri_recon_mc <in.raw> <motion.raw> <out.raw> F C H W maxIntIter maxIter stoptest miniter
//   mri_recon_mc <in.raw> register <out.raw> F C H W maxIntIter maxIter stoptest miniter [motionIters [radius [W]]]
// in.raw: y [F][C][H][W] complex64, S [C][H][W] complex64, M [F][H][W] uint8 (as mri_recon reads it).
// motion.raw: xn [2][F][H][W] float32, mesh [2][W][H] int32, bound box 4 x uint32 (1-based).
// With "register" the deformation comes from GroupwiseRegistration of the latest reconstruction, motionIters times
// (default 1), with tests/MRIRecon.cpp:122-130's parameters (radius 40 and W 1 unless given); the coefficients of
// one motion iteration start the next.  It writes out.raw.xn<i> (xn [2][F][H][W] float32 of motion iteration i) and
// out.raw.rec<i> (the reconstruction of motion iteration i), and prints one line per registration:
// "registration <i> iters <n> H <first> <last>".
// Writes out.raw (the motion-compensated reconstruction), out.raw.nomc (the first reconstruction), and, with a motion
// file, out.raw.mc / out.raw.mcadj (MotionCompensation / AdjointMotionCompensation of the first reconstruction).
#include <OpenCLIPER/processes/AdjointMotionCompensation.hpp>
#include <OpenCLIPER/processes/GroupwiseRegistration.hpp>
#include <OpenCLIPER/processes/MotionCompensation.hpp>
#include <OpenCLIPER/processes/nesta/NestaUp.hpp>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>

using namespace OpenCLIPER;

template <class T> static std::vector<T>* readVec(std::ifstream& in, size_t n) {
    auto* v = new std::vector<T>(n);
    in.read((char*)v->data(), (std::streamsize)(sizeof(T) * n));
    return v;
}

static void writeX(const std::shared_ptr<XData>& x, const std::string& name) {
    x->device2Host();
    std::ofstream out(name, std::ios::binary);
    for (dimIndexType f = 0; f < x->getNumNDArrays(); ++f)
        out.write((const char*)x->getHostBuffer(f), (std::streamsize)(sizeof(complexType) * x->getNDArray(f)->size()));
}

int main(int argc, char* argv[]) {
    if (argc < 12) {
        std::cerr << "usage: mri_recon_mc in.raw motion.raw out.raw F C H W maxIntIter maxIter stoptest miniter\n";
        return 2;
    }
    const int F = atoi(argv[4]), C = atoi(argv[5]), H = atoi(argv[6]), W = atoi(argv[7]);
    const dimIndexType uW = (dimIndexType)W, uH = (dimIndexType)H;
    try {
        auto pCLapp = CLapp::create();
        std::ifstream in(argv[1], std::ios::binary);
        if (!in) throw std::runtime_error("cannot open input");
        auto* karr = new std::vector<NDArray*>();
        for (int i = 0; i < F * C; ++i) {
            auto* dims = new std::vector<dimIndexType>({uW, uH});
            auto* data = readVec<complexType>(in, (size_t)H * W);
            karr->push_back(new ConcreteNDArray<complexType>(dims, data));
        }
        auto* sarr = new std::vector<NDArray*>();
        for (int c = 0; c < C; ++c) {
            auto* dims = new std::vector<dimIndexType>({uW, uH});
            auto* data = readVec<complexType>(in, (size_t)H * W);
            sarr->push_back(new ConcreteNDArray<complexType>(dims, data));
        }
        auto* marr = new std::vector<NDArray*>();
        for (int f = 0; f < F; ++f) {
            auto* dims = new std::vector<dimIndexType>({uW, uH});
            auto* data = readVec<uint8_t>(in, (size_t)H * W);
            marr->push_back(new ConcreteNDArray<uint8_t>(dims, data));
        }
        if (!in) throw std::runtime_error("input file too short");
        auto* dyn = new std::vector<dimIndexType>({(dimIndexType)F});
        auto pKData = std::make_shared<KData>(pCLapp, karr, (numCoilsType)C, dyn);
        SensitivityMapsData* pS = new SensitivityMapsData(pCLapp, sarr, (numCoilsType)C);
        auto* dynm = new std::vector<dimIndexType>({(dimIndexType)F});
        SamplingMasksData* pM = new SamplingMasksData(pCLapp, marr, dynm, uW);
        pKData->setSensitivityMapsData(pS);
        pKData->setSamplingMasksData(pM);
        auto pOutputXData = std::make_shared<XData>(pCLapp, pKData);

        const float lambda_i = 1e-2f, mu_f = 3e-4f, La = 1.f, tolVar = 3e-4f;
        const uint32_t maxIntIter = (uint32_t)atoi(argv[8]), stoptest = (uint32_t)atoi(argv[10]),
                       miniter = (uint32_t)atoi(argv[11]);
        const int maxIter = atoi(argv[9]);
        auto nesta = Process::create<NestaUp>(pCLapp);
        nesta->setForm(NestaUp::FUSED);
        nesta->setInput(pKData);
        nesta->setOutput(pOutputXData);
        nesta->init();

        if (std::strcmp(argv[2], "register") == 0) {
            const int motionIters = argc > 12 ? atoi(argv[12]) : 1;
            // tests/MRIRecon.cpp:122-130
            const unsigned nmax = 100;
            const float et = 0.01f, eh = 0.005f;
            const std::vector<float> lambda = {0.f, 0.005f, 0.f, 0.5f};
            const unsigned radius = argc > 13 ? (unsigned)atoi(argv[13]) : 40;
            const unsigned E = 3;
            int Dp[2] = {4, 4};
            const float Wgt = argc > 14 ? (float)atof(argv[14]) : 1.f;
            const bool flagW = true;
            auto gw = Process::create<GroupwiseRegistration>(pCLapp);
            gw->setInput(pOutputXData);
            gw->setInitParameters(std::make_shared<GroupwiseRegistration::InitParameters>(Wgt, flagW, radius, E, Dp,
                                                                                           nmax, et, eh, lambda));
            gw->init();
            nesta->setLaunchParameters(std::make_shared<NestaUp::LaunchParameters>(
                lambda_i, mu_f, La, maxIntIter, tolVar, 1, maxIter, stoptest, miniter, nullptr));
            nesta->launch();
            writeX(pOutputXData, std::string(argv[3]) + ".nomc");
            std::shared_ptr<Data> TData;
            for (int i = 0; i < motionIters; ++i) {
                ArgumentsMotionCompensation* argsMC = new ArgumentsMotionCompensation();
                gw->setLaunchParameters(std::make_shared<GroupwiseRegistration::LaunchParameters>(argsMC, TData, i));
                gw->launch();
                TData = argsMC->TData;
                const auto& g = gw->getReport();
                std::cout << "registration " << i << " iters " << g.iters << " H " << g.H.front() << " "
                          << g.H.back() << "\n";
                argsMC->xnData->device2Host();
                std::ofstream xo(std::string(argv[3]) + ".xn" + std::to_string(i), std::ios::binary);
                xo.write((const char*)argsMC->xnData->getHostBuffer(0),
                         (std::streamsize)(sizeof(float) * 2 * (size_t)F * H * W));
                nesta->setLaunchParameters(std::make_shared<NestaUp::LaunchParameters>(
                    lambda_i, mu_f, La, maxIntIter, tolVar, 1, maxIter, stoptest, miniter, argsMC));
                nesta->launch();
                writeX(pOutputXData, std::string(argv[3]) + ".rec" + std::to_string(i));
                delete argsMC;   // bound_box belongs to the registration process, as in the reference
            }
            writeX(pOutputXData, argv[3]);
            const auto& r = nesta->getReport();
            std::cout << "normU " << r.normU << " iters";
            for (int n : r.itersPerStage) std::cout << " " << n;
            std::cout << "\n";
            return 0;
        }

        // the deformation, in the Data forms GroupwiseRegistration.cpp:150-164,307-314 uses
        std::ifstream mf(argv[2], std::ios::binary);
        if (!mf) throw std::runtime_error("cannot open motion file");
        auto* xnDims = new std::vector<dimIndexType>({uW, uH, (dimIndexType)(2 * F)});
        auto* xnVals = readVec<float>(mf, (size_t)2 * F * H * W);
        NDArray* xnArr = new ConcreteNDArray<float>(xnDims, xnVals);
        auto* meshDims = new std::vector<dimIndexType>({uH, uW, 2});
        auto* meshVals = readVec<int32_t>(mf, (size_t)2 * H * W);
        NDArray* meshArr = new ConcreteNDArray<int32_t>(meshDims, meshVals);
        std::vector<dimIndexType> bb(4);
        mf.read((char*)bb.data(), sizeof(dimIndexType) * 4);
        if (!mf) throw std::runtime_error("motion file too short");

        ArgumentsMotionCompensation* argsMC = new ArgumentsMotionCompensation();
        argsMC->xnData = std::make_shared<XData>(pCLapp, xnArr, TYPEID_REAL);
        argsMC->xData = std::make_shared<XData>(pCLapp, meshArr, TYPEID_INT32);
        argsMC->bound_box = new std::vector<dimIndexType>(bb);

        // first reconstruction (tests/MRIRecon.cpp:232-235)
        nesta->setLaunchParameters(std::make_shared<NestaUp::LaunchParameters>(lambda_i, mu_f, La, maxIntIter, tolVar, 1,
                                                                                maxIter, stoptest, miniter, nullptr));
        nesta->launch();
        writeX(pOutputXData, std::string(argv[3]) + ".nomc");

        auto mcOut = std::make_shared<XData>(pCLapp, pOutputXData, false);
        auto mc = Process::create<MotionCompensation>(pCLapp, pOutputXData, mcOut);
        mc->setLaunchParameters(std::make_shared<MotionCompensation::LaunchParameters>(argsMC));
        mc->launch();
        writeX(mcOut, std::string(argv[3]) + ".mc");
        auto amc = Process::create<AdjointMotionCompensation>(pCLapp, pOutputXData, mcOut);
        amc->setLaunchParameters(std::make_shared<AdjointMotionCompensation::LaunchParameters>(argsMC));
        amc->launch();
        writeX(mcOut, std::string(argv[3]) + ".mcadj");

        // reconstruction with motion compensation (tests/MRIRecon.cpp:265-270)
        nesta->setLaunchParameters(std::make_shared<NestaUp::LaunchParameters>(lambda_i, mu_f, La, maxIntIter, tolVar, 1,
                                                                                maxIter, stoptest, miniter, argsMC));
        nesta->launch();
        writeX(pOutputXData, argv[3]);
        const auto& r = nesta->getReport();
        std::cout << "normU " << r.normU << " iters";
        for (int n : r.itersPerStage) std::cout << " " << n;
        std::cout << "\n";
        delete argsMC->bound_box;
        delete argsMC;
    } catch (std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
