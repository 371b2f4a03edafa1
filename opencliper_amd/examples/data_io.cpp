// Loads KData or XData from a file through the OpenCLIPER-compatible API, prints what was loaded as one JSON line and
// saves it again:
//   data_io kdata <in.mat|in.cfl> <out.mat> [masks]   matlabSave (masks: also write the sampling masks)
//   data_io kdata <in.mat|in.cfl> <out.cfl>           saveCFLData(<out>)
//   data_io xdata <in.mat> <out.mat>
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/XData.hpp>

#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>

using namespace OpenCLIPER;

static void printDims(const char* key, const std::vector<dimIndexType>& d) {
    std::printf("\"%s\": [", key);
    for (size_t i = 0; i < d.size(); ++i) std::printf("%s%u", i ? ", " : "", d[i]);
    std::printf("]");
}

int main(int argc, char* argv[]) {
    if (argc < 4) {
        std::cerr << "usage: data_io kdata|xdata in.mat out.mat [masks]\n";
        return 2;
    }
    try {
        auto pCLapp = CLapp::create();
        if (!std::strcmp(argv[1], "kdata")) {
            auto k = std::make_shared<KData>(pCLapp, std::string(argv[2]));
            std::printf("{\"kind\": \"kdata\", \"ncoils\": %u, \"narrays\": %u, ", k->getNCoils(), k->getNumNDArrays());
            printDims("spatial", *k->getNDArray(0)->getDims());
            std::printf(", ");
            printDims("temporal", *k->getDynDims());
            std::printf(", \"sens\": %s, \"masks\": %s}\n", k->getSensitivityMapsData() ? "true" : "false",
                        k->getSamplingMasksData() ? "true" : "false");
            const std::string out = argv[3];
            if (out.size() > 4 && out.compare(out.size() - 4, 4, ".cfl") == 0)
                k->saveCFLData(out.substr(0, out.size() - 4));
            else
                k->matlabSave(out, argc > 4 && !std::strcmp(argv[4], "masks"));
        } else if (!std::strcmp(argv[1], "xdata")) {
            auto x = std::make_shared<XData>(pCLapp, std::string(argv[2]));
            std::printf("{\"kind\": \"xdata\", \"narrays\": %u, ", x->getNumNDArrays());
            printDims("spatial", *x->getNDArray(0)->getDims());
            std::printf(", ");
            printDims("temporal", *x->getDynDims());
            std::printf("}\n");
            x->matlabSave(argv[3]);
        } else {
            std::cerr << "data_io: unknown kind\n";
            return 2;
        }
    } catch (const std::exception& e) {
        std::cerr << "data_io: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
