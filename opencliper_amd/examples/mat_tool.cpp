// Host-only MAT v5 utility over OpenCLIPER::mat (no device needed):
//   mat_tool dump <file.mat>                      one JSON line per numeric variable: name, class, complex, dims and
//                                                 the sums of the real parts, imaginary parts and squared magnitudes
//   mat_tool copy <in.mat> <out.mat> <compress>   reads every numeric variable and writes them again (compress 0 / 1)
#include <OpenCLIPER/MatFile.hpp>

#include <cstdio>
#include <cstdlib>
#include <iostream>

using namespace OpenCLIPER;

int main(int argc, char* argv[]) {
    if (argc < 3) {
        std::cerr << "usage: mat_tool dump file.mat | mat_tool copy in.mat out.mat compress\n";
        return 2;
    }
    try {
        const std::string cmd = argv[1];
        if (cmd == "dump") {
            for (const auto& kv : mat::read(argv[2])) {
                const mat::Var& v = kv.second;
                double sre = 0, sim = 0, s2 = 0;
                for (size_t i = 0; i < v.re.size(); ++i) {
                    sre += v.re[i];
                    s2 += v.re[i] * v.re[i];
                    if (v.complex) {
                        sim += v.im[i];
                        s2 += v.im[i] * v.im[i];
                    }
                }
                std::printf("{\"name\": \"%s\", \"class\": %d, \"complex\": %s, \"dims\": [", v.name.c_str(), v.mclass,
                            v.complex ? "true" : "false");
                for (size_t i = 0; i < v.dims.size(); ++i)
                    std::printf("%s%llu", i ? ", " : "", (unsigned long long)v.dims[i]);
                std::printf("], \"sum_re\": %.17g, \"sum_im\": %.17g, \"sum_abs2\": %.17g}\n", sre, sim, s2);
            }
        } else if (cmd == "copy" && argc >= 5) {
            std::vector<mat::Var> vars;
            for (auto& kv : mat::read(argv[2])) vars.push_back(kv.second);
            mat::write(argv[3], vars, std::atoi(argv[4]) != 0);
        } else {
            std::cerr << "mat_tool: unknown command\n";
            return 2;
        }
    } catch (const std::exception& e) {
        std::cerr << "mat_tool: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
