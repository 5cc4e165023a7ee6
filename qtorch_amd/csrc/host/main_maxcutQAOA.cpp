// maxcutQAOA <graph.dgf> <p> <0|1|2> <file> [file] [seconds]: the QAOA MaxCut driver
// (qTorch src/maxcut.cpp:227-328) on the device engine.
//   0: find angles, write them to <file>
//   1: read angles from <file>, write the greedy cut to the second file
//   2: both; the answer goes to <file>
// The seconds argument (the reference's preprocessing budget) is accepted and unused:
// the full-circuit order is planned greedily.
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "maxcut_driver.hpp"

int main(int argc, char** argv) {
    if (argc < 5) {
        std::cout << "Not enough arguments" << std::endl;
        std::cout << "arguments: <GraphFile Path> <p value> <0 for getAngles> <file path to output angle file>\n";
        std::cout << "OR arguments: <GraphFile Path> <p value> <1 for final cut> <input angle file> <output answer file>\n";
        std::cout << "OR arguments: <GraphFile Path> <p value> <2 for both> <output answer file>\n";
        return -1;
    }
    const int p = std::atoi(argv[2]);
    const int mode = std::atoi(argv[3]);
    std::ifstream g(argv[1]);
    if (!g) {
        std::cout << "Could Not Open File" << std::endl;
        return -1;
    }
    std::stringstream ss;
    ss << g.rdbuf();
    const std::string dgf = ss.str();
    try {
        qt_ctx* ctx = qtorch::DefaultContext();
        std::vector<double> angles;
        if (mode == 0 || mode == 2) {
            auto r = qtorch::MaxcutOptimalAngles(ctx, dgf, p, QT_C64_4M, 256, 4000, 0);
            angles = r.angles;
            std::cout << "F_p = " << r.energy << " after " << r.evaluations << " evaluations" << std::endl;
            if (mode == 0) {
                std::ofstream out(argv[4]);
                for (double a : angles) out << a << " ";
                return 0;
            }
        } else if (mode == 1) {
            if (argc < 6) {
                std::cout << "Not enough arguments" << std::endl;
                return -1;
            }
            std::ifstream in(argv[4]);
            for (int i = 0; i < 2 * p; ++i) {
                double z = 0;
                in >> z;
                angles.push_back(z);
            }
        } else {
            return 0;
        }
        int cut = 0;
        const auto bits = qtorch::MaxcutFinalString(ctx, dgf, p, angles, 0, &cut);
        std::ofstream out(mode == 1 ? argv[5] : argv[4]);
        out << argv[1] << std::endl;
        for (int b : bits) out << b << " ";
        out << std::endl << "Cut edges: " << cut << std::endl;
    } catch (std::exception& e) {
        std::cout << e.what() << std::endl;
        return -1;
    }
    return 0;
}
