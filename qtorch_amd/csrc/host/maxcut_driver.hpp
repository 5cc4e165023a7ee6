// QAOA outer loop and final cut on the device (see maxcut_driver.cpp).
#pragma once

#include <string>
#include <vector>

#include "qtorch.hpp"

namespace qtorch {

struct OptimizeResult {
    std::vector<double> angles;  // [betas..., gammas...] as F_p takes them
    double energy = 0.0;
    int evaluations = 0;
};

OptimizeResult MaxcutOptimalAngles(qt_ctx* ctx, const std::string& dgf, int p, int precision, int n_random,
                                   int max_evals, uint64_t seed);
std::vector<int> MaxcutFinalString(qt_ctx* ctx, const std::string& dgf, int p, const std::vector<double>& bg,
                                   uint64_t seed, int* cut_edges,
                                   std::vector<double>* cond_probs = nullptr);

}  // namespace qtorch
