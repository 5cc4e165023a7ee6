// The QAOA outer loop and the final cut (qTorch src/maxcut.cpp:29-140, 145-222) on the
// device.  The reference maximises F_p with nlopt's COBYLA (src/maxcut.cpp:211-213),
// which is not part of this engine; here a seeded batch of random angle vectors is
// evaluated in one launch and the best one is refined by a Nelder-Mead simplex whose
// vertices are also evaluated as batches.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <sstream>

#include "batch.hpp"
#include "maxcut_driver.hpp"

namespace qtorch {

namespace {

std::vector<double> EvalSets(QaoaBatch& b, const std::vector<std::vector<double>>& sets) {
    std::vector<double> flat;
    for (const auto& s : sets) flat.insert(flat.end(), s.begin(), s.end());
    b.SetAngles(int(sets.size()), flat.data());
    if (qt_batch_launch(b.Batch()) != QT_OK) throw DeviceFailure("qt_batch_launch failed");
    std::vector<double> e(sets.size());
    b.Results(e.data(), nullptr);
    return e;
}

}  // namespace

OptimizeResult MaxcutOptimalAngles(qt_ctx* ctx, const std::string& dgf, int p, int precision, int n_random,
                                   int max_evals, uint64_t seed) {
    if (p < 1 || n_random < 1 || max_evals < 1) throw InvalidFunctionInput();
    QaoaBatch batch(ctx, dgf, p, true, precision, kOrderGreedy, 0);
    const int dim = 2 * p;
    OptimizeResult res;
    // the reference's starting point (src/maxcut.cpp:155-157) plus seeded random points:
    // betas in [0, pi), gammas in [-2 pi, 2 pi)
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> ub(0.0, M_PI), ug(-2 * M_PI, 2 * M_PI);
    std::vector<std::vector<double>> pts;
    std::vector<double> start(static_cast<size_t>(dim));
    for (int i = 0; i < p; ++i) {
        start[size_t(i)] = 0.392699;
        start[size_t(i + p)] = 0.785399;
    }
    pts.push_back(start);
    for (int k = 0; k < n_random; ++k) {
        std::vector<double> x(static_cast<size_t>(dim));
        for (int i = 0; i < p; ++i) {
            x[size_t(i)] = ub(gen);
            x[size_t(i + p)] = ug(gen);
        }
        pts.push_back(x);
    }
    auto vals = EvalSets(batch, pts);
    res.evaluations += int(pts.size());
    const size_t best0 = size_t(std::max_element(vals.begin(), vals.end()) - vals.begin());
    // Nelder-Mead (maximisation) from the best sample
    std::vector<std::vector<double>> simplex{pts[best0]};
    for (int i = 0; i < dim; ++i) {
        auto v = pts[best0];
        v[size_t(i)] += 0.1;
        simplex.push_back(v);
    }
    std::vector<double> f = EvalSets(batch, simplex);
    res.evaluations += int(simplex.size());
    while (res.evaluations < max_evals) {
        std::vector<size_t> order(simplex.size());
        std::iota(order.begin(), order.end(), size_t(0));
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return f[a] > f[b]; });
        std::vector<std::vector<double>> s2;
        std::vector<double> f2;
        for (size_t i : order) {
            s2.push_back(simplex[i]);
            f2.push_back(f[i]);
        }
        simplex.swap(s2);
        f.swap(f2);
        if (std::fabs(f.front() - f.back()) < 1e-10 * std::max(1.0, std::fabs(f.front()))) break;
        const size_t worst = simplex.size() - 1;
        std::vector<double> centroid(static_cast<size_t>(dim), 0.0);
        for (size_t i = 0; i < worst; ++i)
            for (int d = 0; d < dim; ++d) centroid[size_t(d)] += simplex[i][size_t(d)] / double(worst);
        auto along = [&](double t) {
            std::vector<double> x(static_cast<size_t>(dim));
            for (int d = 0; d < dim; ++d)
                x[size_t(d)] = centroid[size_t(d)] + t * (simplex[worst][size_t(d)] - centroid[size_t(d)]);
            return x;
        };
        // reflection, expansion and both contractions in one launch
        const std::vector<std::vector<double>> cand{along(-1.0), along(-2.0), along(-0.5), along(0.5)};
        const auto fc = EvalSets(batch, cand);
        res.evaluations += 4;
        if (fc[0] > f[0] && fc[1] > fc[0]) {
            simplex[worst] = cand[1];
            f[worst] = fc[1];
        } else if (fc[0] > f[worst - 1]) {
            simplex[worst] = cand[0];
            f[worst] = fc[0];
        } else if (std::max(fc[2], fc[3]) > f[worst]) {
            const int k = fc[2] > fc[3] ? 2 : 3;
            simplex[worst] = cand[size_t(k)];
            f[worst] = fc[size_t(k)];
        } else {  // shrink towards the best vertex
            for (size_t i = 1; i < simplex.size(); ++i)
                for (int d = 0; d < dim; ++d)
                    simplex[i][size_t(d)] = simplex[0][size_t(d)] + 0.5 * (simplex[i][size_t(d)] - simplex[0][size_t(d)]);
            std::vector<std::vector<double>> rest(simplex.begin() + 1, simplex.end());
            const auto fr = EvalSets(batch, rest);
            res.evaluations += int(rest.size());
            for (size_t i = 1; i < simplex.size(); ++i) f[i] = fr[i - 1];
        }
    }
    const size_t b = size_t(std::max_element(f.begin(), f.end()) - f.begin());
    res.angles = simplex[b];
    res.energy = f[b];
    return res;
}

// maxcutGetFinalString (src/maxcut.cpp:29-140): qubit by qubit, the conditional
// probability of 0 given the bits chosen so far decides the next bit (ties: a seeded
// coin instead of rand()).  Every probability is one full-circuit contraction on the
// device, in the greedy order planned once for the circuit.
std::vector<int> MaxcutFinalString(qt_ctx* ctx, const std::string& dgf, int p, const std::vector<double>& bg,
                                   uint64_t seed, int* cut_edges, std::vector<double>* cond_probs) {
    ExtraData d = ExtraData::FromText(p, dgf, true);
    std::ostringstream q;
    q << d.numQubits << std::endl;
    for (int i = 0; i < d.numQubits; ++i) q << "H " << i << std::endl;
    for (int i = 0; i < p; ++i) {  // applyU_CsThenU_Bs over every graph edge (src/maxcut.h:205-233)
        for (const auto& e : d.pairs) {
            q << "CNOT " << e.first << " " << e.second << std::endl;
            q << "Rz " << -bg[size_t(i + p)] << " " << e.second << std::endl;
            q << "CNOT " << e.first << " " << e.second << std::endl;
        }
        for (int j = 0; j < d.numQubits; ++j) q << "Rx " << bg[size_t(i)] * 2.0 << " " << j << std::endl;
    }
    const std::string qasm = q.str();
    std::mt19937_64 coin(seed);
    std::vector<int> bits;
    double current = 1.0;
    for (int i = 0; i < d.numQubits; ++i) {
        std::string m;
        for (int j = 0; j < d.numQubits; ++j)
            m += j < int(bits.size()) ? (bits[size_t(j)] ? "1 " : "0 ") : (j == i ? "0 " : "T ");
        NetworkOptions opt;
        opt.ctx = ctx;
        auto net = std::make_shared<Network>(qasm, &m, opt);
        GreedyContract(*net);
        const double cond = net->GetFinalValue().real() / current;
        if (cond_probs) cond_probs->push_back(cond);
        if (cond > 0.5) {
            current *= cond;
            bits.push_back(0);
        } else if (cond < 0.5) {
            current *= 1.0 - cond;
            bits.push_back(1);
        } else {
            bits.push_back(int(coin() & 1));
            current *= 0.5;
        }
    }
    int cut = 0;
    for (const auto& e : d.pairs)
        if (bits[size_t(e.first)] != bits[size_t(e.second)]) ++cut;
    if (cut_edges) *cut_edges = cut;
    return bits;
}

}  // namespace qtorch
