// QAOA MaxCut driver pieces (qTorch src/maxcut.h:31-233, src/maxcut.cpp:145-204):
// graph reading, per-edge light cones, the QASM/measurement text F_p writes, and
// the per-edge energy evaluation on the device.
#include <fstream>
#include <sstream>

#include "qtorch.hpp"

namespace qtorch {

ExtraData::ExtraData(int p0, const char* filename) : fileName(filename), p(p0) {
    ReadInData();
    PopulateIterations();
}

ExtraData ExtraData::FromText(int p, const std::string& dgf_text, bool reference_qubit_map) {
    ExtraData d;
    d.p = p;
    d.referenceQubitMap = reference_qubit_map;
    d.ParseData(dgf_text);
    d.PopulateIterations();
    return d;
}

void ExtraData::ReadInData() {  // src/maxcut.h:56-103
    std::ifstream in(fileName);
    if (!in.is_open()) throw InvalidFile();
    std::stringstream ss;
    ss << in.rdbuf();
    ParseData(ss.str());
}

// 'c' starts a comment line, 'e u v' is an edge; numQubits = max vertex + 1.
void ExtraData::ParseData(const std::string& text) {
    std::istringstream in(text);
    pairs.clear();
    int mx = 0;
    char ch;
    std::string rest;
    while (in >> ch) {
        if (ch != 'e') {  // 'c' comments and malformed lines are skipped
            std::getline(in, rest);
            continue;
        }
        int one, two;
        if (!(in >> one >> two)) throw InvalidFileFormat();
        mx = std::max(mx, std::max(one, two));
        pairs.push_back({one, two});
    }
    numQubits = mx + 1;
    adjacencyLists.assign(size_t(numQubits), {});
    for (const auto& e : pairs) {
        adjacencyLists[size_t(e.first)].push_back(e.second);
        adjacencyLists[size_t(e.second)].push_back(e.first);
    }
}

// PopulateIterations + Helper (src/maxcut.h:105-186): a p-level breadth-first light
// cone per edge; each visited (vertex, unchecked neighbour) pair is one ZZ term, and
// vertices are renumbered in first-visit order.
void ExtraData::PopulateIterations() {
    const size_t n = size_t(numQubits);
    qubitsNeeded.assign(pairs.size(), 0);
    iterations.assign(pairs.size(), {});
    realIterations.assign(pairs.size(), {});
    for (size_t e = 0; e < pairs.size(); ++e) {
        const auto [first, second] = pairs[e];
        std::vector<int> map(n, -1);
        map[size_t(first)] = 0;
        qubitsNeeded[e] = 1;
        if (!referenceQubitMap) {
            map[size_t(second)] = 1;
            qubitsNeeded[e] = 2;
        }
        std::vector<int> working{first, second};
        std::vector<bool> checked(n, false), in_working(n, false), fresh(n, false);
        in_working[size_t(first)] = in_working[size_t(second)] = true;
        for (int level = 0; level < p; ++level) {
            for (int v : working) {
                for (int u : adjacencyLists[size_t(v)]) {
                    if (checked[size_t(u)]) continue;
                    if (map[size_t(u)] == -1) map[size_t(u)] = qubitsNeeded[e]++;
                    iterations[e].push_back({v, u});
                    if (!in_working[size_t(u)]) fresh[size_t(u)] = true;
                }
                checked[size_t(v)] = true;
            }
            working.clear();
            std::fill(in_working.begin(), in_working.end(), false);
            for (size_t i = 0; i < n; ++i)
                if (fresh[i]) {
                    working.push_back(int(i));
                    in_working[i] = true;
                }
            std::fill(fresh.begin(), fresh.end(), false);
        }
        for (const auto& it : iterations[e])
            realIterations[e].push_back({map[size_t(it.first)], map[size_t(it.second)]});
    }
}

// src/maxcut.h:194-233 and src/maxcut.cpp:172-178
std::string QaoaEdgeQasm(const ExtraData& d, int e, const std::vector<double>& bg) {
    std::ostringstream op;
    const int nq = d.qubitsNeeded[size_t(e)];
    op << nq << std::endl;
    for (int i = 0; i < nq; ++i) op << "H " << i << std::endl;
    for (int i = 0; i < d.p; ++i) {
        const double gamma = bg[size_t(i + d.p)], beta = bg[size_t(i)];
        for (const auto& pr : d.realIterations[size_t(e)]) {
            op << "CNOT " << pr.first << " " << pr.second << std::endl;
            op << "Rz " << -gamma << " " << pr.second << std::endl;
            op << "CNOT " << pr.first << " " << pr.second << std::endl;
        }
        for (int j = 0; j < nq; ++j) op << "Rx " << beta * 2.0 << " " << j << std::endl;
    }
    return op.str();
}

std::string QaoaEdgeMeasure(const ExtraData& d, int e) {  // src/maxcut.cpp:179-188
    std::string m;
    for (int i = 0; i < d.qubitsNeeded[size_t(e)]; ++i) m += (i == 0 || i == 1) ? "Z " : "T ";
    return m;
}

double QaoaEnergyPlain(qt_ctx* ctx, const ExtraData& d, const std::vector<double>& bg, uint64_t seed,
                       std::vector<double>* zz) {
    double f = 0.0;
    if (zz) zz->assign(d.pairs.size(), 0.0);
    for (size_t e = 0; e < d.pairs.size(); ++e) {
        const std::string q = QaoaEdgeQasm(d, int(e), bg);
        const std::string m = QaoaEdgeMeasure(d, int(e));
        NetworkOptions opt;
        opt.ctx = ctx;
        auto net = std::make_shared<Network>(q, &m, opt);
        ContractionTools ct(net);
        ct.SetSeed(seed + e);
        ct.Contract(Stochastic);
        const double v = ct.GetFinalVal().real();
        if (zz) (*zz)[e] = v;
        f += 0.5 * (1.0 - v);  // src/maxcut.cpp:196
    }
    return f;
}

}  // namespace qtorch
