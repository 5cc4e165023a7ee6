// LineGraph ordering (qTorch src/LineGraph.h:44-410).  The reference obtains its
// elimination order from the external QuickBB binary; this engine writes the same
// lg.cnf and produces the order itself (greedy min-fill elimination of the line
// graph) in qbb.out's format, so the LGContract path runs end to end on any host.
#include <algorithm>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <sys/stat.h>

#include "qtorch.hpp"

namespace qtorch {

LineGraph::LineGraph(std::shared_ptr<Network> net) : mNet(std::move(net)) { Build(); }

void LineGraph::Reset(std::shared_ptr<Network> net) {
    if (!net) {
        mNet->Reset();
        return;
    }
    mNet = std::move(net);
    mGraphWires.clear();
    mEdges.clear();
    Build();
}

// src/LineGraph.h:106-153: wire ids in (uncontracted node, wire) order; every pair
// of wires on a node is an edge of the line graph.
void LineGraph::Build() {
    for (const auto& node : mNet->GetUncontractedNodes()) {
        const auto wires = node->GetWires();
        for (size_t wid = 0; wid < wires.size(); ++wid) {
            const auto& w = wires[wid];
            if (std::find(mGraphWires.begin(), mGraphWires.end(), w) == mGraphWires.end()) {
                w->SetWireID(int(mGraphWires.size()));
                mGraphWires.push_back(w);
            }
            for (size_t wid2 = 0; wid2 < wid; ++wid2) mEdges.push_back({wires[wid2], w});
        }
    }
}

// CNF text of src/LineGraph.h:251-261 (1-based wire ids).
std::string LineGraph::CnfText() const {
    std::ostringstream os;
    os << "p cnf " << mGraphWires.size() << " " << mEdges.size() << std::endl;
    for (const auto& e : mEdges) os << e.first->GetWireID() + 1 << " " << e.second->GetWireID() + 1 << " " << 0 << std::endl;
    return os.str();
}

// Greedy min-fill elimination order of the line graph; ties go to the lowest id.
std::vector<int> LineGraph::MinFillOrder() const {
    const int n = int(mGraphWires.size());
    std::vector<std::set<int>> adj(static_cast<size_t>(n));
    for (const auto& e : mEdges) {
        const int a = e.first->GetWireID(), b = e.second->GetWireID();
        if (a == b) continue;
        adj[size_t(a)].insert(b);
        adj[size_t(b)].insert(a);
    }
    std::vector<bool> alive(size_t(n), true);
    std::vector<int> order;
    order.reserve(size_t(n));
    for (int step = 0; step < n; ++step) {
        int best = -1;
        long best_fill = -1;
        for (int v = 0; v < n; ++v) {
            if (!alive[size_t(v)]) continue;
            const std::vector<int> nb(adj[size_t(v)].begin(), adj[size_t(v)].end());
            long fill = 0;
            for (size_t i = 0; i < nb.size(); ++i)
                for (size_t j = i + 1; j < nb.size(); ++j)
                    if (!adj[size_t(nb[i])].count(nb[j])) ++fill;
            if (best < 0 || fill < best_fill) {
                best = v;
                best_fill = fill;
            }
        }
        const std::vector<int> nb(adj[size_t(best)].begin(), adj[size_t(best)].end());
        for (size_t i = 0; i < nb.size(); ++i)
            for (size_t j = i + 1; j < nb.size(); ++j) {
                adj[size_t(nb[i])].insert(nb[j]);
                adj[size_t(nb[j])].insert(nb[i]);
            }
        for (int u : nb) adj[size_t(u)].erase(best);
        adj[size_t(best)].clear();
        alive[size_t(best)] = false;
        order.push_back(best + 1);
    }
    return order;
}

bool LineGraph::runQuickBB(int, void*, bool) {
    mkdir("output", 0755);
    std::remove(qbbOutName.c_str());
    {
        std::ofstream cnf(cnfName);
        cnf << CnfText();
    }
    const auto order = MinFillOrder();
    std::ofstream out(qbbOutName);
    if (!out) throw QbbFailure();
    out << " Ordering: greedy min-fill elimination of the line graph\n";
    out << " The optimal ordering is \n";
    for (int o : order) out << o << " ";
    out << "\n";
    return true;
}

// LGContract (src/LineGraph.h:296-410) with the order already parsed.
std::complex<double> LineGraph::ContractInOrder(const std::vector<int>& order) {
    for (int o : order) {
        if (o < 1 || o > int(mGraphWires.size())) throw InvalidFileFormat();
        const auto& w = mGraphWires[size_t(o - 1)];
        if (!w->IsContracted()) mNet->ContractNodes(w->GetNodeA().lock(), w->GetNodeB().lock(), 100);
    }
    const auto rem = mNet->GetUncontractedNodes();
    if (rem.size() != 1) throw ContractionFailure();
    const auto& last = rem[0];
    if (mNet->IsPlanOnly()) {
        if (last->mRank != 0) throw ContractionFailure();
        mResult = mNet->GetFinalValue();
        return mResult;
    }
    const auto vals = last->GetTensorVals();
    if (vals.size() != 1) throw ContractionFailure();
    mResult = vals[0];
    return mResult;
}

bool LineGraph::LGContract() {
    std::ifstream in(qbbOutName);
    if (!in) throw QbbFailure();
    std::vector<int> order;
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        if (line == " The optimal ordering is ") {
            std::getline(in, line);
            std::stringstream ss(line);
            for (size_t i = 0; i < mGraphWires.size(); ++i) {
                int w = 0;
                ss >> w;
                order.push_back(w);
            }
            found = true;
            break;
        }
    }
    if (!found) return false;
    ContractInOrder(order);
    return true;
}

}  // namespace qtorch
