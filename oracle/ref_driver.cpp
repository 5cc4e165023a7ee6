// Test-infrastructure driver over the qTorch reference headers (compiled from the
// reference sources where they lie; see build_ref.sh).  It is the "oracle/_ref"
// checker: it never ships and the product never links it.
//
// Commands (all results printed on stdout with %.17g precision; the reference's
// own std::cout chatter is discarded):
//   stoch  <qasm> <meas>                 ContractionTools::Contract(Stochastic)
//                                        (src/ContractionTools.h:137-148,219-388),
//                                        prints VALUE, FLOPS and the recorded
//                                        mCreatedFrom sequence (src/preprocess.h:41-46)
//   given  <qasm> <meas> <seqfile>       ContractGivenSequence (src/ContractionTools.h:392-411)
//   user   <qasm> <meas> <seqfile>       ContractUserDefinedSequenceOfWires (:154-214)
//   rstoch <qasm> <meas>                 ReduceCircuit (src/Network.h:1015) then Stochastic
//   rgiven <qasm> <meas> <seqfile>       ReduceCircuit then replay a sequence
//   lg     <qasm> <meas> <qbbout> <0|1>  optional ReduceCircuit, then LineGraph::LGContract
//                                        (src/LineGraph.h:296-410) on a given qbb.out
//   nwires <qasm> <meas> <0|1>           number of line-graph nodes (distinct wires)
//   time   <listfile> <reps>             wall time of Contract(Stochastic) over every
//                                        "<qasm> <meas>" line of listfile (CPU baseline)
//   cones  <dgf> <p> <bg_0> .. <bg_2p-1> ExtraData(p, dgf) (src/maxcut.h:31-186): per
//                                        edge CONE e qubitsNeeded, its realIterations
//                                        as IT a b, and the QASM text F_p writes for it
//                                        (src/maxcut.cpp:172-178) produced by the
//                                        reference's own outputInitialPlusStateToFile
//                                        and applyU_CsThenU_Bs (src/maxcut.h:191-233)
//                                        between QASM_BEGIN / QASM_END
#include "ContractionTools.h"
#include "LineGraph.h"
#include "maxcut.h"   // needs nlopt-2.4.2/api on the include path (headers only)

#include <chrono>
#include <cstdlib>
#include <stdexcept>
#include <unistd.h>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

using namespace qtorch;

namespace {

std::vector<std::pair<int, int>> read_seq(const std::string& path) {
    std::vector<std::pair<int, int>> seq;
    std::ifstream in(path);
    int a, b;
    while (in >> a >> b) seq.push_back({a, b});
    return seq;
}

void print_value(std::complex<double> v) {
    std::printf("VALUE %.17g %.17g\n", v.real(), v.imag());
}

void print_record(const std::shared_ptr<Network>& net) {
    for (const auto& node : net->GetAllNodes()) {
        if (!(node->mCreatedFrom.first == 0 && node->mCreatedFrom.second == 0))
            std::printf("SEQ %d %d\n", node->mCreatedFrom.first, node->mCreatedFrom.second);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: ref_driver <cmd> args...\n");
        return 2;
    }
    std::streambuf* saved = std::cout.rdbuf();
    std::ofstream devnull("/dev/null");
    std::cout.rdbuf(devnull.rdbuf());
    const std::string cmd = argv[1];
    int rc = 0;
    try {
        if (cmd == "stoch" && argc >= 4) {
            ContractionTools ct(argv[2], argv[3]);
            auto net = ct.Contract(Stochastic);
            print_value(ct.GetFinalVal());
            std::printf("FLOPS %lld\n", net->getNumFloatOps());
            print_record(net);
        } else if (cmd == "given" && argc >= 5) {
            ContractionTools ct(argv[2], argv[3]);
            auto net = ct.ContractGivenSequence(read_seq(argv[4]));
            print_value(ct.GetFinalVal());
            std::printf("FLOPS %lld\n", net->getNumFloatOps());
        } else if (cmd == "user" && argc >= 5) {
            ContractionTools ct(argv[2], argv[3]);
            auto net = ct.ContractUserDefinedSequenceOfWires(argv[4]);
            print_value(ct.GetFinalVal());
            std::printf("FLOPS %lld\n", net->getNumFloatOps());
        } else if ((cmd == "rstoch" && argc >= 4) || (cmd == "rgiven" && argc >= 5)) {
            auto net = std::make_shared<Network>(argv[2], argv[3]);
            net->ReduceCircuit();
            std::printf("REDUCED_NODES %zu\n", net->GetUncontractedNodes().size());
            ContractionTools ct(net);
            if (cmd == "rstoch") {
                ct.Contract(Stochastic);
                print_value(ct.GetFinalVal());
                std::printf("FLOPS %lld\n", net->getNumFloatOps());
                print_record(net);
            } else {
                ct.ContractGivenSequence(read_seq(argv[4]));
                print_value(ct.GetFinalVal());
                std::printf("FLOPS %lld\n", net->getNumFloatOps());
            }
        } else if (cmd == "lg" && argc >= 6) {
            auto net = std::make_shared<Network>(argv[2], argv[3]);
            if (std::atoi(argv[5])) net->ReduceCircuit();
            LineGraph lg(net);
            lg.SetQBBOutFiles("/dev/null", argv[4], "/dev/null");
            lg.LGContract();
            auto rem = net->GetUncontractedNodes();
            print_value(rem[0]->GetTensorVals()[0]);
            std::printf("FLOPS %lld\n", net->getNumFloatOps());
        } else if (cmd == "nwires" && argc >= 5) {
            auto net = std::make_shared<Network>(argv[2], argv[3]);
            if (std::atoi(argv[4])) net->ReduceCircuit();
            std::vector<std::shared_ptr<Wire>> seen;
            for (const auto& n : net->GetUncontractedNodes())
                for (const auto& w : n->GetWires())
                    if (std::find(seen.begin(), seen.end(), w) == seen.end()) seen.push_back(w);
            std::printf("NWIRES %zu\n", seen.size());
        } else if (cmd == "time" && argc >= 4) {
            std::vector<std::pair<std::string, std::string>> jobs;
            std::ifstream in(argv[2]);
            std::string q, m;
            while (in >> q >> m) jobs.push_back({q, m});
            const int reps = std::atoi(argv[3]);
            double sum = 0.0;
            auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < reps; ++r)
                for (const auto& j : jobs) {
                    ContractionTools ct(j.first, j.second);
                    ct.Contract(Stochastic);
                    sum += ct.GetFinalVal().real();
                }
            auto t1 = std::chrono::steady_clock::now();
            std::printf("TIME %.9f\nJOBS %zu\nCHECKSUM %.17g\n",
                        std::chrono::duration<double>(t1 - t0).count(), jobs.size() * reps, sum);
        } else if (cmd == "cones" && argc >= 4) {
            const int p = std::atoi(argv[3]);
            if (argc < 4 + 2 * p) throw std::runtime_error("cones needs 2p angles");
            std::vector<double> bg;
            for (int i = 0; i < 2 * p; ++i) bg.push_back(std::strtod(argv[4 + i], nullptr));
            ExtraData ed(p, argv[2]);
            char tmpl[] = "/tmp/ref_cones_XXXXXX";
            const int fd = mkstemp(tmpl);
            if (fd < 0) throw std::runtime_error("mkstemp failed");
            close(fd);
            for (size_t e = 0; e < ed.pairs.size(); ++e) {
                const int nq = ed.qubitsNeeded[e];
                std::printf("CONE %zu %d\n", e, nq);
                for (const auto& it : ed.realIterations[e]) std::printf("IT %d %d\n", it.first, it.second);
                {
                    // the same calls F_p makes per edge (src/maxcut.cpp:172-178)
                    std::ofstream q(tmpl);
                    q << nq << std::endl;
                    outputInitialPlusStateToFile(q, nq);
                    applyU_CsThenU_Bs(ed.realIterations[e], ed.p, bg, nq, q);
                }
                std::ifstream back(tmpl);
                std::stringstream text;
                text << back.rdbuf();
                std::printf("QASM_BEGIN\n%sQASM_END\n", text.str().c_str());
            }
            std::remove(tmpl);
        } else {
            std::fprintf(stderr, "bad command\n");
            rc = 2;
        }
    } catch (std::exception& e) {
        std::printf("EXCEPTION %s\n", e.what());
        rc = 3;
    }
    std::cout.rdbuf(saved);
    return rc;
}
