// qtorch <script.inp>: the measure-file driver (qTorch src/main.cpp:42-309) on the
// device engine.  Script keys (leviParser, src/leviParser.hpp:41-104):
//   >string qasm <file>   >string measurement <file>   >string outputpath <file>
//   >string contractmethod linegraph-qbb | simple-stoch | user-defined
//   >string user-contract-seq <file>   >int threads n   >int quickbbseconds s
//   >bool qbbonly b   >bool readqbbresonly b   >int seed n (this engine: stochastic seed)
// The output file gets the value, the complex-MAC count and the elapsed time, as the
// reference writes them.  linegraph-qbb writes lg.cnf and a min-fill qbb.out itself.
#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "qtorch.hpp"

namespace {

struct Script {
    std::map<std::string, std::string> s;
    std::map<std::string, bool> b;
    std::map<std::string, int> i;
    std::map<std::string, double> d;
};

// leviParser::readInputFile: lines ">type name value"
bool ReadScript(const std::string& path, Script& sc) {
    std::ifstream f(path);
    if (!f.is_open()) return false;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] != '>') continue;
        std::stringstream ss(line);
        std::string type, name, value;
        ss >> type >> name >> value;
        if (type == ">string") {
            sc.s[name] = value;
        } else if (type == ">bool") {
            if (value == "1" || value == "true" || value == "True" || value == "yes" || value == "Yes") sc.b[name] = true;
            else if (value == "0" || value == "false" || value == "False" || value == "no" || value == "No") sc.b[name] = false;
            else std::cout << "Error in leviParser, " << name << ". bool inputs must be 1, true, True, yes, or Yes." << std::endl;
        } else if (type == ">int") {
            sc.i[name] = std::atoi(value.c_str());
        } else if (type == ">double") {
            sc.d[name] = std::atof(value.c_str());
        } else {
            std::cout << "Error. Only the following types are supported in leviParser: string, bool, int, double." << std::endl;
        }
    }
    return true;
}

std::string Formatted(double t) { return " { " + std::to_string(t) + " } "; }

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << "Usage:\nExecutable <input file>\n";
        return -1;
    }
    Script sc;
    sc.i["quickbbseconds"] = 20;  // setInpDefaultsTN (src/main.cpp:316-336)
    sc.i["threads"] = 2;
    sc.b["qbbonly"] = false;
    sc.b["readqbbresonly"] = false;
    sc.s["outputpath"] = "output/qtorch.out";
    if (!ReadScript(argv[1], sc)) std::cout << "Unable to open file.";
    mkdir("output", 0755);
    std::ofstream out(sc.s["outputpath"]);
    if (!out) {
        std::cout << "Invalid Output File Path" << std::endl;
        return -1;
    }
    std::shared_ptr<qtorch::Network> net;
    try {
        net = std::make_shared<qtorch::Network>(sc.s["qasm"], sc.s["measurement"]);
    } catch (std::exception& e) {
        std::cout << e.what() << std::endl;
        out << e.what() << std::endl;
        return -1;
    }
    net->SetNumThreads(sc.i["threads"]);
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    bool succ = false;
    try {
        net->ReduceCircuit();  // src/main.cpp:112
        const std::string method = sc.s["contractmethod"];
        std::complex<double> value;
        if (method == "linegraph-qbb") {
            qtorch::LineGraph lg(net);
            if (sc.b["qbbonly"]) {
                lg.runQuickBB(sc.i["quickbbseconds"]);
                std::cout << "QuickBB-format ordering written. Set qbbonly=false and readqbbresonly=true to contract.\n";
                return 0;
            }
            if (!sc.b["readqbbresonly"]) lg.runQuickBB(sc.i["quickbbseconds"]);
            succ = lg.LGContract();
            value = net->GetFinalValue();
        } else if (method == "simple-stoch" || (method == "user-defined" && !sc.s.count("user-contract-seq"))) {
            qtorch::ContractionTools ct(net);
            if (sc.i.count("seed")) ct.SetSeed(uint64_t(sc.i["seed"]));
            succ = ct.Contract(qtorch::Stochastic) != nullptr;
            value = ct.GetFinalVal();
        } else if (method == "user-defined") {
            qtorch::ContractionTools ct(net);
            succ = ct.ContractUserDefinedSequenceOfWires(sc.s["user-contract-seq"]) != nullptr;
            value = ct.GetFinalVal();
        } else {
            std::cout << "Error. 'contractmethod' bad option.\n";
            return -1;
        }
        std::cout.precision(17);
        out.precision(17);
        std::cout << "Result of Contraction: " << value << std::endl;
        out << "Result of Contraction: " << value << std::endl;
    } catch (std::exception& e) {
        std::cout << e.what() << std::endl;
        out << e.what() << std::endl;
        succ = false;
    }
    if (!succ) {
        std::cout << "ERROR. ABORTING.\n" << Formatted(elapsed()) << "\n";
        out << "ERROR. ABORTING.\n" << Formatted(elapsed()) << "\n";
        return -1;
    }
    std::cout << "Number of floating point ops in full contraction: " << net->getNumFloatOps() << "\n";
    out << "Number of floating point ops in full contraction: " << net->getNumFloatOps() << "\n";
    std::cout << "Contraction complete. " << Formatted(elapsed()) << "\n";
    out << "Contraction complete. " << Formatted(elapsed()) << "\n";
    return 0;
}
