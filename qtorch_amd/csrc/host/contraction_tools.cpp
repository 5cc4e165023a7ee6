// ContractionTools: stochastic, given and user-defined orderings
// (qTorch src/ContractionTools.h:52-411,1261-1279).  Each ordering decides only
// which pairs to contract; Network::ContractNodes runs the contraction on the device.
#include <cmath>
#include <fstream>
#include <sstream>

#include "qtorch.hpp"

namespace qtorch {

ContractionTools::ContractionTools(const std::string& inputFile, const std::string& measureFile, int numThreads)
    : mString(inputFile), mMeasureFile(measureFile), mRandGen(std::random_device{}()),
      mNumThreadsInNetwork(numThreads) {}

ContractionTools::ContractionTools(std::shared_ptr<Network> network)
    : mNetwork(std::move(network)), mRandGen(std::random_device{}()), mCopyCreated(true) {}

void ContractionTools::Reset() {
    mNetwork = nullptr;
    mCopyCreated = false;
}

void ContractionTools::Reset(const std::string& inputFile, const std::string& measureFile, int numThreads) {
    mNetwork = nullptr;
    mCopyCreated = false;
    mString = inputFile;
    mMeasureFile = measureFile;
    mNumThreadsInNetwork = numThreads;
}

std::shared_ptr<Network> ContractionTools::MakeNetwork() {
    if (mCopyCreated) return mNetwork;
    auto net = std::make_shared<Network>(mString, mMeasureFile);
    net->SetNumThreads(mNumThreadsInNetwork);
    return net;
}

std::shared_ptr<Network> ContractionTools::Contract(ContractionType type, int, int) {
    if (type != Stochastic) throw InvalidContractionMethod();
    return ParallelContract(mRandGen);
}

namespace {

// Replaces the two contracted entries by the result the way the reference does
// (swap-with-back pops, src/ContractionTools.h:280-294).
void Splice(std::vector<std::shared_ptr<Node>>& v, int one, int two, const std::shared_ptr<Node>& result) {
    if (two == int(v.size()) - 1) {
        v[size_t(two)] = v.back();
        v.pop_back();
        v[size_t(one)] = v.back();
        v.pop_back();
    } else {
        v[size_t(one)] = v.back();
        v.pop_back();
        v[size_t(two)] = v.back();
        v.pop_back();
    }
    v.push_back(result);
}

}  // namespace

// ParallelContract (src/ContractionTools.h:219-388).  The partitions of
// CreateChunksOfNodes (:1261-1279) are contracted one after another with one seeded
// generator (the reference runs them on two threads sharing one unseeded
// generator, so its order cannot be reproduced; any order gives the same value to
// rounding).  Rank-threshold rules, failure counters and the final pass are the
// reference's.
std::shared_ptr<Network> ContractionTools::ParallelContract(std::mt19937& gen) {
    std::shared_ptr<Network> net = MakeNetwork();
    if (!mCopyCreated) mNetwork = net;
    if (net->HasFailed()) return nullptr;

    // CreateChunksOfNodes
    std::vector<std::vector<std::shared_ptr<Node>>> parts;
    {
        const auto& un = net->GetUncontractedNodes();
        const int size = int(un.size());
        int skip = size / 2, counter = 0;
        size_t pos = 0;
        if (skip == 0) {
            parts.push_back(un);
        } else {
            while (pos < un.size()) {
                std::vector<std::shared_ptr<Node>> chunk(un.begin() + long(pos), un.begin() + long(pos + size_t(skip)));
                pos += size_t(skip);
                parts.push_back(std::move(chunk));
                counter += skip;
                if (size - counter < skip) skip = size - counter;
                if (skip == 0) break;
            }
        }
    }

    // contractPieceFunction, threshold 1 (src/ContractionTools.h:243-304)
    for (auto& part : parts) {
        int failures = 0;
        while (part.size() != 1 && double(failures) < std::pow(double(part.size()), 2)) {
            std::uniform_int_distribution<> dist(0, int(part.size()) - 1);
            const int one = dist(gen);
            const int two = dist(gen);
            if (one == two) continue;
            if (totTimer.getElapsed() > maxTime) throw ContractionFailure();  // :271, :380-386
            auto result = net->ContractNodes(part[size_t(one)], part[size_t(two)], 1);
            if (result) {
                Splice(part, one, two, result);
                failures = 0;
            } else if (!net->IsDone()) {
                ++failures;
            } else {
                break;
            }
        }
    }

    // contractNodesLeftFunction, threshold -1 raised after size^2 failures (:309-356)
    std::vector<std::shared_ptr<Node>> left;
    for (auto& part : parts) left.insert(left.end(), part.begin(), part.end());
    int threshold = -1;
    int fails = 0;
    while (!net->IsDone()) {
        if (double(fails) > std::pow(double(left.size()), 2)) {
            ++threshold;
            fails = 0;
        }
        std::uniform_int_distribution<> dist(0, int(left.size()) - 1);
        const int one = dist(gen);
        const int two = dist(gen);
        if (one == two || left[size_t(one)]->mContracted || left[size_t(two)]->mContracted) continue;
        if (totTimer.getElapsed() > maxTime) throw ContractionFailure();  // :332, :380-386
        auto result = net->ContractNodes(left[size_t(one)], left[size_t(two)], threshold);
        if (result) {
            Splice(left, one, two, result);
            fails = 0;
            threshold = -1;
        } else {
            ++fails;
        }
    }
    if (!net->IsDone()) throw ContractionFailure();
    mFinalVal = net->GetFinalValue();
    return net;
}

// ContractGivenSequence (src/ContractionTools.h:392-411)
std::shared_ptr<Network> ContractionTools::ContractGivenSequence(const std::vector<std::pair<int, int>>& sequence) {
    if (!mCopyCreated) {
        mNetwork = std::make_shared<Network>(mString, mMeasureFile);
        mNetwork->SetNumThreads(mNumThreadsInNetwork);
    }
    if (mNetwork->HasFailed()) return nullptr;
    const auto& all = mNetwork->GetAllNodes();
    for (const auto& p : sequence) {
        if (p.first < 0 || p.second < 0 || size_t(p.first) >= all.size() || size_t(p.second) >= all.size())
            throw InvalidFunctionInput();
        mNetwork->ContractNodes(all[size_t(p.first)], all[size_t(p.second)], 100);
    }
    if (!mNetwork->IsDone()) throw ContractionFailure();
    mFinalVal = mNetwork->GetFinalValue();
    return mNetwork;
}

// ContractUserDefinedSequenceOfWires (src/ContractionTools.h:154-214): pairs of
// original node indices; after each contraction both indices map to the result.
std::shared_ptr<Network> ContractionTools::ContractUserDefinedSequence(const std::vector<std::pair<int, int>>& seq) {
    std::shared_ptr<Network> net = MakeNetwork();
    if (!mCopyCreated) mNetwork = net;
    if (net->HasFailed()) return nullptr;
    const int n_all = int(net->GetAllNodes().size());
    for (const auto& p : seq)
        if (p.first < 0 || p.first > n_all || p.second < 0 || p.second > n_all) throw InvalidFileFormat();
    std::vector<int> map(static_cast<size_t>(n_all));
    for (int i = 0; i < n_all; ++i) map[size_t(i)] = i;
    for (const auto& p : seq) {
        const auto& all = net->GetAllNodes();
        if (size_t(p.first) >= map.size() || size_t(p.second) >= map.size()) throw InvalidFileFormat();
        const auto a = all[size_t(map[size_t(p.first)])];
        const auto b = all[size_t(map[size_t(p.second)])];
        if (a->mContracted || b->mContracted) continue;
        net->ContractNodes(a, b, 10000);
        map[size_t(p.first)] = int(net->GetAllNodes().size()) - 1;
        map[size_t(p.second)] = int(net->GetAllNodes().size()) - 1;
    }
    if (!net->IsDone()) throw InvalidUserContractionSequence();
    mFinalVal = net->GetFinalValue();
    return net;
}

std::shared_ptr<Network> ContractionTools::ContractUserDefinedSequenceOfWires(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw InvalidFile();
    std::vector<std::pair<int, int>> seq;
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        int a, b;
        if (ss >> a >> b) seq.push_back({a, b});
    }
    return ContractUserDefinedSequence(seq);
}

// preProcess (src/preprocess.h:28-52)
bool preProcess(const std::string& qasm_text, const std::string* measure_text, qt_ctx* ctx, uint64_t seed,
                std::vector<std::pair<int, int>>& optimalContractionSequence, double timeThreshold) {
    maxTime = timeThreshold;
    for (int i = 0; i < 100; ++i) {
        totTimer = Timer();
        totTimer.start();
        NetworkOptions opt;
        opt.ctx = ctx;
        ContractionTools p(std::make_shared<Network>(qasm_text, measure_text, opt));
        p.SetSeed(seed + uint64_t(i));
        std::shared_ptr<Network> temp = p.Contract(Stochastic);
        if (totTimer.getElapsed() <= timeThreshold) {
            const auto r = RecordSequence(*temp);
            optimalContractionSequence.insert(optimalContractionSequence.end(), r.begin(), r.end());
            totTimer.reset();
            return true;
        }
    }
    return false;
}

bool preProcess(const std::string& fileName, std::vector<std::pair<int, int>>& optimalContractionSequence,
                double timeThreshold) {
    std::ifstream in(fileName);
    if (!in) throw InvalidFile();
    std::stringstream ss;
    ss << in.rdbuf();
    // the reference measures with "measureTest.txt", which it never writes: every qubit traced
    return preProcess(ss.str(), nullptr, DefaultContext(), std::random_device{}(), optimalContractionSequence,
                      timeThreshold);
}

}  // namespace qtorch
