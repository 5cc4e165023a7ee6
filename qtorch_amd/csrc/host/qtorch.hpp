// qTorch's C++ host surface (Network / Node / Wire / ContractionTools / LineGraph /
// maxcut) over the MI355X contraction engine (libqtorch_hip.so, include/qtorch_hip.h).
//
// Class and method names, argument meaning and exception types follow the
// reference (src/Network.h, src/Node.h, src/Wire.h, src/ContractionTools.h,
// src/LineGraph.h, src/maxcut.h) so that code written against qTorch compiles
// against this header.  What differs: node tensors live in device buffers
// (qt_tensor) and every pairwise contraction runs as HIP kernels; a network can
// also be built in plan-only mode, which runs the same ordering logic and records
// each step without computing values (used to build device programs and in CPU
// tests).  Stochastic ordering takes an explicit seed.
#pragma once

#include <chrono>
#include <complex>
#include <cstdint>
#include <ctime>
#include <exception>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "qtorch_hip.h"

namespace qtorch {

// ---- exceptions (src/Exceptions.h:24-59) -----------------------------------------
class InvalidFile : public std::exception {
public:
    const char* what() const noexcept override { return "Invalid Input or Output File Path"; }
};
class InvalidFileFormat : public std::exception {
public:
    const char* what() const noexcept override { return "Invalid File Format."; }
};
class ContractionFailure : public std::exception {
public:
    const char* what() const noexcept override { return "Contraction Failed."; }
};
class InvalidUserContractionSequence : public std::exception {
public:
    const char* what() const noexcept override { return "Invalid User Defined Contraction Sequence."; }
};
class InvalidFunctionInput : public std::exception {
public:
    const char* what() const noexcept override { return "Input to Function Invalid."; }
};
class InvalidContractionMethod : public std::exception {
public:
    const char* what() const noexcept override { return "Invalid Contraction Method."; }
};
class QbbFailure : public std::exception {
public:
    const char* what() const noexcept override { return "Quick BB Failure."; }
};
// Raised when the engine reports a device error (no reference counterpart: the
// reference never leaves the CPU).
class DeviceFailure : public std::exception {
public:
    explicit DeviceFailure(std::string m) : msg_(std::move(m)) {}
    const char* what() const noexcept override { return msg_.c_str(); }

private:
    std::string msg_;
};

// src/Timer.h:24-67: wall-clock seconds since start() (0 until started).
class Timer {
public:
    void start() {
        mStart = std::chrono::steady_clock::now();
        mCPUClockStart = std::clock();
        mStarted = true;
    }
    void reset() { mStarted = false; }
    double getElapsed() const {
        if (!mStarted) return 0.0;
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - mStart).count();
    }
    double getCPUElapsed() const { return double(std::clock() - mCPUClockStart) / CLOCKS_PER_SEC; }

private:
    std::chrono::steady_clock::time_point mStart{};
    std::clock_t mCPUClockStart = 0;
    bool mStarted = false;
};

// The reference's process-wide contraction timer and limit (src/Network.h:50-51).
// Once totTimer has been started and getElapsed() > maxTime, ContractNodes throws
// ContractionFailure before it changes the network, i.e. between pairwise steps
// (the reference tests them per MAC inside ContractIndices, src/Network.h:899-910,
// and before each stochastic pick, src/ContractionTools.h:271,332, and ends in
// ContractionFailure, :380-386).  The timer is only started by preProcess or by
// the caller.
extern Timer totTimer;
extern double maxTime;

enum class GateType {
    CNOT, SWAP, HADAMARD, RX, RY, RZ, X, Y, Z, PHASE, DEPOLARIZER, CRK, CZ, CPHASE,
    INITSTATE, MEASURETRACE, INTERMEDIATESTATE, ARBITRARYONEQUBITUNITARY, ARBITRARYTWOQUBITUNITARY
};

class Node;
class Network;

// src/Wire.h:26-60
class Wire {
public:
    Wire(std::shared_ptr<Node> a, std::shared_ptr<Node> b, int qubit)
        : mNodeA(a), mNodeB(b), mQubitNumber(qubit) {}
    void SetNodeA(std::shared_ptr<Node> n) { mNodeA = n; }
    void SetNodeB(std::shared_ptr<Node> n) { mNodeB = n; }
    std::weak_ptr<Node> GetNodeA() { return mNodeA; }
    std::weak_ptr<Node> GetNodeB() { return mNodeB; }
    int GetQubitNumber() const { return mQubitNumber; }
    void SetQubitNumber(int q) { mQubitNumber = q; }
    void SetWireID(int id) { mWireID = id; }
    int GetWireID() const { return mWireID; }
    void SetIsContracted(bool v) { mContracted = v; }
    bool IsContracted() const { return mContracted; }

private:
    std::weak_ptr<Node> mNodeA, mNodeB;
    int mQubitNumber;
    int mWireID = -1;
    bool mContracted = false;
};

// src/Node.h:108-166.  The tensor is a device buffer; host values exist only for
// leaf tensors until they are uploaded.
class Node {
public:
    explicit Node(int rank) : mRank(rank) {}
    virtual ~Node();
    Node(const Node&) = delete;
    Node& operator=(const Node&) = delete;

    int mRank;
    int mDimLog2 = 2;  // wire dimension 2^mDimLog2 (4: Liouville, 2: amplitudes)
    int mID = -1;
    int mIndexOfPreviousNode = -1;
    bool mContracted = false;
    std::pair<int, int> mCreatedFrom{0, 0};
    bool mSelectedInCostContractionAlgorithm = false;

    std::vector<std::shared_ptr<Wire>>& GetWires() { return mWires; }
    const std::vector<int>& GetWireNumber() const { return mWireNumbers; }
    void AddWireNumber(int n) { mWireNumbers.push_back(n); }
    void SetWireNumber(int i, int n) { mWireNumbers[size_t(i)] = n; }
    GateType GetTypeOfNode() const { return mType; }
    void SetTypeOfNode(GateType t) { mType = t; }
    const std::string& GetTypeOfNodeString() const { return mStringType; }
    void SetTypeOfNodeString(const std::string& s) { mStringType = s; }

    // Node::Index(vector) (src/Node.h:178-186) on the host copy of a leaf tensor.
    std::complex<double>& Index(const std::vector<int>& digits);
    // Values of the tensor (downloaded from the device when needed): the data the
    // reference exposes through GetTensorVals (src/Node.h:151).
    std::vector<std::complex<double>> GetTensorVals() const;
    // Node::ClearNodeData (src/Node.h:137): releases the device buffer.
    void ClearNodeData();
    bool HasData() const { return mHostVals.size() > 0 || mDev != nullptr || mPlanned; }

    // engine state
    std::vector<std::complex<double>> mHostVals;  // leaf values before upload
    qt_tensor* mDev = nullptr;
    qt_ctx* mCtx = nullptr;
    bool mPlanned = false;  // plan-only networks: a placeholder for data

protected:
    GateType mType = GateType::INTERMEDIATESTATE;
    std::string mStringType = "INTERMEDIATESTATE";
    std::vector<std::shared_ptr<Wire>> mWires;
    std::vector<int> mWireNumbers;
};

// One recorded pairwise contraction (the unit the device programs replay).
struct PlanStep {
    int a_id, b_id, c_id;  // node ids; c_id = -1 for the final scalar of a done network
    int rank_a, rank_b, k, rank_c;
    std::vector<int32_t> pair_a, pair_b;  // toSumOn
    std::vector<int32_t> free_src;        // toNotSumOn: axis | QT_FROM_A
};

struct NetworkOptions {
    qt_ctx* ctx = nullptr;     // nullptr: plan-only (structure, no values)
    // 2: qTorch's Liouville picture (dimension-4 wires, values Tr(M rho));
    // 1: amplitude picture (dimension-2 wires, value <x|U|0...0> for the measured bitstring x)
    int dim_log2 = 2;
    std::string base_dir;      // resolves relative def1/def2 gate-file paths ("" = cwd)
    bool verbose = false;      // the reference's std::cout progress messages
    // Amplitude picture only: every diagonal two-qubit gate diag(1, 1, 1, d) (CZ,
    // CPHASE, CRk) becomes two rank-3 tensors joined by a dimension-2 bond (the
    // reference has no such split): a cut through the gate then costs one index bit
    // instead of two.  Node ids after the first such gate differ from the reference's.
    bool split_diagonal = false;
    // Qubits whose output wire stays open (no measurement or trace node; no reference
    // counterpart, where every qubit ends in one): the contraction ends in a tensor
    // over these wires instead of a scalar (GetFinalTensor).
    std::vector<int> open_qubits;
};

// src/Network.h:54-140
class Network {
public:
    Network();
    // Reads the QASM and measurement files (src/Network.h:148-151) on the process
    // default device context.
    Network(const std::string& inputFile, const std::string& measureFile);
    // From in-memory text.  measure_text == nullptr behaves like a missing file.
    Network(const std::string& qasm_text, const std::string* measure_text, const NetworkOptions& opt);
    ~Network();

    std::shared_ptr<Node> ContractNodes(std::shared_ptr<Node> nodeA, std::shared_ptr<Node> nodeB,
                                        int threshold);
    const std::complex<double>& GetFinalValue() const noexcept { return mFinalVal; }
    void ContractNetworkLinearly();
    const std::vector<std::shared_ptr<Node>>& GetAllNodes() const noexcept { return mAllNodes; }
    void SetNumThreads(int n) noexcept { mNumberOfThreads = n; }
    int GetNumQubits() const noexcept { return mNumberOfQubits; }
    const std::vector<std::shared_ptr<Node>>& GetUncontractedNodes() const noexcept { return mUncontracted; }
    bool IsDone() const noexcept { return mDone; }
    void ReduceCircuit();
    bool HasFailed() const noexcept { return mFailure; }
    void Reset();
    const std::string& GetInputQasm() const noexcept { return mInputFile; }
    void resetFloatCounter() noexcept { mNumFloatOps = 0; }
    long long getNumFloatOps() const noexcept { return mNumFloatOps; }

    // engine extensions
    // Values of the tensor a network with open qubits ends in, and the qubit of each
    // of its axes (digit i of the flat index belongs to axis_qubits[i]); a finished
    // network without open qubits gives its scalar.
    std::vector<std::complex<double>> GetFinalTensor(std::vector<int>* axis_qubits) const;
    const std::vector<PlanStep>& GetPlan() const noexcept { return mPlan; }
    bool IsPlanOnly() const noexcept { return mCtx == nullptr; }
    qt_ctx* GetContext() const noexcept { return mCtx; }
    const std::vector<std::vector<std::shared_ptr<Node>>>& GetNodesByWire() const { return mNodesByWire; }
    int GetNumOriginalNodes() const noexcept { return mNumOriginal; }
    int GetDimLog2() const noexcept { return mOpt.dim_log2; }

private:
    void ParseNetwork(const std::string& qasm_text, const std::string* measure_text);
    void ParseNode(const std::string& line);
    void AttachOneQubit(const std::shared_ptr<Node>& n, int q);
    void AttachTwoQubit(const std::shared_ptr<Node>& n, int q1, int q2);
    void AttachSplitDiagonal(GateType type, const std::string& name, std::complex<double> d, int q1, int q2);
    // The arithmetic of one step on the device (no bookkeeping: it runs outside
    // mLocker); *value receives the scalar of a rank-0 result.
    void ContractIndices(const std::vector<std::pair<bool, int>>& toNotSumOn,
                         const std::vector<std::pair<int, int>>& toSumOn,
                         const std::shared_ptr<Node>& A, const std::shared_ptr<Node>& B,
                         const std::shared_ptr<Node>& C, PlanStep* step, std::complex<double>* value);
    void EnsureOnDevice(const std::shared_ptr<Node>& n);
    void UploadLeaves();
    void FindAndRemove(const std::shared_ptr<Node>& n);
    void FindAndReplace(const std::shared_ptr<Node>& from, const std::shared_ptr<Node>& to);
    static void FindAndReplace2D(std::vector<std::vector<std::shared_ptr<Node>>>& t,
                                 const std::shared_ptr<Node>& from, const std::shared_ptr<Node>& to);

    std::string mInputFile, mMeasureFile;
    std::string mQasmText, mMeasureText;
    bool mHasMeasure = false;
    NetworkOptions mOpt;
    qt_ctx* mCtx = nullptr;
    std::complex<double> mFinalVal{0.0, 0.0};
    int mNumberOfQubits = 0;
    bool mDone = false;
    bool mFailure = false;
    int mNumOriginal = 0;
    std::vector<std::shared_ptr<Node>> mAllNodes;
    std::vector<std::shared_ptr<Node>> mUncontracted;
    std::vector<std::vector<std::shared_ptr<Node>>> mNodesByWire;
    std::vector<std::shared_ptr<Wire>> mHanging;
    std::unordered_map<std::string, std::string> mOneQubitGates, mTwoQubitGates;
    long long mNumFloatOps = 0;
    int mNumberOfThreads = 8;
    std::vector<PlanStep> mPlan;
    std::shared_ptr<Node> mFinalTensor;  // open qubits: the tensor the network ends in
    // guards the network's bookkeeping (wires, node lists, ids, plan, final value)
    // when several threads contract disjoint pairs (src/Network.h:718-862)
    std::mutex mLocker;
};

enum ContractionType { Stochastic, FromEdges, CostContractSimple, CostContractBruteForce };

// src/ContractionTools.h:56-119
class ContractionTools {
public:
    ContractionTools(const std::string& inputFile, const std::string& measureFile, int numThreads = 8);
    explicit ContractionTools(std::shared_ptr<Network> network);
    std::complex<double> GetFinalVal() const noexcept { return mFinalVal; }
    // Only Stochastic is provided (the reference's guide recommends it; the cost
    // based orderings are outside this engine) — others throw InvalidContractionMethod.
    std::shared_ptr<Network> Contract(ContractionType type, int pValue = 1, int numSamples = 1);
    std::shared_ptr<Network> ContractUserDefinedSequenceOfWires(const std::string& userInputFilePath);
    std::shared_ptr<Network> ContractUserDefinedSequence(const std::vector<std::pair<int, int>>& seq);
    std::shared_ptr<Network> ContractGivenSequence(const std::vector<std::pair<int, int>>& sequence);
    void Reset(const std::string& inputFile, const std::string& measureFile, int numThreads = 8);
    void Reset();
    void Reset(std::shared_ptr<Network> network) {
        mNetwork = network;
        mCopyCreated = true;
    }
    // Seeds the stochastic ordering (the reference seeds from std::random_device).
    void SetSeed(uint64_t seed) { mRandGen.seed(static_cast<std::mt19937::result_type>(seed)); }

private:
    std::shared_ptr<Network> MakeNetwork();
    std::shared_ptr<Network> ParallelContract(std::mt19937& gen);

    std::string mString, mMeasureFile;
    std::complex<double> mFinalVal{0.0, 0.0};
    std::shared_ptr<Network> mNetwork;
    std::mt19937 mRandGen;
    bool mCopyCreated = false;
    int mNumThreadsInNetwork = 8;
};

// The mCreatedFrom record of a contracted network (src/preprocess.h:41-46).
std::vector<std::pair<int, int>> RecordSequence(const Network& net);

// preProcess (src/preprocess.h:28-52): up to 100 stochastic contractions of the
// circuit, each timed from a fresh start of totTimer with maxTime = timeThreshold;
// the first that finishes within timeThreshold seconds gives its mCreatedFrom record
// in optimalContractionSequence and true (totTimer is then reset), else false.  A
// contraction that runs out of time throws ContractionFailure, as the reference's
// does.  maxTime keeps timeThreshold afterwards, as in the reference.
bool preProcess(const std::string& fileName, std::vector<std::pair<int, int>>& optimalContractionSequence,
                double timeThreshold);
// The same on in-memory text and a context (nullptr: plan-only), with seeded orders
// (attempt i uses seed + i); measure_text nullptr traces every qubit, as the
// reference's missing measureTest.txt does.
bool preProcess(const std::string& qasm_text, const std::string* measure_text, qt_ctx* ctx, uint64_t seed,
                std::vector<std::pair<int, int>>& optimalContractionSequence, double timeThreshold);

// src/LineGraph.h:44-89
class LineGraph {
public:
    explicit LineGraph(std::shared_ptr<Network> net);
    void Reset(std::shared_ptr<Network> net = nullptr);
    // Writes lg.cnf (src/LineGraph.h:251-261) and, in place of the external QuickBB
    // binary, a min-fill elimination order of the line graph in qbb.out's format.
    bool runQuickBB(int maxTimeInSec, void* timer = nullptr, bool sixtyFourBit = true);
    // Contracts in the order read from qbb.out (src/LineGraph.h:296-410).
    bool LGContract();
    void SetQBBOutDirectory(const std::string& dir) {
        cnfName = dir + "lg.cnf";
        qbbOutName = dir + "qbb.out";
        qbbStatsName = dir + "qbb-stats.out";
    }
    void SetQBBOutFiles(const std::string& cnf, const std::string& qbb, const std::string& stats) {
        cnfName = cnf;
        qbbOutName = qbb;
        qbbStatsName = stats;
    }

    // engine extensions
    std::string CnfText() const;
    std::vector<int> MinFillOrder() const;  // 1-based wire ids
    std::complex<double> ContractInOrder(const std::vector<int>& order_1based);
    int NumWires() const { return int(mGraphWires.size()); }
    std::complex<double> GetResult() const { return mResult; }

private:
    void Build();
    std::shared_ptr<Network> mNet;
    std::vector<std::shared_ptr<Wire>> mGraphWires;
    std::vector<std::pair<std::shared_ptr<Wire>, std::shared_ptr<Wire>>> mEdges;
    std::complex<double> mResult{0.0, 0.0};
    std::string cnfName = "output/lg.cnf";
    std::string qbbOutName = "output/qbb.out";
    std::string qbbStatsName = "output/qbb-stats.out";
};

// ---- QAOA MaxCut (src/maxcut.h:31-233, src/maxcut.cpp:145-204) --------------------
struct ExtraData {
    ExtraData() = default;
    ExtraData(int p0, const char* filename);
    // From .dgf text; reference_qubit_map reproduces the reference's light-cone map
    // (only pair.first -> 0 fixed, src/maxcut.h:119-120); false also fixes second -> 1.
    static ExtraData FromText(int p, const std::string& dgf_text, bool reference_qubit_map = true);

    std::string fileName;
    std::vector<std::pair<int, int>> pairs;
    std::vector<std::vector<int>> adjacencyLists;
    std::vector<int> qubitsNeeded;
    std::string outputFile;
    int numQubits = 0;
    int p = 1;
    bool referenceQubitMap = true;
    std::vector<std::vector<std::pair<int, int>>> iterations;
    std::vector<std::vector<std::pair<int, int>>> realIterations;
    void ReadInData();
    void ParseData(const std::string& text);
    void PopulateIterations();
};

// QASM text F_p writes for edge `e` (src/maxcut.cpp:172-178): H on every qubit,
// then p x [CNOT, Rz -gamma, CNOT per light-cone edge; Rx 2 beta per qubit] with
// the angles printed at the default ostream precision.
std::string QaoaEdgeQasm(const ExtraData& d, int e, const std::vector<double>& betas_gammas);
// Measurement text of F_p (src/maxcut.cpp:179-188): Z on qubits 0 and 1, T elsewhere.
std::string QaoaEdgeMeasure(const ExtraData& d, int e);

// F_p (src/maxcut.cpp:162-204) for one set of angles, on the device, each edge in
// its own network contracted stochastically (seed + e).  Returns the energy and
// fills per-edge <ZZ> values.
double QaoaEnergyPlain(qt_ctx* ctx, const ExtraData& d, const std::vector<double>& betas_gammas,
                       uint64_t seed, std::vector<double>* zz);

// Process default context (device 0), created on first use.
qt_ctx* DefaultContext();

}  // namespace qtorch
