// Network, Node and gate tensors over the device engine.
// Semantics follow qTorch's src/Network.h and src/Node.h (cited per function).
#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

#include "qtorch.hpp"

namespace qtorch {

Timer totTimer;
double maxTime = 60.0;  // src/Network.h:51

namespace {

constexpr double kPi = 3.14159265358979323846;  // src/Node.h:45
using cd = std::complex<double>;

std::shared_ptr<Node> Leaf(int rank, GateType type, const std::string& name, int dim_log2 = 2) {
    auto n = std::make_shared<Node>(rank);
    n->mDimLog2 = dim_log2;
    n->SetTypeOfNode(type);
    n->SetTypeOfNodeString(name);
    n->mHostVals.assign(size_t(1) << (dim_log2 * rank), cd(0.0, 0.0));
    return n;
}

// ---- amplitude picture (dimension-2 wires) -------------------------------------------
// The same gates as unitaries U acting on kets, in the conventions the superoperator
// tables encode: Rx(t) = exp(-i t X / 2), Ry(t) = [[c, -s], [s, c]], Rz(t) = PHASE(t) =
// diag(e^{-it/2}, e^{it/2}), CPHASE(t) = diag(1, 1, 1, e^{it}), CRk c t = CPHASE(2 pi /
// 2^(c+1)); two-qubit matrices act on the basis 2 q1 + q2.  A one-qubit tensor over
// wires [in, out] holds U[out][in] at in + 2 out; a two-qubit one over
// [in1, in2, out1, out2] holds U[2 o1 + o2][2 i1 + i2] at i1 + 2 i2 + 4 o1 + 8 o2.
std::shared_ptr<Node> AmpOne(GateType type, const std::string& name, const std::array<cd, 4>& u) {
    auto n = Leaf(2, type, name, 1);
    for (int i = 0; i < 2; ++i)
        for (int o = 0; o < 2; ++o) n->mHostVals[size_t(i + 2 * o)] = u[size_t(2 * o + i)];
    return n;
}

std::shared_ptr<Node> AmpTwo(GateType type, const std::string& name, const std::array<cd, 16>& u) {
    auto n = Leaf(4, type, name, 1);
    for (int i1 = 0; i1 < 2; ++i1)
        for (int i2 = 0; i2 < 2; ++i2)
            for (int o1 = 0; o1 < 2; ++o1)
                for (int o2 = 0; o2 < 2; ++o2)
                    n->mHostVals[size_t(i1 + 2 * i2 + 4 * o1 + 8 * o2)] = u[size_t(4 * (2 * o1 + o2) + 2 * i1 + i2)];
    return n;
}

std::shared_ptr<Node> AmpRotation(const std::string& g, double t) {
    const double c = std::cos(t / 2.0), s = std::sin(t / 2.0);
    if (g == "Rx" || g == "RX") return AmpOne(GateType::RX, "Rx", {cd(c, 0), cd(0, -s), cd(0, -s), cd(c, 0)});
    if (g == "Ry" || g == "RY") return AmpOne(GateType::RY, "Ry", {cd(c, 0), cd(-s, 0), cd(s, 0), cd(c, 0)});
    const GateType ty = g == "PHASE" ? GateType::PHASE : GateType::RZ;
    return AmpOne(ty, g == "PHASE" ? "Phase" : "Rz", {cd(c, -s), cd(0), cd(0), cd(c, s)});
}

std::shared_ptr<Node> AmpFixed(const std::string& g) {
    const double h = 1.0 / std::sqrt(2.0);
    if (g == "H") return AmpOne(GateType::HADAMARD, "H", {cd(h), cd(h), cd(h), cd(-h)});
    if (g == "X") return AmpOne(GateType::X, "X", {cd(0), cd(1), cd(1), cd(0)});
    if (g == "Y") return AmpOne(GateType::Y, "Y", {cd(0), cd(0, -1), cd(0, 1), cd(0)});
    return AmpOne(GateType::Z, "Z", {cd(1), cd(0), cd(0), cd(-1)});
}

std::array<cd, 16> Diag4(cd d3) {
    std::array<cd, 16> u{};
    u[0] = u[5] = u[10] = 1.0;
    u[15] = d3;
    return u;
}

std::shared_ptr<Node> AmpTwoQubitGate(const std::string& g, double angle, int control) {
    if (g == "CNOT") {
        std::array<cd, 16> u{};
        u[0] = u[5] = u[4 * 2 + 3] = u[4 * 3 + 2] = 1.0;
        return AmpTwo(GateType::CNOT, "CNOT", u);
    }
    if (g == "SWAP") {
        std::array<cd, 16> u{};
        u[0] = u[4 * 1 + 2] = u[4 * 2 + 1] = u[15] = 1.0;
        return AmpTwo(GateType::SWAP, "SWAP", u);
    }
    if (g == "CZ") return AmpTwo(GateType::CZ, "CZ", Diag4(cd(-1.0)));
    if (g == "CRk") {
        const double th = 2.0 * kPi / std::pow(2, control + 1.0);
        return AmpTwo(GateType::CRK, "CRk", Diag4(cd(std::cos(th), std::sin(th))));
    }
    return AmpTwo(GateType::CPHASE, "CPhase", Diag4(cd(std::cos(angle), std::sin(angle))));
}

// ---- superoperator tables (src/Node.h:197-898) ----------------------------------
// Digit d = 2*ket + bra of a wire's density-matrix element; gate wires are
// [inputs..., outputs...].  Each table lists the same entries, computed with the
// same expressions, as the reference, so the doubles are bit-identical.

std::shared_ptr<Node> CnotLeaf() {  // src/Node.h:197-220: the control's ket/bra bits flip the target's
    auto n = Leaf(4, GateType::CNOT, "CNOT");
    for (int c = 0; c < 4; ++c)
        for (int t = 0; t < 4; ++t) n->Index({c, t, c, t ^ c}) = 1.0;
    return n;
}

std::shared_ptr<Node> SwapLeaf() {  // src/Node.h:515-542
    auto n = Leaf(4, GateType::SWAP, "SWAP");
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) n->Index({a, b, b, a}) = 1.0;
    return n;
}

// Diagonal two-qubit phase gates: entry (a,b,a,b) depends on whether the |11>
// component is on the ket side only, the bra side only, both or neither.
std::shared_ptr<Node> DiagTwoQubit(GateType type, const std::string& name, cd ket_only, cd bra_only,
                                   cd both) {
    auto n = Leaf(4, type, name);
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            const bool ket = (a >> 1) & (b >> 1), bra = a & b & 1;
            n->Index({a, b, a, b}) = ket && bra ? both : ket ? ket_only : bra ? bra_only : cd(1.0);
        }
    return n;
}

std::shared_ptr<Node> CzLeaf() {  // src/Node.h:449-473
    return DiagTwoQubit(GateType::CZ, "CZ", cd(-1.0), cd(-1.0), cd(1.0));
}

std::shared_ptr<Node> CPhaseLeaf(double t) {  // src/Node.h:475-498
    return DiagTwoQubit(GateType::CPHASE, "CPhase", cd(std::cos(t), std::sin(t)),
                        cd(std::cos(t), -1.0 * std::sin(t)), cd(1.0));
}

std::shared_ptr<Node> CrkLeaf(int control) {  // src/Node.h:424-447: phase 2 pi / 2^(control+1)
    const cd em = std::exp(2.0 * kPi * cd(0, -1) / std::pow(2, control + 1.0));
    const cd ep = std::exp(2.0 * kPi * cd(0, 1) / std::pow(2, control + 1.0));
    return DiagTwoQubit(GateType::CRK, "CRk", ep, em, cd(1.0));
}

// One-qubit gates from a 4x4 symbol table over (in, out).
enum Sym : int8_t { Z0, ONE, NONE, C2, S2, NS2, H, NH, IH, NIH, HALF, NHALF, EM, EP };

std::shared_ptr<Node> OneQubitTable(GateType type, const std::string& name, const int8_t (&tab)[4][4],
                                    double t) {
    auto n = Leaf(2, type, name);
    const double c2 = std::pow(std::cos(t / 2.0), 2);
    const double s2 = std::pow(std::sin(t / 2.0), 2);
    const double h = std::sin(t) / 2.0;
    for (int i = 0; i < 4; ++i)
        for (int o = 0; o < 4; ++o) {
            cd v(0.0);
            switch (tab[i][o]) {
                case ONE: v = 1.0; break;
                case NONE: v = -1.0; break;
                case C2: v = c2; break;
                case S2: v = s2; break;
                case NS2: v = -1.0 * std::pow(std::sin(t / 2), 2); break;
                case H: v = h; break;
                case NH: v = -1.0 * std::sin(t) / 2.0; break;
                case IH: v = cd(0.0, h); break;
                case NIH: v = cd(0.0, -1.0 * (std::sin(t) / 2.0)); break;
                case HALF: v = 1.0 / 2.0; break;
                case NHALF: v = -1.0 / 2.0; break;
                case EM: v = cd(std::cos(t), -1.0 * std::sin(t)); break;
                case EP: v = cd(std::cos(t), std::sin(t)); break;
                default: break;
            }
            n->Index({i, o}) = v;
        }
    return n;
}

constexpr int8_t kRx[4][4] = {{C2, IH, NIH, S2}, {IH, C2, S2, NIH}, {NIH, S2, C2, IH}, {S2, NIH, IH, C2}};
constexpr int8_t kRy[4][4] = {{C2, H, H, S2}, {NH, C2, NS2, H}, {NH, NS2, C2, H}, {S2, NH, NH, C2}};
constexpr int8_t kPhase[4][4] = {{ONE, Z0, Z0, Z0}, {Z0, EM, Z0, Z0}, {Z0, Z0, EP, Z0}, {Z0, Z0, Z0, ONE}};
constexpr int8_t kH[4][4] = {{HALF, HALF, HALF, HALF}, {HALF, NHALF, HALF, NHALF},
                             {HALF, HALF, NHALF, NHALF}, {HALF, NHALF, NHALF, HALF}};
constexpr int8_t kX[4][4] = {{Z0, Z0, Z0, ONE}, {Z0, Z0, ONE, Z0}, {Z0, ONE, Z0, Z0}, {ONE, Z0, Z0, Z0}};
constexpr int8_t kY[4][4] = {{Z0, Z0, Z0, ONE}, {Z0, Z0, NONE, Z0}, {Z0, NONE, Z0, Z0}, {ONE, Z0, Z0, Z0}};
constexpr int8_t kZ[4][4] = {{ONE, Z0, Z0, Z0}, {Z0, NONE, Z0, Z0}, {Z0, Z0, NONE, Z0}, {Z0, Z0, Z0, ONE}};

// Rank-1 state / measurement vectors (src/Node.h:356-421) by digit 0..3.
std::shared_ptr<Node> Vec4(GateType type, const std::string& name, cd d0, cd d1, cd d2, cd d3) {
    auto n = Leaf(1, type, name);
    n->mHostVals = {d0, d1, d2, d3};
    return n;
}

std::shared_ptr<Node> MeasureLeaf(char ch) {  // src/Network.h:199-228
    switch (ch) {
        case 'X': return Vec4(GateType::MEASURETRACE, "X measure", 0.0, 1.0, 1.0, 0.0);
        case 'Y': return Vec4(GateType::MEASURETRACE, "Y measure", 0.0, cd(0, 1.0), cd(0, -1.0), 0.0);
        case 'Z': return Vec4(GateType::MEASURETRACE, "Z measure", 1.0, 0.0, 0.0, -1.0);
        case '0': return Vec4(GateType::MEASURETRACE, "|0><0| measure", 1.0, 0.0, 0.0, 0.0);
        case '1': return Vec4(GateType::MEASURETRACE, "|1><1| measure", 0.0, 0.0, 0.0, 1.0);
        default: return Vec4(GateType::MEASURETRACE, "Trace", 1.0, 0.0, 0.0, 1.0);
    }
}

std::vector<cd> ReadComplexes(const std::string& path, int count) {
    std::ifstream in(path);
    if (!in.is_open()) throw InvalidFile();
    std::vector<cd> v(static_cast<size_t>(count));
    for (auto& x : v) {
        if (in.eof()) throw InvalidFileFormat();
        in >> x;
    }
    return v;
}

// src/Node.h:544-586: entry (in, out) = U[ket_out][ket_in] * conj(U[bra_out][bra_in])
std::shared_ptr<Node> ArbitraryOneQubit(const std::string& path, const std::string& name) {
    const auto u = ReadComplexes(path, 4);
    auto n = Leaf(2, GateType::ARBITRARYONEQUBITUNITARY, name);
    for (int i = 0; i < 4; ++i)
        for (int o = 0; o < 4; ++o)
            n->Index({i, o}) = u[size_t(2 * (o >> 1) + (i >> 1))] * std::conj(u[size_t(2 * (o & 1) + (i & 1))]);
    return n;
}

// src/Node.h:588-898: U is 4x4 row-major over 2*q1 + q2; wires [in q1, in q2, out q1, out q2]
std::shared_ptr<Node> ArbitraryTwoQubit(const std::string& path, const std::string& name) {
    const auto u = ReadComplexes(path, 16);
    auto n = Leaf(4, GateType::ARBITRARYTWOQUBITUNITARY, name);
    for (int d0 = 0; d0 < 4; ++d0)
        for (int d1 = 0; d1 < 4; ++d1)
            for (int d2 = 0; d2 < 4; ++d2)
                for (int d3 = 0; d3 < 4; ++d3) {
                    const int ki = 2 * (d0 >> 1) + (d1 >> 1), bi = 2 * (d0 & 1) + (d1 & 1);
                    const int ko = 2 * (d2 >> 1) + (d3 >> 1), bo = 2 * (d2 & 1) + (d3 & 1);
                    n->Index({d0, d1, d2, d3}) = u[size_t(4 * ko + ki)] * std::conj(u[size_t(4 * bo + bi)]);
                }
    return n;
}

// def1 / def2 gates in the amplitude picture: U itself (row-major in the file)
std::shared_ptr<Node> AmpArbitrary(const std::string& path, const std::string& name, int qubits) {
    const auto u = ReadComplexes(path, qubits == 1 ? 4 : 16);
    if (qubits == 1)
        return AmpOne(GateType::ARBITRARYONEQUBITUNITARY, name, {u[0], u[1], u[2], u[3]});
    std::array<cd, 16> m;
    for (size_t i = 0; i < 16; ++i) m[i] = u[i];
    return AmpTwo(GateType::ARBITRARYTWOQUBITUNITARY, name, m);
}

// Network::ParseTokens (src/Network.h:976-992): '#' only at column 0; split on single
// spaces; an empty final piece is dropped.
std::vector<std::string> Tokens(const std::string& line) {
    std::vector<std::string> out;
    if (line.empty() || line.find('#') == 0) return out;
    size_t start = 0;
    while (true) {
        const size_t sp = line.find(' ', start);
        if (sp == std::string::npos) {
            if (start < line.size()) out.push_back(line.substr(start));
            break;
        }
        out.push_back(line.substr(start, sp - start));
        start = sp + 1;
    }
    return out;
}

void Check(qt_ctx* ctx, qt_status s, const char* what) {
    if (s == QT_OK) return;
    if (s == QT_INVALID_INPUT) throw InvalidFunctionInput();
    if (s == QT_TIMEOUT) throw ContractionFailure();
    throw DeviceFailure(std::string(what) + ": " + qt_last_error(ctx));
}

}  // namespace

Node::~Node() {
    if (mDev && mCtx) qt_tensor_free(mCtx, mDev);
}

std::complex<double>& Node::Index(const std::vector<int>& digits) {
    size_t idx = 0, mult = 1;
    for (int d : digits) {
        idx += size_t(d) * mult;
        mult <<= mDimLog2;
    }
    return mHostVals[idx];
}

std::vector<std::complex<double>> Node::GetTensorVals() const {
    if (mDev && mCtx) {
        std::vector<cd> v(size_t(qt_tensor_numel(mDev)));
        Check(mCtx, qt_tensor_download(mCtx, mDev, v.data(), int64_t(v.size() * sizeof(cd))), "download");
        return v;
    }
    return mHostVals;
}

void Node::ClearNodeData() {
    mHostVals = std::vector<cd>();
    if (mDev && mCtx) qt_tensor_free(mCtx, mDev);
    mDev = nullptr;
    mPlanned = false;
}

qt_ctx* DefaultContext() {
    static qt_ctx* ctx = [] {
        qt_ctx* c = nullptr;
        if (qt_ctx_create(0, &c) != QT_OK) c = nullptr;
        return c;
    }();
    if (!ctx) throw DeviceFailure("qtorch: no HIP device available (the contraction engine runs on the GPU)");
    return ctx;
}

// ---- Network ------------------------------------------------------------------------

Network::Network() = default;

Network::Network(const std::string& inputFile, const std::string& measureFile)
    : mInputFile(inputFile), mMeasureFile(measureFile) {
    std::ifstream q(inputFile);
    if (!q.is_open()) {
        mFailure = true;
        throw InvalidFile();  // src/Network.h:249-253
    }
    std::stringstream qs;
    qs << q.rdbuf();
    mQasmText = qs.str();
    std::ifstream m(measureFile);
    mHasMeasure = m.is_open();
    if (mHasMeasure) {
        std::stringstream ms;
        ms << m.rdbuf();
        mMeasureText = ms.str();
    }
    mOpt.ctx = DefaultContext();
    mCtx = mOpt.ctx;
    ParseNetwork(mQasmText, mHasMeasure ? &mMeasureText : nullptr);
}

Network::Network(const std::string& qasm_text, const std::string* measure_text, const NetworkOptions& opt)
    : mQasmText(qasm_text), mOpt(opt), mCtx(opt.ctx) {
    mHasMeasure = measure_text != nullptr;
    if (mHasMeasure) mMeasureText = *measure_text;
    ParseNetwork(mQasmText, mHasMeasure ? &mMeasureText : nullptr);
}

Network::~Network() = default;

void Network::Reset() {  // src/Network.h:154-171
    mFinalVal = cd(0.0);
    mNumberOfQubits = 0;
    mDone = false;
    mFailure = false;
    mAllNodes.clear();
    mUncontracted.clear();
    mNodesByWire.clear();
    mHanging.clear();
    mOneQubitGates.clear();
    mTwoQubitGates.clear();
    mPlan.clear();
    mFinalTensor.reset();
    ParseNetwork(mQasmText, mHasMeasure ? &mMeasureText : nullptr);
}

// ParseNetwork (src/Network.h:246-304) with CreateInitialStates (:177-194) and
// AddMeasurementsOrTrace (:199-240).
void Network::ParseNetwork(const std::string& qasm, const std::string* measure) {
    std::vector<std::string> lines;
    {
        size_t start = 0;
        while (true) {
            const size_t nl = qasm.find('\n', start);
            if (nl == std::string::npos) {
                lines.push_back(qasm.substr(start));
                break;
            }
            lines.push_back(qasm.substr(start, nl - start));
            start = nl + 1;
        }
    }
    mNumberOfQubits = std::stoi(lines[0]);
    if (mNumberOfQubits < 0) throw InvalidFileFormat();
    mNodesByWire.assign(size_t(mNumberOfQubits), {});
    for (int i = 0; i < mNumberOfQubits; ++i) {
        std::shared_ptr<Node> n;
        if (mOpt.dim_log2 == 2) {
            n = Vec4(GateType::INITSTATE, "|0><0|", 1.0, 0.0, 0.0, 0.0);
        } else {
            n = Leaf(1, GateType::INITSTATE, "|0>", 1);
            n->mHostVals = {1.0, 0.0};
        }
        if (mOpt.verbose) std::cout << "Creating qubit " << i << " in the initial state: |0><0|" << std::endl;
        auto w = std::make_shared<Wire>(n, nullptr, i);
        n->GetWires().push_back(w);
        mHanging.push_back(w);
        mAllNodes.push_back(n);
        n->AddWireNumber(i);
        mNodesByWire[size_t(i)].push_back(n);
        n->mID = int(mAllNodes.size()) - 1;
    }
    for (size_t i = 1; i < lines.size(); ++i) ParseNode(lines[i]);
    std::vector<char> meas(size_t(mNumberOfQubits), 'T');
    if (measure) {
        size_t q = 0;
        for (char ch : *measure) {
            if (q >= meas.size()) break;
            if (std::isspace(static_cast<unsigned char>(ch))) continue;
            meas[q++] = ch;
        }
    }
    std::vector<char> open(size_t(mNumberOfQubits), 0);
    for (int q : mOpt.open_qubits) {
        if (q < 0 || q >= mNumberOfQubits || open[size_t(q)]) throw InvalidFunctionInput();
        open[size_t(q)] = 1;
    }
    for (int i = 0; i < mNumberOfQubits; ++i) {
        if (open[size_t(i)]) continue;  // the hanging wire stays an open axis
        std::shared_ptr<Node> n;
        if (mOpt.dim_log2 == 2) {
            n = MeasureLeaf(meas[size_t(i)]);
        } else {
            // amplitude picture: the measurement string is the output bitstring
            const char ch = meas[size_t(i)];
            if (ch != '0' && ch != '1') throw InvalidFileFormat();
            n = Leaf(1, GateType::MEASURETRACE, ch == '0' ? "<0|" : "<1|", 1);
            n->mHostVals = {ch == '0' ? 1.0 : 0.0, ch == '1' ? 1.0 : 0.0};
        }
        mHanging[size_t(i)]->SetNodeB(n);
        n->GetWires().push_back(mHanging[size_t(i)]);
        mAllNodes.push_back(n);
        n->AddWireNumber(i);
        mNodesByWire[size_t(i)].push_back(n);
        n->mID = int(mAllNodes.size()) - 1;
    }
    mHanging.clear();
    mUncontracted = mAllNodes;
    mNumOriginal = int(mAllNodes.size());
    if (mCtx) UploadLeaves();
}

void Network::AttachOneQubit(const std::shared_ptr<Node>& n, int q) {
    if (q > mNumberOfQubits - 1 || q < 0) throw InvalidFileFormat();
    n->GetWires().push_back(mHanging[size_t(q)]);
    mHanging[size_t(q)]->SetNodeB(n);
    auto w = std::make_shared<Wire>(n, nullptr, q);
    mHanging[size_t(q)] = w;
    n->GetWires().push_back(w);
    n->AddWireNumber(q);
    n->mIndexOfPreviousNode = int(mNodesByWire[size_t(q)].size()) - 1;
    mNodesByWire[size_t(q)].push_back(n);
}

void Network::AttachTwoQubit(const std::shared_ptr<Node>& n, int q1, int q2) {
    if (q1 > mNumberOfQubits - 1 || q2 > mNumberOfQubits - 1 || q1 == q2 || q1 < 0 || q2 < 0)
        throw InvalidFileFormat();
    n->GetWires().push_back(mHanging[size_t(q1)]);
    mHanging[size_t(q1)]->SetNodeB(n);
    n->GetWires().push_back(mHanging[size_t(q2)]);
    mHanging[size_t(q2)]->SetNodeB(n);
    auto w1 = std::make_shared<Wire>(n, nullptr, q1);
    mHanging[size_t(q1)] = w1;
    n->GetWires().push_back(w1);
    auto w2 = std::make_shared<Wire>(n, nullptr, q2);
    mHanging[size_t(q2)] = w2;
    n->GetWires().push_back(w2);
    n->AddWireNumber(q1);
    n->AddWireNumber(q2);
    mNodesByWire[size_t(q1)].push_back(n);
    mNodesByWire[size_t(q2)].push_back(n);
}

// diag(1, 1, 1, d) on (q1, q2) = sum_k |k><k| (x) diag(1, d^k): on q1 a rank-3 tensor
// [in, out, k] = delta(in, out) delta(in, k), on q2 [in, out, k] = delta(in, out) d^(in k);
// the bond k joins them.  Both are leaves, q1's first.
void Network::AttachSplitDiagonal(GateType type, const std::string& name, cd d, int q1, int q2) {
    if (q1 > mNumberOfQubits - 1 || q2 > mNumberOfQubits - 1 || q1 == q2 || q1 < 0 || q2 < 0)
        throw InvalidFileFormat();
    auto a = Leaf(3, type, name + " (control half)", 1);
    auto b = Leaf(3, type, name + " (target half)", 1);
    a->mHostVals[0 + 2 * 0 + 4 * 0] = 1.0;
    a->mHostVals[1 + 2 * 1 + 4 * 1] = 1.0;
    for (int x = 0; x < 2; ++x)
        for (int k = 0; k < 2; ++k) b->mHostVals[size_t(x + 2 * x + 4 * k)] = (x && k) ? d : cd(1.0);
    auto bond = std::make_shared<Wire>(a, b, q1);
    for (auto [n, q] : {std::make_pair(a, q1), std::make_pair(b, q2)}) {
        n->GetWires().push_back(mHanging[size_t(q)]);
        mHanging[size_t(q)]->SetNodeB(n);
        auto w = std::make_shared<Wire>(n, nullptr, q);
        mHanging[size_t(q)] = w;
        n->GetWires().push_back(w);
        n->GetWires().push_back(bond);
        n->AddWireNumber(q);
        n->mIndexOfPreviousNode = int(mNodesByWire[size_t(q)].size()) - 1;
        mNodesByWire[size_t(q)].push_back(n);
        mAllNodes.push_back(n);
        n->mID = int(mAllNodes.size()) - 1;
    }
}

// ParseNode (src/Network.h:326-707)
void Network::ParseNode(const std::string& line) {
    const auto t = Tokens(line);
    if (t.empty()) return;
    const std::string& g = t[0];
    std::shared_ptr<Node> n;
    auto arg = [&](size_t i) -> const std::string& {
        if (i >= t.size()) throw InvalidFileFormat();
        return t[i];
    };
    if (g == "Rx" || g == "RX" || g == "Ry" || g == "RY" || g == "Rz" || g == "RZ" || g == "PHASE") {
        const float angle = std::stof(arg(1));  // FP32 parse, src/Network.h:342,383,402,422
        if (mOpt.dim_log2 == 1) n = AmpRotation(g, angle);
        else if (g == "Rx" || g == "RX") n = OneQubitTable(GateType::RX, "Rx", kRx, angle);
        else if (g == "Ry" || g == "RY") n = OneQubitTable(GateType::RY, "Ry", kRy, angle);
        else if (g == "Rz" || g == "RZ") n = OneQubitTable(GateType::RZ, "Rz", kPhase, angle);
        else n = OneQubitTable(GateType::PHASE, "Phase", kPhase, angle);
        AttachOneQubit(n, std::stoi(arg(2)));
    } else if (g == "H" || g == "X" || g == "Y" || g == "Z") {
        if (mOpt.dim_log2 == 1) n = AmpFixed(g);
        else if (g == "H") n = OneQubitTable(GateType::HADAMARD, "H", kH, 0.0);
        else if (g == "X") n = OneQubitTable(GateType::X, "X", kX, 0.0);
        else if (g == "Y") n = OneQubitTable(GateType::Y, "Y", kY, 0.0);
        else n = OneQubitTable(GateType::Z, "Z", kZ, 0.0);
        AttachOneQubit(n, std::stoi(arg(1)));
    } else if (g == "CNOT" || g == "SWAP" || g == "CZ" || g == "CRk") {
        const int q1 = std::stoi(arg(1)), q2 = std::stoi(arg(2));
        if (mOpt.dim_log2 == 1 && mOpt.split_diagonal && (g == "CZ" || g == "CRk")) {
            if (g == "CZ") {
                AttachSplitDiagonal(GateType::CZ, "CZ", cd(-1.0), q1, q2);
            } else {
                if (q1 > mNumberOfQubits - 1 || q2 > mNumberOfQubits - 1 || q1 == q2 || q1 < 0) throw InvalidFileFormat();
                const double th = 2.0 * kPi / std::pow(2, q1 + 1.0);  // as AmpTwoQubitGate
                AttachSplitDiagonal(GateType::CRK, "CRk", cd(std::cos(th), std::sin(th)), q1, q2);
            }
            return;
        }
        if (mOpt.dim_log2 == 1) {
            if (g == "CRk" && (q1 > mNumberOfQubits - 1 || q2 > mNumberOfQubits - 1 || q1 == q2))
                throw InvalidFileFormat();
            n = AmpTwoQubitGate(g, 0.0, q1);
        } else if (g == "CNOT") n = CnotLeaf();
        else if (g == "SWAP") n = SwapLeaf();
        else if (g == "CZ") n = CzLeaf();
        else {
            if (q1 > mNumberOfQubits - 1 || q2 > mNumberOfQubits - 1 || q1 == q2) throw InvalidFileFormat();
            n = CrkLeaf(q1);
        }
        AttachTwoQubit(n, q1, q2);
    } else if (g == "CPHASE") {
        const double angle = std::stod(arg(1));  // src/Network.h:615
        if (mOpt.dim_log2 == 1 && mOpt.split_diagonal) {
            AttachSplitDiagonal(GateType::CPHASE, "CPhase", cd(std::cos(angle), std::sin(angle)), std::stoi(arg(2)),
                                std::stoi(arg(3)));
            return;
        }
        n = mOpt.dim_log2 == 1 ? AmpTwoQubitGate("CPHASE", angle, 0) : CPhaseLeaf(angle);
        AttachTwoQubit(n, std::stoi(arg(2)), std::stoi(arg(3)));
    } else if (g == "def1") {
        mOneQubitGates.insert({arg(1), arg(2)});
        return;
    } else if (g == "def2") {
        mTwoQubitGates.insert({arg(1), arg(2)});
        return;
    } else if (mOneQubitGates.count(g)) {
        std::string path = mOneQubitGates[g];
        if (!mOpt.base_dir.empty() && !path.empty() && path[0] != '/') path = mOpt.base_dir + "/" + path;
        n = mOpt.dim_log2 == 1 ? AmpArbitrary(path, g, 1) : ArbitraryOneQubit(path, g);
        AttachOneQubit(n, std::stoi(arg(1)));
    } else if (mTwoQubitGates.count(g)) {
        std::string path = mTwoQubitGates[g];
        if (!mOpt.base_dir.empty() && !path.empty() && path[0] != '/') path = mOpt.base_dir + "/" + path;
        n = mOpt.dim_log2 == 1 ? AmpArbitrary(path, g, 2) : ArbitraryTwoQubit(path, g);
        AttachTwoQubit(n, std::stoi(arg(1)), std::stoi(arg(2)));
    } else {
        throw InvalidFileFormat();  // src/Network.h:694-702
    }
    mAllNodes.push_back(n);
    n->mID = int(mAllNodes.size()) - 1;
}

void Network::UploadLeaves() {
    std::vector<int> ranks;
    std::vector<const void*> host;
    std::vector<Node*> nodes;
    for (auto& n : mAllNodes) {
        if (n->mDev || n->mHostVals.empty()) continue;
        ranks.push_back(n->mRank);
        host.push_back(n->mHostVals.data());
        nodes.push_back(n.get());
    }
    if (nodes.empty()) return;
    std::vector<qt_tensor*> out(nodes.size(), nullptr);
    Check(mCtx, qt_tensor_create_many(mCtx, int(nodes.size()), ranks.data(), mOpt.dim_log2, QT_DTYPE_C128, host.data(),
                                      out.data()),
          "leaf upload");
    for (size_t i = 0; i < nodes.size(); ++i) {
        nodes[i]->mDev = out[i];
        nodes[i]->mCtx = mCtx;
        nodes[i]->mHostVals = std::vector<cd>();
    }
}

void Network::EnsureOnDevice(const std::shared_ptr<Node>& n) {
    if (n->mDev || !mCtx) return;
    if (n->mHostVals.empty()) throw InvalidFunctionInput();
    qt_tensor* t = nullptr;
    Check(mCtx, qt_tensor_alloc(mCtx, n->mRank, mOpt.dim_log2, QT_DTYPE_C128, &t), "alloc");
    n->mDev = t;
    n->mCtx = mCtx;
    Check(mCtx, qt_tensor_upload(mCtx, t, n->mHostVals.data(), int64_t(n->mHostVals.size() * sizeof(cd))),
          "upload");
    n->mHostVals = std::vector<cd>();
}

// ContractNetworkLinearly (src/Network.h:309-322)
void Network::ContractNetworkLinearly() {
    while (!mDone) {
        auto& w = mUncontracted[0]->GetWires()[0];
        auto a = w->GetNodeA().lock(), b = w->GetNodeB().lock();
        if (!a || !b) throw InvalidContractionMethod();
        ContractNodes(a, b, 1000);
    }
}

// ContractNodes (src/Network.h:715-864).  The bookkeeping runs under mLocker, the
// device step outside it, as the reference runs ContractIndices unlocked
// (:831-837), so threads contracting disjoint pairs of one network overlap.
std::shared_ptr<Node> Network::ContractNodes(std::shared_ptr<Node> A, std::shared_ptr<Node> B, int threshold) {
    if (!A || !B) throw InvalidFunctionInput();
    std::unique_lock<std::mutex> lock(mLocker);
    if (A->mContracted || B->mContracted) return nullptr;
    auto touches = [](const std::shared_ptr<Wire>& w, const std::shared_ptr<Node>& n) {
        return w->GetNodeB().lock() == n || w->GetNodeA().lock() == n;
    };
    std::vector<int> ia, ib;
    std::vector<std::pair<bool, int>> ic;
    std::vector<std::shared_ptr<Wire>> connected, remaining;
    const auto& wa = A->GetWires();
    const auto& wb = B->GetWires();
    for (size_t i = 0; i < wa.size(); ++i) {
        if (touches(wa[i], B)) {
            ia.push_back(int(i));
            connected.push_back(wa[i]);
        } else {
            ic.push_back({true, int(i)});
            remaining.push_back(wa[i]);
        }
    }
    for (const auto& cw : connected)
        for (size_t k = 0; k < wb.size(); ++k)
            if (cw == wb[k]) ib.push_back(int(k));
    for (size_t j = 0; j < wb.size(); ++j)
        if (!touches(wb[j], A)) {
            ic.push_back({false, int(j)});
            remaining.push_back(wb[j]);
        }
    // the reference compares size_t against int: a negative limit wraps to a huge one
    const size_t limit = static_cast<size_t>(std::max(A->mRank, B->mRank) + threshold);
    if ((ia.empty() && !ic.empty()) || remaining.size() > limit) return nullptr;
    // out of time: stop before the network changes
    if (totTimer.getElapsed() > maxTime) throw ContractionFailure();
    if (!A->HasData() || !B->HasData()) throw InvalidFunctionInput();
    std::vector<std::pair<int, int>> pairs;
    for (size_t i = 0; i < ia.size() && i < ib.size(); ++i) pairs.push_back({ia[i], ib[i]});
    auto C = std::make_shared<Node>(int(ic.size()));
    for (const auto& w : remaining) {
        C->GetWires().push_back(w);
        const auto na = w->GetNodeA().lock();
        if (na == A || na == B) {
            w->SetNodeA(C);
        } else {
            const auto nb = w->GetNodeB().lock();
            if (nb == A || nb == B) w->SetNodeB(C);
        }
    }
    if (mOpt.verbose && C->mRank >= 8)
        std::cout << "Contracting Nodes of Rank " << A->mRank << " and " << B->mRank
                  << " to get a Node of Rank: " << C->mRank << " Hold On....." << std::endl;
    B->mContracted = true;
    A->mContracted = true;
    for (auto& w : connected) w->SetIsContracted(true);
    lock.unlock();
    PlanStep st;
    cd value(std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::quiet_NaN());
    ContractIndices(ic, pairs, A, B, C, &st, &value);
    lock.lock();
    // the counter, the final-value rule and the done test of ContractIndices
    // (src/Network.h:884-885, 961-969)
    mNumFloatOps += 1LL << (mOpt.dim_log2 * (ic.size() + pairs.size()));
    st.c_id = int(mAllNodes.size());
    mPlan.push_back(std::move(st));
    if (ic.empty()) {
        if ((std::abs(mFinalVal.real()) <= 1.0e-30 && std::abs(mFinalVal.imag()) <= 1.0e-30) ||
            (A->mRank == 0 && B->mRank == 0))
            mFinalVal = value;
        if (mUncontracted.size() == 2) {
            mDone = true;
            mPlan.back().c_id = -1;
        }
    } else if (!mOpt.open_qubits.empty() && mUncontracted.size() == 2) {
        mDone = true;  // open qubits: the last step leaves their tensor (kept in C)
        mPlan.back().c_id = -1;
    }
    if (mDone) {
        auto t = std::make_shared<Node>(0);
        t->mID = int(mAllNodes.size());
        t->mCreatedFrom = {A->mID, B->mID};
        mAllNodes.push_back(t);
        FindAndRemove(A);
        FindAndReplace(B, C);
        A->ClearNodeData();
        B->ClearNodeData();
        if (!ic.empty()) mFinalTensor = C;
        return nullptr;
    }
    C->mID = int(mAllNodes.size());
    C->mCreatedFrom = {A->mID, B->mID};
    mAllNodes.push_back(C);
    A->ClearNodeData();
    B->ClearNodeData();
    FindAndRemove(A);
    FindAndReplace(B, C);
    return C;
}

// ContractIndices (src/Network.h:876-971): the arithmetic runs on the device.
void Network::ContractIndices(const std::vector<std::pair<bool, int>>& free,
                              const std::vector<std::pair<int, int>>& sum, const std::shared_ptr<Node>& A,
                              const std::shared_ptr<Node>& B, const std::shared_ptr<Node>& C, PlanStep* step,
                              cd* value) {
    PlanStep& st = *step;
    st.a_id = A->mID;
    st.b_id = B->mID;
    st.rank_a = A->mRank;
    st.rank_b = B->mRank;
    st.k = int(sum.size());
    st.rank_c = int(free.size());
    for (const auto& p : sum) {
        st.pair_a.push_back(p.first);
        st.pair_b.push_back(p.second);
    }
    for (const auto& f : free) st.free_src.push_back(f.first ? (f.second | QT_FROM_A) : f.second);
    if (!mCtx) {
        C->mPlanned = true;
        return;
    }
    EnsureOnDevice(A);
    EnsureOnDevice(B);
    qt_tensor* t = nullptr;
    Check(mCtx, qt_tensor_alloc_empty(mCtx, C->mRank, mOpt.dim_log2, QT_DTYPE_C128, &t), "alloc C");
    C->mDev = t;
    C->mCtx = mCtx;
    Check(mCtx,
          qt_contract_pair(mCtx, A->mDev, A->mRank, B->mDev, B->mRank, st.k, st.pair_a.data(), st.pair_b.data(),
                           st.rank_c, st.free_src.data(), t, QT_PLAIN_C128, nullptr),
          "contract");
    if (free.empty()) {
        double re = 0, im = 0;
        Check(mCtx, qt_tensor_read_scalar(mCtx, t, &re, &im), "scalar");
        *value = cd(re, im);
    }
}

void Network::FindAndRemove(const std::shared_ptr<Node>& n) {  // src/Network.h:1269-1282
    auto it = std::find(mUncontracted.begin(), mUncontracted.end(), n);
    if (it != mUncontracted.end()) mUncontracted.erase(it);
}

void Network::FindAndReplace(const std::shared_ptr<Node>& from, const std::shared_ptr<Node>& to) {
    auto it = std::find(mUncontracted.begin(), mUncontracted.end(), from);  // src/Network.h:1253-1265
    if (it != mUncontracted.end()) *it = to;
}

// src/Network.h:1233-1249, including its early exit: after two replacements every
// later row is inspected at its first entry only.
void Network::FindAndReplace2D(std::vector<std::vector<std::shared_ptr<Node>>>& t,
                               const std::shared_ptr<Node>& from, const std::shared_ptr<Node>& to) {
    int count = 0;
    for (auto& row : t)
        for (auto& n : row) {
            if (n == from) {
                n = to;
                ++count;
            }
            if (count >= 2) break;
        }
}

// ReduceCircuit (src/Network.h:1015-1118): absorb every rank-2 node into its
// predecessor on the wire, then merge successive rank-4 nodes on the same pair.
void Network::ReduceCircuit() {
    const size_t nq = size_t(mNumberOfQubits);
    std::vector<std::vector<std::shared_ptr<Node>>> place(nq);
    for (size_t i = 0; i < nq; ++i) {
        for (size_t idx = 0; idx < mNodesByWire[i].size(); ++idx) {
            auto node = mNodesByWire[i][idx];
            if (node->mRank == 2) {
                auto prev = place[i].back();
                auto tmp = ContractNodes(prev, node, 0);
                if (!tmp) throw ContractionFailure();
                auto src = node->GetWires()[0]->GetNodeA().lock();
                tmp->AddWireNumber(src->GetWireNumber()[0]);
                if (src->mRank > 2) tmp->AddWireNumber(src->GetWireNumber()[1]);
                if (prev->GetTypeOfNode() == GateType::INITSTATE) {
                    tmp->SetTypeOfNode(GateType::INITSTATE);
                    tmp->SetTypeOfNodeString("INITSTATE(Manipulated)");
                }
                FindAndReplace2D(mNodesByWire, prev, tmp);
                FindAndReplace2D(place, prev, tmp);
            } else {
                place[i].push_back(node);
            }
        }
    }
    mNodesByWire = std::move(place);

    std::vector<std::vector<std::shared_ptr<Node>>> out(nq);
    std::vector<std::shared_ptr<Node>> cur(nq);
    std::vector<size_t> cidx(nq, 0);
    std::vector<bool> updated(nq, false);
    bool flag = false, found = false;
    auto& nbw = mNodesByWire;
    while (!flag) {
        for (size_t j = 0; j < nq; ++j) {
            if (cidx[j] >= nbw[j].size() || updated[j]) continue;
            auto t = nbw[j][cidx[j]];
            if (t->mRank == 1) {
                cur[j] = t;
                updated[j] = true;
                ++cidx[j];
                continue;
            }
            const size_t w0 = size_t(t->GetWireNumber()[0]), w1 = size_t(t->GetWireNumber()[1]);
            if (!updated[w1] && !updated[w0]) {
                if (cidx[w0] < nbw[w0].size() && cidx[w1] < nbw[w1].size() &&
                    nbw[w0][cidx[w0]] == nbw[w1][cidx[w1]]) {
                    if (cur[w0] == cur[w1]) {
                        cur[w0] = ContractNodes(t, cur[w0], 0);
                        if (!cur[w0]) throw ContractionFailure();
                        cur[w1] = cur[w0];
                        cur[w0]->AddWireNumber(int(w0));
                        cur[w0]->AddWireNumber(int(w1));
                        found = true;
                    } else {
                        cur[w0] = t;
                        cur[w1] = t;
                    }
                    ++cidx[w0];
                    ++cidx[w1];
                    updated[w0] = true;
                    updated[w1] = true;
                }
            }
        }
        for (size_t j = 0; j < nq; ++j) {
            if (updated[j]) {
                if (found) out[j].back() = cur[j];
                else out[j].push_back(cur[j]);
            } else if (cidx[j] < nbw[j].size() && !found) {
                out[j].push_back(nullptr);
            }
        }
        if (std::none_of(updated.begin(), updated.end(), [](bool b) { return b; })) flag = true;
        std::fill(updated.begin(), updated.end(), false);
        found = false;
    }
    mNodesByWire = std::move(out);
}

std::vector<std::complex<double>> Network::GetFinalTensor(std::vector<int>* axis_qubits) const {
    if (!mDone) throw InvalidFunctionInput();
    if (axis_qubits) axis_qubits->clear();
    if (!mFinalTensor) return {mFinalVal};
    if (axis_qubits)
        for (const auto& w : mFinalTensor->GetWires()) axis_qubits->push_back(w->GetQubitNumber());
    if (!mCtx) {
        const double nan = std::numeric_limits<double>::quiet_NaN();
        return std::vector<cd>(size_t(1) << (mOpt.dim_log2 * mFinalTensor->mRank), cd(nan, nan));
    }
    return mFinalTensor->GetTensorVals();
}

std::vector<std::pair<int, int>> RecordSequence(const Network& net) {
    std::vector<std::pair<int, int>> seq;
    for (const auto& n : net.GetAllNodes())
        if (!(n->mCreatedFrom.first == 0 && n->mCreatedFrom.second == 0)) seq.push_back(n->mCreatedFrom);
    return seq;
}

}  // namespace qtorch
