// OpenQASM 2.0 export/import (reference src/io/qasm.cpp). The text format
// (header lines, %.17g parameters, "name(p,...) q[a],q[b];", measure and
// barrier forms) and the accepted subset follow the reference so that files
// round-trip between the two; the lexer and parser here are this library's own:
// source -> token list -> statements.
#include "lqcs/io/qasm.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numbers>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace lqcs::io {

// ---- export ---------------------------------------------------------------------------

namespace {

// qelib1.inc spelling of (gate type, number of controls); the reference's table
// (qasm.cpp:35-59).
struct QasmName {
    GateType type;
    std::size_t controls;
    const char* name;
};

constexpr QasmName kNames[] = {
    {GateType::I, 0, "id"},    {GateType::X, 0, "x"},     {GateType::X, 1, "cx"},    {GateType::X, 2, "ccx"},
    {GateType::Y, 0, "y"},     {GateType::Y, 1, "cy"},    {GateType::Z, 0, "z"},     {GateType::Z, 1, "cz"},
    {GateType::H, 0, "h"},     {GateType::H, 1, "ch"},    {GateType::S, 0, "s"},     {GateType::Sdg, 0, "sdg"},
    {GateType::T, 0, "t"},     {GateType::Tdg, 0, "tdg"}, {GateType::SX, 0, "sx"},   {GateType::RX, 0, "rx"},
    {GateType::RX, 1, "crx"},  {GateType::RY, 0, "ry"},   {GateType::RY, 1, "cry"},  {GateType::RZ, 0, "rz"},
    {GateType::RZ, 1, "crz"},  {GateType::P, 0, "u1"},    {GateType::P, 1, "cu1"},   {GateType::U, 0, "u3"},
    {GateType::SWAP, 0, "swap"}, {GateType::SWAP, 1, "cswap"},
};

const char* qasm_name(const Gate& g) {
    for (const QasmName& n : kNames)
        if (n.type == g.type && n.controls == g.n_controls) return n.name;
    return nullptr;
}

void put_double(std::string& out, double v) {
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    out += buf;
}

void put_qubit(std::string& out, qubit_t q) { out += "q[" + std::to_string(q) + "]"; }

}  // namespace

std::string to_qasm(const QuantumCircuit& circuit) {
    std::string out = "OPENQASM 2.0;\ninclude \"qelib1.inc\";\n";
    out += "qreg q[" + std::to_string(circuit.num_qubits()) + "];\n";
    if (circuit.num_clbits() > 0) out += "creg c[" + std::to_string(circuit.num_clbits()) + "];\n";
    for (const Instruction& inst : circuit.instructions()) {
        const Gate& g = inst.gate;
        switch (g.type) {
            case GateType::Barrier:
                out += "barrier q;\n";
                continue;
            case GateType::Measure:
                out += "measure ";
                put_qubit(out, inst.qubits[0]);
                out += " -> c[" + std::to_string(inst.clbits[0]) + "];\n";
                continue;
            case GateType::Reset:
                out += "reset ";
                put_qubit(out, inst.qubits[0]);
                out += ";\n";
                continue;
            default:
                break;
        }
        const char* name = qasm_name(g);
        if (!name)
            throw std::invalid_argument("to_qasm: gate '" + g.name() +
                                        "' is not representable in the OpenQASM 2.0 qelib1 gate set");
        out += name;
        if (!g.params.empty()) {
            out += '(';
            for (std::size_t i = 0; i < g.params.size(); ++i) {
                if (i) out += ',';
                put_double(out, g.params[i]);
            }
            out += ')';
        }
        for (std::size_t i = 0; i < inst.qubits.size(); ++i) {
            out += i ? ',' : ' ';
            put_qubit(out, inst.qubits[i]);
        }
        out += ";\n";
    }
    return out;
}

// ---- import ---------------------------------------------------------------------------

namespace {

enum class Tok { Ident, Number, String, Symbol, Arrow, End };

struct Token {
    Tok kind;
    std::string text;
    double value = 0.0;
    std::size_t line = 1, col = 1;
};

[[noreturn]] void syntax_error(const std::string& msg, const Token& at) {
    throw std::invalid_argument("from_qasm: " + msg + " at line " + std::to_string(at.line) + ", column " +
                                std::to_string(at.col));
}

// Splits the source into identifiers, numbers (with their literal text),
// quoted strings, "->" and one-character symbols; // comments are dropped.
std::vector<Token> tokenize(std::string_view src) {
    std::vector<Token> toks;
    std::size_t i = 0, line = 1, line_start = 0;
    auto here = [&](Tok k) {
        Token t{k, {}};
        t.line = line;
        t.col = i - line_start + 1;
        return t;
    };
    while (i < src.size()) {
        const char c = src[i];
        if (c == '\n') {
            ++line;
            line_start = ++i;
            continue;
        }
        if (std::isspace(static_cast<unsigned char>(c))) {
            ++i;
            continue;
        }
        if (c == '/' && i + 1 < src.size() && src[i + 1] == '/') {
            while (i < src.size() && src[i] != '\n') ++i;
            continue;
        }
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            Token t = here(Tok::Ident);
            while (i < src.size() && (std::isalnum(static_cast<unsigned char>(src[i])) || src[i] == '_'))
                t.text += src[i++];
            toks.push_back(std::move(t));
            continue;
        }
        if (std::isdigit(static_cast<unsigned char>(c)) || (c == '.' && i + 1 < src.size() &&
                                                              std::isdigit(static_cast<unsigned char>(src[i + 1])))) {
            Token t = here(Tok::Number);
            const std::string rest(src.substr(i, std::min<std::size_t>(64, src.size() - i)));
            char* end = nullptr;
            t.value = std::strtod(rest.c_str(), &end);
            const std::size_t len = static_cast<std::size_t>(end - rest.c_str());
            t.text = rest.substr(0, len);
            i += len;
            toks.push_back(std::move(t));
            continue;
        }
        if (c == '"') {
            Token t = here(Tok::String);
            const std::size_t close = src.find('"', i + 1);
            if (close == std::string_view::npos) syntax_error("unterminated string", t);
            t.text = std::string(src.substr(i + 1, close - i - 1));
            i = close + 1;
            toks.push_back(std::move(t));
            continue;
        }
        if (c == '-' && i + 1 < src.size() && src[i + 1] == '>') {
            Token t = here(Tok::Arrow);
            t.text = "->";
            i += 2;
            toks.push_back(std::move(t));
            continue;
        }
        Token t = here(Tok::Symbol);
        t.text = std::string(1, c);
        ++i;
        toks.push_back(std::move(t));
    }
    Token end = here(Tok::End);
    toks.push_back(end);
    return toks;
}

class Reader {
public:
    explicit Reader(std::vector<Token> toks) : t_(std::move(toks)) {}

    QuantumCircuit read() {
        if (!(peek().kind == Tok::Ident && peek().text == "OPENQASM")) syntax_error("expected 'OPENQASM'", peek());
        next();
        if (!(peek().kind == Tok::Number && peek().text == "2.0")) syntax_error("expected version 2.0", peek());
        next();
        symbol(";");
        while (peek().kind != Tok::End) statement();
        if (!qc_) syntax_error("missing qreg declaration", peek());
        return std::move(*qc_);
    }

private:
    const Token& peek() const { return t_[p_]; }
    const Token& next() { return t_[p_ < t_.size() - 1 ? p_++ : p_]; }
    bool at_symbol(const char* s) const { return peek().kind == Tok::Symbol && peek().text == s; }
    bool take_symbol(const char* s) {
        if (!at_symbol(s)) return false;
        next();
        return true;
    }
    void symbol(const char* s) {
        if (!take_symbol(s)) syntax_error(std::string("expected '") + s + "'", peek());
    }
    std::string ident() {
        if (peek().kind != Tok::Ident) syntax_error("expected identifier", peek());
        return next().text;
    }

    // expr := term {(+|-) term}; term := factor {(*|/) factor};
    // factor := -factor | (expr) | number | pi
    double expr() {
        double v = term();
        for (;;) {
            if (take_symbol("+")) v += term();
            else if (take_symbol("-")) v -= term();
            else return v;
        }
    }
    double term() {
        double v = factor();
        for (;;) {
            if (take_symbol("*")) v *= factor();
            else if (take_symbol("/")) v /= factor();
            else return v;
        }
    }
    double factor() {
        if (take_symbol("-")) return -factor();
        if (take_symbol("(")) {
            const double v = expr();
            symbol(")");
            return v;
        }
        if (peek().kind == Tok::Number) return next().value;
        if (peek().kind == Tok::Ident) {
            if (peek().text != "pi") syntax_error("unknown symbol '" + peek().text + "' in expression", peek());
            next();
            return std::numbers::pi;
        }
        syntax_error("expected a number", peek());
    }

    std::size_t index_expr() {
        const Token& at = peek();
        const double v = expr();
        if (!(v >= 0.0) || v != std::floor(v)) syntax_error("index must be a non-negative integer", at);
        return static_cast<std::size_t>(v);
    }

    // reg or reg[index]; the register must be `want`.
    std::optional<std::size_t> ref(const std::string& want) {
        const Token& at = peek();
        const std::string reg = ident();
        if (reg != want) syntax_error("unknown register '" + reg + "'", at);
        if (!take_symbol("[")) return std::nullopt;
        const std::size_t i = index_expr();
        symbol("]");
        return i;
    }

    void statement() {
        const Token& at = peek();
        const std::string head = ident();
        if (head == "include") {
            if (peek().kind != Tok::String) syntax_error("expected a quoted file name", peek());
            next();
            symbol(";");
            return;
        }
        if (head == "qreg" || head == "creg") {
            const std::string name = ident();
            symbol("[");
            const std::size_t size = index_expr();
            symbol("]");
            symbol(";");
            declare(head == "qreg", name, size, at);
            return;
        }
        if (!qc_) syntax_error("statement before qreg declaration", at);
        if (head == "measure") return measure(at);
        if (head == "barrier") {  // operands ignored: a barrier on every qubit
            while (!at_symbol(";") && peek().kind != Tok::End) next();
            symbol(";");
            qc_->barrier();
            return;
        }
        if (head == "reset") {
            const auto q = ref(qreg_);
            symbol(";");
            if (!q) syntax_error("reset requires an indexed qubit", at);
            qc_->reset(static_cast<qubit_t>(*q));
            return;
        }
        gate(head, at);
    }

    void declare(bool quantum, const std::string& name, std::size_t size, const Token& at) {
        if (quantum) {
            if (qc_) syntax_error("only one qreg is supported", at);
            qreg_ = name;
            nq_ = size;
        } else {
            if (nc_ > 0) syntax_error("only one creg is supported", at);
            if (qc_ && qc_->num_instructions() > 0) syntax_error("creg must be declared before any gate", at);
            creg_ = name;
            nc_ = size;
        }
        if (nq_ > 0) qc_ = std::make_unique<QuantumCircuit>(nq_, nc_);
    }

    void measure(const Token& at) {
        const auto q = ref(qreg_);
        if (peek().kind != Tok::Arrow) syntax_error("expected '->'", peek());
        next();
        const auto c = ref(creg_);
        symbol(";");
        if (q && c) {
            qc_->measure(static_cast<qubit_t>(*q), static_cast<clbit_t>(*c));
        } else if (!q && !c) {
            if (nq_ != nc_) syntax_error("whole-register measure needs equal register sizes", at);
            for (std::size_t i = 0; i < nq_; ++i) qc_->measure(static_cast<qubit_t>(i), static_cast<clbit_t>(i));
        } else {
            syntax_error("mixed whole-register and indexed measure", at);
        }
    }

    void gate(const std::string& name, const Token& at) {
        std::vector<double> p;
        if (take_symbol("(")) {
            p.push_back(expr());
            while (take_symbol(",")) p.push_back(expr());
            symbol(")");
        }
        std::vector<qubit_t> q;
        do {
            const auto i = ref(qreg_);
            if (!i) syntax_error("gate operands must be indexed qubits", at);
            q.push_back(static_cast<qubit_t>(*i));
        } while (take_symbol(","));
        symbol(";");
        if (!apply(name, p, q))
            syntax_error("unsupported gate '" + name + "' with " + std::to_string(p.size()) + " params / " +
                             std::to_string(q.size()) + " qubits",
                         at);
    }

    // qelib1.inc gates by (name, parameter count, qubit count).
    bool apply(const std::string& n, const std::vector<double>& p, const std::vector<qubit_t>& q) {
        QuantumCircuit& c = *qc_;
        const std::size_t np = p.size(), nq = q.size();
        using One = QuantumCircuit& (QuantumCircuit::*)(qubit_t);
        using OneP = QuantumCircuit& (QuantumCircuit::*)(double, qubit_t);
        using Two = QuantumCircuit& (QuantumCircuit::*)(qubit_t, qubit_t);
        using TwoP = QuantumCircuit& (QuantumCircuit::*)(double, qubit_t, qubit_t);
        static const std::pair<const char*, One> one[] = {
            {"id", &QuantumCircuit::i}, {"x", &QuantumCircuit::x},     {"y", &QuantumCircuit::y},
            {"z", &QuantumCircuit::z},  {"h", &QuantumCircuit::h},     {"s", &QuantumCircuit::s},
            {"sdg", &QuantumCircuit::sdg}, {"t", &QuantumCircuit::t}, {"tdg", &QuantumCircuit::tdg},
            {"sx", &QuantumCircuit::sx}};
        static const std::pair<const char*, OneP> one_p[] = {
            {"rx", &QuantumCircuit::rx}, {"ry", &QuantumCircuit::ry}, {"rz", &QuantumCircuit::rz},
            {"p", &QuantumCircuit::p},   {"u1", &QuantumCircuit::p}};
        static const std::pair<const char*, Two> two[] = {
            {"cx", &QuantumCircuit::cx}, {"cy", &QuantumCircuit::cy}, {"cz", &QuantumCircuit::cz},
            {"ch", &QuantumCircuit::ch}, {"swap", &QuantumCircuit::swap}};
        static const std::pair<const char*, TwoP> two_p[] = {
            {"cp", &QuantumCircuit::cp},   {"cu1", &QuantumCircuit::cp}, {"crx", &QuantumCircuit::crx},
            {"cry", &QuantumCircuit::cry}, {"crz", &QuantumCircuit::crz}};
        if (np == 0 && nq == 1)
            for (const auto& [name, f] : one)
                if (n == name) return (c.*f)(q[0]), true;
        if (np == 1 && nq == 1)
            for (const auto& [name, f] : one_p)
                if (n == name) return (c.*f)(p[0], q[0]), true;
        if (np == 0 && nq == 2)
            for (const auto& [name, f] : two)
                if (n == name) return (c.*f)(q[0], q[1]), true;
        if (np == 1 && nq == 2)
            for (const auto& [name, f] : two_p)
                if (n == name) return (c.*f)(p[0], q[0], q[1]), true;
        if (n == "u2" && np == 2 && nq == 1) return c.u(std::numbers::pi / 2, p[0], p[1], q[0]), true;
        if ((n == "u" || n == "u3") && np == 3 && nq == 1) return c.u(p[0], p[1], p[2], q[0]), true;
        if (n == "ccx" && np == 0 && nq == 3) return c.ccx(q[0], q[1], q[2]), true;
        if (n == "cswap" && np == 0 && nq == 3) return c.cswap(q[0], q[1], q[2]), true;
        return false;
    }

    std::vector<Token> t_;
    std::size_t p_ = 0;
    std::unique_ptr<QuantumCircuit> qc_;
    std::string qreg_, creg_;
    std::size_t nq_ = 0, nc_ = 0;
};

}  // namespace

QuantumCircuit from_qasm(std::string_view source) { return Reader(tokenize(source)).read(); }

}  // namespace lqcs::io
