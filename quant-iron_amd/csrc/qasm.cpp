// qasm.cpp — OpenQASM 3.0 text as an input format for Circuit::execute and the
// batched engine: a host-side reader that turns a program into qi_gate
// records (include/qiron.h, qi_qasm_parse).
//
// It reads the programs quant-iron's exporter writes (src/compiler/qasm.rs,
// src/compiler/ir.rs) and the same gate set in standard OpenQASM 3 syntax:
//   - header `OPENQASM 3[.0];`, `include "...";` (ignored), the exporter's
//     `def xmeasure(...)` / `def ymeasure(...)` definitions (recognised, not
//     executed as code: `xmeasure` / `ymeasure` calls are measurements in the
//     X / Y basis, MeasurementBasis::X / Y, measurement.rs:76-86);
//   - declarations `qubit[n] q;`, `qubit q;`, `qreg q[n];`, `bit[n] c;`,
//     `bit c;`, `creg c[n];`, and circuit parameters `input float[64] t;` /
//     `input angle[32] t;` / `input float t;` (slots in declaration order);
//   - the stdgates.inc gates id, x, y, z, h, s, sdg, t, tdg, sx, p/phase/u1,
//     rx, ry, rz, U/u3, u2, cx/CX, cy, cz, ch, cp/cphase, crx, cry, crz, cu,
//     swap, ccx, cswap, with the modifiers `ctrl @`, `ctrl(k) @`,
//     `negctrl[(k)] @` and `inv @`, register broadcasting (`h q;`), and the
//     exporter's control form `ctrl(k) @ q[c1], ..., q[ck] g q[t]`;
//   - measurements `c[i] = measure q[j];`, `c = measure q;`,
//     `measure q[j] -> c[i];` (runs of consecutive measurements in one basis
//     with no gate between them are one joint Gate::Measurement, gate.rs:103);
//   - `barrier` (ignored); statements may omit the final `;` when a line
//     ends (the exporter writes gate lines without one); `//` and `/* */`
//     comments.
// Angles are expressions over numbers, pi/π, tau/τ, euler/ℇ, circuit
// parameters (only as a whole angle, which becomes an angle slot), + - * / ^,
// unary minus, parentheses and sin, cos, tan, exp, ln, sqrt, arcsin, arccos,
// arctan.
//
// Gate semantics: U(θ, φ, λ) = [[cos θ/2, -e^{iλ} sin θ/2], [e^{iφ} sin θ/2,
// e^{i(φ+λ)} cos θ/2]] is emitted exactly as P(λ), RY(θ), P(φ) (with the same
// controls), so it needs no unitarity tolerance; u2(φ, λ) = U(π/2, φ, λ);
// cu(θ, φ, λ, γ) adds P(γ) on the control. inv @ reverses a gate's
// decomposition and inverts each part (angles negated, S <-> S†, T <-> T†).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/qiron.h"

namespace {

struct ParseError {
    size_t line;
    std::string msg;
};

struct Reg {
    uint64_t offset, size;
    bool is_qubit;
};

struct Operand {  // one qubit, or a whole register (broadcast)
    uint64_t first, count;
    bool whole;
};

// Reader-only gate kinds beside qi_gate_kind: U (decomposed) and sqrt(X).
constexpr int kU = 1000;
constexpr int kSX = 1001;

class Reader {
  public:
    Reader(const char *src, size_t len, const double *values, size_t nvalues)
        : s_(src, len), values_(values), nvalues_(nvalues) {}

    void run() {
        skip_ws();
        if (peek_word("OPENQASM")) {
            word();
            skip_ws_inline();
            while (pos_ < s_.size() && s_[pos_] != ';' && s_[pos_] != '\n') pos_++;
            end_statement();
        }
        while (true) {
            skip_ws();
            if (pos_ >= s_.size()) break;
            statement();
        }
        flush_measure();
    }

    std::vector<qi_gate> gates;
    std::vector<int32_t> slots;
    uint64_t nqubits = 0;
    uint32_t nparams = 0;

  private:
    std::string s_;
    size_t pos_ = 0;
    const double *values_;
    size_t nvalues_;
    std::map<std::string, Reg> regs_;
    std::map<std::string, uint32_t> params_;
    // the pending joint measurement
    std::vector<uint64_t> mq_;
    int mbasis_ = -1;
    std::string mreg_;

    [[noreturn]] void fail(const std::string &m) const { throw ParseError{line(), m}; }
    size_t line() const {
        size_t l = 1;
        for (size_t i = 0; i < pos_ && i < s_.size(); i++) l += s_[i] == '\n';
        return l;
    }

    // ---- lexing
    void skip_comment_or(bool newlines) {
        while (pos_ < s_.size()) {
            char c = s_[pos_];
            if (c == '/' && pos_ + 1 < s_.size() && s_[pos_ + 1] == '/') {
                while (pos_ < s_.size() && s_[pos_] != '\n') pos_++;
            } else if (c == '/' && pos_ + 1 < s_.size() && s_[pos_ + 1] == '*') {
                size_t e = s_.find("*/", pos_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                pos_ = e + 2;
            } else if (c == ' ' || c == '\t' || c == '\r' || (newlines && c == '\n')) {
                pos_++;
            } else {
                break;
            }
        }
    }
    void skip_ws() { skip_comment_or(true); }
    void skip_ws_inline() { skip_comment_or(false); }

    bool at(char c) { skip_ws_inline(); return pos_ < s_.size() && s_[pos_] == c; }
    void expect(char c) {
        skip_ws_inline();
        if (pos_ >= s_.size() || s_[pos_] != c) fail(std::string("expected '") + c + "'");
        pos_++;
    }
    // A statement ends at ';' or at the end of its line (the exporter's gate lines).
    void end_statement() {
        skip_ws_inline();
        if (pos_ < s_.size() && s_[pos_] == ';') { pos_++; return; }
        if (pos_ >= s_.size() || s_[pos_] == '\n') return;
        fail("expected ';' or the end of the line");
    }
    static bool ident_char(unsigned char c, bool first) {
        return std::isalpha(c) || c == '_' || c >= 0x80 || (!first && std::isdigit(c));
    }
    bool peek_word(const char *w) {
        skip_ws_inline();
        size_t n = strlen(w);
        return s_.compare(pos_, n, w) == 0 && (pos_ + n >= s_.size() || !ident_char((unsigned char)s_[pos_ + n], false));
    }
    std::string word() {
        skip_ws_inline();
        size_t b = pos_;
        while (pos_ < s_.size() && ident_char((unsigned char)s_[pos_], pos_ == b)) pos_++;
        if (b == pos_) fail("expected a name");
        return s_.substr(b, pos_ - b);
    }
    uint64_t integer() {
        skip_ws_inline();
        size_t b = pos_;
        while (pos_ < s_.size() && std::isdigit((unsigned char)s_[pos_])) pos_++;
        if (b == pos_) fail("expected an integer");
        return std::stoull(s_.substr(b, pos_ - b));
    }

    // ---- expressions (angles)
    struct Val {
        double v;
        int param;  // >= 0: the value is exactly this parameter
    };
    Val expr() {
        Val a = term();
        while (true) {
            if (at('+')) { pos_++; a = combine(a, term(), '+'); }
            else if (at('-')) { pos_++; a = combine(a, term(), '-'); }
            else return a;
        }
    }
    Val term() {
        Val a = power();
        while (true) {
            if (at('*')) { pos_++; a = combine(a, power(), '*'); }
            else if (at('/')) { pos_++; a = combine(a, power(), '/'); }
            else return a;
        }
    }
    Val power() {
        Val a = unary();
        if (at('^') || (at('*') && pos_ + 1 < s_.size() && s_[pos_ + 1] == '*')) {
            pos_ += s_[pos_] == '^' ? 1 : 2;
            a = combine(a, power(), '^');
        }
        return a;
    }
    Val unary() {
        if (at('-')) { pos_++; Val v = unary(); if (v.param >= 0) fail("a circuit parameter may only be a whole angle"); return {-v.v, -1}; }
        if (at('+')) { pos_++; return unary(); }
        return primary();
    }
    Val combine(Val a, Val b, char op) {
        if (a.param >= 0 || b.param >= 0) fail("a circuit parameter may only be a whole angle");
        switch (op) {
        case '+': return {a.v + b.v, -1};
        case '-': return {a.v - b.v, -1};
        case '*': return {a.v * b.v, -1};
        case '/': return {a.v / b.v, -1};
        default: return {std::pow(a.v, b.v), -1};
        }
    }
    Val primary() {
        skip_ws_inline();
        if (pos_ >= s_.size()) fail("expected an expression");
        char c = s_[pos_];
        if (c == '(') { pos_++; Val v = expr(); expect(')'); return v; }
        if (std::isdigit((unsigned char)c) || c == '.') {
            const char *b = s_.c_str() + pos_;
            char *e = nullptr;
            double v = std::strtod(b, &e);
            if (e == b) fail("bad number");
            pos_ += size_t(e - b);
            return {v, -1};
        }
        std::string w = word();
        if (w == "pi" || w == "\xcf\x80") return {M_PI, -1};
        if (w == "tau" || w == "\xcf\x84") return {2.0 * M_PI, -1};
        if (w == "euler" || w == "\xe2\x84\x87") return {M_E, -1};
        static const std::map<std::string, double (*)(double)> fns = {
            {"sin", std::sin}, {"cos", std::cos}, {"tan", std::tan}, {"exp", std::exp}, {"ln", std::log},
            {"sqrt", std::sqrt}, {"arcsin", std::asin}, {"arccos", std::acos}, {"arctan", std::atan}};
        auto f = fns.find(w);
        if (f != fns.end()) {
            expect('(');
            Val v = expr();
            expect(')');
            if (v.param >= 0) fail("a circuit parameter may only be a whole angle");
            return {f->second(v.v), -1};
        }
        auto p = params_.find(w);
        if (p != params_.end()) {
            double v = p->second < nvalues_ && values_ ? values_[p->second] : 0.0;
            return {v, (int)p->second};
        }
        fail("unknown name '" + w + "' in an expression");
    }

    // ---- operands
    Operand operand() {
        std::string name = word();
        auto r = regs_.find(name);
        if (r == regs_.end() || !r->second.is_qubit) fail("'" + name + "' is not a qubit register");
        if (at('[')) {
            pos_++;
            uint64_t i = integer();
            expect(']');
            if (i >= r->second.size) fail("index " + std::to_string(i) + " out of range for '" + name + "'");
            return {r->second.offset + i, 1, false};
        }
        return {r->second.offset, r->second.size, r->second.size != 1};
    }
    std::vector<Operand> operand_list() {
        std::vector<Operand> v{operand()};
        while (at(',')) { pos_++; v.push_back(operand()); }
        return v;
    }

    // ---- statements
    void statement() {
        if (peek_word("include")) { word(); skip_ws_inline(); if (at('"')) { pos_++; while (pos_ < s_.size() && s_[pos_] != '"') pos_++; expect('"'); } end_statement(); return; }
        if (peek_word("def")) { skip_def(); return; }
        if (peek_word("qubit")) { declare(true, true); return; }
        if (peek_word("bit")) { declare(false, true); return; }
        if (peek_word("qreg")) { declare(true, false); return; }
        if (peek_word("creg")) { declare(false, false); return; }
        if (peek_word("input")) { declare_input(); return; }
        if (peek_word("barrier")) { word(); while (pos_ < s_.size() && s_[pos_] != ';' && s_[pos_] != '\n') pos_++; end_statement(); return; }
        if (peek_word("measure")) { word(); measure_arrow(); return; }
        if (peek_word("reset")) fail("reset is not supported");
        for (const char *kw : {"gate", "if", "for", "while", "let", "const", "output", "box", "extern", "cal", "defcal"})
            if (peek_word(kw)) fail(std::string("'") + kw + "' statements are not supported");
        // `c[i] = measure q[j];` / `c = measure q;` or a gate call
        size_t save = pos_;
        std::string w = word();
        auto r = regs_.find(w);
        if (r != regs_.end() && !r->second.is_qubit) {
            pos_ = save;
            measure_assign();
            return;
        }
        pos_ = save;
        gate_call();
    }

    void skip_def() {
        word();
        std::string name = word();
        if (name != "xmeasure" && name != "ymeasure") fail("subroutine definitions are not supported ('" + name + "')");
        int depth = 0;
        while (pos_ < s_.size()) {
            char c = s_[pos_++];
            if (c == '{') depth++;
            else if (c == '}' && --depth == 0) break;
        }
        if (depth != 0) fail("unterminated definition");
        end_statement();
    }

    void declare(bool qubit, bool v3) {
        word();
        uint64_t size = 1;
        std::string name;
        if (v3) {
            if (at('[')) { pos_++; size = integer(); expect(']'); }
            name = word();
        } else {
            name = word();
            expect('[');
            size = integer();
            expect(']');
        }
        if (regs_.count(name) || params_.count(name)) fail("'" + name + "' is declared twice");
        if (qubit) {
            regs_[name] = {nqubits, size, true};
            nqubits += size;
            if (nqubits > QI_MAX_QUBITS) fail("more qubits than the engine supports");
        } else {
            regs_[name] = {0, size, false};
        }
        end_statement();
    }

    void declare_input() {
        word();
        std::string type = word();
        if (type != "float" && type != "angle") fail("circuit parameters must be float or angle");
        if (at('[')) { pos_++; integer(); expect(']'); }
        std::string name = word();
        if (regs_.count(name) || params_.count(name)) fail("'" + name + "' is declared twice");
        params_[name] = nparams++;
        end_statement();
    }

    // ---- measurements
    void add_measure(uint64_t q, int basis, const std::string &reg) {
        if (mbasis_ >= 0 && (mbasis_ != basis || mreg_ != reg)) flush_measure();
        mreg_ = reg;
        for (uint64_t x : mq_)
            if (x == q) { flush_measure(); break; }
        mbasis_ = basis;
        mq_.push_back(q);
    }
    void flush_measure() {
        if (mq_.empty()) return;
        if (mq_.size() > QI_MAX_CONTROLS) fail("a joint measurement of more than 62 qubits");
        qi_gate g;
        memset(&g, 0, sizeof g);
        g.kind = QI_GATE_MEASURE;
        g.basis = (uint32_t)mbasis_;
        g.ncontrols = (uint32_t)mq_.size();
        for (size_t i = 0; i < mq_.size(); i++) g.controls[i] = mq_[i];
        g.params[0] = std::nan("");  // drawn when the gate runs
        gates.push_back(g);
        slots.push_back(-1);
        mq_.clear();
        mbasis_ = -1;
    }
    // the classical target: one bit or a whole register (its size)
    uint64_t bit_target(std::string &name) {
        name = word();
        auto r = regs_.find(name);
        if (r == regs_.end() || r->second.is_qubit) fail("'" + name + "' is not a bit register");
        if (at('[')) {
            pos_++;
            uint64_t i = integer();
            expect(']');
            if (i >= r->second.size) fail("bit index out of range");
            return 1;
        }
        return r->second.size;
    }
    void measured(Operand q, uint64_t nbits, int basis, const std::string &reg) {
        if (q.count != nbits) fail("measured qubits and bits differ in number");
        for (uint64_t k = 0; k < q.count; k++) add_measure(q.first + k, basis, reg);
    }
    void measure_assign() {
        std::string reg;
        uint64_t nb = bit_target(reg);
        expect('=');
        std::string m = word();
        int basis = m == "measure" ? QI_BASIS_COMPUTATIONAL : m == "xmeasure" ? QI_BASIS_X : m == "ymeasure" ? QI_BASIS_Y : -1;
        if (basis < 0) fail("expected measure, xmeasure or ymeasure");
        measured(operand(), nb, basis, reg);
        end_statement();
    }
    void measure_arrow() {
        Operand q = operand();
        std::string reg;
        if (at('-')) {
            pos_++;
            expect('>');
            uint64_t nb = bit_target(reg);
            measured(q, nb, QI_BASIS_COMPUTATIONAL, reg);
        } else {
            measured(q, q.count, QI_BASIS_COMPUTATIONAL, reg);
        }
        end_statement();
    }

    // ---- gates
    struct GateInfo {
        int kind;      // qi_gate_kind, kU or kSX
        int nparams;
        int ncontrols;  // built-in controls (cx: 1, ccx: 2)
        int ntargets;
    };
    static const std::map<std::string, GateInfo> &gate_table() {
        static const std::map<std::string, GateInfo> t = {
            {"id", {QI_GATE_I, 0, 0, 1}},    {"x", {QI_GATE_X, 0, 0, 1}},      {"y", {QI_GATE_Y, 0, 0, 1}},
            {"z", {QI_GATE_Z, 0, 0, 1}},     {"h", {QI_GATE_H, 0, 0, 1}},      {"s", {QI_GATE_S, 0, 0, 1}},
            {"sdg", {QI_GATE_SDAG, 0, 0, 1}}, {"t", {QI_GATE_T, 0, 0, 1}},     {"tdg", {QI_GATE_TDAG, 0, 0, 1}},
            {"sx", {kSX, 0, 0, 1}},           {"p", {QI_GATE_P, 1, 0, 1}},      {"phase", {QI_GATE_P, 1, 0, 1}},
            {"u1", {QI_GATE_P, 1, 0, 1}},     {"rx", {QI_GATE_RX, 1, 0, 1}},   {"ry", {QI_GATE_RY, 1, 0, 1}},
            {"rz", {QI_GATE_RZ, 1, 0, 1}},   {"U", {kU, 3, 0, 1}},             {"u3", {kU, 3, 0, 1}},
            {"u", {kU, 3, 0, 1}},             {"u2", {kU, 2, 0, 1}},            {"cx", {QI_GATE_X, 0, 1, 1}},
            {"CX", {QI_GATE_X, 0, 1, 1}},     {"cy", {QI_GATE_Y, 0, 1, 1}},     {"cz", {QI_GATE_Z, 0, 1, 1}},
            {"ch", {QI_GATE_H, 0, 1, 1}},     {"cp", {QI_GATE_P, 1, 1, 1}},     {"cphase", {QI_GATE_P, 1, 1, 1}},
            {"crx", {QI_GATE_RX, 1, 1, 1}},   {"cry", {QI_GATE_RY, 1, 1, 1}},   {"crz", {QI_GATE_RZ, 1, 1, 1}},
            {"cu", {kU, 4, 1, 1}},            {"swap", {QI_GATE_SWAP, 0, 0, 2}}, {"ccx", {QI_GATE_X, 0, 2, 1}},
            {"cswap", {QI_GATE_SWAP, 0, 1, 2}}};
        return t;
    }

    void gate_call() {
        // modifiers
        std::vector<int> ctrl_signs;  // +1 ctrl, -1 negctrl, in order
        bool inverse = false;
        std::vector<Operand> lead_controls;  // the exporter's `ctrl(k) @ q[c], ... g q[t]`
        while (true) {
            size_t save = pos_;
            if (!peek_word("ctrl") && !peek_word("negctrl") && !peek_word("inv") && !peek_word("pow")) break;
            std::string m = word();
            if (m == "pow") fail("the pow modifier is not supported");
            uint64_t k = 1;
            if (m != "inv" && at('(')) { pos_++; k = integer(); expect(')'); }
            if (!at('@')) { pos_ = save; break; }
            pos_++;
            if (m == "inv") { inverse = !inverse; continue; }
            for (uint64_t i = 0; i < k; i++) ctrl_signs.push_back(m == "ctrl" ? 1 : -1);
            // exporter form: the control operands follow the modifier
            size_t save2 = pos_;
            std::string nxt = word();
            pos_ = save2;
            auto r = regs_.find(nxt);
            if (r != regs_.end() && r->second.is_qubit) {
                std::vector<Operand> cs = operand_list();
                if (cs.size() != k) fail("ctrl(" + std::to_string(k) + ") takes " + std::to_string(k) + " control qubits");
                lead_controls.insert(lead_controls.end(), cs.begin(), cs.end());
            }
        }
        std::string name = word();
        auto gi = gate_table().find(name);
        if (gi == gate_table().end()) fail("unknown gate '" + name + "'");
        const GateInfo &G = gi->second;
        Val a[4] = {{0, -1}, {0, -1}, {0, -1}, {0, -1}};
        int np = 0;
        if (at('(')) {
            pos_++;
            if (!at(')')) {
                a[np++] = expr();
                while (at(',')) { pos_++; if (np == 4) fail("too many gate arguments"); a[np++] = expr(); }
            }
            expect(')');
        }
        if (np != G.nparams) fail("gate '" + name + "' takes " + std::to_string(G.nparams) + " arguments");
        std::vector<Operand> ops = operand_list();
        end_statement();
        const size_t nlead = lead_controls.size();
        const size_t nmod = ctrl_signs.size() - nlead;  // modifier controls taken from the operand list
        if (ops.size() != nmod + size_t(G.ncontrols + G.ntargets)) fail("gate '" + name + "' has the wrong number of qubits");
        std::vector<Operand> all = lead_controls;
        all.insert(all.end(), ops.begin(), ops.end());
        // broadcast over whole registers of equal size
        uint64_t width = 1;
        for (const Operand &o : all)
            if (o.whole) {
                if (width != 1 && o.count != width) fail("registers of different sizes in one gate");
                width = o.count;
            }
        flush_measure();
        for (uint64_t b = 0; b < width; b++) {
            std::vector<uint64_t> qs;
            for (const Operand &o : all) qs.push_back(o.whole ? o.first + b : o.first);
            const size_t nctl = ctrl_signs.size() + size_t(G.ncontrols);
            std::vector<uint64_t> controls(qs.begin(), qs.begin() + (long)nctl), targets(qs.begin() + (long)nctl, qs.end());
            std::vector<uint64_t> negs;
            for (size_t i = 0; i < ctrl_signs.size(); i++)
                if (ctrl_signs[i] < 0) negs.push_back(controls[i]);
            for (uint64_t q : negs) emit_simple(QI_GATE_X, {q}, {});
            emit(G, a, controls, targets, inverse);
            for (uint64_t q : negs) emit_simple(QI_GATE_X, {q}, {});
        }
    }

    void emit_simple(int kind, std::vector<uint64_t> t, std::vector<uint64_t> c, double p0 = 0.0, int slot = -1,
                     const double *mat = nullptr) {
        if (c.size() > QI_MAX_CONTROLS) fail("too many controls");
        qi_gate g;
        memset(&g, 0, sizeof g);
        g.kind = kind;
        g.ntargets = (uint32_t)t.size();
        for (size_t i = 0; i < t.size() && i < QI_MAX_TARGETS; i++) g.targets[i] = t[i];
        g.ncontrols = (uint32_t)c.size();
        for (size_t i = 0; i < c.size(); i++) g.controls[i] = c[i];
        g.params[0] = p0;
        if (mat) memcpy(g.matrix, mat, sizeof g.matrix);
        gates.push_back(g);
        slots.push_back(slot);
    }

    void emit(const GateInfo &G, const Val *a, const std::vector<uint64_t> &controls, const std::vector<uint64_t> &targets,
              bool inverse) {
        const double sg = inverse ? -1.0 : 1.0;
        auto rot = [&](int kind, const Val &v) {
            // a parameter slot carries the angle as given; an inverse would need -theta
            if (v.param >= 0 && inverse) fail("inv @ on a gate whose angle is a circuit parameter");
            emit_simple(kind, targets, controls, sg * v.v, v.param);
        };
        switch (G.kind) {
        case kU: {
            Val th = a[0], ph = a[1], la = a[2];
            if (G.nparams == 2) { th = {M_PI / 2, -1}; ph = a[0]; la = a[1]; }
            for (const Val *v : {&th, &ph, &la})
                if (v->param >= 0) fail("circuit parameters in U / u2 / cu are not supported");
            if (G.nparams == 4 && !inverse) emit_simple(QI_GATE_P, {controls.back()}, std::vector<uint64_t>(controls.begin(), controls.end() - 1), a[3].v);
            // U(θ, φ, λ) = P(φ) RY(θ) P(λ); its inverse is P(-λ) RY(-θ) P(-φ)
            if (!inverse) {
                emit_simple(QI_GATE_P, targets, controls, la.v);
                emit_simple(QI_GATE_RY, targets, controls, th.v);
                emit_simple(QI_GATE_P, targets, controls, ph.v);
            } else {
                emit_simple(QI_GATE_P, targets, controls, -ph.v);
                emit_simple(QI_GATE_RY, targets, controls, -th.v);
                emit_simple(QI_GATE_P, targets, controls, -la.v);
            }
            if (G.nparams == 4 && inverse) emit_simple(QI_GATE_P, {controls.back()}, std::vector<uint64_t>(controls.begin(), controls.end() - 1), -a[3].v);
            return;
        }
        case kSX: {  // sqrt(X) = [[1+i, 1-i], [1-i, 1+i]] / 2; its inverse conjugates every entry
            const double im = inverse ? -0.5 : 0.5;
            const double m[8] = {0.5, im, 0.5, -im, 0.5, -im, 0.5, im};
            emit_simple(QI_GATE_U2, targets, controls, 0.0, -1, m);
            return;
        }
        case QI_GATE_RX: case QI_GATE_RY: case QI_GATE_RZ: case QI_GATE_P:
            rot(G.kind, a[0]);
            return;
        case QI_GATE_S: emit_simple(inverse ? QI_GATE_SDAG : QI_GATE_S, targets, controls); return;
        case QI_GATE_SDAG: emit_simple(inverse ? QI_GATE_S : QI_GATE_SDAG, targets, controls); return;
        case QI_GATE_T: emit_simple(inverse ? QI_GATE_TDAG : QI_GATE_T, targets, controls); return;
        case QI_GATE_TDAG: emit_simple(inverse ? QI_GATE_T : QI_GATE_TDAG, targets, controls); return;
        default:  // I, X, Y, Z, H, SWAP: their own inverses
            emit_simple(G.kind, targets, controls);
            return;
        }
    }
};

}  // namespace

extern "C" qi_status qi_qasm_parse(const char *src, size_t len, const double *param_values, size_t nvalues,
                                   qi_gate *gates, int32_t *angle_slots, size_t capacity, size_t *ngates,
                                   uint32_t *nqubits, uint32_t *nparams, qi_error *err) {
    auto fail = [&](int code, uint64_t a, const std::string &msg) -> qi_status {
        if (err) {
            memset(err, 0, sizeof *err);
            err->code = code;
            err->a = a;
            snprintf(err->msg, sizeof err->msg, "%s", msg.c_str());
        }
        return (qi_status)code;
    };
    if (!src || !ngates) return fail(QI_ERR_INVALID_ARGUMENT, 0, "Invalid argument: src or ngates is NULL");
    Reader r(src, len, param_values, nvalues);
    try {
        r.run();
    } catch (const ParseError &e) {
        return fail(QI_ERR_INVALID_INPUT_VALUE, e.line, "Invalid input value: qasm line " + std::to_string(e.line) + ": " + e.msg);
    }
    *ngates = r.gates.size();
    if (nqubits) *nqubits = (uint32_t)r.nqubits;
    if (nparams) *nparams = r.nparams;
    if (gates) {
        if (capacity < r.gates.size()) return fail(QI_ERR_INVALID_ARGUMENT, r.gates.size(), "Invalid argument: gate capacity too small");
        memcpy(gates, r.gates.data(), r.gates.size() * sizeof(qi_gate));
        if (angle_slots) memcpy(angle_slots, r.slots.data(), r.slots.size() * sizeof(int32_t));
    }
    if (err) memset(err, 0, sizeof *err);
    return QI_OK;
}
