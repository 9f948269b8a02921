// qasm.cpp -- OpenQASM 3 subset lowered straight to the op array, so that a
// serving front end can submit circuit text without a Python round trip.
//
// The statements and angle forms of the reference importer
// (parse_qasm3_string / parse_angle, butterfly/utils/qasm3_import.mojo:7-123:
// qubit[n], h, x, p(a), cp(a); angles pi, c*pi, pi/d, c/pi, numbers) keep its
// results and error messages.  Extensions: named registers (q[i] resolves to
// the register's offset + i), y z s sdg t tdg rx ry rz u3 U u2 u1 phase cx cy cz
// crx cry crz ccx cphase cu1 swap mcp, the ctrl(k) @ modifier, measurements,
// barrier / bit declarations (ignored), statements split on ';', and an
// arithmetic angle expression (+ - * / parentheses, pi) for the forms the
// reference rejects.  Gates lower exactly as lowering.py lowers the
// QuantumCircuit the Python mirror builds (same kinds, angles, and matrices
// formed with CPython's complex product, so the op arrays match bit for bit).
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "bfly_plan.h"

namespace bfly {
int fail(int code, const std::string& msg);
}

namespace {

using bfly::fail;

struct Cx {
    double r, i;
};
// CPython's _Py_c_prod (a float operand enters as (x, 0.0)).
Cx cprod(Cx a, Cx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
Cx cis(double t) { return {std::cos(t), std::sin(t)}; }

struct Err {
    int code;
    std::string msg;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }
bool ends(const std::string& s, const char* p) {
    const size_t n = std::strlen(p);
    return s.size() >= n && s.compare(s.size() - n, n, p) == 0;
}

// Float64(String): the whole (trimmed) token must be a number.
double to_f64(const std::string& t) {
    const std::string s = trim(t);
    if (s.empty()) throw Err{BFLY_EINVAL, "Invalid number: " + t};
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size()) throw Err{BFLY_EINVAL, "Invalid number: " + s};
    return v;
}

// Arithmetic angle expression: + - * / unary minus, parentheses, pi, numbers.
struct Expr {
    const std::string& s;
    size_t i = 0;
    explicit Expr(const std::string& str) : s(str) {}
    void ws() { while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i; }
    double primary() {
        ws();
        if (i < s.size() && s[i] == '(') {
            ++i;
            const double v = sum();
            ws();
            if (i >= s.size() || s[i] != ')') throw Err{BFLY_EINVAL, "Invalid angle: " + s};
            ++i;
            return v;
        }
        if (i < s.size() && s[i] == '-') { ++i; return -primary(); }
        if (i < s.size() && s[i] == '+') { ++i; return primary(); }
        if (s.compare(i, 2, "pi") == 0) { i += 2; return M_PI; }
        const size_t b = i;
        while (i < s.size() && (std::isdigit(static_cast<unsigned char>(s[i])) || s[i] == '.' || s[i] == 'e' ||
                                s[i] == 'E' || ((s[i] == '-' || s[i] == '+') && i > b && (s[i - 1] == 'e' || s[i - 1] == 'E'))))
            ++i;
        if (i == b) throw Err{BFLY_EINVAL, "Invalid angle: " + s};
        return to_f64(s.substr(b, i - b));
    }
    double product() {
        double v = primary();
        for (;;) {
            ws();
            if (i < s.size() && s[i] == '*') { ++i; v = v * primary(); }
            else if (i < s.size() && s[i] == '/') { ++i; v = v / primary(); }
            else return v;
        }
    }
    double sum() {
        double v = product();
        for (;;) {
            ws();
            if (i < s.size() && s[i] == '+') { ++i; v = v + product(); }
            else if (i < s.size() && s[i] == '-') { ++i; v = v - product(); }
            else return v;
        }
    }
};

// parse_angle (qasm3_import.mojo:7-30), then the expression form.
double parse_angle(const std::string& expr) {
    const std::string trimmed = trim(expr);
    if (trimmed.find("pi") != std::string::npos) {
        std::string s;
        for (char ch : trimmed)
            if (ch != ' ') s += ch;
        double sign = 1.0;
        if (starts(s, "-")) { sign = -1.0; s = s.substr(1); }
        try {
            if (s == "pi") return sign * M_PI;
            if (ends(s, "*pi")) return sign * to_f64(s.substr(0, s.size() - 3)) * M_PI;
            if (starts(s, "pi/")) return sign * M_PI / to_f64(s.substr(3));
            const size_t p = s.find("/pi");
            if (p != std::string::npos && s.find("/pi", p + 1) == std::string::npos && p + 3 == s.size())
                return sign * to_f64(s.substr(0, p)) / M_PI;
        } catch (const Err&) {
        }
    } else {
        try {
            return to_f64(trimmed);
        } catch (const Err&) {
        }
    }
    Expr e(trimmed);
    const double v = e.sum();
    e.ws();
    if (e.i != trimmed.size()) throw Err{BFLY_EINVAL, "Invalid angle: " + trimmed};
    return v;
}

struct Program {
    int n = 0;
    bool has_circuit = false;
    std::map<std::string, std::pair<int, int>> regs;   // name -> (offset, size)
    std::vector<bfly_op> ops;
    std::vector<double> udata;                        // (qubit, value) pairs of measurements
};

// parse_qubit_index (qasm3_import.mojo:43-50) with the register's offset.
int qubit_index(const Program& P, const std::string& token) {
    std::string t = trim(token);
    if (ends(t, ";")) t = trim(t.substr(0, t.size() - 1));
    if (ends(t, ",")) t = trim(t.substr(0, t.size() - 1));
    const size_t lb = t.find('['), rb = t.find(']');
    if (lb == std::string::npos || rb == std::string::npos || rb <= lb) throw Err{BFLY_EINVAL, "Invalid qubit token: " + t};
    const std::string name = trim(t.substr(0, lb));
    const std::string idx = trim(t.substr(lb + 1, rb - lb - 1));
    char* end = nullptr;
    const long v = std::strtol(idx.c_str(), &end, 10);
    if (idx.empty() || end != idx.c_str() + idx.size()) throw Err{BFLY_EINVAL, "Invalid qubit token: " + t};
    int offset = 0;
    if (P.regs.size() > 1) {
        auto it = P.regs.find(name);
        if (it == P.regs.end()) throw Err{BFLY_EINVAL, "Unknown register: " + name};
        offset = it->second.first;
    }
    return offset + static_cast<int>(v);
}

std::vector<std::string> split_operands(const std::string& rest) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : rest) {
        if (ch == ',') { out.push_back(trim(cur)); cur.clear(); }
        else cur += ch;
    }
    if (!trim(cur).empty() || !out.empty()) out.push_back(trim(cur));
    return out;
}

void check_q(const Program& P, int q, const char* label) {
    if (q < 0 || q >= P.n) throw Err{BFLY_EINVAL, std::string(label) + " out of bounds: " + std::to_string(q)};
}

// GateTransformation -> GATE record, as lowering._lower_gate does.
void emit_gate(Program& P, const std::vector<int>& controls, int target, int kind, const double m[8], double theta) {
    check_q(P, target, "Target");
    for (int c : controls) check_q(P, c, "Control");
    for (int c : controls)
        if (c == target) return;   // never fires in the reference's scalar kernels
    bfly_op op{};
    op.kind = BFLY_OP_GATE;
    op.gate = kind;
    op.target = target;
    for (int c : controls) op.controls |= 1ull << c;
    op.theta = theta;
    std::memcpy(op.mat, m, sizeof op.mat);
    P.ops.push_back(op);
}

void mat_fixed(int kind, double m[8]) {
    const double h = std::sqrt(0.5);
    const double H[8] = {h, 0.0, h, 0.0, h, 0.0, -h, 0.0};
    const double X[8] = {0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0};
    const double Y[8] = {0.0, 0.0, -0.0, -1.0, 0.0, 1.0, 0.0, 0.0};
    const double Z[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, -1.0, 0.0};
    const double* src = kind == BFLY_GATE_H ? H : kind == BFLY_GATE_X ? X : kind == BFLY_GATE_Y ? Y : Z;
    std::memcpy(m, src, 8 * sizeof(double));
}

void mat_param(int kind, double t, double m[8]) {
    if (kind == BFLY_GATE_P) {
        const Cx z = cis(t);
        const double v[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, z.r, z.i};
        std::memcpy(m, v, sizeof v);
        return;
    }
    const double c = std::cos(t / 2), s = std::sin(t / 2);
    if (kind == BFLY_GATE_RX) {
        const double v[8] = {c, 0.0, 0.0, -s, 0.0, -s, c, 0.0};
        std::memcpy(m, v, sizeof v);
    } else if (kind == BFLY_GATE_RY) {
        const double v[8] = {c, 0.0, -s, 0.0, s, 0.0, c, 0.0};
        std::memcpy(m, v, sizeof v);
    } else {   // RZ
        const double v[8] = {c, -s, 0.0, 0.0, 0.0, 0.0, c, s};
        std::memcpy(m, v, sizeof v);
    }
}

// U3 (gates.mojo:42-53) as the Python mirror computes it.
void mat_u3(double theta, double phi, double lam, double m[8]) {
    const double c = std::cos(theta / 2), s = std::sin(theta / 2);
    const Cx l = cis(lam);
    const Cx a = cprod(Cx{-l.r, -l.i}, Cx{s, 0.0});
    const Cx b = cprod(cis(phi), Cx{s, 0.0});
    const Cx d = cprod(cis(phi + lam), Cx{c, 0.0});
    const double v[8] = {c, 0.0, a.r, a.i, b.r, b.i, d.r, d.i};
    std::memcpy(m, v, sizeof v);
}

std::vector<double> parse_args(const std::string& inside) {
    std::vector<double> out;
    if (trim(inside).empty()) return out;
    int depth = 0;
    std::string cur;
    for (char ch : inside) {
        if (ch == '(') ++depth;
        if (ch == ')') --depth;
        if (ch == ',' && depth == 0) { out.push_back(parse_angle(cur)); cur.clear(); }
        else cur += ch;
    }
    out.push_back(parse_angle(cur));
    return out;
}

void statement(Program& P, std::string s) {
    if (s.empty() || starts(s, "OPENQASM") || starts(s, "include")) return;
    if (starts(s, "qubit") || starts(s, "qreg")) {
        // qubit[n] name | qreg name[n]
        const size_t lb = s.find('['), rb = s.find(']');
        if (lb == std::string::npos || rb == std::string::npos) throw Err{BFLY_EINVAL, "Invalid qubit declaration: " + s};
        const int size = static_cast<int>(to_f64(s.substr(lb + 1, rb - lb - 1)));
        std::string name = starts(s, "qreg") ? trim(s.substr(4, lb - 4)) : trim(s.substr(rb + 1));
        if (P.regs.count(name)) throw Err{BFLY_EINVAL, "Register declared twice: " + name};
        P.regs[name] = {P.n, size};
        P.n += size;
        P.has_circuit = true;
        if (P.n > 64) throw Err{BFLY_EINVAL, "more than 64 qubits"};
        return;
    }
    // measurements: measure q[i] | c[i] = measure q[i] | measure q[i] -> c[i] |
    // bit[k] c = measure q (a whole register)
    {
        std::string m = s;
        const size_t eq = m.find('=');
        if (eq != std::string::npos && trim(m.substr(eq + 1)).rfind("measure", 0) == 0) m = trim(m.substr(eq + 1));
        if (starts(m, "measure ") || starts(m, "measure\t")) {
            if (!P.has_circuit) throw Err{BFLY_EINVAL, "QASM missing qubit declaration before gates."};
            std::string rest = trim(m.substr(7));
            const size_t arrow = rest.find("->");
            if (arrow != std::string::npos) rest = trim(rest.substr(0, arrow));
            std::vector<int> targets;
            if (rest.find('[') == std::string::npos) {
                auto it = P.regs.find(rest);
                if (it == P.regs.end()) throw Err{BFLY_EINVAL, "Invalid qubit token: " + rest};
                for (int i = 0; i < it->second.second; ++i) targets.push_back(it->second.first + i);
            } else {
                targets.push_back(qubit_index(P, rest));
            }
            bfly_op op{};
            op.kind = BFLY_OP_MEASURE;
            op.arg_i = static_cast<int32_t>(targets.size());
            op.aux = P.udata.size() / 2;
            for (int q : targets) {
                check_q(P, q, "Target");
                P.udata.push_back(static_cast<double>(q));
                P.udata.push_back(-1.0);
            }
            P.ops.push_back(op);
            return;
        }
    }
    if (starts(s, "bit") || starts(s, "creg") || starts(s, "barrier")) return;
    if (!P.has_circuit) throw Err{BFLY_EINVAL, "QASM missing qubit declaration before gates."};

    // [ctrl(k) @] name[(args)] operands
    int nctrl = 0;
    if (starts(s, "ctrl")) {
        const size_t at = s.find('@');
        if (at == std::string::npos) throw Err{BFLY_EINVAL, "Unsupported QASM line: " + s};
        std::string mod = trim(s.substr(4, at - 4));
        if (mod.empty()) nctrl = 1;
        else if (mod.front() == '(' && mod.back() == ')') nctrl = static_cast<int>(to_f64(mod.substr(1, mod.size() - 2)));
        else throw Err{BFLY_EINVAL, "Unsupported QASM line: " + s};
        s = trim(s.substr(at + 1));
    }
    size_t k = 0;
    while (k < s.size() && (std::isalnum(static_cast<unsigned char>(s[k])) || s[k] == '_')) ++k;
    const std::string name = s.substr(0, k);
    std::string rest = s.substr(k);
    std::vector<double> args;
    if (!trim(rest).empty() && trim(rest)[0] == '(') {
        rest = trim(rest);
        int depth = 0;
        size_t close = std::string::npos;
        for (size_t i = 0; i < rest.size(); ++i) {
            if (rest[i] == '(') ++depth;
            if (rest[i] == ')' && --depth == 0) { close = i; break; }
        }
        if (close == std::string::npos) throw Err{BFLY_EINVAL, "Invalid " + name + " gate: " + s};
        args = parse_args(rest.substr(1, close - 1));
        rest = rest.substr(close + 1);
    }
    std::vector<std::string> operands = split_operands(rest);
    std::vector<int> q;
    for (const std::string& o : operands) q.push_back(qubit_index(P, o));

    auto need = [&](size_t nargs, size_t nq) {
        if (args.size() != nargs) throw Err{BFLY_EINVAL, "Invalid " + name + " gate: " + s};
        if (q.size() != nq + static_cast<size_t>(nctrl)) {
            if (name == "cp" && nctrl == 0) throw Err{BFLY_EINVAL, "Invalid cp operands: " + trim(rest)};
            throw Err{BFLY_EINVAL, "Invalid " + name + " operands: " + trim(rest)};
        }
    };
    std::vector<int> ctl(q.begin(), q.begin() + std::min<size_t>(q.size(), static_cast<size_t>(nctrl)));
    double m[8];
    auto one = [&](int kind, double theta) {   // single-qubit gate, optional ctrl @ controls
        emit_gate(P, ctl, q.back(), kind, m, theta);
    };
    auto ctl_plus = [&](size_t extra) {          // leading operands that are controls
        std::vector<int> c = ctl;
        for (size_t i = 0; i < extra; ++i) c.push_back(q[static_cast<size_t>(nctrl) + i]);
        return c;
    };
    if (name == "h" || name == "x" || name == "y" || name == "z") {
        need(0, 1);
        const int kind = name == "h" ? BFLY_GATE_H : name == "x" ? BFLY_GATE_X : name == "y" ? BFLY_GATE_Y : BFLY_GATE_Z;
        mat_fixed(kind, m);
        one(kind, 0.0);
    } else if (name == "p" || name == "phase" || name == "u1" || name == "s" || name == "sdg" || name == "t" ||
               name == "tdg") {
        double th;
        if (name == "s" || name == "sdg" || name == "t" || name == "tdg") {
            need(0, 1);
            th = name == "s" ? M_PI / 2 : name == "sdg" ? -M_PI / 2 : name == "t" ? M_PI / 4 : -M_PI / 4;
        } else {
            need(1, 1);
            th = args[0];
        }
        mat_param(BFLY_GATE_P, th, m);
        one(BFLY_GATE_P, th);
    } else if (name == "rx" || name == "ry" || name == "rz") {
        need(1, 1);
        const int kind = name == "rx" ? BFLY_GATE_RX : name == "ry" ? BFLY_GATE_RY : BFLY_GATE_RZ;
        mat_param(kind, args[0], m);
        one(kind, args[0]);
    } else if (name == "u3" || name == "U" || name == "u" || name == "u2") {
        if (name == "u2") {
            need(2, 1);
            mat_u3(M_PI / 2, args[0], args[1], m);
        } else {
            need(3, 1);
            mat_u3(args[0], args[1], args[2], m);
        }
        one(BFLY_GATE_CUSTOM, 0.0);
    } else if (name == "cx" || name == "CX" || name == "cnot" || name == "cy" || name == "cz") {
        need(0, 2);
        const int kind = name == "cy" ? BFLY_GATE_Y : name == "cz" ? BFLY_GATE_Z : BFLY_GATE_X;
        mat_fixed(kind, m);
        emit_gate(P, ctl_plus(1), q.back(), kind, m, 0.0);
    } else if (name == "cp" || name == "cphase" || name == "cu1") {
        need(1, 2);
        mat_param(BFLY_GATE_P, args[0], m);
        emit_gate(P, ctl_plus(1), q.back(), BFLY_GATE_P, m, args[0]);
    } else if (name == "crx" || name == "cry" || name == "crz") {
        need(1, 2);
        const int kind = name == "crx" ? BFLY_GATE_RX : name == "cry" ? BFLY_GATE_RY : BFLY_GATE_RZ;
        mat_param(kind, args[0], m);
        emit_gate(P, ctl_plus(1), q.back(), kind, m, args[0]);
    } else if (name == "ccx" || name == "toffoli") {
        need(0, 3);
        mat_fixed(BFLY_GATE_X, m);
        emit_gate(P, ctl_plus(2), q.back(), BFLY_GATE_X, m, 0.0);
    } else if (name == "mcp") {
        if (args.size() != 1 || q.size() < 2) throw Err{BFLY_EINVAL, "Invalid mcp gate: " + s};
        mat_param(BFLY_GATE_P, args[0], m);
        emit_gate(P, std::vector<int>(q.begin(), q.end() - 1), q.back(), BFLY_GATE_P, m, args[0]);
    } else if (name == "swap" && nctrl == 0) {
        need(0, 2);
        check_q(P, q[0], "Target");
        check_q(P, q[1], "Target");
        bfly_op op{};
        op.kind = BFLY_OP_SWAP;
        op.target = q[0];
        op.arg_i = q[1];
        P.ops.push_back(op);
    } else {
        throw Err{BFLY_EINVAL, "Unsupported QASM line: " + s};
    }
}

}  // namespace

extern "C" int bfly_qasm3_lower(const char* text, int32_t* num_qubits, bfly_op* ops, uint64_t ops_cap, uint64_t* nops,
                                double* udata, uint64_t udata_cap, uint64_t* udata_len) {
    if (!text || !num_qubits || !nops || !udata_len) return fail(BFLY_EINVAL, "null argument");
    Program P;
    try {
        // lines (comments end at the line), then statements split on ';'
        const std::string all(text);
        size_t pos = 0;
        while (pos <= all.size()) {
            size_t nl = all.find('\n', pos);
            if (nl == std::string::npos) nl = all.size();
            std::string line = all.substr(pos, nl - pos);
            pos = nl + 1;
            const size_t cm = line.find("//");
            if (cm != std::string::npos) line = line.substr(0, cm);
            size_t a = 0;
            while (a <= line.size()) {
                size_t b = line.find(';', a);
                if (b == std::string::npos) b = line.size();
                statement(P, trim(line.substr(a, b - a)));
                a = b + 1;
            }
        }
        if (!P.has_circuit) throw Err{BFLY_EINVAL, "No qubit declaration found in QASM."};
    } catch (const Err& e) {
        return fail(e.code, e.msg);
    }
    *num_qubits = P.n;
    *nops = P.ops.size();
    *udata_len = P.udata.size() / 2;
    if (ops) {
        if (ops_cap < P.ops.size()) return fail(BFLY_EINVAL, "op buffer too small");
        std::memcpy(ops, P.ops.data(), P.ops.size() * sizeof(bfly_op));
    }
    if (udata && !P.udata.empty()) {
        if (udata_cap < P.udata.size() / 2) return fail(BFLY_EINVAL, "udata buffer too small");
        std::memcpy(udata, P.udata.data(), P.udata.size() * sizeof(double));
    }
    return BFLY_OK;
}
