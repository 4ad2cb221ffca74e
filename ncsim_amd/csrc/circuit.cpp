// Circuits and the Stim-like text format (src/circuit.rs): storage with the
// push checks of Circuit::push (177-229), the parser (459-691) and the
// printer (693-728).
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <sstream>

#include "ncsv_internal.hpp"

namespace ncsv {

// Circuit::push (circuit.rs:177-229).
void circuit_push(ncsv_circuit *c, COp op) {
    switch (op.kind) {
    case CKind::Gate: {
        check_qubits(op.qubits.data(), (uint32_t)op.qubits.size(), c->n);
        const uint32_t ar = gate_arity(op.g);
        if (op.qubits.size() != ar)
            fail(NCSV_ARITY, "gate " + gate_name(op.g) + " acts on " + std::to_string(ar) + " qubits, " +
                                 std::to_string(op.qubits.size()) + " given");
        break;
    }
    case CKind::Measure:
        if (!pauli_is_hermitian(op.x, op.z, op.ph)) fail(NCSV_INVALID, "measured operator is not Hermitian");
        if (c->n < 64 && ((op.x | op.z) >> c->n)) fail(NCSV_INVALID, "measurement on the wrong register");
        c->measurements++;
        break;
    case CKind::Reset: check_qubits(op.qubits.data(), 1, c->n); break;
    case CKind::Noise: {
        check_qubits(op.qubits.data(), (uint32_t)op.qubits.size(), c->n);
        if (op.qubits.size() != noise_arity(op.ch))
            fail(NCSV_ARITY, "gate " + noise_name(op.ch, op.p) + " acts on " + std::to_string(noise_arity(op.ch)) +
                                 " qubits, " + std::to_string(op.qubits.size()) + " given");
        double tot = 0;
        for (auto &e : noise_errors(op.ch, op.p)) {
            if (!(e.p >= 0.0 && e.p <= 1.0)) fail(NCSV_INVALID, noise_name(op.ch, op.p) + " is not a probability distribution");
            tot += e.p;
        }
        if (tot > 1.0 + 1e-12) fail(NCSV_INVALID, noise_name(op.ch, op.p) + " is not a probability distribution");
        break;
    }
    case CKind::Detector:
    case CKind::Observable:
        for (uint32_t k : op.lookback)
            if (k == 0 || k > c->measurements)
                fail(NCSV_INVALID, "rec[-" + std::to_string(k) + "] with " + std::to_string(c->measurements) +
                                       " measurements so far");
        break;
    case CKind::Tick: break;
    }
    c->ops.push_back(std::move(op));
}

namespace {

enum class PK { Gate, MBasis, Mpp, Reset, Noise, Detector, Observable, Tick };
struct Parsed {
    PK kind;
    GateDesc g;
    char basis = 'Z';
    std::vector<uint32_t> qs;
    std::vector<PauliM> paulis;
    ncsv_noise ch = NCSV_NOISE_X_ERROR;
    double p[3] = {0, 0, 0};
    std::vector<uint32_t> recs;
    uint32_t index = 0;
};

std::string upper(std::string s) {
    for (auto &c : s) c = (char)toupper((unsigned char)c);
    return s;
}

bool parse_f64(const std::string &s, double *v) {
    if (s.empty()) return false;
    char *end = nullptr;
    *v = strtod(s.c_str(), &end);
    return end && *end == 0;
}

bool parse_i64(const std::string &s, int64_t *v) {
    if (s.empty()) return false;
    char *end = nullptr;
    *v = strtoll(s.c_str(), &end, 10);
    return end && *end == 0;
}

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// parse_args (circuit.rs:557-569)
void parse_args(const std::string &s, std::string *name, std::vector<std::string> *args) {
    const size_t i = s.find('(');
    if (i == std::string::npos) {
        *name = upper(s);
        return;
    }
    *name = upper(s.substr(0, i));
    const std::string rest = s.substr(i + 1);
    const size_t close = rest.rfind(')');
    if (close == std::string::npos) throw std::string("missing ')' in " + s);
    std::string inner = rest.substr(0, close), cur;
    for (char ch : inner) {
        if (ch == ',') {
            args->push_back(trim(cur));
            cur.clear();
        } else
            cur.push_back(ch);
    }
    args->push_back(trim(cur));
}

// parse_line (circuit.rs:571-691)
Parsed parse_line(const std::vector<std::string> &toks) {
    std::string name;
    std::vector<std::string> args;
    parse_args(toks[0], &name, &args);
    auto qubits = [&]() {
        std::vector<uint32_t> v;
        for (size_t i = 1; i < toks.size(); i++) {
            int64_t q;
            if (!parse_i64(toks[i], &q) || q < 0 || q > 0xffffffffLL) throw std::string("bad qubit '" + toks[i] + "'");
            v.push_back((uint32_t)q);
        }
        return v;
    };
    auto int_arg = [&](size_t i) {
        int64_t v;
        if (i >= args.size() || !parse_i64(args[i], &v))
            throw std::string(name + " needs integer argument " + std::to_string(i + 1));
        return v;
    };
    auto float_arg = [&](size_t i) {
        double v;
        if (i >= args.size() || !parse_f64(args[i], &v)) throw std::string(name + " needs a numeric argument");
        return v;
    };
    Parsed p;
    ncsv_gate id;
    if (name != "UNITARY" && name != "PAULI" && gate_lookup(name, &id)) {
        p.kind = PK::Gate;
        if (id == NCSV_GATE_PHASE || id == NCSV_GATE_CPHASE) {
            const int64_t ip[2] = {int_arg(0), int_arg(1)};
            try {
                p.g = make_gate(id, ip, nullptr, 0);
            } catch (Error &e) {
                throw e.msg;
            }
        } else if (id == NCSV_GATE_RX || id == NCSV_GATE_RY || id == NCSV_GATE_RZ) {
            const double fp[1] = {float_arg(0)};
            p.g = make_gate(id, nullptr, fp, 0);
        } else {
            p.g = make_gate(id, nullptr, nullptr, 0);
        }
        p.qs = qubits();
        return p;
    }
    if (name == "M" || name == "MZ" || name == "MX" || name == "MY") {
        p.kind = PK::MBasis;
        p.basis = name == "MX" ? 'X' : (name == "MY" ? 'Y' : 'Z');
        p.qs = qubits();
        return p;
    }
    if (name == "MPP") {
        p.kind = PK::Mpp;
        for (size_t i = 1; i < toks.size(); i++) {
            try {
                p.paulis.push_back(pauli_parse(toks[i], -1));
            } catch (Error &e) {
                throw e.msg;
            }
        }
        return p;
    }
    if (name == "R") {
        p.kind = PK::Reset;
        p.qs = qubits();
        return p;
    }
    if (name == "DETECTOR" || name == "OBSERVABLE_INCLUDE") {
        for (size_t i = 1; i < toks.size(); i++) {
            const std::string &t = toks[i];
            int64_t k;
            if (t.size() < 7 || t.compare(0, 5, "rec[-") != 0 || t.back() != ']' ||
                !parse_i64(t.substr(5, t.size() - 6), &k) || k < 0)
                throw std::string("bad target '" + t + "' (expected rec[-k])");
            p.recs.push_back((uint32_t)k);
        }
        if (name == "DETECTOR") {
            p.kind = PK::Detector;
        } else {
            p.kind = PK::Observable;
            p.index = (uint32_t)int_arg(0);
        }
        return p;
    }
    struct {
        const char *nm;
        ncsv_noise ch;
    } chans[] = {{"X_ERROR", NCSV_NOISE_X_ERROR},
                 {"Y_ERROR", NCSV_NOISE_Y_ERROR},
                 {"Z_ERROR", NCSV_NOISE_Z_ERROR},
                 {"DEPOLARIZE1", NCSV_NOISE_DEPOLARIZE1},
                 {"DEPOLARIZE2", NCSV_NOISE_DEPOLARIZE2}};
    for (auto &c : chans)
        if (name == c.nm) {
            p.kind = PK::Noise;
            p.ch = c.ch;
            p.p[0] = float_arg(0);
            p.qs = qubits();
            return p;
        }
    if (name == "PAULI_CHANNEL_1") {
        p.kind = PK::Noise;
        p.ch = NCSV_NOISE_PAULI_CHANNEL_1;
        for (int i = 0; i < 3; i++) {
            double v;
            if ((size_t)i >= args.size() || !parse_f64(args[i], &v))
                throw std::string("PAULI_CHANNEL_1 needs three probabilities");
            p.p[i] = v;
        }
        p.qs = qubits();
        return p;
    }
    if (name == "TICK") {
        p.kind = PK::Tick;
        return p;
    }
    throw std::string("unknown instruction '" + name + "'");
}

}  // namespace

}  // namespace ncsv

using namespace ncsv;

// Circuit::parse (circuit.rs:459-555), into an allocated circuit.
ncsv_circuit *ncsv_parse_text(const char *text) {
    std::istringstream in(text ? text : "");
    std::string raw;
    std::vector<std::pair<size_t, std::vector<std::string>>> lines;
    bool have_decl = false;
    uint32_t declared = 0;
    size_t ln = 0;
    auto lerr = [](size_t l, const std::string &m) { fail(NCSV_PARSE, "line " + std::to_string(l + 1) + ": " + m); };
    while (std::getline(in, raw)) {
        std::string line = trim(raw.substr(0, raw.find('#')));
        if (!line.empty()) {
            std::vector<std::string> toks;
            std::istringstream ts(line);
            std::string t;
            while (ts >> t) toks.push_back(t);
            if (upper(toks[0]) == "QUBITS") {
                int64_t nv;
                if (toks.size() < 2 || !parse_i64(toks[1], &nv) || nv < 0) lerr(ln, "QUBITS n");
                declared = (uint32_t)nv;
                have_decl = true;
            } else {
                lines.push_back({ln, toks});
            }
        }
        ln++;
    }
    uint32_t max_q = 0;
    std::vector<std::pair<size_t, Parsed>> parsed;
    for (auto &l : lines) {
        Parsed p;
        try {
            p = parse_line(l.second);
        } catch (std::string &m) {
            lerr(l.first, m);
        }
        if (p.kind == PK::Gate || p.kind == PK::MBasis || p.kind == PK::Reset || p.kind == PK::Noise)
            for (uint32_t q : p.qs) max_q = std::max(max_q, q + 1);
        if (p.kind == PK::Mpp)
            for (auto &pp : p.paulis) max_q = std::max(max_q, pp.n);
        parsed.push_back({l.first, std::move(p)});
    }
    const uint32_t n = have_decl ? declared : max_q;
    if (n < max_q) fail(NCSV_PARSE, "QUBITS " + std::to_string(n) + " but qubit " + std::to_string(max_q - 1) + " is used");
    auto c = std::make_unique<ncsv_circuit>();
    c->n = n;
    for (auto &lp : parsed) {
        Parsed &p = lp.second;
        try {
            switch (p.kind) {
            case PK::Gate: {
                const uint32_t k = gate_arity(p.g);
                if (k == 0 || p.qs.size() % k)
                    lerr(lp.first, gate_name(p.g) + " takes groups of " + std::to_string(k) + " qubits");
                for (size_t i = 0; i < p.qs.size(); i += k) {
                    COp op;
                    op.kind = CKind::Gate;
                    op.g = p.g;
                    op.qubits.assign(p.qs.begin() + i, p.qs.begin() + i + k);
                    circuit_push(c.get(), op);
                }
                break;
            }
            case PK::MBasis:
                for (uint32_t q : p.qs) {
                    COp op;
                    op.kind = CKind::Measure;
                    if (q >= 64) lerr(lp.first, "qubit beyond 64");
                    if (p.basis != 'Z') op.x = 1ull << q;
                    if (p.basis != 'X') op.z = 1ull << q;
                    op.ph = p.basis == 'Y' ? 1 : 0;
                    circuit_push(c.get(), op);
                }
                break;
            case PK::Mpp:
                for (auto &pp : p.paulis) {
                    // Pauli::from_string(n, p.to_sparse_string()) (circuit.rs:511-519)
                    PauliM q = pauli_parse(pauli_sparse_string(pp), (int)n);
                    COp op;
                    op.kind = CKind::Measure;
                    op.x = q.x;
                    op.z = q.z;
                    op.ph = q.ph;
                    circuit_push(c.get(), op);
                }
                break;
            case PK::Reset:
                for (uint32_t q : p.qs) {
                    COp op;
                    op.kind = CKind::Reset;
                    op.qubits = {q};
                    circuit_push(c.get(), op);
                }
                break;
            case PK::Noise: {
                const uint32_t k = noise_arity(p.ch);
                if (p.qs.size() % k)
                    lerr(lp.first, noise_name(p.ch, p.p) + " takes groups of " + std::to_string(k) + " qubits");
                for (size_t i = 0; i < p.qs.size(); i += k) {
                    COp op;
                    op.kind = CKind::Noise;
                    op.ch = p.ch;
                    std::copy(p.p, p.p + 3, op.p);
                    op.qubits.assign(p.qs.begin() + i, p.qs.begin() + i + k);
                    circuit_push(c.get(), op);
                }
                break;
            }
            case PK::Detector:
            case PK::Observable: {
                COp op;
                op.kind = p.kind == PK::Detector ? CKind::Detector : CKind::Observable;
                op.lookback = p.recs;
                op.index = p.index;
                circuit_push(c.get(), op);
                break;
            }
            case PK::Tick: {
                COp op;
                op.kind = CKind::Tick;
                circuit_push(c.get(), op);
                break;
            }
            }
        } catch (Error &e) {
            if (e.code == NCSV_PARSE) throw;
            lerr(lp.first, e.msg);
        }
    }
    return c.release();
}

// Circuit's Display (circuit.rs:693-728).
std::string ncsv_circuit_text(const ncsv_circuit *c) {
    std::ostringstream o;
    o << "QUBITS " << c->n << "\n";
    for (auto &op : c->ops) {
        switch (op.kind) {
        case CKind::Gate:
            o << gate_name(op.g);
            for (auto q : op.qubits) o << " " << q;
            o << "\n";
            break;
        case CKind::Measure: {
            const uint64_t sup = op.x | op.z;
            const uint32_t ys = (uint32_t)(__builtin_popcountll(op.x & op.z) % 4);
            const uint32_t disp = (op.ph + 4 - ys) & 3;
            if (__builtin_popcountll(sup) == 1 && disp == 0) {
                const int q = __builtin_ctzll(sup);
                const bool xb = (op.x >> q) & 1, zb = (op.z >> q) & 1;
                o << "M" << ((xb && zb) ? 'Y' : (xb ? 'X' : 'Z')) << " " << q << "\n";
            } else {
                PauliM p;
                p.n = c->n;
                p.x = op.x;
                p.z = op.z;
                p.ph = op.ph;
                o << "MPP " << pauli_sparse_string(p) << "\n";
            }
            break;
        }
        case CKind::Reset: o << "R " << op.qubits[0] << "\n"; break;
        case CKind::Noise:
            o << noise_name(op.ch, op.p);
            for (auto q : op.qubits) o << " " << q;
            o << "\n";
            break;
        case CKind::Detector:
            o << "DETECTOR";
            for (auto k : op.lookback) o << " rec[-" << k << "]";
            o << "\n";
            break;
        case CKind::Observable:
            o << "OBSERVABLE_INCLUDE(" << op.index << ")";
            for (auto k : op.lookback) o << " rec[-" << k << "]";
            o << "\n";
            break;
        case CKind::Tick: o << "TICK\n"; break;
        }
    }
    return o.str();
}
