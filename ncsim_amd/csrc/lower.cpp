// Host-side gate views and lowering: the named-gate table of src/gate.rs
// restated as what the dense kernels consume (dense matrix, diagonal phases,
// permutation with phases, Pauli masks), plus Pauli parsing and the noise
// tables.
#include <cmath>
#include <cstring>
#include <sstream>

#include "ncsv_internal.hpp"

namespace ncsv {

[[noreturn]] void fail(ncsv_status code, const std::string &msg) { throw Error{code, msg}; }

// quarter_turn (scalar.rs:101-108)
static cd quarter_turn(int64_t j) {
    switch (((j % 4) + 4) % 4) {
    case 0: return cd(1.0, 0.0);
    case 1: return cd(0.0, 1.0);
    case 2: return cd(-1.0, 0.0);
    default: return cd(0.0, -1.0);
    }
}

// C64::root_of_unity (scalar.rs:112-132): exact at multiples of pi/4.
cd root_of_unity(int64_t k, uint32_t bits) {
    if (bits > 62) fail(NCSV_INVALID, "phase bits must be at most 62");
    const __int128 m = (__int128)1 << bits;
    __int128 kk = ((__int128)k % m + m) % m;
    if ((4 * kk) % m == 0) return quarter_turn((int64_t)(4 * kk / m));
    if ((8 * kk) % m == 0) {
        const double h = 0.70710678118654752440;  // FRAC_1_SQRT_2
        switch ((int)(8 * kk / m)) {
        case 1: return cd(h, h);
        case 3: return cd(-h, h);
        case 5: return cd(-h, -h);
        default: return cd(h, -h);
        }
    }
    const double th = 2.0 * M_PI * (double)(int64_t)kk / (double)(int64_t)m;
    return cd(std::cos(th), std::sin(th));
}

// C64::sqrt2_pow (scalar.rs:133-140).
cd sqrt2_pow(int64_t k) {
    int64_t half = k >= 0 ? k / 2 : -((-k + 1) / 2);
    double v = std::ldexp(1.0, (int)half);
    if (((k % 2) + 2) % 2 == 1) v *= 1.41421356237309504880;  // SQRT_2
    return cd(v, 0.0);
}

// The reference's complex product, (a.re b.re - a.im b.im, a.re b.im + a.im b.re).
static inline cd rmul(cd a, cd b) {
    return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(NCSV_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

static const char *NAMES[NCSV_GATE_COUNT] = {
    "I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG", "T", "T_DAG", "CX",
    "CY", "CZ", "SWAP", "CS", "CS_DAG", "CCZ", "CCX", "PHASE", "CPHASE", "RX", "RY", "RZ", "UNITARY", "PAULI"};

bool gate_lookup(const std::string &u, ncsv_gate *id) {
    if (u == "CNOT") {
        *id = NCSV_GATE_CX;
        return true;
    }
    if (u == "TOFFOLI") {
        *id = NCSV_GATE_CCX;
        return true;
    }
    for (int i = 0; i < NCSV_GATE_UNITARY; i++)
        if (u == NAMES[i]) {
            *id = (ncsv_gate)i;
            return true;
        }
    return false;
}

GateDesc make_gate(ncsv_gate id, const int64_t *ip, const double *fp, uint32_t nq) {
    GateDesc g;
    if ((int)id < 0 || id >= NCSV_GATE_COUNT) fail(NCSV_INVALID, "unknown gate id");
    g.id = id;
    switch (id) {
    case NCSV_GATE_PHASE:
    case NCSV_GATE_CPHASE:
        if (!ip) fail(NCSV_INVALID, "PHASE/CPHASE need (k, bits)");
        g.k = ip[0];
        if (ip[1] < 1 || ip[1] > 62) fail(NCSV_INVALID, "phase bits must be in 1..=62");
        g.bits = (uint32_t)ip[1];
        break;
    case NCSV_GATE_RX:
    case NCSV_GATE_RY:
    case NCSV_GATE_RZ:
        if (!fp) fail(NCSV_INVALID, "rotations need theta");
        g.theta = fp[0];
        break;
    case NCSV_GATE_UNITARY: {
        if (!fp) fail(NCSV_INVALID, "UNITARY needs a matrix");
        if (nq > 12) fail(NCSV_UNSUPPORTED, "UNITARY on more than 12 qubits");
        g.mk = nq;
        const size_t d = size_t(1) << nq;
        g.mat.resize(d * d);
        for (size_t i = 0; i < d * d; i++) g.mat[i] = cd(fp[2 * i], fp[2 * i + 1]);
        break;
    }
    case NCSV_GATE_PAULI:
        if (!ip) fail(NCSV_INVALID, "PAULI needs (x, z, phase)");
        g.pn = nq;
        g.px = (uint64_t)ip[0] & (nq >= 64 ? ~0ull : ((1ull << nq) - 1));
        g.pz = (uint64_t)ip[1] & (nq >= 64 ? ~0ull : ((1ull << nq) - 1));
        g.pph = (uint32_t)(ip[2] & 3);
        break;
    default: break;
    }
    return g;
}

// Gate::arity (gate.rs:447-476).
uint32_t gate_arity(const GateDesc &g) {
    switch (g.id) {
    case NCSV_GATE_CX: case NCSV_GATE_CY: case NCSV_GATE_CZ: case NCSV_GATE_SWAP: case NCSV_GATE_CS:
    case NCSV_GATE_CS_DAG: case NCSV_GATE_CPHASE:
        return 2;
    case NCSV_GATE_CCZ: case NCSV_GATE_CCX:
        return 3;
    case NCSV_GATE_UNITARY: return g.mk;
    case NCSV_GATE_PAULI: return g.pn;
    default: return 1;
    }
}

static std::string fmt_double(double v) {
    std::ostringstream o;
    o.precision(17);
    o << v;
    return o.str();
}

// Gate::name (gate.rs:479-515).
std::string gate_name(const GateDesc &g) {
    switch (g.id) {
    case NCSV_GATE_PHASE: return "PHASE(" + std::to_string(g.k) + "," + std::to_string(g.bits) + ")";
    case NCSV_GATE_CPHASE: return "CPHASE(" + std::to_string(g.k) + "," + std::to_string(g.bits) + ")";
    case NCSV_GATE_RX: return "RX(" + fmt_double(g.theta) + ")";
    case NCSV_GATE_RY: return "RY(" + fmt_double(g.theta) + ")";
    case NCSV_GATE_RZ: return "RZ(" + fmt_double(g.theta) + ")";
    case NCSV_GATE_UNITARY: return "UNITARY[" + std::to_string(g.mk) + " qubits]";
    case NCSV_GATE_PAULI: return "PAULI[" + std::to_string(g.pn) + " qubits]";
    default: return NAMES[g.id];
    }
}

// check_qubits (error.rs:69-82).
void check_qubits(const uint32_t *qubits, uint32_t nq, uint32_t n) {
    for (uint32_t i = 0; i < nq; i++) {
        if (qubits[i] >= n)
            fail(NCSV_QUBIT_OUT_OF_RANGE,
                 "qubit " + std::to_string(qubits[i]) + " out of range for " + std::to_string(n) + " qubits");
        for (uint32_t j = 0; j < i; j++)
            if (qubits[j] == qubits[i]) fail(NCSV_DUPLICATE_QUBIT, "qubit " + std::to_string(qubits[i]) + " listed twice");
    }
}

// A diagonal gate with phase `c` on the all-ones local state (Z, S, T, CZ,
// CS, CCZ, PHASE, CPHASE: gate.rs:525-535 via mono_phase 437-443).
static LOp diag_on_ones(const uint32_t *q, uint32_t k, cd c) {
    LOp op;
    op.kind = OpKind::Diag;
    op.qubits.assign(q, q + k);
    op.m.assign(size_t(1) << k, cd(1.0, 0.0));
    // from_phasor: root_of_unity * sqrt2_pow(0) (scalar.rs:39-52)
    op.m.back() = rmul(c, cd(1.0, 0.0));
    return op;
}

static LOp perm_op(const uint32_t *q, uint32_t k, std::vector<uint32_t> img, std::vector<cd> ph) {
    LOp op;
    op.kind = OpKind::Perm;
    op.qubits.assign(q, q + k);
    op.img = std::move(img);
    op.m = std::move(ph);
    return op;
}

static LOp dense_op(const uint32_t *q, uint32_t k, std::vector<cd> m) {
    LOp op;
    op.kind = OpKind::Dense;
    op.qubits.assign(q, q + k);
    op.m = std::move(m);
    return op;
}

// apply_gate's routing (vector.rs:364-395): Pauli -> monomial view
// (gate.rs:519-564) -> matrix view (gate.rs:670-747).  CCX's reference matrix
// is a permutation matrix applied with exact-zero skips, so it is lowered to
// the same permutation (bit-exact moves).
LOp lower_gate(const GateDesc &g, const uint32_t *q, uint32_t nq, uint32_t n) {
    check_qubits(q, nq, n);
    const uint32_t ar = gate_arity(g);
    if (nq != ar)
        fail(NCSV_ARITY, "gate " + gate_name(g) + " acts on " + std::to_string(ar) + " qubits, " +
                             std::to_string(nq) + " given");
    const cd one(1.0, 0.0), I(0.0, 1.0);
    switch (g.id) {
    case NCSV_GATE_PAULI: {
        LOp op;
        op.kind = OpKind::Pauli;
        op.qubits.assign(q, q + nq);
        for (uint32_t j = 0; j < nq; j++) {  // Pauli::embed (pauli/mod.rs:283-292)
            op.x |= ((g.px >> j) & 1ull) << q[j];
            op.z |= ((g.pz >> j) & 1ull) << q[j];
        }
        op.ph = g.pph;
        return op;
    }
    case NCSV_GATE_I: {
        LOp op;
        op.kind = OpKind::Nop;
        op.qubits.assign(q, q + nq);
        return op;
    }
    case NCSV_GATE_X: return perm_op(q, 1, {1, 0}, {one, one});
    case NCSV_GATE_Y:
        // MonomialGate::pauli(Y): scalar i, phase (-1)^x, shift 1 (gate.rs:99-108)
        return perm_op(q, 1, {1, 0}, {rmul(cd(0.0, 1.0), one), rmul(cd(0.0, -1.0), one)});
    case NCSV_GATE_Z: return diag_on_ones(q, 1, root_of_unity(1, 1));
    case NCSV_GATE_S: return diag_on_ones(q, 1, root_of_unity(1, 2));
    case NCSV_GATE_S_DAG: return diag_on_ones(q, 1, root_of_unity(3, 2));
    case NCSV_GATE_T: return diag_on_ones(q, 1, root_of_unity(1, 3));
    case NCSV_GATE_T_DAG: return diag_on_ones(q, 1, root_of_unity(7, 3));
    case NCSV_GATE_PHASE: return diag_on_ones(q, 1, root_of_unity(g.k, g.bits));
    case NCSV_GATE_CPHASE: return diag_on_ones(q, 2, root_of_unity(g.k, g.bits));
    case NCSV_GATE_CZ: return diag_on_ones(q, 2, root_of_unity(1, 1));
    case NCSV_GATE_CS: return diag_on_ones(q, 2, root_of_unity(1, 2));
    case NCSV_GATE_CS_DAG: return diag_on_ones(q, 2, root_of_unity(3, 2));
    case NCSV_GATE_CCZ: return diag_on_ones(q, 3, root_of_unity(1, 1));
    case NCSV_GATE_CX: return perm_op(q, 2, {0, 3, 2, 1}, {one, one, one, one});
    case NCSV_GATE_CY:
        // f = x0 + 2 x0 x1 mod 4 (gate.rs:543-555): x=1 -> i, x=3 -> -i
        return perm_op(q, 2, {0, 3, 2, 1}, {one, rmul(root_of_unity(1, 2), one), one, rmul(root_of_unity(3, 2), one)});
    case NCSV_GATE_SWAP: return perm_op(q, 2, {0, 2, 1, 3}, {one, one, one, one});
    case NCSV_GATE_CCX: {
        std::vector<uint32_t> img(8);
        for (uint32_t x = 0; x < 8; x++) img[x] = (x & 3) == 3 ? x ^ 4 : x;  // gate.rs:697-705
        return perm_op(q, 3, img, std::vector<cd>(8, one));
    }
    case NCSV_GATE_H: {
        const cd r = sqrt2_pow(-1);
        return dense_op(q, 1, {r, r, r, -r});
    }
    case NCSV_GATE_SQRT_X:
    case NCSV_GATE_SQRT_X_DAG:
    case NCSV_GATE_SQRT_Y:
    case NCSV_GATE_SQRT_Y_DAG: {
        const cd half(0.5, 0.0);
        const cd p = rmul(one + I, half), mneg = rmul(one - I, half);
        if (g.id == NCSV_GATE_SQRT_X) return dense_op(q, 1, {p, mneg, mneg, p});
        if (g.id == NCSV_GATE_SQRT_X_DAG) return dense_op(q, 1, {mneg, p, p, mneg});
        if (g.id == NCSV_GATE_SQRT_Y) return dense_op(q, 1, {p, -p, p, p});
        return dense_op(q, 1, {mneg, mneg, -mneg, mneg});
    }
    case NCSV_GATE_RX:
    case NCSV_GATE_RY:
    case NCSV_GATE_RZ: {
        const double t = g.theta, c = std::cos(t / 2.0), s = std::sin(t / 2.0);
        if (g.id == NCSV_GATE_RX) return dense_op(q, 1, {cd(c, 0), cd(0, -s), cd(0, -s), cd(c, 0)});
        if (g.id == NCSV_GATE_RY) return dense_op(q, 1, {cd(c, 0), cd(-s, 0), cd(s, 0), cd(c, 0)});
        // Rz is applied through the matrix view in the reference (a dense
        // diagonal); as a diagonal op the arithmetic per entry is the same
        // single complex product.
        LOp op;
        op.kind = OpKind::Diag;
        op.qubits.assign(q, q + 1);
        op.m = {cd(std::cos(-t / 2.0), std::sin(-t / 2.0)), cd(std::cos(t / 2.0), std::sin(t / 2.0))};
        return op;
    }
    case NCSV_GATE_UNITARY: return dense_op(q, nq, g.mat);
    default: fail(NCSV_NOT_REPRESENTABLE, "gate " + gate_name(g) + " in the scalar type");
    }
}

// ------------------------------------------------------------------ Paulis

bool pauli_is_hermitian(uint64_t x, uint64_t z, uint32_t ph) {
    return (ph & 1) == (uint32_t)(__builtin_popcountll(x & z) & 1);  // ops.rs:22-27
}

static void set_kind(PauliM &p, uint32_t q, char kind) {
    // Pauli::set_kind (pauli/mod.rs:192-205): keeps the letter phase
    const bool wasy = ((p.x >> q) & 1) && ((p.z >> q) & 1);
    if (wasy) p.ph = (p.ph + 3) & 3;
    const uint64_t b = 1ull << q;
    p.x &= ~b;
    p.z &= ~b;
    if (kind == 'X' || kind == 'Y') p.x |= b;
    if (kind == 'Z' || kind == 'Y') p.z |= b;
    if (kind == 'Y') p.ph = (p.ph + 1) & 3;
}

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

PauliM pauli_parse(const std::string &s, int n) {
    auto err = [&](const std::string &m) { fail(NCSV_PARSE, "Pauli '" + s + "': " + m); };
    std::string rest = trim(s);
    uint32_t phase = 0;
    const char *pres[5] = {"-i", "+i", "-", "+", "i"};
    const uint32_t phs[5] = {3, 1, 2, 0, 1};
    for (int i = 0; i < 5; i++) {
        const std::string pre = pres[i];
        if (rest.compare(0, pre.size(), pre) == 0) {
            std::string r = trim(rest.substr(pre.size()));
            if (!r.empty() && r[0] == '*') r = trim(r.substr(1));
            phase = phs[i];
            rest = r;
            break;
        }
    }
    bool dense = !rest.empty();
    for (char c : rest)
        if (!strchr("IXYZ_", c)) dense = false;
    PauliM p;
    if (dense) {
        if (rest.size() > 64) err("more than 64 qubits");
        if (n >= 0 && (size_t)n != rest.size())
            err(std::to_string(rest.size()) + " letters for " + std::to_string(n) + " qubits");
        p.n = (uint32_t)rest.size();
        for (uint32_t q = 0; q < p.n; q++) set_kind(p, q, rest[q]);
        p.ph = (p.ph + phase) & 3;
        return p;
    }
    std::vector<std::pair<uint32_t, char>> sites;
    std::string tok;
    auto flush = [&]() {
        if (tok.empty()) return;
        if (tok != "1") {
            const char k = tok[0];
            if (!strchr("IXYZ_", k)) err("token '" + tok + "'");
            const std::string num = tok.substr(1);
            if (num.empty() || num.find_first_not_of("0123456789") != std::string::npos) err("token '" + tok + "'");
            const unsigned long idx = std::stoul(num);
            for (auto &st : sites)
                if (st.first == idx) err("qubit " + std::to_string(idx) + " repeated");
            sites.push_back({(uint32_t)idx, k});
        }
        tok.clear();
    };
    for (char c : rest) {
        if (c == '*' || isspace((unsigned char)c))
            flush();
        else
            tok.push_back(c);
    }
    flush();
    uint32_t nn = 0;
    if (n >= 0)
        nn = (uint32_t)n;
    else
        for (auto &st : sites) nn = std::max(nn, st.first + 1);
    if (nn > 64) err("more than 64 qubits");
    p.n = nn;
    for (auto &st : sites) {
        if (st.first >= nn) err("qubit " + std::to_string(st.first) + " out of range for " + std::to_string(nn) + " qubits");
        set_kind(p, st.first, st.second);
    }
    p.ph = (p.ph + phase) & 3;
    return p;
}

std::string pauli_sparse_string(const PauliM &p) {
    const uint32_t ys = (uint32_t)(__builtin_popcountll(p.x & p.z) % 4);
    const uint32_t disp = (p.ph + 4 - ys) & 3;
    static const char *pre[4] = {"", "i*", "-", "-i*"};
    std::string body;
    for (uint32_t q = 0; q < p.n; q++) {
        const bool xb = (p.x >> q) & 1, zb = (p.z >> q) & 1;
        if (!xb && !zb) continue;
        if (!body.empty()) body += "*";
        body += (xb && zb) ? 'Y' : (xb ? 'X' : 'Z');
        body += std::to_string(q);
    }
    return std::string(pre[disp]) + (body.empty() ? "1" : body);
}

// ------------------------------------------------------------------- noise

std::vector<NoiseErr> noise_errors(ncsv_noise ch, const double *p) {
    switch (ch) {
    case NCSV_NOISE_X_ERROR: return {{p[0], 1, 0}};
    case NCSV_NOISE_Y_ERROR: return {{p[0], 1, 1}};
    case NCSV_NOISE_Z_ERROR: return {{p[0], 0, 1}};
    case NCSV_NOISE_PAULI_CHANNEL_1: return {{p[0], 1, 0}, {p[1], 1, 1}, {p[2], 0, 1}};
    case NCSV_NOISE_DEPOLARIZE1: return {{p[0] / 3.0, 1, 0}, {p[0] / 3.0, 1, 1}, {p[0] / 3.0, 0, 1}};
    case NCSV_NOISE_DEPOLARIZE2: {
        std::vector<NoiseErr> v;
        for (uint32_t k = 1; k < 16; k++) v.push_back({p[0] / 15.0, k & 3, k >> 2});
        return v;
    }
    default: fail(NCSV_INVALID, "unknown noise channel");
    }
}

uint32_t noise_arity(ncsv_noise ch) { return ch == NCSV_NOISE_DEPOLARIZE2 ? 2 : 1; }

std::string noise_name(ncsv_noise ch, const double *p) {
    switch (ch) {
    case NCSV_NOISE_X_ERROR: return "X_ERROR(" + fmt_double(p[0]) + ")";
    case NCSV_NOISE_Y_ERROR: return "Y_ERROR(" + fmt_double(p[0]) + ")";
    case NCSV_NOISE_Z_ERROR: return "Z_ERROR(" + fmt_double(p[0]) + ")";
    case NCSV_NOISE_PAULI_CHANNEL_1:
        return "PAULI_CHANNEL_1(" + fmt_double(p[0]) + "," + fmt_double(p[1]) + "," + fmt_double(p[2]) + ")";
    case NCSV_NOISE_DEPOLARIZE1: return "DEPOLARIZE1(" + fmt_double(p[0]) + ")";
    default: return "DEPOLARIZE2(" + fmt_double(p[0]) + ")";
    }
}

}  // namespace ncsv
