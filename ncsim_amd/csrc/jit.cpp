// Sweep kernels specialized per plan: the stage/op list of one tile sweep
// written out as straight-line HIP source and compiled at run time with
// hiprtc for the device's architecture.  Register bits, masks, tile and
// outer qubit positions and the LDS layouts of every stage become
// constants; the gate coefficients stay data (read from a device array with
// constant offsets).  X-type moves without conditions (X, CX-free swaps,
// SWAP) cost nothing: the generator renames registers.  The compiled
// kernels are cached by source text for the life of the process.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <future>
#include <list>
#include <map>
#include <mutex>
#include <sstream>

#include "jit.hpp"

namespace ncsv {
namespace jit {

namespace {

struct Gen {
    std::ostringstream o;
    std::vector<double> *coef;
    std::vector<int> name;  // logical register r -> variable index
    uint32_t LT, R, L, NR;
    std::vector<uint32_t> tq;
    std::vector<uint32_t> oq;

    std::string var(uint32_t r) const { return "a" + std::to_string(name[r]); }

    bool literal = false;  // coefficients as constants in the code instead of loads from c[]
    bool dbl = true;
    double scale = 1.0;  // common factor taken out of the real-form 2x2s, applied at the store
    bool opaque = false;  // exchange addresses recomputed where used (kernels that loop over tiles)
    int S = 4;            // LDS swizzle width (16-byte slots: 4; 8-byte: 5)
    int mfma_blocks = 0;  // FP32 MFMA dense blocks emitted
    const char *bar = "__syncthreads()";  // the exchanges' workgroup barrier
    int half = 0;         // exchanges through half a tile of LDS, in two passes (2, 3: test the
                          // wavefront-turn and intermediate-layout forms wherever they apply)

    uint32_t push(double v) {
        coef->push_back(v);
        return (uint32_t)coef->size() - 1;
    }
    std::string C(uint32_t i) const {
        if (!literal) return "c[" + std::to_string(i) + "]";
        // exact hexadecimal literal of the value the kernel would have loaded
        std::ostringstream e;
        e << std::hexfloat;
        if (dbl)
            e << "(" << (*coef)[i] << ")";
        else
            e << "(" << (double)(float)(*coef)[i] << "f)";
        return e.str();
    }
};

// Coefficients compiled in as constants (default; NCSV_JIT_LITERALS=0 loads
// them from the coefficient buffer so that one kernel serves every angle).
bool use_literals() {
    const char *e = getenv("NCSV_JIT_LITERALS");
    return !(e && atoi(e) == 0);
}

// Sweep kernel shape knobs: exchanges go through half a tile of LDS in two
// passes by default (NCSV_JIT_HALF=0: the whole tile at once), so three or
// four workgroups share a CU instead of two (30-qubit rotation brickwork:
// complex64 sweeps 4.41 -> 3.89 ms, complex128 7.28 -> 7.15 ms);
// NCSV_JIT_WAVES=w asks the compiler for w wavefronts per SIMD (registers
// capped at 512 / w).
int use_half() {
    const char *e = getenv("NCSV_JIT_HALF");
    return e ? atoi(e) : 1;
}
int jit_waves() {
    const char *e = getenv("NCSV_JIT_WAVES");
    return e ? std::max(0, std::min(8, atoi(e))) : 0;
}
// Amplitudes are read and written once per sweep, so the sweeps load and
// store them with nontemporal hints (NCSV_JIT_NT: bit 0 stores, bit 1 loads;
// default both).  On MI355X this took the 30-qubit complex128 brickwork's
// sweeps from 7.10 to 6.91 ms and its expectation passes from 3.31 to
// 3.16 ms, and the 33-qubit complex64 QFT's sweeps from 26.3 to 24.5 ms.
int jit_nt() {
    const char *e = getenv("NCSV_JIT_NT");
    return e ? atoi(e) & 3 : 3;
}
const char *NT_HELPERS =
    "typedef T NV __attribute__((ext_vector_type(2)));\n"
    "__device__ __forceinline__ V ld_nt(const V *p) { NV t = __builtin_nontemporal_load((const NV *)p); V v; "
    "v.x = t.x; v.y = t.y; return v; }\n"
    "__device__ __forceinline__ void st_nt(V *p, V v) { NV t; t.x = v.x; t.y = v.y; "
    "__builtin_nontemporal_store(t, (NV *)p); }\n";

// Expression depositing bits 0..k-1 of `src` on the bit positions `dst`
// (ascending source bits), grouping runs of consecutive destinations.
std::string deposit(const std::string &src, const std::vector<uint32_t> &dst, bool wide) {
    std::ostringstream e;
    bool first = true;
    size_t i = 0;
    while (i < dst.size()) {
        size_t j = i + 1;
        while (j < dst.size() && dst[j] == dst[j - 1] + 1) j++;
        const uint32_t len = (uint32_t)(j - i);
        const uint64_t mask = (len >= 64) ? ~0ull : ((1ull << len) - 1);
        if (!first) e << " | ";
        first = false;
        e << "(((" << src << " >> " << i << ") & " << mask << (wide ? "ull" : "u") << ") << " << dst[i] << ")";
        i = j;
    }
    if (first) e << (wide ? "0ull" : "0u");
    return e.str();
}

std::vector<uint32_t> positions(uint32_t mask, uint32_t L, bool set) {
    std::vector<uint32_t> v;
    for (uint32_t p = 0; p < L; p++)
        if ((((mask >> p) & 1) != 0) == set) v.push_back(p);
    return v;
}

// LDS slot of tile index (thread part | off) under the XOR swizzle of the
// kernels, swz(t) = t ^ (((t >> S) ^ (t >> 2S)) & (2^S - 1)).  swz is linear
// over XOR and the two parts have disjoint bits, so the slot is
// swz(thread part) ^ swz(off): with `sw` = swz(thread part) in a register
// (bits only on `tbits` and the low S bits), the constant's other bits can
// be added (an immediate offset of the LDS instruction) and only the rest is
// XORed.
uint32_t swz_host(uint32_t t, int S) { return t ^ (((t >> S) ^ (t >> (2 * S))) & ((1u << S) - 1)); }
std::string lidx(int S, const char *sw, uint32_t tbits, uint32_t off) {
    const uint32_t c = swz_host(off, S);
    const uint32_t live = tbits | ((1u << S) - 1);
    const uint32_t lo = c & live, hi = c & ~live;
    std::ostringstream e;
    if (lo)
        e << "(" << sw << " ^ " << lo << "u)";
    else
        e << sw;
    if (hi) e << " + " << hi << "u";
    return e.str();
}
uint32_t bits_of(const std::vector<uint32_t> &pos) {
    uint32_t m = 0;
    for (uint32_t q : pos) m |= 1u << q;
    return m;
}

// i^k * (variable) as an expression pair
void ipow_assign(std::ostringstream &o, const std::string &dst, const std::string &src, int k) {
    switch (k & 3) {
    case 0: o << dst << " = " << src << ";"; break;
    case 1: o << dst << ".x = -" << src << ".y; " << dst << ".y = " << src << ".x;"; break;
    case 2: o << dst << ".x = -" << src << ".x; " << dst << ".y = -" << src << ".y;"; break;
    default: o << dst << ".x = " << src << ".y; " << dst << ".y = -" << src << ".x;"; break;
    }
}

// Condition on the global index bits outside the registers, in the
// current layout: thread bits come from tid, outer bits from the base.
std::string cond_expr(const Gen &g, uint64_t gmask, uint32_t cur) {
    if (!gmask) return "";
    const std::vector<uint32_t> tp = positions(cur, g.L, false);
    uint64_t tmask = 0, bmask = 0;
    for (uint64_t a = gmask; a; a &= a - 1) {
        const uint32_t q = (uint32_t)__builtin_ctzll(a);
        int pos = -1;
        for (uint32_t p = 0; p < g.L; p++)
            if (g.tq[p] == q) pos = (int)p;
        if (pos < 0) {
            bmask |= 1ull << q;
        } else {
            for (size_t i = 0; i < tp.size(); i++)
                if (tp[i] == (uint32_t)pos) tmask |= 1ull << i;
        }
    }
    std::ostringstream e;
    e << "(";
    bool first = true;
    if (tmask) {
        e << "((tid & " << tmask << "u) == " << tmask << "u)";
        first = false;
    }
    if (bmask) {
        if (!first) e << " && ";
        e << "((base & " << bmask << "ull) == " << bmask << "ull)";
    }
    e << ")";
    return e.str();
}

// The exchange between register layouts `from` and `to` through half a tile
// of LDS, for layouts sharing a register position p: the amplitudes whose
// tile position p holds h go in pass h, each thread writing the half of its
// registers with that bit and reading the half of the new layout with it.
// Each pass reads into the variables its own writes just freed (the
// generator renames registers), so no value is overwritten before it is
// stored and no extra registers are needed.  Layouts with no common
// register position go through an intermediate layout sharing one with
// each.  LDS use halves at the price of two more barriers per exchange.
// A tile position that is the same wavefront-index bit of the thread
// index in both layouts: whole wavefronts write and then read in the pass
// of their bit, so none overwrites a register it has not stored.  -1 if none.
int common_wave_bit(const Gen &g, uint32_t from, uint32_t to) {
    const std::vector<uint32_t> ta = positions(from, g.L, false), tb = positions(to, g.L, false);
    for (int i = (int)g.LT - 1; i >= 6; i--)
        if (ta[i] == tb[i]) return i;
    return -1;
}

void emit_exchange_wave_halves(Gen &g, uint32_t from, uint32_t to, int i) {
    auto &o = g.o;
    const std::vector<uint32_t> ta = positions(from, g.L, false), tb = positions(to, g.L, false);
    const std::vector<uint32_t> ra = positions(from, g.L, true), rb = positions(to, g.L, true);
    const uint32_t p = ta[i];
    auto squeeze = [&](uint32_t q) { return q < p ? q : q - 1; };
    auto tpart = [&](const std::vector<uint32_t> &tpos) {
        std::vector<uint32_t> d;
        for (size_t k = 0; k < tpos.size(); k++)
            if ((int)k != i) d.push_back(squeeze(tpos[k]));
        // the thread index without bit i
        return deposit("((tid & " + std::to_string((1u << i) - 1) + "u) | ((tid >> " + std::to_string(i + 1) +
                           "u) << " + std::to_string(i) + "u))",
                       d, false);
    };
    auto roffs = [&](const std::vector<uint32_t> &rpos, uint32_t r) {
        uint32_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if ((r >> j) & 1) off |= 1u << squeeze(rpos[j]);
        return off;
    };
    o << "  { // exchange in halves on tile position " << p << " (wavefronts take turns)\n";
    o << "    unsigned ta = " << tpart(ta) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(ta));\n";
    o << "    unsigned tb = " << tpart(tb) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(tb));\n";
    o << "    const unsigned sa = swz(ta), sb = swz(tb);\n";
    std::vector<uint32_t> sqa, sqb;
    for (size_t k = 0; k < ta.size(); k++)
        if ((int)k != i) {
            sqa.push_back(squeeze(ta[k]));
            sqb.push_back(squeeze(tb[k]));
        }
    const uint32_t ma = bits_of(sqa), mb = bits_of(sqb);
    for (uint32_t h = 0; h < 2; h++) {
        o << "    if (((tid >> " << i << ") & 1u) == " << h << "u) {\n";
        for (uint32_t r = 0; r < g.NR; r++)
            o << "      lds[" << lidx(g.S, "sa", ma, roffs(ra, r)) << "] = " << g.var(r) << ";\n";
        o << "    }\n    " << g.bar << ";\n";
        o << "    if (((tid >> " << i << ") & 1u) == " << h << "u) {\n";
        for (uint32_t r = 0; r < g.NR; r++)
            o << "      " << g.var(r) << " = lds[" << lidx(g.S, "sb", mb, roffs(rb, r)) << "];\n";
        o << "    }\n    " << g.bar << ";\n";
    }
    o << "  }\n";
}

void emit_exchange_half(Gen &g, uint32_t from, uint32_t to) {
    if (!(from & to) || (g.half >= 2 && from != to)) {
        const int i = g.half == 3 ? -1 : common_wave_bit(g, from, to);
        if (i >= 0) {
            emit_exchange_wave_halves(g, from, to, i);
            return;
        }
        if (!(from & to)) {
            const uint32_t hi = 31u - (uint32_t)__builtin_clz(from);
            const uint32_t mid = (from & ~(1u << hi)) | (1u << (uint32_t)__builtin_ctz(to));
            const int keep = g.half;
            g.half = 1;
            emit_exchange_half(g, from, mid);
            emit_exchange_half(g, mid, to);
            g.half = keep;
            return;
        }
    }
    auto &o = g.o;
    const std::vector<uint32_t> ta = positions(from, g.L, false), tb = positions(to, g.L, false);
    const std::vector<uint32_t> ra = positions(from, g.L, true), rb = positions(to, g.L, true);
    const uint32_t p = (uint32_t)__builtin_ctz(from & to);
    uint32_t ja = 0, jb = 0;
    for (uint32_t j = 0; j < g.R; j++) {
        if (ra[j] == p) ja = j;
        if (rb[j] == p) jb = j;
    }
    auto squeeze = [&](uint32_t q) { return q < p ? q : q - 1; };
    auto tpart = [&](const std::vector<uint32_t> &tpos) {
        std::vector<uint32_t> d;
        for (uint32_t q : tpos) d.push_back(squeeze(q));
        return deposit("tid", d, false);
    };
    auto roffs = [&](const std::vector<uint32_t> &rpos, uint32_t jp, uint32_t r) {
        uint32_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if (j != jp && ((r >> j) & 1)) off |= 1u << squeeze(rpos[j]);
        return off;
    };
    std::vector<int> renamed(g.NR, -1);
    o << "  { // exchange in halves on tile position " << p << "\n";
    o << "    unsigned ta = " << tpart(ta) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(ta));\n";
    o << "    unsigned tb = " << tpart(tb) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(tb));\n";
    o << "    const unsigned sa = swz(ta), sb = swz(tb);\n";
    uint32_t ma = 0, mb = 0;
    for (uint32_t q : ta) ma |= 1u << squeeze(q);
    for (uint32_t q : tb) mb |= 1u << squeeze(q);
    for (uint32_t h = 0; h < 2; h++) {
        std::vector<int> freed;
        for (uint32_t r = 0; r < g.NR; r++) {
            if (((r >> ja) & 1u) != h) continue;
            o << "    lds[" << lidx(g.S, "sa", ma, roffs(ra, ja, r)) << "] = " << g.var(r) << ";\n";
            freed.push_back(g.name[r]);
        }
        o << "    " << g.bar << ";\n";
        size_t k = 0;
        for (uint32_t r = 0; r < g.NR; r++) {
            if (((r >> jb) & 1u) != h) continue;
            renamed[r] = freed[k++];
            o << "    a" << renamed[r] << " = lds[" << lidx(g.S, "sb", mb, roffs(rb, jb, r)) << "];\n";
        }
        o << "    " << g.bar << ";\n";
    }
    o << "  }\n";
    g.name = renamed;
}

void emit_exchange(Gen &g, uint32_t from, uint32_t to) {
    if (g.half) {
        emit_exchange_half(g, from, to);
        return;
    }
    auto &o = g.o;
    const std::vector<uint32_t> ta = positions(from, g.L, false), tb = positions(to, g.L, false);
    const std::vector<uint32_t> ra = positions(from, g.L, true), rb = positions(to, g.L, true);
    o << "  {\n    unsigned ta = " << deposit("tid", ta, false) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(ta));\n";
    o << "    const unsigned sa = swz(ta);\n";
    for (uint32_t r = 0; r < g.NR; r++) {
        uint32_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if ((r >> j) & 1) off |= 1u << ra[j];
        o << "    lds[" << lidx(g.S, "sa", bits_of(ta), off) << "] = " << g.var(r) << ";\n";
    }
    o << "    " << g.bar << ";\n    unsigned tb = " << deposit("tid", tb, false) << ";\n";
    if (g.opaque) o << "    asm volatile(\"\" : \"+v\"(tb));\n";
    o << "    const unsigned sb = swz(tb);\n";
    for (uint32_t r = 0; r < g.NR; r++) {
        uint32_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if ((r >> j) & 1) off |= 1u << rb[j];
        o << "    " << g.var(r) << " = lds[" << lidx(g.S, "sb", bits_of(tb), off) << "];\n";
    }
    o << "    " << g.bar << ";\n  }\n";
}

bool is_exact(cd c, double re, double im) { return c.real() == re && c.imag() == im; }

// A real 2x2 r as f * q with as many entries of q equal to +-1 as one common
// factor allows (two for rotations and reflections, four for H), |f| largest
// among the choices that reach that count, so |q| <= 1 for those.  The
// factor is multiplied into the sweep's scale and applied once at the store;
// a +-1 entry costs an add instead of a multiply-add.
struct Factored {
    double f = 1.0;
    double q[4];
};
Factored factor_2x2(const double r[4]) {
    auto units = [](const double *q) {
        int u = 0;
        for (int j = 0; j < 4; j++) u += (q[j] == 1.0 || q[j] == -1.0);
        return u;
    };
    Factored best;
    for (int j = 0; j < 4; j++) best.q[j] = r[j];
    int bu = units(best.q);
    for (int i = 0; i < 4; i++) {
        if (r[i] == 0.0) continue;
        Factored c;
        c.f = r[i];
        for (int j = 0; j < 4; j++) c.q[j] = r[j] / c.f;
        const int u = units(c.q);
        if (u > bu || (u == bu && u > 0 && std::fabs(c.f) > std::fabs(best.f))) {
            best = c;
            bu = u;
        }
    }
    return best;
}

// q * v as a term of a sum: "v", "-v" or "C*v"
std::string term(Gen &g, double q, const std::string &v, bool first) {
    if (q == 1.0) return (first ? "" : " + ") + v;
    if (q == -1.0) return (first ? "-" : " - ") + v;
    return (first ? "" : " + ") + g.C(g.push(q)) + "*" + v;
}

void emit_op(Gen &g, const TileOpH &op, uint32_t cur) {
    auto &o = g.o;
    const uint32_t NR = g.NR;
    switch (op.kind) {
    case TK_U1R:
    case TK_U1X: {
        // real form: U1R [[r0, r1], [r2, r3]]; U1X [[r0, i r1], [i r2, r3]]
        const uint32_t J = op.j0;
        const bool xk = op.kind == TK_U1X;
        const double r[4] = {op.m[0].real(), xk ? op.m[1].imag() : op.m[1].real(), xk ? op.m[2].imag() : op.m[2].real(),
                             op.m[3].real()};
        const Factored fq = factor_2x2(r);
        g.scale *= fq.f;
        o << "  {\n";
        for (uint32_t i = 0; i < NR; i++) {
            if (i & (1u << J)) continue;
            const std::string a = g.var(i), b = g.var(i | (1u << J));
            o << "    { const V x = " << a << ", y = " << b << ";\n";
            if (!xk) {
                o << "      " << a << ".x = " << term(g, fq.q[0], "x.x", true) << term(g, fq.q[1], "y.x", false) << "; ";
                o << a << ".y = " << term(g, fq.q[0], "x.y", true) << term(g, fq.q[1], "y.y", false) << ";\n";
                o << "      " << b << ".x = " << term(g, fq.q[2], "x.x", true) << term(g, fq.q[3], "y.x", false) << "; ";
                o << b << ".y = " << term(g, fq.q[2], "x.y", true) << term(g, fq.q[3], "y.y", false) << "; }\n";
            } else {
                // a = q0 x + i q1 y, b = i q2 x + q3 y
                o << "      " << a << ".x = " << term(g, fq.q[0], "x.x", true) << term(g, -fq.q[1], "y.y", false) << "; ";
                o << a << ".y = " << term(g, fq.q[0], "x.y", true) << term(g, fq.q[1], "y.x", false) << ";\n";
                o << "      " << b << ".x = " << term(g, fq.q[3], "y.x", true) << term(g, -fq.q[2], "x.y", false) << "; ";
                o << b << ".y = " << term(g, fq.q[3], "y.y", true) << term(g, fq.q[2], "x.x", false) << "; }\n";
            }
        }
        o << "  }\n";
        break;
    }
    case TK_U1: {
        const uint32_t J = op.j0;
        // m = [a0 a1 | b0 b1 | c0 c1 | d0 d1] at k0..k0+7
        const uint32_t k0 = g.push(op.m[0].real());
        for (int t = 0; t < 4; t++) {
            if (t) g.push(op.m[t].real());
            g.push(op.m[t].imag());
        }
        auto c = [&](int t) { return g.C(k0 + t); };
        o << "  {\n";
        for (uint32_t i = 0; i < NR; i++) {
            if (i & (1u << J)) continue;
            const std::string a = g.var(i), b = g.var(i | (1u << J));
            o << "    { const V x = " << a << ", y = " << b << ";\n";
            o << "      " << a << ".x = " << c(0) << "*x.x - " << c(1) << "*x.y + " << c(2) << "*y.x - " << c(3)
              << "*y.y;\n";
            o << "      " << a << ".y = " << c(0) << "*x.y + " << c(1) << "*x.x + " << c(2) << "*y.y + " << c(3)
              << "*y.x;\n";
            o << "      " << b << ".x = " << c(4) << "*x.x - " << c(5) << "*x.y + " << c(6) << "*y.x - " << c(7)
              << "*y.y;\n";
            o << "      " << b << ".y = " << c(4) << "*x.y + " << c(5) << "*x.x + " << c(6) << "*y.y + " << c(7)
              << "*y.x; }\n";
        }
        o << "  }\n";
        break;
    }
    case TK_U2: {
        const uint32_t A = op.j0, B = op.j1;
        const uint32_t k0 = g.push(op.m[0].real());
        g.push(op.m[0].imag());
        for (int t = 1; t < 16; t++) {
            g.push(op.m[t].real());
            g.push(op.m[t].imag());
        }
        o << "  {\n";
        for (uint32_t i = 0; i < NR; i++) {
            if (i & ((1u << A) | (1u << B))) continue;
            const uint32_t ix[4] = {i, i | (1u << A), i | (1u << B), i | (1u << A) | (1u << B)};
            o << "    { const V x0 = " << g.var(ix[0]) << ", x1 = " << g.var(ix[1]) << ", x2 = " << g.var(ix[2])
              << ", x3 = " << g.var(ix[3]) << ";\n";
            for (int a = 0; a < 4; a++) {
                std::ostringstream re, im;
                for (int b = 0; b < 4; b++) {
                    const std::string mr = g.C(k0 + 2 * (4 * a + b)), mi = g.C(k0 + 2 * (4 * a + b) + 1);
                    re << (b ? " + " : "") << mr << "*x" << b << ".x - " << mi << "*x" << b << ".y";
                    im << (b ? " + " : "") << mr << "*x" << b << ".y + " << mi << "*x" << b << ".x";
                }
                o << "      " << g.var(ix[a]) << ".x = " << re.str() << ";\n";
                o << "      " << g.var(ix[a]) << ".y = " << im.str() << ";\n";
            }
            o << "    }\n";
        }
        o << "  }\n";
        break;
    }
    case TK_DM: {
        const std::string cond = cond_expr(g, op.gmask, cur);
        const cd ph = op.m[0];
        std::string pre = "  ";
        if (!cond.empty()) {
            o << "  if " << cond << " {\n";
            pre = "    ";
        }
        uint32_t kr = 0;
        const bool neg = is_exact(ph, -1.0, 0.0), pi = is_exact(ph, 0.0, 1.0), mi = is_exact(ph, 0.0, -1.0);
        if (!neg && !pi && !mi) {
            kr = g.push(ph.real());
            g.push(ph.imag());
        }
        for (uint32_t i = 0; i < NR; i++) {
            if ((i & op.rmask) != op.rmask) continue;
            const std::string a = g.var(i);
            if (neg)
                o << pre << a << ".x = -" << a << ".x; " << a << ".y = -" << a << ".y;\n";
            else if (pi || mi) {
                o << pre << "{ const V x = " << a << "; ";
                ipow_assign(o, a, "x", pi ? 1 : 3);
                o << " }\n";
            } else
                o << pre << "{ const V x = " << a << "; " << a << ".x = " << g.C(kr) << "*x.x - " << g.C(kr + 1)
                  << "*x.y; " << a << ".y = " << g.C(kr) << "*x.y + " << g.C(kr + 1) << "*x.x; }\n";
        }
        if (!cond.empty()) o << "  }\n";
        break;
    }
    case TK_PERM1: {
        const std::string cond = cond_expr(g, op.gmask, cur);
        const uint32_t J = op.j0;
        if (cond.empty() && op.k0 == 0 && op.k1 == 0) {
            // a plain move: rename the registers
            for (uint32_t i = 0; i < NR; i++) {
                if (i & (1u << J)) continue;
                if ((i & op.rmask) != op.rmask) continue;
                std::swap(g.name[i], g.name[i | (1u << J)]);
            }
            break;
        }
        std::string pre = "  ";
        if (!cond.empty()) {
            o << "  if " << cond << " {\n";
            pre = "    ";
        }
        for (uint32_t i = 0; i < NR; i++) {
            if (i & (1u << J)) continue;
            if ((i & op.rmask) != op.rmask) continue;
            const std::string a = g.var(i), b = g.var(i | (1u << J));
            o << pre << "{ const V x = " << a << ", y = " << b << "; ";
            ipow_assign(o, a, "y", op.k1);
            o << " ";
            ipow_assign(o, b, "x", op.k0);
            o << " }\n";
        }
        if (!cond.empty()) o << "  }\n";
        break;
    }
    case TK_SWAPR:
        for (uint32_t i = 0; i < NR; i++)
            if ((i & (1u << op.j0)) && !(i & (1u << op.j1)))
                std::swap(g.name[i], g.name[(i & ~(1u << op.j0)) | (1u << op.j1)]);
        break;
    default: fail(NCSV_INVALID, "jit: unknown op");
    }
}

// A run of consecutive diagonal ops (phase c where the register bits hold
// rmask and the other index bits hold gmask) as one phase per amplitude:
// per register mask m, P_m = (host product of the unconditional phases) x
// (thread product of the conditional ones), and every amplitude r is
// multiplied by the P_m with m inside r.  Products by 1, -1, i, -i stay
// exact.
void emit_diag_run(Gen &g, const std::vector<const TileOpH *> &run, uint32_t cur) {
    auto &o = g.o;
    std::map<uint32_t, cd> uncond;
    std::map<std::pair<uint32_t, uint64_t>, cd> cond;
    for (const TileOpH *op : run) {
        if (op->gmask == 0) {
            auto it = uncond.find(op->rmask);
            uncond[op->rmask] = it == uncond.end() ? op->m[0] : it->second * op->m[0];
        } else {
            const auto key = std::make_pair(op->rmask, op->gmask);
            auto it = cond.find(key);
            cond[key] = it == cond.end() ? op->m[0] : it->second * op->m[0];
        }
    }
    std::vector<uint32_t> masks;
    for (auto &kv : uncond)
        if (kv.second != cd(1.0, 0.0)) masks.push_back(kv.first);
    for (auto &kv : cond)
        if (std::find(masks.begin(), masks.end(), kv.first.first) == masks.end()) masks.push_back(kv.first.first);
    if (masks.empty()) return;
    o << "  { // diagonal run of " << run.size() << " ops\n";
    auto cm = [&](const std::string &p, const std::string &re, const std::string &im) {
        o << "    { const T pr = " << p << ".x, pi = " << p << ".y; " << p << ".x = pr*" << re << " - pi*" << im << "; "
          << p << ".y = pr*" << im << " + pi*" << re << "; }\n";
    };
    for (uint32_t m : masks) {
        auto it = uncond.find(m);
        const cd u = it == uncond.end() ? cd(1.0, 0.0) : it->second;
        if (u == cd(1.0, 0.0)) {
            o << "    V P" << m << "; P" << m << ".x = 1; P" << m << ".y = 0;\n";
        } else {
            const uint32_t k = g.push(u.real());
            g.push(u.imag());
            o << "    V P" << m << "; P" << m << ".x = " << g.C(k) << "; P" << m << ".y = " << g.C(k + 1) << ";\n";
        }
    }
    for (auto &kv : cond) {
        if (kv.second == cd(1.0, 0.0)) continue;
        const uint32_t m = kv.first.first;
        const std::string ce = cond_expr(g, kv.first.second, cur);
        const uint32_t k = g.push(kv.second.real());
        g.push(kv.second.imag());
        o << "    if " << ce << " {\n  ";
        cm("P" + std::to_string(m), g.C(k), g.C(k + 1));
        o << "    }\n";
    }
    for (uint32_t r = 0; r < g.NR; r++) {
        for (uint32_t m : masks) {
            if ((r & m) != m) continue;
            const std::string a = g.var(r);
            o << "    { const V x = " << a << "; " << a << ".x = P" << m << ".x*x.x - P" << m << ".y*x.y; " << a
              << ".y = P" << m << ".x*x.y + P" << m << ".y*x.x; }\n";
        }
    }
    o << "  }\n";
}

// Folding a diagonal run pays when it needs fewer complex products per thread
// than applying its ops one by one (sign flips and quarter turns count a
// third of a product).
bool diag_run_pays(const std::vector<const TileOpH *> &run, uint32_t NR) {
    auto nreg = [&](uint32_t m) {
        uint32_t k = 0;
        for (uint32_t r = 0; r < NR; r++) k += (r & m) == m;
        return k;
    };
    double single = 0.0;
    std::map<uint32_t, int> masks;
    std::map<std::pair<uint32_t, uint64_t>, int> conds;
    for (const TileOpH *op : run) {
        const cd c = op->m[0];
        const bool special = c == cd(-1, 0) || c == cd(0, 1) || c == cd(0, -1);
        single += nreg(op->rmask) * (special ? 0.34 : 1.0);
        masks[op->rmask] = 1;
        if (op->gmask) conds[{op->rmask, op->gmask}] = 1;
    }
    double folded = (double)conds.size();
    for (auto &kv : masks) folded += nreg(kv.first);
    return folded < single;
}

// Diagonal ops applied lazily: a diagonal phase on the registers holding
// mask m (with conditions on thread or outer bits) commutes with every op
// that does not act non-diagonally on a register bit of m, and with every
// other diagonal op.  So each op only multiplies a pending factor for its
// mask (a compile-time constant, times a per-thread value when conditioned),
// and the factors are applied to the amplitudes when an op is about to act
// on one of their bits, before an LDS exchange, or at the store.  Factors on
// the same mask from many ops (a QFT's controlled phases between two
// Hadamards) then cost one complex product per amplitude; factors that
// multiply to 1, -1 or +-i cost nothing or an exact sign or swap.
bool use_lazy_diagonals() {
    const char *e = getenv("NCSV_JIT_LAZY");
    return !(e && atoi(e) == 0);
}

struct Pending {
    cd k = cd(1.0, 0.0);  // compile-time part
    int rt = -1;          // per-thread part: variable D<rt>, or none
    int sg = -1;          // per-thread sign (conditioned -1s): sign-bit mask S<sg>, or none
    std::vector<std::pair<uint64_t, cd>> conds;  // conditioned phases (condition bits, phase)
};

void emit_cmul_var(std::ostringstream &o, const std::string &v, const std::string &re, const std::string &im,
                   const char *pre) {
    o << pre << "{ const T pr = " << v << ".x, pi = " << v << ".y; " << v << ".x = pr*" << re << " - pi*" << im << "; "
      << v << ".y = pr*" << im << " + pi*" << re << "; }\n";
}

// The thread-index bits (tmask) and block-index bits (bmask: bit i of b is
// the outer qubit oq[i]) a condition on global index bits tests.
void split_cond(const Gen &g, uint64_t gmask, uint32_t cur, uint32_t *tmask, uint64_t *bmask) {
    const std::vector<uint32_t> tp = positions(cur, g.L, false);
    *tmask = 0;
    *bmask = 0;
    for (uint64_t a = gmask; a; a &= a - 1) {
        const uint32_t q = (uint32_t)__builtin_ctzll(a);
        int pos = -1;
        for (uint32_t p = 0; p < g.L; p++)
            if (g.tq[p] == q) pos = (int)p;
        if (pos >= 0) {
            for (size_t i = 0; i < tp.size(); i++)
                if (tp[i] == (uint32_t)pos) *tmask |= 1u << i;
        } else {
            for (size_t i = 0; i < g.oq.size(); i++)
                if (g.oq[i] == q) *bmask |= 1ull << i;
        }
    }
}

// The per-thread product of conditioned phases as V D<id>: conditions on
// thread-index bits only go into one table over the thread index (a vector
// load), conditions within one group of 8 block-index bits into a table
// over that group (a load at a workgroup-uniform address), and the rest are
// conditioned products.  Tables sit in the coefficient buffer and are used
// where they replace two or more products.
std::string build_runtime(Gen &g, const std::vector<std::pair<uint64_t, cd>> &conds, uint32_t cur, int id) {
    auto &o = g.o;
    const std::string d = "D" + std::to_string(id);
    std::vector<std::pair<uint32_t, cd>> thr;                       // (thread-index mask, phase)
    std::map<uint32_t, std::vector<std::pair<uint32_t, cd>>> grp;  // group -> (mask in the group, phase)
    std::vector<std::pair<uint64_t, cd>> rest;                      // conditioned products
    std::vector<std::pair<uint64_t, cd>> thr_src;
    std::map<uint32_t, std::vector<std::pair<uint64_t, cd>>> grp_src;
    for (auto &kv : conds) {
        uint32_t tm;
        uint64_t bm;
        split_cond(g, kv.first, cur, &tm, &bm);
        if (!bm) {
            thr.push_back({tm, kv.second});
            thr_src.push_back(kv);
            continue;
        }
        const uint32_t gi = (uint32_t)__builtin_ctzll(bm) / 8;
        if (!tm && (bm >> (8 * gi)) < 256) {
            grp[gi].push_back({(uint32_t)(bm >> (8 * gi)), kv.second});
            grp_src[gi].push_back(kv);
            continue;
        }
        rest.push_back(kv);
    }
    auto table = [&](uint32_t bits, const std::vector<std::pair<uint32_t, cd>> &ops) {
        if (g.coef->size() & 1) g.push(0.0);
        const uint32_t i0 = (uint32_t)g.coef->size();
        for (uint32_t v = 0; v < (1u << bits); v++) {
            cd t(1.0, 0.0);
            for (auto &op : ops)
                if ((v & op.first) == op.first) t *= op.second;
            g.push(t.real());
            g.push(t.imag());
        }
        return i0;
    };
    if (thr.size() >= 2) {
        const uint32_t i0 = table(g.LT, thr);
        o << "  V " << d << "; " << d << ".x = c[" << i0 << "u + 2u*tid]; " << d << ".y = c[" << i0 << "u + 2u*tid + 1u];\n";
    } else {
        rest.insert(rest.end(), thr_src.begin(), thr_src.end());
        o << "  V " << d << "; " << d << ".x = 1; " << d << ".y = 0;\n";
    }
    for (auto &kv : grp) {
        if (kv.second.size() < 2) {
            rest.insert(rest.end(), grp_src[kv.first].begin(), grp_src[kv.first].end());
            continue;
        }
        const uint32_t gi = kv.first;
        const uint32_t bits = std::min<uint32_t>(8, (uint32_t)g.oq.size() - 8 * gi);
        const uint32_t i0 = table(bits, kv.second);
        o << "  { const unsigned ig = (unsigned)(b >> " << 8 * gi << ") & " << ((1u << bits) - 1) << "u; const T tr = c["
          << i0 << "u + 2u*ig], ti = c[" << i0 << "u + 2u*ig + 1u];\n";
        emit_cmul_var(o, d, "tr", "ti", "    ");
        o << "  }\n";
    }
    for (auto &kv : rest) {
        const cd c = kv.second;
        o << "  if " << cond_expr(g, kv.first, cur) << " {\n";
        if (c == cd(0.0, 1.0) || c == cd(0.0, -1.0)) {
            o << "    { const V x = " << d << "; ";
            ipow_assign(o, d, "x", c.imag() > 0 ? 1 : 3);
            o << " }\n";
        } else {
            const uint32_t kc = g.push(c.real());
            g.push(c.imag());
            emit_cmul_var(o, d, g.C(kc), g.C(kc + 1), "    ");
        }
        o << "  }\n";
    }
    return d;
}

// XOR of a sign-bit mask into both components of v (exact negation where set).
void emit_sign_xor(Gen &g, const std::string &v, const std::string &mask, const char *pre) {
    auto &o = g.o;
    if (g.dbl) {
        for (const char *c : {".x", ".y"})
            o << pre << v << c << " = __hiloint2double(__double2hiint(" << v << c << ") ^ (int)" << mask << ", __double2loint("
              << v << c << "));\n";
    } else {
        for (const char *c : {".x", ".y"})
            o << pre << v << c << " = __int_as_float(__float_as_int(" << v << c << ") ^ (int)" << mask << ");\n";
    }
}


// Multiplies every register holding mask m by the pending factor p.
void apply_pending(Gen &g, uint32_t m, const Pending &p, uint32_t cur) {
    auto &o = g.o;
    auto exact = [](cd c, double re, double im) { return c.real() == re && c.imag() == im; };
    if (p.rt < 0 && p.sg >= 0 && (exact(p.k, 1.0, 0.0) || exact(p.k, -1.0, 0.0))) {
        // a sign per thread (and a constant sign): one XOR per component
        std::string mask = "S" + std::to_string(p.sg);
        if (exact(p.k, -1.0, 0.0)) {
            o << "  " << mask << " ^= 0x80000000u;\n";
        }
        for (uint32_t r = 0; r < g.NR; r++)
            if ((r & m) == m) emit_sign_xor(g, g.var(r), mask, "  ");
        return;
    }
    if (p.rt < 0 && p.sg >= 0) {
        // a general constant with a per-thread sign: the signed constant per thread
        const uint32_t k = g.push(p.k.real());
        g.push(p.k.imag());
        o << "  V DS" << p.sg << "; DS" << p.sg << ".x = " << g.C(k) << "; DS" << p.sg << ".y = " << g.C(k + 1) << ";\n";
        emit_sign_xor(g, "DS" + std::to_string(p.sg), "S" + std::to_string(p.sg), "  ");
        for (uint32_t r = 0; r < g.NR; r++) {
            if ((r & m) != m) continue;
            const std::string a = g.var(r), d = "DS" + std::to_string(p.sg);
            o << "  { const V x = " << a << "; " << a << ".x = " << d << ".x*x.x - " << d << ".y*x.y; " << a << ".y = "
              << d << ".x*x.y + " << d << ".y*x.x; }\n";
        }
        return;
    }
    if (p.rt < 0) {
        const cd c = p.k;
        if (exact(c, 1.0, 0.0)) return;
        const bool neg = exact(c, -1.0, 0.0), pi = exact(c, 0.0, 1.0), mi = exact(c, 0.0, -1.0);
        uint32_t kr = 0;
        if (!neg && !pi && !mi) {
            kr = g.push(c.real());
            g.push(c.imag());
        }
        for (uint32_t r = 0; r < g.NR; r++) {
            if ((r & m) != m) continue;
            const std::string a = g.var(r);
            if (neg)
                o << "  " << a << ".x = -" << a << ".x; " << a << ".y = -" << a << ".y;\n";
            else if (pi || mi) {
                o << "  { const V x = " << a << "; ";
                ipow_assign(o, a, "x", pi ? 1 : 3);
                o << " }\n";
            } else
                o << "  { const V x = " << a << "; " << a << ".x = " << g.C(kr) << "*x.x - " << g.C(kr + 1) << "*x.y; "
                  << a << ".y = " << g.C(kr) << "*x.y + " << g.C(kr + 1) << "*x.x; }\n";
        }
        return;
    }
    const std::string d = build_runtime(g, p.conds, cur, p.rt);
    if (p.sg >= 0) emit_sign_xor(g, d, "S" + std::to_string(p.sg), "  ");
    if (!exact(p.k, 1.0, 0.0)) {
        const uint32_t k = g.push(p.k.real());
        g.push(p.k.imag());
        emit_cmul_var(o, d, g.C(k), g.C(k + 1), "  ");
    }
    for (uint32_t r = 0; r < g.NR; r++) {
        if ((r & m) != m) continue;
        const std::string a = g.var(r);
        o << "  { const V x = " << a << "; " << a << ".x = " << d << ".x*x.x - " << d << ".y*x.y; " << a << ".y = "
          << d << ".x*x.y + " << d << ".y*x.x; }\n";
    }
}

uint32_t swap_bits(uint32_t m, uint32_t a, uint32_t b) {
    const uint32_t ba = (m >> a) & 1u, bb = (m >> b) & 1u;
    if (ba == bb) return m;
    return m ^ ((1u << a) | (1u << b));
}

// ---------------------------------------------- FP32 MFMA dense blocks (complex64)
//
// A run of a stage's gate ops on at most four register bits becomes one
// 16x16 complex unitary U, composed on the host in FP64, and is applied with
// v_mfma_f32_32x32x2f32 as the real 32x32 form [[Re U, -Im U], [Im U, Re U]]:
// one block register bit b3 is first exchanged with lane bit 5 (one
// v_permlane32_swap per register pair and component), so each 32-lane column
// n = lane & 31 holds a 16-amplitude vector as 32 reals, 16 per lane half
// (slot s = component + 2 * the bits b0..b2, half h = the bit b3).  The
// instruction's B operand j is slot j of each lane (K = 2j + h), its D
// register i is the output row 8(i>>2) + 4h + (i&3), so with A's rows and
// columns ordered accordingly every lane gets back the slots it supplied; the
// four groups of the other two register bits take 16 instructions each.  A
// lane's A operand j is A[lane & 31][2j + (lane >> 5)], read from the
// coefficient buffer (16 coalesced loads per block).  Lane layouts checked on
// MI355X (gfx950): D register i, lane l = D[8(i/4) + 4(l/32) + i%4][l%32];
// permlane32_swap(x, y) returns (x lanes 0-31 | y lanes 0-31) and
// (x lanes 32-63 | y lanes 32-63).
//
// At the rated 64 cycles per instruction a block's 16 x 4 MFMAs per wavefront
// take the time of ~32 FP32 vector instructions per amplitude, so runs whose
// ops cost more (NCSV_JIT_MFMA_COST, default 36: three generic 4x4s) would
// gain.  Measured on MI355X they do not: on a 30-qubit random SU(4)
// brickwork (complex64, 25 sweeps, blocks of three 4x4s) the sweeps take
// 6.37 ms with blocks against 4.87 ms without (blocks of four 4x4s: 4.92
// against 4.87 ms), so blocks are off by default.  NCSV_JIT_MFMA: 0 never
// (default), 1 by the cost above, 2 for every run with arithmetic.
int mfma_mode() {
    const char *e = getenv("NCSV_JIT_MFMA");
    return e ? atoi(e) : 0;
}

const char *MFMA_HELPERS =
    "typedef float f16v __attribute__((ext_vector_type(16)));  // @mfma\n"
    "__device__ __forceinline__ f16v mfma32(float a, float b, f16v c) { return "
    "__builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }  // @mfma\n"
    "__device__ __forceinline__ void swap32(float &x, float &y) { auto r = "
    "__builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false); "
    "x = __uint_as_float(r[0]); y = __uint_as_float(r[1]); }  // @mfma\n";

bool mfma_absorbable(const TileOpH &op) {
    switch (op.kind) {
    case TK_U1R:
    case TK_U1X:
    case TK_U1:
    case TK_U2:
    case TK_SWAPR: return true;
    case TK_PERM1:
    case TK_DM: return op.gmask == 0;
    default: return false;
    }
}

uint32_t mfma_op_bits(const TileOpH &op) {
    switch (op.kind) {
    case TK_U1R:
    case TK_U1X:
    case TK_U1: return 1u << op.j0;
    case TK_U2:
    case TK_SWAPR: return (1u << op.j0) | (1u << op.j1);
    case TK_PERM1: return (1u << op.j0) | op.rmask;
    default: return op.rmask;
    }
}

// Vector instructions per amplitude an op costs in the register sweeps.
double valu_cost(const TileOpH &op) {
    switch (op.kind) {
    case TK_U1R:
    case TK_U1X: return 2.0;
    case TK_U1: return 8.0;
    case TK_U2: return 16.0;
    case TK_PERM1: return (op.k0 == 0 && op.k1 == 0 && op.rmask == 0) ? 0.0 : 1.0;
    case TK_SWAPR: return 0.0;
    default: return 0.5;
    }
}
double mfma_block_cost() {
    const char *e = getenv("NCSV_JIT_MFMA_COST");
    return e ? atof(e) : 36.0;
}

// U (16x16, U[row * 16 + col]) of the ops on register bits bb[0..3] (local
// index bit k = register bit bb[k]), composed in FP64 in program order.
void mfma_block_matrix(const std::vector<const TileOpH *> &ops, const uint32_t bb[4], cd *U) {
    auto lb = [&](uint32_t j) {
        for (int k = 0; k < 4; k++)
            if (bb[k] == j) return (uint32_t)k;
        fail(NCSV_INVALID, "jit: MFMA block op outside the block");
        return 0u;
    };
    auto lmask = [&](uint32_t rmask) {
        uint32_t m = 0;
        for (uint32_t a = rmask; a; a &= a - 1) m |= 1u << lb((uint32_t)__builtin_ctz(a));
        return m;
    };
    auto ip = [](int k, cd a) {
        switch (k & 3) {
        case 0: return a;
        case 1: return cd(-a.imag(), a.real());
        case 2: return -a;
        default: return cd(a.imag(), -a.real());
        }
    };
    for (uint32_t col = 0; col < 16; col++) {
        cd v[16];
        for (uint32_t a = 0; a < 16; a++) v[a] = a == col ? cd(1.0, 0.0) : cd(0.0, 0.0);
        for (const TileOpH *o : ops) {
            const TileOpH &op = *o;
            switch (op.kind) {
            case TK_U1R:
            case TK_U1X:
            case TK_U1: {
                const uint32_t J = 1u << lb(op.j0);
                for (uint32_t a = 0; a < 16; a++) {
                    if (a & J) continue;
                    const cd x = v[a], y = v[a | J];
                    v[a] = op.m[0] * x + op.m[1] * y;
                    v[a | J] = op.m[2] * x + op.m[3] * y;
                }
                break;
            }
            case TK_U2: {
                const uint32_t A = 1u << lb(op.j0), B = 1u << lb(op.j1);
                for (uint32_t a = 0; a < 16; a++) {
                    if (a & (A | B)) continue;
                    const uint32_t ix[4] = {a, a | A, a | B, a | A | B};
                    cd x[4], y[4];
                    for (int t = 0; t < 4; t++) x[t] = v[ix[t]];
                    for (int t = 0; t < 4; t++) {
                        y[t] = 0.0;
                        for (int u = 0; u < 4; u++) y[t] += op.m[4 * t + u] * x[u];
                    }
                    for (int t = 0; t < 4; t++) v[ix[t]] = y[t];
                }
                break;
            }
            case TK_PERM1: {
                const uint32_t J = 1u << lb(op.j0), cm = lmask(op.rmask);
                for (uint32_t a = 0; a < 16; a++) {
                    if ((a & J) || (a & cm) != cm) continue;
                    const cd x = v[a], y = v[a | J];
                    v[a] = ip(op.k1, y);
                    v[a | J] = ip(op.k0, x);
                }
                break;
            }
            case TK_SWAPR: {
                const uint32_t A = 1u << lb(op.j0), B = 1u << lb(op.j1);
                for (uint32_t a = 0; a < 16; a++)
                    if ((a & A) && !(a & B)) std::swap(v[a], v[(a & ~A) | B]);
                break;
            }
            case TK_DM: {
                const uint32_t cm = lmask(op.rmask);
                for (uint32_t a = 0; a < 16; a++)
                    if ((a & cm) == cm) v[a] *= op.m[0];
                break;
            }
            default: fail(NCSV_INVALID, "jit: op kind not composable into an MFMA block");
            }
        }
        for (uint32_t r = 0; r < 16; r++) U[r * 16 + col] = v[r];
    }
}

// Emits the block of `ops` on register bits `bits` (at most four; padded with
// other register bits).  Complex64 tiles with 32 registers per thread only.
void emit_mfma_block(Gen &g, const std::vector<const TileOpH *> &ops, uint32_t bits) {
    auto &o = g.o;
    uint32_t bb[4], nb = 0;
    for (uint32_t j = 0; j < g.R && nb < 4; j++)
        if ((bits >> j) & 1) bb[nb++] = j;
    for (uint32_t j = 0; j < g.R && nb < 4; j++)
        if (!((bits >> j) & 1)) bits |= 1u << (bb[nb++] = j);
    std::sort(bb, bb + 4);
    uint32_t u = 0;
    for (uint32_t j = 0; j < g.R; j++)
        if (!((bits >> j) & 1)) u = j;
    const uint32_t bs = bb[3];
    cd U[256];
    mfma_block_matrix(ops, bb, U);
    // register of (group gi over bits u, bs) and block index t over b0..b2
    auto reg = [&](uint32_t gi, uint32_t t) {
        uint32_t r = 0;
        for (int k = 0; k < 3; k++)
            if ((t >> k) & 1) r |= 1u << bb[k];
        if (gi & 1) r |= 1u << u;
        if (gi & 2) r |= 1u << bs;
        return r;
    };
    // real element (h, s): amplitude a = (s >> 1) | (h << 3), component s & 1
    auto ureal = [&](uint32_t h, uint32_t s, uint32_t h2, uint32_t s2) {
        const cd z = U[(((s >> 1) | (h << 3)) * 16) + ((s2 >> 1) | (h2 << 3))];
        const uint32_t c = s & 1, c2 = s2 & 1;
        if (c == 0 && c2 == 0) return z.real();
        if (c == 0 && c2 == 1) return -z.imag();
        if (c == 1 && c2 == 0) return z.imag();
        return z.real();
    };
    while (g.coef->size() % 64) g.push(0.0);
    const uint32_t off = (uint32_t)g.coef->size();
    for (uint32_t j = 0; j < 16; j++)
        for (uint32_t l = 0; l < 64; l++) {
            const uint32_t M = l & 31, h = (M >> 2) & 1, s = (M & 3) | ((M >> 3) << 2);
            g.push(ureal(h, s, l >> 5, j));
        }
    auto swap_bs = [&]() {
        for (uint32_t r = 0; r < g.NR; r++) {
            if ((r >> bs) & 1) continue;
            const std::string a = g.var(r), b = g.var(r | (1u << bs));
            o << "    swap32(" << a << ".x, " << b << ".x); swap32(" << a << ".y, " << b << ".y);\n";
        }
    };
    o << "  { // MFMA block on register bits " << bb[0] << " " << bb[1] << " " << bb[2] << " " << bb[3] << " (" << ops.size()
      << " ops)\n";
    swap_bs();
    o << "    const unsigned ln = tid & 63u;\n";
    for (uint32_t j = 0; j < 16; j++) o << "    const float A" << j << " = c[" << off + 64 * j << "u + ln];\n";
    for (uint32_t gi = 0; gi < 4; gi++) {
        o << "    { f16v m = {};\n";
        for (uint32_t j = 0; j < 16; j++)
            o << "      m = mfma32(A" << j << ", " << g.var(reg(gi, j >> 1)) << ((j & 1) ? ".y" : ".x") << ", m);\n";
        for (uint32_t i = 0; i < 16; i++)
            o << "      " << g.var(reg(gi, i >> 1)) << ((i & 1) ? ".y" : ".x") << " = m[" << i << "];\n";
        o << "    }\n";
    }
    swap_bs();
    o << "  }\n";
    g.mfma_blocks++;
}

void emit_stages_lazy(Gen &g, const SweepH &sw, uint32_t &cur, uint32_t memr) {
    auto &o = g.o;
    std::map<uint32_t, Pending> pend;
    int nvars = 0;
    auto flush = [&](uint32_t touched, bool all) {
        for (auto it = pend.begin(); it != pend.end();) {
            if (!all && !(it->first & touched)) {
                ++it;
                continue;
            }
            apply_pending(g, it->first, it->second, cur);
            it = pend.erase(it);
        }
    };
    for (auto &St : sw.stages) {
        if (St.rmask_pos != cur) {
            flush(0, true);
            emit_exchange(g, cur, St.rmask_pos);
            cur = St.rmask_pos;
        }
        auto process = [&](const TileOpH &op) {
            switch (op.kind) {
            case TK_DM: {
                Pending &p = pend[op.rmask];
                if (!op.gmask) {
                    p.k *= op.m[0];
                    break;
                }
                const cd c = op.m[0];
                if (c == cd(-1.0, 0.0)) {
                    // a conditioned sign flip: a sign-bit mask per thread
                    if (p.sg < 0) {
                        p.sg = nvars++;
                        o << "  unsigned S" << p.sg << " = 0u;\n";
                    }
                    o << "  if " << cond_expr(g, op.gmask, cur) << " S" << p.sg << " ^= 0x80000000u;\n";
                    break;
                }
                // a conditioned phase: multiplied into the factor when it is applied
                if (p.rt < 0) p.rt = nvars++;
                p.conds.push_back({op.gmask, c});
                break;
            }
            case TK_SWAPR: {
                // the pending masks follow the renamed register bits
                std::map<uint32_t, Pending> moved;
                for (auto &kv : pend) moved[swap_bits(kv.first, op.j0, op.j1)] = kv.second;
                pend.swap(moved);
                emit_op(g, op, cur);
                break;
            }
            case TK_U2: flush((1u << op.j0) | (1u << op.j1), false); emit_op(g, op, cur); break;
            default: flush(1u << op.j0, false); emit_op(g, op, cur); break;
            }
        };
        // runs of ops on at most four register bits, and the conditioned
        // phases on other register bits met inside them (those commute with
        // the run), as MFMA blocks where that is cheaper
        const bool mf = !g.dbl && g.R == 5 && mfma_mode() > 0;
        std::vector<const TileOpH *> run;
        uint32_t ubits = 0, forbid = 0;
        double vcost = 0.0;
        auto close_run = [&]() {
            if (run.empty()) return;
            const bool use = mfma_mode() == 2 ? vcost > 0.0 : vcost > mfma_block_cost();
            if (use) {
                std::vector<const TileOpH *> blk;
                for (const TileOpH *op : run) {
                    if (mfma_absorbable(*op))
                        blk.push_back(op);
                    else
                        process(*op);  // a conditioned phase off the block's bits: pending
                }
                flush(ubits, false);
                emit_mfma_block(g, blk, ubits);
            } else {
                for (const TileOpH *op : run) process(*op);
            }
            run.clear();
            ubits = forbid = 0;
            vcost = 0.0;
        };
        for (uint32_t k = St.op_begin; k < St.op_end; k++) {
            const TileOpH &op = sw.ops[k];
            if (mf) {
                if (mfma_absorbable(op)) {
                    const uint32_t b = mfma_op_bits(op);
                    if ((b & forbid) || __builtin_popcount(ubits | b) > 4) close_run();
                    run.push_back(&op);
                    ubits |= b;
                    vcost += valu_cost(op);
                    continue;
                }
                if (op.kind == TK_DM && op.gmask && !(op.rmask & ubits) && !run.empty()) {
                    run.push_back(&op);
                    forbid |= op.rmask;
                    continue;
                }
                close_run();
            }
            process(op);
        }
        close_run();
    }
    flush(0, true);
    if (cur != memr) {
        emit_exchange(g, cur, memr);
        cur = memr;
    }
}

}  // namespace

uint32_t default_memr(const TileGeom &geo) { return ((1u << geo.R) - 1) << geo.LT; }

// Workgroup order over the tiles: XCD-chunked by default (NCSV_JIT_BORDER=0:
// ascending block index).  Workgroups are dispatched to the 8 XCDs in turn,
// so with the block index b taken as ((b & 7) << (nb - 3)) | (b >> 3) each
// XCD walks one contiguous run of tiles and the workgroups in flight on it
// touch neighbouring tiles.  On MI355X this took the 30-qubit complex128
// brickwork's sweeps from 6.97 to 6.74 ms and the 33-qubit complex64 QFT's
// from 24.6 to 21.9 ms.
int default_border() {
    static const int v = getenv("NCSV_JIT_BORDER") ? atoi(getenv("NCSV_JIT_BORDER")) : 1;
    return v;
}

std::string generate(const SweepH &sw, uint32_t n, const TileGeom &geo, bool dbl, std::vector<double> *coef,
                     bool from_basis, const MemShape *ms) {
    Gen g;
    g.coef = coef;
    g.literal = use_literals();
    g.dbl = dbl;
    g.LT = geo.LT;
    g.R = geo.R;
    g.L = geo.L();
    g.NR = 1u << geo.R;
    g.tq = sw.tq;
    g.name.resize(g.NR);
    for (uint32_t r = 0; r < g.NR; r++) g.name[r] = (int)r;
    uint64_t in = 0;
    for (uint32_t q : sw.tq) in |= 1ull << q;
    for (uint32_t q = 0; q < n; q++)
        if (!((in >> q) & 1)) g.oq.push_back(q);
    // the register positions of the loads and of the stores (no register on
    // the fixed low positions, so those stay thread bits 0..c-1 and every
    // wavefront access covers runs of 2^c amplitudes) and the block order
    const uint32_t memr = ms && ms->load ? ms->load : default_memr(geo);
    const uint32_t mems = ms && ms->store ? ms->store : default_memr(geo);
    if ((memr | mems) & ((1u << geo.c) - 1) || (uint32_t)__builtin_popcount(memr) != g.R ||
        (uint32_t)__builtin_popcount(mems) != g.R || (memr | mems) >> g.L)
        fail(NCSV_INVALID, "jit: bad memory layout");
    const int border = ms && ms->border >= 0 ? ms->border : default_border();
    const uint32_t nb = n - g.L;
    bool lds = false;
    {
        uint32_t cur = memr;
        for (auto &S : sw.stages) {
            if (S.rmask_pos != cur) lds = true;
            cur = S.rmask_pos;
        }
        if (cur != mems) lds = true;
    }
    auto &o = g.o;
    const int S = dbl ? 4 : 5;
    g.S = S;
    g.half = use_half();
    const int waves = jit_waves();
    const int nt = jit_nt();
    o << "typedef " << (dbl ? "double" : "float") << " T;\ntypedef " << (dbl ? "double2" : "float2") << " V;\n";
    if (nt) o << NT_HELPERS;
    o << "__device__ __forceinline__ unsigned swz(unsigned t) { return t ^ (((t >> " << S << ") ^ (t >> " << 2 * S
      << ")) & " << ((1u << S) - 1) << "u); }\n";
    const std::string header = o.str();  // the MFMA helpers go here if the body uses them
    o.str("");
    o << "extern \"C\" __global__ __launch_bounds__(" << (1u << g.LT);
    if (waves) o << ", " << waves;
    // a regular sweep reads amp and writes out (the same buffer, or the
    // state's second one when it writes out of place: no restrict on either)
    if (from_basis)
        o << ") void ncsv_sweep_from_basis(V *__restrict__ amp, const T *__restrict__ c, unsigned long long binit) {\n";
    else
        o << ") void ncsv_sweep(V *amp, const T *__restrict__ c, V *out) {\n";
    if (lds) o << "  __shared__ V lds[" << (1u << (g.half ? g.L - 1 : g.L)) << "];\n";
    o << "  const unsigned tid = threadIdx.x;\n";
    if (border == 1 && nb >= 3)
        // XCD-chunked order: workgroups go to the 8 XCDs in turn, so XCD x
        // takes the contiguous run of tiles x * 2^(nb-3) ...
        o << "  const unsigned long long bx = blockIdx.x;\n  const unsigned long long b = ((bx & 7ull) << " << nb - 3
          << ") | (bx >> 3);\n";
    else
        o << "  const unsigned long long b = blockIdx.x;\n";
    o << "  const unsigned long long base = " << deposit("b", g.oq, true) << ";\n";
    // thread bit i holds the i-th non-register position of a layout, register j its j-th register position
    auto thread_qubits = [&](uint32_t lay) {
        std::vector<uint32_t> q;
        for (uint32_t p : positions(lay, g.L, false)) q.push_back(sw.tq[p]);
        return q;
    };
    const std::vector<uint32_t> low = thread_qubits(memr), low_s = thread_qubits(mems);
    const std::vector<uint32_t> rpl = positions(memr, g.L, true), rps = positions(mems, g.L, true);
    {
        std::ostringstream e;
        e << "base";
        for (uint32_t i = 0; i < g.LT; i++) e << " | (((unsigned long long)(tid >> " << i << ") & 1ull) << " << low[i] << ")";
        o << "  const unsigned long long gm = " << e.str() << ";\n";
    }
    auto roff_of = [&](const std::vector<uint32_t> &rp, uint32_t r) {
        uint64_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if ((r >> j) & 1) off |= 1ull << sw.tq[rp[j]];
        return off;
    };
    auto roff = [&](uint32_t r) { return roff_of(rpl, r); };
    if (from_basis) {
        // A tile without the basis index starts at zero and stays zero (the
        // sweep's non-diagonal ops act on tile qubits only; its phases map 0
        // to 0), so only the basis index's tile does arithmetic.  The zeros
        // are written in address order instead of tile order: workgroup b
        // clears the contiguous run of 2^L amplitudes from b * 2^L (one
        // contiguous 2^LT run per store), skipping the basis tile's own
        // amplitudes, which that tile's workgroup writes (a workgroup-uniform
        // branch after the clearing).
        uint64_t outer = 0;
        for (uint32_t q : g.oq) outer |= 1ull << q;
        o << "  {\n    const unsigned long long cb = b << " << g.L << ", bo = binit & " << outer
          << "ull;\n    V z; z.x = (T)0; z.y = (T)0;\n";
        for (uint32_t r = 0; r < g.NR; r++) {
            o << "    { const unsigned long long i = cb | " << ((uint64_t)r << g.LT) << "ull | tid; if ((i & " << outer
              << "ull) != bo) ";
            if (nt & 1)
                o << "st_nt(&amp[i], z); }\n";
            else
                o << "amp[i] = z; }\n";
        }
        o << "    if (base != bo) return;\n  }\n";
    }
    for (uint32_t r = 0; r < g.NR; r++) {
        if (from_basis)
            o << "  V a" << r << "; a" << r << ".x = (T)((gm | " << roff(r) << "ull) == binit ? 1 : 0); a" << r
              << ".y = (T)0;\n";
        else if (nt & 2)
            o << "  V a" << r << " = ld_nt(&amp[gm | " << roff(r) << "ull]);\n";
        else
            o << "  V a" << r << " = amp[gm | " << roff(r) << "ull];\n";
    }
    // kernels with MFMA blocks keep every tile load ahead of the arithmetic
    // (the scheduler otherwise sinks loads into the MFMA chains and waits on
    // each; the vector-only kernels issue them all first by themselves)
    o << "/*@loads-issued@*/\n";
    uint32_t cur = memr;
    if (use_lazy_diagonals()) {
        emit_stages_lazy(g, sw, cur, mems);
    } else {
    for (auto &St : sw.stages) {
        if (St.rmask_pos != cur) {
            emit_exchange(g, cur, St.rmask_pos);
            cur = St.rmask_pos;
        }
        std::vector<const TileOpH *> run;
        auto flush = [&]() {
            if (run.size() > 1 && diag_run_pays(run, g.NR)) {
                emit_diag_run(g, run, cur);
            } else {
                for (const TileOpH *op : run) emit_op(g, *op, cur);
            }
            run.clear();
        };
        for (uint32_t k = St.op_begin; k < St.op_end; k++) {
            if (sw.ops[k].kind == TK_DM) {
                run.push_back(&sw.ops[k]);
                continue;
            }
            flush();
            emit_op(g, sw.ops[k], cur);
        }
        flush();
    }
    if (cur != mems) emit_exchange(g, cur, mems);
    }
    if (g.scale != 1.0) {
        // the factors taken out of the real-form 2x2s
        const std::string S = g.C(g.push(g.scale));
        for (uint32_t r = 0; r < g.NR; r++)
            o << "  " << g.var(r) << ".x *= " << S << "; " << g.var(r) << ".y *= " << S << ";\n";
    }
    {
        // store addresses recomputed from an opaque copy of the thread index,
        // so the load addresses are not held in registers across the sweep
        std::ostringstream e;
        e << "base";
        for (uint32_t i = 0; i < g.LT; i++)
            e << " | (((unsigned long long)(tids >> " << i << ") & 1ull) << " << low_s[i] << ")";
        o << "  unsigned tids = tid;\n  asm volatile(\"\" : \"+v\"(tids));\n";
        o << "  const unsigned long long gs = " << e.str() << ";\n";
    }
    for (uint32_t r = 0; r < g.NR; r++) {
        if (nt & 1)
            o << "  st_nt(&" << (from_basis ? "amp" : "out") << "[gs | " << roff_of(rps, r) << "ull], " << g.var(r) << ");\n";
        else
            o << "  " << (from_basis ? "amp" : "out") << "[gs | " << roff_of(rps, r) << "ull] = " << g.var(r) << ";\n";
    }
    o << "}\n";
    {
        std::string body = o.str();
        const std::string mark = "/*@loads-issued@*/\n";
        const size_t at = body.find(mark);
        if (at != std::string::npos)
            body.replace(at, mark.size(), g.mfma_blocks ? "  __builtin_amdgcn_sched_barrier(0);\n" : "");
        o.str("");
        o << header;
        if (g.mfma_blocks) o << MFMA_HELPERS;
        o << body;
    }
    if (const char *dir = getenv("NCSV_JIT_DUMP")) {
        // the sweep's coefficients beside its dumped source (doubles, in order)
        const std::string path =
            std::string(dir) + "/ncsv_" + std::to_string(std::hash<std::string>()(o.str())) + ".coef";
        if (FILE *f = fopen(path.c_str(), "wb")) {
            fwrite(coef->data(), sizeof(double), coef->size(), f);
            fclose(f);
        }
        // and the order the sweeps were generated in
        if (FILE *f = fopen((std::string(dir) + "/order.txt").c_str(), "a")) {
            fprintf(f, "%zu %d\n", std::hash<std::string>()(o.str()), from_basis ? 1 : 0);
            fclose(f);
        }
    }
    return o.str();
}

// ------------------------------------------------------- expectation kernels

// Group kernel for Paulis whose X part lies on register positions of one of
// the group's stages.  For each Pauli, per register pair (r, r ^ xr):
// P = conj(psi(c^x)) psi(c); term(c) = (-1)^pt i^(ph + 2 pr(r)) P and
// term(c^x) = (-1)^pt i^(ph + 2 pr(r^xr)) conj(P), pt the Z parity of the
// thread and outer bits, pr that of the register bits.  One workgroup walks
// tiles with a grid stride; partial sums go to part[k * gridDim.x + block].
// Complex128 expectation kernels load each tile's successor while working
// on it (30 qubits, 64 Paulis: 3.149 -> 3.072 ms per pass on MI355X); in
// complex64 the second set of registers costs more than it hides (groups of
// 13 Paulis: 2.128 -> 2.686 ms), so those take one tile at a time.
// NCSV_EXPECT_PREFETCH=0/1 forces either.
bool expect_prefetch(bool dbl) {
    const char *e = getenv("NCSV_EXPECT_PREFETCH");
    return e ? atoi(e) != 0 : dbl;
}

std::string generate_expect(const ExpGroup &eg, uint32_t n, const TileGeom &geo, bool dbl) {
    Gen g;
    // each exchange's LDS addresses are computed where it runs, not held across the tile loop
    g.opaque = true;
    // NCSV_EXPECT_HALF=1: exchanges through half a tile of LDS, as NCSV_JIT_HALF for sweeps
    if (const char *e = getenv("NCSV_EXPECT_HALF")) g.half = atoi(e);
    std::vector<double> dummy;
    g.coef = &dummy;
    g.LT = geo.LT;
    g.R = geo.R;
    g.L = geo.L();
    g.NR = 1u << geo.R;
    g.tq = eg.tq;
    g.name.resize(g.NR);
    for (uint32_t r = 0; r < g.NR; r++) g.name[r] = (int)r;
    uint64_t in = 0;
    for (uint32_t q : eg.tq) in |= 1ull << q;
    for (uint32_t q = 0; q < n; q++)
        if (!((in >> q) & 1)) g.oq.push_back(q);
    std::vector<int> pos_of(64, -1);
    for (uint32_t p = 0; p < g.L; p++) pos_of[eg.tq[p]] = (int)p;
    // loads in the first stage's layout where it leaves the fixed low
    // positions to the threads (an exchange fewer per tile), else the default
    uint32_t memr = ((1u << g.R) - 1) << g.LT;
    if (!eg.stages.empty() && !(eg.stages[0].rpos & ((1u << geo.c) - 1)) &&
        (uint32_t)__builtin_popcount(eg.stages[0].rpos) == g.R)
        memr = eg.stages[0].rpos;
    const uint32_t nb = n - g.L;
    const uint32_t cnt = (uint32_t)eg.paulis.size();
    bool lds = false;
    {
        uint32_t cur = memr;
        for (auto &st : eg.stages) {
            if (st.rpos != cur) lds = true;
            cur = st.rpos;
        }
    }
    auto &o = g.o;
    const int S = dbl ? 4 : 5;
    g.S = S;
    o << "typedef " << (dbl ? "double" : "float") << " T;\ntypedef " << (dbl ? "double2" : "float2") << " V;\n";
    if (jit_nt() & 2) o << NT_HELPERS;
    o << "__device__ __forceinline__ unsigned swz(unsigned t) { return t ^ (((t >> " << S << ") ^ (t >> " << 2 * S
      << ")) & " << ((1u << S) - 1) << "u); }\n";
    o << "__device__ __forceinline__ double wsum(double v) { for (int k = 32; k > 0; k >>= 1) v += __shfl_down(v, k, 64); return v; }\n";
    // a barrier for the LDS exchanges that does not wait for the next tile's
    // loads in flight (__syncthreads waits for every outstanding memory op)
    const bool pf = expect_prefetch(dbl);
    if (pf) {
        o << "__device__ __forceinline__ void lds_bar() { asm volatile(\"s_waitcnt lgkmcnt(0)\\n\\ts_barrier\" ::: \"memory\"); }\n";
        g.bar = "lds_bar()";
    }
    o << "extern \"C\" __global__ __launch_bounds__(" << (1u << g.LT)
      << ") void ncsv_expect(const V *__restrict__ amp, double *__restrict__ part, unsigned long long ntiles) {\n";
    if (lds) o << "  __shared__ V lds[" << (1u << (g.half ? g.L - 1 : g.L)) << "];\n";
    o << "  __shared__ double red[" << ((1u << g.LT) / 64) << "][" << 2 * cnt << "];\n";
    o << "  const unsigned tid = threadIdx.x;\n";
    for (uint32_t k = 0; k < cnt; k++) o << "  double sr" << k << " = 0.0, si" << k << " = 0.0;\n";
    std::vector<uint32_t> low;
    for (uint32_t p : positions(memr, g.L, false)) low.push_back(eg.tq[p]);
    const std::vector<uint32_t> rpl = positions(memr, g.L, true);
    auto roff = [&](uint32_t r) {
        uint64_t off = 0;
        for (uint32_t j = 0; j < g.R; j++)
            if ((r >> j) & 1) off |= 1ull << eg.tq[rpl[j]];
        return off;
    };
    // tile index -> block index b (XCD-chunked order: the grid is a multiple
    // of 8 workgroups, which go to the 8 XCDs in turn, so XCD x walks the
    // contiguous run of tiles x * ntiles/8 ...) -> base -> loads into `dst`
    auto emit_tile = [&](const std::string &ti, const std::string &sfx, const std::string &dst, const char *pre) {
        if (nb >= 3 && default_border() != 0)
            o << pre << "const unsigned long long b" << sfx << " = ((" << ti << " & 7ull) << " << nb - 3 << ") | (" << ti
              << " >> 3);\n";
        else
            o << pre << "const unsigned long long b" << sfx << " = " << ti << ";\n";
        o << pre << "const unsigned long long base" << sfx << " = " << deposit("b" + sfx, g.oq, true) << ";\n";
        if (dst.empty()) return;
        std::ostringstream e;
        e << "base" << sfx;
        for (uint32_t i = 0; i < g.LT; i++) e << " | (((unsigned long long)(tid >> " << i << ") & 1ull) << " << low[i] << ")";
        o << pre << "const unsigned long long gm" << sfx << " = " << e.str() << ";\n";
        for (uint32_t r = 0; r < g.NR; r++) {
            if (jit_nt() & 2)
                o << pre << dst << r << " = ld_nt(&amp[gm" << sfx << " | " << roff(r) << "ull]);\n";
            else
                o << pre << dst << r << " = amp[gm" << sfx << " | " << roff(r) << "ull];\n";
        }
    };
    if (pf) {
        // software pipelined: the next tile's loads go out before this tile's
        // exchanges and products, into a second set of registers
        o << "  V";
        for (uint32_t r = 0; r < g.NR; r++) o << (r ? ", a" : " a") << r;
        o << ";\n  {\n";
        emit_tile("(unsigned long long)blockIdx.x", "0", "a", "    ");
        o << "  }\n";
        o << "  for (unsigned long long bi = blockIdx.x; bi < ntiles; bi += gridDim.x) {\n";
        emit_tile("bi", "", "", "  ");
        o << "  V";
        for (uint32_t r = 0; r < g.NR; r++) o << (r ? ", p" : " p") << r;
        // unconditional (the last tile of a workgroup loads itself again, 1/128
        // of the reads at 30 qubits): a conditional load leaves the compiler
        // waiting for every load, the prefetch included, where the paths join
        o << ";\n  {\n    const unsigned long long bj = bi + gridDim.x < ntiles ? bi + gridDim.x : bi;\n";
        emit_tile("bj", "1", "p", "    ");
        o << "  }\n";
    } else {
        o << "  for (unsigned long long bi = blockIdx.x; bi < ntiles; bi += gridDim.x) {\n";
        emit_tile("bi", "", "", "  ");
        o << "  V";
        for (uint32_t r = 0; r < g.NR; r++) o << (r ? ", a" : " a") << r;
        o << ";\n  {\n";
        emit_tile("bi", "0", "a", "    ");
        o << "  }\n";
    }
    // which of each Pauli's sums (bit 0 real, bit 1 imaginary) receives any
    // term: the other stays out of the loop (adding its signed zeros would
    // keep a register and an add per Pauli live)
    std::vector<uint8_t> used(cnt, 0);
    uint32_t cur = memr;
    for (auto &st : eg.stages) {
        if (st.rpos != cur) {
            emit_exchange(g, cur, st.rpos);
            cur = st.rpos;
        }
        // global bits of the thread in this layout: base | thread positions
        const std::vector<uint32_t> tp = positions(cur, g.L, false), rp = positions(cur, g.L, true);
        {
            std::ostringstream e;
            e << "base";
            for (size_t i = 0; i < tp.size(); i++)
                e << " | (((unsigned long long)(tid >> " << i << ") & 1ull) << " << eg.tq[tp[i]] << ")";
            o << "  {\n  const unsigned long long g = " << e.str() << ";\n";
        }
        for (uint32_t k : st.paulis) {
            const ExpPauli &P = eg.paulis[k];
            // register masks of x and z; z on the rest
            uint32_t xr = 0, zr = 0;
            uint64_t zrest = P.z;
            for (uint32_t j = 0; j < g.R; j++) {
                const uint32_t q = eg.tq[rp[j]];
                if ((P.x >> q) & 1) xr |= 1u << j;
                if ((P.z >> q) & 1) {
                    zr |= 1u << j;
                    zrest &= ~(1ull << q);
                }
            }
            o << "  { const bool odd = __popcll(g & " << zrest << "ull) & 1; double tr = 0.0, ti = 0.0;\n";
            auto add = [&](int kk, const std::string &re, const std::string &im) {
                // (tr, ti) += i^kk (re + i im); an empty im is zero
                const bool hi = !im.empty();
                used[k] |= hi ? 3 : (kk & 1) ? 2 : 1;
                switch (kk & 3) {
                case 0: o << "    tr += " << re << ";" << (hi ? " ti += " + im + ";" : "") << "\n"; break;
                case 1: o << (hi ? "    tr -= " + im + ";" : "   ") << " ti += " << re << ";\n"; break;
                case 2: o << "    tr -= " << re << ";" << (hi ? " ti -= " + im + ";" : "") << "\n"; break;
                default: o << (hi ? "    tr += " + im + ";" : "   ") << " ti -= " << re << ";\n"; break;
                }
            };
            for (uint32_t r = 0; r < g.NR; r++) {
                const uint32_t r2 = r ^ xr;
                if (xr && r2 < r) continue;
                const int kr = (int)(P.ph + 2 * __builtin_popcount(zr & r)) & 3;
                const std::string w = g.var(r), u = g.var(r2);
                if (!xr) {
                    o << "    { const T x = " << w << ".x, y = " << w << ".y; const double m = (double)x*x + (double)y*y;\n";
                    add(kr, "m", "");
                    o << "    }\n";
                } else {
                    // i^kr z + i^kr2 conj(z) with z = conj(psi(c^x)) psi(c) and kr2 - kr in {0, 2}:
                    // 2 i^kr Re z, or 2 i^(kr+1) Im z; the factor 2 is applied per thread below
                    const int kr2 = (int)(P.ph + 2 * __builtin_popcount(zr & r2)) & 3;
                    if (kr2 == kr) {
                        o << "    { const double v = (double)" << u << ".x*" << w << ".x + (double)" << u << ".y*" << w
                          << ".y;\n";
                        add(kr, "v", "");
                    } else {
                        o << "    { const double v = (double)" << u << ".x*" << w << ".y - (double)" << u << ".y*" << w
                          << ".x;\n";
                        add(kr + 1, "v", "");
                    }
                    o << "    }\n";
                }
            }
            const std::string ar = (used[k] & 1) ? " sr" + std::to_string(k) + (xr ? " += f2*tr;" : " += tr;") : "",
                              ai = (used[k] & 2) ? " si" + std::to_string(k) + (xr ? " += f2*ti;" : " += ti;") : "";
            if (xr)
                o << "    const double f2 = odd ? -2.0 : 2.0;\n   " << ar << ai << " }\n";
            else
                o << "    if (odd) { tr = -tr; ti = -ti; }\n   " << ar << ai << " }\n";
            // one Pauli at a time: interleaving them raised the registers past two wavefronts per SIMD
            o << "    __builtin_amdgcn_sched_barrier(0);\n";
        }
        o << "  }\n";
    }
    if (lds) o << "  " << g.bar << ";\n";
    if (pf) {
        // the prefetched tile becomes the current one (the loop body starts
        // from logical register r in variable a<r>)
        for (uint32_t r = 0; r < g.NR; r++) o << "  a" << r << " = p" << r << ";\n";
    }
    o << "  }\n";
    o << "  const int lane = tid & 63, wv = tid >> 6;\n";
    for (uint32_t k = 0; k < cnt; k++) {
        const std::string a = (used[k] & 1) ? "wsum(sr" + std::to_string(k) + ")" : "0.0",
                          b2 = (used[k] & 2) ? "wsum(si" + std::to_string(k) + ")" : "0.0";
        o << "  { const double a = " << a << ", b2 = " << b2 << "; if (lane == 0) { red[wv][" << 2 * k << "] = a; red[wv]["
          << 2 * k + 1 << "] = b2; } }\n";
    }
    o << "  __syncthreads();\n  if (tid < " << 2 * cnt << ") { double s = 0.0; for (int i = 0; i < " << ((1u << g.LT) / 64)
      << "; i++) s += red[i][tid]; part[(unsigned long long)tid * gridDim.x + blockIdx.x] = s; }\n}\n";
    return o.str();
}

// ------------------------------------------------------------------ compile

namespace {

struct Entry {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    std::string err;
    // the module goes when the last plan or cache slot holding it does
    ~Entry() {
        if (mod) (void)hipModuleUnload(mod);
    }
};

// Compiled kernels by source text, least recently used first out: at most
// NCSV_JIT_CACHE (default 1024) finished entries stay cached.  Plans hold
// their own references, so an evicted kernel in use stays loaded until its
// plan goes; expectation kernels (whose sources hold the Paulis' masks) of
// observables no longer asked for are unloaded instead of accumulating.
std::mutex g_mu;
struct Slot {
    std::shared_future<std::shared_ptr<Entry>> f;
    std::list<std::string>::iterator lru;
};
// held for the life of the process (no module unloads during static
// destruction, after the runtime may have gone)
std::map<std::string, Slot> &g_cache = *new std::map<std::string, Slot>();
std::list<std::string> &g_lru = *new std::list<std::string>();  // front: most recently used

size_t cache_cap() {
    const char *e = getenv("NCSV_JIT_CACHE");
    return e ? (size_t)std::max(1, atoi(e)) : 1024;
}

std::string arch_name() {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
        std::string a = p.gcnArchName;
        const size_t c = a.find(':');
        return c == std::string::npos ? a : a.substr(0, c);
    }
    return "gfx950";
}

}  // namespace

std::vector<char> compile_to_code(const std::string &src, const std::string &arch, std::string *log) {
    if (const char *dir = getenv("NCSV_JIT_DUMP")) {
        // keep the generated source for inspection
        const std::string path = std::string(dir) + "/ncsv_" + std::to_string(std::hash<std::string>()(src)) + ".hip";
        if (FILE *f = fopen(path.c_str(), "w")) {
            fwrite(src.data(), 1, src.size(), f);
            fclose(f);
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "ncsv_sweep.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *log = "hiprtcCreateProgram failed";
        return {};
    }
    const std::string a = "--offload-arch=" + arch;
    const char *opts[] = {a.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (ls > 1) {
        std::string l(ls, '\0');
        hiprtcGetProgramLog(prog, &l[0]);
        *log = l;
    }
    std::vector<char> code;
    if (r == HIPRTC_SUCCESS) {
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
    }
    hiprtcDestroyProgram(&prog);
    return code;
}

std::shared_future<std::shared_ptr<Entry>> get_or_start(const std::string &src, const char *fname = "ncsv_sweep") {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(src);
    if (it != g_cache.end()) {
        g_lru.splice(g_lru.begin(), g_lru, it->second.lru);
        return it->second.f;
    }
    // evict the least recently used finished compiles beyond the cap
    for (auto r = g_lru.end(); g_cache.size() >= cache_cap() && r != g_lru.begin();) {
        --r;
        auto c = g_cache.find(*r);
        if (c->second.f.wait_for(std::chrono::seconds(0)) != std::future_status::ready) continue;
        g_cache.erase(c);
        r = g_lru.erase(r);
    }
    const std::string arch = arch_name();
    const std::string fn = fname;
    std::shared_future<std::shared_ptr<Entry>> f = std::async(std::launch::async, [src, arch, fn]() {
                                                         auto e = std::make_shared<Entry>();
                                                         std::string log;
                                                         std::vector<char> code = compile_to_code(src, arch, &log);
                                                         if (code.empty()) {
                                                             e->err = "sweep kernel compile failed: " + log;
                                                             return e;
                                                         }
                                                         if (hipModuleLoadData(&e->mod, code.data()) != hipSuccess ||
                                                             hipModuleGetFunction(&e->fn, e->mod, fn.c_str()) !=
                                                                 hipSuccess) {
                                                             e->err = "sweep kernel load failed";
                                                             e->fn = nullptr;
                                                         }
                                                         return e;
                                                     }).share();
    g_lru.push_front(src);
    g_cache[src] = Slot{f, g_lru.begin()};
    return f;
}

void *prepare(const std::string &src) {
    auto f = get_or_start(src);
    return (void *)new std::shared_future<std::shared_ptr<Entry>>(f);
}

void *prepare_named(const std::string &src, const char *fname) {
    auto f = get_or_start(src, fname);
    return (void *)new std::shared_future<std::shared_ptr<Entry>>(f);
}

void release(void *handle) { delete (std::shared_future<std::shared_ptr<Entry>> *)handle; }

// NCSV_JIT_OOP=1: sweeps read one buffer and write the other (the state's
// second buffer, allocated on first use; in place when it cannot be had).
bool out_of_place() {
    const char *e = getenv("NCSV_JIT_OOP");
    return e && atoi(e) != 0;
}

void launch(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, const void *dcoef) {
    auto &f = *(std::shared_future<std::shared_ptr<Entry>> *)handle;
    const std::shared_ptr<Entry> e = f.get();
    if (!e->fn) fail(NCSV_DEVICE, e->err);
    if (out_of_place() && !s->spare && !s->spare_failed) {
        bool ok = true;
        if (const char *cap = getenv("NCSV_DEVICE_MEMORY_LIMIT_GB"))
            ok = 2.0 * (double)s->amp_bytes() <= atof(cap) * (double)(1ull << 30);
        if (ok && hipMalloc(&s->spare, s->amp_bytes()) != hipSuccess) {
            (void)hipGetLastError();  // the refused allocation's error is not the caller's
            s->spare = nullptr;
            ok = false;
        }
        s->spare_failed = !ok;
    }
    void *amp = s->amp;
    void *out = s->spare ? s->spare : s->amp;
    const void *c = dcoef;
    void *args[] = {&amp, &c, &out};
    const uint64_t tiles = 1ull << (s->n - L);
    KTimer kt(s, "sweep_jit", 2.0 * (double)s->amp_bytes());
    check_hip(hipModuleLaunchKernel(e->fn, (uint32_t)tiles, 1, 1, 1u << LT, 1, 1, 0, s->stream, args, nullptr),
              "hipModuleLaunchKernel(sweep)");
    if (out != amp) std::swap(s->amp, s->spare);
    s->sweeps++;
}

void launch_from_basis(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, const void *dcoef, uint64_t binit) {
    auto &f = *(std::shared_future<std::shared_ptr<Entry>> *)handle;
    const std::shared_ptr<Entry> e = f.get();
    if (!e->fn) fail(NCSV_DEVICE, e->err);
    void *amp = s->amp;
    const void *c = dcoef;
    unsigned long long b = binit;
    void *args[] = {&amp, &c, &b};
    const uint64_t tiles = 1ull << (s->n - L);
    KTimer kt(s, "sweep_jit_from_basis", (double)s->amp_bytes());
    check_hip(hipModuleLaunchKernel(e->fn, (uint32_t)tiles, 1, 1, 1u << LT, 1, 1, 0, s->stream, args, nullptr),
              "hipModuleLaunchKernel(sweep from a basis state)");
    s->sweeps++;
}

void launch_expect(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, uint32_t grid, double *dpart) {
    auto &f = *(std::shared_future<std::shared_ptr<Entry>> *)handle;
    const std::shared_ptr<Entry> e = f.get();
    if (!e->fn) fail(NCSV_DEVICE, e->err);
    const void *amp = s->amp;
    unsigned long long ntiles = 1ull << (s->n - L);
    void *args[] = {&amp, &dpart, &ntiles};
    KTimer kt(s, "expect_jit", (double)s->amp_bytes());
    check_hip(hipModuleLaunchKernel(e->fn, grid, 1, 1, 1u << LT, 1, 1, 0, s->stream, args, nullptr),
              "hipModuleLaunchKernel(expect)");
}

bool available() {
    static int ok = -1;
    if (ok < 0) {
        const char *e = getenv("NCSV_JIT");
        ok = (e && atoi(e) == 0) ? 0 : 1;
    }
    return ok == 1;
}

// Expectation values through specialized group kernels.  Returns false
// (nothing done) when some Pauli's X part does not fit one tile's register
// stage.
// Most Paulis per group kernel: every Pauli holds two FP64 accumulators per
// thread.  On 64 random Paulis at 30 qubits: complex128 tiles (16 registers
// per thread) group best by 12 (six passes of 3.5 ms; 13 still needs six
// passes, 16 takes 3.7 ms per pass); complex64 tiles (32 registers per thread,
// so more X parts share a stage) by 13 (five passes of 2.4 ms against six of
// 2.2 ms at 12 and five of 3.0 ms at 16, where 16 Paulis take the kernel to
// 256 registers and one wavefront per SIMD).
// Complex128 groups of 13 are taken only where they save a pass (with three
// fixed low qubits, so nine free tile positions, and the order search below,
// the 64 Paulis fit five groups of 13: 5 x 3.38 ms against 6 x 3.06 ms; where
// they save none, 13 cost 2.6 % per pass).
uint32_t group_max(const TileGeom &geo) {
    const char *e = getenv("NCSV_EXPECT_GMAX");
    if (e) return (uint32_t)std::min(16, std::max(1, atoi(e)));
    return geo.R >= 5 ? 13u : 12u;
}

// Orders of the first fit tried: the X-weight order, plus (from 26 qubits,
// where a pass over the state costs a fraction of a millisecond or more) 511
// orders with the weights jittered; NCSV_EXPECT_ORDERS sets the count.
uint32_t expect_orders(uint32_t n) {
    if (const char *e = getenv("NCSV_EXPECT_ORDERS")) return (uint32_t)std::max(1, atoi(e));
    return n >= 26 ? 512u : 1u;
}

bool plan_expect_g(uint32_t n, const TileGeom &geo, uint32_t count, const uint64_t *x, const uint64_t *z,
                   const uint32_t *ph, uint32_t gmax, std::vector<ExpGroup> *out_groups) {
    const uint32_t L = geo.L(), R = geo.R, c = std::min(geo.c, L), LT = geo.LT;
    if (n < L) return false;
    ncsv_state fake;
    fake.n = n;
    ncsv_state *s = &fake;
    const uint64_t low = (1ull << c) - 1;
    constexpr uint32_t GMAX = 16;
    for (uint32_t i = 0; i < count; i++)
        if ((uint32_t)__builtin_popcountll(x[i]) > R || (uint32_t)__builtin_popcountll(x[i] & ~low) > L - c) return false;
    // tiles: first fit on the union of the X qubits above the low ones
    struct G0 {
        uint64_t T;
        std::vector<uint32_t> idx;
    };
    // first fit, widest X parts first (fewer tiles, so fewer passes over the state)
    auto first_fit = [&](const std::vector<uint32_t> &order) {
        std::vector<G0> g0;
        for (uint32_t i : order) {
            const uint64_t need = x[i] & ~low;
            bool placed = false;
            for (auto &g : g0)
                if (g.idx.size() < gmax && (uint32_t)__builtin_popcountll(g.T | need) <= L) {
                    g.T |= need;
                    g.idx.push_back(i);
                    placed = true;
                    break;
                }
            if (!placed) g0.push_back({low | need, {i}});
        }
        return g0;
    };
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return __builtin_popcountll(x[a] & ~low) > __builtin_popcountll(x[b] & ~low);
    });
    std::vector<G0> g0 = first_fit(order);
    // more orders (X weight plus a fixed pseudo-random jitter of up to 2.5),
    // keeping the first with the fewest groups
    const uint32_t orders = expect_orders(n);
    const size_t fewest = (count + gmax - 1) / gmax;
    for (uint32_t t = 1; t < orders && g0.size() > fewest; t++) {
        std::vector<double> key(count);
        for (uint32_t i = 0; i < count; i++) {
            uint64_t h = ((uint64_t)t << 32 | i) * 0x9E3779B97F4A7C15ull;
            h ^= h >> 29;
            h *= 0xBF58476D1CE4E5B9ull;
            h ^= h >> 32;
            key[i] = -(double)__builtin_popcountll(x[i] & ~low) + 2.5 * (double)(h >> 11) / 9007199254740992.0;
        }
        std::vector<uint32_t> o2(count);
        for (uint32_t i = 0; i < count; i++) o2[i] = i;
        std::stable_sort(o2.begin(), o2.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
        std::vector<G0> g2 = first_fit(o2);
        if (g2.size() < g0.size()) g0 = std::move(g2);
    }
    std::vector<ExpGroup> groups;
    for (auto &g : g0) {
        ExpGroup eg;
        eg.index = g.idx;
        // register stages over qubits: union of X qubits <= R
        std::vector<std::pair<uint64_t, std::vector<uint32_t>>> st;
        for (uint32_t k = 0; k < g.idx.size(); k++) {
            const uint64_t xq = x[g.idx[k]];
            bool placed = false;
            for (auto &sg : st)
                if ((uint32_t)__builtin_popcountll(sg.first | xq) <= R) {
                    sg.first |= xq;
                    sg.second.push_back(k);
                    placed = true;
                    break;
                }
            if (!placed) st.push_back({xq, {k}});
            eg.paulis.push_back({x[g.idx[k]], z[g.idx[k]], ph[g.idx[k]] & 3});
        }
        // tile qubits: the low ones, every X qubit, then the lowest others
        uint64_t T = g.T;
        for (auto &sg : st) T |= sg.first;
        for (uint32_t q = 0; q < s->n && (uint32_t)__builtin_popcountll(T) < L; q++) T |= 1ull << q;
        // positions: low qubits fixed; stage 0's other X qubits on the top
        eg.tq.assign(L, ~0u);
        std::vector<int> pos_of(64, -1);
        for (uint32_t p = 0; p < c; p++) {
            eg.tq[p] = p;
            pos_of[p] = (int)p;
        }
        uint32_t top = L - R;
        for (uint64_t a = st[0].first & ~low; a; a &= a - 1) {
            const uint32_t q = (uint32_t)__builtin_ctzll(a);
            eg.tq[top] = q;
            pos_of[q] = (int)top++;
        }
        uint32_t next = c;
        for (uint64_t a = T; a; a &= a - 1) {
            const uint32_t q = (uint32_t)__builtin_ctzll(a);
            if (pos_of[q] >= 0) continue;
            while (eg.tq[next] != ~0u) next++;
            eg.tq[next] = q;
            pos_of[q] = (int)next;
        }
        uint32_t prev = ((1u << R) - 1) << LT;
        for (auto &sg : st) {
            uint32_t rpos = 0;
            for (uint64_t a = sg.first; a; a &= a - 1) rpos |= 1u << pos_of[__builtin_ctzll(a)];
            for (uint32_t p = L; p-- > 0 && (uint32_t)__builtin_popcount(rpos) < R;)
                if ((prev >> p) & 1) rpos |= 1u << p;
            for (uint32_t p = L; p-- > 0 && (uint32_t)__builtin_popcount(rpos) < R;) rpos |= 1u << p;
            eg.stages.push_back({rpos, sg.second});
            prev = rpos;
        }
        groups.push_back(std::move(eg));
    }
    *out_groups = std::move(groups);
    return true;
}

// The groups of a Pauli list, planned once per list, tile shape and state
// size (the order search makes planning cost up to a millisecond).
bool plan_expect(uint32_t n, const TileGeom &geo, uint32_t count, const uint64_t *x, const uint64_t *z,
                 const uint32_t *ph, std::vector<ExpGroup> *out_groups) {
    static std::mutex mu;
    static std::map<std::string, std::vector<ExpGroup>> cache;
    std::string key;
    {
        const uint32_t head[7] = {n, geo.LT, geo.R, geo.c, count, group_max(geo), expect_orders(n)};
        key.append((const char *)head, sizeof head);
        key.append((const char *)x, count * sizeof(uint64_t));
        key.append((const char *)z, count * sizeof(uint64_t));
        key.append((const char *)ph, count * sizeof(uint32_t));
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            *out_groups = it->second;
            return true;
        }
    }
    const uint32_t gmax = group_max(geo);
    std::vector<ExpGroup> groups;
    if (!plan_expect_g(n, geo, count, x, z, ph, gmax, &groups)) return false;
    if (gmax == 12 && geo.R < 5 && !getenv("NCSV_EXPECT_GMAX") && groups.size() > 1) {
        std::vector<ExpGroup> g13;
        if (plan_expect_g(n, geo, count, x, z, ph, 13, &g13) && g13.size() < groups.size()) groups = std::move(g13);
    }
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() >= 256) cache.clear();
    cache.emplace(key, groups);
    *out_groups = std::move(groups);
    return true;
}

// The expectation kernels' tile shape: the sweeps' default, but complex128
// tiles fix three low qubits instead of four (128-byte runs, the same pass
// time: 3.05 against 3.06 ms at 30 qubits), leaving nine free positions for
// the groups' X qubits; NCSV_EXPECT_C sets the fixed low qubits.
TileGeom expect_geom(ncsv_precision prec) {
    TileGeom geo = dev::tile_geom(prec);
    if (prec == NCSV_C128 && !getenv("NCSV_TILE_C")) geo.c = 3;
    if (const char *e = getenv("NCSV_EXPECT_C")) geo.c = (uint32_t)atoi(e);
    return geo;
}

bool expectations(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                  double *out) {
    const TileGeom geo = expect_geom(s->prec);
    const uint32_t L = geo.L(), LT = geo.LT;
    constexpr uint32_t GMAX = 16;
    std::vector<ExpGroup> groups;
    if (!plan_expect(s->n, geo, count, x, z, ph, &groups)) return false;
    const bool dbl = s->prec == NCSV_C128;
    std::vector<void *> kern;
    for (auto &eg : groups) kern.push_back(prepare_named(generate_expect(eg, s->n, geo, dbl), "ncsv_expect"));
    const uint64_t tiles = 1ull << (s->n - L);
    // persistent grid (a multiple of 8 keeps each XCD on its own run of
    // tiles); NCSV_EXPECT_GRID sets it
    uint64_t want = 2048;
    if (const char *e = getenv("NCSV_EXPECT_GRID")) want = std::max<uint64_t>(8, strtoull(e, nullptr, 10) & ~7ull);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(tiles, want);
    dev::ensure_part(s, (size_t)2 * GMAX * grid + 2 * GMAX + 64);
    std::vector<double> res(2 * GMAX);
    try {
        for (size_t gi = 0; gi < groups.size(); gi++) {
            const uint32_t cnt = (uint32_t)groups[gi].paulis.size();
            launch_expect(s, kern[gi], L, LT, grid, s->d_part);
            dev::rowsum(s, s->d_part, 2 * cnt, grid, res.data());
            for (uint32_t k = 0; k < cnt; k++) {
                out[2 * groups[gi].index[k]] = res[2 * k];
                out[2 * groups[gi].index[k] + 1] = res[2 * k + 1];
            }
        }
    } catch (...) {
        for (void *k : kern) release(k);
        throw;
    }
    for (void *k : kern) release(k);
    return true;
}

}  // namespace jit
}  // namespace ncsv

extern "C" {

// Diagnostics: the specialized expectation kernels for `count` Paulis on n
// qubits, compiled for `arch` without a device; *ngroups = kernels.
ncsv_status ncsv_jit_check_expect(uint32_t n, ncsv_precision prec, uint32_t count, const uint64_t *x,
                                  const uint64_t *z, const uint32_t *ph, const char *arch, uint32_t *ngroups) {
    try {
        using namespace ncsv;
        const TileGeom geo = jit::expect_geom(prec);
        std::vector<jit::ExpGroup> groups;
        if (!jit::plan_expect(n, geo, count, x, z, ph, &groups)) return NCSV_UNSUPPORTED;
        for (auto &eg : groups) {
            std::string log;
            if (jit::compile_to_code(jit::generate_expect(eg, n, geo, prec == NCSV_C128), arch ? arch : "gfx950", &log)
                    .empty())
                fail(NCSV_DEVICE, "compile: " + log);
        }
        if (ngroups) *ngroups = (uint32_t)groups.size();
        return NCSV_OK;
    } catch (ncsv::Error &e) {
        return e.code;
    }
}


// Diagnostics: the specialized source of every tile sweep of a unitary
// circuit's plan, compiled for `arch` (e.g. "gfx950") without a device.
// Returns NCSV_OK when all compile; *nsweeps receives the count.
ncsv_status ncsv_jit_check(const ncsv_circuit *c, ncsv_precision prec, const char *arch, uint32_t *nsweeps) {
    try {
        using namespace ncsv;
        TileGeom geo = dev::tile_geom(prec);
        std::vector<LOp> ops;
        for (const COp &op : c->ops) {
            if (op.kind == CKind::Gate)
                ops.push_back(lower_gate(op.g, op.qubits.data(), (uint32_t)op.qubits.size(), c->n));
            else if (op.kind != CKind::Tick)
                fail(NCSV_INVALID, "jit check takes unitary circuits");
        }
        const std::vector<SegItem> items = plan_segment_tuned(ops, c->n, &geo, true);
        uint32_t k = 0;
        for (const SegItem &it : items) {
            if (!it.tile) continue;
            std::vector<double> coef;
            const jit::MemShape ms = forced_memory(it.sw, geo);
            const std::string src = jit::generate(it.sw, c->n, geo, prec == NCSV_C128, &coef, false, &ms);
            std::string log;
            if (jit::compile_to_code(src, arch ? arch : "gfx950", &log).empty()) fail(NCSV_DEVICE, "compile: " + log);
            if (k == 0) {
                // the first sweep's form for runs from a basis state
                std::vector<double> cb;
                const std::string sb = jit::generate(it.sw, c->n, geo, prec == NCSV_C128, &cb, true, &ms);
                if (jit::compile_to_code(sb, arch ? arch : "gfx950", &log).empty()) fail(NCSV_DEVICE, "compile: " + log);
            }
            k++;
        }
        if (nsweeps) *nsweeps = k;
        return NCSV_OK;
    } catch (ncsv::Error &e) {
        return e.code;
    }
}

}  // extern "C"
