// Gate-segment execution and the fusion planner.
//
// Fusion mode 0 applies one device sweep per gate, the reference's schedule
// (vector.rs:364-395 per op).  Mode 1 plans tile sweeps: gates are taken in
// circuit order as long as they do not depend on a gate left for a later
// sweep (two gates commute when every qubit they share is one both act on
// diagonally), and each sweep's tile qubits are chosen greedily to admit the
// most gates.  Within a sweep the same rule, with R register qubits instead
// of L tile qubits, splits the gates into stages.
#include "fuse.hpp"

#include <cstdlib>

#include "jit.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace ncsv {

namespace {

// A lowered gate as a tile op on global qubits, plus its dependency masks.
struct Cls {
    bool tileable = false;
    uint64_t act = 0;  // qubits it acts on non-diagonally (must be register qubits)
    uint64_t sup = 0;  // all its qubits
    TileKind kind = TK_U1;
    uint32_t q0 = 0, q1 = 0;
    uint64_t cmask = 0;  // DM: qubits that must all be 1; PERM1: controls
    cd m[16];
    uint8_t k0 = 0, k1 = 0;
    uint32_t slot = 0;  // NSLOT
};

int quarter(cd c) {
    if (c == cd(1, 0)) return 0;
    if (c == cd(0, 1)) return 1;
    if (c == cd(-1, 0)) return 2;
    if (c == cd(0, -1)) return 3;
    return -1;
}

Cls classify(const LOp &op) {
    Cls c;
    const uint32_t k = op.k();
    for (uint32_t q : op.qubits) c.sup |= 1ull << q;
    switch (op.kind) {
    case OpKind::Dense:
        if (k == 1) {
            c.tileable = true;
            c.kind = TK_U1;
            c.q0 = op.qubits[0];
            c.act = c.sup;
            for (int i = 0; i < 4; i++) c.m[i] = op.m[i];
        } else if (k == 2) {
            c.tileable = true;
            c.kind = TK_U2;
            c.q0 = op.qubits[0];
            c.q1 = op.qubits[1];
            c.act = c.sup;
            for (int i = 0; i < 16; i++) c.m[i] = op.m[i];
        }
        return c;
    case OpKind::Diag: {
        const uint32_t d = 1u << k;
        bool mask_form = true;
        for (uint32_t x = 0; x + 1 < d; x++)
            if (op.m[x] != cd(1.0, 0.0)) mask_form = false;
        if (mask_form) {
            c.tileable = true;
            c.kind = TK_DM;
            c.cmask = c.sup;
            c.m[0] = op.m[d - 1];
        } else if (k == 1) {
            c.tileable = true;
            c.kind = TK_D1R;  // resolved to D1R / D1T per stage
            c.q0 = op.qubits[0];
            c.m[0] = op.m[0];
            c.m[1] = op.m[1];
        }
        return c;
    }
    case OpKind::Slot:
        c.tileable = true;
        c.kind = TK_NSLOT;
        c.q0 = op.qubits[0];
        c.act = c.sup;
        c.slot = op.ph;
        return c;
    case OpKind::Perm: {
        const uint32_t d = 1u << k;
        bool all_one = true;
        for (uint32_t x = 0; x < d; x++)
            if (op.m[x] != cd(1.0, 0.0)) all_one = false;
        if (k == 2 && all_one && op.img[0] == 0 && op.img[1] == 2 && op.img[2] == 1 && op.img[3] == 3) {
            c.tileable = true;
            c.kind = TK_SWAPR;
            c.q0 = op.qubits[0];
            c.q1 = op.qubits[1];
            c.act = c.sup;
            return c;
        }
        for (uint32_t t = 0; t < k; t++) {
            const uint32_t tb = 1u << t;
            bool ok = true;
            uint32_t cm = ~0u & (d - 1) & ~tb, moved = 0;
            for (uint32_t x = 0; x < d && ok; x++) {
                if (op.img[x] == x) {
                    if (op.m[x] != cd(1.0, 0.0)) ok = false;
                } else if (op.img[x] == (x ^ tb)) {
                    cm &= x;
                    moved++;
                } else {
                    ok = false;
                }
            }
            if (!ok || moved == 0) continue;
            cm &= ~tb;
            // every state with the controls set moves, with the same phases
            if (moved != (d >> __builtin_popcount(cm))) continue;
            const int k0 = quarter(op.m[cm]), k1 = quarter(op.m[cm | tb]);
            if (k0 < 0 || k1 < 0) continue;
            bool same = true;
            for (uint32_t x = 0; x < d; x++)
                if (op.img[x] != x && op.m[x] != op.m[(x & tb) ? (cm | tb) : cm]) same = false;
            if (!same) continue;
            c.tileable = true;
            c.kind = TK_PERM1;
            c.q0 = op.qubits[t];
            c.act = 1ull << c.q0;
            for (uint32_t j = 0; j < k; j++)
                if ((cm >> j) & 1) c.cmask |= 1ull << op.qubits[j];
            c.k0 = (uint8_t)k0;
            c.k1 = (uint8_t)k1;
            return c;
        }
        return c;
    }
    default: return c;
    }
}

// The ops of `rem` (in order) that can run when the qubits `avail` may be
// acted on non-diagonally: those that do not depend on a skipped op.
void pick(const std::vector<Cls> &cl, const std::vector<uint32_t> &rem, uint64_t avail, std::vector<uint32_t> *out,
          size_t *count) {
    uint64_t ba = 0, bn = 0;
    size_t cnt = 0;
    for (uint32_t i : rem) {
        const Cls &c = cl[i];
        const bool dep = (c.act & ba) || ((c.sup & ~c.act) & bn);
        if (!dep && (c.act & ~avail) == 0) {
            cnt++;
            if (out) out->push_back(i);
        } else {
            ba |= c.sup;
            bn |= c.act;
        }
    }
    if (count) *count = cnt;
}

// The qubits in the order the remaining ops first need them (non-diagonally).
std::vector<uint32_t> need_order(const std::vector<Cls> &cl, const std::vector<uint32_t> &rem, uint64_t have) {
    std::vector<uint32_t> v;
    uint64_t seen = have;
    for (uint32_t i : rem) {
        uint64_t a = cl[i].act & ~seen;
        while (a) {
            const int q = __builtin_ctzll(a);
            a &= a - 1;
            v.push_back((uint32_t)q);
            seen |= 1ull << q;
        }
    }
    return v;
}

// Greedy choice of `cap` qubits (starting from `fixed`, drawn from `pool`)
// that admit the most ops of `rem`.
// variant > 0: the candidate qubits in a permuted order (a fixed function of
// variant and the set so far), so ties and near-equal choices go elsewhere
uint64_t choose_set(const std::vector<Cls> &cl, const std::vector<uint32_t> &rem, uint64_t fixed, uint64_t pool,
                    uint32_t cap, uint32_t variant = 0) {
    uint64_t T = fixed;
    size_t cur = 0;
    pick(cl, rem, T, nullptr, &cur);
    while ((uint32_t)__builtin_popcountll(T) < cap) {
        uint64_t cand = pool & ~T;
        if (!cand) break;
        const std::vector<uint32_t> order = need_order(cl, rem, T);
        int best = -1;
        size_t bs = cur;
        // candidates in need order first, so ties go to the earliest needed
        std::vector<uint32_t> cands;
        for (uint32_t q : order)
            if ((cand >> q) & 1) cands.push_back(q);
        for (uint64_t a = cand; a; a &= a - 1) {
            const uint32_t q = (uint32_t)__builtin_ctzll(a);
            if (std::find(cands.begin(), cands.end(), q) == cands.end()) cands.push_back(q);
        }
        if (variant) {
            // Fisher-Yates with a SplitMix64 stream seeded by (variant, T)
            uint64_t x = (uint64_t)variant * 0x9E3779B97F4A7C15ull + T;
            for (size_t i = cands.size(); i > 1; i--) {
                x += 0x9E3779B97F4A7C15ull;
                uint64_t z = x;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                std::swap(cands[i - 1], cands[z % i]);
            }
        }
        for (uint32_t q : cands) {
            size_t s = 0;
            pick(cl, rem, T | (1ull << q), nullptr, &s);
            if (s > bs) {
                bs = s;
                best = (int)q;
            }
        }
        if (best < 0) {
            // no single qubit helps: take the first needed (in need order, not
            // the variant's permuted order)
            best = (int)cands[0];
            for (uint32_t q : order)
                if ((cand >> q) & 1) {
                    best = (int)q;
                    break;
                }
        }
        T |= 1ull << best;
        pick(cl, rem, T, nullptr, &cur);
    }
    return T;
}

// Register bit of a qubit in a stage, or -1.
int reg_bit(uint32_t q, const uint32_t *pos_of, uint32_t rpos) {
    const uint32_t p = pos_of[q];
    if (p == ~0u || !((rpos >> p) & 1)) return -1;
    return __builtin_popcount(rpos & ((1u << p) - 1));
}

// Peephole pass over a sweep's ops: one-qubit diagonals diag(d0, d1) become
// a global phase d0 (folded into one dense op, which touches every
// amplitude) times a phase d1/d0 on the amplitudes with the bit set; 2x2
// matrices with real entries, or with a real diagonal and an imaginary
// off-diagonal, get their cheaper forms.
void finish_sweep(SweepH *sw) {
    // 2x2s left diagonal by merging go back to the diagonal form
    for (TileOpH &o : sw->ops) {
        if (o.kind == TK_U1 && o.m[1] == cd(0, 0) && o.m[2] == cd(0, 0)) {
            o.kind = TK_D1R;
            o.m[1] = o.m[3];
            o.rmask = 0;
            o.gmask = 0;
        }
    }
    cd gphase(1.0, 0.0);
    for (TileOpH &o : sw->ops) {
        if (o.kind == TK_D1R || o.kind == TK_D1T) {
            const cd d0 = o.m[0], d1 = o.m[1];
            if (d0 != cd(1.0, 0.0)) gphase *= d0;
            const cd ph = d0 == cd(1.0, 0.0) ? d1 : d1 / d0;
            if (o.kind == TK_D1R) {
                o.rmask = 1u << o.j0;
                o.gmask = 0;
            } else {
                o.rmask = 0;
            }
            o.kind = TK_DM;
            o.m[0] = ph;
        }
    }
    if (gphase != cd(1.0, 0.0)) {
        bool folded = false;
        for (TileOpH &o : sw->ops) {
            if (o.kind == TK_U1 || o.kind == TK_U2) {
                const int nm = o.kind == TK_U1 ? 4 : 16;
                for (int t = 0; t < nm; t++) o.m[t] *= gphase;
                folded = true;
                break;
            }
        }
        if (!folded) {
            // a phase on every amplitude: DM with empty masks, first op
            TileOpH g;
            g.kind = TK_DM;
            g.m[0] = gphase;
            sw->ops.insert(sw->ops.begin(), g);
            for (size_t s = 0; s < sw->stages.size(); s++) {
                if (s > 0) sw->stages[s].op_begin++;
                sw->stages[s].op_end++;
            }
        }
    }
    for (TileOpH &o : sw->ops) {
        if (o.kind != TK_U1) continue;
        bool real = true, xform = true;
        for (int t = 0; t < 4; t++) {
            if (o.m[t].imag() != 0.0) real = false;
            if ((t == 0 || t == 3) ? o.m[t].imag() != 0.0 : o.m[t].real() != 0.0) xform = false;
        }
        if (real)
            o.kind = TK_U1R;
        else if (xform)
            o.kind = TK_U1X;
    }
}

SweepH build_sweep(const std::vector<Cls> &cl, const std::vector<uint32_t> &chosen, uint64_t T, uint32_t n,
                   const TileGeom &geo) {
    const uint32_t L = geo.L(), R = geo.R, c = std::min(geo.c, n);
    // stages over qubits
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> st;
    std::vector<uint32_t> rem = chosen;
    while (!rem.empty()) {
        const uint64_t Rq = choose_set(cl, rem, 0, T, R);
        std::vector<uint32_t> got;
        pick(cl, rem, Rq, &got, nullptr);
        if (got.empty()) fail(NCSV_INVALID, "tile planner made no progress");
        std::vector<uint32_t> left;
        size_t gi = 0;
        for (uint32_t i : rem) {
            if (gi < got.size() && got[gi] == i)
                gi++;
            else
                left.push_back(i);
        }
        st.push_back({Rq, got});
        rem.swap(left);
    }
    // positions: low qubits fixed; stage 0's register qubits on the top
    // (memory-layout register) positions; the rest ascending
    SweepH sw;
    sw.tq.assign(L, ~0u);
    std::vector<uint32_t> pos_of(64, ~0u);
    for (uint32_t p = 0; p < c; p++) {
        sw.tq[p] = p;
        pos_of[p] = p;
    }
    uint32_t top = L - R;
    for (uint64_t a = st.empty() ? 0 : st[0].first; a; a &= a - 1) {
        const uint32_t q = (uint32_t)__builtin_ctzll(a);
        while (top < L && sw.tq[top] != ~0u) top++;  // R = L: skip the fixed low positions
        if (pos_of[q] != ~0u || top >= L) continue;
        sw.tq[top] = q;
        pos_of[q] = top++;
    }
    uint32_t next = c;
    for (uint64_t a = T; a; a &= a - 1) {
        const uint32_t q = (uint32_t)__builtin_ctzll(a);
        if (pos_of[q] != ~0u) continue;
        while (next < L && sw.tq[next] != ~0u) next++;
        sw.tq[next] = q;
        pos_of[q] = next;
    }
    for (uint32_t p = 0; p < L; p++)
        if (sw.tq[p] == ~0u) fail(NCSV_INVALID, "tile planner left a position empty");
    // stage layouts and ops
    uint32_t prev = ((1u << R) - 1) << (L - R);  // memory layout
    std::vector<int> last(64, -1);               // last op index touching each qubit (U1 merging)
    for (auto &sg : st) {
        uint32_t need = 0;
        for (uint64_t a = sg.first; a; a &= a - 1) need |= 1u << pos_of[__builtin_ctzll(a)];
        uint32_t rpos = need;
        // the needed positions, padded from the previous layout, then from the top
        for (uint32_t p = L; p-- > 0 && (uint32_t)__builtin_popcount(rpos) < R;)
            if ((prev >> p) & 1) rpos |= 1u << p;
        for (uint32_t p = L; p-- > 0 && (uint32_t)__builtin_popcount(rpos) < R;) rpos |= 1u << p;
        StageH S;
        S.rmask_pos = rpos;
        S.op_begin = (uint32_t)sw.ops.size();
        std::fill(last.begin(), last.end(), -1);
        for (uint32_t i : sg.second) {
            const Cls &g = cl[i];
            TileOpH o;
            o.kind = g.kind;
            // conditions on the global mask split into register bits and the rest
            auto split = [&](uint64_t mask) {
                uint32_t rm = 0;
                uint64_t gm = 0;
                for (uint64_t a = mask; a; a &= a - 1) {
                    const uint32_t q = (uint32_t)__builtin_ctzll(a);
                    const int b = reg_bit(q, pos_of.data(), rpos);
                    if (b >= 0)
                        rm |= 1u << b;
                    else
                        gm |= 1ull << q;
                }
                o.rmask = rm;
                o.gmask = gm;
            };
            switch (g.kind) {
            case TK_U1: {
                o.j0 = (uint8_t)reg_bit(g.q0, pos_of.data(), rpos);
                for (int t = 0; t < 4; t++) o.m[t] = g.m[t];
            merge_u1:
                const int l = last[g.q0];
                if (l >= 0 && sw.ops[l].kind == TK_U1 && sw.ops[l].j0 == o.j0) {
                    // merge: later x earlier (FP64 on the host)
                    cd *e = sw.ops[l].m, r[4];
                    r[0] = o.m[0] * e[0] + o.m[1] * e[2];
                    r[1] = o.m[0] * e[1] + o.m[1] * e[3];
                    r[2] = o.m[2] * e[0] + o.m[3] * e[2];
                    r[3] = o.m[2] * e[1] + o.m[3] * e[3];
                    for (int t = 0; t < 4; t++) e[t] = r[t];
                    continue;
                }
                break;
            }
            case TK_U2: {
                int b0 = reg_bit(g.q0, pos_of.data(), rpos), b1 = reg_bit(g.q1, pos_of.data(), rpos);
                if (b0 < b1) {
                    for (int t = 0; t < 16; t++) o.m[t] = g.m[t];
                } else {
                    // swap the two local bits of the matrix
                    auto sw2 = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
                    for (int a = 0; a < 4; a++)
                        for (int b = 0; b < 4; b++) o.m[sw2(a) * 4 + sw2(b)] = g.m[a * 4 + b];
                    std::swap(b0, b1);
                }
                o.j0 = (uint8_t)b0;
                o.j1 = (uint8_t)b1;
                break;
            }
            case TK_SWAPR: {
                int b0 = reg_bit(g.q0, pos_of.data(), rpos), b1 = reg_bit(g.q1, pos_of.data(), rpos);
                o.j0 = (uint8_t)std::min(b0, b1);
                o.j1 = (uint8_t)std::max(b0, b1);
                break;
            }
            case TK_D1R: {
                const int b = reg_bit(g.q0, pos_of.data(), rpos);
                o.m[0] = g.m[0];
                o.m[1] = g.m[1];
                if (b >= 0 && geo.merge_diag) {
                    // a one-qubit diagonal on a register bit: a 2x2 that merges
                    // with its neighbours (split again by the peephole pass)
                    o.kind = TK_U1;
                    o.j0 = (uint8_t)b;
                    o.m[0] = g.m[0];
                    o.m[1] = cd(0, 0);
                    o.m[2] = cd(0, 0);
                    o.m[3] = g.m[1];
                    goto merge_u1;
                } else if (b >= 0) {
                    o.j0 = (uint8_t)b;
                } else {
                    o.kind = TK_D1T;
                    o.gmask = 1ull << g.q0;
                }
                break;
            }
            case TK_DM:
                split(g.cmask);
                o.m[0] = g.m[0];
                if (geo.merge_diag && o.gmask == 0 && __builtin_popcount(o.rmask) == 1 && g.cmask == g.sup &&
                    __builtin_popcountll(g.sup) == 1) {
                    // phase on one register bit: a diagonal 2x2, as above
                    o.kind = TK_U1;
                    o.j0 = (uint8_t)__builtin_ctz(o.rmask);
                    o.rmask = 0;
                    o.m[0] = cd(1, 0);
                    o.m[1] = cd(0, 0);
                    o.m[2] = cd(0, 0);
                    o.m[3] = g.m[0];
                    goto merge_u1;
                }
                break;
            case TK_PERM1:
                o.j0 = (uint8_t)reg_bit(g.q0, pos_of.data(), rpos);
                split(g.cmask);
                o.k0 = g.k0;
                o.k1 = g.k1;
                break;
            case TK_NSLOT:
                o.j0 = (uint8_t)reg_bit(g.q0, pos_of.data(), rpos);
                o.aux = g.slot;
                break;
            default: fail(NCSV_INVALID, "tile planner: unknown op");
            }
            for (uint64_t a = g.sup; a; a &= a - 1) last[__builtin_ctzll(a)] = (int)sw.ops.size();
            sw.ops.push_back(o);
        }
        S.op_end = (uint32_t)sw.ops.size();
        sw.stages.push_back(S);
        prev = rpos;
    }
    sw.gates = (uint32_t)chosen.size();
    finish_sweep(&sw);
    return sw;
}

// Vector instructions per amplitude a gate costs in a register sweep (the
// planner's balance estimate): 2x2s with real entries or a real diagonal
// and imaginary off-diagonal ~2.5, other 2x2s ~8.5, 4x4s ~16.5, phases and
// flips (applied lazily, or renamed) ~0.5.
double op_weight(const Cls &c) {
    switch (c.kind) {
    case TK_U1: {
        auto re = [](cd z) { return z.imag() == 0.0; };
        auto im = [](cd z) { return z.real() == 0.0; };
        const bool real = re(c.m[0]) && re(c.m[1]) && re(c.m[2]) && re(c.m[3]);
        const bool xform = re(c.m[0]) && re(c.m[3]) && im(c.m[1]) && im(c.m[2]);
        return (real || xform) ? 2.5 : 8.5;
    }
    case TK_U2: return 16.5;
    case TK_SWAPR: return 0.0;
    default: return 0.5;
    }
}

// cap > 0: a sweep admits its gates in program order only while their
// summed weight stays within cap (at least one gate); the rest wait for a
// later sweep (a prefix of the admissible gates is closed under their
// dependencies).
std::vector<SweepH> plan_sweeps(const std::vector<Cls> &cl, std::vector<uint32_t> rem, uint32_t n,
                                const TileGeom &geo, uint64_t first_fixed = 0, uint32_t variant = 0,
                                double cap = 0.0, std::vector<std::vector<uint32_t>> *gots = nullptr,
                                std::vector<uint64_t> *sets = nullptr) {
    std::vector<SweepH> out;
    const uint32_t L = geo.L();
    const uint64_t fixed = (1ull << std::min(geo.c, n)) - 1;
    const uint64_t pool = n >= 64 ? ~0ull : (1ull << n) - 1;
    while (!rem.empty()) {
        const uint64_t T = choose_set(cl, rem, fixed | (out.empty() ? first_fixed : 0), pool, L, variant);
        std::vector<uint32_t> got;
        pick(cl, rem, T, &got, nullptr);
        if (got.empty()) fail(NCSV_INVALID, "sweep planner made no progress");
        if (cap > 0.0) {
            double w = 0.0;
            size_t keep = 0;
            while (keep < got.size() && (keep == 0 || w + op_weight(cl[got[keep]]) <= cap)) w += op_weight(cl[got[keep++]]);
            got.resize(keep);
        }
        out.push_back(build_sweep(cl, got, T, n, geo));
        if (gots) gots->push_back(got);
        if (sets) sets->push_back(T);
        std::vector<uint32_t> left;
        size_t gi = 0;
        for (uint32_t i : rem) {
            if (gi < got.size() && got[gi] == i)
                gi++;
            else
                left.push_back(i);
        }
        rem.swap(left);
    }
    return out;
}

// A cached plan: the segment's items and their device programs (the
// interpreter's op lists, or the specialized kernels' coefficients).
struct Plan {
    std::vector<SegItem> items;
    std::vector<size_t> prog_off;  // offset of each tile item's program in dbuf
    std::vector<void *> kern;      // specialized kernel per tile item (JIT plans)
    void *kern_from_basis = nullptr;  // the first item's kernel started from a basis state (made with the plan)
    void *dbuf = nullptr;
    TileGeom geo;  // the tile shape the plan was made for
    std::vector<jit::MemShape> mem;  // memory layout and block order of each specialized sweep
    ~Plan() {
        for (void *k : kern)
            if (k) jit::release(k);
        if (kern_from_basis) jit::release(kern_from_basis);
        if (dbuf) (void)hipFree(dbuf);
    }
};



}  // namespace

// The memory layouts a specialized sweep may load and store in: the default
// (the top R tile positions), the R positions holding the lowest and the
// highest qubits, every other position by qubit order (both phases), and the
// first and last register stage's layouts (an exchange fewer) where they
// leave the fixed low positions to the threads.
std::vector<uint32_t> mem_candidates(const SweepH &sw, const TileGeom &geo) {
    const uint32_t L = geo.L(), R = geo.R, c = geo.c;
    std::vector<uint32_t> out;
    auto add = [&](uint32_t m) {
        if (!m || (m & ((1u << c) - 1)) || (uint32_t)__builtin_popcount(m) != R || (m >> L)) return;
        if (std::find(out.begin(), out.end(), m) == out.end()) out.push_back(m);
    };
    add(jit::default_memr(geo));
    std::vector<uint32_t> P;
    for (uint32_t p = c; p < L; p++) P.push_back(p);
    std::sort(P.begin(), P.end(), [&](uint32_t a, uint32_t b) { return sw.tq[a] < sw.tq[b]; });
    uint32_t lo = 0, hi = 0;
    for (uint32_t i = 0; i < R; i++) {
        lo |= 1u << P[i];
        hi |= 1u << P[P.size() - 1 - i];
    }
    add(lo);
    add(hi);
    for (uint32_t ph = 0; ph < 2; ph++) {
        uint32_t m = 0, k = 0;
        for (uint32_t i = ph; i < P.size() && k < R; i += 2, k++) m |= 1u << P[i];
        for (uint32_t i = 1 - ph; i < P.size() && k < R; i += 2, k++) m |= 1u << P[i];
        add(m);
    }
    if (!sw.stages.empty()) {
        add(sw.stages.front().rmask_pos);
        add(sw.stages.back().rmask_pos);
    }
    return out;
}

// The first and last register stage's layouts as the load and store layouts
// where they leave the fixed low positions to the threads (an exchange fewer
// each), else the default.
jit::MemShape stage_memory(const SweepH &sw, const TileGeom &geo) {
    jit::MemShape m;
    if (sw.stages.empty()) return m;
    const uint32_t low = (1u << geo.c) - 1;
    const uint32_t f = sw.stages.front().rmask_pos, l = sw.stages.back().rmask_pos;
    if (!(f & low) && (uint32_t)__builtin_popcount(f) == geo.R) m.load = f;
    if (!(l & low) && (uint32_t)__builtin_popcount(l) == geo.R) m.store = l;
    return m;
}

jit::MemShape forced_memory(const SweepH &sw, const TileGeom &geo) {
    jit::MemShape m;
    const char *a = getenv("NCSV_JIT_MEMSHAPE"), *b = getenv("NCSV_JIT_MEMSHAPE_STORE");
    if (!a && !b) return stage_memory(sw, geo);
    if (a && !strcmp(a, "default")) return m;
    const std::vector<uint32_t> c = mem_candidates(sw, geo);
    if (a) m.load = m.store = c[(size_t)atoi(a) % c.size()];
    if (b) m.store = c[(size_t)atoi(b) % c.size()];
    return m;
}

namespace {

}  // namespace

std::vector<SweepH> plan_lds_sweeps(const std::vector<LOp> &ops, uint32_t n, uint32_t L, uint32_t c,
                                    uint64_t first_fixed) {
    if (L > n || c > L || (uint32_t)__builtin_popcountll(first_fixed | ((1ull << c) - 1)) > L)
        fail(NCSV_INVALID, "tile planner: bad tile shape");
    std::vector<Cls> cl;
    std::vector<uint32_t> run;
    for (const LOp &op : ops) {
        if (op.kind == OpKind::Nop) continue;
        Cls k = classify(op);
        if (!k.tileable) fail(NCSV_UNSUPPORTED, "batched form: a gate on " + std::to_string(op.k()) +
                                                     " qubits that is not a tile op (dense on more than 2 qubits)");
        cl.push_back(k);
        run.push_back((uint32_t)cl.size() - 1);
    }
    TileGeom geo;
    geo.LT = 0;
    geo.R = L;
    geo.c = c;
    geo.merge_diag = true;
    std::vector<SweepH> out = plan_sweeps(cl, run, n, geo, first_fixed);
    // states of 2^(L + 8) amplitudes or more (a step over them moves hundreds
    // of tiles): the greedy tile choice and 7 permuted variants of it, keeping
    // the fewest steps (4096-circuit complex128 mix: 60,883 -> 59,185 steps,
    // kernel 773.1 -> 766.9 ms; 24 variants: 764.8 ms but 8.4 s of planning);
    // NCSV_BATCH_VARIANTS sets the count
    uint32_t variants = n >= L + 8 ? 8u : 1u;
    if (const char *e = getenv("NCSV_BATCH_VARIANTS")) variants = (uint32_t)std::max(1, atoi(e));
    for (uint32_t v = 1; v < variants && n > L; v++) {
        std::vector<SweepH> alt;
        try {
            alt = plan_sweeps(cl, run, n, geo, first_fixed, v);
        } catch (const Error &) {
            continue;
        }
        if (alt.size() < out.size()) out.swap(alt);
    }
    if (out.empty() && first_fixed) {
        // no gates: an empty sweep whose tile holds the fixed qubits
        SweepH sw;
        uint64_t T = ((1ull << c) - 1) | first_fixed;
        for (uint32_t q = 0; q < n && (uint32_t)__builtin_popcountll(T) < L; q++) T |= 1ull << q;
        for (uint32_t q = 0; q < c; q++) sw.tq.push_back(q);
        for (uint64_t a = T & ~((1ull << c) - 1); a; a &= a - 1) sw.tq.push_back((uint32_t)__builtin_ctzll(a));
        out.push_back(sw);
    }
    return out;
}

// Plans the segment; for specialized kernels also with one and two more
// fixed low qubits (512 B and 1 KiB runs of complex128, 256 B to 1 KiB of
// complex64) and keeps the plan whose sweeps cost least: a sweep over runs
// that long reads and writes the state in 0.94 and 0.86 of the time of the
// default runs on MI355X (30-qubit rotation brickwork and 33-qubit QFT
// sweeps), but fewer free tile qubits may need more sweeps.  NCSV_TILE_C
// fixes the count instead.
std::vector<SegItem> plan_segment_tuned(const std::vector<LOp> &ops, uint32_t n, TileGeom *geo, bool specialized) {
    std::vector<SegItem> best = plan_segment(ops, n, *geo);
    if (!specialized || getenv("NCSV_TILE_C")) return best;
    static const double run_cost[3] = {1.0, 0.94, 0.86};
    auto cost = [&](const std::vector<SegItem> &items, uint32_t dc) {
        double t = 0.0;
        for (const SegItem &it : items) t += it.tile ? run_cost[dc] : 1.0;
        return t;
    };
    double bc = cost(best, 0);
    const TileGeom g0 = *geo;
    for (uint32_t dc = 1; dc <= 2; dc++) {
        TileGeom g = g0;
        g.c = g0.c + dc;
        if (g.c > g.LT || g.c > n) break;
        std::vector<SegItem> items = plan_segment(ops, n, g);
        const double c = cost(items, dc);
        if (c < bc - 1e-9) {
            bc = c;
            best.swap(items);
            *geo = g;
        }
    }
    return best;
}

bool use_specialized(const ncsv_state *s) {
    if (!jit::available() || s->fusion == 0) return false;
    if (s->fusion == 2) return true;
    int minq = 20;
    if (const char *e = getenv("NCSV_JIT_MIN_QUBITS")) minq = atoi(e);
    return (int)s->n >= minq;
}

// Whether the planner evens out the arithmetic of a plan's sweeps
// (NCSV_PLAN_BALANCE=0: never).
bool balance_plans() {
    const char *e = getenv("NCSV_PLAN_BALANCE");
    return e ? atoi(e) != 0 : true;
}

// Evens out a plan without changing its sweeps' tiles or count: a sweep
// whose gates cost more arithmetic than a pass over memory hides (W0 vector
// instructions per amplitude) hands gates to the next sweep, latest first,
// while the pair's larger cost goes down.  A gate may move when it commutes
// with every later gate of its sweep (so the gates left stay closed under
// their dependencies) and the next sweep's tile holds the qubits it acts on
// non-diagonally; it commutes with the next sweep's earlier gates by
// construction (those were left out of its sweep, and it was not).  Returns
// whether anything moved.
bool balance_sweeps(const std::vector<Cls> &cl, uint32_t n, const TileGeom &geo,
                    std::vector<std::vector<uint32_t>> *gots, const std::vector<uint64_t> &sets) {
    (void)n;
    const double W0 = geo.floor_ops >= 100 ? 128.0 : 64.0;
    auto weight = [&](const std::vector<uint32_t> &g) {
        double w = 0.0;
        for (uint32_t i : g) w += op_weight(cl[i]);
        return w;
    };
    auto commute = [&](const Cls &a, const Cls &b) { return !(a.act & b.sup) && !(b.act & a.sup); };
    bool any = false;
    for (int pass = 0; pass < 8; pass++) {
        bool moved = false;
        for (size_t k = 0; k + 1 < gots->size(); k++) {
            std::vector<uint32_t> &A = (*gots)[k], &B = (*gots)[k + 1];
            double wa = weight(A), wb = weight(B);
            while (wa > W0) {
                // the latest movable gate of A
                int pick = -1;
                for (int j = (int)A.size() - 1; j >= 0 && pick < 0; j--) {
                    const Cls &g = cl[A[(size_t)j]];
                    if (g.act & ~sets[k + 1]) continue;
                    const double w = op_weight(g);
                    if (w <= 0.0 || wb + w >= wa) continue;
                    bool ok = true;
                    for (size_t t = (size_t)j + 1; t < A.size() && ok; t++) ok = commute(g, cl[A[t]]);
                    if (ok) pick = j;
                }
                if (pick < 0) break;
                const uint32_t gi = A[(size_t)pick];
                A.erase(A.begin() + pick);
                B.insert(std::lower_bound(B.begin(), B.end(), gi), gi);
                wa -= op_weight(cl[gi]);
                wb += op_weight(cl[gi]);
                moved = any = true;
            }
        }
        if (!moved) break;
    }
    return any;
}

std::vector<SegItem> plan_segment(const std::vector<LOp> &ops, uint32_t n, const TileGeom &geo) {
    std::vector<SegItem> items;
    std::vector<Cls> cl;
    std::vector<uint32_t> run;
    auto flush = [&]() {
        if (run.empty()) return;
        // the greedy tile choice, and on states of 2^20 amplitudes or more
        // (where a sweep costs milliseconds) 23 permuted variants of it: the
        // plan with the fewest sweeps, then the least estimated time (a sweep
        // costs the larger of its ops and geo.floor_ops, the ops its memory
        // traffic takes as long as), then the fewest register stages
        std::vector<SweepH> best = plan_sweeps(cl, run, n, geo);
        auto score = [&](const std::vector<SweepH> &p) {
            size_t t = 0, k = 0;
            for (const SweepH &sw : p) {
                t += std::max<size_t>(sw.ops.size(), geo.floor_ops);
                k += sw.stages.size();
            }
            return std::make_tuple(p.size(), t, k);
        };
        if (n >= 20) {
            auto bs = score(best);
            uint32_t bv = 0;
            for (uint32_t v = 1; v < 24; v++) {
                std::vector<SweepH> c;
                try {
                    c = plan_sweeps(cl, run, n, geo, 0, v);
                } catch (const Error &) {
                    continue;  // a variant that stalls is only a worse choice
                }
                const auto cs = score(c);
                if (cs < bs) {
                    best.swap(c);
                    bs = cs;
                    bv = v;
                }
            }
            if (balance_plans()) {
                std::vector<std::vector<uint32_t>> gots;
                std::vector<uint64_t> sets;
                (void)plan_sweeps(cl, run, n, geo, 0, bv, 0.0, &gots, &sets);
                if (balance_sweeps(cl, n, geo, &gots, sets))
                    for (size_t k = 0; k < gots.size(); k++) best[k] = build_sweep(cl, gots[k], sets[k], n, geo);
            }
        }
        for (auto &sw : best) {
            SegItem it;
            it.tile = true;
            it.sw = std::move(sw);
            items.push_back(std::move(it));
        }
        run.clear();
    };
    for (const LOp &op : ops) {
        if (op.kind == OpKind::Nop) continue;
        Cls c = classify(op);
        if (!c.tileable || n < geo.L()) {
            flush();
            SegItem it;
            it.tile = false;
            it.op = op;
            items.push_back(std::move(it));
            continue;
        }
        cl.push_back(c);
        run.push_back((uint32_t)cl.size() - 1);
    }
    flush();
    return items;
}

void run_gate_segment(ncsv_state *s, ncsv_circuit *c, size_t i, size_t j) {
    const TileGeom geo0 = dev::tile_geom(s->prec);
    if (s->fusion == 0 || s->n < geo0.L()) {
        materialize(s);
        for (size_t t = i; t < j; t++) {
            const COp &op = c->ops[t];
            if (op.kind != CKind::Gate) continue;
            dev::apply_lop(s, lower_gate(op.g, op.qubits.data(), (uint32_t)op.qubits.size(), s->n));
        }
        return;
    }
    const bool jt = use_specialized(s);
    const int64_t key = ((int64_t)i << 21) | ((int64_t)geo0.key() << 3) | ((int64_t)jt << 2) | (int64_t)s->prec;
    std::shared_ptr<Plan> plan;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->plans.find(key);
        if (it != c->plans.end()) plan = std::static_pointer_cast<Plan>(it->second);
    }
    if (!plan) {
        plan = std::make_shared<Plan>();
        std::vector<LOp> ops;
        for (size_t t = i; t < j; t++) {
            const COp &op = c->ops[t];
            if (op.kind == CKind::Gate) ops.push_back(lower_gate(op.g, op.qubits.data(), (uint32_t)op.qubits.size(), s->n));
        }
        plan->geo = geo0;
        plan->items = plan_segment_tuned(ops, s->n, &plan->geo, jt);
        const TileGeom &geo = plan->geo;
        size_t total = 0;
        std::vector<char> host;
        plan->kern.assign(plan->items.size(), nullptr);
        plan->mem.assign(plan->items.size(), jit::MemShape());
        if (jt)
            for (size_t k = 0; k < plan->items.size(); k++)
                if (plan->items[k].tile) plan->mem[k] = forced_memory(plan->items[k].sw, geo);
        if (jt) {
            // specialized kernels: sources compile in parallel; coefficients as T
            const size_t tsz = s->prec == NCSV_C128 ? 8 : 4;
            for (size_t k = 0; k < plan->items.size(); k++) {
                plan->prog_off.push_back(total);
                if (!plan->items[k].tile) continue;
                std::vector<double> coef;
                const std::string src =
                    jit::generate(plan->items[k].sw, s->n, geo, s->prec == NCSV_C128, &coef, false, &plan->mem[k]);
                plan->kern[k] = jit::prepare(src);
                host.resize(total + (coef.size() * tsz + 255) / 256 * 256 + 256);
                for (size_t t = 0; t < coef.size(); t++) {
                    if (tsz == 8)
                        memcpy(host.data() + total + 8 * t, &coef[t], 8);
                    else {
                        const float f = (float)coef[t];
                        memcpy(host.data() + total + 4 * t, &f, 4);
                    }
                }
                total = host.size();
            }
            // the first sweep's form for runs from a deferred basis state,
            // compiled alongside the others
            if (!plan->items.empty() && plan->items[0].tile) {
                std::vector<double> cb;
                plan->kern_from_basis = jit::prepare_named(
                    jit::generate(plan->items[0].sw, s->n, geo, s->prec == NCSV_C128, &cb, true, &plan->mem[0]),
                    "ncsv_sweep_from_basis");
            }
        } else {
            for (auto &it : plan->items) {
                plan->prog_off.push_back(total);
                if (it.tile) total += dev::tile_prog_bytes(it.sw, s->prec);
            }
            host.resize(total);
            for (size_t k = 0; k < plan->items.size(); k++)
                if (plan->items[k].tile) dev::tile_prog_pack(plan->items[k].sw, s->prec, host.data() + plan->prog_off[k]);
        }
        if (total) {
            check_hip(hipMalloc(&plan->dbuf, total), "hipMalloc(tile programs)");
            check_hip(hipMemcpy(plan->dbuf, host.data(), total, hipMemcpyHostToDevice), "upload tile programs");
        }
        std::lock_guard<std::mutex> lk(c->mu);
        c->plans[key] = plan;
    }
    const TileGeom &geo = plan->geo;
    for (size_t k = 0; k < plan->items.size(); k++) {
        const SegItem &it = plan->items[k];
        if (s->pending_basis >= 0) {
            if (k == 0 && it.tile && plan->kern[0]) {
                // from a deferred basis state: the first sweep computes its
                // input in registers instead of reading a filled state (its
                // kernel is made with the plan; made here for a plan without it)
                void *kb;
                {
                    std::lock_guard<std::mutex> lk(c->mu);
                    if (!plan->kern_from_basis) {
                        std::vector<double> cb;
                        plan->kern_from_basis = jit::prepare_named(
                            jit::generate(it.sw, s->n, geo, s->prec == NCSV_C128, &cb, true, &plan->mem[0]),
                            "ncsv_sweep_from_basis");
                    }
                    kb = plan->kern_from_basis;
                }
                // the state counts as written only once the launch went out
                const uint64_t b = (uint64_t)s->pending_basis;
                jit::launch_from_basis(s, kb, geo.L(), geo.LT, (const char *)plan->dbuf + plan->prog_off[0], b);
                s->pending_basis = -1;
                continue;
            }
            materialize(s);
        }
        if (it.tile && plan->kern[k])
            jit::launch(s, plan->kern[k], geo.L(), geo.LT, (const char *)plan->dbuf + plan->prog_off[k]);
        else if (it.tile)
            dev::run_tile_sweep(s, it.sw, (const char *)plan->dbuf + plan->prog_off[k], geo);
        else
            dev::apply_lop(s, it.op);
    }
}

void fuse_release(ncsv_state *) {}

}  // namespace ncsv

// ---------------------------------------------------------------------------
// Host emulation of the fused schedule (diagnostics): applies a unitary
// circuit's plan to host amplitudes in FP64 with the kernels' op semantics
// (register bits per stage, global-index conditions), so the planner can be
// checked without a device.
namespace ncsv {
namespace {

using cdv = std::vector<cd>;

cd ipow_h(int k, cd a) {
    switch (k & 3) {
    case 0: return a;
    case 1: return cd(-a.imag(), a.real());
    case 2: return cd(-a.real(), -a.imag());
    default: return cd(a.imag(), -a.real());
    }
}

uint64_t gmask_all(const TileOpH &) { return 0; }

void emu_op(const TileOpH &o, cd *v, uint32_t NR, uint64_t g) {
    switch (o.kind) {
    case TK_U1:
        for (uint32_t i = 0; i < NR; i++) {
            if (i & (1u << o.j0)) continue;
            const cd x = v[i], y = v[i | (1u << o.j0)];
            v[i] = o.m[0] * x + o.m[1] * y;
            v[i | (1u << o.j0)] = o.m[2] * x + o.m[3] * y;
        }
        break;
    case TK_U1R:
    case TK_U1X: {
        TileOpH g = o;
        g.kind = TK_U1;
        emu_op(g, v, NR, gmask_all(g));
        break;
    }
    case TK_U2:
        for (uint32_t i = 0; i < NR; i++) {
            if (i & ((1u << o.j0) | (1u << o.j1))) continue;
            const uint32_t ix[4] = {i, i | (1u << o.j0), i | (1u << o.j1), i | (1u << o.j0) | (1u << o.j1)};
            cd x[4], y[4];
            for (int t = 0; t < 4; t++) x[t] = v[ix[t]];
            for (int a = 0; a < 4; a++) {
                y[a] = 0;
                for (int b = 0; b < 4; b++) y[a] += o.m[4 * a + b] * x[b];
            }
            for (int t = 0; t < 4; t++) v[ix[t]] = y[t];
        }
        break;
    case TK_D1R:
        for (uint32_t i = 0; i < NR; i++) v[i] *= ((i >> o.j0) & 1) ? o.m[1] : o.m[0];
        break;
    case TK_D1T:
        for (uint32_t i = 0; i < NR; i++) v[i] *= (g & o.gmask) ? o.m[1] : o.m[0];
        break;
    case TK_DM:
        if ((g & o.gmask) != o.gmask) break;
        for (uint32_t i = 0; i < NR; i++)
            if ((i & o.rmask) == o.rmask) v[i] *= o.m[0];
        break;
    case TK_PERM1:
        if ((g & o.gmask) != o.gmask) break;
        for (uint32_t i = 0; i < NR; i++) {
            if (i & (1u << o.j0)) continue;
            if ((i & o.rmask) != o.rmask) continue;
            const cd x = v[i], y = v[i | (1u << o.j0)];
            v[i] = ipow_h(o.k1, y);
            v[i | (1u << o.j0)] = ipow_h(o.k0, x);
        }
        break;
    case TK_SWAPR:
        for (uint32_t i = 0; i < NR; i++)
            if ((i & (1u << o.j0)) && !(i & (1u << o.j1))) std::swap(v[i], v[(i & ~(1u << o.j0)) | (1u << o.j1)]);
        break;
    }
}

void emu_sweep(const SweepH &sw, cd *amp, uint32_t n, const TileGeom &geo) {
    const uint32_t L = geo.L(), R = geo.R, NR = 1u << R;
    uint64_t in = 0;
    for (uint32_t q : sw.tq) in |= 1ull << q;
    std::vector<uint32_t> oq;
    for (uint32_t q = 0; q < n; q++)
        if (!((in >> q) & 1)) oq.push_back(q);
    std::vector<cd> loc(size_t(1) << L);
    for (uint64_t b = 0; b < (1ull << (n - L)); b++) {
        uint64_t base = 0;
        for (size_t i = 0; i < oq.size(); i++) base |= ((b >> i) & 1ull) << oq[i];
        auto gidx = [&](uint32_t t) {
            uint64_t g = base;
            for (uint32_t p = 0; p < L; p++)
                if ((t >> p) & 1) g |= 1ull << sw.tq[p];
            return g;
        };
        for (uint32_t t = 0; t < (1u << L); t++) loc[t] = amp[gidx(t)];
        for (const StageH &S : sw.stages) {
            std::vector<uint32_t> rp, tp;
            for (uint32_t p = 0; p < L; p++) ((S.rmask_pos >> p) & 1 ? rp : tp).push_back(p);
            for (uint32_t th = 0; th < (1u << tp.size()); th++) {
                uint32_t tt = 0;
                for (size_t i = 0; i < tp.size(); i++)
                    if ((th >> i) & 1) tt |= 1u << tp[i];
                const uint64_t g = gidx(tt);
                std::vector<cd> v(NR);
                std::vector<uint32_t> ti(NR);
                for (uint32_t r = 0; r < NR; r++) {
                    uint32_t t = tt;
                    for (uint32_t j = 0; j < R; j++)
                        if ((r >> j) & 1) t |= 1u << rp[j];
                    ti[r] = t;
                    v[r] = loc[t];
                }
                for (uint32_t o = S.op_begin; o < S.op_end; o++) emu_op(sw.ops[o], v.data(), NR, g);
                for (uint32_t r = 0; r < NR; r++) loc[ti[r]] = v[r];
            }
        }
        for (uint32_t t = 0; t < (1u << L); t++) amp[gidx(t)] = loc[t];
    }
}

}  // namespace
}  // namespace ncsv

extern ncsv_circuit *ncsv_parse_text(const char *text);

extern "C" {

// Diagnostics: the fused schedule of a unitary circuit (gates only) for a
// state of `prec`, applied on the host to amp (2^n interleaved doubles), and
// its shape: sweeps, stages, ops, standalone gates.
ncsv_status ncsv_plan_emulate(const ncsv_circuit *c, ncsv_precision prec, double *amp, uint64_t *stats) {
    try {
        using namespace ncsv;
        TileGeom geo = dev::tile_geom(prec);
        std::vector<LOp> ops;
        for (const COp &op : c->ops) {
            if (op.kind == CKind::Gate) ops.push_back(lower_gate(op.g, op.qubits.data(), (uint32_t)op.qubits.size(), c->n));
            else if (op.kind != CKind::Tick) fail(NCSV_INVALID, "plan emulation takes unitary circuits");
        }
        // the plan a fusion-1 run makes (specialized kernels from 20 qubits on)
        const std::vector<SegItem> items = plan_segment_tuned(ops, c->n, &geo, c->n >= 20 && jit::available());
        uint64_t st[4] = {0, 0, 0, 0};
        cd *a = (cd *)amp;
        for (const SegItem &it : items) {
            if (it.tile) {
                st[0]++;
                st[1] += it.sw.stages.size();
                st[2] += it.sw.ops.size();
                if (a) emu_sweep(it.sw, a, c->n, geo);
            } else {
                st[3]++;
                if (a) fail(NCSV_UNSUPPORTED, "plan emulation: standalone gate in the plan");
            }
        }
        if (stats)
            for (int i = 0; i < 4; i++) stats[i] = st[i];
        return NCSV_OK;
    } catch (ncsv::Error &e) {
        return e.code;
    }
}

}  // extern "C"
