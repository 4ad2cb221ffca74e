// Host side of the batched form: per-circuit programs (planned once and
// shared by every entry running the same circuit), the queue of (entry,
// step) records, chunking by device memory, the launch, and the FP64
// results.  A host emulator runs the same programs on FP64 amplitudes so the
// planner can be checked without a device.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "batch.hpp"
#include "fuse.hpp"
#include "ncsv_internal.hpp"

namespace ncsv {
namespace batch {

uint32_t max_tile_bits(ncsv_precision prec);
int resident_blocks(ncsv_precision prec, uint32_t threads);
void launch(const BArgs &A, uint32_t nrec, ncsv_precision prec, uint32_t threads, int blocks, hipStream_t stream);
void launch_finish(const double *part, const BSumD *sums, uint32_t count, double *out, hipStream_t stream);
void simulate(const BArgs &A, ncsv_precision prec, uint32_t threads);

struct StepH {
    uint32_t L = 0;
    std::vector<uint8_t> qt;  // tile qubits, then outer qubits ascending
    std::vector<BOpD> ops;    // as planned
    std::vector<BOpD> gops, gopsc;  // grouped; grouped without noise slots
    std::vector<uint32_t> slots;    // noise slots of this step
    std::vector<std::vector<BOpD>> gops1;  // grouped keeping only slot k (empty: none built)
    std::vector<cd> coef;
    std::vector<BPauliD> paulis;
    std::vector<BMGroupD> mgroups;
    uint32_t flags = 0;
    bool modifies = false;
};

struct NoiseRef {
    ncsv_noise ch;
    double p[3];
    uint32_t slot0, k;
};

// where one Pauli's per-tile partials live
struct ExpRef {
    uint64_t off;
    uint32_t K, k, tiles;
};

struct Program {
    uint32_t n = 0, L = 0;
    std::vector<StepH> steps;
    uint32_t nmeas = 0, nslots = 0;
    uint64_t npart = 0, nexp = 0;     // partial doubles per entry
    std::vector<uint8_t> draw_kind;   // 0 noise, 1 measure, 2 reset
    std::vector<uint32_t> draw_ref;   // noise index or measurement ordinal
    std::vector<NoiseRef> noise;
    std::vector<ExpRef> exps;         // the Paulis, then the identity (norm)
    uint64_t gates = 0;               // unfused gates
    uint64_t items = 0;               // tiles over all steps
    uint64_t hbm_elems = 0;           // amplitudes read + written over all steps
    // placement in the device tables
    uint32_t step0 = 0;
    uint32_t tiles(uint32_t s) const { return 1u << (n - steps[s].L); }
    // per step, what each run's queue records need (kept apart from the step
    // lists, which the host would otherwise walk once per record): log2 of
    // the tiles, and whether the step holds noise slots
    std::vector<uint8_t> step_lg, step_noisy;
};

namespace {

uint64_t splitmix_next(uint64_t *s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double splitmix_f64(uint64_t *s) { return (double)(splitmix_next(s) >> 11) * (1.0 / 9007199254740992.0); }

// A Pauli gate as one-qubit tile ops: Z factors, then X factors, then i^ph
// (apply_pauli's action, vector.rs:397-410).
void expand_pauli(const LOp &p, std::vector<LOp> *out) {
    for (uint64_t a = p.z; a; a &= a - 1) {
        LOp d;
        d.kind = OpKind::Diag;
        d.qubits = {(uint32_t)__builtin_ctzll(a)};
        d.m = {cd(1, 0), cd(-1, 0)};
        out->push_back(d);
    }
    for (uint64_t a = p.x; a; a &= a - 1) {
        LOp d;
        d.kind = OpKind::Perm;
        d.qubits = {(uint32_t)__builtin_ctzll(a)};
        d.m = {cd(1, 0), cd(1, 0)};
        d.img = {1, 0};
        out->push_back(d);
    }
    if (p.ph & 3) {
        static const cd ip[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
        LOp d;
        d.kind = OpKind::Diag;
        d.m = {ip[p.ph & 3]};
        out->push_back(d);
    }
}

struct Ev {
    bool meas = false;
    std::vector<LOp> ops;
    bool reset = false;
    uint64_t z = 0;
    uint32_t ph = 0, q = 0, ord = 0;
};

// Consecutive measurement events decided from one reduction.
struct MGroup {
    std::vector<uint64_t> masks;  // distinct Z masks (at most 4)
    std::vector<Ev> evs;          // at most 16
    uint64_t resets = 0;          // reset qubits
    bool accepts(const Ev &e) const {
        if (evs.size() >= 16) return false;
        if (e.z & resets) return false;  // a reset qubit is not measured again within the run
        if (std::find(masks.begin(), masks.end(), e.z) == masks.end() && masks.size() >= 4) return false;
        return true;
    }
    void add(const Ev &e) {
        if (std::find(masks.begin(), masks.end(), e.z) == masks.end()) masks.push_back(e.z);
        evs.push_back(e);
        if (e.reset) resets |= 1ull << e.q;
    }
};

int pos_in(const StepH &s, uint32_t q) {
    for (uint32_t p = 0; p < s.L; p++)
        if (s.qt[p] == q) return (int)p;
    return -1;
}

void split_mask(const StepH &s, uint64_t m, uint32_t *lm, uint64_t *om) {
    *lm = 0;
    *om = 0;
    for (uint64_t a = m; a; a &= a - 1) {
        const uint32_t q = (uint32_t)__builtin_ctzll(a);
        const int p = pos_in(s, q);
        if (p >= 0)
            *lm |= 1u << p;
        else
            *om |= 1ull << q;
    }
}

// The group's descriptor for a step's tile layout (index into s->mgroups);
// reset positions only where the step applies the resets.
uint32_t group_desc(StepH *s, const MGroup &g, bool resets = true) {
    BMGroupD d{};
    d.D = (uint32_t)g.masks.size();
    d.E = (uint32_t)g.evs.size();
    for (uint32_t i = 0; i < d.D; i++) split_mask(*s, g.masks[i], &d.lm[i], &d.om[i]);
    for (uint32_t i = 0; i < d.E; i++) {
        const Ev &e = g.evs[i];
        d.ev[i].d = (uint8_t)(std::find(g.masks.begin(), g.masks.end(), e.z) - g.masks.begin());
        d.ev[i].sign = (uint8_t)(e.ph >> 1);
        d.ev[i].ord = e.ord;
        d.ev[i].rpos = 0xff;
        if (e.reset && resets) {
            const int p = pos_in(*s, e.q);
            if (p < 0) fail(NCSV_INVALID, "batched planner: reset qubit outside the tile");
            d.ev[i].rpos = (uint8_t)p;
        }
    }
    s->mgroups.push_back(d);
    return (uint32_t)s->mgroups.size() - 1;
}

void append_tile_ops(StepH *s, const SweepH &sw) {
    for (const TileOpH &o : sw.ops) {
        BOpD d{};
        d.p0 = o.j0;
        d.p1 = o.j1;
        d.k0 = o.k0;
        d.k1 = o.k1;
        d.lmask = o.rmask;
        d.omask = o.gmask;
        auto put = [&](int cnt) {
            d.coef = (uint32_t)s->coef.size();
            for (int t = 0; t < cnt; t++) s->coef.push_back(o.m[t]);
        };
        switch (o.kind) {
        case TK_U1: d.kind = BK_U1; put(4); break;
        case TK_U1R: d.kind = BK_U1R; put(4); break;
        case TK_U1X: d.kind = BK_U1X; put(4); break;
        case TK_U2: d.kind = BK_U2; put(16); break;
        case TK_DM: d.kind = BK_DM; put(1); break;
        case TK_PERM1: d.kind = BK_PERM1; break;
        case TK_SWAPR: d.kind = BK_SWAP; break;
        case TK_NSLOT:
            d.kind = BK_NSLOT;
            d.a = o.aux;
            s->slots.push_back(o.aux);
            break;
        default: fail(NCSV_INVALID, "batched form: unexpected tile op");
        }
        s->ops.push_back(d);
    }
}

// The 2x2 of an unconditional one-position op, or false.
bool as_2x2(const StepH &s, const BOpD &o, uint32_t *pos, cd m[4]) {
    static const cd ip[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
    switch (o.kind) {
    case BK_U1:
    case BK_U1R:
    case BK_U1X:
        if (o.omask) return false;  // two 2x2s chosen by the outer condition
        *pos = o.p0;
        for (int t = 0; t < 4; t++) m[t] = s.coef[o.coef + t];
        return true;
    case BK_DM:
        if (o.omask || __builtin_popcount(o.lmask) != 1) return false;
        *pos = (uint32_t)__builtin_ctz(o.lmask);
        m[0] = cd(1, 0);
        m[1] = m[2] = cd(0, 0);
        m[3] = s.coef[o.coef];
        return true;
    case BK_PERM1:
        if (o.omask || o.lmask) return false;
        *pos = o.p0;
        m[0] = m[3] = cd(0, 0);
        m[1] = ip[o.k1 & 3];
        m[2] = ip[o.k0 & 3];
        return true;
    default: return false;
    }
}

// The tile positions an op acts on, when it can join a dense block (at most
// three positions; a condition on outer qubits is allowed); 0 otherwise.
uint32_t block_support(const BOpD &o) {
    uint32_t m = 0;
    switch (o.kind) {
    case BK_U1:
    case BK_U1R:
    case BK_U1X: m = 1u << o.p0; break;
    case BK_U2:
    case BK_SWAP: m = (1u << o.p0) | (1u << o.p1); break;
    case BK_DM: m = o.lmask; break;
    case BK_PERM1: m = (1u << o.p0) | o.lmask; break;
    default: return 0;
    }
    return __builtin_popcount(m) <= 3 ? m : 0u;
}

// Applies one op to the 2^np amplitudes of positions pos[0..np) (local bit i
// <-> position pos[i]), following the kernel's meaning of each op kind, in a
// tile whose outer qubits meet the op's outer condition (outer_true) or not.
void apply_local(const StepH &s, const BOpD &o, const uint32_t *pos, uint32_t np, cd *v, bool outer_true) {
    if (o.omask && !outer_true) return;  // the tile's outer qubits fail the op's condition
    static const cd ip[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
    auto lb = [&](uint32_t p) -> uint32_t {
        for (uint32_t i = 0; i < np; i++)
            if (pos[i] == p) return 1u << i;
        fail(NCSV_INVALID, "batched planner: op outside its block");
        return 0;
    };
    auto lmask = [&](uint32_t m) {
        uint32_t r = 0;
        for (; m; m &= m - 1) r |= lb((uint32_t)__builtin_ctz(m));
        return r;
    };
    const uint32_t NA = 1u << np;
    switch (o.kind) {
    case BK_U1:
    case BK_U1R:
    case BK_U1X: {
        const uint32_t b = lb(o.p0);
        const cd *m = &s.coef[o.coef];
        for (uint32_t x = 0; x < NA; x++) {
            if (x & b) continue;
            const cd a = v[x], c = v[x | b];
            v[x] = m[0] * a + m[1] * c;
            v[x | b] = m[2] * a + m[3] * c;
        }
        break;
    }
    case BK_U2: {
        const uint32_t b0 = lb(o.p0), b1 = lb(o.p1);
        const cd *m = &s.coef[o.coef];
        for (uint32_t x = 0; x < NA; x++) {
            if (x & (b0 | b1)) continue;
            const uint32_t xs[4] = {x, x | b0, x | b1, x | b0 | b1};
            cd w[4];
            for (int r = 0; r < 4; r++) w[r] = v[xs[r]];
            for (int r = 0; r < 4; r++) {
                cd acc = 0;
                for (int c = 0; c < 4; c++) acc += m[4 * r + c] * w[c];
                v[xs[r]] = acc;
            }
        }
        break;
    }
    case BK_DM: {
        const uint32_t lm = lmask(o.lmask);
        for (uint32_t x = 0; x < NA; x++)
            if ((x & lm) == lm) v[x] *= s.coef[o.coef];
        break;
    }
    case BK_PERM1: {
        const uint32_t b = lb(o.p0), lm = lmask(o.lmask);
        for (uint32_t x = 0; x < NA; x++) {
            if ((x & b) || (x & lm) != lm) continue;
            const cd a = v[x], c = v[x | b];
            v[x | b] = ip[o.k0 & 3] * a;
            v[x] = ip[o.k1 & 3] * c;
        }
        break;
    }
    case BK_SWAP: {
        const uint32_t b0 = lb(o.p0), b1 = lb(o.p1);
        for (uint32_t x = 0; x < NA; x++)
            if ((x & b0) && !(x & b1)) std::swap(v[x], v[x ^ b0 ^ b1]);
        break;
    }
    default: fail(NCSV_INVALID, "batched planner: op cannot join a block");
    }
}

bool block_fusion_enabled() {
    const char *e = getenv("NCSV_BATCH_MFMA");
    return !(e && e[0] == '0');
}

// Complex128 dense blocks on two positions run as 4x4s on the VALU unless
// NCSV_BATCH_U2=0 (then as 8x8s with an identity factor, on the MFMA).
bool u2_blocks() {
    const char *e = getenv("NCSV_BATCH_U2");
    return !(e && e[0] == '0');
}

// Ops fused into dense 8x8 blocks on three tile positions (BK_U3: FP32 MFMA
// in complex64 programs, FP64 MFMA in complex128 ones).  Each position belongs to at most one open
// block; an op whose positions fit one block (together with the blocks it
// touches) joins it, anything else closes the blocks it touches first, so a
// block is emitted at its close with every op it commutes past acting on other
// positions.  A closed block becomes one BK_U3 when it has two or more
// positions, three or more ops and one op on two or more positions (a run of
// one-position ops is cheaper as grouped 2x2s; in complex128 a block on two
// positions becomes one BK_U2 4x4); otherwise its ops stay as they were.
std::vector<BOpD> fuse_blocks(StepH *s, const std::vector<BOpD> &in, bool fp32_copy) {
    struct Blk {
        uint32_t mask;
        std::vector<BOpD> ops;
        uint64_t omask = 0;  // the one outer condition of its ops (0: none)
    };
    // blocks may join when their outer conditions agree or one has none
    auto agree = [](uint64_t a, uint64_t b) { return !a || !b || a == b; };
    std::vector<BOpD> out;
    std::vector<Blk> open;
    // the last emitted block while nothing was emitted after it: its ops fill
    // out[start..), as one dense block or as they were
    struct Prev {
        bool valid = false, dense = false;
        size_t start = 0;
        Blk b{0, {}};
    } prev;
    auto worthy = [&](const Blk &b) {
        bool multi = false;
        for (const BOpD &o : b.ops) multi |= __builtin_popcount(block_support(o)) >= 2;
        return __builtin_popcount(b.mask) >= 2 && b.ops.size() >= 3 && multi && s->L >= 3;
    };
    auto emit = [&](const Blk &b0) {
        Blk b = b0;
        bool dense = worthy(b);
        if (prev.valid && !(prev.b.mask & b.mask) && __builtin_popcount(prev.b.mask | b.mask) <= 3 &&
            agree(prev.b.omask, b.omask) && (prev.dense || dense)) {
            // the previous block acts on other positions and nothing ran in
            // between: one dense block for both (they commute)
            b.ops.insert(b.ops.begin(), prev.b.ops.begin(), prev.b.ops.end());
            b.mask |= prev.b.mask;
            b.omask |= prev.b.omask;
            out.resize(prev.start);
            dense = true;
        }
        prev.valid = true;
        prev.dense = dense;
        prev.start = out.size();
        prev.b = b;
        if (!dense && __builtin_popcount(b.mask) == 1 && b.omask && b.ops.size() >= 2) {
            // ops on one position, some conditioned on outer qubits: one 2x2
            // per outcome of the condition, in one pass
            const uint32_t p = (uint32_t)__builtin_ctz(b.mask);
            cd M[8];
            for (uint32_t var = 0; var < 2; var++)
                for (uint32_t c = 0; c < 2; c++) {
                    cd v[2] = {};
                    v[c] = 1;
                    for (const BOpD &o : b.ops) apply_local(*s, o, &p, 1, v, var == 1);
                    M[4 * var + c] = v[0];
                    M[4 * var + 2 + c] = v[1];
                }
            BOpD d{};
            d.kind = BK_U1;
            d.p0 = (uint8_t)p;
            d.omask = b.omask;
            d.k0 = M[0] == cd(1, 0) && M[1] == cd(0, 0) && M[2] == cd(0, 0) && M[3] == cd(1, 0);
            d.coef = (uint32_t)s->coef.size();
            s->coef.insert(s->coef.end(), M, M + 8);
            out.push_back(d);
            return;
        }
        if (!dense) {
            out.insert(out.end(), b.ops.begin(), b.ops.end());
            return;
        }
        uint32_t pos[3], k = 0;
        for (uint32_t m = b.mask; m; m &= m - 1) pos[k++] = (uint32_t)__builtin_ctz(m);
        if (k == 2 && !b.omask && !fp32_copy && u2_blocks()) {
            // complex128: a block on two positions as one 4x4 on the VALU, half
            // the arithmetic of the 8x8 with an identity factor (FP64 MFMA runs
            // at the FP64 VALU rate on this part; 21-22-qubit circuits of the
            // mix: 674.5 -> 664.2 ms)
            cd M[16];
            for (uint32_t c = 0; c < 4; c++) {
                cd v[4] = {};
                v[c] = 1;
                for (const BOpD &o : b.ops) apply_local(*s, o, pos, 2, v, false);
                for (uint32_t r = 0; r < 4; r++) M[4 * r + c] = v[r];
            }
            BOpD d{};
            d.kind = BK_U2;
            d.p0 = (uint8_t)pos[0];
            d.p1 = (uint8_t)pos[1];
            d.coef = (uint32_t)s->coef.size();
            s->coef.insert(s->coef.end(), M, M + 16);
            out.push_back(d);
            return;
        }
        for (uint32_t p = 0; k < 3; p++)
            if (!((b.mask >> p) & 1)) pos[k++] = p;  // identity on an unused position
        BOpD d{};
        d.kind = BK_U3;
        d.p0 = (uint8_t)pos[0];
        d.p1 = (uint8_t)pos[1];
        d.p2 = (uint8_t)pos[2];
        d.omask = b.omask;
        // one matrix, or two when some ops are conditioned on outer qubits
        // (variant 1 for tiles whose outer qubits meet the condition)
        const uint32_t nvar = b.omask ? 2 : 1;
        std::vector<cd> M(64 * nvar);
        for (uint32_t var = 0; var < nvar; var++)
            for (uint32_t c = 0; c < 8; c++) {
                cd v[8] = {};
                v[c] = 1;
                for (const BOpD &o : b.ops) apply_local(*s, o, pos, 3, v, var == 1);
                for (uint32_t r = 0; r < 8; r++) M[64 * var + 8 * r + c] = v[r];
            }
        d.coef = (uint32_t)s->coef.size();
        s->coef.insert(s->coef.end(), M.begin(), M.end());
        // the kernel's A operands per variant, in lane order: lane l (i = l % 16,
        // q = l / 16) holds the real-form entries of row ai (U[ai][2q], U[ai][2q + 1])
        // that its four MFMAs take, in the operand's precision (complex64: one
        // 16-byte slot per lane; complex128: two), so the kernel loads them as they
        // are and its first MFMA is the first use of the load
        d.a = (uint32_t)s->coef.size();
        for (uint32_t var = 0; var < nvar; var++)
            for (uint32_t l = 0; l < 64; l++) {
                const uint32_t i = l & 15, q = l >> 4;
                const uint32_t ai = fp32_copy ? 2 * (i >> 2) + ((i >> 1) & 1) : 2 * (i & 3) + (i >> 3);
                const bool ci = fp32_copy ? (i & 1) != 0 : ((i >> 2) & 1) != 0;
                const cd u0 = M[64 * var + 8 * ai + 2 * q], u1 = M[64 * var + 8 * ai + 2 * q + 1];
                const double av[4] = {ci ? u0.imag() : u0.real(), ci ? u0.real() : -u0.imag(),
                                      ci ? u1.imag() : u1.real(), ci ? u1.real() : -u1.imag()};
                if (fp32_copy) {
                    const float f[4] = {(float)av[0], (float)av[1], (float)av[2], (float)av[3]};
                    cd slot;
                    static_assert(sizeof(slot) == sizeof(f), "four floats per complex slot");
                    memcpy((void *)&slot, f, sizeof(f));
                    s->coef.push_back(slot);
                } else {
                    s->coef.push_back(cd(av[0], av[1]));
                    s->coef.push_back(cd(av[2], av[3]));
                }
            }
        out.push_back(d);
    };
    auto close_if = [&](uint32_t m) {
        for (size_t i = 0; i < open.size();) {
            if (open[i].mask & m) {
                emit(open[i]);
                open.erase(open.begin() + (long)i);
            } else {
                i++;
            }
        }
    };
    for (const BOpD &o : in) {
        const uint32_t sup = block_support(o);
        if (!sup) {
            // outer-conditioned ops and slots act on their own positions; the
            // rest (measurement groups, expectations) on the whole tile
            uint32_t m = ~0u;
            if (o.kind == BK_NSLOT) m = 1u << o.p0;
            else if (o.kind <= BK_SWAP) m = (1u << o.p0) | (1u << o.p1) | o.lmask;
            if (o.kind == BK_DM) m = o.lmask;
            close_if(m);
            out.push_back(o);
            prev.valid = false;
            continue;
        }
        uint32_t u = sup;
        uint64_t om = o.omask;
        bool ok = true;
        for (const Blk &b : open)
            if (b.mask & sup) {
                u |= b.mask;
                ok = ok && agree(om, b.omask);
                om |= b.omask;
            }
        if (ok && __builtin_popcount(u) <= 3) {
            Blk nb{u, {}, om};
            for (size_t i = 0; i < open.size();) {
                if (open[i].mask & sup) {
                    nb.ops.insert(nb.ops.end(), open[i].ops.begin(), open[i].ops.end());
                    open.erase(open.begin() + (long)i);
                } else {
                    i++;
                }
            }
            nb.ops.push_back(o);
            open.push_back(std::move(nb));
        } else {
            close_if(sup);
            open.push_back(Blk{sup, {o}, o.omask});
        }
    }
    for (const Blk &b : open) emit(b);
    return out;
}

// Runs of unconditional one-position ops become one pass over up to three
// distinct positions (a later op on a position already in the run is
// multiplied into its 2x2; the run's positions commute).  Noise slots are
// kept, or dropped for the variant run by entries with no error here.
std::vector<BOpD> group_ops(StepH *s, const std::vector<BOpD> &in, bool drop_slots) {
    std::vector<BOpD> out;
    std::vector<uint32_t> pos;
    std::vector<std::array<cd, 4>> mats;
    std::vector<BOpD> orig;
    auto flush = [&]() {
        if (pos.size() >= 2) {
            BOpD d{};
            d.kind = BK_G3;
            d.k1 = (uint8_t)pos.size();
            d.p0 = (uint8_t)pos[0];
            d.p1 = (uint8_t)pos[1];
            d.p2 = (uint8_t)(pos.size() > 2 ? pos[2] : 0);
            d.coef = (uint32_t)s->coef.size();
            for (auto &m : mats)
                for (int t = 0; t < 4; t++) s->coef.push_back(m[t]);
            out.push_back(d);
        } else {
            out.insert(out.end(), orig.begin(), orig.end());
        }
        pos.clear();
        mats.clear();
        orig.clear();
    };
    for (const BOpD &o : in) {
        if (drop_slots && o.kind == BK_NSLOT) continue;
        uint32_t p;
        cd m[4];
        if (!as_2x2(*s, o, &p, m)) {
            flush();
            out.push_back(o);
            continue;
        }
        const auto it = std::find(pos.begin(), pos.end(), p);
        if (it != pos.end()) {
            // later x earlier on the same position
            std::array<cd, 4> &e = mats[it - pos.begin()];
            const std::array<cd, 4> r = {m[0] * e[0] + m[1] * e[2], m[0] * e[1] + m[1] * e[3],
                                         m[2] * e[0] + m[3] * e[2], m[2] * e[1] + m[3] * e[3]};
            e = r;
            orig.clear();  // the run no longer equals its ops one by one
            if (pos.size() == 1) {
                // a single merged 2x2: keep it as a U1
                BOpD d{};
                d.kind = BK_U1;
                d.p0 = (uint8_t)p;
                d.coef = (uint32_t)s->coef.size();
                for (int t = 0; t < 4; t++) s->coef.push_back(e[t]);
                orig.push_back(d);
            }
            continue;
        }
        if (pos.size() == 3) flush();
        if (pos.size() == 1 && orig.empty()) {
            // the first position's op was a merged 2x2: restore its record
            BOpD d{};
            d.kind = BK_U1;
            d.p0 = (uint8_t)pos[0];
            d.coef = (uint32_t)s->coef.size();
            for (int t = 0; t < 4; t++) s->coef.push_back(mats[0][t]);
            orig.push_back(d);
        }
        pos.push_back(p);
        mats.push_back({m[0], m[1], m[2], m[3]});
        orig.push_back(o);
    }
    flush();
    return out;
}

}  // namespace

// Plans one circuit (plus its Paulis) for tiles of at most Lb qubits.
// one_error_budget: complex coefficients the program may spend on op lists
// keeping one noise slot each, planned per step while the budget lasts
// (worth their planning and table space when many entries share the
// program, as noise trajectories do; 0: none)
std::shared_ptr<Program> build_program(const ncsv_circuit *c, const std::vector<uint64_t> &X,
                                       const std::vector<uint64_t> &Z, const std::vector<uint32_t> &PH,
                                       ncsv_precision prec, uint32_t Lb, int low_bits, size_t one_error_budget) {
    auto P = std::make_shared<Program>();
    const uint32_t n = c->n;
    P->n = n;
    const uint32_t L = std::min(n, Lb);
    P->L = L;
    // low qubits fixed in every gate tile (256-byte runs), leaving room for a
    // two-qubit gate plus a reset qubit on small tiles
    uint32_t cl = std::min(low_bits >= 0 ? (uint32_t)low_bits : (prec == NCSV_C64 ? 5u : 4u), L > 3 ? L - 3 : 0u);

    // the circuit as gate runs and measurement events (ticks and annotations dropped)
    std::vector<Ev> evs;
    auto run = [&]() -> std::vector<LOp> & {
        if (evs.empty() || evs.back().meas) evs.emplace_back();
        return evs.back().ops;
    };
    for (const COp &op : c->ops) {
        switch (op.kind) {
        case CKind::Gate: {
            LOp lo = lower_gate(op.g, op.qubits.data(), (uint32_t)op.qubits.size(), n);
            P->gates++;
            if (lo.kind == OpKind::Pauli)
                expand_pauli(lo, &run());
            else if (lo.kind != OpKind::Nop)
                run().push_back(std::move(lo));
            break;
        }
        case CKind::Noise: {
            NoiseRef nr;
            nr.ch = op.ch;
            memcpy(nr.p, op.p, sizeof nr.p);
            nr.slot0 = P->nslots;
            nr.k = (uint32_t)op.qubits.size();
            for (uint32_t j = 0; j < nr.k; j++) {
                LOp s;
                s.kind = OpKind::Slot;
                s.qubits = {op.qubits[j]};
                s.ph = P->nslots + j;
                run().push_back(s);
            }
            P->nslots += nr.k;
            P->draw_kind.push_back(0);
            P->draw_ref.push_back((uint32_t)P->noise.size());
            P->noise.push_back(nr);
            break;
        }
        case CKind::Measure:
        case CKind::Reset: {
            Ev e;
            e.meas = true;
            if (op.kind == CKind::Reset) {
                e.reset = true;
                e.q = op.qubits[0];
                e.z = 1ull << e.q;
            } else {
                if (op.x) fail(NCSV_UNSUPPORTED, "batched form: measurements of Paulis with X or Y factors");
                if (op.ph & 1) fail(NCSV_INVALID, "measured operator is not Hermitian");
                e.z = op.z;
                e.ph = op.ph & 3;
            }
            e.ord = P->nmeas++;
            P->draw_kind.push_back(e.reset ? 2 : 1);
            P->draw_ref.push_back(e.ord);
            evs.push_back(std::move(e));
            break;
        }
        default: break;
        }
    }

    // Pauli groups: up to 9 per EXPECT op, x-supports (plus the low qubits) within a tile
    std::vector<BPauliD> dummy;
    struct Pg {
        std::vector<uint32_t> idx;  // into the Pauli list (npauli = identity)
        uint64_t xs = 0;
    };
    const uint32_t np = (uint32_t)X.size();
    std::vector<Pg> groups;
    {
        for (uint32_t k = 0; k <= np; k++) {
            const uint64_t xs = k < np ? X[k] : 0;
            bool placed = false;
            for (Pg &g : groups) {
                if (g.idx.size() >= 9) continue;
                if ((uint32_t)__builtin_popcountll(g.xs | xs) > L) continue;
                g.idx.push_back(k);
                g.xs |= xs;
                placed = true;
                break;
            }
            if (!placed) {
                if ((uint32_t)__builtin_popcountll(xs) > L)
                    fail(NCSV_UNSUPPORTED, "batched form: a Pauli's X/Y support does not fit a tile");
                Pg g;
                g.idx.push_back(k);
                g.xs = xs;
                groups.push_back(g);
            }
        }
    }
    P->exps.resize(np + 1);
    auto add_expect = [&](StepH &s, const Pg &g, uint32_t tiles) {
        BOpD d{};
        d.kind = BK_EXPECT;
        d.a = (uint32_t)P->nexp;
        d.b = (uint32_t)g.idx.size();
        d.c = (uint32_t)s.paulis.size();
        for (uint32_t kk = 0; kk < g.idx.size(); kk++) {
            const uint32_t k = g.idx[kk];
            BPauliD pd{};
            if (k < np) {
                uint32_t xl = 0, zl = 0;
                uint64_t xo = 0, zo = 0;
                split_mask(s, X[k], &xl, &xo);
                split_mask(s, Z[k], &zl, &zo);
                if (xo) fail(NCSV_INVALID, "batched planner: Pauli X support outside the tile");
                pd.xl = xl;
                pd.zl = zl;
                pd.zo = zo;
                pd.ph = PH[k] & 3;
            }
            s.paulis.push_back(pd);
            P->exps[k] = ExpRef{P->nexp, (uint32_t)g.idx.size(), kk, tiles};
        }
        P->nexp += (uint64_t)2 * g.idx.size() * tiles;
        s.ops.push_back(d);
    };

    // gate runs and measurement groups
    struct Seg {
        bool meas = false;
        std::vector<LOp> ops;
        MGroup grp;
    };
    std::vector<Seg> segs;
    for (Ev &e : evs) {
        if (!e.meas) {
            if (e.ops.empty()) continue;
            Seg g;
            g.ops = std::move(e.ops);
            segs.push_back(std::move(g));
            continue;
        }
        if (segs.empty() || !segs.back().meas || !segs.back().grp.accepts(e)) {
            Seg g;
            g.meas = true;
            segs.push_back(std::move(g));
        }
        segs.back().grp.add(e);
    }

    if (n <= Lb) {
        // one step on one tile: the state stays in LDS
        StepH s;
        s.L = n;
        for (uint32_t q = 0; q < n; q++) s.qt.push_back((uint8_t)q);
        s.flags = BF_INIT;
        s.modifies = true;
        for (const Seg &g : segs) {
            if (!g.meas) {
                for (const SweepH &sw : plan_lds_sweeps(g.ops, n, n, cl, 0)) {
                    for (uint32_t p = 0; p < n; p++)
                        if (sw.tq[p] != p) fail(NCSV_INVALID, "batched planner: whole-state tile not in order");
                    append_tile_ops(&s, sw);
                }
            } else {
                BOpD d{};
                d.kind = BK_MEASURE;
                d.c = group_desc(&s, g.grp);
                s.ops.push_back(d);
            }
        }
        P->steps.push_back(std::move(s));
        for (const Pg &g : groups) add_expect(P->steps.back(), g, 1);
    } else {
        // a group reduced at the end of one step is collapsed at the start of the next
        struct Pend {
            bool on = false;
            const MGroup *grp = nullptr;
            uint32_t off = 0, tiles = 0;
        } pend;
        const uint32_t tiles = 1u << (n - L);
        // the needed qubits, filled up from qubit 0 (the low qubits keep loads coalesced)
        auto default_tq = [&](uint64_t need) {
            uint64_t T = need;
            for (uint32_t q = 0; q < n && (uint32_t)__builtin_popcountll(T) < L; q++) T |= 1ull << q;
            std::vector<uint32_t> tq;
            for (uint64_t a = T; a; a &= a - 1) tq.push_back((uint32_t)__builtin_ctzll(a));
            return tq;
        };
        auto pend_resets = [&]() { return pend.on ? pend.grp->resets : 0ull; };
        auto make_step = [&](const std::vector<uint32_t> &tq) -> StepH & {
            StepH s;
            s.L = L;
            uint64_t in = 0;
            for (uint32_t q : tq) {
                s.qt.push_back((uint8_t)q);
                in |= 1ull << q;
            }
            for (uint32_t q = 0; q < n; q++)
                if (!((in >> q) & 1)) s.qt.push_back((uint8_t)q);
            if (s.qt.size() != n || tq.size() != L) fail(NCSV_INVALID, "batched planner: bad tile");
            s.flags = P->steps.empty() ? BF_INIT : BF_LOAD;
            s.modifies = P->steps.empty();
            if (pend.on) {
                BOpD d{};
                d.kind = BK_COLLAPSE;
                d.a = pend.off;
                d.b = pend.tiles;
                d.c = group_desc(&s, *pend.grp);
                s.ops.push_back(d);
                s.modifies = true;
                pend.on = false;
            }
            P->steps.push_back(std::move(s));
            return P->steps.back();
        };
        for (const Seg &g : segs) {
            if (!g.meas) {
                for (const SweepH &sw : plan_lds_sweeps(g.ops, n, L, cl, pend_resets())) {
                    StepH &s = make_step(sw.tq);
                    append_tile_ops(&s, sw);
                    s.modifies = true;
                }
            } else {
                if ((uint32_t)__builtin_popcountll(g.grp.resets) > L)
                    fail(NCSV_UNSUPPORTED, "batched form: more reset qubits in one run than tile qubits");
                if (P->steps.empty() || pend.on) make_step(default_tq(pend_resets()));
                StepH &s = P->steps.back();
                BOpD d{};
                d.kind = BK_REDUCE;
                d.a = (uint32_t)P->npart;
                d.c = group_desc(&s, g.grp, false);
                s.ops.push_back(d);
                pend.on = true;
                pend.grp = &g.grp;
                pend.off = (uint32_t)P->npart;
                pend.tiles = tiles;
                P->npart += (uint64_t)(1u << g.grp.masks.size()) * tiles;
            }
        }
        auto fits_last = [&](uint64_t xs) {
            if (pend.on || P->steps.empty()) return false;
            for (uint64_t a = xs; a; a &= a - 1)
                if (pos_in(P->steps.back(), (uint32_t)__builtin_ctzll(a)) < 0) return false;
            return true;
        };
        for (const Pg &g : groups) {
            if (!fits_last(g.xs)) {
                uint64_t need = g.xs;
                if (pend.on && pend.grp->resets) {
                    const uint64_t rq = pend.grp->resets;
                    if ((uint32_t)__builtin_popcountll(need | rq) > L)
                        make_step(default_tq(rq));  // the resets' collapse on its own
                    else
                        need |= rq;
                }
                if (!fits_last(g.xs)) make_step(default_tq(need));
            }
            add_expect(P->steps.back(), g, tiles);
        }
        for (size_t i = 0; i + 1 < P->steps.size(); i++)
            if (P->steps[i].modifies) P->steps[i].flags |= BF_STORE;
    }
    const bool fuse = block_fusion_enabled();
    const bool f32 = prec == NCSV_C64;
    for (StepH &st : P->steps) {
        auto plan = [&](const std::vector<BOpD> &all, bool drop_slots) {
            // phases on outer qubits only multiply whole tiles and commute with
            // every op of the step: they leave the op passes and are taken at
            // the store, as one phase per tile
            std::vector<BOpD> ops, tph;
            for (BOpD o : all) {
                if (o.kind == BK_DM && !o.lmask && o.omask) {
                    o.kind = BK_OPHASE;
                    tph.push_back(o);
                } else {
                    ops.push_back(o);
                }
            }
            std::vector<BOpD> out;
            if (!fuse) {
                out = group_ops(&st, ops, drop_slots);
            } else if (!drop_slots) {
                out = group_ops(&st, fuse_blocks(&st, ops, f32), false);
            } else {
                std::vector<BOpD> clean;
                for (const BOpD &o : ops)
                    if (o.kind != BK_NSLOT) clean.push_back(o);
                out = group_ops(&st, fuse_blocks(&st, clean, f32), true);
            }
            out.insert(out.end(), tph.begin(), tph.end());
            return out;
        };
        st.gops = plan(st.ops, false);
        st.gopsc = plan(st.ops, true);
        // one list per noise slot keeping only that slot: a step where one
        // error was drawn runs it, with the other slots' sites fused over
        if (one_error_budget && st.slots.size() >= 2 && st.slots.size() <= 256) {
            const size_t before = st.coef.size();
            st.gops1.resize(st.slots.size());
            for (size_t k = 0; k < st.slots.size(); k++) {
                std::vector<BOpD> one;
                for (const BOpD &o : st.ops)
                    if (o.kind != BK_NSLOT || o.a == st.slots[k]) one.push_back(o);
                st.gops1[k] = plan(one, false);
            }
            if (st.coef.size() - before > one_error_budget) {
                st.gops1.clear();
                st.coef.resize(before);
            } else {
                one_error_budget -= st.coef.size() - before;
            }
        }
    }
    if (getenv("NCSV_BATCH_DUMP")) {
        // programs are planned on several threads: each one's report is
        // assembled first and written in one piece
        static const char *nm[] = {"U1", "U1R", "U1X", "U2", "DM", "PERM1", "SWAP", "NSLOT",
                                   "MEAS", "RED", "COLL", "EXP", "G3", "U3", "OPH"};
        std::string txt;
        char buf[256];
        for (size_t i = 0; i < P->steps.size(); i++) {
            const StepH &st = P->steps[i];
            int cnt[15] = {0};
            for (const BOpD &o : st.gopsc) cnt[o.kind]++;
            snprintf(buf, sizeof buf, "step %zu L=%u flags=%u ops=%zu clean=%zu slots=%zu one-error-lists=%zu:", i,
                     st.L, st.flags, st.gops.size(), st.gopsc.size(), st.slots.size(), st.gops1.size());
            txt += buf;
            for (int k = 0; k < 15; k++)
                if (cnt[k]) {
                    snprintf(buf, sizeof buf, " %s=%d", nm[k], cnt[k]);
                    txt += buf;
                }
            txt += "\n";
            if (atoi(getenv("NCSV_BATCH_DUMP")) >= 2) {
                // the slot-free op list: kind, positions, in-tile and outer condition masks
                for (const BOpD &o : st.gopsc) {
                    snprintf(buf, sizeof buf, "  %s p=%u,%u,%u l=%x o=%llx\n", nm[o.kind], o.p0, o.p1, o.p2, o.lmask,
                             (unsigned long long)o.omask);
                    txt += buf;
                }
            }
        }
        static std::mutex mu;
        std::lock_guard<std::mutex> lk(mu);
        fputs(txt.c_str(), stderr);
    }
    for (uint32_t s = 0; s < P->steps.size(); s++) {
        P->step_lg.push_back((uint8_t)(n - P->steps[s].L));
        P->step_noisy.push_back(!P->steps[s].slots.empty());
        P->items += P->tiles(s);
        if (n > Lb) {
            const uint64_t N = 1ull << n;
            if (P->steps[s].flags & BF_LOAD) P->hbm_elems += N;
            if (P->steps[s].flags & BF_STORE) P->hbm_elems += N;
        }
    }
    return P;
}

}  // namespace batch
}  // namespace ncsv

using ncsv::cd;
using namespace ncsv::batch;

struct ncsv_batch {
    ncsv_precision prec = NCSV_C128;
    uint32_t tile_bits = 0;  // 0: the largest tile for the precision
    int blocks = 0;          // 0: every resident workgroup
    int low_bits = -1;       // fixed low tile qubits (-1: 5 for complex64, 4 for complex128)
    int items_lg = -1;       // log2 of the most tiles per queue item (-1: 5)
    struct Entry {
        const ncsv_circuit *c = nullptr;
        std::vector<uint64_t> x, z;
        std::vector<uint32_t> ph;
        int prog = -1;
        uint32_t ndraws = 0;
    };
    std::vector<Entry> entries;
    std::vector<std::shared_ptr<Program>> progs;
    uint32_t planned_bits = 0;
    int planned_low = -1;
    // device program tables
    void *d_tab = nullptr;
    size_t tab_bytes = 0;
    size_t off_steps = 0, off_ops = 0, off_coef = 0, off_paulis = 0, off_mgroups = 0, off_qtab = 0;
    bool tables_dirty = true;               // device copy out of date
    std::shared_ptr<void> flat;             // host tables of the current programs
    bool flat_dirty = true;
    // per-chunk device buffers (grow only)
    void *d_amp = nullptr;
    size_t amp_cap = 0;
    void *d_work = nullptr;
    size_t work_cap = 0;
    int resident = 0;               // resident workgroups of the build in use
    uint32_t resident_threads = 0;  // ... for this workgroup size
    // results of the last run
    std::vector<std::vector<double>> expv;
    std::vector<double> norm;
    std::vector<std::vector<uint8_t>> outc;
    double kernel_ms = 0.0, plan_ms = 0.0;
    uint64_t items = 0, launches = 0;
    double amp_updates = 0.0, hbm_bytes = 0.0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the launch (counter reset + persistent kernel) captured as a graph, replayed
    // while the chunk's device layout and arguments stay the same
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> graph_key;
    uint64_t graph_replays = 0;
    std::mutex mu;
};

namespace ncsv {
namespace batch {
namespace {

uint32_t count_draws(const ncsv_circuit *c) {
    uint32_t d = 0;
    for (const COp &op : c->ops)
        if (op.kind == CKind::Measure || op.kind == CKind::Reset || op.kind == CKind::Noise) d++;
    return d;
}

uint32_t tile_bits_of(const ncsv_batch *b) {
    const uint32_t mx = max_tile_bits(b->prec);
    if (b->tile_bits) return std::min(b->tile_bits, mx);
    if (const char *e = getenv("NCSV_BATCH_TILE_BITS")) {
        const int v = atoi(e);
        if (v >= 6) return std::min((uint32_t)v, mx);
    }
    return mx;
}

// Threads per workgroup of the persistent kernel for tiles of Lb qubits:
// 256 below the largest tile (half the LDS per workgroup, so twice as many
// workgroups stay resident and one's loads overlap another's ops), else 512.
// NCSV_BATCH_THREADS=512 keeps 512 for every tile width.
uint32_t workgroup_threads(ncsv_precision prec, uint32_t Lb) {
    const char *e = getenv("NCSV_BATCH_THREADS");
    if (e && atoi(e) == 512) return 512;
    return Lb < max_tile_bits(prec) ? 256u : 512u;
}

// Plans every entry's program (entries running the same circuit with the
// same Paulis share one), in parallel host threads.
void plan_all(ncsv_batch *b) {
    const uint32_t Lb = tile_bits_of(b);
    if (Lb != b->planned_bits || b->low_bits != b->planned_low) {
        b->planned_low = b->low_bits;
        for (auto &e : b->entries) e.prog = -1;
        b->progs.clear();
        b->planned_bits = Lb;
        b->tables_dirty = true;
        b->flat_dirty = true;
    }
    std::vector<size_t> todo;
    for (size_t i = 0; i < b->entries.size(); i++) {
        auto &e = b->entries[i];
        if (e.prog >= 0) continue;
        // share with an earlier entry of the same circuit and Paulis
        for (size_t j = 0; j < i; j++) {
            const auto &f = b->entries[j];
            if (f.c == e.c && f.x == e.x && f.z == e.z && f.ph == e.ph && (f.prog >= 0 || j < i)) {
                e.prog = f.prog >= 0 ? f.prog : -2 - (int)j;
                break;
            }
        }
        if (e.prog == -1) todo.push_back(i);
    }
    if (todo.empty() && std::all_of(b->entries.begin(), b->entries.end(), [](const ncsv_batch::Entry &e) {
            return e.prog >= 0;
        }))
        return;
    // entries sharing each program to be built: single-error lists pay only
    // for programs that 16 or more entries run; each such program may spend
    // 2^22 complex coefficients (64 MiB) on them, and all of them together
    // 2^24
    std::vector<uint32_t> sharers(todo.size(), 1);
    {
        std::unordered_map<size_t, size_t> at;
        for (size_t k = 0; k < todo.size(); k++) at[todo[k]] = k;
        for (const auto &e : b->entries)
            if (e.prog <= -2) {
                const auto it = at.find((size_t)(-2 - e.prog));
                if (it != at.end()) sharers[it->second]++;
            }
    }
    const size_t nshared = (size_t)std::count_if(sharers.begin(), sharers.end(), [](uint32_t v) { return v >= 16; });
    const size_t budget = nshared ? std::min<size_t>((size_t)1 << 22, ((size_t)1 << 24) / nshared) : 0;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::shared_ptr<Program>> made(todo.size());
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<size_t>(nt, todo.size());
    std::vector<std::future<void>> fs;
    std::atomic<size_t> next{0};
    for (unsigned t = 0; t < nt; t++) {
        fs.push_back(std::async(std::launch::async, [&]() {
            for (;;) {
                const size_t k = next.fetch_add(1);
                if (k >= todo.size()) return;
                const auto &e = b->entries[todo[k]];
                made[k] = build_program(e.c, e.x, e.z, e.ph, b->prec, Lb, b->low_bits, sharers[k] >= 16 ? budget : 0);
            }
        }));
    }
    for (auto &f : fs) f.get();  // rethrows a planning error
    for (size_t k = 0; k < todo.size(); k++) {
        b->entries[todo[k]].prog = (int)b->progs.size();
        b->progs.push_back(made[k]);
    }
    for (auto &e : b->entries)
        if (e.prog <= -2) e.prog = b->entries[(size_t)(-2 - e.prog)].prog;
    b->tables_dirty = true;
    b->flat_dirty = true;
    b->plan_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <typename T> T *at(void *base, size_t off) { return (T *)((char *)base + off); }
size_t up256(size_t v) { return (v + 255) / 256 * 256; }

// The programs' steps, ops, coefficients, Paulis and qubit tables, uploaded once.
// The slot (index in the step's slot list) of the one error drawn in this
// step when the step has single-slot lists and exactly one error; else -1.
int single_error(const StepH &s, const uint8_t *errs) {
    if (s.gops1.empty()) return -1;
    int k = -1;
    for (size_t i = 0; i < s.slots.size(); i++)
        if (errs[s.slots[i]]) {
            if (k >= 0) return -1;
            k = (int)i;
        }
    return k;
}

struct Tables {
    std::vector<BStepD> steps;
    std::vector<BOpD> ops;
    std::vector<double> coef;
    std::vector<BPauliD> paulis;
    std::vector<BMGroupD> mgroups;
    std::vector<uint8_t> qtab;
};

Tables flatten(ncsv_batch *b) {
    Tables t;
    for (auto &P : b->progs) {
        P->step0 = (uint32_t)t.steps.size();
        for (const StepH &s : P->steps) {
            BStepD d{};
            d.L = s.L;
            d.n = P->n;
            d.flags = s.flags;
            d.qoff = (uint32_t)t.qtab.size();
            for (uint32_t i = 0; i < s.qt.size(); i++) {
                if (i < s.L && i < 16) d.tpos[i] = s.qt[i];
                if (i >= s.L) d.omask |= 1u << s.qt[i];
            }
            if (!s.slots.empty()) {
                const auto mm = std::minmax_element(s.slots.begin(), s.slots.end());
                d.slot_lo = *mm.first;
                d.slot_n = *mm.second - *mm.first + 1;
            }
            const uint32_t c0 = (uint32_t)(t.coef.size() / 2), p0 = (uint32_t)t.paulis.size(),
                           g0 = (uint32_t)t.mgroups.size();
            auto put = [&](const std::vector<BOpD> &ops) {
                for (BOpD o : ops) {
                    if (o.kind <= BK_DM || o.kind == BK_G3 || o.kind == BK_U3 || o.kind == BK_OPHASE) o.coef += c0;
                    if (o.kind == BK_U3) o.a += c0;
                    if (o.kind == BK_EXPECT) o.c += p0;
                    if (o.kind == BK_MEASURE || o.kind == BK_REDUCE || o.kind == BK_COLLAPSE) o.c += g0;
                    t.ops.push_back(o);
                }
            };
            d.op0 = (uint32_t)t.ops.size();
            d.nops = (uint32_t)s.gops.size();
            put(s.gops);
            d.op0c = (uint32_t)t.ops.size();
            d.nopsc = (uint32_t)s.gopsc.size();
            put(s.gopsc);
            if (!s.gops1.empty()) {
                // headers first (never run as ops), then the lists
                d.var0 = (uint32_t)t.ops.size();
                t.ops.resize(t.ops.size() + s.gops1.size(), BOpD{});
                for (size_t k = 0; k < s.gops1.size(); k++) {
                    BOpD &h = t.ops[d.var0 + k];
                    h.kind = 0xff;
                    h.a = (uint32_t)t.ops.size();
                    h.b = (uint32_t)s.gops1[k].size();
                    put(s.gops1[k]);
                }
            }
            for (const cd &c : s.coef) {
                t.coef.push_back(c.real());
                t.coef.push_back(c.imag());
            }
            t.paulis.insert(t.paulis.end(), s.paulis.begin(), s.paulis.end());
            t.mgroups.insert(t.mgroups.end(), s.mgroups.begin(), s.mgroups.end());
            t.qtab.insert(t.qtab.end(), s.qt.begin(), s.qt.end());
            t.steps.push_back(d);
        }
    }
    return t;
}

void upload_tables(ncsv_batch *b, const Tables &t) {
    size_t off = 0;
    b->off_steps = off;
    off = up256(off + t.steps.size() * sizeof(BStepD));
    b->off_ops = off;
    off = up256(off + t.ops.size() * sizeof(BOpD));
    b->off_coef = off;
    off = up256(off + t.coef.size() * sizeof(double));
    b->off_paulis = off;
    off = up256(off + t.paulis.size() * sizeof(BPauliD));
    b->off_mgroups = off;
    off = up256(off + t.mgroups.size() * sizeof(BMGroupD));
    b->off_qtab = off;
    off = up256(off + t.qtab.size() + 1);
    if (off > b->tab_bytes) {
        if (b->d_tab) check_hip(hipFree(b->d_tab), "hipFree");
        b->d_tab = nullptr;
        check_hip(hipMalloc(&b->d_tab, off), "hipMalloc(batch programs)");
        b->tab_bytes = off;
    }
    std::vector<char> h(off, 0);
    memcpy(h.data() + b->off_steps, t.steps.data(), t.steps.size() * sizeof(BStepD));
    memcpy(h.data() + b->off_ops, t.ops.data(), t.ops.size() * sizeof(BOpD));
    memcpy(h.data() + b->off_coef, t.coef.data(), t.coef.size() * sizeof(double));
    memcpy(h.data() + b->off_paulis, t.paulis.data(), t.paulis.size() * sizeof(BPauliD));
    memcpy(h.data() + b->off_mgroups, t.mgroups.data(), t.mgroups.size() * sizeof(BMGroupD));
    memcpy(h.data() + b->off_qtab, t.qtab.data(), t.qtab.size());
    check_hip(hipMemcpy(b->d_tab, h.data(), off, hipMemcpyHostToDevice), "upload batch programs");
    b->tables_dirty = false;
}

// One chunk's host-side arrays.
struct Chunk {
    std::vector<uint32_t> ent;  // batch entry of each chunk entry
    std::vector<BEntryD> ed;
    std::vector<BRecD> recs;
    std::vector<uint32_t> rec_blk;
    uint64_t total = 0, amp_elems = 0, npart = 0, nexp = 0;
    uint32_t nmeas = 0, nslots = 0;
    std::vector<double> uni, lsc;
    std::vector<uint8_t> forced, errs, outc;
    bool any_forced = false;
    std::vector<double> pexp;
    // every entry's expectation sums (Paulis, then the norm; real and
    // imaginary parts), where their per-tile partials sit in pexp, and their
    // values before the entry's scale
    std::vector<BSumD> sums;
    std::vector<double> esum;
};

// The sums of per-tile partials on the host (the device's k_batch_finish
// sequence): in tile order, Neumaier-compensated.
void sum_partials(Chunk *ch) {
    ch->esum.resize(ch->sums.size());
    for (size_t i = 0; i < ch->sums.size(); i++) {
        const BSumD &r = ch->sums[i];
        double s = 0.0, comp = 0.0;
        for (uint32_t t = 0; t < r.tiles; t++) {
            const double x = ch->pexp[r.off + (uint64_t)r.stride * t];
            const double y = s + x;
            comp += std::fabs(s) >= std::fabs(x) ? (s - y) + x : (x - y) + s;
            s = y;
        }
        ch->esum[i] = s + comp;
    }
}

// Draws of one entry: uniforms from the caller's array or its SplitMix64
// seed, in program order; noise draws become per-slot error codes
// (Noise::sample, circuit.rs:136-146, with the error applied as raw x/z bits).
void fill_draws(const Program &P, const double *u, uint64_t *seed, const uint8_t *forced, Chunk *ch,
                const BEntryD &ed) {
    for (size_t d = 0; d < P.draw_kind.size(); d++) {
        double r = 0.0;
        if (u)
            r = u[d];
        else if (seed)
            r = splitmix_f64(seed);
        if (P.draw_kind[d] == 0) {
            const NoiseRef &nr = P.noise[P.draw_ref[d]];
            for (const NoiseErr &e : noise_errors(nr.ch, nr.p)) {
                if (r < e.p) {
                    for (uint32_t j = 0; j < nr.k; j++)
                        ch->errs[ed.err0 + nr.slot0 + j] = (uint8_t)(((e.x >> j) & 1) | (((e.z >> j) & 1) << 1));
                    break;
                }
                r -= e.p;
            }
        } else {
            const uint32_t o = ed.m0 + P.draw_ref[d];
            ch->uni[o] = r;
            if (forced && forced[d] != 0xff) {
                ch->forced[o] = forced[d] & 1;
                ch->any_forced = true;
            }
        }
    }
}

// spread: the fewest queue items a depth of the queue is cut into (0: no cap)
Chunk make_chunk(ncsv_batch *b, const std::vector<uint32_t> &ent, const double *uniforms,
                 const std::vector<size_t> &draw_off, const uint64_t *seeds, const uint8_t *forced,
                 uint64_t spread) {
    // tiles per queue item: fewer, longer items mean fewer device-scope releases
    uint32_t lg_tiles_per_item = b->items_lg >= 0 ? (uint32_t)b->items_lg : 5u;
    if (const char *e = getenv("NCSV_BATCH_TILES_PER_ITEM")) {
        const int v = atoi(e);
        lg_tiles_per_item = 0;
        while (v > 1 && (1 << (lg_tiles_per_item + 1)) <= v) lg_tiles_per_item++;
    }
    Chunk ch;
    ch.ent = ent;
    const size_t ne = ent.size();
    ch.ed.resize(ne);
    for (size_t i = 0; i < ne; i++) {
        const Program &P = *b->progs[b->entries[ent[i]].prog];
        BEntryD &d = ch.ed[i];
        d.amp_off = ch.amp_elems;
        if (P.n > P.L) ch.amp_elems += 1ull << P.n;
        d.part0 = ch.npart;
        ch.npart += P.npart;
        d.exp0 = ch.nexp;
        ch.nexp += P.nexp;
        d.m0 = ch.nmeas;
        ch.nmeas += P.nmeas;
        d.err0 = ch.nslots;
        ch.nslots += P.nslots;
    }
    for (size_t i = 0; i < ne; i++) {
        const auto &E = b->entries[ent[i]];
        const Program &P = *b->progs[E.prog];
        for (size_t k = 0; k <= E.x.size(); k++) {
            const ExpRef &r = P.exps[k];
            for (uint32_t c = 0; c < 2; c++)
                ch.sums.push_back(BSumD{ch.ed[i].exp0 + r.off + 2ull * r.k + c, 2 * r.K, r.tiles});
        }
    }
    ch.uni.assign(ch.nmeas + 1, 0.0);
    ch.forced.assign(ch.nmeas + 1, 0xff);
    ch.errs.assign(ch.nslots + 1, 0);
    ch.lsc.assign(ne, 1.0);
    for (size_t i = 0; i < ne; i++) {
        const uint32_t e = ent[i];
        const Program &P = *b->progs[b->entries[e].prog];
        uint64_t seed = seeds ? seeds[e] : 0;
        fill_draws(P, uniforms ? uniforms + draw_off[e] : nullptr, seeds ? &seed : nullptr,
                   forced ? forced + draw_off[e] : nullptr, &ch, ch.ed[i]);
    }
    // depth-major queue, deeper programs first within a depth
    std::vector<const Program *> prog(ne);
    std::vector<uint32_t> nst(ne), order(ne);
    size_t nrec = 0;
    for (uint32_t i = 0; i < ne; i++) {
        prog[i] = b->progs[b->entries[ent[i]].prog].get();
        nst[i] = (uint32_t)prog[i]->steps.size();
        nrec += nst[i];
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return nst[a] > nst[c]; });
    std::vector<uint32_t> last(ne, ~0u);
    ch.recs.reserve(nrec + 1);
    const uint32_t maxd = ne ? nst[order[0]] : 0;
    // depths where little work is left (the last steps of the deepest
    // programs) get shorter items, so that each depth still spreads over at
    // least `spread` queue items (twice the resident workgroups on the device:
    // 4096-circuit mix, complex128 766.7 -> 757.1 ms, complex64 447.2 -> 443.1).
    // Only in queues of at least 16 x spread items: a small batch (such as an
    // autotuning sample) is thin at every depth, and cutting all of its items
    // short changes how tile shapes compare on it.
    auto items_at = [&](uint32_t d, uint32_t cap) {
        uint64_t items = 0;
        for (uint32_t i : order) {
            if (d >= nst[i]) break;
            const uint32_t lg = prog[i]->step_lg[d];
            items += 1ull << (lg - std::min(lg > 3 ? lg - 3 : 0u, cap));
        }
        return items;
    };
    if (spread) {
        uint64_t total = 0;
        for (uint32_t d = 0; d < maxd; d++) total += items_at(d, lg_tiles_per_item);
        if (total < 16 * spread) spread = 0;
    }
    for (uint32_t d = 0; d < maxd; d++) {
        uint32_t cap = lg_tiles_per_item;
        if (spread) {
            for (; cap > 0; cap--)
                if (items_at(d, cap) >= spread) break;
        }
        for (uint32_t i : order) {
            const Program &P = *prog[i];
            if (d >= nst[i]) break;  // sorted by depth
            BRecD r{};
            r.first = ch.total;
            r.entry = i;
            r.step = P.step0 + d;
            r.prev = last[i];
            r.prev_tiles = d ? 1u << P.step_lg[d - 1] : 0;
            r.flags = d + 1 < nst[i] ? RF_RELEASE : 0u;
            if (P.step_noisy[d]) {
                uint32_t err = 0;
                for (uint32_t sl : P.steps[d].slots) err |= ch.errs[ch.ed[i].err0 + sl];
                if (err) r.flags |= RF_ERR;
                const int one = single_error(P.steps[d], &ch.errs[ch.ed[i].err0]);
                if (one >= 0) {
                    r.flags |= RF_ONE;
                    r.var = (uint32_t)one;
                }
            }
            // at least 8 items per step where the step has that many tiles
            const uint32_t lg = P.step_lg[d];
            r.lgt = std::min(lg > 3 ? lg - 3 : 0u, cap);
            last[i] = (uint32_t)ch.recs.size();
            ch.recs.push_back(r);
            ch.total += (1ull << lg) >> r.lgt;
        }
    }
    if (getenv("NCSV_BATCH_DUMP")) {
        size_t err = 0, one = 0;
        for (const BRecD &r : ch.recs) {
            err += (r.flags & RF_ERR) != 0;
            one += (r.flags & RF_ONE) != 0;
        }
        fprintf(stderr, "records=%zu with-errors=%zu one-error=%zu\n", ch.recs.size(), err, one);
    }
    BRecD sentinel{};
    sentinel.first = ch.total;
    sentinel.prev = ~0u;
    ch.recs.push_back(sentinel);
    ch.rec_blk.assign((ch.total + 15) / 16 + 1, 0);
    for (uint32_t r = 0; r + 1 < ch.recs.size(); r++)
        for (uint64_t blk = (ch.recs[r].first + 15) / 16; blk * 16 < ch.recs[r + 1].first; blk++) ch.rec_blk[blk] = r;
    return ch;
}

void ensure(void **p, size_t *cap, size_t bytes, const char *what) {
    if (bytes <= *cap) return;
    if (*p) check_hip(hipFree(*p), "hipFree");
    *p = nullptr;
    *cap = 0;
    check_hip(hipMalloc(p, bytes), what);
    *cap = bytes;
}

void run_chunk_device(ncsv_batch *b, Chunk &ch) {
    const size_t es = b->prec == NCSV_C128 ? 16 : 8;
    ensure(&b->d_amp, &b->amp_cap, std::max<size_t>(256, ch.amp_elems * es), "hipMalloc(batch states)");
    // work buffer layout
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = up256(off + bytes);
        return o;
    };
    const size_t o_ent = take(ch.ed.size() * sizeof(BEntryD)), o_rec = take(ch.recs.size() * sizeof(BRecD)),
                 o_blk = take(ch.rec_blk.size() * 4), o_part = take((ch.npart + 16) * 8),
                 o_exp = take((ch.nexp + 1) * 8), o_lsc = take(ch.lsc.size() * 8), o_uni = take(ch.uni.size() * 8),
                 o_forced = take(ch.forced.size()), o_outc = take(ch.nmeas + 1), o_errs = take(ch.errs.size()),
                 o_done = take(ch.recs.size() * 4), o_next = take(8), o_flag = take(4),
                 o_sums = take(ch.sums.size() * sizeof(BSumD)), o_esum = take(ch.sums.size() * 8 + 8);
    ensure(&b->d_work, &b->work_cap, off, "hipMalloc(batch work)");
    char *w = (char *)b->d_work;
    hipStream_t st = b->stream;
    auto up = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) check_hip(hipMemcpyAsync(w + o, src, bytes, hipMemcpyHostToDevice, st), "upload batch");
    };
    up(o_ent, ch.ed.data(), ch.ed.size() * sizeof(BEntryD));
    up(o_rec, ch.recs.data(), ch.recs.size() * sizeof(BRecD));
    up(o_blk, ch.rec_blk.data(), ch.rec_blk.size() * 4);
    up(o_lsc, ch.lsc.data(), ch.lsc.size() * 8);
    up(o_uni, ch.uni.data(), ch.uni.size() * 8);
    up(o_forced, ch.forced.data(), ch.forced.size());
    up(o_errs, ch.errs.data(), ch.errs.size());
    up(o_sums, ch.sums.data(), ch.sums.size() * sizeof(BSumD));
    BArgs A{};
    A.steps = at<const BStepD>(b->d_tab, b->off_steps);
    A.ops = at<const BOpD>(b->d_tab, b->off_ops);
    A.coef = at<const double>(b->d_tab, b->off_coef);
    A.paulis = at<const BPauliD>(b->d_tab, b->off_paulis);
    A.mgroups = at<const BMGroupD>(b->d_tab, b->off_mgroups);
    A.qtab = at<const uint8_t>(b->d_tab, b->off_qtab);
    A.entries = (const BEntryD *)(w + o_ent);
    A.recs = (const BRecD *)(w + o_rec);
    A.rec_blk = (const uint32_t *)(w + o_blk);
    A.total = ch.total;
    A.amp = b->d_amp;
    A.part = (double *)(w + o_part);
    A.pexp = (double *)(w + o_exp);
    A.lsc = (double *)(w + o_lsc);
    A.uni = (const double *)(w + o_uni);
    A.forced = ch.any_forced ? (const uint8_t *)(w + o_forced) : nullptr;
    A.outc = (uint8_t *)(w + o_outc);
    A.errs = (const uint8_t *)(w + o_errs);
    A.done = (uint32_t *)(w + o_done);
    A.next = (unsigned long long *)(w + o_next);
    A.flag = (uint32_t *)(w + o_flag);
    A.spin_limit = 1u << 22;  // about a second of waiting per item
    const uint32_t threads = workgroup_threads(b->prec, b->planned_bits);
    int blocks = b->blocks > 0 ? b->blocks : resident_blocks(b->prec, threads);
    if ((uint64_t)blocks > ch.total) blocks = (int)std::max<uint64_t>(1, ch.total);
    // the same arguments as the captured launch: replay it
    std::vector<char> key(sizeof(BArgs) + 3 * sizeof(uint64_t));
    {
        const uint64_t extra[3] = {(uint64_t)ch.recs.size(), (uint64_t)blocks, (uint64_t)b->prec | (uint64_t)threads << 8};
        memcpy(key.data(), &A, sizeof(BArgs));
        memcpy(key.data() + sizeof(BArgs), extra, sizeof extra);
    }
    const bool use_graph = !getenv("NCSV_BATCH_NO_GRAPH");
    if (use_graph && (!b->gexec || key != b->graph_key)) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->gexec = nullptr;
        b->graph = nullptr;
        check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch(A, (uint32_t)ch.recs.size(), b->prec, threads, blocks, st);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(st, &b->graph), "hipStreamEndCapture");
        check_hip(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        b->graph_key = key;
    } else if (use_graph) {
        b->graph_replays++;
    }
    check_hip(hipEventRecord(b->ev0, st), "hipEventRecord");
    if (use_graph)
        check_hip(hipGraphLaunch(b->gexec, st), "hipGraphLaunch");
    else
        launch(A, (uint32_t)ch.recs.size(), b->prec, threads, blocks, st);
    // the per-tile partials summed on the device: only the sums come back
    launch_finish(A.pexp, (const BSumD *)(w + o_sums), (uint32_t)ch.sums.size(), (double *)(w + o_esum), st);
    check_hip(hipEventRecord(b->ev1, st), "hipEventRecord");
    ch.esum.assign(ch.sums.size(), 0.0);
    ch.outc.assign(ch.nmeas + 1, 0xff);
    uint32_t flag = 0;
    check_hip(hipMemcpyAsync(ch.esum.data(), w + o_esum, ch.esum.size() * 8, hipMemcpyDeviceToHost, st),
              "download");
    check_hip(hipMemcpyAsync(ch.lsc.data(), w + o_lsc, ch.lsc.size() * 8, hipMemcpyDeviceToHost, st), "download");
    check_hip(hipMemcpyAsync(ch.outc.data(), w + o_outc, ch.outc.size(), hipMemcpyDeviceToHost, st), "download");
    check_hip(hipMemcpyAsync(&flag, w + o_flag, 4, hipMemcpyDeviceToHost, st), "download");
    check_hip(hipStreamSynchronize(st), "batch run");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime");
    b->kernel_ms += ms;
    b->launches++;
    b->items += ch.total;
    if (flag & 1) fail(NCSV_DEVICE, "batched run: a step waited too long for its predecessor");
}

// The kernel body itself on host memory (one simulated workgroup).
void run_chunk_sim(ncsv_batch *b, const Tables &t, Chunk &ch) {
    const size_t es = b->prec == NCSV_C128 ? 16 : 8;
    std::vector<double> amp(std::max<size_t>(2, ch.amp_elems * es / 8), 0.0);
    std::vector<double> part(ch.npart + 16, 0.0);
    ch.pexp.assign(ch.nexp + 1, 0.0);
    ch.outc.assign(ch.nmeas + 1, 0xff);
    std::vector<uint32_t> done(ch.recs.size(), 0);
    unsigned long long next = 0;
    uint32_t flag = 0;
    BArgs A{};
    A.steps = t.steps.data();
    A.ops = t.ops.data();
    A.coef = t.coef.data();
    A.paulis = t.paulis.data();
    A.mgroups = t.mgroups.data();
    A.qtab = t.qtab.data();
    A.entries = ch.ed.data();
    A.recs = ch.recs.data();
    A.rec_blk = ch.rec_blk.data();
    A.total = ch.total;
    A.amp = amp.data();
    A.part = part.data();
    A.pexp = ch.pexp.data();
    A.lsc = ch.lsc.data();
    A.uni = ch.uni.data();
    A.forced = ch.any_forced ? ch.forced.data() : nullptr;
    A.outc = ch.outc.data();
    A.errs = ch.errs.data();
    A.done = done.data();
    A.next = &next;
    A.flag = &flag;
    A.spin_limit = 1u << 20;
    simulate(A, b->prec, workgroup_threads(b->prec, b->planned_bits));
    b->launches++;
    b->items += ch.total;
    if (flag & 1) fail(NCSV_DEVICE, "batched simulation: a step waited too long for its predecessor");
}

// ---------------------------------------------------------------- emulator
// The kernel's op semantics on FP64 host amplitudes.
using cdv = std::vector<cd>;

cd ipow_c(uint32_t k, cd v) {
    switch (k & 3) {
    case 1: return cd(-v.imag(), v.real());
    case 2: return -v;
    case 3: return cd(v.imag(), -v.real());
    default: return v;
    }
}
uint32_t ins0h(uint32_t j, uint32_t p) {
    const uint32_t lo = j & ((1u << p) - 1);
    return ((j ^ lo) << 1) | lo;
}
uint32_t parity(uint64_t v) { return (uint32_t)__builtin_popcountll(v) & 1u; }

void emulate_chunk(ncsv_batch *b, const Tables &t, Chunk &ch) {
    ch.pexp.assign(ch.nexp + 1, 0.0);
    ch.outc.assign(ch.nmeas + 1, 0xff);
    std::vector<double> part(ch.npart + 16, 0.0);
    const bool c64 = b->prec == NCSV_C64;
    for (size_t i = 0; i < ch.ent.size(); i++) {
        const Program &P = *b->progs[b->entries[ch.ent[i]].prog];
        const BEntryD &en = ch.ed[i];
        const uint32_t n = P.n;
        cdv amp(size_t(1) << n, cd(0, 0));
        for (uint32_t si = 0; si < P.steps.size(); si++) {
            const BStepD &st = t.steps[P.step0 + si];
            const uint32_t L = st.L, NL = 1u << L;
            const uint8_t *qt = t.qtab.data() + st.qoff;
            for (uint64_t tileno = 0; tileno < (1ull << (n - L)); tileno++) {
                uint64_t base = 0;
                for (uint32_t bb = 0; bb + L < n; bb++)
                    if ((tileno >> bb) & 1) base |= 1ull << qt[L + bb];
                std::vector<uint64_t> addr(NL);
                for (uint32_t x = 0; x < NL; x++) {
                    uint64_t off = base;
                    for (uint32_t p = 0; p < L; p++)
                        if ((x >> p) & 1) off |= 1ull << qt[p];
                    addr[x] = off;
                }
                cdv tl(NL);
                for (uint32_t x = 0; x < NL; x++)
                    tl[x] = (st.flags & BF_INIT) ? cd((x == 0 && tileno == 0) ? 1.0 : 0.0, 0.0) : amp[addr[x]];
                uint32_t err = 0;
                for (uint32_t sl : P.steps[si].slots) err |= ch.errs[en.err0 + sl];
                uint32_t op0 = err ? st.op0 : st.op0c, op1 = op0 + (err ? st.nops : st.nopsc);
                const int one = single_error(P.steps[si], &ch.errs[en.err0]);
                if (one >= 0) {
                    op0 = t.ops[st.var0 + (uint32_t)one].a;
                    op1 = op0 + t.ops[st.var0 + (uint32_t)one].b;
                }
                for (uint32_t oi = op0; oi < op1; oi++) {
                    const BOpD &op = t.ops[oi];
                    auto C = [&](uint32_t k) { return cd(t.coef[2 * (op.coef + k)], t.coef[2 * (op.coef + k) + 1]); };
                    switch (op.kind) {
                    case BK_G3: {
                        const uint32_t ps[3] = {op.p0, op.p1, op.p2};
                        for (uint32_t i = 0; i < op.k1; i++)
                            for (uint32_t x0 = 0; x0 < NL; x0++) {
                                if ((x0 >> ps[i]) & 1) continue;
                                const uint32_t x1 = x0 | (1u << ps[i]);
                                const cd a = tl[x0], v = tl[x1];
                                tl[x0] = C(4 * i) * a + C(4 * i + 1) * v;
                                tl[x1] = C(4 * i + 2) * a + C(4 * i + 3) * v;
                            }
                        break;
                    }
                    case BK_U3: {
                        const uint32_t pb[3] = {1u << op.p0, 1u << op.p1, 1u << op.p2};
                        const uint32_t vo = (op.omask && (base & op.omask) == op.omask) ? 64 : 0;
                        for (uint32_t x = 0; x < NL; x++) {
                            if (x & (pb[0] | pb[1] | pb[2])) continue;
                            uint32_t xs[8];
                            cd v[8];
                            for (uint32_t a = 0; a < 8; a++) {
                                xs[a] = x | ((a & 1) ? pb[0] : 0) | ((a & 2) ? pb[1] : 0) | ((a & 4) ? pb[2] : 0);
                                v[a] = tl[xs[a]];
                            }
                            for (uint32_t r = 0; r < 8; r++) {
                                cd acc = 0;
                                for (uint32_t c = 0; c < 8; c++) acc += C(vo + 8 * r + c) * v[c];
                                tl[xs[r]] = acc;
                            }
                        }
                        break;
                    }
                    case BK_U1:
                    case BK_U1R:
                    case BK_U1X: {
                        // a U1 with an outer condition: the second 2x2 where it holds
                        const uint32_t vo = op.omask && (base & op.omask) == op.omask ? 4u : 0u;
                        if (op.omask && !vo && op.k0) break;
                        for (uint32_t j = 0; j < NL / 2; j++) {
                            const uint32_t x0 = ins0h(j, op.p0), x1 = x0 | (1u << op.p0);
                            const cd a = tl[x0], v = tl[x1];
                            tl[x0] = C(vo) * a + C(vo + 1) * v;
                            tl[x1] = C(vo + 2) * a + C(vo + 3) * v;
                        }
                        break;
                    }
                    case BK_U2:
                        for (uint32_t j = 0; j < NL / 4; j++) {
                            const uint32_t x = ins0h(ins0h(j, op.p0), op.p1);
                            const uint32_t xs[4] = {x, x | (1u << op.p0), x | (1u << op.p1),
                                                    x | (1u << op.p0) | (1u << op.p1)};
                            cd v[4], r[4];
                            for (int k = 0; k < 4; k++) v[k] = tl[xs[k]];
                            for (int rr = 0; rr < 4; rr++) {
                                r[rr] = 0;
                                for (int cc = 0; cc < 4; cc++) r[rr] += C(4 * rr + cc) * v[cc];
                            }
                            for (int k = 0; k < 4; k++) tl[xs[k]] = r[k];
                        }
                        break;
                    case BK_DM:
                    case BK_OPHASE:  // lmask 0: the whole tile
                        if ((base & op.omask) != op.omask) break;
                        for (uint32_t x = 0; x < NL; x++)
                            if ((x & op.lmask) == op.lmask) tl[x] *= C(0);
                        break;
                    case BK_PERM1:
                        if ((base & op.omask) != op.omask) break;
                        for (uint32_t x0 = 0; x0 < NL; x0++) {
                            if ((x0 >> op.p0) & 1 || (x0 & op.lmask) != op.lmask) continue;
                            const uint32_t x1 = x0 | (1u << op.p0);
                            const cd a = tl[x0], v = tl[x1];
                            tl[x1] = ipow_c(op.k0, a);
                            tl[x0] = ipow_c(op.k1, v);
                        }
                        break;
                    case BK_SWAP:
                        for (uint32_t x = 0; x < NL; x++)
                            if (((x >> op.p0) & 1) && !((x >> op.p1) & 1))
                                std::swap(tl[x], tl[x ^ (1u << op.p0) ^ (1u << op.p1)]);
                        break;
                    case BK_NSLOT: {
                        const uint32_t code = ch.errs[en.err0 + op.a];
                        if (!code) break;
                        for (uint32_t j = 0; j < NL / 2; j++) {
                            const uint32_t x0 = ins0h(j, op.p0), x1 = x0 | (1u << op.p0);
                            cd a = tl[x0], v = tl[x1];
                            if (code & 2) v = -v;
                            if (code & 1) std::swap(a, v);
                            tl[x0] = a;
                            tl[x1] = v;
                        }
                        break;
                    }
                    case BK_MEASURE:
                    case BK_REDUCE:
                    case BK_COLLAPSE: {
                        const BMGroupD &G = t.mgroups[op.c];
                        const uint32_t NC = 1u << G.D;
                        uint32_t ocell = 0;
                        for (uint32_t d = 0; d < G.D; d++) ocell |= parity(base & G.om[d]) << d;
                        auto cell = [&](uint32_t x) {
                            uint32_t c = ocell;
                            for (uint32_t d = 0; d < G.D; d++) c ^= parity(x & G.lm[d]) << d;
                            return c;
                        };
                        double W[16] = {0};
                        if (op.kind == BK_COLLAPSE) {
                            for (uint32_t tt = 0; tt < op.b; tt++)
                                for (uint32_t c = 0; c < NC; c++) W[c] += part[en.part0 + op.a + NC * tt + c];
                        } else {
                            for (uint32_t x = 0; x < NL; x++) W[cell(x)] += std::norm(tl[x]);
                        }
                        if (op.kind == BK_REDUCE) {
                            for (uint32_t c = 0; c < NC; c++) part[en.part0 + op.a + NC * tileno + c] = W[c];
                            break;
                        }
                        uint32_t allow = (1u << NC) - 1, flip = 0;
                        double scale = 1.0, lfac = 1.0;
                        for (uint32_t e = 0; e < G.E; e++) {
                            const auto &ev = G.ev[e];
                            double nn = 0, ee = 0;
                            for (uint32_t c = 0; c < NC; c++) {
                                if (!((allow >> c) & 1)) continue;
                                nn += W[c];
                                ee += (((c >> ev.d) & 1) ^ ev.sign) ? -W[c] : W[c];
                            }
                            const uint32_t ord = en.m0 + ev.ord;
                            // NaN (zero state) kept as by f64::clamp: a draw, outcome 0
                            double p_plus = (1.0 + ee / nn) / 2.0;
                            if (p_plus < 0.0) p_plus = 0.0;
                            if (p_plus > 1.0) p_plus = 1.0;
                            uint32_t o;
                            if (ch.forced[ord] != 0xff)
                                o = ch.forced[ord] & 1;
                            else if (p_plus > 1.0 - 1e-12)
                                o = 0;
                            else if (p_plus < 1e-12)
                                o = 1;
                            else
                                o = ch.uni[ord] >= p_plus ? 1 : 0;
                            for (uint32_t c = 0; c < NC; c++)
                                if ((((c >> ev.d) & 1) ^ ev.sign) != o) allow &= ~(1u << c);
                            const double kept = 0.5 * (nn + (o ? -ee : ee)) * scale * scale;
                            if (c64 && kept > 0.0 && kept < 0.25) {
                                scale /= std::sqrt(kept);
                                lfac *= std::sqrt(kept);
                            }
                            if (tileno == 0) ch.outc[ord] = (uint8_t)o;
                            if (ev.rpos != 0xff && o) flip ^= 1u << ev.rpos;
                        }
                        if (tileno == 0) ch.lsc[i] *= lfac;
                        for (uint32_t x = 0; x < NL; x++) tl[x] = ((allow >> cell(x)) & 1) ? tl[x] * scale : cd(0, 0);
                        if (flip) {
                            cdv nt(NL);
                            for (uint32_t x = 0; x < NL; x++) nt[x] = tl[x ^ flip];
                            tl.swap(nt);
                        }
                        break;
                    }
                    case BK_EXPECT: {
                        for (uint32_t k = 0; k < op.b; k++) {
                            const BPauliD &Pd = t.paulis[op.c + k];
                            cd s(0, 0);
                            for (uint32_t x = 0; x < NL; x++) {
                                const uint32_t pw = (Pd.ph + 2 * (parity(Pd.zl & x) ^ parity(Pd.zo & base))) & 3;
                                s += ipow_c(pw, std::conj(tl[x ^ Pd.xl]) * tl[x]);
                            }
                            ch.pexp[en.exp0 + op.a + 2ull * op.b * tileno + 2 * k] = s.real();
                            ch.pexp[en.exp0 + op.a + 2ull * op.b * tileno + 2 * k + 1] = s.imag();
                        }
                        break;
                    }
                    default: fail(NCSV_INVALID, "batched emulator: unknown op");
                    }
                }
                if (st.flags & BF_STORE || n == L)
                    for (uint32_t x = 0; x < NL; x++) amp[addr[x]] = tl[x];
            }
        }
    }
}

void collect(ncsv_batch *b, const Chunk &ch) {
    size_t si = 0;  // the entries' sums follow one another in ch.esum
    for (size_t i = 0; i < ch.ent.size(); i++) {
        const uint32_t e = ch.ent[i];
        const auto &E = b->entries[e];
        const Program &P = *b->progs[E.prog];
        const BEntryD &en = ch.ed[i];
        const double l2 = ch.lsc[i] * ch.lsc[i];
        const size_t np = E.x.size();
        std::vector<double> v(2 * (np + 1));
        for (size_t k = 0; k < 2 * (np + 1); k++) v[k] = ch.esum[si++] * l2;
        b->norm[e] = v[2 * np];
        v.resize(2 * np);
        b->expv[e] = std::move(v);
        std::vector<uint8_t> oc(P.draw_kind.size(), 0xff);
        for (size_t d = 0; d < oc.size(); d++)
            if (P.draw_kind[d]) oc[d] = ch.outc[en.m0 + P.draw_ref[d]];
        b->outc[e] = std::move(oc);
    }
}

// mode 0: device, 1: host emulator, 2: the kernel body on the host
void run_batch(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced, int mode) {
    const bool emulate = mode != 0;
    if (b->entries.empty()) return;
    for (const auto &e : b->entries)
        if (e.ndraws && !uniforms && !seeds) fail(NCSV_INVALID, "batched run: circuits with draws need uniforms or seeds");
    plan_all(b);
    if (!b->flat || b->flat_dirty) {
        b->flat = std::make_shared<Tables>(flatten(b));
        b->flat_dirty = false;
        b->tables_dirty = true;
    }
    const Tables &t = *std::static_pointer_cast<Tables>(b->flat);
    if (!emulate && b->tables_dirty) upload_tables(b, t);
    std::vector<size_t> draw_off(b->entries.size() + 1, 0);
    for (size_t e = 0; e < b->entries.size(); e++) draw_off[e + 1] = draw_off[e] + b->entries[e].ndraws;
    b->expv.assign(b->entries.size(), {});
    b->norm.assign(b->entries.size(), 0.0);
    b->outc.assign(b->entries.size(), {});
    b->kernel_ms = 0.0;
    b->items = 0;
    b->launches = 0;
    b->amp_updates = 0.0;
    b->hbm_bytes = 0.0;
    const size_t es = b->prec == NCSV_C128 ? 16 : 8;
    // chunks that fit the state pool
    size_t limit = (size_t)64 << 30;
    if (const char *e = getenv("NCSV_BATCH_MEM_GB")) limit = (size_t)(atof(e) * (double)(1ull << 30));
    if (!emulate) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr) limit = std::min(limit, fr / 10 * 6);
    }
    std::vector<uint32_t> cur;
    size_t cur_bytes = 0;
    // NCSV_BATCH_TIMING=1: the host's share of each launch on stderr
    static const bool timing = getenv("NCSV_BATCH_TIMING") != nullptr;
    uint64_t spread = 0;
    if (mode == 0) {
        const uint32_t threads = workgroup_threads(b->prec, b->planned_bits);
        if (b->resident_threads != threads) {
            b->resident = resident_blocks(b->prec, threads);
            b->resident_threads = threads;
        }
        spread = 2ull * (uint64_t)(b->blocks > 0 ? b->blocks : b->resident);
        if (const char *e = getenv("NCSV_BATCH_SPREAD")) spread = (uint64_t)atoll(e);
    }
    auto flush = [&]() {
        if (cur.empty()) return;
        const auto t0 = std::chrono::steady_clock::now();
        Chunk ch = make_chunk(b, cur, uniforms, draw_off, seeds, forced, spread);
        const auto t1 = std::chrono::steady_clock::now();
        if (mode == 1) {
            emulate_chunk(b, t, ch);
            sum_partials(&ch);
        } else if (mode == 2) {
            run_chunk_sim(b, t, ch);
            sum_partials(&ch);
        } else {
            run_chunk_device(b, ch);
        }
        const auto t2 = std::chrono::steady_clock::now();
        collect(b, ch);
        const auto t3 = std::chrono::steady_clock::now();
        if (timing) {
            auto ms = [](auto a, auto c) { return std::chrono::duration<double, std::milli>(c - a).count(); };
            fprintf(stderr, "batch launch: %zu entries, chunk %.2f ms, run %.2f ms (kernel %.2f ms), collect %.2f ms\n",
                    cur.size(), ms(t0, t1), ms(t1, t2), b->kernel_ms, ms(t2, t3));
        }
        cur.clear();
        cur_bytes = 0;
    };
    for (uint32_t e = 0; e < b->entries.size(); e++) {
        const Program &P = *b->progs[b->entries[e].prog];
        const size_t bytes = P.n > P.L ? (size_t(1) << P.n) * es : 0;
        if (bytes > limit) fail(NCSV_DEVICE, "batched run: one state exceeds the memory limit");
        if (cur_bytes + bytes > limit) flush();
        cur.push_back(e);
        cur_bytes += bytes;
        b->amp_updates += (double)P.gates * (double)(1ull << P.n);
        b->hbm_bytes += (double)P.hbm_elems * (double)es;
    }
    flush();
}

}  // namespace
}  // namespace batch
}  // namespace ncsv

// ---------------------------------------------------------------- C ABI
namespace ncsv {
ncsv_status guard_status(const std::function<void()> &f);
}

extern "C" {

ncsv_batch *ncsv_batch_new(ncsv_precision precision) {
    ncsv_batch *out = nullptr;
    const ncsv_status st = ncsv::guard_status([&]() {
        if (precision != NCSV_C128 && precision != NCSV_C64) ncsv::fail(NCSV_INVALID, "unknown precision");
        auto b = std::make_unique<ncsv_batch>();
        b->prec = precision;
        out = b.release();
    });
    return st == NCSV_OK ? out : nullptr;
}

void ncsv_batch_destroy(ncsv_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->d_tab) (void)hipFree(b->d_tab);
    if (b->d_amp) (void)hipFree(b->d_amp);
    if (b->d_work) (void)hipFree(b->d_work);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

ncsv_status ncsv_batch_add(ncsv_batch *b, const ncsv_circuit *c, uint32_t count, const uint64_t *x,
                           const uint64_t *z, const uint32_t *phase, uint32_t *index) {
    return ncsv::guard_status([&]() {
        if (!b || !c) ncsv::fail(NCSV_INVALID, "null batch or circuit");
        if (count && (!x || !z || !phase)) ncsv::fail(NCSV_INVALID, "null Pauli arrays");
        if (c->n == 0 || c->n > 32) ncsv::fail(NCSV_INVALID, "batched form: 1..32 qubits per circuit");
        const uint64_t full = c->n >= 64 ? ~0ull : (1ull << c->n) - 1;
        for (uint32_t k = 0; k < count; k++)
            if ((x[k] | z[k]) & ~full) ncsv::fail(NCSV_INVALID, "Pauli acts outside the circuit's qubits");
        std::lock_guard<std::mutex> lk(b->mu);
        ncsv_batch::Entry e;
        e.c = c;
        e.x.assign(x, x + count);
        e.z.assign(z, z + count);
        e.ph.assign(phase, phase + count);
        e.ndraws = ncsv::batch::count_draws(c);
        if (index) *index = (uint32_t)b->entries.size();
        b->entries.push_back(std::move(e));
    });
}

uint32_t ncsv_batch_size(const ncsv_batch *b) { return b ? (uint32_t)b->entries.size() : 0; }

ncsv_status ncsv_batch_draws(const ncsv_batch *b, uint32_t entry, uint32_t *count) {
    return ncsv::guard_status([&]() {
        if (!b || !count || entry >= b->entries.size()) ncsv::fail(NCSV_INVALID, "bad batch entry");
        *count = b->entries[entry].ndraws;
    });
}

ncsv_status ncsv_batch_set_tile_bits(ncsv_batch *b, uint32_t bits) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        if (bits && (bits < 6 || bits > ncsv::batch::max_tile_bits(b->prec)))
            ncsv::fail(NCSV_INVALID, "tile bits out of range");
        std::lock_guard<std::mutex> lk(b->mu);
        b->tile_bits = bits;
    });
}

ncsv_status ncsv_batch_set_shape(ncsv_batch *b, int32_t low_bits, int32_t tiles_per_item) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        if (low_bits > 8 || tiles_per_item > 1024) ncsv::fail(NCSV_INVALID, "shape out of range");
        std::lock_guard<std::mutex> lk(b->mu);
        b->low_bits = low_bits < 0 ? -1 : low_bits;
        int lg = -1;
        if (tiles_per_item > 0) {
            lg = 0;
            while ((2 << lg) <= tiles_per_item) lg++;
        }
        b->items_lg = lg;
    });
}

ncsv_status ncsv_batch_set_blocks(ncsv_batch *b, int32_t blocks) {
    return ncsv::guard_status([&]() {
        if (!b || blocks < 0) ncsv::fail(NCSV_INVALID, "bad batch or block count");
        std::lock_guard<std::mutex> lk(b->mu);
        b->blocks = blocks;
    });
}

ncsv_status ncsv_batch_run(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        std::lock_guard<std::mutex> lk(b->mu);
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) ncsv::fail(NCSV_DEVICE, "no GPU device visible");
        if (!b->stream) {
            ncsv::check_hip(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
            ncsv::check_hip(hipEventCreate(&b->ev0), "hipEventCreate");
            ncsv::check_hip(hipEventCreate(&b->ev1), "hipEventCreate");
        }
        ncsv::batch::run_batch(b, uniforms, seeds, forced, 0);
    });
}

ncsv_status ncsv_batch_plan(ncsv_batch *b) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        std::lock_guard<std::mutex> lk(b->mu);
        ncsv::batch::plan_all(b);
    });
}

ncsv_status ncsv_batch_emulate(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        std::lock_guard<std::mutex> lk(b->mu);
        ncsv::batch::run_batch(b, uniforms, seeds, forced, 1);
    });
}

ncsv_status ncsv_batch_simulate(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced) {
    return ncsv::guard_status([&]() {
        if (!b) ncsv::fail(NCSV_INVALID, "null batch");
        std::lock_guard<std::mutex> lk(b->mu);
        ncsv::batch::run_batch(b, uniforms, seeds, forced, 2);
    });
}

ncsv_status ncsv_batch_result(const ncsv_batch *b, uint32_t entry, double *expect, double *norm_sq,
                              uint8_t *outcomes) {
    return ncsv::guard_status([&]() {
        if (!b || entry >= b->entries.size() || entry >= b->expv.size()) ncsv::fail(NCSV_INVALID, "bad batch entry");
        const auto &v = b->expv[entry];
        if (expect) std::copy(v.begin(), v.end(), expect);
        if (norm_sq) *norm_sq = b->norm[entry];
        if (outcomes) std::copy(b->outc[entry].begin(), b->outc[entry].end(), outcomes);
    });
}

ncsv_status ncsv_batch_stats(const ncsv_batch *b, double *out) {
    return ncsv::guard_status([&]() {
        if (!b || !out) ncsv::fail(NCSV_INVALID, "null batch or output");
        uint64_t steps = 0;
        for (auto &P : b->progs) steps += P->steps.size();
        out[0] = b->kernel_ms;
        out[1] = b->plan_ms;
        out[2] = (double)b->items;
        out[3] = (double)b->launches;
        out[4] = b->amp_updates;
        out[5] = b->hbm_bytes;
        out[6] = (double)b->progs.size();
        out[7] = (double)steps;
        out[8] = (double)b->graph_replays;
    });
}

}  // extern "C"
