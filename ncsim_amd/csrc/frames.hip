// Pauli-frame sampling of noisy Clifford circuits on the GPU (ncsim's
// FrameSampler, src/frame.rs, over the column layout of
// src/pauli/columns.rs).
//
// Host: the Clifford check (frame.rs:42-52, gate.rs:658-666), the noiseless
// reference record from a stabilizer tableau (frame.rs:54-56; Aaronson and
// Gottesman's CHP with general Pauli measurements), and the circuit compiled
// to a stream of frame ops.  Device: one thread per word of 64 shots runs
// the whole op stream on its shots' frames, held by qubit in HBM as
// x[q * W + w], z[q * W + w] (W words), so every access of a wavefront is
// one contiguous 512-byte run; records go to rec[k * W + w].  A second
// kernel forms detectors and observables.  Gates are GF(2) maps of a
// qubit's or a pair's frame bits (signs do not matter for frames); a
// measurement of P records the parity of the frame's anticommutation with
// P XOR the reference bit and multiplies P into the shots of a random mask;
// a reset clears X and randomizes Z; a noise channel hits each shot with the
// channel's total probability and applies the error chosen by a second draw.
//
// Draws are counter-based so results do not depend on the launch shape:
// rnd(seed, a, b) = mix(mix(seed ^ mix(a)) ^ b), mix = SplitMix64's output
// step on v + 0x9E3779B97F4A7C15; a names the site (1<<56 | q initial Z of
// qubit q, 2<<56 | k measurement k, 3<<56 | r reset r, 4<<56 | j and
// 5<<56 | j hit and error of noise application j), b the word or the shot.
// The host stores mix(a) per site; the device forms K = mix(seed ^ mix(a))
// once per op and mix(K ^ b) per word or shot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "ncsv_internal.hpp"

namespace ncsv {
ncsv_status guard_status(const std::function<void()> &f);
}

namespace ncsv {
namespace frames {

constexpr uint64_t GAMMA = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t mix(uint64_t v) {
    uint64_t z = v + GAMMA;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t site(uint64_t kind, uint64_t index) { return (kind << 56) | index; }

// Op stream (32-bit words).
enum : uint32_t { F_LIN1 = 1, F_LIN2 = 2, F_MEAS = 3, F_RESET = 4, F_NOISE = 5 };

__host__ __device__ inline uint64_t u64at(const uint32_t *p) { return (uint64_t)p[0] | ((uint64_t)p[1] << 32); }

struct Args {
    const uint32_t *ops;   // the op stream
    uint32_t nwords;       // its length
    const uint64_t *hinit; // mix(1<<56 | q) per qubit
    uint32_t n;
    uint64_t W;            // words per column
    uint64_t shots;
    uint64_t seed;
    uint64_t *x, *z, *rec; // frames and records, by column
};

// The whole circuit on the 64 shots of word w.
__host__ __device__ inline void run_word(const Args &a, uint64_t w) {
    const uint64_t W = a.W;
    const uint64_t last = (w + 1) * 64 <= a.shots ? ~0ull : ((1ull << (a.shots - w * 64)) - 1);
    for (uint32_t q = 0; q < a.n; q++) {
        a.x[q * W + w] = 0;
        a.z[q * W + w] = mix(mix(a.seed ^ a.hinit[q]) ^ w) & last;
    }
    uint32_t i = 0;
    while (i < a.nwords) {
        const uint32_t *o = a.ops + i;
        switch (o[0]) {
        case F_LIN1: {
            const uint64_t q = o[1], m = o[2];
            const uint64_t x = a.x[q * W + w], z = a.z[q * W + w];
            a.x[q * W + w] = ((m & 1) ? x : 0) ^ ((m & 2) ? z : 0);
            a.z[q * W + w] = ((m & 4) ? x : 0) ^ ((m & 8) ? z : 0);
            i += 3;
            break;
        }
        case F_LIN2: {
            // bit (out * 4 + in) of m; bits in/out ordered x0, x1, z0, z1
            const uint64_t q0 = o[1], q1 = o[2], m = o[3];
            const uint64_t in[4] = {a.x[q0 * W + w], a.x[q1 * W + w], a.z[q0 * W + w], a.z[q1 * W + w]};
            uint64_t out[4];
            for (int t = 0; t < 4; t++) {
                uint64_t v = 0;
                for (int s = 0; s < 4; s++)
                    if ((m >> (t * 4 + s)) & 1) v ^= in[s];
                out[t] = v;
            }
            a.x[q0 * W + w] = out[0];
            a.x[q1 * W + w] = out[1];
            a.z[q0 * W + w] = out[2];
            a.z[q1 * W + w] = out[3];
            i += 4;
            break;
        }
        case F_MEAS: {
            // [3, k, ref, hsite lo, hi, nterms, terms (q | x << 30 | z << 31)...]
            const uint64_t k = o[1];
            const uint32_t nt = o[5];
            uint64_t r = 0;
            for (uint32_t t = 0; t < nt; t++) {
                const uint32_t e = o[6 + t];
                const uint64_t q = e & 0x3fffffffu;
                if (e >> 31) r ^= a.x[q * W + w];  // Z part of P anticommutes with X of the frame
                if ((e >> 30) & 1) r ^= a.z[q * W + w];
            }
            if (o[2]) r = ~r;
            a.rec[k * W + w] = r & last;
            const uint64_t msk = mix(mix(a.seed ^ u64at(o + 3)) ^ w) & last;
            for (uint32_t t = 0; t < nt; t++) {
                const uint32_t e = o[6 + t];
                const uint64_t q = e & 0x3fffffffu;
                if ((e >> 30) & 1) a.x[q * W + w] ^= msk;
                if (e >> 31) a.z[q * W + w] ^= msk;
            }
            i += 6 + nt;
            break;
        }
        case F_RESET: {
            const uint64_t q = o[1];
            a.x[q * W + w] = 0;
            a.z[q * W + w] = mix(mix(a.seed ^ u64at(o + 2)) ^ w) & last;
            i += 4;
            break;
        }
        case F_NOISE: {
            // [5, nq, q0, q1, hhit lo, hi, hwhich lo, hi, T lo, hi, always, nerr, (bits, thr lo, thr hi)...]
            const uint32_t nq = o[1], nerr = o[11];
            const uint64_t kh = mix(a.seed ^ u64at(o + 4)), kw = mix(a.seed ^ u64at(o + 6)), T = u64at(o + 8);
            const bool always = o[10] != 0;
            uint64_t fx0 = 0, fz0 = 0, fx1 = 0, fz1 = 0;
            for (uint32_t b = 0; b < 64; b++) {
                const uint64_t s = w * 64 + b;
                if (s >= a.shots) break;
                if (!always && !(mix(kh ^ s) < T)) continue;
                const uint64_t v = mix(kw ^ s);
                uint32_t pick = nerr - 1;
                for (uint32_t e = 0; e < nerr; e++)
                    if (v < u64at(o + 13 + 3 * e)) {
                        pick = e;
                        break;
                    }
                const uint32_t bits = o[12 + 3 * pick];
                const uint64_t bit = 1ull << b;
                if (bits & 1) fx0 ^= bit;
                if (bits & 2) fx1 ^= bit;
                if ((bits >> 8) & 1) fz0 ^= bit;
                if ((bits >> 8) & 2) fz1 ^= bit;
            }
            const uint64_t q0 = o[2], q1 = o[3];
            if (fx0 | fz0) {
                a.x[q0 * W + w] ^= fx0;
                a.z[q0 * W + w] ^= fz0;
            }
            if (nq > 1 && (fx1 | fz1)) {
                a.x[q1 * W + w] ^= fx1;
                a.z[q1 * W + w] ^= fz1;
            }
            i += 12 + 3 * nerr;
            break;
        }
        default: return;  // not produced by the compiler
        }
    }
}

__global__ __launch_bounds__(256) void k_frames(Args a) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < a.W) run_word(a, w);
}

// Detectors and observables: out[d * W + w] = parity of the listed records,
// complemented where the reference's parity is odd (flip[d]), masked to the
// shots.  idx[off[d] .. off[d+1]) are measurement indices.
__global__ __launch_bounds__(256) void k_parities(const uint64_t *rec, uint64_t W, uint64_t shots,
                                                  const uint32_t *off, const uint32_t *idx, const uint8_t *flip,
                                                  uint32_t count, uint64_t *out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)count * W) return;
    const uint64_t d = t / W, w = t % W;
    uint64_t v = 0;
    for (uint32_t e = off[d]; e < off[d + 1]; e++) v ^= rec[(uint64_t)idx[e] * W + w];
    if (flip[d]) v = ~v;
    const uint64_t last = (w + 1) * 64 <= shots ? ~0ull : ((1ull << (shots - w * 64)) - 1);
    out[d * W + w] = v & last;
}

// --------------------------------------------------------------- host side
// A Pauli on the register as bit vectors with phase i^ph X^x Z^z.
struct Row {
    std::vector<uint64_t> x, z;
    uint32_t ph = 0;
};

// Local image of a generator: bits on the gate's k qubits and a phase.
struct Img {
    uint32_t x = 0, z = 0, ph = 0;
};
struct Images {
    std::vector<Img> xs, zs;
};

Img parse_signed(const char *s) {
    Img r;
    if (*s == '-') {
        r.ph = 2;
        s++;
    }
    while (*s) {
        const char l = *s++;
        uint32_t q = 0;
        while (*s >= '0' && *s <= '9') q = q * 10 + (uint32_t)(*s++ - '0');
        if (l == 'X' || l == 'Y') r.x |= 1u << q;
        if (l == 'Y' || l == 'Z') r.z |= 1u << q;
        if (l == 'Y') r.ph += 1;
        if (*s == '*') s++;
    }
    r.ph &= 3;
    return r;
}

Images named(const std::vector<const char *> &xs, const std::vector<const char *> &zs) {
    Images im;
    for (auto s : xs) im.xs.push_back(parse_signed(s));
    for (auto s : zs) im.zs.push_back(parse_signed(s));
    return im;
}

// The images of X_j and Z_j under a Clifford gate (gate.rs:587-604 for the
// named gates; PHASE / CPHASE when their phase is a power of i / +-1), or
// false for a non-Clifford gate.
bool clifford_images(const GateDesc &g, Images *out) {
    switch (g.id) {
    case NCSV_GATE_I: *out = named({"X0"}, {"Z0"}); return true;
    case NCSV_GATE_X: *out = named({"X0"}, {"-Z0"}); return true;
    case NCSV_GATE_Y: *out = named({"-X0"}, {"-Z0"}); return true;
    case NCSV_GATE_Z: *out = named({"-X0"}, {"Z0"}); return true;
    case NCSV_GATE_H: *out = named({"Z0"}, {"X0"}); return true;
    case NCSV_GATE_S: *out = named({"Y0"}, {"Z0"}); return true;
    case NCSV_GATE_S_DAG: *out = named({"-Y0"}, {"Z0"}); return true;
    case NCSV_GATE_SQRT_X: *out = named({"X0"}, {"-Y0"}); return true;
    case NCSV_GATE_SQRT_X_DAG: *out = named({"X0"}, {"Y0"}); return true;
    case NCSV_GATE_SQRT_Y: *out = named({"-Z0"}, {"X0"}); return true;
    case NCSV_GATE_SQRT_Y_DAG: *out = named({"Z0"}, {"-X0"}); return true;
    case NCSV_GATE_CX: *out = named({"X0*X1", "X1"}, {"Z0", "Z0*Z1"}); return true;
    case NCSV_GATE_CY: *out = named({"X0*Y1", "Z0*X1"}, {"Z0", "Z0*Z1"}); return true;
    case NCSV_GATE_CZ: *out = named({"X0*Z1", "Z0*X1"}, {"Z0", "Z1"}); return true;
    case NCSV_GATE_SWAP: *out = named({"X1", "X0"}, {"Z1", "Z0"}); return true;
    case NCSV_GATE_PAULI: {
        Images im;
        for (uint32_t q = 0; q < g.pn; q++) {
            im.xs.push_back({1u << q, 0, ((g.pz >> q) & 1) ? 2u : 0u});
            im.zs.push_back({0, 1u << q, ((g.px >> q) & 1) ? 2u : 0u});
        }
        *out = im;
        return true;
    }
    case NCSV_GATE_PHASE:
    case NCSV_GATE_CPHASE: {
        // zeta_{2^bits}^k: a power of i when 4k = 0 mod 2^bits
        if (g.bits > 62) return false;
        const int64_t m = (int64_t)1 << g.bits;
        const int64_t km = ((g.k % m) + m) % m;
        if ((km * 4) % m) return false;
        const int64_t quarter = km * 4 / m;
        if (g.id == NCSV_GATE_PHASE) {
            static const char *xi[4] = {"X0", "Y0", "-X0", "-Y0"};
            *out = named({xi[quarter]}, {"Z0"});
            return true;
        }
        if (quarter == 0) {
            *out = named({"X0", "X1"}, {"Z0", "Z1"});
            return true;
        }
        if (quarter == 2) {
            *out = named({"X0*Z1", "Z0*X1"}, {"Z0", "Z1"});
            return true;
        }
        return false;  // a controlled +-i phase (CS) is not Clifford
    }
    default: return false;
    }
}

// (i^pa X^xa Z^za)(i^pb X^xb Z^zb) = i^(pa+pb+2 popcount(za & xb)) X^(xa^xb) Z^(za^zb)
void rowmul(Row &a, const Row &b) {
    uint32_t c = 0;
    for (size_t i = 0; i < a.x.size(); i++) {
        c += (uint32_t)__builtin_popcountll(a.z[i] & b.x[i]);
        a.x[i] ^= b.x[i];
        a.z[i] ^= b.z[i];
    }
    a.ph = (a.ph + b.ph + 2 * c) & 3;
}

bool anticommutes(const Row &a, const Row &b) {
    uint32_t c = 0;
    for (size_t i = 0; i < a.x.size(); i++)
        c += (uint32_t)__builtin_popcountll((a.x[i] & b.z[i]) ^ (a.z[i] & b.x[i]));
    return c & 1;
}

inline bool getb(const std::vector<uint64_t> &v, uint32_t q) { return (v[q >> 6] >> (q & 63)) & 1; }
inline void setb(std::vector<uint64_t> &v, uint32_t q, bool b) {
    if (b)
        v[q >> 6] |= 1ull << (q & 63);
    else
        v[q >> 6] &= ~(1ull << (q & 63));
}

Img local_product(const Images &im, uint32_t lx, uint32_t lz) {
    Img acc;
    auto mul = [&](const Img &b) {
        acc.ph = (acc.ph + b.ph + 2 * (uint32_t)__builtin_popcount(acc.z & b.x)) & 3;
        acc.x ^= b.x;
        acc.z ^= b.z;
    };
    for (size_t j = 0; j < im.xs.size(); j++)
        if ((lx >> j) & 1) mul(im.xs[j]);
    for (size_t j = 0; j < im.zs.size(); j++)
        if ((lz >> j) & 1) mul(im.zs[j]);
    return acc;
}

// A stabilizer tableau: 2n rows (destabilizers, then stabilizers).
struct Tableau {
    uint32_t n;
    std::vector<Row> rows;
    explicit Tableau(uint32_t n_) : n(n_) {
        const size_t nw = (n + 63) / 64;
        rows.resize(2 * n);
        for (uint32_t i = 0; i < 2 * n; i++) {
            rows[i].x.assign(nw, 0);
            rows[i].z.assign(nw, 0);
            if (i < n)
                setb(rows[i].x, i, true);
            else
                setb(rows[i].z, i - n, true);
        }
    }
    void apply(const Images &im, const std::vector<uint32_t> &qs) {
        for (Row &r : rows) {
            uint32_t lx = 0, lz = 0;
            for (size_t j = 0; j < qs.size(); j++) {
                lx |= (uint32_t)getb(r.x, qs[j]) << j;
                lz |= (uint32_t)getb(r.z, qs[j]) << j;
            }
            if (!lx && !lz) continue;
            const Img m = local_product(im, lx, lz);
            for (size_t j = 0; j < qs.size(); j++) {
                setb(r.x, qs[j], (m.x >> j) & 1);
                setb(r.z, qs[j], (m.z >> j) & 1);
            }
            r.ph = (r.ph + m.ph) & 3;
        }
    }
    // Measures the Hermitian Pauli p; draw() gives a random outcome bit.
    template <typename D> uint32_t measure(const Row &p, D draw) {
        int piv = -1;
        for (uint32_t i = n; i < 2 * n && piv < 0; i++)
            if (anticommutes(rows[i], p)) piv = (int)i;
        if (piv >= 0) {
            for (uint32_t j = 0; j < 2 * n; j++)
                if ((int)j != piv && anticommutes(rows[j], p)) rowmul(rows[j], rows[piv]);
            rows[piv - n] = rows[piv];
            const uint32_t o = draw();
            rows[piv] = p;
            rows[piv].ph = (p.ph + 2 * o) & 3;
            return o;
        }
        Row acc;
        acc.x.assign(p.x.size(), 0);
        acc.z.assign(p.z.size(), 0);
        for (uint32_t i = 0; i < n; i++)
            if (anticommutes(rows[i], p)) rowmul(acc, rows[i + n]);
        if (acc.x != p.x || acc.z != p.z) fail(NCSV_INVALID, "frame sampler: tableau measurement inconsistent");
        const uint32_t d = (acc.ph - p.ph) & 3;
        return d == 2 ? 1 : 0;
    }
};

}  // namespace frames
}  // namespace ncsv

using namespace ncsv;
using namespace ncsv::frames;

struct ncsv_frames {
    uint32_t n = 0;
    uint64_t m = 0;                      // measurements
    std::vector<uint8_t> reference;      // noiseless record
    std::vector<uint32_t> ops;           // the frame op stream
    std::vector<uint64_t> hinit;         // mix(1<<56 | q)
    std::vector<std::vector<uint32_t>> dets, obss;  // absolute measurement indices
    // device copies
    uint32_t *d_ops = nullptr;
    uint64_t *d_hinit = nullptr;
    uint64_t *d_buf = nullptr;
    size_t buf_words = 0;
    hipStream_t stream = nullptr;
    bool ready = false;                  // device copies uploaded
    // pinned staging for large downloads into the caller's memory
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t staged[2] = {nullptr, nullptr};
    // the parity kernel's tables and output, kept between calls (grown on demand)
    void *d_par = nullptr;
    size_t par_cap = 0;
    ~ncsv_frames() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (d_ops) (void)hipFree(d_ops);
        if (d_hinit) (void)hipFree(d_hinit);
        if (d_buf) (void)hipFree(d_buf);
        for (int k = 0; k < 2; k++) {
            if (stage[k]) (void)hipHostFree(stage[k]);
            if (staged[k]) (void)hipEventDestroy(staged[k]);
        }
        if (d_par) (void)hipFree(d_par);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void push64(std::vector<uint32_t> &v, uint64_t x) {
    v.push_back((uint32_t)x);
    v.push_back((uint32_t)(x >> 32));
}

// FrameSampler::new (frame.rs:41-58): Clifford check, reference record,
// and the op stream.
ncsv_frames *compile(const ncsv_circuit *c, uint64_t reference_seed) {
    auto f = std::make_unique<ncsv_frames>();
    const uint32_t n = c->n;
    f->n = n;
    Tableau tab(n);
    uint64_t rs = reference_seed;  // SplitMix64 stream of the reference's random outcomes
    auto draw = [&]() -> uint32_t {
        rs += GAMMA;
        uint64_t z = rs;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (uint32_t)((z ^ (z >> 31)) >> 63);
    };
    const size_t nw = (n + 63) / 64;
    uint64_t k = 0, r = 0, j = 0;
    for (uint32_t q = 0; q < n; q++) f->hinit.push_back(mix(site(1, q)));
    for (const COp &op : c->ops) {
        switch (op.kind) {
        case CKind::Gate: {
            Images im;
            if (!clifford_images(op.g, &im))
                fail(NCSV_UNSUPPORTED, "FrameSampler: non-Clifford gate " + gate_name(op.g));
            tab.apply(im, op.qubits);
            const uint32_t kq = (uint32_t)op.qubits.size();
            if (op.g.id == NCSV_GATE_PAULI || op.g.id == NCSV_GATE_I || op.g.id == NCSV_GATE_X ||
                op.g.id == NCSV_GATE_Y || op.g.id == NCSV_GATE_Z)
                break;  // a Pauli leaves frames alone (signs do not matter)
            if (kq == 1) {
                const uint32_t m = (im.xs[0].x & 1) | ((im.zs[0].x & 1) << 1) | ((im.xs[0].z & 1) << 2) |
                                   ((im.zs[0].z & 1) << 3);
                if (m == 9) break;  // identity on frames (S^2 = Z and the like)
                f->ops.insert(f->ops.end(), {F_LIN1, op.qubits[0], m});
            } else if (kq == 2) {
                // bit (out * 4 + in), in/out ordered x0, x1, z0, z1
                uint32_t m = 0;
                const Img in[4] = {im.xs[0], im.xs[1], im.zs[0], im.zs[1]};
                for (int s = 0; s < 4; s++)
                    for (int t = 0; t < 4; t++) {
                        const uint32_t bit = t < 2 ? (in[s].x >> t) & 1 : (in[s].z >> (t - 2)) & 1;
                        if (bit) m |= 1u << (t * 4 + s);
                    }
                if (m == 0x8421) break;  // identity
                f->ops.insert(f->ops.end(), {F_LIN2, op.qubits[0], op.qubits[1], m});
            } else {
                fail(NCSV_UNSUPPORTED, "FrameSampler: Clifford gates on more than 2 qubits");
            }
            break;
        }
        case CKind::Measure:
        case CKind::Reset: {
            Row p;
            p.x.assign(nw, 0);
            p.z.assign(nw, 0);
            uint32_t q0 = 0;
            if (op.kind == CKind::Measure) {
                for (uint32_t q = 0; q < n && q < 64; q++) {
                    setb(p.x, q, (op.x >> q) & 1);
                    setb(p.z, q, (op.z >> q) & 1);
                }
                p.ph = op.ph;
            } else {
                q0 = op.qubits[0];
                setb(p.z, q0, true);
            }
            const uint32_t o = tab.measure(p, draw);
            if (op.kind == CKind::Reset) {
                if (o) tab.apply(named({"X0"}, {"-Z0"}), {q0});
                f->ops.insert(f->ops.end(), {F_RESET, q0});
                push64(f->ops, mix(site(3, r)));
                r++;
                break;
            }
            f->reference.push_back((uint8_t)o);
            std::vector<uint32_t> terms;
            for (uint32_t q = 0; q < n && q < 64; q++) {
                const uint32_t xb = (op.x >> q) & 1, zb = (op.z >> q) & 1;
                if (xb || zb) terms.push_back(q | (xb << 30) | (zb << 31));
            }
            f->ops.insert(f->ops.end(), {F_MEAS, (uint32_t)k, o});
            push64(f->ops, mix(site(2, k)));
            f->ops.push_back((uint32_t)terms.size());
            f->ops.insert(f->ops.end(), terms.begin(), terms.end());
            k++;
            break;
        }
        case CKind::Noise: {
            const std::vector<NoiseErr> errs = noise_errors(op.ch, op.p);
            double total = 0.0;
            for (auto &e : errs) total += e.p;
            if (total > 0.0) {
                f->ops.insert(f->ops.end(), {F_NOISE, (uint32_t)op.qubits.size(), op.qubits[0],
                                             op.qubits.size() > 1 ? op.qubits[1] : 0u});
                push64(f->ops, mix(site(4, j)));
                push64(f->ops, mix(site(5, j)));
                const bool always = total >= 1.0;
                push64(f->ops, always ? ~0ull : (uint64_t)(total * 18446744073709551616.0));
                f->ops.push_back(always ? 1u : 0u);
                f->ops.push_back((uint32_t)errs.size());
                double acc = 0.0;
                for (auto &e : errs) {
                    acc += e.p;
                    const double fr = acc / total;
                    f->ops.push_back(e.x | (e.z << 8));
                    push64(f->ops, fr >= 1.0 ? ~0ull : (uint64_t)(fr * 18446744073709551616.0));
                }
            }
            j++;
            break;
        }
        case CKind::Detector: {
            std::vector<uint32_t> d;
            for (uint32_t lb : op.lookback) d.push_back((uint32_t)(k - lb));
            f->dets.push_back(d);
            break;
        }
        case CKind::Observable: {
            if (f->obss.size() <= op.index) f->obss.resize(op.index + 1);
            for (uint32_t lb : op.lookback) f->obss[op.index].push_back((uint32_t)(k - lb));
            break;
        }
        case CKind::Tick: break;
        }
    }
    if (n > 0x3fffffffu) fail(NCSV_INVALID, "FrameSampler: too many qubits");
    f->m = k;
    return f.release();
}

Args host_args(const ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *x, uint64_t *z, uint64_t *rec) {
    Args a;
    a.ops = f->ops.data();
    a.nwords = (uint32_t)f->ops.size();
    a.hinit = f->hinit.data();
    a.n = f->n;
    a.W = (shots + 63) / 64;
    a.shots = shots;
    a.seed = seed;
    a.x = x;
    a.z = z;
    a.rec = rec;
    return a;
}

void ensure_device(ncsv_frames *f) {
    if (f->ready) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(NCSV_DEVICE, "no GPU device visible");
    // Each step is kept by the sampler as it succeeds (the destructor frees
    // it); `ready` is set only once all of them have.
    if (!f->stream) check_hip(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!f->d_ops) check_hip(hipMalloc(&f->d_ops, std::max<size_t>(1, f->ops.size()) * 4), "hipMalloc(frame ops)");
    if (!f->d_hinit)
        check_hip(hipMalloc(&f->d_hinit, std::max<size_t>(1, f->hinit.size()) * 8), "hipMalloc(frame sites)");
    if (!f->ops.empty())
        check_hip(hipMemcpy(f->d_ops, f->ops.data(), f->ops.size() * 4, hipMemcpyHostToDevice), "upload ops");
    if (!f->hinit.empty())
        check_hip(hipMemcpy(f->d_hinit, f->hinit.data(), f->hinit.size() * 8, hipMemcpyHostToDevice), "upload sites");
    f->ready = true;
}

// Parity lists of detectors and observables in CSR form with the reference flips.
void parity_lists(const ncsv_frames *f, const std::vector<std::vector<uint32_t>> &sets, std::vector<uint32_t> *off,
                  std::vector<uint32_t> *idx, std::vector<uint8_t> *flip) {
    off->assign(1, 0);
    for (auto &s : sets) {
        uint8_t fl = 0;
        for (uint32_t kk : s) {
            idx->push_back(kk);
            fl ^= f->reference[kk];
        }
        off->push_back((uint32_t)idx->size());
        flip->push_back(fl);
    }
}

// Runs the sampler on the device; records stay in f->d_buf after the frames.
uint64_t *sample_device(ncsv_frames *f, uint64_t shots, uint64_t seed) {
    ensure_device(f);
    const uint64_t W = (shots + 63) / 64;
    const size_t need = (size_t)(2 * (uint64_t)f->n + f->m) * W + 1;
    if (f->buf_words < need) {
        if (f->d_buf) (void)hipFree(f->d_buf);
        f->d_buf = nullptr;
        f->buf_words = 0;
        check_hip(hipMalloc(&f->d_buf, need * 8), "hipMalloc(frames)");
        f->buf_words = need;
    }
    uint64_t *x = f->d_buf, *z = x + (size_t)f->n * W, *rec = z + (size_t)f->n * W;
    Args a = host_args(f, shots, seed, x, z, rec);
    a.ops = f->d_ops;
    a.hinit = f->d_hinit;
    if (W) {
        hipLaunchKernelGGL(k_frames, dim3((uint32_t)((W + 255) / 256)), dim3(256), 0, f->stream, a);
        check_hip(hipGetLastError(), "k_frames");
    }
    return rec;
}

// A device-to-host copy into the caller's (pageable) memory, done when it
// returns.  Large ones go through two pinned staging buffers: one chunk
// copies in while host threads spread the previous one into the destination
// (a pageable destination is otherwise copied a page-locked piece at a time).
void download(ncsv_frames *f, void *dst, const void *src, size_t bytes) {
    constexpr size_t CH = size_t(32) << 20;
    const char *e = getenv("NCSV_FRAMES_STAGE");  // 0: one plain copy (for comparison)
    if (bytes < 4 * CH || (e && e[0] == '0')) {
        check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, f->stream), "download");
        check_hip(hipStreamSynchronize(f->stream), "hipStreamSynchronize");
        return;
    }
    for (int k = 0; k < 2; k++) {
        if (!f->stage[k] && hipHostMalloc(&f->stage[k], CH, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();  // no pinned memory to be had: one plain copy
            f->stage[k] = nullptr;
            check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, f->stream), "download");
            check_hip(hipStreamSynchronize(f->stream), "hipStreamSynchronize");
            return;
        }
        if (!f->staged[k]) check_hip(hipEventCreateWithFlags(&f->staged[k], hipEventDisableTiming), "hipEventCreate");
    }
    const unsigned T = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    auto spread = [&](size_t c) {
        const size_t off = c * CH, len = std::min(CH, bytes - off), part = (len + T - 1) / T;
        const char *from = (const char *)f->stage[c & 1];
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; t++) {
            const size_t a = std::min(len, t * part), b = std::min(len, a + part);
            if (a < b) th.emplace_back([=]() { memcpy((char *)dst + off + a, from + a, b - a); });
        }
        for (auto &x : th) x.join();
    };
    const size_t nch = (bytes + CH - 1) / CH;
    for (size_t c = 0; c < nch; c++) {
        const size_t off = c * CH, len = std::min(CH, bytes - off);
        // stage c & 1 was last spread in the iteration before the previous one
        check_hip(hipMemcpyAsync(f->stage[c & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, f->stream),
                  "download");
        check_hip(hipEventRecord(f->staged[c & 1], f->stream), "hipEventRecord");
        if (c > 0) {
            check_hip(hipEventSynchronize(f->staged[(c - 1) & 1]), "hipEventSynchronize");
            spread(c - 1);
        }
    }
    check_hip(hipEventSynchronize(f->staged[(nch - 1) & 1]), "hipEventSynchronize");
    spread(nch - 1);
}

void parities_device(ncsv_frames *f, const uint64_t *rec, uint64_t shots, const std::vector<std::vector<uint32_t>> &sets,
                     uint64_t *host_out) {
    if (sets.empty()) return;
    const uint64_t W = (shots + 63) / 64;
    std::vector<uint32_t> off, idx;
    std::vector<uint8_t> flip;
    parity_lists(f, sets, &off, &idx, &flip);
    // one device buffer for the output (8-byte words first), the offsets, the
    // indices and the flips, kept on the sampler between calls
    const uint64_t items = sets.size() * W;
    auto up256 = [](size_t v) { return (v + 255) & ~size_t(255); };
    const size_t o_off = up256(std::max<uint64_t>(1, items) * 8), o_idx = o_off + up256(off.size() * 4),
                 o_flip = o_idx + up256(std::max<size_t>(1, idx.size()) * 4), total = o_flip + up256(flip.size() + 1);
    if (total > f->par_cap) {
        if (f->d_par) check_hip(hipFree(f->d_par), "hipFree");
        f->d_par = nullptr;
        f->par_cap = 0;
        check_hip(hipMalloc(&f->d_par, total), "hipMalloc(parities)");
        f->par_cap = total;
    }
    uint64_t *d_out = (uint64_t *)f->d_par;
    uint32_t *d_off = (uint32_t *)((char *)f->d_par + o_off), *d_idx = (uint32_t *)((char *)f->d_par + o_idx);
    uint8_t *d_flip = (uint8_t *)((char *)f->d_par + o_flip);
    check_hip(hipMemcpyAsync(d_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, f->stream), "upload");
    if (!idx.empty())
        check_hip(hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, f->stream), "upload");
    check_hip(hipMemcpyAsync(d_flip, flip.data(), flip.size(), hipMemcpyHostToDevice, f->stream), "upload");
    if (items) {
        hipLaunchKernelGGL(k_parities, dim3((uint32_t)((items + 255) / 256)), dim3(256), 0, f->stream, rec, W, shots,
                           d_off, d_idx, d_flip, (uint32_t)sets.size(), d_out);
        check_hip(hipGetLastError(), "k_parities");
        download(f, host_out, d_out, items * 8);
    }
    check_hip(hipStreamSynchronize(f->stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" {

ncsv_status ncsv_frames_new(const ncsv_circuit *c, uint64_t reference_seed, ncsv_frames **out) {
    return guard_status([&] {
        if (!c || !out) fail(NCSV_INVALID, "null argument");
        *out = compile(c, reference_seed);
    });
}

ncsv_status ncsv_frames_destroy(ncsv_frames *f) {
    return guard_status([&] { delete f; });
}

ncsv_status ncsv_frames_shape(const ncsv_frames *f, uint64_t *measurements, uint64_t *detectors, uint64_t *observables) {
    return guard_status([&] {
        if (!f) fail(NCSV_INVALID, "null sampler");
        if (measurements) *measurements = f->m;
        if (detectors) *detectors = f->dets.size();
        if (observables) *observables = f->obss.size();
    });
}

ncsv_status ncsv_frames_reference(const ncsv_frames *f, uint8_t *record, uint64_t cap) {
    return guard_status([&] {
        if (!f) fail(NCSV_INVALID, "null sampler");
        if (f->m && !record) fail(NCSV_INVALID, "null record");
        for (uint64_t i = 0; i < f->m && i < cap; i++) record[i] = f->reference[i];
    });
}

ncsv_status ncsv_frames_sample_packed(ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *out) {
    return guard_status([&] {
        if (!f) fail(NCSV_INVALID, "null sampler");
        const uint64_t W = (shots + 63) / 64;
        if (f->m && W && !out) fail(NCSV_INVALID, "null output");
        uint64_t *rec = sample_device(f, shots, seed);
        if (f->m && W) download(f, out, rec, f->m * W * 8);
        check_hip(hipStreamSynchronize(f->stream), "hipStreamSynchronize");
    });
}

ncsv_status ncsv_frames_sample_detectors(ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *det, uint64_t *obs) {
    return guard_status([&] {
        if (!f) fail(NCSV_INVALID, "null sampler");
        const uint64_t W = (shots + 63) / 64;
        if (W && ((!f->dets.empty() && !det) || (!f->obss.empty() && !obs))) fail(NCSV_INVALID, "null output");
        const uint64_t *rec = sample_device(f, shots, seed);
        parities_device(f, rec, shots, f->dets, det);
        parities_device(f, rec, shots, f->obss, obs);
        check_hip(hipStreamSynchronize(f->stream), "hipStreamSynchronize");
    });
}

ncsv_status ncsv_frames_emulate_packed(const ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *out) {
    return guard_status([&] {
        if (!f) fail(NCSV_INVALID, "null sampler");
        const uint64_t W = (shots + 63) / 64;
        std::vector<uint64_t> x((size_t)f->n * W + 1), z((size_t)f->n * W + 1);
        std::vector<uint64_t> rec((size_t)f->m * W + 1);
        const Args a = host_args(f, shots, seed, x.data(), z.data(), rec.data());
        for (uint64_t w = 0; w < W; w++) run_word(a, w);
        if (f->m && W) memcpy(out, rec.data(), f->m * W * 8);
    });
}

}  // extern "C"
