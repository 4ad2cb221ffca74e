// The batched form's persistent kernel body, written once for two builds:
// the gfx950 device code (batch.hip) and a host build in which one
// workgroup runs as BT host threads with real barriers (batch_sim.cpp), so
// the exact kernel logic can be run and checked without a device.  The
// including file defines the primitives:
//   NCSV_DEV, NCSV_KERNEL, NCSV_SHARED, NCSV_TID, NCSV_SYNC(),
//   NCSV_SHFL_XOR(x, mask), NCSV_ATOMIC_ADD(p, v), NCSV_ATOMIC_OR(p, v),
//   NCSV_LOAD_AGENT(p), NCSV_ADD_AGENT(p, v), NCSV_FENCE_ACQUIRE(),
//   NCSV_FENCE_RELEASE(), NCSV_SLEEP(), NCSV_UNIFORM32(v) / NCSV_UNIFORM64(v)
//   (the value of the wavefront's first lane), NCSV_LANE0_STORE(dst, v)
//   (dst = lane 0's v), NCSV_OPAQUE(x) (x unknown to the optimizer from
//   here on), NCSV_MFMA16(a, b, acc) (one v_mfma_f32_16x16x4_f32: lane l
//   holds A[l % 16][l / 16] and B[l / 16][l % 16]; acc[v] is D[4 (l / 16) + v][l % 16]),
//   NCSV_MFMA16D(a, b, acc) (v_mfma_f64_16x16x4_f64: A and B as above, acc[v] is
//   D[4 v + l / 16][l % 16]), V2<T>::type (float2 / double2), M4<T>::type (4 floats / doubles).
//
// The tile sits in LDS at swizzled slots: slot(x) = x ^ f(x >> S), an
// XOR-linear bijection that spreads the strided patterns of pair, quad and
// group accesses over the banks.  Every index used by an op is a per-thread
// part XOR a wave-uniform part (both swizzled), so an access costs one XOR.
// The read-only tables arrive as separate restrict-qualified arguments so
// their wave-uniform reads go through the scalar cache.
#pragma once

#include "batch.hpp"

namespace ncsv {
namespace batch {
namespace kern {

constexpr uint32_t CMAX = 18;       // values per block reduction (9 Paulis)
constexpr uint32_t SMAX = 4096;     // noise codes staged in LDS per record

template <typename V> NCSV_DEV inline V cm(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename V> NCSV_DEV inline V cadd(V a, V b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
// i^k v, exact
template <typename V> NCSV_DEV inline V ipow(uint32_t k, V v) {
    V r = v;
    switch (k & 3) {
    case 1: r.x = -v.y; r.y = v.x; break;
    case 2: r.x = -v.x; r.y = -v.y; break;
    case 3: r.x = v.y; r.y = -v.x; break;
    default: break;
    }
    return r;
}
template <typename V> NCSV_DEV inline V coef(const double *c, uint32_t i) {
    V r;
    r.x = (decltype(r.x))c[2 * i];
    r.y = (decltype(r.y))c[2 * i + 1];
    return r;
}
NCSV_DEV inline uint32_t ins0(uint32_t j, uint32_t p) {
    const uint32_t lo = j & ((1u << p) - 1);
    return ((j ^ lo) << 1) | lo;
}
// j with a bit inserted at each position of `mask` (ascending), set where `ones` is
NCSV_DEV inline uint32_t deposit(uint32_t j, uint32_t mask, uint32_t ones) {
    for (uint32_t m = mask; m; m &= m - 1) {
        const uint32_t p = (uint32_t)__builtin_ctz(m);
        j = ins0(j, p) | (ones & (1u << p));
    }
    return j;
}
template <uint32_t N> struct KTag {
    static constexpr uint32_t value = N;
};

// One 16x16x4 MFMA of the calling wavefront in the element type.
NCSV_DEV inline M4<float>::type mfma16(float a, float b, M4<float>::type acc) { return NCSV_MFMA16(a, b, acc); }
NCSV_DEV inline M4<double>::type mfma16(double a, double b, M4<double>::type acc) { return NCSV_MFMA16D(a, b, acc); }

NCSV_DEV inline uint32_t par32(uint32_t v) { return (uint32_t)__builtin_popcount(v) & 1u; }
NCSV_DEV inline uint32_t par64(uint64_t v) { return (uint32_t)__builtin_popcountll(v) & 1u; }

// LDS slot of tile index x (rows of 256 bytes: 32 float2 or 16 double2)
template <typename V> NCSV_DEV inline uint32_t sw(uint32_t x) {
    constexpr uint32_t S = sizeof(V) == 16 ? 4u : 5u, M = (1u << S) - 1;
    return x ^ (((x >> S) ^ (x >> (2 * S))) & M);
}

// f(x, slot) for every x = deposit(j, mask, ones), j < NL >> popcount(mask).
// deposit and the swizzle are linear over disjoint bits, so x splits into
// the thread's part (j's low log2(BT) bits) and a wave-uniform part per iteration.
template <typename V, uint32_t BT, typename F>
NCSV_DEV inline void visit(uint32_t NL, uint32_t tid, uint32_t mask, uint32_t ones, F &&f) {
    const uint32_t cnt = NL >> __builtin_popcount(mask);
    const uint32_t xt = deposit(tid, mask, ones);
    if (cnt <= BT) {
        if (tid < cnt) f(xt, sw<V>(xt));
        return;
    }
    const uint32_t st = sw<V>(xt);
    // h = deposit(i * BT, mask, 0) by a masked increment: the positions of
    // mask and of the thread's bits are held at one while adding
    const uint32_t fill = mask | deposit(BT - 1, mask, 0);
    uint32_t h = 0;
#pragma unroll 2
    for (uint32_t i = 0; i < cnt / BT; i++) {
        f(xt | h, st ^ sw<V>(h));
        h = ((h | fill) + 1) & ~fill;
    }
}

// Sums v[0..cnt) over the workgroup (fixed order) into bc[0..cnt), which
// stays valid until the next barrier-separated reduction (NW wavefronts).
template <uint32_t NW>
NCSV_DEV inline void block_sum(double (&v)[CMAX], uint32_t cnt, double *red, double *bc) {
    const uint32_t tid = NCSV_TID, lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (uint32_t k = 0; k < CMAX; k++) {
        if (k < cnt) {
            double x = v[k];
            for (uint32_t o = 32; o; o >>= 1) x += NCSV_SHFL_XOR(x, o);
            v[k] = x;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (uint32_t k = 0; k < CMAX; k++)
            if (k < cnt) red[w * CMAX + k] = v[k];
    }
    NCSV_SYNC();
    if (tid < cnt) {
        double s = 0.0;
        for (uint32_t i = 0; i < NW; i++) s += red[i * CMAX + tid];
        bc[tid] = s;
    }
    NCSV_SYNC();
}

// QuantumState::measure's decision (mod.rs:113-131) from FP64 sums.
NCSV_DEV inline uint32_t decide(const double *__restrict__ uni, const uint8_t *__restrict__ forced, uint32_t ordinal,
                                double nn, double e) {
    // a NaN (zero state) stays NaN as under f64::clamp: a draw, outcome 0
    double p_plus = (1.0 + e / nn) / 2.0;
    if (p_plus < 0.0) p_plus = 0.0;
    if (p_plus > 1.0) p_plus = 1.0;
    const uint8_t f = forced ? forced[ordinal] : (uint8_t)0xff;
    if (f != 0xff) return f & 1u;
    if (p_plus > 1.0 - 1e-12) return 0;
    if (p_plus < 1e-12) return 1;
    return uni[ordinal] >= p_plus ? 1u : 0u;
}

// BT threads per workgroup (512 or 256); tiles of up to 2^LB amplitudes.
template <typename T, uint32_t LB, uint32_t BT>
NCSV_KERNEL void k_batch(const BStepD *__restrict__ steps, const BOpD *__restrict__ ops,
                         const double *__restrict__ cf, const BPauliD *__restrict__ paulis,
                         const BMGroupD *__restrict__ mgroups, const BEntryD *__restrict__ entries,
                         const BRecD *__restrict__ recs, const uint32_t *__restrict__ rec_blk,
                         const double *__restrict__ uni, const uint8_t *__restrict__ forced,
                         const uint8_t *__restrict__ errs, BArgs A) {
    using V = typename V2<T>::type;
    static_assert(BT == 512 || BT == 256, "workgroups of 512 or 256 threads");
    constexpr uint32_t NW = BT / 64;                 // wavefronts per workgroup
    constexpr uint32_t LOB = BT == 512 ? 9u : 8u;    // log2(BT): tile bits taken from the thread index
    NCSV_SHARED V tile[1u << LB];
    NCSV_SHARED double red[NW * CMAX];
    NCSV_SHARED double bc[CMAX];
    NCSV_SHARED unsigned long long s_claim;
    NCSV_SHARED uint8_t scode[SMAX];  // the record's noise codes (steps with an error drawn)
    const uint32_t tid = NCSV_TID;
    // Queue traffic (claim, wait, completion) is done by the whole first
    // wavefront under a wave-uniform branch, and the claim is read back as a
    // wave-uniform value: a lane-divergent branch next to the loop's
    // barriers lets the compiler split one lane from the rest of its wave
    // across iterations, which deadlocks the workgroup.
    const uint32_t wave = NCSV_UNIFORM32(tid >> 6);
    constexpr uint32_t KH = (1u << LB) / BT;  // amplitudes per thread in a full tile
    constexpr uint32_t HB = LB - LOB;
    const uint32_t st_tid = sw<V>(tid);
    for (;;) {
        if (wave == 0) {
            const unsigned long long old = NCSV_ATOMIC_ADD(A.next, tid == 0 ? 1ull : 0ull);
            NCSV_LANE0_STORE(s_claim, old);
        }
        NCSV_SYNC();
        const uint64_t item = NCSV_UNIFORM64(s_claim);
        if (item >= A.total) break;
        uint32_t r = rec_blk[item >> 4];
        while (recs[r + 1].first <= item) r++;
        const BRecD rec = recs[r];
        const uint32_t ntl = 1u << rec.lgt;
        const uint64_t t0 = (item - rec.first) << rec.lgt;
        if (rec.prev != ~0u) {
            if (wave == 0) {
                uint32_t spins = 0;
                for (;;) {
                    const uint32_t d = NCSV_UNIFORM32(NCSV_LOAD_AGENT(A.done + rec.prev));
                    if (d >= rec.prev_tiles) break;
                    NCSV_SLEEP();
                    if (++spins > A.spin_limit) {
                        NCSV_ATOMIC_OR(A.flag, tid == 0 ? 1u : 0u);
                        break;
                    }
                }
            }
            NCSV_SYNC();
            NCSV_FENCE_ACQUIRE();
        }

        const BStepD st = steps[rec.step];
        const BEntryD en = entries[rec.entry];
        const uint32_t L = st.L, NL = 1u << L;
        V *g = (V *)A.amp + en.amp_off;
        // element offsets within a state fit 32 bits (at most 32 qubits): the
        // thread's own bits, and the wave-uniform strides of the tile
        // positions above the thread bits
        uint32_t offlo = 0;
#pragma unroll
        for (uint32_t p = 0; p < LOB; p++)
            if (p < L && ((tid >> p) & 1)) offlo |= 1u << st.tpos[p];
        uint32_t wq[HB];
#pragma unroll
        for (uint32_t b = 0; b < HB; b++) wq[b] = (b + LOB < L) ? 1u << st.tpos[LOB + b] : 0u;
        const bool full = L >= LOB;  // every thread holds 2^(L - LOB) amplitudes
        const uint32_t kcnt = full ? 1u << (L - LOB) : 0u;
        constexpr uint32_t CH = sizeof(T) == 4 ? 8u : 4u;
        // the memory offset of the high tile bits k
        auto hibits = [&](uint32_t k) {
            uint32_t h = 0;
#pragma unroll
            for (uint32_t b = 0; b < HB; b++)
                if ((k >> b) & 1) h |= wq[b];
            return h;
        };
        uint32_t op0 = (rec.flags & RF_ERR) ? st.op0 : st.op0c;
        uint32_t op1 = op0 + ((rec.flags & RF_ERR) ? st.nops : st.nopsc);
        if (rec.flags & RF_ONE) {
            // one error drawn in this step: the list keeping only its slot
            const BOpD h = ops[st.var0 + rec.var];
            op0 = h.a;
            op1 = h.a + h.b;
        }
        if (rec.flags & RF_ERR) {
            // this entry's noise codes for the step's slots (read after the first
            // tile's load barrier; the previous record ended with a barrier)
            const uint32_t ns = st.slot_n < SMAX ? st.slot_n : SMAX;
            for (uint32_t t = tid; t < ns; t += BT) scode[t] = errs[en.err0 + st.slot_lo + t];
        }
        // each thread reads back only the LDS slots it wrote, so consecutive
        // tiles need no barrier between the store of one and the load of the next
        for (uint32_t ti = 0; ti < ntl; ti++) {
            const uint64_t tileno = t0 + ti;
            // the outer qubits' offset: tile number bits deposited on omask
            uint32_t base = 0;
            {
                uint64_t tn = tileno;
                for (uint32_t m = st.omask; m; m &= m - 1, tn >>= 1)
                    if (tn & 1) base |= m & (0u - m);
            }

            // the tile's first amplitude; the thread's offset recomputed per tile
            // so that no per-load address stays live across the ops
            const V *gt = g + base;
            uint32_t lo = offlo;
            NCSV_OPAQUE(lo);
            if (st.flags & BF_INIT) {
                V z;
                z.x = (T)0;
                z.y = (T)0;
                if (full) {
                    for (uint32_t k = 0; k < kcnt; k++) tile[st_tid ^ sw<V>(BT * k)] = z;
                } else if (tid < NL) {
                    tile[st_tid] = z;
                }
                if (tid == 0 && tileno == 0) tile[0].x = (T)1;
            } else if (kcnt >= CH) {
                // loads issued in chunks of CH before any of them is waited
                // on; the chunk loop and every address term but the thread's
                // own are wave-uniform
                for (uint32_t k0 = 0; k0 < kcnt; k0 += CH) {
                    const uint32_t h0 = lo | hibits(k0);
                    V v[CH];
#pragma unroll
                    for (uint32_t c = 0; c < CH; c++) v[c] = NCSV_LDG(&gt[h0 | hibits(c)]);
#pragma unroll
                    for (uint32_t c = 0; c < CH; c++) tile[st_tid ^ sw<V>(BT * (k0 + c))] = v[c];
                }
            } else if (full) {
                for (uint32_t k = 0; k < kcnt; k++) tile[st_tid ^ sw<V>(BT * k)] = NCSV_LDG(&gt[lo | hibits(k)]);
            } else if (tid < NL) {
                tile[st_tid] = NCSV_LDG(&gt[lo]);
            }
            NCSV_SYNC();

            // a program's first step starts from |0...0>: tiles other than the
            // first hold zeros throughout it (no op moves amplitudes between
            // tiles), so their gate ops are skipped; reductions still run and
            // write their zero partials
            const bool zero_tile = (st.flags & BF_INIT) && tileno != 0;
            // the product of the step's phases on outer qubits met by this tile,
            // multiplied into the amplitudes at the store
            double tpr = 1.0, tpi = 0.0;
            for (uint32_t oi = op0; oi < op1; oi++) {
                const BOpD op = ops[oi];
                if (zero_tile && (op.kind <= BK_NSLOT || op.kind >= BK_G3)) continue;  // the tile is unchanged: no barrier
                if (op.kind == BK_OPHASE) {
                    if ((base & op.omask) == op.omask) {
                        const double cr = cf[2 * op.coef], ci = cf[2 * op.coef + 1];
                        const double r = tpr * cr - tpi * ci;
                        tpi = tpr * ci + tpi * cr;
                        tpr = r;
                    }
                    continue;  // no pass: no barrier
                }
                switch (op.kind) {
                case BK_U1: {
                    // with an outer condition: the second 2x2 for tiles meeting it; k0 = 1
                    // leaves the other tiles as they are
                    uint32_t c0 = op.coef;
                    if (op.omask) {
                        if ((base & op.omask) == op.omask)
                            c0 += 4;
                        else if (op.k0)
                            continue;  // not this tile: no barrier
                    }
                    const V m0 = coef<V>(cf, c0), m1 = coef<V>(cf, c0 + 1), m2 = coef<V>(cf, c0 + 2), m3 = coef<V>(cf, c0 + 3);
                    const uint32_t sb = sw<V>(1u << op.p0);
                    visit<V, BT>(NL, tid, 1u << op.p0, 0, [&](uint32_t, uint32_t s0) {
                        const uint32_t s1 = s0 ^ sb;
                        const V a = tile[s0], b = tile[s1];
                        tile[s0] = cadd(cm(m0, a), cm(m1, b));
                        tile[s1] = cadd(cm(m2, a), cm(m3, b));
                    });
                    break;
                }
                case BK_U1R: {
                    const T r0 = (T)cf[2 * op.coef], r1 = (T)cf[2 * op.coef + 2], r2 = (T)cf[2 * op.coef + 4],
                            r3 = (T)cf[2 * op.coef + 6];
                    const uint32_t sb = sw<V>(1u << op.p0);
                    visit<V, BT>(NL, tid, 1u << op.p0, 0, [&](uint32_t, uint32_t s0) {
                        const uint32_t s1 = s0 ^ sb;
                        const V a = tile[s0], b = tile[s1];
                        V u, w;
                        u.x = r0 * a.x + r1 * b.x;
                        u.y = r0 * a.y + r1 * b.y;
                        w.x = r2 * a.x + r3 * b.x;
                        w.y = r2 * a.y + r3 * b.y;
                        tile[s0] = u;
                        tile[s1] = w;
                    });
                    break;
                }
                case BK_U1X: {
                    // [[d0, i o1], [i o2, d1]]
                    const T d0 = (T)cf[2 * op.coef], o1 = (T)cf[2 * op.coef + 3], o2 = (T)cf[2 * op.coef + 5],
                            d1 = (T)cf[2 * op.coef + 6];
                    const uint32_t sb = sw<V>(1u << op.p0);
                    visit<V, BT>(NL, tid, 1u << op.p0, 0, [&](uint32_t, uint32_t s0) {
                        const uint32_t s1 = s0 ^ sb;
                        const V a = tile[s0], b = tile[s1];
                        V u, w;
                        u.x = d0 * a.x - o1 * b.y;
                        u.y = d0 * a.y + o1 * b.x;
                        w.x = d1 * b.x - o2 * a.y;
                        w.y = d1 * b.y + o2 * a.x;
                        tile[s0] = u;
                        tile[s1] = w;
                    });
                    break;
                }
                case BK_U2: {
                    const uint32_t b0 = 1u << op.p0, b1 = 1u << op.p1;
                    const uint32_t c1 = sw<V>(b0), c2 = sw<V>(b1), c3 = c1 ^ c2;
                    visit<V, BT>(NL, tid, b0 | b1, 0, [&](uint32_t, uint32_t s) {
                        const uint32_t xs[4] = {s, s ^ c1, s ^ c2, s ^ c3};
                        V v[4];
#pragma unroll
                        for (int t = 0; t < 4; t++) v[t] = tile[xs[t]];
#pragma unroll
                        for (int r2 = 0; r2 < 4; r2++) {
                            V acc = cm(coef<V>(cf, op.coef + 4 * r2), v[0]);
#pragma unroll
                            for (int c = 1; c < 4; c++) acc = cadd(acc, cm(coef<V>(cf, op.coef + 4 * r2 + c), v[c]));
                            tile[xs[r2]] = acc;
                        }
                    });
                    break;
                }
                case BK_G3: {
                    // K unconditional 2x2s on distinct positions, 2^K amplitudes at a time
                    const uint32_t K = op.k1;
                    const uint32_t pb[3] = {1u << op.p0, 1u << op.p1, K == 3 ? 1u << op.p2 : 0u};
                    const uint32_t sp[3] = {sw<V>(pb[0]), sw<V>(pb[1]), sw<V>(pb[2])};
                    V m[3][4];
#pragma unroll
                    for (uint32_t i = 0; i < 3; i++)
#pragma unroll
                        for (uint32_t t = 0; t < 4; t++) m[i][t] = coef<V>(cf, op.coef + 4 * (i < K ? i : 0) + t);
                    auto body = [&](uint32_t s, auto kk) {
                        constexpr uint32_t KK = decltype(kk)::value, NA = 1u << KK;
                        uint32_t xs[NA];
                        V v[NA];
#pragma unroll
                        for (uint32_t a = 0; a < NA; a++) {
                            uint32_t q = s;
#pragma unroll
                            for (uint32_t i = 0; i < KK; i++)
                                if ((a >> i) & 1) q ^= sp[i];
                            xs[a] = q;
                            v[a] = tile[q];
                        }
#pragma unroll
                        for (uint32_t i = 0; i < KK; i++) {
#pragma unroll
                            for (uint32_t a = 0; a < NA; a++) {
                                if ((a >> i) & 1) continue;
                                const V u = v[a], w = v[a | (1u << i)];
                                v[a] = cadd(cm(m[i][0], u), cm(m[i][1], w));
                                v[a | (1u << i)] = cadd(cm(m[i][2], u), cm(m[i][3], w));
                            }
                        }
#pragma unroll
                        for (uint32_t a = 0; a < NA; a++) tile[xs[a]] = v[a];
                    };
                    if (K == 3)
                        visit<V, BT>(NL, tid, pb[0] | pb[1] | pb[2], 0, [&](uint32_t, uint32_t s) { body(s, KTag<3>{}); });
                    else
                        visit<V, BT>(NL, tid, pb[0] | pb[1], 0, [&](uint32_t, uint32_t s) { body(s, KTag<2>{}); });
                    break;
                }
                case BK_U3: {
                    // dense 8x8 on positions p0, p1, p2 (local bit i <-> p_i): per 16 groups
                    // of 8 amplitudes, D (16 x 16) = A (16 x 16) x B (16 x 16) in four MFMAs
                    // (FP32 for complex64, FP64 for complex128), A the real form of the 8x8.
                    // Column 4 kb + q of B <-> amplitude 2q + (kb >> 1), real part for even
                    // kb, imaginary for odd; lane l (group l % 16, q = l / 16) supplies and
                    // receives amplitudes 2q and 2q + 1 of its group and nothing else.
                    // A[i][k] sits in lane i + 16 (k % 4), register k / 4 (both sizes); the
                    // row i of output register v in lane l is 4 q + v (FP32) or 4 v + q
                    // (FP64), so A's rows are ordered to match.  The planner lays each
                    // lane's four A values out at op.a (per variant, in lane order).
                    const uint32_t pb[3] = {1u << op.p0, 1u << op.p1, 1u << op.p2};
                    const uint32_t mask = pb[0] | pb[1] | pb[2];
                    const uint32_t lane = tid & 63, q = lane >> 4, wv = tid >> 6;
                    // two matrices when ops of the block are conditioned on outer qubits:
                    // the second for tiles that meet the condition
                    const uint32_t var = (op.omask && (base & (uint32_t)op.omask) == (uint32_t)op.omask) ? 1u : 0u;
                    const T *ap = (const T *)(cf + 2 * op.a) + 4 * (64 * var + lane);
                    const T av[4] = {ap[0], ap[1], ap[2], ap[3]};
                    auto off = [&](uint32_t a) {
                        return sw<V>(((a & 1) ? pb[0] : 0u) | ((a & 2) ? pb[1] : 0u) | ((a & 4) ? pb[2] : 0u));
                    };
                    const uint32_t sa0 = off(2 * q), sa1 = off(2 * q + 1);
                    const uint32_t ngroups = NL >> 3, gl = (lane & 15) + 16 * wv;
                    using A4 = typename M4<T>::type;
                    auto mult = [&](V in0, V in1) {
                        A4 acc = {(T)0, (T)0, (T)0, (T)0};
                        acc = mfma16(av[0], in0.x, acc);
                        acc = mfma16(av[1], in0.y, acc);
                        acc = mfma16(av[2], in1.x, acc);
                        return mfma16(av[3], in1.y, acc);
                    };
                    auto put = [&](uint32_t sh, const A4 &acc) {
                        V o0, o1;
                        o0.x = acc[0];
                        o0.y = acc[1];
                        o1.x = acc[2];
                        o1.y = acc[3];
                        tile[sh ^ sa0] = o0;
                        tile[sh ^ sa1] = o1;
                    };
                    if (ngroups >= 32 * NW) {
                        // every group valid, two rounds at a time (two independent MFMA
                        // chains); round rd covers groups 16 NW rd + gl, its uniform part
                        // stepped by a masked increment
                        const uint32_t st = sw<V>(deposit(gl, mask, 0));
                        const uint32_t fill = mask | deposit(16 * NW - 1, mask, 0);
                        const uint32_t npairs = ngroups / (32 * NW);
                        uint32_t h = 0;
                        for (uint32_t rp = 0; rp < npairs; rp++) {
                            const uint32_t h1 = ((h | fill) + 1) & ~fill;
                            const uint32_t sh0 = st ^ sw<V>(h), sh1 = st ^ sw<V>(h1);
                            const V a0 = tile[sh0 ^ sa0], a1 = tile[sh0 ^ sa1];
                            const V b0 = tile[sh1 ^ sa0], b1 = tile[sh1 ^ sa1];
                            const A4 ra = mult(a0, a1), rb = mult(b0, b1);
                            put(sh0, ra);
                            put(sh1, rb);
                            h = ((h1 | fill) + 1) & ~fill;
                        }
                    } else {
                        // one round; the same MFMAs in every wavefront
                        const bool valid = gl < ngroups;
                        const uint32_t sh = sw<V>(deposit(valid ? gl : 0u, mask, 0));
                        V in0 = tile[sh ^ sa0], in1 = tile[sh ^ sa1];
                        if (!valid) {
                            in0.x = in0.y = in1.x = in1.y = (T)0;
                        }
                        const A4 r = mult(in0, in1);
                        if (valid) put(sh, r);
                    }
                    break;
                }
                case BK_DM: {
                    if ((base & op.omask) != op.omask) continue;  // not this tile: no barrier
                    const V ph = coef<V>(cf, op.coef);
                    visit<V, BT>(NL, tid, op.lmask, op.lmask, [&](uint32_t, uint32_t s) { tile[s] = cm(ph, tile[s]); });
                    break;
                }
                case BK_PERM1: {
                    if ((base & op.omask) != op.omask) continue;  // not this tile: no barrier
                    const uint32_t bp = 1u << op.p0, sb = sw<V>(bp);
                    const uint32_t k0 = op.k0, k1 = op.k1;
                    visit<V, BT>(NL, tid, op.lmask | bp, op.lmask, [&](uint32_t, uint32_t s0) {
                        const uint32_t s1 = s0 ^ sb;
                        const V a = tile[s0], b = tile[s1];
                        tile[s1] = ipow(k0, a);
                        tile[s0] = ipow(k1, b);
                    });
                    break;
                }
                case BK_SWAP: {
                    const uint32_t b0 = 1u << op.p0, b1 = 1u << op.p1, sb = sw<V>(b0 | b1);
                    visit<V, BT>(NL, tid, b0 | b1, b0, [&](uint32_t, uint32_t s) {
                        const uint32_t s2 = s ^ sb;
                        const V a = tile[s], b = tile[s2];
                        tile[s] = b;
                        tile[s2] = a;
                    });
                    break;
                }
                case BK_NSLOT: {
                    const uint32_t si = op.a - st.slot_lo;
                    const uint32_t code = NCSV_UNIFORM32(si < SMAX ? (uint32_t)scode[si] : (uint32_t)errs[en.err0 + op.a]);
                    if (!code) continue;  // no error at this site: the tile is unchanged, no barrier
                    const uint32_t sb = sw<V>(1u << op.p0);
                    visit<V, BT>(NL, tid, 1u << op.p0, 0, [&](uint32_t, uint32_t s0) {
                        const uint32_t s1 = s0 ^ sb;
                        V a = tile[s0], b = tile[s1];
                        if (code & 2) {
                            b.x = -b.x;
                            b.y = -b.y;
                        }
                        if (code & 1) {
                            const V t = a;
                            a = b;
                            b = t;
                        }
                        tile[s0] = a;
                        tile[s1] = b;
                    });
                    break;
                }
                case BK_MEASURE:
                case BK_REDUCE:
                case BK_COLLAPSE: {
                    const BMGroupD &G = mgroups[op.c];
                    const uint32_t D = G.D, NC = 1u << D;
                    uint32_t ocell = 0;
                    for (uint32_t d = 0; d < D; d++) ocell |= par64(base & G.om[d]) << d;
                    auto cell = [&](uint32_t x) {
                        uint32_t c = ocell;
                        for (uint32_t d = 0; d < D; d++) c ^= par32(x & G.lm[d]) << d;
                        return c;
                    };
                    double v[CMAX];
#pragma unroll
                    for (uint32_t c = 0; c < CMAX; c++) v[c] = 0.0;
                    if (op.kind == BK_COLLAPSE) {
                        // the cell weights from every tile's partials, same order in each tile
                        // 16 loads per tile without per-cell branches (the partials
                        // are padded by 16 doubles; cells from NC on are dropped)
                        const double *pp = A.part + en.part0 + op.a;
                        for (uint32_t t = tid; t < op.b; t += BT) {
                            double x[16];
#pragma unroll
                            for (uint32_t c = 0; c < 16; c++) x[c] = pp[NC * t + c];
#pragma unroll
                            for (uint32_t c = 0; c < 16; c++) v[c] += c < NC ? x[c] : 0.0;
                        }
                    } else {
                        visit<V, BT>(NL, tid, 0, 0, [&](uint32_t x, uint32_t s) {
                            const V a = tile[s];
                            const double w = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
                            const uint32_t cx = cell(x);
#pragma unroll
                            for (uint32_t c = 0; c < 16; c++) v[c] += (c == cx) ? w : 0.0;
                        });
                    }
                    block_sum<NW>(v, NC, red, bc);
                    if (op.kind == BK_REDUCE) {
                        if (tid < NC) A.part[en.part0 + op.a + (uint64_t)NC * tileno + tid] = bc[tid];
                        break;
                    }
                    double W[16];
#pragma unroll
                    for (uint32_t c = 0; c < 16; c++) W[c] = c < NC ? bc[c] : 0.0;
                    // the events in order, each restricting the kept cells (the
                    // same arithmetic in every thread and every tile)
                    uint32_t allow = (1u << NC) - 1, flip = 0;
                    double scale = 1.0, lfac = 1.0;
                    for (uint32_t e = 0; e < G.E; e++) {
                        const uint32_t ed = G.ev[e].d, es = G.ev[e].sign, rp = G.ev[e].rpos;
                        double nn = 0.0, ee = 0.0;
#pragma unroll
                        for (uint32_t c = 0; c < 16; c++) {
                            const double w = ((allow >> c) & 1) ? W[c] : 0.0;
                            nn += w;
                            ee += (((c >> ed) & 1) ^ es) ? -w : w;
                        }
                        const uint32_t ord = en.m0 + G.ev[e].ord;
                        const uint32_t o = decide(uni, forced, ord, nn, ee);
#pragma unroll
                        for (uint32_t c = 0; c < 16; c++)
                            if ((((c >> ed) & 1) ^ es) != o) allow &= ~(1u << c);
                        // complex64: renormalize small branches, the scale kept in FP64 per entry
                        const double kept = 0.5 * (nn + (o ? -ee : ee)) * scale * scale;
                        if (sizeof(T) == 4 && kept > 0.0 && kept < 0.25) {
                            scale /= sqrt(kept);
                            lfac *= sqrt(kept);
                        }
                        if (tid == 0 && tileno == 0) A.outc[ord] = (uint8_t)o;
                        if (rp != 0xff && o) flip ^= 1u << rp;
                    }
                    if (tid == 0 && tileno == 0 && lfac != 1.0) A.lsc[rec.entry] *= lfac;
                    const T sc = (T)scale;
                    visit<V, BT>(NL, tid, 0, 0, [&](uint32_t x, uint32_t s) {
                        V a = tile[s];
                        if (!((allow >> cell(x)) & 1)) {
                            a.x = (T)0;
                            a.y = (T)0;
                        } else if (scale != 1.0) {
                            a.x *= sc;
                            a.y *= sc;
                        }
                        tile[s] = a;
                    });
                    if (flip) {
                        // resets with outcome 1: x -> x ^ flip
                        NCSV_SYNC();
                        const uint32_t sf = sw<V>(flip);
                        visit<V, BT>(NL, tid, flip & (0u - flip), 0, [&](uint32_t, uint32_t s) {
                            const uint32_t s2 = s ^ sf;
                            const V a = tile[s], b = tile[s2];
                            tile[s] = b;
                            tile[s2] = a;
                        });
                    }
                    break;
                }
                case BK_EXPECT: {
                    const uint32_t K = op.b;
                    double v[CMAX];
#pragma unroll
                    for (uint32_t k = 0; k < CMAX; k++) v[k] = 0.0;
                    visit<V, BT>(NL, tid, 0, 0, [&](uint32_t x, uint32_t s) {
                        const V a = tile[s];
                        const double ax = (double)a.x, ay = (double)a.y;
#pragma unroll
                        for (uint32_t k = 0; k < CMAX / 2; k++) {
                            if (k < K) {
                                const BPauliD P = paulis[op.c + k];
                                const V b = tile[s ^ sw<V>(P.xl)];
                                const double bx = (double)b.x, by = (double)b.y;
                                // conj(b) a
                                const double re = bx * ax + by * ay, im = bx * ay - by * ax;
                                const uint32_t pw = (P.ph + 2 * (par32(P.zl & x) ^ par64(P.zo & base))) & 3;
                                double tr = re, ti = im;
                                if (pw == 1) {
                                    tr = -im;
                                    ti = re;
                                } else if (pw == 2) {
                                    tr = -re;
                                    ti = -im;
                                } else if (pw == 3) {
                                    tr = im;
                                    ti = -re;
                                }
                                v[2 * k] += tr;
                                v[2 * k + 1] += ti;
                            }
                        }
                    });
                    block_sum<NW>(v, 2 * K, red, bc);
                    if (tid < 2 * K) A.pexp[en.exp0 + op.a + (uint64_t)2 * K * tileno + tid] = bc[tid];
                    break;
                }
                default: break;
                }
                NCSV_SYNC();
            }

            if (st.flags & BF_STORE) {
                V *gs = g + base;
                uint32_t los = offlo;
                NCSV_OPAQUE(los);  // recomputed here, not held live across the ops
                V pv;
                pv.x = (T)tpr;
                pv.y = (T)tpi;
                auto store = [&](auto tag) {
                    constexpr bool PH = decltype(tag)::value != 0;  // multiply by the tile's phase
                    auto val = [&](uint32_t s) { return PH ? cm(pv, tile[s]) : tile[s]; };
                    if (kcnt >= CH) {
                        for (uint32_t k0 = 0; k0 < kcnt; k0 += CH) {
                            const uint32_t h0 = los | hibits(k0);
#pragma unroll
                            for (uint32_t c = 0; c < CH; c++) NCSV_STG(&gs[h0 | hibits(c)], val(st_tid ^ sw<V>(BT * (k0 + c))));
                        }
                    } else if (full) {
                        for (uint32_t k = 0; k < kcnt; k++) NCSV_STG(&gs[los | hibits(k)], val(st_tid ^ sw<V>(BT * k)));
                    } else if (tid < NL) {
                        NCSV_STG(&gs[los], val(st_tid));
                    }
                };
                if (tpr != 1.0 || tpi != 0.0)
                    store(KTag<1>{});
                else
                    store(KTag<0>{});
            }
        }  // tiles of the item
        if (rec.flags & RF_RELEASE) {
            NCSV_FENCE_RELEASE();
            NCSV_SYNC();
            if (wave == 0) NCSV_ADD_AGENT(A.done + r, tid == 0 ? ntl : 0u);
        } else {
            NCSV_SYNC();
        }
    }
}

}  // namespace kern
}  // namespace batch
}  // namespace ncsv
