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
//   (dst = lane 0's v), V2<T>::type (float2 / double2).
#pragma once

#include "batch.hpp"

namespace ncsv {
namespace batch {
namespace kern {


constexpr uint32_t BT = 512;        // threads per workgroup
constexpr uint32_t NW = BT / 64;    // wavefronts per workgroup
constexpr uint32_t LOB = 9;         // log2(BT): tile bits taken from the thread index
constexpr uint32_t CMAX = 18;       // values per block reduction (9 Paulis)


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
NCSV_DEV inline uint32_t par32(uint32_t v) { return (uint32_t)__builtin_popcount(v) & 1u; }
NCSV_DEV inline uint32_t par64(uint64_t v) { return (uint32_t)__builtin_popcountll(v) & 1u; }

// Sums v[0..cnt) over the workgroup (fixed order) into bc[0..cnt), which
// stays valid until the next barrier-separated reduction.
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
NCSV_DEV inline uint32_t decide(const BArgs &A, uint32_t ordinal, double nn, double e) {
    double p_plus = (1.0 + e / nn) / 2.0;
    if (!(p_plus >= 0.0)) p_plus = 0.0;
    if (p_plus > 1.0) p_plus = 1.0;
    const uint8_t f = A.forced ? A.forced[ordinal] : (uint8_t)0xff;
    if (f != 0xff) return f & 1u;
    if (p_plus > 1.0 - 1e-12) return 0;
    if (p_plus < 1e-12) return 1;
    return A.uni[ordinal] >= p_plus ? 1u : 0u;
}

template <typename T, uint32_t LB>
NCSV_KERNEL void k_batch(BArgs A) {
    using V = typename V2<T>::type;
    NCSV_SHARED V tile[1u << LB];
    NCSV_SHARED double red[NW * CMAX];
    NCSV_SHARED double bc[CMAX];
    NCSV_SHARED unsigned long long s_claim;
    const uint32_t tid = NCSV_TID;
    // Queue traffic (claim, wait, completion) is done by the whole first
    // wavefront under a wave-uniform branch, and the claim is read back as a
    // wave-uniform value: a lane-divergent branch next to the loop's
    // barriers lets the compiler split one lane from the rest of its wave
    // across iterations, which deadlocks the workgroup.
    const uint32_t wave = NCSV_UNIFORM32(tid >> 6);
    for (;;) {
        if (wave == 0) {
            const unsigned long long old = NCSV_ATOMIC_ADD(A.next, tid == 0 ? 1ull : 0ull);
            NCSV_LANE0_STORE(s_claim, old);
        }
        NCSV_SYNC();
        const uint64_t item = NCSV_UNIFORM64(s_claim);
        if (item >= A.total) break;
        uint32_t r = A.rec_blk[item >> 4];
        while (A.recs[r + 1].first <= item) r++;
        const BRecD rec = A.recs[r];
        const uint64_t tileno = item - rec.first;
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

        const BStepD st = A.steps[rec.step];
        const BEntryD en = A.entries[rec.entry];
        const uint32_t L = st.L, NL = 1u << L;
        const uint8_t *qt = A.qtab + st.qoff;
        V *g = (V *)A.amp + en.amp_off;
        uint64_t base = 0;
        for (uint32_t b = 0; b + L < st.n; b++)
            if ((tileno >> b) & 1) base |= 1ull << qt[L + b];
        uint64_t offlo = 0;
        for (uint32_t p = 0; p < L && p < LOB; p++)
            if ((tid >> p) & 1) offlo |= 1ull << qt[p];

        if (st.flags & BF_INIT) {
            for (uint32_t x = tid; x < NL; x += BT) {
                V v;
                v.x = (x == 0 && tileno == 0) ? (T)1 : (T)0;
                v.y = (T)0;
                tile[x] = v;
            }
        } else {
            uint32_t k = 0;
            for (uint32_t x = tid; x < NL; x += BT, k++) {
                uint64_t off = offlo;
                for (uint32_t b = 0; b + LOB < L; b++)
                    if ((k >> b) & 1) off |= 1ull << qt[LOB + b];
                tile[x] = g[base + off];
            }
        }
        NCSV_SYNC();

        for (uint32_t oi = st.op0; oi < st.op0 + st.nops; oi++) {
            const BOpD op = A.ops[oi];
            switch (op.kind) {
            case BK_U1: {
                const V m0 = coef<V>(A.coef, op.coef), m1 = coef<V>(A.coef, op.coef + 1),
                        m2 = coef<V>(A.coef, op.coef + 2), m3 = coef<V>(A.coef, op.coef + 3);
                const uint32_t p = op.p0;
                for (uint32_t j = tid; j < NL / 2; j += BT) {
                    const uint32_t x0 = ins0(j, p), x1 = x0 | (1u << p);
                    const V a = tile[x0], b = tile[x1];
                    tile[x0] = cadd(cm(m0, a), cm(m1, b));
                    tile[x1] = cadd(cm(m2, a), cm(m3, b));
                }
                break;
            }
            case BK_U1R: {
                const T r0 = (T)A.coef[2 * op.coef], r1 = (T)A.coef[2 * op.coef + 2], r2 = (T)A.coef[2 * op.coef + 4],
                        r3 = (T)A.coef[2 * op.coef + 6];
                const uint32_t p = op.p0;
                for (uint32_t j = tid; j < NL / 2; j += BT) {
                    const uint32_t x0 = ins0(j, p), x1 = x0 | (1u << p);
                    const V a = tile[x0], b = tile[x1];
                    V u, w;
                    u.x = r0 * a.x + r1 * b.x;
                    u.y = r0 * a.y + r1 * b.y;
                    w.x = r2 * a.x + r3 * b.x;
                    w.y = r2 * a.y + r3 * b.y;
                    tile[x0] = u;
                    tile[x1] = w;
                }
                break;
            }
            case BK_U1X: {
                // [[d0, i o1], [i o2, d1]]
                const T d0 = (T)A.coef[2 * op.coef], o1 = (T)A.coef[2 * op.coef + 3], o2 = (T)A.coef[2 * op.coef + 5],
                        d1 = (T)A.coef[2 * op.coef + 6];
                const uint32_t p = op.p0;
                for (uint32_t j = tid; j < NL / 2; j += BT) {
                    const uint32_t x0 = ins0(j, p), x1 = x0 | (1u << p);
                    const V a = tile[x0], b = tile[x1];
                    V u, w;
                    u.x = d0 * a.x - o1 * b.y;
                    u.y = d0 * a.y + o1 * b.x;
                    w.x = d1 * b.x - o2 * a.y;
                    w.y = d1 * b.y + o2 * a.x;
                    tile[x0] = u;
                    tile[x1] = w;
                }
                break;
            }
            case BK_U2: {
                V m[16];
#pragma unroll
                for (int t = 0; t < 16; t++) m[t] = coef<V>(A.coef, op.coef + t);
                const uint32_t p0 = op.p0, p1 = op.p1;
                for (uint32_t j = tid; j < NL / 4; j += BT) {
                    const uint32_t x = ins0(ins0(j, p0), p1);
                    const uint32_t xs[4] = {x, x | (1u << p0), x | (1u << p1), x | (1u << p0) | (1u << p1)};
                    V v[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) v[t] = tile[xs[t]];
#pragma unroll
                    for (int r2 = 0; r2 < 4; r2++) {
                        V s = cm(m[4 * r2], v[0]);
#pragma unroll
                        for (int c2 = 1; c2 < 4; c2++) s = cadd(s, cm(m[4 * r2 + c2], v[c2]));
                        tile[xs[r2]] = s;
                    }
                }
                break;
            }
            case BK_DM: {
                if ((base & op.omask) != op.omask) break;
                const V ph = coef<V>(A.coef, op.coef);
                const uint32_t cnt = NL >> __builtin_popcount(op.lmask);
                for (uint32_t j = tid; j < cnt; j += BT) {
                    const uint32_t x = deposit(j, op.lmask, op.lmask);
                    tile[x] = cm(ph, tile[x]);
                }
                break;
            }
            case BK_PERM1: {
                if ((base & op.omask) != op.omask) break;
                const uint32_t p = op.p0;
                const uint32_t cnt = NL >> (1 + __builtin_popcount(op.lmask));
                for (uint32_t j = tid; j < cnt; j += BT) {
                    const uint32_t x0 = deposit(j, op.lmask | (1u << p), op.lmask), x1 = x0 | (1u << p);
                    const V a = tile[x0], b = tile[x1];
                    tile[x1] = ipow(op.k0, a);
                    tile[x0] = ipow(op.k1, b);
                }
                break;
            }
            case BK_SWAP: {
                const uint32_t b0 = 1u << op.p0, b1 = 1u << op.p1;
                for (uint32_t j = tid; j < NL / 4; j += BT) {
                    const uint32_t x = deposit(j, b0 | b1, b0), y = x ^ (b0 | b1);
                    const V a = tile[x], b = tile[y];
                    tile[x] = b;
                    tile[y] = a;
                }
                break;
            }
            case BK_NSLOT: {
                const uint32_t code = A.errs[en.err0 + op.a];
                if (!code) break;
                const uint32_t p = op.p0;
                for (uint32_t j = tid; j < NL / 2; j += BT) {
                    const uint32_t x0 = ins0(j, p), x1 = x0 | (1u << p);
                    V a = tile[x0], b = tile[x1];
                    if (code & 2) {
                        b.x = -b.x;
                        b.y = -b.y;
                    }
                    if (code & 1) {
                        const V t = a;
                        a = b;
                        b = t;
                    }
                    tile[x0] = a;
                    tile[x1] = b;
                }
                break;
            }
            case BK_MEASURE:
            case BK_REDUCE:
            case BK_COLLAPSE: {
                const uint32_t sgn = op.k0 ^ par64(base & op.omask);
                double v[CMAX];
                if (op.kind == BK_COLLAPSE) {
                    // the measured weights from every tile's partials, same order in each tile
                    double a0 = 0.0, a1 = 0.0;
                    const double *pp = A.part + en.part0 + op.a;
                    for (uint32_t t = tid; t < op.b; t += BT) {
                        a0 += pp[2 * t];
                        a1 += pp[2 * t + 1];
                    }
                    v[0] = a0;
                    v[1] = a1;
                } else {
                    double a0 = 0.0, a1 = 0.0;
                    for (uint32_t x = tid; x < NL; x += BT) {
                        const V a = tile[x];
                        const double w = (double)a.x * (double)a.x + (double)a.y * (double)a.y;
                        a0 += w;
                        a1 += (par32(x & op.lmask) ^ sgn) ? -w : w;
                    }
                    v[0] = a0;
                    v[1] = a1;
                }
                block_sum(v, 2, red, bc);
                const double snn = bc[0], se = bc[1];
                if (op.kind == BK_REDUCE) {
                    if (tid == 0) {
                        A.part[en.part0 + op.a + 2 * tileno] = snn;
                        A.part[en.part0 + op.a + 2 * tileno + 1] = se;
                    }
                    break;
                }
                const uint32_t ord = en.m0 + op.c;
                const uint32_t o = decide(A, ord, snn, se);
                const double kept = 0.5 * (snn + (o ? -se : se));
                // complex64: renormalize small branches, the scale kept in FP64 per entry
                double sc = 1.0;
                if (sizeof(T) == 4 && kept > 0.0 && kept < 0.25) sc = 1.0 / sqrt(kept);
                if (tid == 0 && tileno == 0) {
                    A.outc[ord] = (uint8_t)o;
                    if (sc != 1.0) A.lsc[rec.entry] *= sqrt(kept);
                }
                const T s = (T)sc;
                if (op.p0 == 0xff) {
                    for (uint32_t x = tid; x < NL; x += BT) {
                        V a = tile[x];
                        if ((par32(x & op.lmask) ^ sgn) != o) {
                            a.x = (T)0;
                            a.y = (T)0;
                        } else if (sc != 1.0) {
                            a.x *= s;
                            a.y *= s;
                        }
                        tile[x] = a;
                    }
                } else {
                    // reset: keep the outcome's half and move it to bit p0 = 0
                    const uint32_t p = op.p0;
                    for (uint32_t j = tid; j < NL / 2; j += BT) {
                        const uint32_t x0 = ins0(j, p), x1 = x0 | (1u << p);
                        V a = o ? tile[x1] : tile[x0];
                        if (sc != 1.0) {
                            a.x *= s;
                            a.y *= s;
                        }
                        V z;
                        z.x = (T)0;
                        z.y = (T)0;
                        tile[x0] = a;
                        tile[x1] = z;
                    }
                }
                break;
            }
            case BK_EXPECT: {
                const uint32_t K = op.b;
                double v[CMAX];
#pragma unroll
                for (uint32_t k = 0; k < CMAX; k++) v[k] = 0.0;
                for (uint32_t x = tid; x < NL; x += BT) {
                    const V a = tile[x];
                    const double ax = (double)a.x, ay = (double)a.y;
#pragma unroll
                    for (uint32_t k = 0; k < CMAX / 2; k++) {
                        if (k < K) {
                            const BPauliD P = A.paulis[op.c + k];
                            const V b = tile[x ^ P.xl];
                            const double bx = (double)b.x, by = (double)b.y;
                            // conj(b) a
                            double re = bx * ax + by * ay, im = bx * ay - by * ax;
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
                }
                block_sum(v, 2 * K, red, bc);
                if (tid < 2 * K) A.pexp[en.exp0 + op.a + (uint64_t)2 * K * tileno + tid] = bc[tid];
                break;
            }
            default: break;
            }
            NCSV_SYNC();
        }

        if (st.flags & BF_STORE) {
            uint32_t k = 0;
            for (uint32_t x = tid; x < NL; x += BT, k++) {
                uint64_t off = offlo;
                for (uint32_t b = 0; b + LOB < L; b++)
                    if ((k >> b) & 1) off |= 1ull << qt[LOB + b];
                g[base + off] = tile[x];
            }
        }
        NCSV_FENCE_RELEASE();
        NCSV_SYNC();
        if (wave == 0) NCSV_ADD_AGENT(A.done + r, tid == 0 ? 1u : 0u);
    }
}

}  // namespace kern
}  // namespace batch
}  // namespace ncsv
