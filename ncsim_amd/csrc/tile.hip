// The fused tile sweep for gfx950.
//
// One workgroup per tile of 2^L amplitudes (L = LT + R): 2^LT threads, each
// holding 2^R amplitudes in registers.  Loads and stores use the memory
// layout (thread bits on tile positions 0..LT-1, which start with the
// contiguous low qubits; register bits on positions LT..L-1), so a run of
// 2^c lanes reads 2^c consecutive amplitudes.  The program is a list of
// stages; a stage names its R register positions and its ops, and the tile
// is re-laid out through LDS (XOR-swizzled) when the register positions
// change.  Ops on register bits are unrolled per bit index; diagonal ops and
// controls test the thread's global index bits.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "fuse.hpp"

namespace ncsv {
namespace {

template <typename T> struct VecOf;
template <> struct VecOf<double> { using type = double2; };
template <> struct VecOf<float> { using type = float2; };

struct StageD {
    uint32_t rpos, op_begin, op_end, pad;
};

// One op: 24-byte header and 8 coefficients (U2 keeps its 16 complex
// entries in a side array at index xidx).
template <typename T> struct alignas(8) TileOpD {
    uint8_t kind, j0, j1, k0, k1, pad0, pad1, pad2;
    uint32_t rmask, xidx;
    uint64_t gmask;
    T m[8];
};

struct SweepArgs {
    uint32_t n, nstages;
    uint8_t tq[16];  // global qubit of each tile position
    uint8_t oq[48];  // outer qubits, ascending
};

template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// i^k * a, exact
template <typename V> __device__ __forceinline__ V ipow(int k, V a) {
    V r;
    switch (k & 3) {
    case 0: r = a; break;
    case 1: r.x = -a.y; r.y = a.x; break;
    case 2: r.x = -a.x; r.y = -a.y; break;
    default: r.x = a.y; r.y = -a.x; break;
    }
    return r;
}

// XOR swizzle of the LDS slot: folds the higher slot bits onto the bank bits
// so that lanes differing in high tile positions hit different banks.
template <typename V> __device__ __forceinline__ uint32_t swz(uint32_t t) {
    constexpr int S = sizeof(V) == 16 ? 4 : 5;
    constexpr uint32_t M = (1u << S) - 1;
    return t ^ (((t >> S) ^ (t >> (2 * S))) & M);
}

// thread bits deposited on the positions not in `rpos` (ascending)
template <int L> __device__ __forceinline__ uint32_t thread_local_index(uint32_t tid, uint32_t rpos) {
    uint32_t t = 0;
    int i = 0;
#pragma unroll
    for (int p = 0; p < L; p++) {
        if (!((rpos >> p) & 1)) {
            t |= ((tid >> i) & 1u) << p;
            i++;
        }
    }
    return t;
}

template <typename V, int LT, int R>
__device__ __forceinline__ void exchange(V *lds, V (&v)[1 << R], uint32_t tid, uint32_t from, uint32_t to) {
    constexpr int L = LT + R, NR = 1 << R;
    uint32_t pa[R], pb[R];
    {
        int j = 0;
#pragma unroll
        for (int p = 0; p < L; p++)
            if ((from >> p) & 1) {
                if (j < R) pa[j] = 1u << p;
                j++;
            }
        j = 0;
#pragma unroll
        for (int p = 0; p < L; p++)
            if ((to >> p) & 1) {
                if (j < R) pb[j] = 1u << p;
                j++;
            }
    }
    const uint32_t ta = thread_local_index<L>(tid, from);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint32_t o = ta;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((r >> j) & 1) o |= pa[j];
        lds[swz<V>(o)] = v[r];
    }
    __syncthreads();
    const uint32_t tb = thread_local_index<L>(tid, to);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint32_t o = tb;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((r >> j) & 1) o |= pb[j];
        v[r] = lds[swz<V>(o)];
    }
    __syncthreads();
}

// ------------------------------------------------------------ register ops
template <int J, int NR, typename V, typename T>
__device__ __forceinline__ void u1(V (&v)[NR], const T *m) {
    const T a0 = m[0], a1 = m[1], b0 = m[2], b1 = m[3], c0 = m[4], c1 = m[5], d0 = m[6], d1 = m[7];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (i & (1 << J)) continue;
        const V x = v[i], y = v[i | (1 << J)];
        V p, q;
        p.x = a0 * x.x - a1 * x.y + b0 * y.x - b1 * y.y;
        p.y = a0 * x.y + a1 * x.x + b0 * y.y + b1 * y.x;
        q.x = c0 * x.x - c1 * x.y + d0 * y.x - d1 * y.y;
        q.y = c0 * x.y + c1 * x.x + d0 * y.y + d1 * y.x;
        v[i] = p;
        v[i | (1 << J)] = q;
    }
}

// real 2x2 [[a, b], [c, d]]
template <int J, int NR, typename V, typename T>
__device__ __forceinline__ void u1r(V (&v)[NR], T a, T b, T c, T d) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (i & (1 << J)) continue;
        const V x = v[i], y = v[i | (1 << J)];
        V p, q;
        p.x = a * x.x + b * y.x;
        p.y = a * x.y + b * y.y;
        q.x = c * x.x + d * y.x;
        q.y = c * x.y + d * y.y;
        v[i] = p;
        v[i | (1 << J)] = q;
    }
}

// [[a, i b], [i c, d]] with a, b, c, d real
template <int J, int NR, typename V, typename T>
__device__ __forceinline__ void u1x(V (&v)[NR], T a, T b, T c, T d) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (i & (1 << J)) continue;
        const V x = v[i], y = v[i | (1 << J)];
        V p, q;
        p.x = a * x.x - b * y.y;
        p.y = a * x.y + b * y.x;
        q.x = d * y.x - c * x.y;
        q.y = d * y.y + c * x.x;
        v[i] = p;
        v[i | (1 << J)] = q;
    }
}

template <int A, int B, int NR, typename V, typename T>
__device__ __forceinline__ void u2(V (&v)[NR], const T *m) {
    T mm[32];
#pragma unroll
    for (int t = 0; t < 32; t++) mm[t] = m[t];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (i & ((1 << A) | (1 << B))) continue;
        const int ix[4] = {i, i | (1 << A), i | (1 << B), i | (1 << A) | (1 << B)};
        V x[4], y[4];
#pragma unroll
        for (int t = 0; t < 4; t++) x[t] = v[ix[t]];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            T re = 0, im = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const T mr = mm[2 * (4 * a + b)], mi = mm[2 * (4 * a + b) + 1];
                re += mr * x[b].x - mi * x[b].y;
                im += mr * x[b].y + mi * x[b].x;
            }
            y[a].x = re;
            y[a].y = im;
        }
#pragma unroll
        for (int t = 0; t < 4; t++) v[ix[t]] = y[t];
    }
}

template <int J, int NR, typename V> __device__ __forceinline__ void d1r(V (&v)[NR], V d0, V d1) {
#pragma unroll
    for (int i = 0; i < NR; i++) v[i] = cmul((i >> J) & 1 ? d1 : d0, v[i]);
}

template <int J, int NR, typename V>
__device__ __forceinline__ void perm1(V (&v)[NR], bool ct, uint32_t rmask, int k0, int k1) {
    if (!ct) return;
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (i & (1 << J)) continue;
        if ((i & rmask) != rmask) continue;
        const V x = v[i], y = v[i | (1 << J)];
        v[i] = ipow(k1, y);
        v[i | (1 << J)] = ipow(k0, x);
    }
}

template <int A, int B, int NR, typename V> __device__ __forceinline__ void swapr(V (&v)[NR]) {
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if ((i & (1 << A)) && !(i & (1 << B))) {
            const int j = (i & ~(1 << A)) | (1 << B);
            const V t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
}

#define PAIR_CASES(FN)                                   \
    case 0 * 8 + 1: FN(0, 1); break;                     \
    case 0 * 8 + 2: if constexpr (R > 2) FN(0, 2); break; \
    case 0 * 8 + 3: if constexpr (R > 3) FN(0, 3); break; \
    case 0 * 8 + 4: if constexpr (R > 4) FN(0, 4); break; \
    case 1 * 8 + 2: if constexpr (R > 2) FN(1, 2); break; \
    case 1 * 8 + 3: if constexpr (R > 3) FN(1, 3); break; \
    case 1 * 8 + 4: if constexpr (R > 4) FN(1, 4); break; \
    case 2 * 8 + 3: if constexpr (R > 3) FN(2, 3); break; \
    case 2 * 8 + 4: if constexpr (R > 4) FN(2, 4); break; \
    case 3 * 8 + 4: if constexpr (R > 4) FN(3, 4); break;

#define BIT_CASES(FN)                        \
    case 0: FN(0); break;                    \
    case 1: FN(1); break;                    \
    case 2: if constexpr (R > 2) FN(2); break; \
    case 3: if constexpr (R > 3) FN(3); break; \
    case 4: if constexpr (R > 4) FN(4); break;

template <typename T, int R>
__device__ __forceinline__ void apply_op(const TileOpD<T> &op, typename VecOf<T>::type (&v)[1 << R], uint64_t g,
                                         const T *__restrict__ u2m) {
    using V = typename VecOf<T>::type;
    constexpr int NR = 1 << R;
    const TileOpD<T> *opp = &op;
    const int kind = opp->kind;
    const int j0 = opp->j0;
    switch (kind) {
    case TK_U1: {
#define U1C(J) u1<J, NR>(v, opp->m)
        switch (j0) { BIT_CASES(U1C) }
#undef U1C
        break;
    }
    case TK_U1R: {
        const T a = opp->m[0], b = opp->m[1], c = opp->m[2], d = opp->m[3];
#define U1RC(J) u1r<J, NR>(v, a, b, c, d)
        switch (j0) { BIT_CASES(U1RC) }
#undef U1RC
        break;
    }
    case TK_U1X: {
        const T a = opp->m[0], b = opp->m[1], c = opp->m[2], d = opp->m[3];
#define U1XC(J) u1x<J, NR>(v, a, b, c, d)
        switch (j0) { BIT_CASES(U1XC) }
#undef U1XC
        break;
    }
    case TK_U2: {
        const T *m2 = u2m + 32 * (size_t)opp->xidx;
#define U2C(A, B) u2<A, B, NR>(v, m2)
        switch (j0 * 8 + opp->j1) { PAIR_CASES(U2C) }
#undef U2C
        break;
    }
    case TK_D1R: {
        V d0, d1;
        d0.x = opp->m[0];
        d0.y = opp->m[1];
        d1.x = opp->m[2];
        d1.y = opp->m[3];
#define D1C(J) d1r<J, NR>(v, d0, d1)
        switch (j0) { BIT_CASES(D1C) }
#undef D1C
        break;
    }
    case TK_D1T: {
        V d;
        const int s = (g & opp->gmask) ? 2 : 0;
        d.x = opp->m[s];
        d.y = opp->m[s + 1];
#pragma unroll
        for (int i = 0; i < NR; i++) v[i] = cmul(d, v[i]);
        break;
    }
    case TK_DM: {
        const uint64_t gm = opp->gmask;
        if ((g & gm) != gm) break;
        const uint32_t rm = opp->rmask;
        V c;
        c.x = opp->m[0];
        c.y = opp->m[1];
#pragma unroll
        for (int i = 0; i < NR; i++)
            if (((uint32_t)i & rm) == rm) v[i] = cmul(c, v[i]);
        break;
    }
    case TK_PERM1: {
        const bool ct = (g & opp->gmask) == opp->gmask;
        const uint32_t rm = opp->rmask;
        const int k0 = opp->k0, k1 = opp->k1;
#define P1C(J) perm1<J, NR>(v, ct, rm, k0, k1)
        switch (j0) { BIT_CASES(P1C) }
#undef P1C
        break;
    }
    case TK_SWAPR: {
#define SWC(A, B) swapr<A, B, NR>(v)
        switch (j0 * 8 + opp->j1) { PAIR_CASES(SWC) }
#undef SWC
        break;
    }
    default: break;
    }
}

template <typename T, int LT, int R, bool PIPE>
__global__ __launch_bounds__(1 << LT) void k_tile(typename VecOf<T>::type *__restrict__ amp, SweepArgs a,
                                                  const StageD *__restrict__ st,
                                                  const TileOpD<T> *__restrict__ ops, const T *__restrict__ u2m) {
    using V = typename VecOf<T>::type;
    constexpr int L = LT + R, NR = 1 << R;
    __shared__ V lds[1 << L];
    const uint32_t tid = threadIdx.x;
    const uint64_t b = blockIdx.x;
    uint64_t base = 0;
    const int nout = (int)a.n - L;
    for (int i = 0; i < nout; i++) base |= ((b >> i) & 1ull) << a.oq[i];
    uint64_t gmem = base;
#pragma unroll
    for (int p = 0; p < LT; p++) gmem |= (uint64_t)((tid >> p) & 1u) << a.tq[p];
    uint64_t rq[R];
#pragma unroll
    for (int j = 0; j < R; j++) rq[j] = 1ull << a.tq[LT + j];
    V v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint64_t o = gmem;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((r >> j) & 1) o |= rq[j];
        v[r] = amp[o];
    }
    const uint32_t memr = ((1u << R) - 1) << LT;
    uint32_t cur = memr;
    const uint32_t ns = a.nstages;
    for (uint32_t s = 0; s < ns; s++) {
        const StageD S = st[s];
        if (S.rpos != cur) {
            exchange<V, LT, R>(lds, v, tid, cur, S.rpos);
            cur = S.rpos;
        }
        // the thread's global index bits (tile base + thread positions)
        uint64_t g = base;
        {
            int i = 0;
#pragma unroll
            for (int p = 0; p < L; p++) {
                if (!((cur >> p) & 1)) {
                    g |= (uint64_t)((tid >> i) & 1u) << a.tq[p];
                    i++;
                }
            }
        }
        // PIPE: the next op's record is loaded while the current one is applied
        const uint32_t cnt = S.op_end - S.op_begin;
        if (!PIPE) {
            for (uint32_t o = S.op_begin; o < S.op_end; o++) apply_op<T, R>(ops[o], v, g, u2m);
        } else if (cnt) {
            const TileOpD<T> *op = ops + S.op_begin;
            TileOpD<T> nxt = op[0];
            for (uint32_t o = 0; o < cnt; o++) {
                const TileOpD<T> cur = nxt;
                if (o + 1 < cnt) nxt = op[o + 1];
                apply_op<T, R>(cur, v, g, u2m);
            }
        }
    }
    if (cur != memr) exchange<V, LT, R>(lds, v, tid, cur, memr);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint64_t o = gmem;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((r >> j) & 1) o |= rq[j];
        amp[o] = v[r];
    }
}

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

size_t n_u2(const SweepH &sw) {
    size_t k = 0;
    for (auto &o : sw.ops) k += o.kind == TK_U2;
    return k;
}

// program: stages | ops | U2 coefficients
template <typename T> size_t ops_off(const SweepH &sw) { return align16(sw.stages.size() * sizeof(StageD)); }
template <typename T> size_t u2_off(const SweepH &sw) {
    return ops_off<T>(sw) + align16(sw.ops.size() * sizeof(TileOpD<T>));
}
template <typename T> size_t prog_bytes_t(const SweepH &sw) { return u2_off<T>(sw) + n_u2(sw) * 32 * sizeof(T); }

template <typename T> void pack_t(const SweepH &sw, void *dst) {
    char *p = (char *)dst;
    memset(dst, 0, prog_bytes_t<T>(sw));
    for (size_t s = 0; s < sw.stages.size(); s++) {
        StageD d;
        d.rpos = sw.stages[s].rmask_pos;
        d.op_begin = sw.stages[s].op_begin;
        d.op_end = sw.stages[s].op_end;
        d.pad = 0;
        memcpy(p + s * sizeof(StageD), &d, sizeof(d));
    }
    char *po = (char *)dst + ops_off<T>(sw);
    T *pu = (T *)((char *)dst + u2_off<T>(sw));
    uint32_t nu = 0;
    for (size_t o = 0; o < sw.ops.size(); o++) {
        const TileOpH &h = sw.ops[o];
        TileOpD<T> d;
        memset(&d, 0, sizeof(d));
        d.kind = h.kind;
        d.j0 = h.j0;
        d.j1 = h.j1;
        d.k0 = h.k0;
        d.k1 = h.k1;
        d.rmask = h.rmask;
        d.gmask = h.gmask;
        switch (h.kind) {
        case TK_U1:
            for (int t = 0; t < 4; t++) {
                d.m[2 * t] = (T)h.m[t].real();
                d.m[2 * t + 1] = (T)h.m[t].imag();
            }
            break;
        case TK_U1R:
            for (int t = 0; t < 4; t++) d.m[t] = (T)h.m[t].real();
            break;
        case TK_U1X:
            d.m[0] = (T)h.m[0].real();
            d.m[1] = (T)h.m[1].imag();
            d.m[2] = (T)h.m[2].imag();
            d.m[3] = (T)h.m[3].real();
            break;
        case TK_U2:
            d.xidx = nu;
            for (int t = 0; t < 16; t++) {
                pu[32 * nu + 2 * t] = (T)h.m[t].real();
                pu[32 * nu + 2 * t + 1] = (T)h.m[t].imag();
            }
            nu++;
            break;
        default:
            for (int t = 0; t < 2; t++) {
                d.m[2 * t] = (T)h.m[t].real();
                d.m[2 * t + 1] = (T)h.m[t].imag();
            }
            break;
        }
        memcpy(po + o * sizeof(d), &d, sizeof(d));
    }
}

template <typename T, int LT, int R, bool PIPE> void launch_t(ncsv_state *s, const SweepH &sw, const void *dprog) {
    using V = typename VecOf<T>::type;
    constexpr int L = LT + R;
    SweepArgs a;
    memset(&a, 0, sizeof(a));
    a.n = s->n;
    a.nstages = (uint32_t)sw.stages.size();
    uint64_t in = 0;
    for (int p = 0; p < L; p++) {
        a.tq[p] = (uint8_t)sw.tq[p];
        in |= 1ull << sw.tq[p];
    }
    int k = 0;
    for (uint32_t q = 0; q < s->n; q++)
        if (!((in >> q) & 1)) a.oq[k++] = (uint8_t)q;
    const StageD *st = (const StageD *)dprog;
    const TileOpD<T> *ops = (const TileOpD<T> *)((const char *)dprog + ops_off<T>(sw));
    const T *u2m = (const T *)((const char *)dprog + u2_off<T>(sw));
    const uint64_t tiles = 1ull << (s->n - L);
    KTimer kt(s, "tile", 2.0 * (double)s->amp_bytes());
    hipLaunchKernelGGL((k_tile<T, LT, R, PIPE>), dim3((uint32_t)tiles), dim3(1 << LT), 0, s->stream, (V *)s->amp, a, st,
                       ops, u2m);
    check_hip(hipGetLastError(), "k_tile");
}

}  // namespace

namespace dev {

// complex128: 256 threads x 16 amplitudes (64 KiB of LDS), 16 contiguous
// amplitudes (256 B) per run.  complex64: 256 threads x 32 amplitudes
// (64 KiB), 32 contiguous (256 B).  NCSV_TILE_LT / NCSV_TILE_R / NCSV_TILE_C
// select another built shape (see run_tile_sweep).
TileGeom tile_geom(ncsv_precision prec) {
    TileGeom g;
    if (prec == NCSV_C128) {
        g.LT = 8;
        g.R = 4;
        g.c = 4;
    } else {
        g.LT = 8;
        g.R = 5;
        g.c = 5;
        g.floor_ops = 58;  // half the bytes, FP32 at the FP64 issue rate
    }
    if (const char *e = getenv("NCSV_TILE_LT")) g.LT = (uint32_t)atoi(e);
    if (const char *e = getenv("NCSV_TILE_R")) g.R = (uint32_t)atoi(e);
    if (const char *e = getenv("NCSV_TILE_C")) g.c = (uint32_t)atoi(e);
    if (const char *e = getenv("NCSV_TILE_PIPE")) g.pipe = atoi(e) != 0;
    if (const char *e = getenv("NCSV_FLOOR_OPS")) g.floor_ops = (uint32_t)atoi(e);
    return g;
}

size_t tile_prog_bytes(const SweepH &sw, ncsv_precision prec) {
    return prec == NCSV_C128 ? prog_bytes_t<double>(sw) : prog_bytes_t<float>(sw);
}

void tile_prog_pack(const SweepH &sw, ncsv_precision prec, void *dst) {
    if (prec == NCSV_C128)
        pack_t<double>(sw, dst);
    else
        pack_t<float>(sw, dst);
}

void run_tile_sweep(ncsv_state *s, const SweepH &sw, const void *dprog, const TileGeom &geo) {
#define TILE_SHAPE(T, LT_, R_)                                   \
    if (geo.LT == LT_ && geo.R == R_) {                          \
        if (geo.pipe)                                            \
            launch_t<T, LT_, R_, true>(s, sw, dprog);            \
        else                                                     \
            launch_t<T, LT_, R_, false>(s, sw, dprog);           \
        s->sweeps++;                                             \
        return;                                                  \
    }
    if (s->prec == NCSV_C128) {
        TILE_SHAPE(double, 8, 4)
        TILE_SHAPE(double, 9, 4)
        TILE_SHAPE(double, 9, 3)
    } else {
        TILE_SHAPE(float, 8, 5)
        TILE_SHAPE(float, 9, 5)
        TILE_SHAPE(float, 9, 4)
    }
#undef TILE_SHAPE
    fail(NCSV_INVALID, "unsupported tile shape");
}

}  // namespace dev
}  // namespace ncsv

// ---------------------------------------------------------------------------
// Expectation values of a group of Paulis whose X parts lie inside one tile:
// each workgroup reads its tile once (to registers and LDS); for every Pauli
// the partner amplitude psi(c ^ x) comes from LDS.  Sums are FP64.
namespace ncsv {
namespace {

constexpr int EXP_MAX = 16;  // Paulis per group

struct ExpArgs {
    SweepArgs a;
    uint32_t count, pad;
};

// i^k * a with selects instead of branches (exact)
template <typename V> __device__ __forceinline__ V ipow_sel(int k, V a) {
    V r;
    const bool odd = k & 1, neg = k & 2;
    r.x = odd ? -a.y : a.x;
    r.y = odd ? a.x : a.y;
    if (neg) {
        r.x = -r.x;
        r.y = -r.y;
    }
    return r;
}

// per-group Pauli parameters, in device memory (uniform loads)
struct ExpParams {
    uint64_t z[EXP_MAX];     // Z part on qubits
    uint32_t xloc[EXP_MAX];  // X part on tile positions
    uint32_t ph[EXP_MAX];
};

__device__ __forceinline__ double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <typename T, int LT, int R>
__global__ __launch_bounds__(1 << LT) void k_expect_tile(const typename VecOf<T>::type *__restrict__ amp, ExpArgs e,
                                                         const ExpParams *__restrict__ P, double *part) {
    using V = typename VecOf<T>::type;
    constexpr int L = LT + R, NR = 1 << R, NT = 1 << LT;
    __shared__ V lds[1 << L];
    __shared__ double red[NT / 64][2 * EXP_MAX];
    const uint32_t tid = threadIdx.x;
    const uint64_t b = blockIdx.x;
    uint64_t base = 0;
    const int nout = (int)e.a.n - L;
    for (int i = 0; i < nout; i++) base |= ((b >> i) & 1ull) << e.a.oq[i];
    uint64_t gmem = base;
#pragma unroll
    for (int p = 0; p < LT; p++) gmem |= (uint64_t)((tid >> p) & 1u) << e.a.tq[p];
    uint64_t rq[R];
#pragma unroll
    for (int j = 0; j < R; j++) rq[j] = 1ull << e.a.tq[LT + j];
    V v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        uint64_t o = gmem;
#pragma unroll
        for (int j = 0; j < R; j++)
            if ((r >> j) & 1) o |= rq[j];
        v[r] = amp[o];
        lds[swz<V>(tid | ((uint32_t)r << LT))] = v[r];
    }
    __syncthreads();
    const int lane = tid & 63, wv = tid >> 6;
    const uint32_t cnt = e.count;
    for (uint32_t p = 0; p < cnt; p++) {
        const uint32_t xl = P->xloc[p];
        const uint64_t zm = P->z[p];
        const int ph = (int)P->ph[p];
        // Z parity of the thread part, then per register the register bits
        const int kt = ph + 2 * __popcll(zm & gmem);
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            const uint32_t t = tid | ((uint32_t)r << LT);
            const V w = lds[swz<V>(t ^ xl)];  // psi(c ^ x); psi(c) when x = 0
            int kr = 0;
#pragma unroll
            for (int j = 0; j < R; j++)
                if ((r >> j) & 1) kr += (zm & rq[j]) ? 1 : 0;
            const V tv = ipow_sel((kt + 2 * kr) & 3, v[r]);
            const double tr = (double)tv.x, ti = (double)tv.y, xr = (double)w.x, xi = (double)w.y;
            sr += xr * tr + xi * ti;  // conj(psi(c^x)) i^k psi(c)
            si += xr * ti - xi * tr;
        }
        sr = wsum(sr);
        si = wsum(si);
        if (lane == 0) {
            red[wv][2 * p] = sr;
            red[wv][2 * p + 1] = si;
        }
    }
    __syncthreads();
    if (tid < 2 * e.count) {
        double s = 0.0;
        for (int i = 0; i < NT / 64; i++) s += red[i][tid];
        part[b * (2 * EXP_MAX) + tid] = s;
    }
}

// sums nb x (2 EXP_MAX) partials into out[2 count], one block per value
__global__ void k_expect_finish(const double *part, uint64_t nb, uint32_t count, double *out) {
    __shared__ double red[256];
    const uint32_t k = blockIdx.x;
    double s = 0.0;
    for (uint64_t i = threadIdx.x; i < nb; i += blockDim.x) s += part[i * (2 * EXP_MAX) + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

template <typename T, int LT, int R>
void expect_group_t(ncsv_state *s, const std::vector<uint32_t> &tq, const uint64_t *x, const uint64_t *z,
                    const uint32_t *ph, const std::vector<uint32_t> &idx, double *out) {
    using V = typename VecOf<T>::type;
    constexpr int L = LT + R;
    ExpArgs e;
    memset(&e, 0, sizeof(e));
    e.a.n = s->n;
    uint64_t in = 0;
    std::vector<int> pos_of(64, -1);
    for (int p = 0; p < L; p++) {
        e.a.tq[p] = (uint8_t)tq[p];
        in |= 1ull << tq[p];
        pos_of[tq[p]] = p;
    }
    int k = 0;
    for (uint32_t q = 0; q < s->n; q++)
        if (!((in >> q) & 1)) e.a.oq[k++] = (uint8_t)q;
    e.count = (uint32_t)idx.size();
    ExpParams P;
    memset(&P, 0, sizeof(P));
    for (size_t i = 0; i < idx.size(); i++) {
        uint32_t xl = 0;
        for (uint64_t a = x[idx[i]]; a; a &= a - 1) xl |= 1u << pos_of[__builtin_ctzll(a)];
        P.xloc[i] = xl;
        P.z[i] = z[idx[i]];
        P.ph[i] = ph[idx[i]] & 3;
    }
    dev::ensure_tab(s, sizeof(ExpParams));
    check_hip(hipMemcpyAsync(s->d_tab, &P, sizeof(P), hipMemcpyHostToDevice, s->stream), "upload Paulis");
    const uint64_t tiles = 1ull << (s->n - L);
    dev::ensure_part(s, tiles * 2 * EXP_MAX + 2 * EXP_MAX + 64);
    double *fin = s->d_part + tiles * 2 * EXP_MAX;
    {
        KTimer kt(s, "expect_tile", (double)s->amp_bytes());
        hipLaunchKernelGGL((k_expect_tile<T, LT, R>), dim3((uint32_t)tiles), dim3(1 << LT), 0, s->stream,
                           (const V *)s->amp, e, (const ExpParams *)s->d_tab, s->d_part);
        check_hip(hipGetLastError(), "k_expect_tile");
    }
    hipLaunchKernelGGL(k_expect_finish, dim3(2 * e.count), dim3(256), 0, s->stream, (const double *)s->d_part, tiles,
                       e.count, fin);
    check_hip(hipGetLastError(), "k_expect_finish");
    check_hip(hipMemcpyAsync(s->h_part, fin, 2 * e.count * sizeof(double), hipMemcpyDeviceToHost, s->stream),
              "readback");
    dev::sync(s);
    for (size_t i = 0; i < idx.size(); i++) {
        out[2 * idx[i]] = s->h_part[2 * i];
        out[2 * idx[i] + 1] = s->h_part[2 * i + 1];
    }
}


__global__ void k_rowsum(const double *part, uint32_t cols, double *out) {
    __shared__ double red[256];
    const uint64_t row = blockIdx.x;
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < cols; i += blockDim.x) s += part[row * cols + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[row] = red[0];
}

}  // namespace

namespace dev {

void rowsum(ncsv_state *s, const double *part, uint32_t rows, uint32_t cols, double *host_out) {
    if (rows > 64) fail(NCSV_INVALID, "rowsum: too many rows");
    double *fin = nullptr;
    check_hip(hipMallocAsync((void **)&fin, rows * sizeof(double), s->stream), "hipMallocAsync(rowsum)");
    hipLaunchKernelGGL(k_rowsum, dim3(rows), dim3(256), 0, s->stream, part, cols, fin);
    check_hip(hipGetLastError(), "k_rowsum");
    check_hip(hipMemcpyAsync(s->h_part, fin, rows * sizeof(double), hipMemcpyDeviceToHost, s->stream), "readback");
    check_hip(hipFreeAsync(fin, s->stream), "hipFreeAsync");
    sync(s);
    memcpy(host_out, s->h_part, rows * sizeof(double));
}

}  // namespace dev

namespace dev {

// Groups the Paulis so each group's X qubits fit in one tile (first fit),
// then one read of the state per group.  Returns false (nothing done) when
// the state is smaller than a tile or some Pauli's X part does not fit.
bool expectations_tiled(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                        double *out) {
    const TileGeom geo = tile_geom(s->prec);
    const uint32_t L = geo.L(), c = geo.c;
    if (s->n < L || s->n > 40) return false;
    const uint64_t low = (1ull << c) - 1;
    struct Group {
        uint64_t T;
        std::vector<uint32_t> idx;
    };
    std::vector<Group> groups;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t need = x[i] & ~low;
        if ((uint32_t)__builtin_popcountll(need) > L - c) return false;
        bool placed = false;
        for (auto &g : groups) {
            if (g.idx.size() < (size_t)EXP_MAX && (uint32_t)__builtin_popcountll(g.T | need) <= L) {
                g.T |= need;
                g.idx.push_back(i);
                placed = true;
                break;
            }
        }
        if (!placed) groups.push_back({low | need, {i}});
    }
    for (auto &g : groups) {
        // fill the tile with the lowest unused qubits
        for (uint32_t q = 0; q < s->n && (uint32_t)__builtin_popcountll(g.T) < L; q++) g.T |= 1ull << q;
        std::vector<uint32_t> tq;
        for (uint32_t q = 0; q < s->n; q++)
            if ((g.T >> q) & 1) tq.push_back(q);
        if (s->prec == NCSV_C128)
            expect_group_t<double, 8, 4>(s, tq, x, z, ph, g.idx, out);
        else
            expect_group_t<float, 8, 5>(s, tq, x, z, ph, g.idx, out);
    }
    return true;
}

}  // namespace dev
}  // namespace ncsv
