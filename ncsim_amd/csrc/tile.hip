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

#include <cstring>

#include "fuse.hpp"

namespace ncsv {
namespace {

template <typename T> struct VecOf;
template <> struct VecOf<double> { using type = double2; };
template <> struct VecOf<float> { using type = float2; };

struct StageD {
    uint32_t rpos, op_begin, op_end, pad;
};

template <typename T> struct alignas(16) TileOpD {
    uint8_t kind, j0, j1, k0, k1, pad0, pad1, pad2;
    uint32_t rmask, pad3;
    uint64_t gmask;
    T m[32];  // interleaved complex
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
__device__ __forceinline__ void apply_op(const TileOpD<T> *__restrict__ opp, typename VecOf<T>::type (&v)[1 << R],
                                         uint64_t g) {
    using V = typename VecOf<T>::type;
    constexpr int NR = 1 << R;
    const int kind = opp->kind;
    const int j0 = opp->j0;
    switch (kind) {
    case TK_U1: {
#define U1C(J) u1<J, NR>(v, opp->m)
        switch (j0) { BIT_CASES(U1C) }
#undef U1C
        break;
    }
    case TK_U2: {
#define U2C(A, B) u2<A, B, NR>(v, opp->m)
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

template <typename T, int LT, int R>
__global__ __launch_bounds__(1 << LT) void k_tile(typename VecOf<T>::type *__restrict__ amp, SweepArgs a,
                                                  const StageD *__restrict__ st,
                                                  const TileOpD<T> *__restrict__ ops) {
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
        for (uint32_t o = S.op_begin; o < S.op_end; o++) apply_op<T, R>(ops + o, v, g);
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

template <typename T> size_t prog_bytes_t(const SweepH &sw) {
    return align16(sw.stages.size() * sizeof(StageD)) + sw.ops.size() * sizeof(TileOpD<T>);
}

template <typename T> void pack_t(const SweepH &sw, void *dst) {
    char *p = (char *)dst;
    for (size_t s = 0; s < sw.stages.size(); s++) {
        StageD d;
        d.rpos = sw.stages[s].rmask_pos;
        d.op_begin = sw.stages[s].op_begin;
        d.op_end = sw.stages[s].op_end;
        d.pad = 0;
        memcpy(p + s * sizeof(StageD), &d, sizeof(d));
    }
    p += align16(sw.stages.size() * sizeof(StageD));
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
        for (int t = 0; t < 16; t++) {
            d.m[2 * t] = (T)h.m[t].real();
            d.m[2 * t + 1] = (T)h.m[t].imag();
        }
        memcpy(p + o * sizeof(d), &d, sizeof(d));
    }
}

template <typename T, int LT, int R> void launch_t(ncsv_state *s, const SweepH &sw, const void *dprog) {
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
    const TileOpD<T> *ops = (const TileOpD<T> *)((const char *)dprog + align16(sw.stages.size() * sizeof(StageD)));
    const uint64_t tiles = 1ull << (s->n - L);
    KTimer kt(s, "tile", 2.0 * (double)s->amp_bytes());
    hipLaunchKernelGGL((k_tile<T, LT, R>), dim3((uint32_t)tiles), dim3(1 << LT), 0, s->stream, (V *)s->amp, a, st,
                       ops);
    check_hip(hipGetLastError(), "k_tile");
}

}  // namespace

namespace dev {

// complex128: 256 threads x 16 amplitudes (64 KiB of LDS), 16 contiguous
// amplitudes (256 B) per run.  complex64: 256 threads x 32 amplitudes
// (64 KiB), 32 contiguous (256 B).
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
    }
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
    if (s->prec == NCSV_C128) {
        if (geo.LT != 8 || geo.R != 4) fail(NCSV_INVALID, "unsupported tile shape");
        launch_t<double, 8, 4>(s, sw, dprog);
    } else {
        if (geo.LT != 8 || geo.R != 5) fail(NCSV_INVALID, "unsupported tile shape");
        launch_t<float, 8, 5>(s, sw, dprog);
    }
    s->sweeps++;
}

}  // namespace dev
}  // namespace ncsv
