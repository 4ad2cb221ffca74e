// HIP kernels for gfx950: one device sweep per lowered gate (dense k-qubit,
// diagonal, permutation, Pauli), FP64 reductions (norm, expectations, inner),
// fused collapse + weight, tensor products and prefix-sum sampling.  Each
// kernel is templated on the storage type: double2 (complex128) or float2
// (complex64, reductions still in FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "jit.hpp"
#include "ncsv_internal.hpp"

namespace ncsv {
namespace {

constexpr int TPB = 256;  // threads per block for the streaming kernels
constexpr int SAMPLE_BLOCK = 4096;  // vector.rs:192 (BLOCK) — sampling blocks

template <typename T> struct VecOf;
template <> struct VecOf<double> { using type = double2; };
template <> struct VecOf<float> { using type = float2; };

template <typename V> __device__ __forceinline__ V cmulv(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename V> __device__ __forceinline__ V caddv(V a, V b) {
    V r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename V> __device__ __forceinline__ bool czero(V a) { return a.x == 0 && a.y == 0; }

// i^k * a, exactly (k mod 4): the products by 0 and +-1 of the reference's
// complex multiply are exact, so this equals it up to the sign of zero.
template <typename V> __device__ __forceinline__ V ipow_mul(int k, V a) {
    V r;
    switch (k & 3) {
    case 0: r = a; break;
    case 1: r.x = -a.y; r.y = a.x; break;
    case 2: r.x = -a.x; r.y = -a.y; break;
    default: r.x = a.y; r.y = -a.x; break;
    }
    return r;
}

__device__ __forceinline__ uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// Insert zero bits at the ascending positions sq[0..k) of a fiber number.
__device__ __forceinline__ uint64_t deposit_zeros(uint64_t t, const int *sq, int k) {
    for (int j = 0; j < k; j++) {
        const uint64_t lo = t & ((1ull << sq[j]) - 1);
        t = ((t >> sq[j]) << (sq[j] + 1)) | lo;
    }
    return t;
}

// Small op tables ride in the kernel arguments.
template <typename T> struct OpArgs {
    int k;
    int sq[12];          // gate qubits, ascending
    uint64_t off[16];    // offset of local state x (bit j on qubits[j]), x < 16
    T re[16], im[16];    // dense 2^k x 2^k for k <= 2; phases for diag/perm (k <= 4)
    uint32_t img[16];    // perm images
    uint32_t unit;       // bit x: entry x is exactly 1 (skip / plain move)
};

// ---------------------------------------------------------------- dense
template <typename T, int K>
__global__ __launch_bounds__(TPB) void k_dense_small(typename VecOf<T>::type *a, uint64_t nf, OpArgs<T> g) {
    using V = typename VecOf<T>::type;
    constexpr int D = 1 << K;
    for (uint64_t t = gtid(); t < nf; t += gstride()) {
        const uint64_t base = deposit_zeros(t, g.sq, K);
        V v[D];
#pragma unroll
        for (int b = 0; b < D; b++) v[b] = a[base + g.off[b]];
#pragma unroll
        for (int r = 0; r < D; r++) {
            V acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int c = 0; c < D; c++) {
                V m;
                m.x = g.re[r * D + c];
                m.y = g.im[r * D + c];
                acc = caddv(acc, cmulv(m, v[c]));
            }
            a[base + g.off[r]] = acc;
        }
    }
}

// k = 3..5: matrix and offsets staged in LDS; one fiber per thread.
template <typename T, int K>
__global__ __launch_bounds__(TPB) void k_dense_mid(typename VecOf<T>::type *a, uint64_t nf, OpArgs<T> g,
                                                   const typename VecOf<T>::type *mat, const uint64_t *offs) {
    using V = typename VecOf<T>::type;
    constexpr int D = 1 << K;
    __shared__ V sm[D * D];
    __shared__ uint64_t so[D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sm[i] = mat[i];
    for (int i = threadIdx.x; i < D; i += blockDim.x) so[i] = offs[i];
    __syncthreads();
    for (uint64_t t = gtid(); t < nf; t += gstride()) {
        const uint64_t base = deposit_zeros(t, g.sq, K);
        V v[D];
#pragma unroll
        for (int b = 0; b < D; b++) v[b] = a[base + so[b]];
        for (int r = 0; r < D; r++) {
            V acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int c = 0; c < D; c++) acc = caddv(acc, cmulv(sm[r * D + c], v[c]));
            a[base + so[r]] = acc;
        }
    }
}

// k > 5: out-of-place, one output amplitude per thread.
template <typename T>
__global__ __launch_bounds__(TPB) void k_dense_big(const typename VecOf<T>::type *in, typename VecOf<T>::type *out,
                                                   uint64_t N, int k, const int *q, const typename VecOf<T>::type *mat) {
    using V = typename VecOf<T>::type;
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        uint64_t base = i;
        uint32_t row = 0;
        for (int j = 0; j < k; j++) {
            row |= (uint32_t)((i >> q[j]) & 1) << j;
            base &= ~(1ull << q[j]);
        }
        const uint32_t D = 1u << k;
        V acc;
        acc.x = 0;
        acc.y = 0;
        for (uint32_t c = 0; c < D; c++) {
            uint64_t idx = base;
            for (int j = 0; j < k; j++) idx |= (uint64_t)((c >> j) & 1) << q[j];
            acc = caddv(acc, cmulv(mat[(uint64_t)row * D + c], in[idx]));
        }
        out[i] = acc;
    }
}

// ------------------------------------------------------------ diagonal
// Touches only the amplitudes whose local state has a non-unit phase.
template <typename T>
__global__ __launch_bounds__(TPB) void k_diag(typename VecOf<T>::type *a, uint64_t nf, OpArgs<T> g) {
    using V = typename VecOf<T>::type;
    const int D = 1 << g.k;
    for (uint64_t t = gtid(); t < nf; t += gstride()) {
        const uint64_t base = deposit_zeros(t, g.sq, g.k);
        for (int x = 0; x < D; x++) {
            if ((g.unit >> x) & 1) continue;
            V c;
            c.x = g.re[x];
            c.y = g.im[x];
            const uint64_t i = base + g.off[x];
            const V v = a[i];
            if (!czero(v)) a[i] = cmulv(c, v);
        }
    }
}

// --------------------------------------------------------- permutation
template <typename T>
__global__ __launch_bounds__(TPB) void k_perm(typename VecOf<T>::type *a, uint64_t nf, OpArgs<T> g) {
    using V = typename VecOf<T>::type;
    const int D = 1 << g.k;
    for (uint64_t t = gtid(); t < nf; t += gstride()) {
        const uint64_t base = deposit_zeros(t, g.sq, g.k);
        V v[16];
        for (int x = 0; x < D; x++) v[x] = a[base + g.off[x]];
        for (int x = 0; x < D; x++) {
            V w = v[x];
            if (!((g.unit >> x) & 1)) {
                V c;
                c.x = g.re[x];
                c.y = g.im[x];
                w = cmulv(c, w);
            }
            a[base + g.off[g.img[x]]] = w;
        }
    }
}

// ---------------------------------------------------------------- Pauli
// (P psi)(c) = i^{ph + 2|z & (c^x)|} psi(c ^ x)  (vector.rs:143-164)
template <typename T>
__global__ __launch_bounds__(TPB) void k_pauli_diag(typename VecOf<T>::type *a, uint64_t N, uint64_t z, int ph) {
    for (uint64_t c = gtid(); c < N; c += gstride()) {
        const int k = (ph + 2 * __popcll(z & c)) & 3;
        if (k) a[c] = ipow_mul(k, a[c]);
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_pauli_pair(typename VecOf<T>::type *a, uint64_t np, uint64_t x, uint64_t z,
                                                     int ph, int h) {
    using V = typename VecOf<T>::type;
    for (uint64_t t = gtid(); t < np; t += gstride()) {
        const uint64_t c = ((t >> h) << (h + 1)) | (t & ((1ull << h) - 1));
        const uint64_t d = c ^ x;
        const V va = a[c], vb = a[d];
        a[c] = ipow_mul((ph + 2 * __popcll(z & d)) & 3, vb);
        a[d] = ipow_mul((ph + 2 * __popcll(z & c)) & 3, va);
    }
}

// ------------------------------------------------------------ reductions
// Per-thread Neumaier-compensated FP64 sums, then a fixed tree across the
// block and across blocks (deterministic for a given size).
struct Comp {
    double s, c;
    __device__ void init() { s = 0.0; c = 0.0; }
    __device__ void add(double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ double val() const { return s + c; }
};

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Reduce W values per thread into part[blockIdx.x * W + w].
template <int W> __device__ void block_store(const double (&v)[W], double *part) {
    __shared__ double red[TPB / 64][W];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < W; w++) {
        const double s = wave_sum(v[w]);
        if (lane == 0) red[wv][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < W) {
        double s = 0.0;
        for (int i = 0; i < TPB / 64; i++) s += red[i][threadIdx.x];
        part[(uint64_t)blockIdx.x * W + threadIdx.x] = s;
    }
}

enum RedMode { R_NORM = 0, R_EXPECT = 1, R_INNER = 2, R_MEASURE = 3 };
constexpr int MAXP = 8;  // Paulis per expectation launch

struct RedArgs {
    int count;
    uint64_t x[MAXP], z[MAXP];
    int ph[MAXP];
};

template <typename T, int MODE, int W>
__global__ __launch_bounds__(TPB) void k_reduce(const typename VecOf<T>::type *a, const typename VecOf<T>::type *b,
                                                uint64_t N, RedArgs r, double *part) {
    using V = typename VecOf<T>::type;
    Comp acc[W];
#pragma unroll
    for (int w = 0; w < W; w++) acc[w].init();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = gtid(); c < N; c += stride) {
        const V va = a[c];
        const double ar = (double)va.x, ai = (double)va.y;
        if (MODE == R_NORM) {
            acc[0].add(ar * ar + ai * ai);
        } else if (MODE == R_INNER) {
            const V vb = b[c];
            const double br = (double)vb.x, bi = (double)vb.y;
            acc[0].add(ar * br + ai * bi);  // conj(a) b
            acc[1].add(ar * bi - ai * br);
        } else if (MODE == R_MEASURE) {
            // norm and <P> (real part) together
            acc[0].add(ar * ar + ai * ai);
            const V vx = a[c ^ r.x[0]];
            const int k = (r.ph[0] + 2 * __popcll(r.z[0] & c)) & 3;
            const V t = ipow_mul(k, va);
            acc[1].add((double)vx.x * (double)t.x + (double)vx.y * (double)t.y);
        } else {
#pragma unroll
            for (int p = 0; p < MAXP; p++) {
                if (p < r.count) {
                    const V vx = a[c ^ r.x[p]];
                    const int k = (r.ph[p] + 2 * __popcll(r.z[p] & c)) & 3;
                    const V t = ipow_mul(k, va);
                    // conj(psi(c^x)) * i^k psi(c)
                    const double tr = (double)t.x, ti = (double)t.y, xr = (double)vx.x, xi = (double)vx.y;
                    acc[2 * p].add(xr * tr + xi * ti);
                    acc[2 * p + 1].add(xr * ti - xi * tr);
                }
            }
        }
    }
    double v[W];
#pragma unroll
    for (int w = 0; w < W; w++) v[w] = acc[w].val();
    block_store<W>(v, part);
}

// Sum nblocks x W partials into out[W] (one block; fixed order).
__global__ void k_finish(const double *part, int nblocks, int W, double *out) {
    __shared__ double red[TPB];
    for (int w = 0; w < W; w++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblocks; i += blockDim.x) s += part[(uint64_t)i * W + w];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[w] = red[0];
        __syncthreads();
    }
}

// -------------------------------------------------------------- compare
// Two states of any precisions, scaled by their FP64 factors: sum conj(a) b,
// sum |a|^2, sum |b|^2 and max |a - b| in one read of each.
template <typename VA, typename VB>
__global__ __launch_bounds__(TPB) void k_compare(const VA *a, const VB *b, uint64_t N, double sa, double sb,
                                                 double *part) {
    Comp acc[4];
#pragma unroll
    for (int w = 0; w < 4; w++) acc[w].init();
    double mx = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = gtid(); c < N; c += stride) {
        const VA va = a[c];
        const VB vb = b[c];
        const double ar = (double)va.x * sa, ai = (double)va.y * sa, br = (double)vb.x * sb, bi = (double)vb.y * sb;
        acc[0].add(ar * br + ai * bi);
        acc[1].add(ar * bi - ai * br);
        acc[2].add(ar * ar + ai * ai);
        acc[3].add(br * br + bi * bi);
        const double dr = ar - br, di = ai - bi;
        mx = fmax(mx, sqrt(dr * dr + di * di));
    }
    double v[5] = {acc[0].val(), acc[1].val(), acc[2].val(), acc[3].val(), 0.0};
    // block max of mx in slot 4, sums in 0..3
    __shared__ double red[TPB / 64][5];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int w = 0; w < 4; w++) {
        const double t = wave_sum(v[w]);
        if (lane == 0) red[wv][w] = t;
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    if (lane == 0) red[wv][4] = mx;
    __syncthreads();
    if (threadIdx.x < 5) {
        double t = threadIdx.x == 4 ? 0.0 : 0.0;
        for (int i = 0; i < TPB / 64; i++)
            t = threadIdx.x == 4 ? fmax(t, red[i][4]) : t + red[i][threadIdx.x];
        part[(uint64_t)blockIdx.x * 5 + threadIdx.x] = t;
    }
}

// ------------------------------------------------------------- collapse
// Z-type projector: keep or zero each amplitude (exact), weight of the kept.
template <typename T>
__global__ __launch_bounds__(TPB) void k_project_z(typename VecOf<T>::type *a, uint64_t N, uint64_t z, int ph,
                                                   int outcome, double *part) {
    using V = typename VecOf<T>::type;
    Comp acc;
    acc.init();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = gtid(); c < N; c += stride) {
        // sign of P on |c>: i^{ph + 2|z & c|} with ph even (Hermitian)
        const int k = (ph + 2 * __popcll(z & c)) & 3;
        const bool keep = ((k == 0) != (outcome != 0));
        const V v = a[c];
        if (keep) {
            acc.add((double)v.x * v.x + (double)v.y * v.y);
        } else if (!czero(v)) {
            V zr;
            zr.x = 0;
            zr.y = 0;
            a[c] = zr;
        }
    }
    double vv[1] = {acc.val()};
    block_store<1>(vv, part);
}

// General Hermitian P: psi <- (psi + s P psi)/2 on pairs {c, c^x}
// (vector.rs:470-476), weight accumulated in the same sweep.
template <typename T>
__global__ __launch_bounds__(TPB) void k_project_pair(typename VecOf<T>::type *a, uint64_t np, uint64_t x, uint64_t z,
                                                      int ph, int outcome, int h, double *part) {
    using V = typename VecOf<T>::type;
    Comp acc;
    acc.init();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int neg = outcome ? 2 : 0;
    for (uint64_t t = gtid(); t < np; t += stride) {
        const uint64_t c = ((t >> h) << (h + 1)) | (t & ((1ull << h) - 1));
        const uint64_t d = c ^ x;
        const V va = a[c], vb = a[d];
        V na = caddv(va, ipow_mul((ph + 2 * __popcll(z & d) + neg) & 3, vb));
        V nb = caddv(vb, ipow_mul((ph + 2 * __popcll(z & c) + neg) & 3, va));
        na.x *= (T)0.5;
        na.y *= (T)0.5;
        nb.x *= (T)0.5;
        nb.y *= (T)0.5;
        a[c] = na;
        a[d] = nb;
        acc.add((double)na.x * na.x + (double)na.y * na.y);
        acc.add((double)nb.x * nb.x + (double)nb.y * nb.y);
    }
    double vv[1] = {acc.val()};
    block_store<1>(vv, part);
}

// ------------------------------------------------------------ misc sweeps
template <typename T>
__global__ __launch_bounds__(TPB) void k_scale(typename VecOf<T>::type *a, uint64_t N, T cr, T ci) {
    // products round separately, as in the reference's complex multiply
#pragma clang fp contract(off)
    using V = typename VecOf<T>::type;
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        const V v = a[i];
        if (czero(v)) continue;
        V r;
        r.x = v.x * cr - v.y * ci;
        r.y = v.x * ci + v.y * cr;
        a[i] = r;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_fill(typename VecOf<T>::type *a, uint64_t N, T re, uint64_t one_index) {
    using V = typename VecOf<T>::type;
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        V v;
        v.x = (one_index == ~0ull) ? re : (i == one_index ? (T)1 : (T)0);
        v.y = 0;
        a[i] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_kron(const typename VecOf<T>::type *a, int na,
                                              const typename VecOf<T>::type *b, typename VecOf<T>::type *out,
                                              uint64_t N) {
    using V = typename VecOf<T>::type;
    for (uint64_t i = gtid(); i < N; i += gstride()) {
        const V va = a[i & ((1ull << na) - 1)], vb = b[i >> na];
        V r;
        r.x = 0;
        r.y = 0;
        if (!czero(va) && !czero(vb)) r = cmulv(va, vb);
        out[i] = r;
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(TPB) void k_convert(const TI *in, TO *out, uint64_t n2, double s) {
    for (uint64_t i = gtid(); i < n2; i += gstride()) out[i] = (TO)((double)in[i] * s);
}

// -------------------------------------------------------------- sampling
// QuantumState::sample as vector.rs:494-540 computes it, every floating-point
// operation in the reference's order, so the indices picked are identical
// given the same draws:
//  - a block's weight is the left-to-right sum, in index order, of
//    re*re + im*im (norm_sqr: two rounded products and a rounded sum, no
//    fused multiply-add), one lane per block;
//  - the running prefix acc_b and the total are one sequential fold;
//  - a shot's draw scaled by the total, u, lands in the first block with
//    u < acc_b + w_b (the last block if none) and walks it from acc_b while
//    run + |a_idx|^2 <= u, then steps back over amplitudes that are exactly 0.

template <typename V> __device__ __forceinline__ double abs_sq_f64(V v) {
#pragma clang fp contract(off)
    const double x = (double)v.x, y = (double)v.y;
    return x * x + y * y;
}

// Block weights, one lane per block of `bl` amplitudes, 64 blocks per
// single-wave workgroup.  The wave loads 128-byte runs (C amplitudes of one
// block per C lanes), squares them, and passes them through LDS to the lane
// that owns the block, which adds them in index order.
template <typename T>
__global__ __launch_bounds__(64) void k_block_weights(const typename VecOf<T>::type *a, uint64_t bl, uint64_t nblk,
                                                      uint32_t C, double *w) {
#pragma clang fp contract(off)
    using V = typename VecOf<T>::type;
    __shared__ double sq[64 * 16];
    const uint32_t lane = threadIdx.x;
    const uint64_t b0 = (uint64_t)blockIdx.x * 64;
    const uint32_t per = 64 / C;  // blocks per load instruction
    double s = 0.0;
    for (uint64_t c0 = 0; c0 < bl; c0 += C) {
        for (uint32_t i = 0; i < C; i++) {
            const uint32_t bi = i * per + lane / C;
            const uint64_t b = b0 + bi;
            double p = 0.0;
            if (b < nblk) p = abs_sq_f64(a[b * bl + c0 + lane % C]);
            sq[bi * C + lane % C] = p;
        }
        __syncthreads();
        for (uint32_t j = 0; j < C; j++) s += sq[lane * C + j];
        __syncthreads();
    }
    if (b0 + lane < nblk) w[b0 + lane] = s;
}

// The sequential fold of vector.rs:509-537 (acc = hi after each block) and
// of the total (vector.rs:501): w[b] becomes acc_b + w_b, *total the last.
__global__ void k_fold(double *w, uint64_t nblk, double *total) {
#pragma clang fp contract(off)
    double acc = 0.0;
    uint64_t b = 0;
    for (; b + 8 <= nblk; b += 8) {
        double v[8];
        for (int j = 0; j < 8; j++) v[j] = w[b + j];
        for (int j = 0; j < 8; j++) {
            acc = acc + v[j];
            w[b + j] = acc;
        }
    }
    for (; b < nblk; b++) {
        acc = acc + w[b];
        w[b] = acc;
    }
    *total = acc;
}

// One thread per shot: u = draw * total; the first block whose inclusive
// prefix exceeds u (binary search over the non-decreasing prefixes, the
// same block as the reference's forward scan over the sorted draws), then
// the in-block walk.  The reference walks a block's shots in draw order
// from where the previous one stopped; the run values along the walk are
// the same sequence, so starting each shot at the block start picks the
// same index.
template <typename T>
__global__ __launch_bounds__(TPB) void k_sample_walk(const typename VecOf<T>::type *a, uint64_t N, uint64_t bl,
                                                     uint64_t nblk, const double *hi, const double *total,
                                                     const double *us, uint64_t shots, uint64_t *out) {
#pragma clang fp contract(off)
    using V = typename VecOf<T>::type;
    for (uint64_t s = gtid(); s < shots; s += gstride()) {
        const double u = us[s] * *total;
        uint64_t lo = 0, up = nblk;
        while (lo < up) {
            const uint64_t mid = (lo + up) >> 1;
            if (u < hi[mid])
                up = mid;
            else
                lo = mid + 1;
        }
        const uint64_t b = lo < nblk ? lo : nblk - 1;
        double run = b ? hi[b - 1] : 0.0;
        const uint64_t start = b * bl, end = min(start + bl, N);
        uint64_t idx = start;
        while (idx + 1 < end) {
            const double p = abs_sq_f64(a[idx]);
            if (!(run + p <= u)) break;
            run += p;
            idx++;
        }
        uint64_t pick = idx;
        while (pick > start && czero(a[pick])) pick--;
        out[s] = pick;
    }
}

// ---------------------------------------------------------------- helpers
// Blocks for a grid-stride launch over `items`: at most 2^20 blocks, so
// grid x block stays well inside 32 bits for any state size.
inline uint32_t nblocks(uint64_t items) {
    const uint64_t b = (items + TPB - 1) / TPB;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

inline uint32_t red_blocks(uint64_t N) {
    uint64_t b = (N + TPB * 8 - 1) / (TPB * 8);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 2048));
}

template <typename T> OpArgs<T> make_args(const LOp &op) {
    OpArgs<T> g;
    memset(&g, 0, sizeof(g));
    const int k = (int)op.k();
    g.k = k;
    std::vector<int> s(op.qubits.begin(), op.qubits.end());
    std::sort(s.begin(), s.end());
    for (int j = 0; j < k && j < 12; j++) g.sq[j] = s[j];
    for (uint32_t x = 0; x < 16 && x < (1u << std::min(k, 4)); x++) {
        uint64_t o = 0;
        for (int j = 0; j < k; j++)
            if ((x >> j) & 1) o |= 1ull << op.qubits[j];
        g.off[x] = o;
    }
    return g;
}

void launch_check(const char *what) { check_hip(hipGetLastError(), what); }

}  // namespace

KTimer::KTimer(ncsv_state *s_, const char *name_, double bytes_) : s(s_), name(name_), bytes(bytes_) {
    if (s->profiling) check_hip(hipEventRecord(s->ev0, s->stream), "hipEventRecord");
}
KTimer::~KTimer() {
    if (!s->profiling) return;
    (void)hipEventRecord(s->ev1, s->stream);
    (void)hipEventSynchronize(s->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    auto &st = s->stats[name];
    st.launches++;
    st.ms += ms;
    st.bytes += bytes;
}

namespace dev {

void ensure_part(ncsv_state *s, size_t doubles) {
    if (s->part_cap >= doubles) return;
    if (s->d_part) (void)hipFree(s->d_part);
    s->d_part = nullptr;
    check_hip(hipMalloc(&s->d_part, doubles * sizeof(double)), "hipMalloc(partials)");
    s->part_cap = doubles;
    if (s->h_part) (void)hipHostFree(s->h_part);
    s->h_part = nullptr;
    check_hip(hipHostMalloc(&s->h_part, 64 * sizeof(double)), "hipHostMalloc");
}

void ensure_tab(ncsv_state *s, size_t bytes) {
    if (s->tab_cap >= bytes) return;
    if (s->d_tab) (void)hipFree(s->d_tab);
    s->d_tab = nullptr;
    check_hip(hipMalloc(&s->d_tab, bytes), "hipMalloc(tables)");
    s->tab_cap = bytes;
}

void sync(const ncsv_state *s) { check_hip(hipStreamSynchronize(s->stream), "hipStreamSynchronize"); }

template <typename T> static void fill_t(ncsv_state *s, double re, uint64_t idx) {
    const uint64_t N = 1ull << s->n;
    KTimer kt(s, "fill", (double)s->amp_bytes());
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks(N)), dim3(TPB), 0, s->stream, (typename VecOf<T>::type *)s->amp, N,
                       (T)re, idx);
    launch_check("k_fill");
}

void fill_basis(ncsv_state *s, uint64_t index) {
    s->lscale = 1.0;
    if (s->prec == NCSV_C128)
        fill_t<double>(s, 0.0, index);
    else
        fill_t<float>(s, 0.0, index);
}

void fill_const(ncsv_state *s, double re) {
    if (s->prec == NCSV_C128) {
        s->lscale = 1.0;
        fill_t<double>(s, re, ~0ull);
    } else {
        // complex64: store 1 and keep the constant in the FP64 scale
        s->lscale = re;
        fill_t<float>(s, 1.0, ~0ull);
    }
}

// Copies between the device and the caller's (pageable) memory, done when
// they return.  From 128 MiB on they go through two pinned 32 MiB staging
// buffers: host threads fill or drain one while the other crosses the link
// (a pageable buffer is otherwise copied a page-locked piece at a time).
// NCSV_STAGE=0 keeps one plain copy.
static bool use_staging(const ncsv_state *s, size_t bytes) {
    constexpr size_t CH = size_t(32) << 20;
    const char *e = getenv("NCSV_STAGE");
    if (bytes < 4 * CH || (e && e[0] == '0')) return false;
    auto *m = const_cast<ncsv_state *>(s);
    for (int k = 0; k < 2; k++) {
        if (!m->stage[k] && hipHostMalloc(&m->stage[k], CH, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();  // no pinned memory to be had: one plain copy
            m->stage[k] = nullptr;
            return false;
        }
        if (!m->staged[k]) check_hip(hipEventCreateWithFlags(&m->staged[k], hipEventDisableTiming), "hipEventCreate");
    }
    return true;
}

static void host_copy(void *dst, const void *src, size_t len) {
    const unsigned T = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const size_t part = (len + T - 1) / T;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) {
        const size_t a = std::min(len, t * part), b = std::min(len, a + part);
        if (a < b) th.emplace_back([=]() { memcpy((char *)dst + a, (const char *)src + a, b - a); });
    }
    for (auto &x : th) x.join();
}

static void to_host(const ncsv_state *s, void *dst, const void *src, size_t bytes) {
    constexpr size_t CH = size_t(32) << 20;
    if (!use_staging(s, bytes)) {
        check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s->stream), "download");
        sync(s);
        return;
    }
    const size_t nch = (bytes + CH - 1) / CH;
    auto drain = [&](size_t c) {
        const size_t off = c * CH;
        host_copy((char *)dst + off, s->stage[c & 1], std::min(CH, bytes - off));
    };
    for (size_t c = 0; c < nch; c++) {
        const size_t off = c * CH;
        check_hip(hipMemcpyAsync(s->stage[c & 1], (const char *)src + off, std::min(CH, bytes - off),
                                 hipMemcpyDeviceToHost, s->stream),
                  "download");
        check_hip(hipEventRecord(s->staged[c & 1], s->stream), "hipEventRecord");
        if (c > 0) {
            check_hip(hipEventSynchronize(s->staged[(c - 1) & 1]), "hipEventSynchronize");
            drain(c - 1);
        }
    }
    check_hip(hipEventSynchronize(s->staged[(nch - 1) & 1]), "hipEventSynchronize");
    drain(nch - 1);
}

static void to_device(const ncsv_state *s, void *dst, const void *src, size_t bytes) {
    constexpr size_t CH = size_t(32) << 20;
    if (!use_staging(s, bytes)) {
        check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->stream), "upload");
        sync(s);
        return;
    }
    const size_t nch = (bytes + CH - 1) / CH;
    for (size_t c = 0; c < nch; c++) {
        const size_t off = c * CH, len = std::min(CH, bytes - off);
        // the copy that last read this staging buffer (two chunks back) must be done
        if (c >= 2) check_hip(hipEventSynchronize(s->staged[c & 1]), "hipEventSynchronize");
        host_copy(s->stage[c & 1], (const char *)src + off, len);
        check_hip(hipMemcpyAsync((char *)dst + off, s->stage[c & 1], len, hipMemcpyHostToDevice, s->stream), "upload");
        check_hip(hipEventRecord(s->staged[c & 1], s->stream), "hipEventRecord");
    }
    sync(s);
}

void upload(ncsv_state *s, const double *amp) {
    const uint64_t N = 1ull << s->n;
    s->lscale = 1.0;
    if (s->prec == NCSV_C128) {
        to_device(s, s->amp, amp, s->amp_bytes());
        return;
    }
    // complex64: stage FP64 on the device, convert
    double *tmp = nullptr;
    check_hip(hipMalloc(&tmp, N * 16), "hipMalloc(staging)");
    to_device(s, tmp, amp, N * 16);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(2 * N)), dim3(TPB), 0, s->stream, tmp,
                       (float *)s->amp, 2 * N, 1.0);
    launch_check("k_convert");
    sync(s);
    (void)hipFree(tmp);
}

void download(const ncsv_state *s, double *out) {
    const uint64_t N = 1ull << s->n;
    if (s->prec == NCSV_C128) {
        to_host(s, out, s->amp, s->amp_bytes());
        return;
    }
    double *tmp = nullptr;
    check_hip(hipMalloc(&tmp, N * 16), "hipMalloc(staging)");
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(2 * N)), dim3(TPB), 0, s->stream,
                       (const float *)s->amp, tmp, 2 * N, s->lscale);
    launch_check("k_convert");
    to_host(s, out, tmp, N * 16);
    (void)hipFree(tmp);
}

void read_amp(const ncsv_state *s, uint64_t index, double out[2]) {
    if (s->prec == NCSV_C128) {
        check_hip(hipMemcpyAsync(s->h_part, (const double2 *)s->amp + index, 16, hipMemcpyDeviceToHost, s->stream),
                  "read_amp");
        sync(s);
        out[0] = s->h_part[0];
        out[1] = s->h_part[1];
    } else {
        check_hip(hipMemcpyAsync(s->h_part, (const float2 *)s->amp + index, 8, hipMemcpyDeviceToHost, s->stream),
                  "read_amp");
        sync(s);
        float f[2];
        memcpy(f, s->h_part, 8);
        out[0] = (double)f[0] * s->lscale;
        out[1] = (double)f[1] * s->lscale;
    }
}

void copy_state(const ncsv_state *src, ncsv_state *dst) {
    check_hip(hipMemcpyAsync(dst->amp, src->amp, src->amp_bytes(), hipMemcpyDeviceToDevice, dst->stream), "clone");
    dst->lscale = src->lscale;
    sync(dst);
}

template <typename T> static void apply_lop_t(ncsv_state *s, const LOp &op) {
    using V = typename VecOf<T>::type;
    V *a = (V *)s->amp;
    const int k = (int)op.k();
    const uint64_t N = 1ull << s->n;
    const uint64_t nf = N >> k;
    const double sweep = 2.0 * (double)s->amp_bytes();
    OpArgs<T> g = make_args<T>(op);
    switch (op.kind) {
    case OpKind::Nop: return;
    case OpKind::Pauli: return;
    case OpKind::Slot: fail(NCSV_INVALID, "noise slots exist only in batched programs");
    case OpKind::Diag: {
        if (k > 4) fail(NCSV_UNSUPPORTED, "diagonal op on more than 4 qubits");
        for (int x = 0; x < (1 << k); x++) {
            g.re[x] = (T)op.m[x].real();
            g.im[x] = (T)op.m[x].imag();
            if (op.m[x] == cd(1.0, 0.0)) g.unit |= 1u << x;
        }
        // only the amplitudes of non-unit phases are touched
        const double frac = (double)(k ? __builtin_popcount(~g.unit & ((1u << (1 << k)) - 1)) : 0) / (1 << k);
        KTimer kt(s, "diag", sweep * frac);
        hipLaunchKernelGGL(k_diag<T>, dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g);
        launch_check("k_diag");
        s->sweeps++;
        return;
    }
    case OpKind::Perm: {
        if (k > 4) fail(NCSV_UNSUPPORTED, "permutation op on more than 4 qubits");
        for (int x = 0; x < (1 << k); x++) {
            g.re[x] = (T)op.m[x].real();
            g.im[x] = (T)op.m[x].imag();
            g.img[x] = op.img[x];
            if (op.m[x] == cd(1.0, 0.0)) g.unit |= 1u << x;
        }
        KTimer kt(s, "perm", sweep);
        hipLaunchKernelGGL(k_perm<T>, dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g);
        launch_check("k_perm");
        s->sweeps++;
        return;
    }
    case OpKind::Dense: break;
    }
    const int D = 1 << k;
    if (k <= 2) {
        for (int i = 0; i < D * D; i++) {
            g.re[i] = (T)op.m[i].real();
            g.im[i] = (T)op.m[i].imag();
        }
        KTimer kt(s, k == 1 ? "dense1" : "dense2", sweep);
        if (k == 1)
            hipLaunchKernelGGL((k_dense_small<T, 1>), dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g);
        else
            hipLaunchKernelGGL((k_dense_small<T, 2>), dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g);
        launch_check("k_dense_small");
        s->sweeps++;
        return;
    }
    // upload matrix (+ offsets) to the table buffer
    std::vector<V> mat((size_t)D * D);
    for (size_t i = 0; i < mat.size(); i++) {
        mat[i].x = (T)op.m[i].real();
        mat[i].y = (T)op.m[i].imag();
    }
    if (k <= 5) {
        std::vector<uint64_t> offs(D);
        for (int x = 0; x < D; x++) {
            uint64_t o = 0;
            for (int j = 0; j < k; j++)
                if ((x >> j) & 1) o |= 1ull << op.qubits[j];
            offs[x] = o;
        }
        const size_t mb = mat.size() * sizeof(V);
        ensure_tab(s, mb + D * sizeof(uint64_t));
        check_hip(hipMemcpyAsync(s->d_tab, mat.data(), mb, hipMemcpyHostToDevice, s->stream), "upload matrix");
        check_hip(hipMemcpyAsync((char *)s->d_tab + mb, offs.data(), D * 8, hipMemcpyHostToDevice, s->stream),
                  "upload offsets");
        const V *dm = (const V *)s->d_tab;
        const uint64_t *doff = (const uint64_t *)((char *)s->d_tab + mb);
        KTimer kt(s, "dense_mid", sweep);
        if (k == 3)
            hipLaunchKernelGGL((k_dense_mid<T, 3>), dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g, dm, doff);
        else if (k == 4)
            hipLaunchKernelGGL((k_dense_mid<T, 4>), dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g, dm, doff);
        else
            hipLaunchKernelGGL((k_dense_mid<T, 5>), dim3(nblocks(nf)), dim3(TPB), 0, s->stream, a, nf, g, dm, doff);
        launch_check("k_dense_mid");
        sync(s);  // the host vectors above must outlive the copies
        s->sweeps++;
        return;
    }
    // k > 5: out of place
    std::vector<int> qv(op.qubits.begin(), op.qubits.end());
    const size_t mb = mat.size() * sizeof(V);
    ensure_tab(s, mb + qv.size() * sizeof(int));
    check_hip(hipMemcpyAsync(s->d_tab, mat.data(), mb, hipMemcpyHostToDevice, s->stream), "upload matrix");
    check_hip(hipMemcpyAsync((char *)s->d_tab + mb, qv.data(), qv.size() * sizeof(int), hipMemcpyHostToDevice,
                             s->stream),
              "upload qubits");
    void *out = nullptr;
    check_hip(hipMalloc(&out, s->amp_bytes()), "hipMalloc(out of place)");
    {
        KTimer kt(s, "dense_big", sweep);
        hipLaunchKernelGGL(k_dense_big<T>, dim3(nblocks(N)), dim3(TPB), 0, s->stream, (const V *)s->amp, (V *)out, N,
                           k, (const int *)((char *)s->d_tab + mb), (const V *)s->d_tab);
        launch_check("k_dense_big");
    }
    sync(s);
    (void)hipFree(s->amp);
    s->amp = out;
    s->sweeps++;
}

void apply_lop(ncsv_state *s, const LOp &op) {
    if (op.kind == OpKind::Pauli) {
        apply_pauli(s, op.x, op.z, op.ph);
        return;
    }
    if (s->prec == NCSV_C128)
        apply_lop_t<double>(s, op);
    else
        apply_lop_t<float>(s, op);
}

template <typename T> static void apply_pauli_t(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph) {
    using V = typename VecOf<T>::type;
    const uint64_t N = 1ull << s->n;
    V *a = (V *)s->amp;
    KTimer kt(s, "pauli", 2.0 * (double)s->amp_bytes());
    if (x == 0) {
        hipLaunchKernelGGL(k_pauli_diag<T>, dim3(nblocks(N)), dim3(TPB), 0, s->stream, a, N, z, (int)ph);
    } else {
        const int h = 63 - __builtin_clzll(x);
        hipLaunchKernelGGL(k_pauli_pair<T>, dim3(nblocks(N / 2)), dim3(TPB), 0, s->stream, a, N / 2, x, z, (int)ph,
                           h);
    }
    launch_check("k_pauli");
    s->sweeps++;
}

void apply_pauli(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph) {
    if (s->n < 64) {
        const uint64_t m = (1ull << s->n) - 1;
        x &= m;
        z &= m;
    }
    if (s->prec == NCSV_C128)
        apply_pauli_t<double>(s, x, z, ph & 3);
    else
        apply_pauli_t<float>(s, x, z, ph & 3);
}

// Runs a reduction kernel and sums its partials on the device.
template <typename T, int MODE, int W>
static void reduce_t(ncsv_state *s, const void *b, const RedArgs &r, double *out, const char *name, double bytes) {
    using V = typename VecOf<T>::type;
    const uint64_t N = 1ull << s->n;
    const uint32_t nb = red_blocks(N);
    ensure_part(s, (size_t)nb * W + 64);
    {
        KTimer kt(s, name, bytes);
        hipLaunchKernelGGL((k_reduce<T, MODE, W>), dim3(nb), dim3(TPB), 0, s->stream, (const V *)s->amp,
                           (const V *)b, N, r, s->d_part);
        launch_check("k_reduce");
    }
    double *fin = s->d_part + (size_t)nb * W;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(TPB), 0, s->stream, (const double *)s->d_part, (int)nb, W, fin);
    launch_check("k_finish");
    check_hip(hipMemcpyAsync(s->h_part, fin, W * sizeof(double), hipMemcpyDeviceToHost, s->stream), "readback");
    sync(s);
    for (int w = 0; w < W; w++) out[w] = s->h_part[w];
}

double norm_sq(ncsv_state *s) {
    RedArgs r;
    memset(&r, 0, sizeof(r));
    double out[1];
    if (s->prec == NCSV_C128)
        reduce_t<double, R_NORM, 1>(s, nullptr, r, out, "norm", (double)s->amp_bytes());
    else
        reduce_t<float, R_NORM, 1>(s, nullptr, r, out, "norm", (double)s->amp_bytes());
    return out[0];
}

bool expectations_tiled(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                        double *out);

void expectations(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                  double *out) {
    const uint64_t mask = s->n < 64 ? (1ull << s->n) - 1 : ~0ull;
    {
        std::vector<uint64_t> xm(count), zm(count);
        for (uint32_t i = 0; i < count; i++) {
            xm[i] = x[i] & mask;
            zm[i] = z[i] & mask;
        }
        if (count > 1 && use_specialized(s) && jit::expectations(s, count, xm.data(), zm.data(), ph, out)) return;
        if (count > 1 && expectations_tiled(s, count, xm.data(), zm.data(), ph, out)) return;
    }
    for (uint32_t p0 = 0; p0 < count; p0 += MAXP) {
        RedArgs r;
        memset(&r, 0, sizeof(r));
        r.count = (int)std::min<uint32_t>(MAXP, count - p0);
        int nx = 0;
        for (int i = 0; i < r.count; i++) {
            r.x[i] = x[p0 + i] & mask;
            r.z[i] = z[p0 + i] & mask;
            r.ph[i] = (int)(ph[p0 + i] & 3);
            nx += r.x[i] != 0;
        }
        // one read of psi plus one partner read per Pauli with an X part
        const double bytes = (double)s->amp_bytes() * (1 + nx);
        double tmp[2 * MAXP];
        if (s->prec == NCSV_C128)
            reduce_t<double, R_EXPECT, 2 * MAXP>(s, nullptr, r, tmp, "expect", bytes);
        else
            reduce_t<float, R_EXPECT, 2 * MAXP>(s, nullptr, r, tmp, "expect", bytes);
        for (int i = 0; i < r.count; i++) {
            out[2 * (p0 + i)] = tmp[2 * i];
            out[2 * (p0 + i) + 1] = tmp[2 * i + 1];
        }
    }
}

void expect_and_norm(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, double *e, double *nn) {
    RedArgs r;
    memset(&r, 0, sizeof(r));
    const uint64_t mask = s->n < 64 ? (1ull << s->n) - 1 : ~0ull;
    r.count = 1;
    r.x[0] = x & mask;
    r.z[0] = z & mask;
    r.ph[0] = (int)(ph & 3);
    double out[2];
    const double bytes = (double)s->amp_bytes() * (x ? 2 : 1);
    if (s->prec == NCSV_C128)
        reduce_t<double, R_MEASURE, 2>(s, nullptr, r, out, "measure", bytes);
    else
        reduce_t<float, R_MEASURE, 2>(s, nullptr, r, out, "measure", bytes);
    *nn = out[0];
    *e = out[1];
}

void inner(ncsv_state *a, ncsv_state *b, double out[2]) {
    RedArgs r;
    memset(&r, 0, sizeof(r));
    if (a->prec != b->prec) fail(NCSV_INVALID, "inner: states of different precision");
    sync(b);
    if (a->prec == NCSV_C128)
        reduce_t<double, R_INNER, 2>(a, b->amp, r, out, "inner", 2.0 * (double)a->amp_bytes());
    else
        reduce_t<float, R_INNER, 2>(a, b->amp, r, out, "inner", 2.0 * (double)a->amp_bytes());
}

template <typename VA, typename VB> static void compare_t(ncsv_state *a, ncsv_state *b, double out[5]) {
    const uint64_t N = 1ull << a->n;
    const uint32_t nb = red_blocks(N);
    ensure_part(a, (size_t)nb * 5 + 64);
    {
        KTimer kt(a, "compare", (double)a->amp_bytes() + (double)b->amp_bytes());
        hipLaunchKernelGGL((k_compare<VA, VB>), dim3(nb), dim3(TPB), 0, a->stream, (const VA *)a->amp,
                           (const VB *)b->amp, N, a->lscale, b->lscale, a->d_part);
        launch_check("k_compare");
    }
    std::vector<double> h((size_t)nb * 5);
    check_hip(hipMemcpyAsync(h.data(), a->d_part, h.size() * sizeof(double), hipMemcpyDeviceToHost, a->stream),
              "readback");
    sync(a);
    for (int w = 0; w < 5; w++) out[w] = 0.0;
    for (uint32_t i = 0; i < nb; i++) {
        for (int w = 0; w < 4; w++) out[w] += h[(size_t)i * 5 + w];
        out[4] = std::max(out[4], h[(size_t)i * 5 + 4]);
    }
}

void compare(ncsv_state *a, ncsv_state *b, double out[5]) {
    if (a->n != b->n) fail(NCSV_INVALID, "compare: register sizes differ");
    sync(b);
    if (a->prec == NCSV_C128 && b->prec == NCSV_C128) compare_t<double2, double2>(a, b, out);
    else if (a->prec == NCSV_C128) compare_t<double2, float2>(a, b, out);
    else if (b->prec == NCSV_C128) compare_t<float2, double2>(a, b, out);
    else compare_t<float2, float2>(a, b, out);
}

template <typename T> static double project_t(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, int outcome) {
    using V = typename VecOf<T>::type;
    const uint64_t N = 1ull << s->n;
    const uint32_t nb = red_blocks(N);
    ensure_part(s, (size_t)nb + 64);
    {
        KTimer kt(s, "project", 2.0 * (double)s->amp_bytes());
        if (x == 0) {
            hipLaunchKernelGGL(k_project_z<T>, dim3(nb), dim3(TPB), 0, s->stream, (V *)s->amp, N, z, (int)ph, outcome,
                               s->d_part);
        } else {
            const int h = 63 - __builtin_clzll(x);
            hipLaunchKernelGGL(k_project_pair<T>, dim3(nb), dim3(TPB), 0, s->stream, (V *)s->amp, N / 2, x, z, (int)ph,
                               outcome, h, s->d_part);
        }
        launch_check("k_project");
    }
    double *fin = s->d_part + nb;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(TPB), 0, s->stream, (const double *)s->d_part, (int)nb, 1, fin);
    launch_check("k_finish");
    check_hip(hipMemcpyAsync(s->h_part, fin, sizeof(double), hipMemcpyDeviceToHost, s->stream), "readback");
    sync(s);
    s->sweeps++;
    return s->h_part[0];
}

double project(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, int outcome) {
    const uint64_t mask = s->n < 64 ? (1ull << s->n) - 1 : ~0ull;
    x &= mask;
    z &= mask;
    if (s->prec == NCSV_C128) return project_t<double>(s, x, z, ph & 3, outcome);
    return project_t<float>(s, x, z, ph & 3, outcome);
}

void scale(ncsv_state *s, cd c) {
    const uint64_t N = 1ull << s->n;
    KTimer kt(s, "scale", 2.0 * (double)s->amp_bytes());
    if (s->prec == NCSV_C128)
        hipLaunchKernelGGL(k_scale<double>, dim3(nblocks(N)), dim3(TPB), 0, s->stream, (double2 *)s->amp, N, c.real(),
                           c.imag());
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(nblocks(N)), dim3(TPB), 0, s->stream, (float2 *)s->amp, N,
                           (float)c.real(), (float)c.imag());
    launch_check("k_scale");
    s->sweeps++;
}

void kron(ncsv_state *a, ncsv_state *b, ncsv_state *out) {
    const uint64_t N = 1ull << out->n;
    sync(a);
    sync(b);
    KTimer kt(out, "kron", (double)out->amp_bytes());
    if (out->prec == NCSV_C128)
        hipLaunchKernelGGL(k_kron<double>, dim3(nblocks(N)), dim3(TPB), 0, out->stream, (const double2 *)a->amp,
                           (int)a->n, (const double2 *)b->amp, (double2 *)out->amp, N);
    else
        hipLaunchKernelGGL(k_kron<float>, dim3(nblocks(N)), dim3(TPB), 0, out->stream, (const float2 *)a->amp,
                           (int)a->n, (const float2 *)b->amp, (float2 *)out->amp, N);
    launch_check("k_kron");
    out->lscale = a->lscale * b->lscale;
    sync(out);
}

// Scratch of one sampling call: block prefixes | total (2 doubles) | draws |
// picks.
struct SampleBufs {
    double *hi, *total, *us;
    uint64_t *out;
};

static size_t sample_doubles(uint64_t nblk, uint64_t shots) { return nblk + 2 + 2 * shots; }

static SampleBufs sample_carve(void *base, uint64_t nblk, uint64_t shots) {
    SampleBufs b;
    b.hi = (double *)base;
    b.total = b.hi + nblk;
    b.us = b.total + 2;
    b.out = (uint64_t *)(b.us + shots);
    return b;
}

// The sampling launches over caller-provided device scratch `base` (at
// least sample_doubles() doubles): draws up, block weights, fold, ZeroState
// check, walk, picks down.
template <typename T>
static void sample_on(ncsv_state *s, void *base, uint64_t shots, const double *uniforms, uint64_t *out) {
    using V = typename VecOf<T>::type;
    const uint64_t N = 1ull << s->n;
    const uint64_t bl = std::min<uint64_t>(SAMPLE_BLOCK, N);
    const uint64_t nblk = N / bl;
    const uint32_t C = (uint32_t)std::min<uint64_t>(bl, 128 / sizeof(V));  // 128-byte runs
    const SampleBufs sb = sample_carve(base, nblk, shots);
    if (shots)
        check_hip(hipMemcpyAsync(sb.us, uniforms, shots * sizeof(double), hipMemcpyHostToDevice, s->stream),
                  "upload draws");
    {
        KTimer kt(s, "sample_weights", (double)s->amp_bytes());
        hipLaunchKernelGGL(k_block_weights<T>, dim3((uint32_t)((nblk + 63) / 64)), dim3(64), 0, s->stream,
                           (const V *)s->amp, bl, nblk, C, sb.hi);
        launch_check("k_block_weights");
    }
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(1), 0, s->stream, sb.hi, nblk, sb.total);
    launch_check("k_fold");
    check_hip(hipMemcpyAsync(s->h_part, sb.total, sizeof(double), hipMemcpyDeviceToHost, s->stream), "total");
    sync(s);
    if (!(s->h_part[0] > 0.0)) fail(NCSV_ZERO_STATE, "state vanished: sample");
    if (shots) {
        hipLaunchKernelGGL(k_sample_walk<T>, dim3(nblocks(shots)), dim3(TPB), 0, s->stream, (const V *)s->amp, N, bl,
                           nblk, (const double *)sb.hi, (const double *)sb.total, (const double *)sb.us, shots,
                           sb.out);
        launch_check("k_sample_walk");
        check_hip(hipMemcpyAsync(out, sb.out, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream), "shots");
    }
    sync(s);
}

static uint64_t sample_nblk(const ncsv_state *s) {
    const uint64_t N = 1ull << s->n;
    return N / std::min<uint64_t>(SAMPLE_BLOCK, N);
}

}  // namespace dev
}  // namespace ncsv

// Device scratch of ncsv_sample_noalloc: one array per GPU, held until the
// process ends (32 MiB: block prefixes, the total, draws and picks).  Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the device (the array's name is a host-side stand-in there).
constexpr size_t NOALLOC_DOUBLES = size_t(4) << 20;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace ncsv {
namespace dev {
namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};
// every call on one GPU uses the same array: calls are serialized per device
std::mutex g_scratch_use[MAX_DEVICES];

// The device address of `scratch` on `device`, looked up once per device
// with that device current.
void *scratch_on(int device) {
    if (device < 0 || device >= MAX_DEVICES) fail(NCSV_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        check_hip(hipGetDevice(&cur), "hipGetDevice");
        if (cur != device) check_hip(hipSetDevice(device), "hipSetDevice");
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        check_hip(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace

void sample_noalloc(ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out) {
    const uint64_t nblk = sample_nblk(s);
    if (shots > NOALLOC_DOUBLES || sample_doubles(nblk, shots) > NOALLOC_DOUBLES)
        fail(NCSV_INVALID, "sample_noalloc: " + std::to_string(shots) + " shots on " + std::to_string(s->n) +
                               " qubits need more than the fixed scratch of " + std::to_string(NOALLOC_DOUBLES) +
                               " doubles (block count + 2 + 2 x shots)");
    void *ptr = scratch_on(s->device);
    // nothing in the array must start at zero: every word read is written
    // earlier in the same call
    std::lock_guard<std::mutex> lk(g_scratch_use[s->device]);
    if (s->prec == NCSV_C128)
        sample_on<double>(s, ptr, shots, uniforms, out);
    else
        sample_on<float>(s, ptr, shots, uniforms, out);
}

void sample(ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out) {
    // per-handle scratch, grown on demand and kept with the state
    const size_t bytes = sample_doubles(sample_nblk(s), shots) * sizeof(double);
    if (s->samp_cap < bytes) {
        if (s->d_samp) (void)hipFree(s->d_samp);
        s->d_samp = nullptr;
        s->samp_cap = 0;
        check_hip(hipMalloc(&s->d_samp, bytes), "hipMalloc(sampling scratch)");
        s->samp_cap = bytes;
    }
    if (s->prec == NCSV_C128)
        sample_on<double>(s, s->d_samp, shots, uniforms, out);
    else
        sample_on<float>(s, s->d_samp, shots, uniforms, out);
}

}  // namespace dev
}  // namespace ncsv
