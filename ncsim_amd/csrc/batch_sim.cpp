// Host build of the batched form's kernel body (batch_kernel.hpp): one
// workgroup runs as 512 or 256 host threads, with a real barrier for each
// workgroup barrier and a shared exchange for the wavefront shuffle, so the
// kernel's own logic (queue, dependencies, ops, reductions) can be run and
// checked without a device.  One workgroup claims every item in order, so
// every dependency is met when it is reached.
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "ncsv_internal.hpp"

namespace ncsv {
namespace batch {
namespace kern {
struct hf2 {
    float x, y;
};
struct hd2 {
    double x, y;
};
template <typename T> struct V2;
template <> struct V2<float> { using type = hf2; };
template <> struct V2<double> { using type = hd2; };
template <typename T> struct Q4 {
    T v[4];
    T &operator[](int i) { return v[i]; }
    T operator[](int i) const { return v[i]; }
};
template <typename T> struct M4 {
    using type = Q4<T>;
};
}  // namespace kern

namespace sim {
struct Barrier {
    std::mutex m;
    std::condition_variable cv;
    unsigned n, count = 0, gen = 0;
    explicit Barrier(unsigned n_) : n(n_) {}
    void wait() {
        std::unique_lock<std::mutex> lk(m);
        const unsigned g = gen;
        if (++count == n) {
            count = 0;
            gen++;
            cv.notify_all();
            return;
        }
        cv.wait(lk, [&] { return gen != g; });
    }
};
thread_local unsigned tid = 0;
Barrier *bar = nullptr;
double xbuf[1024];
double ma[1024], mb[1024];
// The 16x16x4 MFMAs of the calling thread's wavefront: lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; acc[v] += D[i][l % 16] with i = 4 (l / 16) + v for FP32
// (v_mfma_f32_16x16x4_f32) and i = 4 v + l / 16 for FP64 (v_mfma_f64_16x16x4_f64).
template <typename T> kern::Q4<T> mfma16(T a, T b, kern::Q4<T> acc) {
    ma[tid] = a;
    mb[tid] = b;
    bar->wait();
    const unsigned w0 = tid & ~63u, l = tid & 63u;
    for (unsigned v = 0; v < 4; v++) {
        const unsigned i = sizeof(T) == 4 ? 4 * (l / 16) + v : 4 * v + l / 16, j = l % 16;
        T d = (T)0;
        for (unsigned k = 0; k < 4; k++) d += (T)ma[w0 + i + 16 * k] * (T)mb[w0 + j + 16 * k];
        acc[v] += d;
    }
    bar->wait();
    return acc;
}
double shfl_xor(double x, unsigned mask) {
    xbuf[tid] = x;
    bar->wait();
    const double r = xbuf[tid ^ mask];
    bar->wait();
    return r;
}
}  // namespace sim
}  // namespace batch
}  // namespace ncsv

#define NCSV_DEV
#define NCSV_KERNEL
#define NCSV_SHARED static
#define NCSV_TID (::ncsv::batch::sim::tid)
#define NCSV_SYNC() ::ncsv::batch::sim::bar->wait()
#define NCSV_SHFL_XOR(x, m) ::ncsv::batch::sim::shfl_xor((x), (m))
#define NCSV_ATOMIC_ADD(p, v) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)
#define NCSV_ATOMIC_OR(p, v) __atomic_fetch_or((p), (v), __ATOMIC_SEQ_CST)
#define NCSV_LOAD_AGENT(p) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define NCSV_ADD_AGENT(p, v) __atomic_fetch_add((p), (v), __ATOMIC_SEQ_CST)
#define NCSV_FENCE_ACQUIRE() __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define NCSV_FENCE_RELEASE() __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define NCSV_SLEEP() std::this_thread::yield()
#define NCSV_UNIFORM64(v) (v)
#define NCSV_UNIFORM32(v) (v)
#define NCSV_OPAQUE(x) ((void)0)
#define NCSV_MFMA16(a, b, acc) ::ncsv::batch::sim::mfma16<float>((a), (b), (acc))
#define NCSV_MFMA16D(a, b, acc) ::ncsv::batch::sim::mfma16<double>((a), (b), (acc))
#define NCSV_LDG(p) (*(p))
#define NCSV_STG(p, v) (*(p) = (v))
#define NCSV_LANE0_STORE(dst, v) \
    do {                         \
        if (NCSV_TID == 0) (dst) = (v); \
    } while (0)

#include "batch_kernel.hpp"

namespace ncsv {
namespace batch {

// Runs the kernel body over the whole queue with one simulated workgroup of
// `threads` threads (512, or 256 for tiles below the largest).
void simulate(const BArgs &A, ncsv_precision prec, uint32_t threads) {
    static std::mutex one;  // the simulated workgroup's LDS is static
    std::lock_guard<std::mutex> lk(one);
    sim::Barrier b(threads);
    sim::bar = &b;
    std::vector<std::thread> th;
    th.reserve(threads);
    for (unsigned t = 0; t < threads; t++) {
        th.emplace_back([&A, prec, t, threads]() {
            sim::tid = t;
#define NCSV_RUN(T, LB, BT) \
    kern::k_batch<T, LB, BT>(A.steps, A.ops, A.coef, A.paulis, A.mgroups, A.entries, A.recs, A.rec_blk, A.uni, A.forced, A.errs, A)
            if (prec == NCSV_C128 && threads == 256)
                NCSV_RUN(double, 11, 256);
            else if (prec == NCSV_C128)
                NCSV_RUN(double, 12, 512);
            else if (threads == 256)
                NCSV_RUN(float, 12, 256);
            else
                NCSV_RUN(float, 13, 512);
#undef NCSV_RUN
        });
    }
    for (auto &x : th) x.join();
    sim::bar = nullptr;
}

}  // namespace batch
}  // namespace ncsv
