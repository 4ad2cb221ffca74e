// The batched form's persistent kernel for gfx950: workgroups claim (step,
// tile) items of many entries from one queue, hold each tile in LDS while
// the step's ops run on it, and order an entry's steps through completion
// counters.  The kernel body is shared with the host simulation.
#include "batch.hpp"
#include "ncsv_internal.hpp"

namespace ncsv {
namespace batch {
namespace kern {
template <typename T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <typename T> struct M4;
template <> struct M4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct M4<double> { typedef double type __attribute__((ext_vector_type(4))); };
}  // namespace kern
}  // namespace batch
}  // namespace ncsv

#define NCSV_DEV __device__
#define NCSV_KERNEL __global__ __launch_bounds__(BT, 4)  // BT: the kernel's template parameter
#define NCSV_SHARED __shared__
#define NCSV_TID threadIdx.x
#define NCSV_SYNC() __syncthreads()
#define NCSV_SHFL_XOR(x, m) __shfl_xor((x), (int)(m))
#define NCSV_ATOMIC_ADD(p, v) atomicAdd((p), (v))
#define NCSV_ATOMIC_OR(p, v) atomicOr((p), (v))
#define NCSV_LOAD_AGENT(p) __hip_atomic_load((p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define NCSV_ADD_AGENT(p, v) __hip_atomic_fetch_add((p), (v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
#define NCSV_FENCE_ACQUIRE() __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent")
#define NCSV_FENCE_RELEASE() __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent")
#define NCSV_SLEEP() __builtin_amdgcn_s_sleep(8)
#define NCSV_UNIFORM64(v) ::ncsv::batch::kern::uniform64(v)
#define NCSV_UNIFORM32(v) __builtin_amdgcn_readfirstlane((uint32_t)(v))
#define NCSV_LANE0_STORE(dst, v) ((dst) = ::ncsv::batch::kern::uniform64(v))
#define NCSV_OPAQUE(x) asm volatile("" : "+v"(x))
#define NCSV_MFMA16(a, b, acc) __builtin_amdgcn_mfma_f32_16x16x4f32((a), (b), (acc), 0, 0, 0)
#define NCSV_MFMA16D(a, b, acc) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (acc), 0, 0, 0)
// tile loads and stores stream through the caches with nontemporal hints
// (4096-circuit mix: complex128 1014 -> 920 ms per run, complex64 510 -> 503)
#define NCSV_LDG(p) ::ncsv::batch::kern::ld_nt(p)
#define NCSV_STG(p, v) ::ncsv::batch::kern::st_nt((p), (v))

namespace ncsv {
namespace batch {
namespace kern {
__device__ inline float2 ld_nt(const float2 *p) {
    typedef float NV __attribute__((ext_vector_type(2)));
    const NV t = __builtin_nontemporal_load((const NV *)p);
    return make_float2(t.x, t.y);
}
__device__ inline double2 ld_nt(const double2 *p) {
    typedef double NV __attribute__((ext_vector_type(2)));
    const NV t = __builtin_nontemporal_load((const NV *)p);
    return make_double2(t.x, t.y);
}
__device__ inline void st_nt(float2 *p, float2 v) {
    typedef float NV __attribute__((ext_vector_type(2)));
    NV t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, (NV *)p);
}
__device__ inline void st_nt(double2 *p, double2 v) {
    typedef double NV __attribute__((ext_vector_type(2)));
    NV t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, (NV *)p);
}
__device__ inline uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
}  // namespace kern
}  // namespace batch
}  // namespace ncsv

#include "batch_kernel.hpp"

namespace ncsv {
namespace batch {
using kern::k_batch;
namespace {

__global__ void k_batch_reset(uint32_t *done, uint32_t nrec, unsigned long long *next, uint32_t *flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrec) done[i] = 0;
    if (i == 0) {
        *next = 0;
        *flag = 0;
    }
}

// Each sum of per-tile partials, one thread per sum, in tile order with
// Neumaier compensation (the same sequence as the host's sum_partials).
__global__ void k_batch_finish(const double *__restrict__ part, const BSumD *__restrict__ sums, uint32_t count,
                               double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const BSumD r = sums[i];
    const double *p = part + r.off;
    double s = 0.0, comp = 0.0;
    for (uint32_t t = 0; t < r.tiles; t++) {
        const double x = p[(uint64_t)r.stride * t];
        const double y = s + x;
        comp += fabs(s) >= fabs(x) ? (s - y) + x : (x - y) + s;
        s = y;
    }
    out[i] = s + comp;
}

}  // namespace

void launch_finish(const double *part, const BSumD *sums, uint32_t count, double *out, hipStream_t stream) {
    if (!count) return;
    hipLaunchKernelGGL(k_batch_finish, dim3((count + 255) / 256), dim3(256), 0, stream, part, sums, count, out);
    check_hip(hipGetLastError(), "k_batch_finish");
}

uint32_t max_tile_bits(ncsv_precision prec) { return prec == NCSV_C128 ? 12u : 13u; }

int resident_blocks(ncsv_precision prec, uint32_t threads) {
    int dev = 0, cus = 0, per = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    if (prec == NCSV_C128 && threads == 256)
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<double, 11, 256>, 256, 0), "occupancy");
    else if (prec == NCSV_C128)
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<double, 12, 512>, 512, 0), "occupancy");
    else if (threads == 256)
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<float, 12, 256>, 256, 0), "occupancy");
    else
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch<float, 13, 512>, 512, 0), "occupancy");
    return cus * std::max(per, 1);
}

// threads == 256: the workgroups of 256 threads, for tiles of at most
// max_tile_bits - 1 qubits (half the LDS, twice the resident workgroups)
void launch(const BArgs &A, uint32_t nrec, ncsv_precision prec, uint32_t threads, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL(k_batch_reset, dim3((nrec + 255) / 256), dim3(256), 0, stream, A.done, nrec, A.next, A.flag);
    check_hip(hipGetLastError(), "k_batch_reset");
#define NCSV_LAUNCH(T, LB, BT)                                                                                        \
    hipLaunchKernelGGL((k_batch<T, LB, BT>), dim3(blocks), dim3(BT), 0, stream, A.steps, A.ops, A.coef, A.paulis,     \
                       A.mgroups, A.entries, A.recs, A.rec_blk, A.uni, A.forced, A.errs, A)
    if (prec == NCSV_C128 && threads == 256)
        NCSV_LAUNCH(double, 11, 256);
    else if (prec == NCSV_C128)
        NCSV_LAUNCH(double, 12, 512);
    else if (threads == 256)
        NCSV_LAUNCH(float, 12, 256);
    else
        NCSV_LAUNCH(float, 13, 512);
#undef NCSV_LAUNCH
    check_hip(hipGetLastError(), "k_batch");
}

}  // namespace batch
}  // namespace ncsv
