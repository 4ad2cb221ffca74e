// Internal declarations shared by the host-side C++ (gate lowering, circuit
// parsing, schedule planning, the C ABI) and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ncsv.h"

namespace ncsv {

using cd = std::complex<double>;

// An error crossing back to the C ABI as a status plus message.
struct Error {
    ncsv_status code;
    std::string msg;
};
[[noreturn]] void fail(ncsv_status code, const std::string &msg);
void check_hip(hipError_t e, const char *what);

// ---------------------------------------------------------------- gates
// A gate as data (src/gate.rs:355-431): the named gates plus UNITARY and
// PAULI.  Phase/CPhase carry (k, bits); rotations theta; UNITARY a matrix;
// PAULI local masks.
struct GateDesc {
    ncsv_gate id = NCSV_GATE_I;
    int64_t k = 0;
    uint32_t bits = 0;
    double theta = 0.0;
    uint32_t mk = 0;         // UNITARY: qubits
    std::vector<cd> mat;     // UNITARY: 2^mk x 2^mk row-major
    uint64_t px = 0, pz = 0; // PAULI: local masks
    uint32_t pph = 0, pn = 0;
};
GateDesc make_gate(ncsv_gate id, const int64_t *iparams, const double *fparams, uint32_t nq);
uint32_t gate_arity(const GateDesc &g);
std::string gate_name(const GateDesc &g);
bool gate_lookup(const std::string &upper_name, ncsv_gate *id);

// C64 constants (src/scalar.rs:101-140).
cd root_of_unity(int64_t k, uint32_t bits);
cd sqrt2_pow(int64_t k);

// A gate lowered to what the dense kernels understand, acting on `qubits`
// of the register (local bit j = qubits[j]).
// Slot: a noise site of the batched form on qubits[0] (slot index in `ph`).
enum class OpKind { Dense, Diag, Perm, Pauli, Nop, Slot };
struct LOp {
    OpKind kind = OpKind::Nop;
    std::vector<uint32_t> qubits;
    std::vector<cd> m;          // Dense: 2^k x 2^k; Diag and Perm: 2^k phases
    std::vector<uint32_t> img;  // Perm: image of each local basis state
    uint64_t x = 0, z = 0;      // Pauli on the register
    uint32_t ph = 0;
    uint32_t k() const { return (uint32_t)qubits.size(); }
};
// check_qubits (src/error.rs:69-82).
void check_qubits(const uint32_t *qubits, uint32_t nq, uint32_t n);
// QuantumState::apply_gate's routing (vector.rs:364-395) as a lowered op.
LOp lower_gate(const GateDesc &g, const uint32_t *qubits, uint32_t nq, uint32_t n);

// ----------------------------------------------------------------- Paulis
struct PauliM {
    uint32_t n = 0;
    uint64_t x = 0, z = 0;
    uint32_t ph = 0;
};
bool pauli_is_hermitian(uint64_t x, uint64_t z, uint32_t ph);
// Pauli::parse / from_string (src/pauli/mod.rs:323-414); n < 0 infers.
PauliM pauli_parse(const std::string &s, int n);
std::string pauli_sparse_string(const PauliM &p);

// --------------------------------------------------------------- circuits
enum class CKind { Gate, Measure, Reset, Noise, Detector, Observable, Tick };
struct COp {
    CKind kind = CKind::Tick;
    GateDesc g;
    std::vector<uint32_t> qubits;
    uint64_t x = 0, z = 0;
    uint32_t ph = 0;
    ncsv_noise ch = NCSV_NOISE_X_ERROR;
    double p[3] = {0, 0, 0};
    std::vector<uint32_t> lookback;
    uint32_t index = 0;
};
// Noise::errors (circuit.rs:119-128): (probability, x bits, z bits).
struct NoiseErr {
    double p;
    uint32_t x, z;
};
std::vector<NoiseErr> noise_errors(ncsv_noise ch, const double *p);
uint32_t noise_arity(ncsv_noise ch);
std::string noise_name(ncsv_noise ch, const double *p);

}  // namespace ncsv

struct ncsv_circuit {
    uint32_t n = 0;
    std::vector<ncsv::COp> ops;
    uint64_t measurements = 0;
    // cached fused schedule, keyed by the state's precision (built on demand)
    std::mutex mu;
    std::map<int64_t, std::shared_ptr<void>> plans;
};

namespace ncsv {
void circuit_push(ncsv_circuit *c, COp op);
}

// -------------------------------------------------------------- the state
struct ncsv_kstat {
    uint64_t launches = 0;
    double ms = 0.0;
    double bytes = 0.0;
};

struct ncsv_state {
    uint32_t n = 0;
    ncsv_precision prec = NCSV_C128;
    int device = 0;
    void *amp = nullptr;  // 2^n complex amplitudes (double2 or float2)
    // a second buffer of the same size for sweeps that write out of place
    // (allocated on the first such sweep; the two swap after each of them);
    // spare_failed: the allocation was refused, sweeps stay in place
    void *spare = nullptr;
    bool spare_failed = false;
    // pinned staging for large copies between the device and the caller's
    // memory (allocated on the first such copy)
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t staged[2] = {nullptr, nullptr};
    double lscale = 1.0;  // logical amplitude = lscale * stored (complex64 form)
    // set_basis deferred: the state is |pending_basis> but the amplitudes are
    // not written yet (-1: they are); a run's first specialized sweep computes
    // them in registers, anything else writes them first (materialize)
    int64_t pending_basis = -1;
    hipStream_t stream = nullptr;
    // reduction partials and small tables
    double *d_part = nullptr;
    size_t part_cap = 0;
    void *d_tab = nullptr;
    size_t tab_cap = 0;
    // sampling scratch (block prefixes, draws, picks), grown on demand
    void *d_samp = nullptr;
    size_t samp_cap = 0;
    double *h_part = nullptr;  // pinned
    size_t h_part_cap = 0;
    int fusion = 1;
    uint64_t sweeps = 0;
    bool profiling = false;
    std::map<std::string, ncsv_kstat> stats;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    mutable std::recursive_mutex mu;

    size_t amp_bytes() const { return (size_t(1) << n) * (prec == NCSV_C128 ? 16 : 8); }
    size_t elem_bytes() const { return prec == NCSV_C128 ? 16 : 8; }
};

namespace ncsv {
// Writes a deferred basis state into the amplitudes (no-op when none).
void materialize(ncsv_state *s);
}  // namespace ncsv

namespace ncsv {
namespace dev {
// Device primitives (kernels.hip).  All run on s->stream.
void ensure_part(ncsv_state *s, size_t doubles);
void ensure_tab(ncsv_state *s, size_t bytes);
void fill_basis(ncsv_state *s, uint64_t index);
void fill_const(ncsv_state *s, double re);
void upload(ncsv_state *s, const double *amp);
void download(const ncsv_state *s, double *out);
void read_amp(const ncsv_state *s, uint64_t index, double out[2]);
void copy_state(const ncsv_state *src, ncsv_state *dst);
void apply_lop(ncsv_state *s, const LOp &op);
void apply_pauli(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph);
// sum |a|^2 of the stored amplitudes (FP64, compensated).
double norm_sq(ncsv_state *s);
// expectation values of `count` Paulis (stored amplitudes), out[2k..2k+1].
void expectations(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                  double *out);
// <Z-type or general P> real part and norm together (measure).
void expect_and_norm(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, double *e, double *nn);
void inner(ncsv_state *a, ncsv_state *b, double out[2]);
// <a|b>, |a|^2, |b|^2 (FP64, logical scale) and max |a - b|, any precisions.
void compare(ncsv_state *a, ncsv_state *b, double out[5]);
// collapse + weight of the stored amplitudes in one sweep.
double project(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, int outcome);
void scale(ncsv_state *s, cd c);
void kron(ncsv_state *a, ncsv_state *b, ncsv_state *out);
void sample(ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out);
// sample() over the library's fixed per-device scratch array (no allocation).
void sample_noalloc(ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out);
void sync(const ncsv_state *s);
}  // namespace dev

// Timing of one launch when profiling is on.
struct KTimer {
    ncsv_state *s;
    const char *name;
    double bytes;
    KTimer(ncsv_state *s_, const char *name_, double bytes_);
    ~KTimer();
};
}  // namespace ncsv
