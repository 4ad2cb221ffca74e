// The C ABI (include/ncsv.h): handles, error translation, and the provided
// methods of QuantumState (src/state/mod.rs:74-233) over the device
// primitives.  Every entry point catches internal errors and returns a
// status with a per-thread message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>

#include "fuse.hpp"
#include "ncsv_internal.hpp"

using namespace ncsv;

ncsv_circuit *ncsv_parse_text(const char *text);
std::string ncsv_circuit_text(const ncsv_circuit *c);

namespace {

thread_local std::string g_last_error;

ncsv_status set_err(ncsv_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

template <typename F> ncsv_status guard(F &&f) {
    try {
        f();
        return NCSV_OK;
    } catch (Error &e) {
        return set_err(e.code, e.msg);
    } catch (std::bad_alloc &) {
        return set_err(NCSV_DEVICE, "out of host memory");
    } catch (std::exception &e) {
        return set_err(NCSV_INVALID, e.what());
    } catch (...) {
        return set_err(NCSV_INVALID, "unknown error");
    }
}

void need(const void *p, const char *what) {
    if (!p) fail(NCSV_INVALID, std::string(what) + " is null");
}

// The mask of the register's qubits.
uint64_t reg_mask(uint32_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

void check_register(const ncsv_state *s, uint64_t x, uint64_t z, const char *what) {
    if ((x | z) & ~reg_mask(s->n)) fail(NCSV_INVALID, std::string(what) + ": operator on the wrong register");
}

// NCSV_DEVICE_MEMORY_LIMIT_GB caps the bytes one state may take (testing the
// out-of-memory path without filling the device); unset, no cap.
// NCSV_DEVICE_MEMORY_LIMIT_GB caps one state's allocation, in units of 2^30
// bytes (a state of exactly the cap is allowed).
void check_memory_cap(size_t bytes) {
    const char *e = getenv("NCSV_DEVICE_MEMORY_LIMIT_GB");
    if (!e) return;
    const double cap = atof(e) * (double)(1ull << 30);
    if ((double)bytes > cap)
        fail(NCSV_DEVICE, "out of device memory: " + std::to_string(bytes) + " bytes requested, NCSV_DEVICE_MEMORY_LIMIT_GB=" +
                              std::string(e));
}

ncsv_state *alloc_state(uint32_t n, ncsv_precision prec) {
    if (n > 34) fail(NCSV_INVALID, "state vector of more than 34 qubits");
    if (prec != NCSV_C128 && prec != NCSV_C64) fail(NCSV_INVALID, "unknown precision");
    check_memory_cap((size_t(1) << n) * (prec == NCSV_C128 ? 16 : 8));
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) fail(NCSV_DEVICE, "no GPU device visible");
    auto s = std::make_unique<ncsv_state>();
    s->n = n;
    s->prec = prec;
    check_hip(hipGetDevice(&s->device), "hipGetDevice");
    check_hip(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreate(&s->ev0), "hipEventCreate");
    check_hip(hipEventCreate(&s->ev1), "hipEventCreate");
    const hipError_t e = hipMalloc(&s->amp, s->amp_bytes());
    if (e != hipSuccess) {
        s->amp = nullptr;
        (void)hipStreamDestroy(s->stream);
        fail(NCSV_DEVICE, "cannot allocate " + std::to_string(s->amp_bytes()) + " bytes of device memory");
    }
    dev::ensure_part(s.get(), 64);
    return s.release();
}

void free_state(ncsv_state *s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->amp) (void)hipFree(s->amp);
    if (s->spare) (void)hipFree(s->spare);
    for (int k = 0; k < 2; k++) {
        if (s->stage[k]) (void)hipHostFree(s->stage[k]);
        if (s->staged[k]) (void)hipEventDestroy(s->staged[k]);
    }
    if (s->d_part) (void)hipFree(s->d_part);
    if (s->d_tab) (void)hipFree(s->d_tab);
    if (s->d_samp) (void)hipFree(s->d_samp);
    if (s->h_part) (void)hipHostFree(s->h_part);
    fuse_release(s);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

double lsq(const ncsv_state *s) { return s->lscale * s->lscale; }

// project with the logical weight; the complex64 form renormalizes its
// stored amplitudes in FP64 terms when the branch weight gets small.
double do_project(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, int outcome) {
    if (!pauli_is_hermitian(x, z, ph)) fail(NCSV_INVALID, "projector of non-Hermitian Pauli");
    check_register(s, x, z, "project");
    const double w = dev::project(s, x, z, ph, outcome);
    const double logical = w * lsq(s);
    if (s->prec == NCSV_C64 && w > 0.0 && w < 0.25) {
        const double r = std::sqrt(w);
        dev::scale(s, cd(1.0 / r, 0.0));
        s->lscale *= r;
    }
    return logical;
}

// QuantumState::measure (mod.rs:113-131).
int do_measure(ncsv_state *s, uint64_t x, uint64_t z, uint32_t ph, ncsv_uniform_fn rng, void *ctx) {
    if (!pauli_is_hermitian(x, z, ph)) fail(NCSV_INVALID, "measured operator is not Hermitian");
    check_register(s, x, z, "measure");
    double e = 0, nn = 0;
    dev::expect_and_norm(s, x, z, ph, &e, &nn);
    // f64::clamp keeps a NaN (a zero state: 0/0), so neither certainty test
    // holds, one uniform is drawn and `u >= NaN` gives outcome 0 (mod.rs:122-128)
    double p_plus = (1.0 + e / nn) / 2.0;
    if (!std::isnan(p_plus)) p_plus = std::min(1.0, std::max(0.0, p_plus));
    int outcome;
    if (p_plus > 1.0 - 1e-12)
        outcome = 0;
    else if (p_plus < 1e-12)
        outcome = 1;
    else {
        if (!rng) fail(NCSV_INVALID, "measure: uncertain outcome and no random source");
        outcome = rng(ctx) >= p_plus ? 1 : 0;
    }
    do_project(s, x, z, ph, outcome);
    return outcome;
}

void do_scale(ncsv_state *s, cd c) {
    if (s->prec == NCSV_C64) {
        const double a = std::abs(c);
        if (a == 0.0) {
            dev::scale(s, cd(0.0, 0.0));
            return;
        }
        s->lscale *= a;
        if (c != cd(a, 0.0)) dev::scale(s, c / a);
        return;
    }
    dev::scale(s, c);
}

// Op::Noise (mod.rs:148-161) with Noise::sample (circuit.rs:136-146).
void do_noise(ncsv_state *s, const COp &op, ncsv_uniform_fn rng, void *ctx) {
    if (!rng) fail(NCSV_INVALID, "noise needs a random source");
    double u = rng(ctx);
    for (auto &e : noise_errors(op.ch, op.p)) {
        if (u < e.p) {
            uint64_t x = 0, z = 0;
            for (size_t j = 0; j < op.qubits.size(); j++) {
                x |= (uint64_t)((e.x >> j) & 1) << op.qubits[j];
                z |= (uint64_t)((e.z >> j) & 1) << op.qubits[j];
            }
            dev::apply_pauli(s, x, z, 0);
            return;
        }
        u -= e.p;
    }
}

}  // namespace

namespace ncsv {
// The guard for entry points defined in other files (the batched form).
ncsv_status guard_status(const std::function<void()> &f) { return guard(f); }
}  // namespace ncsv

namespace ncsv {
void materialize(ncsv_state *s) {
    if (s->pending_basis < 0) return;
    // cleared after the fill went out: a failed fill leaves the basis pending
    dev::fill_basis(s, (uint64_t)s->pending_basis);
    s->pending_basis = -1;
}
}  // namespace ncsv

extern "C" {

size_t ncsv_last_error(char *buf, size_t cap) {
    if (buf && cap) {
        const size_t m = std::min(cap - 1, g_last_error.size());
        memcpy(buf, g_last_error.data(), m);
        buf[m] = 0;
    }
    return g_last_error.size();
}

ncsv_status ncsv_device_arch(char *buf, size_t cap) {
    return guard([&] {
        need(buf, "buffer");
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t p;
        check_hip(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
        const std::string a = std::string(p.gcnArchName) + " " + std::to_string(p.multiProcessorCount) + "cu";
        if (cap == 0 || a.size() + 1 > cap) fail(NCSV_INVALID, "buffer too small");
        memcpy(buf, a.c_str(), a.size() + 1);
    });
}

int ncsv_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

ncsv_status ncsv_create(uint32_t n, ncsv_precision prec, ncsv_state **out) {
    return guard([&] {
        need(out, "out");
        ncsv_state *s = alloc_state(n, prec);
        // |0...0>, written lazily as after set_basis
        s->pending_basis = 0;
        s->lscale = 1.0;
        *out = s;
    });
}

ncsv_status ncsv_destroy(ncsv_state *s) {
    return guard([&] { free_state(s); });
}

ncsv_status ncsv_clone(const ncsv_state *src, ncsv_state **out) {
    return guard([&] {
        need(src, "state");
        need(out, "out");
        std::lock_guard<std::recursive_mutex> lk(src->mu);
        materialize(const_cast<ncsv_state *>(src));
        ncsv_state *d = alloc_state(src->n, src->prec);
        try {
            dev::sync(src);
            dev::copy_state(src, d);
            d->fusion = src->fusion;
        } catch (...) {
            free_state(d);
            throw;
        }
        *out = d;
    });
}

ncsv_status ncsv_num_qubits(const ncsv_state *s, uint32_t *n) {
    return guard([&] {
        need(s, "state");
        need(n, "n");
        *n = s->n;
    });
}

ncsv_status ncsv_get_precision(const ncsv_state *s, ncsv_precision *p) {
    return guard([&] {
        need(s, "state");
        need(p, "precision");
        *p = s->prec;
    });
}

ncsv_status ncsv_set_basis(ncsv_state *s, uint64_t index) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        if (index >> s->n) fail(NCSV_INVALID, "basis index out of range");
        // written by the next run's first sweep, or before anything else reads the state
        s->pending_basis = (int64_t)index;
        s->lscale = 1.0;
    });
}

ncsv_status ncsv_set_plus(ncsv_state *s) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        s->pending_basis = -1;  // overwritten
        dev::fill_const(s, sqrt2_pow(-(int64_t)s->n).real());
        dev::sync(s);
    });
}

ncsv_status ncsv_set_amplitudes(ncsv_state *s, const double *amp, uint64_t len) {
    return guard([&] {
        need(s, "state");
        need(amp, "amplitudes");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        if (len != (1ull << s->n))
            fail(NCSV_INVALID, std::to_string(len) + " amplitudes for " + std::to_string(s->n) + " qubits");
        s->pending_basis = -1;  // overwritten
        dev::upload(s, amp);
    });
}

ncsv_status ncsv_to_vector(const ncsv_state *s, double *out, uint64_t len) {
    return guard([&] {
        need(s, "state");
        need(out, "out");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        if (len != (1ull << s->n)) fail(NCSV_INVALID, "to_vector: length must be 2^n");
        dev::download(s, out);
    });
}

ncsv_status ncsv_amplitude(const ncsv_state *s, uint64_t index, double out[2]) {
    return guard([&] {
        need(s, "state");
        need(out, "out");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        if (s->n < 64 && (index >> s->n)) fail(NCSV_INVALID, "basis index out of range");
        dev::read_amp(s, index, out);
    });
}

ncsv_status ncsv_apply_gate(ncsv_state *s, ncsv_gate gate, const int64_t *ip, const double *fp,
                            const uint32_t *qubits, uint32_t nq) {
    return guard([&] {
        need(s, "state");
        if (nq) need(qubits, "qubits");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        const GateDesc g = make_gate(gate, ip, fp, nq);
        const LOp op = lower_gate(g, qubits, nq, s->n);
        dev::apply_lop(s, op);
    });
}

ncsv_status ncsv_apply_matrix(ncsv_state *s, const double *m, const uint32_t *qubits, uint32_t k) {
    return guard([&] {
        need(s, "state");
        need(m, "matrix");
        if (k) need(qubits, "qubits");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        check_qubits(qubits, k, s->n);
        if (k > 12) fail(NCSV_UNSUPPORTED, "dense matrix on more than 12 qubits");
        LOp op;
        op.kind = OpKind::Dense;
        op.qubits.assign(qubits, qubits + k);
        const size_t d = size_t(1) << k;
        op.m.resize(d * d);
        for (size_t i = 0; i < d * d; i++) op.m[i] = cd(m[2 * i], m[2 * i + 1]);
        dev::apply_lop(s, op);
    });
}

ncsv_status ncsv_apply_monomial(ncsv_state *s, const uint32_t *qubits, uint32_t k, const uint32_t *images,
                                const double *phases) {
    return guard([&] {
        need(s, "state");
        need(images, "images");
        need(phases, "phases");
        if (k) need(qubits, "qubits");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        check_qubits(qubits, k, s->n);
        if (k > 4) fail(NCSV_UNSUPPORTED, "monomial gate on more than 4 qubits");
        const uint32_t d = 1u << k;
        LOp op;
        op.qubits.assign(qubits, qubits + k);
        op.img.assign(images, images + d);
        bool diagonal = true, perm = true;
        std::vector<bool> seen(d, false);
        for (uint32_t x = 0; x < d; x++) {
            if (images[x] >= d || seen[images[x]]) perm = false;
            if (images[x] < d) seen[images[x]] = true;
            if (images[x] != x) diagonal = false;
            op.m.push_back(cd(phases[2 * x], phases[2 * x + 1]));
        }
        if (!perm) fail(NCSV_INVALID, "monomial gate: images are not a permutation");
        op.kind = diagonal ? OpKind::Diag : OpKind::Perm;
        dev::apply_lop(s, op);
    });
}

ncsv_status ncsv_apply_pauli(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        check_register(s, x, z, "apply_pauli");
        dev::apply_pauli(s, x, z, phase & 3);
    });
}

ncsv_status ncsv_norm_sq(const ncsv_state *s, double out[2]) {
    return guard([&] {
        need(s, "state");
        need(out, "out");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        ncsv_state *m = const_cast<ncsv_state *>(s);
        out[0] = dev::norm_sq(m) * lsq(s);
        out[1] = 0.0;
    });
}

ncsv_status ncsv_expectation(const ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, double out[2]) {
    return ncsv_expectation_batch(s, 1, &x, &z, &phase, out);
}

ncsv_status ncsv_expectation_batch(const ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z,
                                   const uint32_t *phase, double *out) {
    return guard([&] {
        need(s, "state");
        if (count) {
            need(x, "x");
            need(z, "z");
            need(phase, "phase");
            need(out, "out");
        }
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        for (uint32_t i = 0; i < count; i++) check_register(s, x[i], z[i], "expectation");
        ncsv_state *m = const_cast<ncsv_state *>(s);
        dev::expectations(m, count, x, z, phase, out);
        for (uint32_t i = 0; i < 2 * count; i++) out[i] *= lsq(s);
    });
}

ncsv_status ncsv_expectation_sum(const ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z,
                                 const uint32_t *phase, const double *coeffs, double out[2]) {
    return guard([&] {
        need(out, "out");
        if (count) need(coeffs, "coefficients");
        std::vector<double> e(2 * (size_t)count);
        const ncsv_status st = ncsv_expectation_batch(s, count, x, z, phase, e.data());
        if (st != NCSV_OK) throw Error{st, g_last_error};
        // c_k <P_k> (mod.rs:98), then the pairwise sum of Amplitude::sum (scalar.rs:80-94)
        std::vector<cd> t(count);
        for (uint32_t i = 0; i < count; i++) {
            const cd c(coeffs[2 * i], coeffs[2 * i + 1]), v(e[2 * i], e[2 * i + 1]);
            t[i] = cd(c.real() * v.real() - c.imag() * v.imag(), c.real() * v.imag() + c.imag() * v.real());
        }
        std::function<cd(size_t, size_t)> rec = [&](size_t lo, size_t hi) -> cd {
            if (hi - lo <= 8) {
                cd acc(0.0, 0.0);
                for (size_t i = lo; i < hi; i++) acc += t[i];
                return acc;
            }
            const size_t mid = lo + (hi - lo) / 2;
            return rec(lo, mid) + rec(mid, hi);
        };
        const cd r = rec(0, count);
        out[0] = r.real();
        out[1] = r.imag();
    });
}

ncsv_status ncsv_inner(const ncsv_state *a, const ncsv_state *b, double out[2]) {
    return guard([&] {
        need(a, "state");
        need(b, "other");
        need(out, "out");
        if (a->n != b->n) fail(NCSV_INVALID, "inner: register sizes differ");
        std::unique_lock<std::recursive_mutex> l1(a->mu, std::defer_lock), l2(b->mu, std::defer_lock);
        if (a == b)
            l1.lock();
        else if (a < b) {
            l1.lock();
            l2.lock();
        } else {
            l2.lock();
            l1.lock();
        }
        materialize(const_cast<ncsv_state *>(a));
        materialize(const_cast<ncsv_state *>(b));
        dev::inner(const_cast<ncsv_state *>(a), const_cast<ncsv_state *>(b), out);
        out[0] *= a->lscale * b->lscale;
        out[1] *= a->lscale * b->lscale;
    });
}

ncsv_status ncsv_compare(const ncsv_state *a, const ncsv_state *b, double out[5]) {
    return guard([&] {
        need(a, "state");
        need(b, "other");
        need(out, "out");
        std::unique_lock<std::recursive_mutex> l1(a->mu, std::defer_lock), l2(b->mu, std::defer_lock);
        if (a == b)
            l1.lock();
        else if (a < b) {
            l1.lock();
            l2.lock();
        } else {
            l2.lock();
            l1.lock();
        }
        materialize(const_cast<ncsv_state *>(a));
        materialize(const_cast<ncsv_state *>(b));
        dev::compare(const_cast<ncsv_state *>(a), const_cast<ncsv_state *>(b), out);
    });
}

ncsv_status ncsv_probability(const ncsv_state *s, uint64_t index, double *out) {
    return guard([&] {
        need(s, "state");
        need(out, "out");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        if (s->n < 64 && (index >> s->n)) fail(NCSV_INVALID, "basis index out of range");
        double a[2];
        dev::read_amp(s, index, a);
        const double nn = dev::norm_sq(const_cast<ncsv_state *>(s)) * lsq(s);
        *out = (a[0] * a[0] + a[1] * a[1]) / nn;
    });
}

ncsv_status ncsv_project(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, int outcome, double weight[2]) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        const double w = do_project(s, x, z, phase & 3, outcome ? 1 : 0);
        if (weight) {
            weight[0] = w;
            weight[1] = 0.0;
        }
    });
}

ncsv_status ncsv_measure(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, ncsv_uniform_fn rng, void *ctx,
                         int *outcome) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        const int o = do_measure(s, x, z, phase & 3, rng, ctx);
        if (outcome) *outcome = o;
    });
}

ncsv_status ncsv_reset(ncsv_state *s, uint32_t qubit, ncsv_uniform_fn rng, void *ctx) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        check_qubits(&qubit, 1, s->n);
        if (do_measure(s, 0, 1ull << qubit, 0, rng, ctx)) dev::apply_pauli(s, 1ull << qubit, 0, 0);
    });
}

ncsv_status ncsv_scale(ncsv_state *s, const double c[2]) {
    return guard([&] {
        need(s, "state");
        need(c, "c");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        do_scale(s, cd(c[0], c[1]));
    });
}

ncsv_status ncsv_normalize(ncsv_state *s) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        const double nn = dev::norm_sq(s) * lsq(s);
        if (nn <= 0.0) fail(NCSV_ZERO_STATE, "state vanished: normalize");
        const double l = -std::log2(nn);
        cd c;
        if (std::fabs(l - std::round(l)) < 1e-9)
            c = sqrt2_pow((int64_t)std::round(l));
        else
            c = cd(1.0 / std::sqrt(nn), 0.0);
        do_scale(s, c);
    });
}

ncsv_status ncsv_sample(const ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out) {
    return guard([&] {
        need(s, "state");
        if (shots) {
            need(uniforms, "uniforms");
            need(out, "out");
        }
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        dev::sample(const_cast<ncsv_state *>(s), shots, uniforms, out);
    });
}

ncsv_status ncsv_sample_noalloc(const ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out) {
    return guard([&] {
        need(s, "state");
        if (shots) {
            need(uniforms, "uniforms");
            need(out, "out");
        }
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        materialize(const_cast<ncsv_state *>(s));
        dev::sample_noalloc(const_cast<ncsv_state *>(s), shots, uniforms, out);
    });
}

ncsv_status ncsv_sample_rng(const ncsv_state *s, uint64_t shots, ncsv_uniform_fn rng, void *ctx, uint64_t *out) {
    return guard([&] {
        if (!rng) fail(NCSV_INVALID, "rng is null");
        std::vector<double> us(shots);
        for (auto &u : us) u = rng(ctx);
        const ncsv_status st = ncsv_sample(s, shots, us.data(), out);
        if (st != NCSV_OK) throw Error{st, g_last_error};
    });
}

ncsv_status ncsv_tensor(const ncsv_state *a, const ncsv_state *b, ncsv_state **out) {
    return guard([&] {
        need(a, "state");
        need(b, "other");
        need(out, "out");
        if (a->prec != b->prec) fail(NCSV_INVALID, "tensor: states of different precision");
        const uint32_t n = a->n + b->n;
        if (n > 34) fail(NCSV_INVALID, "state vector of " + std::to_string(n) + " qubits");
        std::lock_guard<std::recursive_mutex> l1(a->mu);
        std::unique_lock<std::recursive_mutex> l2(b->mu, std::defer_lock);
        if (a != b) l2.lock();
        materialize(const_cast<ncsv_state *>(a));
        materialize(const_cast<ncsv_state *>(b));
        ncsv_state *o = alloc_state(n, a->prec);
        try {
            dev::kron(const_cast<ncsv_state *>(a), const_cast<ncsv_state *>(b), o);
        } catch (...) {
            free_state(o);
            throw;
        }
        *out = o;
    });
}

// ------------------------------------------------------------------ circuits

ncsv_status ncsv_circuit_new(uint32_t n, ncsv_circuit **out) {
    return guard([&] {
        need(out, "out");
        auto c = new ncsv_circuit();
        c->n = n;
        *out = c;
    });
}

ncsv_status ncsv_circuit_parse(const char *text, ncsv_circuit **out) {
    return guard([&] {
        need(text, "text");
        need(out, "out");
        *out = ncsv_parse_text(text);
    });
}

ncsv_status ncsv_circuit_destroy(ncsv_circuit *c) {
    return guard([&] { delete c; });
}

ncsv_status ncsv_circuit_num_qubits(const ncsv_circuit *c, uint32_t *n) {
    return guard([&] {
        need(c, "circuit");
        need(n, "n");
        *n = c->n;
    });
}

ncsv_status ncsv_circuit_num_measurements(const ncsv_circuit *c, uint64_t *m) {
    return guard([&] {
        need(c, "circuit");
        need(m, "m");
        *m = c->measurements;
    });
}

ncsv_status ncsv_circuit_num_gates(const ncsv_circuit *c, uint64_t *g) {
    return guard([&] {
        need(c, "circuit");
        need(g, "g");
        uint64_t k = 0;
        for (auto &op : c->ops) k += op.kind == CKind::Gate;
        *g = k;
    });
}

static void invalidate(ncsv_circuit *c) {
    std::lock_guard<std::mutex> lk(c->mu);
    c->plans.clear();
}

ncsv_status ncsv_circuit_push_gate(ncsv_circuit *c, ncsv_gate gate, const int64_t *ip, const double *fp,
                                   const uint32_t *qubits, uint32_t nq) {
    return guard([&] {
        need(c, "circuit");
        if (nq) need(qubits, "qubits");
        COp op;
        op.kind = CKind::Gate;
        op.g = make_gate(gate, ip, fp, nq);
        op.qubits.assign(qubits, qubits + nq);
        circuit_push(c, op);
        invalidate(c);
    });
}

ncsv_status ncsv_circuit_push_measure(ncsv_circuit *c, uint64_t x, uint64_t z, uint32_t phase) {
    return guard([&] {
        need(c, "circuit");
        COp op;
        op.kind = CKind::Measure;
        op.x = x;
        op.z = z;
        op.ph = phase & 3;
        circuit_push(c, op);
        invalidate(c);
    });
}

ncsv_status ncsv_circuit_push_reset(ncsv_circuit *c, uint32_t qubit) {
    return guard([&] {
        need(c, "circuit");
        COp op;
        op.kind = CKind::Reset;
        op.qubits = {qubit};
        circuit_push(c, op);
        invalidate(c);
    });
}

ncsv_status ncsv_circuit_push_noise(ncsv_circuit *c, ncsv_noise channel, const double *probs,
                                    const uint32_t *qubits, uint32_t nq) {
    return guard([&] {
        need(c, "circuit");
        need(probs, "probabilities");
        if (nq) need(qubits, "qubits");
        if ((int)channel < 0 || channel > NCSV_NOISE_DEPOLARIZE2) fail(NCSV_INVALID, "unknown noise channel");
        COp op;
        op.kind = CKind::Noise;
        op.ch = channel;
        op.p[0] = probs[0];
        if (channel == NCSV_NOISE_PAULI_CHANNEL_1) {
            op.p[1] = probs[1];
            op.p[2] = probs[2];
        }
        op.qubits.assign(qubits, qubits + nq);
        circuit_push(c, op);
        invalidate(c);
    });
}

ncsv_status ncsv_circuit_push_tick(ncsv_circuit *c) {
    return guard([&] {
        need(c, "circuit");
        COp op;
        op.kind = CKind::Tick;
        circuit_push(c, op);
    });
}

size_t ncsv_circuit_to_text(const ncsv_circuit *c, char *buf, size_t cap) {
    if (!c) return 0;
    const std::string t = ncsv_circuit_text(c);
    if (buf && cap) {
        const size_t m = std::min(cap - 1, t.size());
        memcpy(buf, t.data(), m);
        buf[m] = 0;
    }
    return t.size();
}

// QuantumState::run (mod.rs:168-183) with apply_op (134-165).
ncsv_status ncsv_run(ncsv_state *s, const ncsv_circuit *cc, ncsv_uniform_fn rng, void *ctx, uint8_t *record,
                     uint64_t cap, uint64_t *nrecord) {
    return guard([&] {
        need(s, "state");
        need(cc, "circuit");
        ncsv_circuit *c = const_cast<ncsv_circuit *>(cc);
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        if (c->n != s->n)
            fail(NCSV_INVALID,
                 "circuit on " + std::to_string(c->n) + " qubits, state on " + std::to_string(s->n));
        uint64_t m = 0;
        size_t i = 0;
        const size_t nops = c->ops.size();
        while (i < nops) {
            const COp &op = c->ops[i];
            if (op.kind == CKind::Gate) {
                size_t j = i;
                while (j < nops && (c->ops[j].kind == CKind::Gate || c->ops[j].kind == CKind::Tick ||
                                    c->ops[j].kind == CKind::Detector || c->ops[j].kind == CKind::Observable))
                    j++;
                run_gate_segment(s, c, i, j);
                i = j;
                continue;
            }
            materialize(s);  // measurements, resets and noise read the amplitudes
            switch (op.kind) {
            case CKind::Measure: {
                const int o = do_measure(s, op.x, op.z, op.ph, rng, ctx);
                if (record && m < cap) record[m] = (uint8_t)o;
                m++;
                break;
            }
            case CKind::Reset:
                if (do_measure(s, 0, 1ull << op.qubits[0], 0, rng, ctx)) dev::apply_pauli(s, 1ull << op.qubits[0], 0, 0);
                break;
            case CKind::Noise: do_noise(s, op, rng, ctx); break;
            default: break;
            }
            i++;
        }
        dev::sync(s);
        if (nrecord) *nrecord = m;
    });
}

ncsv_status ncsv_set_fusion(ncsv_state *s, int mode) {
    return guard([&] {
        need(s, "state");
        if (mode < 0 || mode > 2) fail(NCSV_INVALID, "fusion mode must be 0, 1 or 2");
        s->fusion = mode;
    });
}

ncsv_status ncsv_profile_enable(ncsv_state *s, int on) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        dev::sync(s);
        s->profiling = on != 0;
    });
}

ncsv_status ncsv_profile_read(ncsv_state *s, ncsv_kernel_stat *out, uint32_t cap, uint32_t *count) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        uint32_t i = 0;
        for (auto &kv : s->stats) {
            if (out && i < cap) {
                memset(&out[i], 0, sizeof(out[i]));
                strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
                out[i].launches = kv.second.launches;
                out[i].total_ms = kv.second.ms;
                out[i].bytes = kv.second.bytes;
            }
            i++;
        }
        if (count) *count = i;
    });
}

ncsv_status ncsv_profile_reset(ncsv_state *s) {
    return guard([&] {
        need(s, "state");
        std::lock_guard<std::recursive_mutex> lk(s->mu);
        s->stats.clear();
        s->sweeps = 0;
    });
}

ncsv_status ncsv_sweep_count(const ncsv_state *s, uint64_t *sweeps) {
    return guard([&] {
        need(s, "state");
        need(sweeps, "sweeps");
        *sweeps = s->sweeps;
    });
}

ncsv_status ncsv_synchronize(const ncsv_state *s) {
    return guard([&] {
        need(s, "state");
        dev::sync(s);
    });
}

}  // extern "C"
