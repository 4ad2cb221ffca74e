/*
 * ncsv.h — C ABI of the MI355X (gfx950) dense state-vector library.
 *
 * Each entry point replaces one method of ncsim's `QuantumState<C64>` trait
 * as implemented by `StateVector<C64>`; the reference interface is cited on
 * every declaration (src/state/mod.rs = trait and provided methods,
 * src/state/vector.rs = the dense implementation).  A Rust `-sys` crate
 * declares these functions verbatim; a safe `HipStateVector` wrapper maps
 * `Gate`/`Pauli`/`Circuit` onto them.
 *
 * Conventions
 *  - Basis index bit q = qubit q (vector.rs:14).  A gate's local index has
 *    bit j on qubits[j], controls first (vector.rs:194-197).  Matrices are
 *    row-major, complex values interleaved (re, im) as double[2].
 *  - Pauli operators are i^phase X^x Z^z with 64-bit masks x, z (the
 *    reference reads word 0 only, vector.rs:352-355) and phase mod 4.
 *  - States are not normalized; `project` returns the branch weight
 *    (mod.rs:9-11, 63-66).  Outcome 1 (true) = eigenvalue -1.
 *  - Every function returns an ncsv_status; on failure a message is kept
 *    per thread (ncsv_last_error).  Nothing throws or aborts across the ABI.
 *  - Randomness: where the reference draws `rng.gen::<f64>()` the library
 *    calls the caller's `ncsv_uniform_fn` at exactly the same points, so a
 *    caller that passes its own generator keeps the reference's stream.
 *  - Threading: calls on one handle are serialized by the handle (the
 *    reference's `&self` methods may be called concurrently, mod.rs:94-99;
 *    here they are safe and run one at a time per handle).
 */
#ifndef NCSV_H
#define NCSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SimError variants (src/error.rs:7-42), plus device failures. */
typedef enum {
    NCSV_OK = 0,
    NCSV_UNSUPPORTED = 1,        /* SimError::Unsupported */
    NCSV_QUBIT_OUT_OF_RANGE = 2, /* SimError::QubitOutOfRange */
    NCSV_ARITY = 3,              /* SimError::Arity */
    NCSV_DUPLICATE_QUBIT = 4,    /* SimError::DuplicateQubit */
    NCSV_NOT_REPRESENTABLE = 5,  /* SimError::NotRepresentable */
    NCSV_ZERO_STATE = 6,         /* SimError::ZeroState */
    NCSV_PARSE = 7,              /* SimError::Parse */
    NCSV_INVALID = 8,            /* SimError::Invalid */
    NCSV_DEVICE = 9              /* no device, out of device memory, HIP error */
} ncsv_status;

/* Storage forms.  C128: complex128 on FP64 (the reference's C64).  C64:
 * complex64 amplitudes with FP64 reductions and a host FP64 scale factor. */
typedef enum { NCSV_C128 = 0, NCSV_C64 = 1 } ncsv_precision;

/* Named gates (src/gate.rs:355-431), in text-format order. */
typedef enum {
    NCSV_GATE_I = 0, NCSV_GATE_X, NCSV_GATE_Y, NCSV_GATE_Z, NCSV_GATE_H,
    NCSV_GATE_S, NCSV_GATE_S_DAG, NCSV_GATE_SQRT_X, NCSV_GATE_SQRT_X_DAG,
    NCSV_GATE_SQRT_Y, NCSV_GATE_SQRT_Y_DAG, NCSV_GATE_T, NCSV_GATE_T_DAG,
    NCSV_GATE_CX, NCSV_GATE_CY, NCSV_GATE_CZ, NCSV_GATE_SWAP, NCSV_GATE_CS,
    NCSV_GATE_CS_DAG, NCSV_GATE_CCZ, NCSV_GATE_CCX,
    NCSV_GATE_PHASE,   /* iparams = {k, bits}: diag(1, zeta_{2^bits}^k), bits <= 62 */
    NCSV_GATE_CPHASE,  /* iparams = {k, bits}: phase on |11> */
    NCSV_GATE_RX,      /* fparams = {theta} */
    NCSV_GATE_RY,
    NCSV_GATE_RZ,
    NCSV_GATE_UNITARY, /* fparams = 2^k x 2^k row-major interleaved matrix */
    NCSV_GATE_PAULI,   /* iparams = {x, z, phase}: local masks over the k qubits */
    NCSV_GATE_COUNT
} ncsv_gate;

/* Noise channels (src/circuit.rs:82-147). */
typedef enum {
    NCSV_NOISE_X_ERROR = 0, NCSV_NOISE_Y_ERROR, NCSV_NOISE_Z_ERROR,
    NCSV_NOISE_PAULI_CHANNEL_1, NCSV_NOISE_DEPOLARIZE1, NCSV_NOISE_DEPOLARIZE2
} ncsv_noise;

typedef struct ncsv_state ncsv_state;
typedef struct ncsv_circuit ncsv_circuit;

/* One uniform draw in [0, 1): the reference's rng.gen::<f64>(). */
typedef double (*ncsv_uniform_fn)(void *ctx);

/* ---- errors ------------------------------------------------------------ */

/* Copies the calling thread's last error message; returns its length. */
size_t ncsv_last_error(char *buf, size_t cap);
/* Number of visible GPUs (0 without a device; never fails). */
int ncsv_device_count(void);
/* The current device's architecture and CU count, e.g. "gfx950:sramecc+:xnack- 256cu"
 * (keys tuning caches; Device without a GPU). */
ncsv_status ncsv_device_arch(char *buf, size_t cap);

/* ---- lifetime and construction (vector.rs:24-64) ------------------------ */

/* StateVector::zero_state (vector.rs:25-27).  n <= 34 (device memory).  The
 * amplitudes are written lazily, as after ncsv_set_basis. */
ncsv_status ncsv_create(uint32_t n, ncsv_precision prec, ncsv_state **out);
ncsv_status ncsv_destroy(ncsv_state *s);
/* Clone (derive(Clone), vector.rs:15): a device-to-device copy. */
ncsv_status ncsv_clone(const ncsv_state *s, ncsv_state **out);
/* QuantumState::num_qubits (mod.rs:46). */
ncsv_status ncsv_num_qubits(const ncsv_state *s, uint32_t *n);
ncsv_status ncsv_get_precision(const ncsv_state *s, ncsv_precision *prec);
/* StateVector::basis_state (vector.rs:29-34).  The amplitudes are written
 * lazily: by the next ncsv_run's first sweep when it starts with gates,
 * otherwise by whichever call reads the state first; results are the same. */
ncsv_status ncsv_set_basis(ncsv_state *s, uint64_t index);
/* StateVector::plus_state (vector.rs:36-42): every amplitude sqrt2^-n. */
ncsv_status ncsv_set_plus(ncsv_state *s);
/* StateVector::from_amplitudes (vector.rs:44-52); Invalid unless len == 2^n. */
ncsv_status ncsv_set_amplitudes(ncsv_state *s, const double *amp, uint64_t len);
/* QuantumState::to_vector (vector.rs:542-544): len must be 2^n. */
ncsv_status ncsv_to_vector(const ncsv_state *s, double *out, uint64_t len);
/* QuantumState::amplitude (vector.rs:408-410). */
ncsv_status ncsv_amplitude(const ncsv_state *s, uint64_t index, double out[2]);

/* ---- gates (vector.rs:67-164, 364-395) ---------------------------------- */

/* QuantumState::apply_gate (vector.rs:364-395): the reference's dispatch
 * (Pauli, monomial view, matrix view).  iparams/fparams as per ncsv_gate. */
ncsv_status ncsv_apply_gate(ncsv_state *s, ncsv_gate gate, const int64_t *iparams, const double *fparams,
                            const uint32_t *qubits, uint32_t nq);
/* StateVector::apply_matrix (vector.rs:67-93): 2^k x 2^k row-major. */
ncsv_status ncsv_apply_matrix(ncsv_state *s, const double *m, const uint32_t *qubits, uint32_t k);
/* StateVector::apply_monomial (vector.rs:96-139) given the images of the
 * local basis states: U|x> = phases[x] |images[x]> (phases interleaved). */
ncsv_status ncsv_apply_monomial(ncsv_state *s, const uint32_t *qubits, uint32_t k, const uint32_t *images,
                                const double *phases);
/* StateVector::apply_pauli (vector.rs:143-164). */
ncsv_status ncsv_apply_pauli(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase);

/* ---- reductions (vector.rs:397-443; mod.rs:93-108) ---------------------- */

/* QuantumState::norm_sq (vector.rs:397-406). */
ncsv_status ncsv_norm_sq(const ncsv_state *s, double out[2]);
/* QuantumState::expectation (vector.rs:412-429). */
ncsv_status ncsv_expectation(const ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, double out[2]);
/* Many expectations over one read of the state: out[2k], out[2k+1] =
 * <P_k>; the batched form of expectation_sum's term loop (mod.rs:93-101). */
ncsv_status ncsv_expectation_batch(const ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z,
                                   const uint32_t *phase, double *out);
/* QuantumState::expectation_sum (mod.rs:93-101): sum_k c_k <P_k>, the
 * coefficients interleaved (2 doubles each), summed pairwise. */
ncsv_status ncsv_expectation_sum(const ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z,
                                 const uint32_t *phase, const double *coeffs, double out[2]);
/* QuantumState::inner (vector.rs:431-443): <a|b>. */
ncsv_status ncsv_inner(const ncsv_state *a, const ncsv_state *b, double out[2]);
/* No counterpart in the reference (for checking the complex64 form): two
 * states of the same size and any precisions, in one read of each:
 * out = {Re<a|b>, Im<a|b>, <a|a>, <b|b>, max_i |a_i - b_i|} (FP64). */
ncsv_status ncsv_compare(const ncsv_state *a, const ncsv_state *b, double out[5]);
/* QuantumState::probability (mod.rs:104-108). */
ncsv_status ncsv_probability(const ncsv_state *s, uint64_t index, double *out);

/* ---- collapse (vector.rs:445-488; mod.rs:113-165, 220-233) -------------- */

/* QuantumState::project (vector.rs:445-479): Invalid for a non-Hermitian P;
 * returns the weight, no renormalization. */
ncsv_status ncsv_project(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, int outcome, double weight[2]);
/* QuantumState::measure (mod.rs:113-131): one draw only when uncertain. */
ncsv_status ncsv_measure(ncsv_state *s, uint64_t x, uint64_t z, uint32_t phase, ncsv_uniform_fn rng, void *ctx,
                         int *outcome);
/* Op::Reset (mod.rs:141-147): measure Z_q, then X if the outcome is 1. */
ncsv_status ncsv_reset(ncsv_state *s, uint32_t qubit, ncsv_uniform_fn rng, void *ctx);
/* QuantumState::scale (vector.rs:481-488). */
ncsv_status ncsv_scale(ncsv_state *s, const double c[2]);
/* QuantumState::normalize (mod.rs:220-233). */
ncsv_status ncsv_normalize(ncsv_state *s);

/* ---- sampling and tensor products ---------------------------------------- */

/* QuantumState::sample (vector.rs:494-540) with the shots' draws given:
 * uniforms[k] in [0, 1) in draw order; out[k] = basis index of shot k. */
ncsv_status ncsv_sample(const ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out);
/* No counterpart in the reference: ncsv_sample for callers that cannot
 * allocate device memory during the call (serving loops, harnesses that
 * forbid hipMalloc).  The same picks as ncsv_sample; the scratch is a fixed
 * 32 MiB array per GPU held by the library until the process ends, so
 * 2^(n-12) + 2 + 2 x shots must be at most 4,194,304 (Invalid otherwise;
 * e.g. about 1.9 million shots at 30 qubits).  Calls on one GPU share the
 * array and are serialized by the library (overlapping calls would race).
 * A deferred basis state is still written first, as by every reader. */
ncsv_status ncsv_sample_noalloc(const ncsv_state *s, uint64_t shots, const double *uniforms, uint64_t *out);
/* The same with the draws taken from `rng` (shots draws, in order). */
ncsv_status ncsv_sample_rng(const ncsv_state *s, uint64_t shots, ncsv_uniform_fn rng, void *ctx, uint64_t *out);
/* QuantumState::tensor / StateVector::kron (vector.rs:168-188, 546-548):
 * a on the low qubits.  Both states must have the same precision. */
ncsv_status ncsv_tensor(const ncsv_state *a, const ncsv_state *b, ncsv_state **out);

/* ---- circuits (src/circuit.rs) and run (mod.rs:134-183) ------------------ */

/* Circuit::new (circuit.rs:156-160). */
ncsv_status ncsv_circuit_new(uint32_t n, ncsv_circuit **out);
/* Circuit::parse (circuit.rs:459-555): the Stim-like text format. */
ncsv_status ncsv_circuit_parse(const char *text, ncsv_circuit **out);
ncsv_status ncsv_circuit_destroy(ncsv_circuit *c);
ncsv_status ncsv_circuit_num_qubits(const ncsv_circuit *c, uint32_t *n);
ncsv_status ncsv_circuit_num_measurements(const ncsv_circuit *c, uint64_t *m);
ncsv_status ncsv_circuit_num_gates(const ncsv_circuit *c, uint64_t *g);
/* Circuit::push with Op::Gate (circuit.rs:177-229): checked. */
ncsv_status ncsv_circuit_push_gate(ncsv_circuit *c, ncsv_gate gate, const int64_t *iparams, const double *fparams,
                                   const uint32_t *qubits, uint32_t nq);
/* Op::Measure of a Hermitian Pauli on the whole register. */
ncsv_status ncsv_circuit_push_measure(ncsv_circuit *c, uint64_t x, uint64_t z, uint32_t phase);
/* Op::Reset. */
ncsv_status ncsv_circuit_push_reset(ncsv_circuit *c, uint32_t qubit);
/* Op::Noise: probs has 1 entry, or 3 for PAULI_CHANNEL_1. */
ncsv_status ncsv_circuit_push_noise(ncsv_circuit *c, ncsv_noise channel, const double *probs,
                                    const uint32_t *qubits, uint32_t nq);
/* Op::Tick. */
ncsv_status ncsv_circuit_push_tick(ncsv_circuit *c);
/* Circuit's Display (circuit.rs:693-728); returns the full length. */
size_t ncsv_circuit_to_text(const ncsv_circuit *c, char *buf, size_t cap);

/* QuantumState::run (mod.rs:168-183): applies every op in order (gates
 * fused into device sweeps between measurements), writes the measurement
 * record (record[i] = outcome of the i-th measurement; cap entries at most)
 * and its length. */
ncsv_status ncsv_run(ncsv_state *s, const ncsv_circuit *c, ncsv_uniform_fn rng, void *ctx, uint8_t *record,
                     uint64_t cap, uint64_t *nrecord);

/* ---- execution options and counters --------------------------------------- */

/* Gate fusion for ncsv_run: 0 = one device sweep per gate (the reference's
 * schedule); 1 = fused tile sweeps (default), run by kernels specialized per
 * sweep (compiled at run time, cached) from 20 qubits on (NCSV_JIT_MIN_QUBITS)
 * and by the op-list interpreter kernel below that; 2 = fused, always
 * specialized.  NCSV_JIT=0 disables specialization. */
ncsv_status ncsv_set_fusion(ncsv_state *s, int mode);
/* Device-kernel timing with HIP events on the state's stream. */
typedef struct {
    char name[32];
    uint64_t launches;
    double total_ms;  /* sum of event-timed durations */
    double bytes;     /* algorithmic HBM bytes of those launches */
} ncsv_kernel_stat;
ncsv_status ncsv_profile_enable(ncsv_state *s, int on);
/* Fills up to cap entries; returns the number of kernel kinds seen in *count. */
ncsv_status ncsv_profile_read(ncsv_state *s, ncsv_kernel_stat *out, uint32_t cap, uint32_t *count);
ncsv_status ncsv_profile_reset(ncsv_state *s);
/* Number of device sweeps (full passes over the amplitudes) run so far. */
ncsv_status ncsv_sweep_count(const ncsv_state *s, uint64_t *sweeps);
/* Waits for the state's stream. */
ncsv_status ncsv_synchronize(const ncsv_state *s);

/* Diagnostics (no counterpart in the reference): the fused schedule that
 * ncsv_run builds for a unitary circuit on a state of `prec`, applied on the
 * host to amp (2^n interleaved doubles, may be NULL) in FP64 with the device
 * ops' semantics; stats[0..3] = sweeps, stages, ops, standalone gates.
 * Lets the planner be checked without a device. */
ncsv_status ncsv_plan_emulate(const ncsv_circuit *c, ncsv_precision prec, double *amp, uint64_t *stats);
/* Diagnostics (no counterpart in the reference): generates the specialized
 * kernel of every tile sweep of a unitary circuit's plan and compiles it for
 * `arch` (e.g. "gfx950") without a device; *nsweeps = kernels compiled. */
ncsv_status ncsv_jit_check(const ncsv_circuit *c, ncsv_precision prec, const char *arch, uint32_t *nsweeps);
/* Diagnostics (no counterpart in the reference): the specialized
 * expectation kernels for `count` Paulis on n qubits, compiled for `arch`
 * without a device; *ngroups = kernels; Unsupported when some X part does
 * not fit a tile's register stage. */
ncsv_status ncsv_jit_check_expect(uint32_t n, ncsv_precision prec, uint32_t count, const uint64_t *x,
                                  const uint64_t *z, const uint32_t *ph, const char *arch, uint32_t *ngroups);

/* ------------------------------------------------------------------ batched form
 * Many circuits or noise trajectories, each run from |0...0> as
 * QuantumState::run does (src/state/mod.rs:168-183 per circuit, with
 * apply_op's Measure/Reset/Noise arms, mod.rs:134-165) and then evaluated on
 * a list of Paulis (StateVector::expectation, vector.rs:412-429), all in one
 * persistent launch per chunk of device memory.  Entries of up to 13 qubits
 * (complex64; 12 for complex128) live in one workgroup's LDS for the whole
 * circuit.
 *
 * Draws: every Measure, Reset and Noise op of an entry takes exactly one
 * uniform in [0, 1), in program order (the single-state run skips the draw
 * of a certain measurement; here the slot is taken either way).  A Noise op
 * picks its error as Noise::sample does (circuit.rs:136-146).  Measurements
 * must be Z-type Pauli products (M, MZ, MPP of Z factors) and gates of more
 * than two qubits must be diagonal or permutations (CCZ, CCX): anything else
 * is Unsupported.  A batch keeps pointers to its circuits until its next run
 * has planned them. */
typedef struct ncsv_batch ncsv_batch;

ncsv_batch *ncsv_batch_new(ncsv_precision precision);
void ncsv_batch_destroy(ncsv_batch *b);
/* Adds a circuit with `count` Paulis (x, z, phase as in ncsv_expectation);
 * *index = its entry number. */
ncsv_status ncsv_batch_add(ncsv_batch *b, const ncsv_circuit *c, uint32_t count, const uint64_t *x,
                           const uint64_t *z, const uint32_t *phase, uint32_t *index);
uint32_t ncsv_batch_size(const ncsv_batch *b);
/* The uniforms entry `entry` takes per run. */
ncsv_status ncsv_batch_draws(const ncsv_batch *b, uint32_t entry, uint32_t *count);
/* Tile size in qubits (6..13 complex64, 6..12 complex128; 0 = largest) and
 * the persistent grid (0 = every resident workgroup). */
ncsv_status ncsv_batch_set_tile_bits(ncsv_batch *b, uint32_t bits);
ncsv_status ncsv_batch_set_blocks(ncsv_batch *b, int32_t blocks);
/* The tile's fixed low qubits (coalescing against fewer sweeps; -1 = 5 for
 * complex64, 4 for complex128) and the most tiles per queue item (-1 = 32). */
ncsv_status ncsv_batch_set_shape(ncsv_batch *b, int32_t low_bits, int32_t tiles_per_item);
/* Runs every entry.  `uniforms` holds each entry's draws back to back (in
 * entry order), or is NULL and entry e's draws come from a SplitMix64 stream
 * seeded with seeds[e] (gen_f64 = (next >> 11) * 2^-53).  `forced`, laid out
 * like the uniforms (or NULL), fixes measurement outcomes: 0 or 1 forces,
 * 0xff draws. */
ncsv_status ncsv_batch_run(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced);
/* Entry `entry` of the last run: expect[2k], expect[2k+1] = Re, Im <P_k>
 * (unnormalized, as StateVector::expectation), *norm_sq = <psi|psi>, and
 * outcomes[d] = the outcome taken at draw d (0xff for noise draws). */
ncsv_status ncsv_batch_result(const ncsv_batch *b, uint32_t entry, double *expect, double *norm_sq,
                              uint8_t *outcomes);
/* out[9]: kernel ms, planning ms, queue items processed, launches, amplitude
 * updates (gates x 2^n summed), HBM bytes of the multi-tile states, programs,
 * program steps, launches replayed from a captured graph. */
ncsv_status ncsv_batch_stats(const ncsv_batch *b, double *out);
/* Plans every entry not planned yet (runs do this themselves; no device). */
ncsv_status ncsv_batch_plan(ncsv_batch *b);
/* Diagnostics: runs the same programs on FP64 host amplitudes (no device). */
ncsv_status ncsv_batch_emulate(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced);
/* Diagnostics: runs the device kernel's own code on the host, one workgroup
 * simulated by host threads (no device; small batches only). */
ncsv_status ncsv_batch_simulate(ncsv_batch *b, const double *uniforms, const uint64_t *seeds, const uint8_t *forced);

/* ------------------------------------------------------------ frame sampler
 * Pauli-frame sampling of noisy Clifford circuits (src/frame.rs
 * FrameSampler over the column layout of src/pauli/columns.rs): one
 * noiseless reference record from a stabilizer tableau, then per shot that
 * record XOR the anticommutation of an accumulated Pauli frame with each
 * measured Pauli; 64 shots per 64-bit word.  Draws are counter-based: the
 * caller's seed plays the part of the reference's `rng.next_u64()` per
 * sample call, and results do not depend on how shots are split over
 * threads.  Measured Paulis lie on qubits 0..63 (as the circuit format
 * allows); gates and noise on any qubit. */
typedef struct ncsv_frames ncsv_frames;

/* FrameSampler::new (frame.rs:41-58): Unsupported for a non-Clifford gate;
 * the reference record's random outcomes take the top bits of successive
 * SplitMix64 outputs of reference_seed. */
ncsv_status ncsv_frames_new(const ncsv_circuit *c, uint64_t reference_seed, ncsv_frames **out);
ncsv_status ncsv_frames_destroy(ncsv_frames *f);
/* Circuit::num_measurements, detectors().len(), observables().len()
 * (circuit.rs:297-327) of the sampler's circuit. */
ncsv_status ncsv_frames_shape(const ncsv_frames *f, uint64_t *measurements, uint64_t *detectors,
                              uint64_t *observables);
/* FrameSampler::reference (frame.rs:66-69): record[k] = outcome k. */
ncsv_status ncsv_frames_reference(const ncsv_frames *f, uint8_t *record, uint64_t cap);
/* FrameSampler::sample_packed (frame.rs:72-92): out[k * W + w] holds
 * measurement k of shots 64w..64w+63 (bit s = shot 64w + s), W =
 * ceil(shots / 64), bits past the last shot zero. */
ncsv_status ncsv_frames_sample_packed(ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *out);
/* FrameSampler::sample_detectors (frame.rs:150-211): detector and
 * observable flips, laid out as sample_packed's records. */
ncsv_status ncsv_frames_sample_detectors(ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *det,
                                         uint64_t *obs);
/* Diagnostics (no counterpart in the reference): sample_packed run on the
 * host by the kernel's own per-word code (no device). */
ncsv_status ncsv_frames_emulate_packed(const ncsv_frames *f, uint64_t shots, uint64_t seed, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif /* NCSV_H */
