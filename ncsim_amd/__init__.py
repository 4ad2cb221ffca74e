"""ncsim_amd — MI355X (gfx950) dense state vectors for ncsim.

`HipStateVector` mirrors ncsim's `QuantumState<C64>` trait as implemented by
`StateVector<C64>` (src/state/mod.rs:41-234, src/state/vector.rs:357-549):
same method names, argument meaning, conventions and error classes.  Every
method calls the C ABI of libncsv.so (include/ncsv.h); the computation runs in
hand-written HIP kernels on the GPU.  There is no CPU path: without the
library or a GPU the calls raise.

Gates are `Gate` values or text-format names; Paulis are any object with
`n`, `x`, `z`, `phase` (i^phase X^x Z^z, bit q = qubit q); circuits are
`Circuit` values (the reference's text format, circuit.rs:4-27); random
sources are objects with `gen_f64()` returning uniforms in [0, 1), drawn
where the reference draws from its rng.
"""
import ctypes

import numpy as np

from . import _lib as L

__all__ = ["HipStateVector", "Circuit", "Gate", "Pauli", "NcsvError", "C128", "C64", "device_count"]

C128, C64 = 0, 1

_ERR = {1: "Unsupported", 2: "QubitOutOfRange", 3: "Arity", 4: "DuplicateQubit", 5: "NotRepresentable",
        6: "ZeroState", 7: "Parse", 8: "Invalid", 9: "Device"}

GATE_IDS = {name: i for i, name in enumerate(
    ["I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG", "T", "T_DAG", "CX",
     "CY", "CZ", "SWAP", "CS", "CS_DAG", "CCZ", "CCX", "PHASE", "CPHASE", "RX", "RY", "RZ", "UNITARY", "PAULI"])}
GATE_IDS["CNOT"], GATE_IDS["TOFFOLI"] = GATE_IDS["CX"], GATE_IDS["CCX"]
NOISE_IDS = {"X_ERROR": 0, "Y_ERROR": 1, "Z_ERROR": 2, "PAULI_CHANNEL_1": 3, "DEPOLARIZE1": 4, "DEPOLARIZE2": 5}


class NcsvError(Exception):
    """A failed call: `kind` names the SimError variant (src/error.rs:7-42)."""

    def __init__(self, code, msg):
        self.code, self.kind = code, _ERR.get(code, str(code))
        super().__init__(f"{self.kind}: {msg}")


def _check(status):
    if status != 0:
        buf = ctypes.create_string_buffer(1024)
        L.lib().ncsv_last_error(buf, 1024)
        raise NcsvError(status, buf.value.decode(errors="replace"))


def device_count():
    return L.lib().ncsv_device_count()


class Pauli:
    """i^phase X^x Z^z on n qubits (src/pauli/mod.rs:79-86)."""

    __slots__ = ("n", "x", "z", "phase")

    def __init__(self, n, x=0, z=0, phase=0):
        self.n, self.x, self.z, self.phase = n, x, z, phase & 3

    @staticmethod
    def z_(n, q):
        return Pauli(n, 0, 1 << q, 0)


class Gate:
    """A gate as data (src/gate.rs:355-431): a text-format name plus its
    parameters (PHASE/CPHASE: k, bits; RX/RY/RZ: theta; UNITARY: matrix;
    PAULI: pauli with n, x, z, phase)."""

    def __init__(self, name, k=None, bits=None, theta=None, matrix=None, pauli=None):
        self.name = name.upper()
        self.k, self.bits, self.theta, self.matrix, self.pauli = k, bits, theta, matrix, pauli

    @staticmethod
    def of(g):
        """Accepts a Gate, a duck-typed gate (name/k/bits/theta/mat/pauli) or a name."""
        if isinstance(g, Gate):
            return g
        if isinstance(g, str):
            return Gate(g)
        return Gate(g.name, getattr(g, "k", None), getattr(g, "bits", None), getattr(g, "theta", None),
                    getattr(g, "mat", getattr(g, "matrix", None)), getattr(g, "pauli", None))

    def encode(self, nq):
        """(gate id, iparams, fparams) for the C ABI."""
        if self.name not in GATE_IDS:
            raise NcsvError(5, f"gate {self.name} has no device lowering")
        gid = GATE_IDS[self.name]
        ip = fp = None
        if self.name in ("PHASE", "CPHASE"):
            ip = (ctypes.c_int64 * 2)(int(self.k), int(self.bits))
        elif self.name in ("RX", "RY", "RZ"):
            fp = (ctypes.c_double * 1)(float(self.theta))
        elif self.name == "UNITARY":
            m = np.ascontiguousarray(self.matrix, dtype=np.complex128).reshape(-1)
            fp = (ctypes.c_double * (2 * len(m)))(*m.view(np.float64))
        elif self.name == "PAULI":
            p = self.pauli
            ip = (ctypes.c_int64 * 3)(int(p.x), int(p.z), int(p.phase))
        return gid, ip, fp


def _qubits(qs):
    qs = [int(q) for q in qs]
    return (ctypes.c_uint32 * max(1, len(qs)))(*qs), len(qs)


class _Rng:
    """Wraps an object with gen_f64() as the C callback (kept alive by the caller)."""

    def __init__(self, rng):
        self.rng = rng
        self.err = None

        def cb(_ctx):
            try:
                return float(rng.gen_f64())
            except Exception as e:  # noqa: BLE001 — surfaced after the call
                self.err = e
                return 0.0

        self.fn = L.UNIFORM_FN(cb) if rng is not None else L.UNIFORM_FN()

    def raise_if(self):
        if self.err is not None:
            raise self.err


def _m64(v):
    return int(v) & 0xFFFFFFFFFFFFFFFF


class Circuit:
    """A circuit (src/circuit.rs:149-565) held by the library: build it with
    the push methods or parse the text format."""

    def __init__(self, n=None, _handle=None):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            _check(L.lib().ncsv_circuit_new(int(n), ctypes.byref(self._h)))

    @staticmethod
    def parse(text):
        h = ctypes.c_void_p()
        _check(L.lib().ncsv_circuit_parse(text.encode(), ctypes.byref(h)))
        return Circuit(_handle=h)

    def __del__(self):
        if getattr(self, "_h", None) and L._lib is not None:
            L._lib.ncsv_circuit_destroy(self._h)
            self._h = None

    def num_qubits(self):
        v = ctypes.c_uint32()
        _check(L.lib().ncsv_circuit_num_qubits(self._h, ctypes.byref(v)))
        return v.value

    def num_measurements(self):
        v = ctypes.c_uint64()
        _check(L.lib().ncsv_circuit_num_measurements(self._h, ctypes.byref(v)))
        return v.value

    def num_gates(self):
        v = ctypes.c_uint64()
        _check(L.lib().ncsv_circuit_num_gates(self._h, ctypes.byref(v)))
        return v.value

    def gate(self, g, qubits):
        g = Gate.of(g)
        qa, nq = _qubits(qubits)
        gid, ip, fp = g.encode(nq)
        _check(L.lib().ncsv_circuit_push_gate(self._h, gid, ip, fp, qa, nq))
        return self

    def measure(self, p):
        _check(L.lib().ncsv_circuit_push_measure(self._h, _m64(p.x), _m64(p.z), p.phase & 3))
        return self

    def measure_z(self, q):
        return self.measure(Pauli.z_(self.num_qubits(), q))

    def reset(self, q):
        _check(L.lib().ncsv_circuit_push_reset(self._h, int(q)))
        return self

    def noise(self, name, probs, qubits):
        qa, nq = _qubits(qubits)
        pa = (ctypes.c_double * 3)(*(list(probs) + [0.0, 0.0, 0.0])[:3])
        _check(L.lib().ncsv_circuit_push_noise(self._h, NOISE_IDS[name.upper()], pa, qa, nq))
        return self

    def tick(self):
        _check(L.lib().ncsv_circuit_push_tick(self._h))
        return self

    def __str__(self):
        n = L.lib().ncsv_circuit_to_text(self._h, None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.lib().ncsv_circuit_to_text(self._h, buf, n + 1)
        return buf.value.decode()


class HipStateVector:
    """A dense state on the GPU: `QuantumState<C64>` for `StateVector<C64>`."""

    BACKEND = "HipStateVector"

    def __init__(self, handle):
        self._h = handle

    # construction (vector.rs:24-52)
    @staticmethod
    def zero_state(n, precision=C128):
        h = ctypes.c_void_p()
        _check(L.lib().ncsv_create(int(n), int(precision), ctypes.byref(h)))
        return HipStateVector(h)

    @staticmethod
    def basis_state(n, x, precision=C128):
        s = HipStateVector.zero_state(n, precision)
        _check(L.lib().ncsv_set_basis(s._h, int(x)))
        return s

    @staticmethod
    def plus_state(n, precision=C128):
        s = HipStateVector.zero_state(n, precision)
        _check(L.lib().ncsv_set_plus(s._h))
        return s

    @staticmethod
    def from_amplitudes(n, amp, precision=C128):
        a = np.ascontiguousarray(amp, dtype=np.complex128)
        if len(a) != 1 << n:
            raise NcsvError(8, f"{len(a)} amplitudes for {n} qubits")
        s = HipStateVector.zero_state(n, precision)
        _check(L.lib().ncsv_set_amplitudes(s._h, a.view(np.float64).ctypes.data_as(L.c_dp), len(a)))
        return s

    def __del__(self):
        if getattr(self, "_h", None) and L._lib is not None:
            L._lib.ncsv_destroy(self._h)
            self._h = None

    def clone(self):
        h = ctypes.c_void_p()
        _check(L.lib().ncsv_clone(self._h, ctypes.byref(h)))
        return HipStateVector(h)

    def num_qubits(self):
        v = ctypes.c_uint32()
        _check(L.lib().ncsv_num_qubits(self._h, ctypes.byref(v)))
        return v.value

    @property
    def n(self):
        return self.num_qubits()

    # gates (vector.rs:67-164, 364-395)
    def apply_gate(self, gate, qubits):
        g = Gate.of(gate)
        qa, nq = _qubits(qubits)
        gid, ip, fp = g.encode(nq)
        _check(L.lib().ncsv_apply_gate(self._h, gid, ip, fp, qa, nq))

    def apply_matrix(self, m, qubits):
        m = np.ascontiguousarray(m, dtype=np.complex128).reshape(-1)
        qa, nq = _qubits(qubits)
        _check(L.lib().ncsv_apply_matrix(self._h, m.view(np.float64).ctypes.data_as(L.c_dp), qa, nq))

    def apply_monomial(self, qubits, images, phases):
        qa, nq = _qubits(qubits)
        im = (ctypes.c_uint32 * len(images))(*[int(v) for v in images])
        ph = np.ascontiguousarray(phases, dtype=np.complex128)
        _check(L.lib().ncsv_apply_monomial(self._h, qa, nq, im, ph.view(np.float64).ctypes.data_as(L.c_dp)))

    def apply_pauli(self, p):
        if p.n != self.num_qubits():
            raise NcsvError(8, "apply_pauli: register size")
        _check(L.lib().ncsv_apply_pauli(self._h, _m64(p.x), _m64(p.z), p.phase & 3))

    # reductions (vector.rs:397-443)
    def norm_sq(self):
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_norm_sq(self._h, out))
        return complex(out[0], out[1])

    def amplitude(self, x):
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_amplitude(self._h, int(x), out))
        return complex(out[0], out[1])

    def expectation(self, p):
        if p.n != self.num_qubits():
            raise NcsvError(8, "expectation: operator on the wrong register")
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_expectation(self._h, _m64(p.x), _m64(p.z), p.phase & 3, out))
        return complex(out[0], out[1])

    def expectations(self, paulis):
        """Many <P_k> over one pass per group of Paulis."""
        k = len(paulis)
        xs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.x) for p in paulis])
        zs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.z) for p in paulis])
        ph = (ctypes.c_uint32 * max(1, k))(*[p.phase & 3 for p in paulis])
        out = (ctypes.c_double * max(2, 2 * k))()
        _check(L.lib().ncsv_expectation_batch(self._h, k, xs, zs, ph, out))
        return [complex(out[2 * i], out[2 * i + 1]) for i in range(k)]

    def expectation_sum(self, terms):
        """sum_k c_k <P_k> for (coefficient, Pauli) terms (mod.rs:93-101)."""
        k = len(terms)
        xs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.x) for _, p in terms])
        zs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.z) for _, p in terms])
        ph = (ctypes.c_uint32 * max(1, k))(*[p.phase & 3 for _, p in terms])
        cs = (ctypes.c_double * max(2, 2 * k))(*[v for c, _ in terms for v in (complex(c).real, complex(c).imag)])
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_expectation_sum(self._h, k, xs, zs, ph, cs, out))
        return complex(out[0], out[1])

    def inner(self, other):
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_inner(self._h, other._h, out))
        return complex(out[0], out[1])

    def compare(self, other):
        """(fidelity, max-abs difference, <self|other>, norms) against a
        state of any precision, computed on the device in FP64."""
        out = (ctypes.c_double * 5)()
        _check(L.lib().ncsv_compare(self._h, other._h, out))
        ip = complex(out[0], out[1])
        fid = abs(ip) ** 2 / (out[2] * out[3]) if out[2] > 0 and out[3] > 0 else 0.0
        return {"fidelity": fid, "max_abs": out[4], "inner": ip, "norm_self": out[2], "norm_other": out[3]}

    def set_basis(self, x):
        _check(L.lib().ncsv_set_basis(self._h, int(x)))

    def probability(self, x):
        out = ctypes.c_double()
        _check(L.lib().ncsv_probability(self._h, int(x), ctypes.byref(out)))
        return out.value

    # collapse (vector.rs:445-488; mod.rs:113-165, 220-233)
    def project(self, p, outcome):
        out = (ctypes.c_double * 2)()
        _check(L.lib().ncsv_project(self._h, _m64(p.x), _m64(p.z), p.phase & 3, 1 if outcome else 0, out))
        return complex(out[0], out[1])

    def measure(self, p, rng):
        r = _Rng(rng)
        o = ctypes.c_int()
        st = L.lib().ncsv_measure(self._h, _m64(p.x), _m64(p.z), p.phase & 3, r.fn, None, ctypes.byref(o))
        r.raise_if()
        _check(st)
        return bool(o.value)

    def reset(self, q, rng):
        r = _Rng(rng)
        st = L.lib().ncsv_reset(self._h, int(q), r.fn, None)
        r.raise_if()
        _check(st)

    def scale(self, c):
        c = complex(c)
        _check(L.lib().ncsv_scale(self._h, (ctypes.c_double * 2)(c.real, c.imag)))

    def normalize(self):
        _check(L.lib().ncsv_normalize(self._h))

    # sampling, tensor, readout
    def sample(self, rng, shots):
        """vector.rs:494-540: `shots` draws, then one pass (basis indices)."""
        r = _Rng(rng)
        out = np.zeros(max(1, shots), dtype=np.uint64)
        st = L.lib().ncsv_sample_rng(self._h, int(shots), r.fn, None, out.ctypes.data_as(L.c_u64p))
        r.raise_if()
        _check(st)
        return out[:shots]

    def sample_with_uniforms(self, us):
        us = np.ascontiguousarray(us, dtype=np.float64)
        out = np.zeros(max(1, len(us)), dtype=np.uint64)
        _check(L.lib().ncsv_sample(self._h, len(us), us.ctypes.data_as(L.c_dp), out.ctypes.data_as(L.c_u64p)))
        return out[:len(us)]

    def tensor(self, other):
        h = ctypes.c_void_p()
        _check(L.lib().ncsv_tensor(self._h, other._h, ctypes.byref(h)))
        return HipStateVector(h)

    kron = tensor

    @staticmethod
    def tensor_all(states):
        """mod.rs:81-90: a balanced tree of tensor products."""
        if not states:
            raise NcsvError(8, "tensor_all of no states")
        if len(states) == 1:
            return states[0].clone()
        k = len(states) // 2
        return HipStateVector.tensor_all(states[:k]).tensor(HipStateVector.tensor_all(states[k:]))

    def to_vector(self):
        n = self.num_qubits()
        out = np.zeros(1 << n, dtype=np.complex128)
        _check(L.lib().ncsv_to_vector(self._h, out.view(np.float64).ctypes.data_as(L.c_dp), 1 << n))
        return out

    # circuits (mod.rs:134-183)
    def run(self, circuit, rng=None):
        """Runs every op of `circuit`; returns the measurement record."""
        r = _Rng(rng)
        m = circuit.num_measurements()
        rec = (ctypes.c_uint8 * max(1, m))()
        nrec = ctypes.c_uint64()
        st = L.lib().ncsv_run(self._h, circuit._h, r.fn, None, rec, m, ctypes.byref(nrec))
        r.raise_if()
        _check(st)
        return [bool(rec[i]) for i in range(nrec.value)]

    # execution options and counters
    def set_fusion(self, mode):
        _check(L.lib().ncsv_set_fusion(self._h, int(mode)))

    def profile(self, on=True):
        _check(L.lib().ncsv_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        _check(L.lib().ncsv_profile_reset(self._h))

    def profile_read(self):
        arr = (L.KernelStat * 64)()
        cnt = ctypes.c_uint32()
        _check(L.lib().ncsv_profile_read(self._h, arr, 64, ctypes.byref(cnt)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms, arr[i].bytes) for i in range(min(64, cnt.value))}

    def sweep_count(self):
        v = ctypes.c_uint64()
        _check(L.lib().ncsv_sweep_count(self._h, ctypes.byref(v)))
        return v.value

    def synchronize(self):
        _check(L.lib().ncsv_synchronize(self._h))


class HipBatch:
    """Many circuits or noise trajectories run by one persistent launch per
    chunk of device memory (include/ncsv.h, batched form).  Each entry is a
    circuit run from |0...0> plus Paulis evaluated at its end; every Measure,
    Reset and Noise op takes one uniform per run, in program order."""

    def __init__(self, precision=C128):
        self.precision = precision
        self._h = L.lib().ncsv_batch_new(int(precision))
        if not self._h:
            _check(8)
        self._circuits = []  # kept alive with the batch
        self._paulis = []
        self._npauli = []
        self._ndraws = []

    def __del__(self):
        if getattr(self, "_h", None) and L._lib is not None:
            L._lib.ncsv_batch_destroy(self._h)
            self._h = None

    def __len__(self):
        return L.lib().ncsv_batch_size(self._h)

    def add(self, circuit, paulis=()):
        k = len(paulis)
        xs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.x) for p in paulis])
        zs = (ctypes.c_uint64 * max(1, k))(*[_m64(p.z) for p in paulis])
        ph = (ctypes.c_uint32 * max(1, k))(*[p.phase & 3 for p in paulis])
        idx = ctypes.c_uint32()
        _check(L.lib().ncsv_batch_add(self._h, circuit._h, k, xs, zs, ph, ctypes.byref(idx)))
        d = ctypes.c_uint32()
        _check(L.lib().ncsv_batch_draws(self._h, idx.value, ctypes.byref(d)))
        self._circuits.append(circuit)
        self._paulis.append(list(paulis))
        self._npauli.append(k)
        self._ndraws.append(d.value)
        return idx.value

    def draws(self, i):
        return self._ndraws[i]

    def set_tile_bits(self, bits):
        _check(L.lib().ncsv_batch_set_tile_bits(self._h, int(bits)))

    def set_blocks(self, blocks):
        _check(L.lib().ncsv_batch_set_blocks(self._h, int(blocks)))

    def set_shape(self, low_bits=-1, tiles_per_item=-1):
        """Fixed low tile qubits and the most tiles per queue item (-1: defaults)."""
        _check(L.lib().ncsv_batch_set_shape(self._h, int(low_bits), int(tiles_per_item)))

    def autotune(self, **kw):
        """Chooses tile width, low qubits and queue item size by timing
        candidates on a sample of this batch (ncsim_amd.autotune)."""
        from .autotune import tune_batch
        return tune_batch(self, **kw)

    def _args(self, uniforms, seeds, forced):
        import numpy as np
        u = s = f = None
        if uniforms is not None:
            u = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in uniforms])
                                     if isinstance(uniforms, (list, tuple)) else uniforms, dtype=np.float64)
            if u.size != sum(self._ndraws):
                raise NcsvError(8, f"{u.size} uniforms for {sum(self._ndraws)} draws")
        if seeds is not None:
            s = np.ascontiguousarray(seeds, dtype=np.uint64)
            if s.size != len(self):
                raise NcsvError(8, "one seed per entry")
        if forced is not None:
            f = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.uint8).ravel() for x in forced])
                                     if isinstance(forced, (list, tuple)) else forced, dtype=np.uint8)
            if f.size != sum(self._ndraws):
                raise NcsvError(8, "forced outcomes are laid out like the uniforms")
        keep = (u, s, f)
        cu = u.ctypes.data_as(L.c_dp) if u is not None else None
        cs = s.ctypes.data_as(L.c_u64p) if s is not None else None
        cf = f.ctypes.data_as(L.c_u8p) if f is not None else None
        return keep, cu, cs, cf

    def plan(self):
        """Plans every entry not planned yet (no device needed)."""
        _check(L.lib().ncsv_batch_plan(self._h))

    def run(self, uniforms=None, seeds=None, forced=None):
        keep, cu, cs, cf = self._args(uniforms, seeds, forced)
        _check(L.lib().ncsv_batch_run(self._h, cu, cs, cf))
        del keep

    def emulate(self, uniforms=None, seeds=None, forced=None):
        """The same programs on FP64 host amplitudes (diagnostics, no device)."""
        keep, cu, cs, cf = self._args(uniforms, seeds, forced)
        _check(L.lib().ncsv_batch_emulate(self._h, cu, cs, cf))
        del keep

    def simulate(self, uniforms=None, seeds=None, forced=None):
        """The device kernel's own code on the host, one workgroup simulated
        by host threads (diagnostics, no device, small batches)."""
        keep, cu, cs, cf = self._args(uniforms, seeds, forced)
        _check(L.lib().ncsv_batch_simulate(self._h, cu, cs, cf))
        del keep

    def result(self, i):
        """(expectations, norm_sq, outcomes per draw (None for noise draws))."""
        k = self._npauli[i]
        out = (ctypes.c_double * max(2, 2 * k))()
        nn = ctypes.c_double()
        oc = (ctypes.c_uint8 * max(1, self._ndraws[i]))()
        _check(L.lib().ncsv_batch_result(self._h, i, out, ctypes.byref(nn), oc))
        outc = [None if oc[d] == 0xff else int(oc[d]) for d in range(self._ndraws[i])]
        return [complex(out[2 * j], out[2 * j + 1]) for j in range(k)], nn.value, outc

    def stats(self):
        out = (ctypes.c_double * 9)()
        _check(L.lib().ncsv_batch_stats(self._h, out))
        keys = ("kernel_ms", "plan_ms", "items", "launches", "amp_updates", "hbm_bytes", "programs", "steps",
                "graph_replays")
        return dict(zip(keys, list(out)))


class HipFrameSampler:
    """FrameSampler (src/frame.rs): Pauli-frame sampling of a noisy Clifford
    circuit, 64 shots per word, on the GPU (include/ncsv.h, frame sampler).
    `seed` arguments replace the reference's `rng` (one u64 per call)."""

    def __init__(self, circuit, reference_seed=0):
        h = ctypes.c_void_p()
        _check(L.lib().ncsv_frames_new(circuit._h, int(reference_seed) & (2 ** 64 - 1), ctypes.byref(h)))
        self._h = h
        m, d, o = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(L.lib().ncsv_frames_shape(self._h, ctypes.byref(m), ctypes.byref(d), ctypes.byref(o)))
        self.num_measurements, self.num_detectors, self.num_observables = m.value, d.value, o.value

    def __del__(self):
        if getattr(self, "_h", None) and L._lib is not None:
            L._lib.ncsv_frames_destroy(self._h)
            self._h = None

    def reference(self):
        rec = (ctypes.c_uint8 * max(1, self.num_measurements))()
        _check(L.lib().ncsv_frames_reference(self._h, rec, self.num_measurements))
        return [int(rec[k]) for k in range(self.num_measurements)]

    def sample_packed(self, shots, seed):
        """Records packed by measurement: a list of uint64 arrays of
        ceil(shots / 64) words (bit s of word w = shot 64 w + s)."""
        import numpy as np
        W = (shots + 63) // 64
        out = np.zeros(max(1, self.num_measurements * W), dtype=np.uint64)
        _check(L.lib().ncsv_frames_sample_packed(self._h, shots, int(seed) & (2 ** 64 - 1), out.ctypes.data_as(L.c_u64p)))
        return [out[k * W:(k + 1) * W] for k in range(self.num_measurements)]

    def emulate_packed(self, shots, seed):
        """sample_packed run on the host by the kernel's per-word code."""
        import numpy as np
        W = (shots + 63) // 64
        out = np.zeros(max(1, self.num_measurements * W), dtype=np.uint64)
        _check(L.lib().ncsv_frames_emulate_packed(self._h, shots, int(seed) & (2 ** 64 - 1),
                                                  out.ctypes.data_as(L.c_u64p)))
        return [out[k * W:(k + 1) * W] for k in range(self.num_measurements)]

    def sample(self, shots, seed):
        """One record per shot (lists of bools), frame.rs:94-107."""
        packed = self.sample_packed(shots, seed)
        return [[bool((int(packed[k][s // 64]) >> (s % 64)) & 1) for k in range(len(packed))] for s in range(shots)]

    def sample_detectors(self, shots, seed):
        """(detectors, observables), each a list of packed uint64 arrays."""
        import numpy as np
        W = (shots + 63) // 64
        det = np.zeros(max(1, self.num_detectors * W), dtype=np.uint64)
        obs = np.zeros(max(1, self.num_observables * W), dtype=np.uint64)
        _check(L.lib().ncsv_frames_sample_detectors(self._h, shots, int(seed) & (2 ** 64 - 1),
                                                    det.ctypes.data_as(L.c_u64p), obs.ctypes.data_as(L.c_u64p)))
        return ([det[k * W:(k + 1) * W] for k in range(self.num_detectors)],
                [obs[k * W:(k + 1) * W] for k in range(self.num_observables)])
