"""ORACLE (test infrastructure only): ncsim's `StateVector<C64>` with the
`QuantumState` provided methods, over the C kernels of sv_oracle.c.

Follows src/state/vector.rs:14-548 (kernels via liboracle.so) and
src/state/mod.rs:74-233 (tensor_all, expectation_sum, probability, measure,
apply_op, run, normalize).  Randomness comes from an object with
`gen_f64()`, drawn at exactly the points the reference draws from its rng.
"""
import ctypes
import math
import os

import numpy as np

from .circuit import noise_sample
from .pauli import Pauli, SimError

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(path)
        dp, ip, u64 = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_uint64
        L.orc_apply_matrix.argtypes = [dp, ctypes.c_int, dp, ip, ctypes.c_int]
        L.orc_apply_monomial.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int, ip, dp]
        L.orc_apply_pauli.argtypes = [dp, ctypes.c_int, u64, u64, ctypes.c_int]
        L.orc_norm_sq.argtypes = [dp, ctypes.c_int, dp]
        L.orc_expectation.argtypes = [dp, ctypes.c_int, u64, u64, ctypes.c_int, dp]
        L.orc_inner.argtypes = [dp, dp, ctypes.c_int, dp]
        L.orc_project.argtypes = [dp, ctypes.c_int, u64, u64, ctypes.c_int, ctypes.c_int, dp]
        L.orc_scale.argtypes = [dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_kron.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int, dp]
        L.orc_sample.argtypes = [dp, ctypes.c_int, dp, ctypes.c_int64, ctypes.POINTER(u64)]
        L.orc_sample.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        L.orc_set_num_threads.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class SplitMix64:
    """A seeded uniform source: gen_f64() = (next_u64 >> 11) * 2^-53, the
    [0, 1) construction of rand's `Standard` f64 distribution."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF
        self.draws = 0

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def gen_f64(self):
        self.draws += 1
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


class StateVector:
    """2^n amplitudes, little-endian (vector.rs:14-19)."""

    BACKEND = "StateVector"

    def __init__(self, n, amp):
        self.n = n
        self.amp = np.ascontiguousarray(amp, dtype=np.complex128)

    # constructors (vector.rs:24-52)
    @staticmethod
    def zero_state(n):
        return StateVector.basis_state(n, 0)

    @staticmethod
    def basis_state(n, x):
        assert n <= 30, "state vector of more than 30 qubits"
        a = np.zeros(1 << n, dtype=np.complex128)
        a[x] = 1.0
        return StateVector(n, a)

    @staticmethod
    def plus_state(n):
        from .scalar import sqrt2_pow

        return StateVector(n, np.full(1 << n, sqrt2_pow(-n), dtype=np.complex128))

    @staticmethod
    def from_amplitudes(n, amp):
        if len(amp) != 1 << n:
            raise SimError("Invalid", f"{len(amp)} amplitudes for {n} qubits")
        return StateVector(n, np.array(amp, dtype=np.complex128))

    def clone(self):
        return StateVector(self.n, self.amp.copy())

    def num_qubits(self):
        return self.n

    # kernels
    def apply_matrix(self, m, qubits):
        """vector.rs:67-93."""
        from .circuit import check_qubits

        check_qubits(qubits, self.n)
        k = len(qubits)
        m = np.ascontiguousarray(m, dtype=np.complex128)
        assert m.shape == (1 << k, 1 << k), "matrix size"
        qs = np.array(qubits, dtype=np.int32)
        lib().orc_apply_matrix(_dp(self.amp), self.n, _dp(m), _ip(qs), k)

    def apply_monomial(self, mono, qubits):
        """vector.rs:96-139."""
        from .circuit import check_qubits

        check_qubits(qubits, self.n)
        if len(qubits) != mono.k:
            raise SimError("Arity", f"gate monomial acts on {mono.k} qubits, {len(qubits)} given")
        im = mono.images()
        y = np.array([t[1] for t in im], dtype=np.int32)
        c = np.array([t[0] for t in im], dtype=np.complex128)
        qs = np.array(qubits, dtype=np.int32)
        lib().orc_apply_monomial(_dp(self.amp), self.n, _ip(qs), len(qubits), _ip(y), _dp(c))

    def apply_pauli(self, p):
        """vector.rs:143-164."""
        assert p.n == self.n, "apply_pauli: register size"
        lib().orc_apply_pauli(_dp(self.amp), self.n, p.x & (2**64 - 1), p.z & (2**64 - 1), p.phase)

    def apply_gate(self, gate, qubits):
        """vector.rs:364-395: Pauli, then monomial, then matrix, then
        decomposition (none is needed for the named gates: every gate with a
        decomposition also has a C64 matrix)."""
        from .circuit import check_qubits

        check_qubits(qubits, self.n)
        if len(qubits) != gate.arity():
            raise SimError("Arity", f"gate {gate!r} acts on {gate.arity()} qubits, {len(qubits)} given")
        if gate.name == "PAULI":
            self.apply_pauli(gate.pauli.embed(self.n, qubits))
            return
        mo = gate.monomial()
        if mo is not None:
            self.apply_monomial(mo, qubits)
            return
        m = gate.matrix()
        if m is not None:
            self.apply_matrix(m, qubits)
            return
        raise SimError("NotRepresentable", f"gate {gate!r} in the scalar type")

    def norm_sq(self):
        out = np.zeros(2)
        lib().orc_norm_sq(_dp(self.amp), self.n, _dp(out))
        return complex(out[0], out[1])

    def amplitude(self, x):
        return complex(self.amp[x])

    def expectation(self, p):
        if p.n != self.n:
            raise SimError("Invalid", "expectation: operator on the wrong register")
        out = np.zeros(2)
        lib().orc_expectation(_dp(self.amp), self.n, p.x & (2**64 - 1), p.z & (2**64 - 1), p.phase, _dp(out))
        return complex(out[0], out[1])

    def inner(self, other):
        if other.n != self.n:
            raise SimError("Invalid", "inner: register sizes differ")
        out = np.zeros(2)
        lib().orc_inner(_dp(self.amp), _dp(other.amp), self.n, _dp(out))
        return complex(out[0], out[1])

    def project(self, p, outcome):
        if not p.is_hermitian():
            raise SimError("Invalid", f"projector of non-Hermitian {p!r}")
        out = np.zeros(2)
        lib().orc_project(_dp(self.amp), self.n, p.x & (2**64 - 1), p.z & (2**64 - 1), p.phase,
                          1 if outcome else 0, _dp(out))
        return complex(out[0], out[1])

    def scale(self, c):
        c = complex(c)
        lib().orc_scale(_dp(self.amp), self.n, c.real, c.imag)

    def kron(self, other):
        n = self.n + other.n
        if n > 30:
            raise SimError("Invalid", f"state vector of {n} qubits")
        out = np.empty(1 << n, dtype=np.complex128)
        lib().orc_kron(_dp(self.amp), self.n, _dp(other.amp), other.n, _dp(out))
        return StateVector(n, out)

    tensor = kron

    def sample(self, rng, shots):
        """vector.rs:494-540: the shots' uniforms are drawn first, in order."""
        us = np.array([rng.gen_f64() for _ in range(shots)], dtype=np.float64)
        return self.sample_with_uniforms(us)

    def sample_with_uniforms(self, us):
        us = np.ascontiguousarray(us, dtype=np.float64)
        out = np.zeros(len(us), dtype=np.uint64)
        rc = lib().orc_sample(_dp(self.amp), self.n, _dp(us), len(us),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc == 6:
            raise SimError("ZeroState", "sample")
        return out

    def to_vector(self):
        return self.amp.copy()

    # provided methods (state/mod.rs)
    @staticmethod
    def tensor_all(states):
        """mod.rs:81-90."""
        if not states:
            raise SimError("Invalid", "tensor_all of no states")
        if len(states) == 1:
            return states[0].clone()
        k = len(states) // 2
        return StateVector.tensor_all(states[:k]).tensor(StateVector.tensor_all(states[k:]))

    def expectation_sum(self, terms):
        """mod.rs:93-101 over (coefficient, Pauli) terms, summed pairwise
        (scalar.rs:80-94)."""
        from .scalar import cmul

        vals = [cmul(complex(c), self.expectation(p)) for c, p in terms]

        def rec(items):
            if len(items) <= 8:
                acc = complex(0.0, 0.0)
                for x in items:
                    acc = acc + x
                return acc
            h = len(items) // 2
            return rec(items[:h]) + rec(items[h:])

        return rec(vals)

    def probability(self, x):
        """mod.rs:104-108."""
        a = self.amplitude(x)
        return (a.real * a.real + a.imag * a.imag) / self.norm_sq().real

    def measure(self, p, rng):
        """mod.rs:113-131."""
        if not p.is_hermitian():
            raise SimError("Invalid", f"measured operator {p!r} is not Hermitian")
        e = self.expectation(p).real
        nn = self.norm_sq().real
        # IEEE division (0/0 = NaN, e/0 = +-inf) and f64::clamp, which keeps a
        # NaN: a zero state draws one uniform and takes outcome 0 (mod.rs:122-128)
        with np.errstate(divide="ignore", invalid="ignore"):
            p_plus = float((1.0 + np.float64(e) / np.float64(nn)) / 2.0)
        if not math.isnan(p_plus):
            p_plus = min(max(p_plus, 0.0), 1.0)
        if p_plus > 1.0 - 1e-12:
            outcome = False
        elif p_plus < 1e-12:
            outcome = True
        else:
            outcome = rng.gen_f64() >= p_plus
        self.project(p, outcome)
        return outcome

    def apply_op(self, op, rng):
        """mod.rs:134-165."""
        kind = op[0]
        if kind == "gate":
            self.apply_gate(op[1], op[2])
            return None
        if kind == "measure":
            return self.measure(op[1], rng)
        if kind == "reset":
            from .gates import Gate

            if self.measure(Pauli.single(self.n, op[1], "Z"), rng):
                self.apply_gate(Gate("X"), [op[1]])
            return None
        if kind == "noise":
            e = noise_sample(op[1][0], op[1][1], rng)
            if e is not None:
                from .gates import Gate

                x, z = e
                qs = op[2]
                local = Pauli(len(qs), x, z, 0)
                self.apply_gate(Gate("PAULI", pauli=local), qs)
            return None
        return None

    def run(self, circuit, rng):
        """mod.rs:168-183."""
        if circuit.n != self.n:
            raise SimError("Invalid", f"circuit on {circuit.n} qubits, state on {self.n}")
        rec = []
        for op in circuit.ops:
            b = self.apply_op(op, rng)
            if b is not None:
                rec.append(b)
        return rec

    def normalize(self):
        """mod.rs:220-233."""
        nn = self.norm_sq().real
        if nn <= 0.0:
            raise SimError("ZeroState", "normalize")
        from .scalar import sqrt2_pow

        l = -math.log2(nn)
        if abs(l - round(l)) < 1e-9:
            c = sqrt2_pow(int(round(l)))
        else:
            c = complex(1.0 / math.sqrt(nn), 0.0)
        self.scale(c)
