"""ORACLE (test infrastructure only): ncsim's gate views, restated.

`Gate` names follow src/gate.rs:355-431 and the text format's names
(gate.rs:479-515, circuit.rs:575-612).  The views the dense vector uses:
`monomial()` (gate.rs:519-564) via `MonomialGate::apply_to_basis` (157-169)
and `PhasePoly::eval` (phasepoly.rs:670-678); `matrix()` (gate.rs:670-747).
Phase bits are allowed up to 62 (the reference's PhasePoly stops at 32,
phasepoly.rs:550; C64::root_of_unity itself allows 62, scalar.rs:113).
"""
import math

import numpy as np

from .pauli import Pauli, SimError
from .scalar import Phasor, cmul, sqrt2_pow

ONE_Q = {"I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG",
         "T", "T_DAG", "PHASE", "RX", "RY", "RZ"}
TWO_Q = {"CX", "CY", "CZ", "SWAP", "CS", "CS_DAG", "CPHASE"}
THREE_Q = {"CCZ", "CCX"}


class PhasePoly:
    """f(x) = c + sum_T c_T prod_{l in T} [l.x] mod 2^bits (phasepoly.rs:536-678)."""

    def __init__(self, n, bits):
        assert 1 <= bits <= 62
        self.n, self.bits, self.constant, self.terms = n, bits, 0, {}

    def add_parity_term(self, c, forms):
        m = 1 << self.bits
        c %= m
        if c == 0:
            return
        key = []
        for f in forms:
            g = f & ((1 << self.n) - 1)
            if g == 0:
                return
            key.append(g)
        key = tuple(sorted(set(key)))
        if not key:
            self.constant = (self.constant + c) % m
            return
        v = (self.terms.get(key, 0) + c) % m
        if v == 0:
            self.terms.pop(key, None)
        else:
            self.terms[key] = v

    def add_monomial(self, c, vars_):
        self.add_parity_term(c, [1 << v for v in vars_])

    def eval(self, x):
        acc = self.constant
        for key, c in self.terms.items():
            if all(bin(l & x).count("1") & 1 for l in key):
                acc = (acc + c) % (1 << self.bits)
        return acc


class Monomial:
    """U|x> = scalar zeta^{f(x)} |M x + shift> (gate.rs:24-37)."""

    def __init__(self, k, rows=None, shift=0, phase=None, scalar=None):
        self.k = k
        self.rows = rows if rows is not None else [1 << i for i in range(k)]
        self.shift = shift
        self.phase = phase if phase is not None else PhasePoly(k, 1)
        self.scalar = scalar if scalar is not None else Phasor.one()

    def apply_to_basis(self, x):
        f = self.phase.eval(x)
        s = self.scalar.mul(Phasor.root(f, self.phase.bits))
        y = self.shift
        for i, r in enumerate(self.rows):
            if bin(r & x).count("1") & 1:
                y ^= 1 << i
        return s, y

    def images(self):
        """(C64 phase, image) of every local basis state (vector.rs:108-112)."""
        out = []
        for x in range(1 << self.k):
            s, y = self.apply_to_basis(x)
            out.append((s.to_c64(), y))
        return out

    def matrix(self):
        d = 1 << self.k
        m = np.zeros((d, d), dtype=complex)
        for x, (c, y) in enumerate(self.images()):
            m[y, x] = c
        return m


def mono_phase(k, bits, terms):
    """gate.rs:437-443."""
    p = PhasePoly(k, bits)
    for c, vs in terms:
        p.add_monomial(c, vs)
    return Monomial(k, phase=p)


def pauli_monomial(p):
    """MonomialGate::pauli (gate.rs:99-108)."""
    ph = PhasePoly(p.n, 1)
    ph.add_parity_term(1, [p.z])
    return Monomial(p.n, shift=p.x, phase=ph, scalar=Phasor.ipow(p.phase))


class Gate:
    """A gate as data.  `name` is the text-format name; PHASE/CPHASE carry
    (k, bits); RX/RY/RZ carry theta; UNITARY carries a dense matrix; DIAG a
    PhasePoly; PAULI a Pauli."""

    def __init__(self, name, k=None, bits=None, theta=None, matrix=None, poly=None, pauli=None, mono=None):
        self.name = name.upper()
        self.k, self.bits, self.theta = k, bits, theta
        self.mat, self.poly, self.pauli, self.mono = matrix, poly, pauli, mono

    def arity(self):
        if self.name in ONE_Q:
            return 1
        if self.name in TWO_Q:
            return 2
        if self.name in THREE_Q:
            return 3
        if self.name == "UNITARY":
            return int(round(math.log2(self.mat.shape[0])))
        if self.name == "DIAG":
            return self.poly.n
        if self.name == "PAULI":
            return self.pauli.n
        if self.name == "MONOMIAL":
            return self.mono.k
        raise SimError("Invalid", f"unknown gate {self.name}")

    def __repr__(self):
        if self.name in ("PHASE", "CPHASE"):
            return f"{self.name}({self.k},{self.bits})"
        if self.name in ("RX", "RY", "RZ"):
            return f"{self.name}({self.theta!r})"
        return self.name

    def monomial(self):
        """gate.rs:519-564."""
        n = self.name
        if n == "I":
            return Monomial(1)
        if n == "X":
            return Monomial(1, shift=1)
        if n == "Y":
            return pauli_monomial(Pauli.single(1, 0, "Y"))
        if n == "Z":
            return mono_phase(1, 1, [(1, [0])])
        if n == "S":
            return mono_phase(1, 2, [(1, [0])])
        if n == "S_DAG":
            return mono_phase(1, 2, [(3, [0])])
        if n == "T":
            return mono_phase(1, 3, [(1, [0])])
        if n == "T_DAG":
            return mono_phase(1, 3, [(7, [0])])
        if n == "PHASE":
            return mono_phase(1, self.bits, [(self.k, [0])])
        if n == "CPHASE":
            return mono_phase(2, self.bits, [(self.k, [0, 1])])
        if n == "CZ":
            return mono_phase(2, 1, [(1, [0, 1])])
        if n == "CS":
            return mono_phase(2, 2, [(1, [0, 1])])
        if n == "CS_DAG":
            return mono_phase(2, 2, [(3, [0, 1])])
        if n == "CCZ":
            return mono_phase(3, 1, [(1, [0, 1, 2])])
        if n == "CX":
            return Monomial(2, rows=[0b01, 0b11])
        if n == "CY":
            p = PhasePoly(2, 2)
            p.add_monomial(1, [0])
            p.add_monomial(2, [0, 1])
            return Monomial(2, rows=[0b01, 0b11], phase=p)
        if n == "SWAP":
            return Monomial(2, rows=[0b10, 0b01])
        if n == "DIAG":
            return Monomial(self.poly.n, phase=self.poly)
        if n == "MONOMIAL":
            return self.mono
        if n == "PAULI":
            return pauli_monomial(self.pauli)
        return None

    def matrix(self):
        """gate.rs:670-747 (C64): row-major, little-endian local basis."""
        n = self.name
        if n == "H":
            r = sqrt2_pow(-1)
            return np.array([[r, r], [r, -r]], dtype=complex)
        half = complex(0.5, 0.0)
        one, i = complex(1, 0), complex(0, 1)
        if n in ("SQRT_X", "SQRT_X_DAG"):
            a, b = cmul(one + i, half), cmul(one - i, half)
            if n == "SQRT_X_DAG":
                a, b = b, a
            return np.array([[a, b], [b, a]], dtype=complex)
        if n == "SQRT_Y":
            a = cmul(one + i, half)
            return np.array([[a, -a], [a, a]], dtype=complex)
        if n == "SQRT_Y_DAG":
            a = cmul(one - i, half)
            return np.array([[a, a], [-a, a]], dtype=complex)
        if n == "CCX":
            m = np.zeros((8, 8), dtype=complex)
            for x in range(8):
                y = x ^ 4 if x & 3 == 3 else x
                m[y, x] = 1.0
            return m
        if n in ("RX", "RY", "RZ"):
            t = self.theta
            c, s = math.cos(t / 2.0), math.sin(t / 2.0)
            if n == "RX":
                return np.array([[complex(c, 0), complex(0, -s)], [complex(0, -s), complex(c, 0)]])
            if n == "RY":
                return np.array([[complex(c, 0), complex(-s, 0)], [complex(s, 0), complex(c, 0)]])
            return np.array([[complex(math.cos(-t / 2.0), math.sin(-t / 2.0)), 0],
                             [0, complex(math.cos(t / 2.0), math.sin(t / 2.0))]], dtype=complex)
        if n == "UNITARY":
            return np.array(self.mat, dtype=complex)
        mo = self.monomial()
        return mo.matrix() if mo is not None else None

    def dagger(self):
        """gate.rs:774-813 for the named gates."""
        n = self.name
        sw = {"S": "S_DAG", "S_DAG": "S", "SQRT_X": "SQRT_X_DAG", "SQRT_X_DAG": "SQRT_X",
              "SQRT_Y": "SQRT_Y_DAG", "SQRT_Y_DAG": "SQRT_Y", "T": "T_DAG", "T_DAG": "T",
              "CS": "CS_DAG", "CS_DAG": "CS"}
        if n in sw:
            return Gate(sw[n])
        if n in ("PHASE", "CPHASE"):
            return Gate(n, k=-self.k, bits=self.bits)
        if n in ("RX", "RY", "RZ"):
            return Gate(n, theta=-self.theta)
        if n == "UNITARY":
            return Gate(n, matrix=np.conj(self.mat).T.copy())
        return Gate(n)

    def is_diagonal(self):
        if self.name == "RZ":
            return True
        mo = self.monomial()
        if mo is None:
            return False
        return all(y == x for x, (_, y) in enumerate(mo.images()))


def named(name, *args):
    """Gate from a text-format name and its arguments."""
    name = name.upper()
    if name in ("PHASE", "CPHASE"):
        return Gate(name, k=int(args[0]), bits=int(args[1]))
    if name in ("RX", "RY", "RZ"):
        return Gate(name, theta=float(args[0]))
    alias = {"CNOT": "CX", "TOFFOLI": "CCX"}
    return Gate(alias.get(name, name))
