"""ORACLE (test infrastructure only): ncsim's Pauli operators, restated.

P = i^phase X^x Z^z with packed masks (bit q = qubit q), Y = i X Z
(src/pauli/ops.rs:5-45, src/pauli/mod.rs:79-414).  Masks are Python ints, so
any register size works here; the dense kernels read only the low 64 bits as
the reference does (vector.rs:352-355).
"""
from .scalar import i_pow

KINDS = {"I": (0, 0), "_": (0, 0), "X": (1, 0), "Y": (1, 1), "Z": (0, 1)}


class SimError(Exception):
    """SimError (src/error.rs:7-42): `kind` is the variant name."""

    def __init__(self, kind, msg=""):
        super().__init__(f"{kind}: {msg}")
        self.kind = kind


class Pauli:
    __slots__ = ("n", "x", "z", "phase")

    def __init__(self, n, x=0, z=0, phase=0):
        mask = (1 << n) - 1
        self.n, self.x, self.z, self.phase = n, x & mask, z & mask, phase & 3

    @staticmethod
    def identity(n):
        return Pauli(n)

    def kind(self, q):
        xb, zb = (self.x >> q) & 1, (self.z >> q) & 1
        return {(0, 0): "I", (1, 0): "X", (1, 1): "Y", (0, 1): "Z"}[(xb, zb)]

    def set_kind(self, q, kind):
        """pauli/mod.rs:192-205: keeps the letter phase."""
        assert q < self.n
        if self.kind(q) == "Y":
            self.phase = (self.phase + 3) & 3
        xb, zb = KINDS[kind]
        self.x = (self.x & ~(1 << q)) | (xb << q)
        self.z = (self.z & ~(1 << q)) | (zb << q)
        if kind == "Y":
            self.phase = (self.phase + 1) & 3

    @staticmethod
    def from_kinds(kinds, k=0):
        p = Pauli(len(kinds))
        for q, c in enumerate(kinds):
            p.set_kind(q, c)
        p.phase = (p.phase + k) & 3
        return p

    @staticmethod
    def single(n, q, kind):
        p = Pauli(n)
        p.set_kind(q, kind)
        return p

    @staticmethod
    def z_(n, q):
        return Pauli.single(n, q, "Z")

    def is_hermitian(self):
        """ops.rs:22-27."""
        return (self.phase & 1) == (bin(self.x & self.z).count("1") & 1)

    def weight(self):
        return bin(self.x | self.z).count("1")

    def support(self):
        v = self.x | self.z
        return [q for q in range(self.n) if (v >> q) & 1]

    def negated(self):
        return Pauli(self.n, self.x, self.z, self.phase + 2)

    def display_phase(self):
        return (self.phase + 4 - (bin(self.x & self.z).count("1") % 4)) & 3

    def restrict(self, qubits):
        p = Pauli(len(qubits), 0, 0, self.phase)
        for i, q in enumerate(qubits):
            p.x |= ((self.x >> q) & 1) << i
            p.z |= ((self.z >> q) & 1) << i
        return p

    def embed(self, n, qubits):
        assert self.n == len(qubits)
        p = Pauli(n, 0, 0, self.phase)
        for i, q in enumerate(qubits):
            p.x |= ((self.x >> i) & 1) << q
            p.z |= ((self.z >> i) & 1) << q
        return p

    def apply_to_basis(self, v):
        """pauli/mod.rs:304-310: P|v> = i^k |v ^ x>."""
        k = (self.phase + 2 * bin(self.z & v).count("1")) & 3
        return k, v ^ self.x

    def apply_dense(self, amp):
        """(P psi) by the definition, for checks (clifford/mod.rs:364-379)."""
        import numpy as np

        out = np.zeros_like(amp)
        for v in range(len(amp)):
            k, w = self.apply_to_basis(v)
            out[w] = i_pow(k) * amp[v]
        return out

    def __eq__(self, o):
        return (self.n, self.x, self.z, self.phase) == (o.n, o.x, o.z, o.phase)

    def __repr__(self):
        pre = ["+", "+i", "-", "-i"][self.display_phase()]
        return pre + "".join(self.kind(q) if self.kind(q) != "I" else "_" for q in range(self.n))

    def to_sparse_string(self):
        pre = ["", "i*", "-", "-i*"][self.display_phase()]
        body = "*".join(f"{self.kind(q)}{q}" for q in self.support())
        return pre + (body if body else "1")

    @staticmethod
    def parse(s, n=None):
        """Pauli::parse / from_string (pauli/mod.rs:323-414): dense `+XZ_Y` or
        sparse `-i*X0*Z1`."""
        t = s.strip()
        phase, rest = 0, t
        for pre, ph in (("-i", 3), ("+i", 1), ("-", 2), ("+", 0), ("i", 1)):
            if rest.startswith(pre):
                r = rest[len(pre):].lstrip()
                if r.startswith("*"):
                    r = r[1:]
                phase, rest = ph, r.lstrip()
                break
        if rest and all(c in "IXYZ_" for c in rest):
            p = Pauli.from_kinds(list(rest), phase)
            if n is not None and n != len(rest):
                raise SimError("Parse", f"Pauli '{s}': {len(rest)} letters for {n} qubits")
            return p
        sites = []
        for tok in [t for t in rest.replace("*", " ").split() if t]:
            if tok == "1":
                continue
            if tok[0] not in "IXYZ_":
                raise SimError("Parse", f"Pauli '{s}': token '{tok}'")
            try:
                idx = int(tok[1:])
            except ValueError:
                raise SimError("Parse", f"Pauli '{s}': token '{tok}'")
            if any(q == idx for q, _ in sites):
                raise SimError("Parse", f"Pauli '{s}': qubit {idx} repeated")
            sites.append((idx, tok[0]))
        if n is None:
            n = max([q + 1 for q, _ in sites], default=0)
        p = Pauli(n)
        for q, c in sites:
            if q >= n:
                raise SimError("Parse", f"Pauli '{s}': qubit {q} out of range for {n} qubits")
            p.set_kind(q, c)
        p.phase = (p.phase + phase) & 3
        return p
