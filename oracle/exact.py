"""ORACLE (test infrastructure only): an exact dense simulator over
D[w] = Z[w, 1/sqrt2], w = e^{i pi/4}, for Clifford+T circuits.

It stands in for the reference's exact `PathSum<Cyclo<8>>` /
`StateVector<Cyclo<8>>` (src/state/pathsum.rs:51-62, 943-997;
src/state/vector.rs:580-633), which cannot run here.  A state is a common
exponent e and integer 4-vectors (a0..a3) per basis state meaning
(a0 + a1 w + a2 w^2 + a3 w^3) / sqrt2^e.  Gates are the named gates whose C64
matrices (src/gate.rs:670-747) and monomial views (519-564) have entries in
D[w]; rotations are refused (NotRepresentable), as in the reference.
"""
from decimal import Decimal, getcontext

from .pauli import Pauli, SimError

getcontext().prec = 60
_SQRT2 = Decimal(2).sqrt()


def zmul(a, b):
    """Product in Z[w] with w^4 = -1."""
    c = [0, 0, 0, 0]
    for i in range(4):
        if a[i] == 0:
            continue
        for j in range(4):
            if b[j] == 0:
                continue
            k = i + j
            if k < 4:
                c[k] += a[i] * b[j]
            else:
                c[k - 4] -= a[i] * b[j]
    return c


def wpow(k):
    """w^k as a Z[w] element."""
    k %= 8
    v = [0, 0, 0, 0]
    v[k % 4] = 1 if k < 4 else -1
    return v


# exact gate tables: (matrix of Z[w] entries, sqrt2 exponent of the scale)
def _dense_exact(name):
    one, zero = [1, 0, 0, 0], [0, 0, 0, 0]
    neg = lambda v: [-x for x in v]
    if name == "H":
        return [[one, one], [one, neg(one)]], 1
    a, b = [1, 0, 1, 0], [1, 0, -1, 0]  # 1 + i, 1 - i (over sqrt2^2)
    if name == "SQRT_X":
        return [[a, b], [b, a]], 2
    if name == "SQRT_X_DAG":
        return [[b, a], [a, b]], 2
    if name == "SQRT_Y":
        return [[a, neg(a)], [a, a]], 2
    if name == "SQRT_Y_DAG":
        return [[b, b], [neg(b), b]], 2
    if name == "CCX":
        m = [[zero] * 8 for _ in range(8)]
        for x in range(8):
            y = x ^ 4 if x & 3 == 3 else x
            m[y][x] = one
        return m, 0
    return None


class ExactState:
    def __init__(self, n):
        self.n = n
        self.e = 0
        self.a = [[0, 0, 0, 0] for _ in range(1 << n)]
        self.a[0] = [1, 0, 0, 0]

    def _reduce(self):
        # divide by 2 while every coefficient is even
        while self.e >= 2 and all(c % 2 == 0 for v in self.a for c in v):
            self.a = [[c // 2 for c in v] for v in self.a]
            self.e -= 2
        # divide by sqrt2 = w - w^3 when it stays integral
        if self.e >= 1:
            s = [0, 1, 0, -1]
            t = [zmul(v, s) for v in self.a]
            if all(c % 2 == 0 for v in t for c in v):
                self.a = [[c // 2 for c in v] for v in t]
                self.e -= 1

    def apply_gate(self, gate, qubits):
        k = len(qubits)
        if k != gate.arity():
            raise SimError("Arity", "exact: arity")
        d = _dense_exact(gate.name)
        if d is None:
            mo = gate.monomial()
            if mo is None:
                raise SimError("NotRepresentable", f"gate {gate!r} in Z[w]")
            m = [[[0, 0, 0, 0] for _ in range(1 << k)] for _ in range(1 << k)]
            for x in range(1 << k):
                # the image's phasor, normalized: a power of w iff bits <= 3
                s, y = mo.apply_to_basis(x)
                if s.bits > 3 or s.p != 0:
                    raise SimError("NotRepresentable", f"gate {gate!r} in Z[w]")
                m[y][x] = wpow(s.phase << (3 - s.bits))
            d = (m, 0)
        mat, eg = d
        gd = 1 << k
        out = [None] * (1 << self.n)
        mask = 0
        for q in qubits:
            mask |= 1 << q
        for base in range(1 << self.n):
            if base & mask:
                continue
            idx = []
            for loc in range(gd):
                i = base
                for j, q in enumerate(qubits):
                    if (loc >> j) & 1:
                        i |= 1 << q
                idx.append(i)
            v = [self.a[i] for i in idx]
            for r in range(gd):
                acc = [0, 0, 0, 0]
                for c in range(gd):
                    if any(mat[r][c]) and any(v[c]):
                        p = zmul(mat[r][c], v[c])
                        acc = [acc[t] + p[t] for t in range(4)]
                out[idx[r]] = acc
        self.a = out
        self.e += eg
        self._reduce()

    def run(self, circuit):
        for op in circuit.ops:
            if op[0] == "gate":
                self.apply_gate(op[1], op[2])
            elif op[0] in ("tick", "detector", "observable"):
                continue
            else:
                raise SimError("Unsupported", "exact: only unitary circuits")
        return self

    def amplitude_exact(self, x):
        """(a0, a1, a2, a3, e)."""
        return tuple(self.a[x]) + (self.e,)

    def to_complex(self, x):
        a0, a1, a2, a3 = (Decimal(c) for c in self.a[x])
        h = 1 / _SQRT2
        re = a0 + (a1 - a3) * h
        im = a2 + (a1 + a3) * h
        s = _SQRT2 ** self.e
        return complex(float(re / s), float(im / s))

    def to_vector(self):
        return [self.to_complex(x) for x in range(1 << self.n)]

    def expectation(self, p):
        """<psi|P|psi> as a complex float from exact arithmetic."""
        from .scalar import i_pow

        tot = complex(0, 0)
        vec = self.to_vector()
        for c in range(1 << self.n):
            k, w = p.apply_to_basis(c)
            tot += vec[w].conjugate() * i_pow(k) * vec[c]
        return tot

    def norm_sq_exact(self):
        """sum |a|^2 as (a0..a3, exponent) in Z[w] / 2^e (conj w = w^7)."""
        tot = [0, 0, 0, 0]
        for v in self.a:
            conj = [v[0], -v[3], -v[2], -v[1]]
            p = zmul(v, conj)
            tot = [tot[t] + p[t] for t in range(4)]
        return tot, self.e
