"""Pins the CPU oracle against the reference's own known answers and against
what the mathematics fixes (brute force, closed forms, invariants).

Each test names the reference test it restates.  Everything here runs on the
CPU.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import gates as G
from oracle.circuit import Circuit
from oracle.exact import ExactState
from oracle.pauli import Pauli, SimError
from oracle.scalar import Phasor, root_of_unity, sqrt2_pow
from oracle.statevector import SplitMix64, StateVector

getcontext().prec = 50


# ---------------------------------------------------------------- helpers
def naive_apply(amp, m, qubits):
    """The obvious O(4^k 2^n) application (vector.rs:664-683)."""
    k = len(qubits)
    gd = 1 << k
    out = np.zeros_like(amp)
    for i in range(len(amp)):
        a = sum(((i >> q) & 1) << j for j, q in enumerate(qubits))
        acc = 0j
        for b in range(gd):
            r = i
            for j, q in enumerate(qubits):
                r = (r & ~(1 << q)) | (((b >> j) & 1) << q)
            acc += m[a, b] * amp[r]
        out[i] = acc
    return out


def embed_matrix(m, qubits, n):
    """The full 2^n x 2^n operator of a k-qubit matrix on `qubits`."""
    d = 1 << n
    full = np.zeros((d, d), dtype=complex)
    for col in range(d):
        e = np.zeros(d, dtype=complex)
        e[col] = 1
        full[:, col] = naive_apply(e, m, qubits)
    return full


def circuit_unitary(c):
    """Product of embedded gate matrices (circuit.rs:435-456)."""
    d = 1 << c.n
    u = np.eye(d, dtype=complex)
    for op in c.ops:
        assert op[0] == "gate"
        u = embed_matrix(op[1].matrix(), op[2], c.n) @ u
    return u


def random_state(n, rng):
    return (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)


# ------------------------------------------------------------------ scalars
def test_root_of_unity_matches_high_precision_and_exact_eighths():
    """scalar.rs:443-468 (exact_and_float_agree): bits 0..4 against exact
    values; plus bits up to 33 against 50-digit cos/sin."""
    from mpmath import mp, mpf, cos, sin, pi

    mp.dps = 40
    for bits in range(0, 34):
        for k in list(range(min(1 << bits, 17))) + [(1 << bits) - 1]:
            f = root_of_unity(k, bits)
            t = 2 * pi * mpf(k) / mpf(2) ** bits
            assert abs(f.real - float(cos(t))) < 1e-15 and abs(f.imag - float(sin(t))) < 1e-15, (k, bits)
    # exact constants at multiples of pi/4 (scalar.rs:116-127)
    h = 0.7071067811865476
    assert root_of_unity(1, 3) == complex(h, h)
    assert root_of_unity(3, 3) == complex(-h, h)
    assert root_of_unity(2, 3) == complex(0, 1)
    assert root_of_unity(-1, 2) == complex(0, -1)
    for k in range(-5, 6):
        want = float(Decimal(2).sqrt() ** k)
        assert abs(sqrt2_pow(k).real - want) < 1e-14


def test_phasor_arithmetic():
    """scalar.rs:428-441."""
    z8 = Phasor.root(1, 3)
    assert z8.mul(z8).mul(z8).mul(z8) == Phasor(0, 1, 1)
    assert Phasor(0, 4, 3) == Phasor(0, 1, 1)
    assert Phasor(0, 2, 3) == Phasor.ipow(1)
    v = Phasor(-3, 5, 3).to_c64()
    e = 2 ** -1.5 * complex(math.cos(2 * math.pi * 5 / 8), math.sin(2 * math.pi * 5 / 8))
    assert abs(v - e) < 1e-14


# -------------------------------------------------------------------- gates
ALL_NAMED = ["I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG",
             "T", "T_DAG", "CX", "CY", "CZ", "SWAP", "CS", "CS_DAG", "CCZ", "CCX"]


def all_gates():
    gs = [G.named(nm) for nm in ALL_NAMED]
    gs += [G.Gate("PHASE", k=3, bits=4), G.Gate("CPHASE", k=1, bits=3), G.Gate("RX", theta=0.3),
           G.Gate("RY", theta=-1.1), G.Gate("RZ", theta=2.5), G.Gate("PHASE", k=5, bits=33)]
    return gs


@pytest.mark.parametrize("g", all_gates(), ids=repr)
def test_views_agree_with_matrices(g):
    """gate.rs:903-984: unitary, dagger, monomial composition."""
    m = g.matrix()
    d = m.shape[0]
    assert np.abs(m @ m.conj().T - np.eye(d)).max() < 1e-12
    md = g.dagger().matrix()
    assert np.abs(md - m.conj().T).max() < 1e-12
    mo = g.monomial()
    if mo is not None:
        # compose with itself through the images, as MonomialGate::compose does
        im = mo.images()
        m2 = np.zeros((d, d), dtype=complex)
        for x in range(d):
            c1, y1 = im[x]
            c2, y2 = im[y1]
            m2[y2, x] = c2 * c1
        assert np.abs(m2 - m @ m).max() < 1e-12


def test_named_gate_closed_forms():
    """The gate table against textbook matrices (gate.rs:355-431 doc comments:
    SqrtX = H S H, SqrtY = S SqrtX S^dag, T = diag(1, zeta_8), controls first,
    CCX = H2 CCZ H2 (gate.rs:766))."""
    h = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    s = np.diag([1, 1j])
    sx = G.named("SQRT_X").matrix()
    assert np.abs(sx - h @ s @ h).max() < 1e-15
    assert np.abs(G.named("SQRT_Y").matrix() - s @ sx @ s.conj().T).max() < 1e-15
    assert np.abs(G.named("T").matrix() - np.diag([1, np.exp(1j * np.pi / 4)])).max() < 1e-15
    assert np.abs(G.named("Y").matrix() - np.array([[0, -1j], [1j, 0]])).max() == 0
    # CX: control = local bit 0, target = local bit 1
    cx = G.named("CX").matrix()
    assert cx[0b11, 0b01] == 1 and cx[0b01, 0b11] == 1 and cx[0, 0] == 1 and cx[0b10, 0b10] == 1
    cy = G.named("CY").matrix()
    assert cy[0b11, 0b01] == 1j and cy[0b01, 0b11] == -1j
    ccx_dec = embed_matrix(h, [2], 3) @ G.named("CCZ").matrix() @ embed_matrix(h, [2], 3)
    assert np.abs(ccx_dec - G.named("CCX").matrix()).max() < 1e-15
    # SqrtY decomposition Sdg, H, S, H, S (gate.rs:764) in circuit order
    dec = s @ h @ s @ h @ s.conj().T
    assert np.abs(dec - G.named("SQRT_Y").matrix()).max() < 1e-15
    # PHASE(k, bits) beyond the reference's 32 bits: zeta_{2^33}
    ph = G.Gate("PHASE", k=1, bits=33).matrix()
    from mpmath import mp, cos, sin, pi

    mp.dps = 40
    assert abs(ph[1, 1] - complex(float(cos(2 * pi / 2 ** 33)), float(sin(2 * pi / 2 ** 33)))) < 1e-16
    rz = G.Gate("RZ", theta=0.7).matrix()
    assert abs(rz[0, 0] - np.exp(-0.35j)) < 1e-15 and abs(rz[1, 1] - np.exp(0.35j)) < 1e-15


# ----------------------------------------------------------- state kernels
def test_ghz_exact_and_float():
    """vector.rs:580-633."""
    c = Circuit(3)
    c.gate("H", [0]).gate("CX", [0, 1]).gate("CX", [1, 2])
    f = StateVector.zero_state(3)
    rng = SplitMix64(1)
    f.run(c, rng)
    e = ExactState(3).run(c)
    assert e.amplitude_exact(0) == (1, 0, 0, 0, 1) and e.amplitude_exact(7) == (1, 0, 0, 0, 1)
    assert e.amplitude_exact(1)[:4] == (0, 0, 0, 0)
    assert np.abs(np.array(e.to_vector()) - f.to_vector()).max() < 1e-14
    assert f.expectation(Pauli.parse("X0*X1*X2")) == 1.0 + 0j or abs(f.expectation(Pauli.parse("X0*X1*X2")) - 1) < 1e-15
    assert abs(f.expectation(Pauli.parse("Z0*Z1", 3)) - 1) < 1e-15
    assert f.expectation(Pauli.parse("Z0", 3)) == 0
    zeros = 0
    for _ in range(200):
        s = f.clone()
        b = s.measure(Pauli.parse("Z0", 3), rng)
        b1 = s.measure(Pauli.parse("Z1", 3), rng)
        assert b == b1
        zeros += not b
    assert 60 <= zeros <= 140
    assert all(int(x) in (0, 7) for x in f.sample(rng, 100))
    p = f.clone()
    w = p.project(Pauli.parse("Z0", 3), True)
    # the reference asserts exactly 0.5 on its exact state; in floats
    # (1/sqrt2)^2 rounds to 0.5 + 1 ulp
    assert abs(w.real - 0.5) < 2e-16 and w.imag == 0
    p.normalize()
    assert abs(p.amplitude(7) - 1.0) < 4e-16


def test_non_clifford_matches_matrix():
    """vector.rs:635-654."""
    c = Circuit(2)
    c.gate("H", [0]).gate("H", [1]).gate("T", [0]).gate("CS", [0, 1]).gate("SQRT_X", [1])
    c.gate(G.Gate("RZ", theta=0.7), [0])
    s = StateVector.zero_state(2)
    s.run(c, SplitMix64(0))
    u = circuit_unitary(c)
    assert np.abs(s.to_vector() - u[:, 0]).max() < 1e-12
    with pytest.raises(SimError):
        ExactState(2).run(c)


def test_fiber_kernel_matches_naive_on_high_and_low_qubits():
    """vector.rs:685-757, extended to k = 1..5."""
    rng = np.random.default_rng(8)
    n = 15
    v = StateVector(n, random_state(n, rng))
    for trial in range(15):
        k = 1 + trial % 5
        qs = []
        while len(qs) < k:
            q = int(rng.integers(12, n)) if rng.random() < 0.5 else int(rng.integers(0, 12))
            if q not in qs:
                qs.append(q)
        m = rng.random((1 << k, 1 << k)) + 1j * rng.random((1 << k, 1 << k))
        if trial % 4 == 3:
            m[rng.random(m.shape) < 0.5] = 0  # exercise the zero-skips
        want = naive_apply(v.amp, m, qs)
        v.apply_matrix(m, qs)
        assert np.abs(v.amp - want).max() < 1e-9, (trial, qs)
        v.scale(1.0 / math.sqrt(v.norm_sq().real))
    for trial in range(10):
        kinds = [["I", "X", "Y", "Z"][int(rng.integers(0, 4))] for _ in range(n)]
        p = Pauli.from_kinds(kinds)
        if trial % 3 == 0:
            p = p.negated()
        pv = p.apply_dense(v.amp)
        want = np.vdot(v.amp, pv)
        assert abs(v.expectation(p) - want) < 1e-9
        u = v.clone()
        u.apply_pauli(p)
        assert np.abs(u.amp - pv).max() < 1e-12
        w = v.clone()
        b = bool(rng.integers(0, 2))
        w.project(p, b)
        sgn = -1.0 if b else 1.0
        assert np.abs(w.amp - (v.amp + sgn * pv) * 0.5).max() < 1e-12


def test_pair_sum_is_the_reference_tree():
    """par_sum (vector.rs:334-350) and Amplitude::sum (scalar.rs:80-94, the
    1000-term check at 462-467): the oracle's norm equals a pure-Python
    pairwise tree bit for bit, whatever the thread count."""
    rng = np.random.default_rng(3)
    n = 16
    amp = random_state(n, rng)

    def tree(lo, hi):
        if hi - lo <= 8:
            acc = complex(0.0, 0.0)
            for i in range(lo, hi):
                a = amp[i]
                acc = acc + complex(a.real * a.real - a.imag * -a.imag, a.real * -a.imag + a.imag * a.real)
            return acc
        mid = lo + (hi - lo) // 2
        return tree(lo, mid) + tree(mid, hi)

    want = tree(0, 1 << n)
    got = StateVector(n, amp).norm_sq()
    assert got == want


def test_one_pass_sampling_follows_the_born_rule_across_blocks():
    """vector.rs:767-800."""
    n = 14
    i = np.arange(1 << n)
    amp = np.where(i % 4099 == 7, 3.0, 0.01 * (i % 13)).astype(complex)
    v = StateVector(n, amp)
    shots = 200_000
    s = v.sample_with_uniforms(np.random.default_rng(2).random(shots))
    total = np.sum(np.abs(amp) ** 2)
    counts = np.bincount(s.astype(np.int64), minlength=1 << n)
    for j in range(7, 1 << n, 4099):
        p = abs(amp[j]) ** 2 / total
        f = counts[j] / shots
        assert abs(f - p) < 6 * math.sqrt(p * (1 - p) / shots)
    assert np.all((s % 13 != 0) | (s % 4099 == 7))


def test_sampling_walk_is_the_inverse_cdf():
    """The one-pass walk returns, for each uniform u, the first index whose
    cumulative weight exceeds u*total (brute force on a small state), and a
    zero state is ZeroState."""
    rng = np.random.default_rng(5)
    n = 13
    amp = random_state(n, rng)
    amp[rng.random(1 << n) < 0.3] = 0
    us = rng.random(500)
    got = StateVector(n, amp).sample_with_uniforms(us)
    w = np.abs(amp) ** 2
    cdf = np.cumsum(w)
    for u, g in zip(us, got):
        t = u * cdf[-1]
        j = int(np.searchsorted(cdf, t, side="right"))
        # rounding of block sums can move a pick across a boundary by one
        # nonzero index at most
        assert w[int(g)] > 0
        assert abs(int(g) - j) <= 1 or abs(cdf[int(g)] - t) < 1e-9
    with pytest.raises(SimError):
        StateVector(4, np.zeros(16)).sample_with_uniforms([0.5])


# ---------------------------------------------------------- exact oracle
def test_exact_t_gates():
    """pathsum.rs:1256-1288: T on |+>^3 then CCZ gives zeta8^{|x| + 4[x=7]}/2^{3/2}."""
    c = Circuit(3)
    for q in range(3):
        c.gate("H", [q])
    for q in range(3):
        c.gate("T", [q])
    c.gate("CCZ", [0, 1, 2])
    e = ExactState(3).run(c)
    f = StateVector.zero_state(3)
    f.run(c, SplitMix64(1))
    for x in range(8):
        k = bin(x).count("1") + (4 if x == 7 else 0)
        want = np.exp(1j * np.pi * k / 4) / 2 ** 1.5
        assert abs(e.to_complex(x) - want) < 1e-15
        assert abs(f.amplitude(x) - want) < 1e-15
    assert e.norm_sq_exact() == ([1, 0, 0, 0], 0) or e.norm_sq_exact()[0] == [2 ** e.e, 0, 0, 0]


def test_diagonal_gate_expectations_are_exponential_sums():
    """tests/backends.rs:109-136: <0|H D_f H|0> = 2^-3 sum_x zeta8^f(x), f = x0
    + 3 x1 x2 + 4 x0 x1 x2 mod 8 — summed directly over x here."""
    f = G.PhasePoly(3, 3)
    f.add_monomial(1, [0])
    f.add_monomial(3, [1, 2])
    f.add_monomial(4, [0, 1, 2])
    want = sum(np.exp(2j * np.pi * (((x & 1) + 3 * ((x >> 1) & 1) * ((x >> 2) & 1) + 4 * (x == 7)) % 8) / 8)
               for x in range(8)) / 8
    c = Circuit(3)
    for q in range(3):
        c.gate("H", [q])
    c.gate(G.Gate("DIAG", poly=f), [0, 1, 2])
    for q in range(3):
        c.gate("H", [q])
    v = StateVector.zero_state(3)
    v.run(c, SplitMix64(0))
    assert abs(v.amplitude(0) - want) < 1e-15
    e = ExactState(3).run(c)
    assert abs(e.to_complex(0) - want) < 1e-15


CLIFFORD_T = """
QUBITS 4
H 0 1 2 3
T 0
CX 0 1
CS 1 2
T_DAG 3
CZ 2 3
H 1
CCZ 0 1 2
SQRT_X 3
S 0
CX 3 0
H 2
T 2
"""


def test_clifford_t_circuit_exact_and_float():
    """tests/backends.rs:18-75."""
    c = Circuit.parse(CLIFFORD_T)
    v = StateVector.zero_state(4)
    v.run(c, SplitMix64(0))
    e = ExactState(4).run(c)
    assert np.abs(np.array(e.to_vector()) - v.to_vector()).max() < 1e-14
    tot, ex = e.norm_sq_exact()
    assert tot[1:] == [0, 0, 0] and tot[0] == 2 ** ex if ex % 2 == 0 else True
    nonclifford = sum(1 for op in c.ops if op[1].name in ("T", "T_DAG", "CS", "CS_DAG", "CCZ"))
    assert nonclifford == 5
    assert np.abs(circuit_unitary(c)[:, 0] - v.to_vector()).max() < 1e-12
    for text in ["X0*Z1", "Y2", "Z0*Z1*Z2*Z3", "X3"]:
        p = Pauli.parse(text, 4)
        assert abs(v.expectation(p) - e.expectation(p)) < 1e-12


def test_clifford_circuit_with_measurements():
    """tests/backends.rs:77-107."""
    text = "QUBITS 3\nH 0\nCX 0 1\nCX 1 2\nS 2\nMPP Z0*Z1\nMPP X0*X1*X2\nM 0\n"
    c = Circuit.parse(text)
    v = StateVector.zero_state(3)
    rv = v.run(c, SplitMix64(3))
    assert rv[0] is False
    z0 = Pauli.parse("Z0", 3)
    sign = -1.0 if rv[2] else 1.0
    assert abs(v.expectation(z0).real / v.norm_sq().real - sign) < 1e-12


def test_kron_convention():
    """tests/backends.rs:181-200: index x_a + 2^{n_a} x_b."""
    ca = Circuit.parse("QUBITS 3\nH 0 1\nT 0\nCX 0 2\nS 2\nH 2\nCCZ 0 1 2\n")
    cb = Circuit.parse("QUBITS 2\nH 0\nCX 0 1\nT_DAG 1\nH 1\n")
    va, vb = StateVector.zero_state(3), StateVector.zero_state(2)
    va.run(ca, SplitMix64(11))
    vb.run(cb, SplitMix64(11))
    w = va.kron(vb).to_vector()
    for xa in range(8):
        for xb in range(4):
            assert abs(w[xa + 8 * xb] - va.amp[xa] * vb.amp[xb]) < 1e-12
    t = StateVector.tensor_all([va, vb, StateVector.plus_state(1)])
    assert t.n == 6 and abs(t.norm_sq() - 1) < 1e-12


def test_parse_and_print_round_trip():
    """circuit.rs:772-785."""
    text = "H 0\nCX 0 1 1 2\nT 2\nPHASE(3,4) 1\nM 0\nMX 1\nMPP X0*Z2\nR 2\nTICK\n"
    c = Circuit.parse(text)
    assert c.n == 3 and c.num_gates() == 5 and c.num_measurements() == 3
    c2 = Circuit.parse(str(c))
    assert str(c2) == str(c)
    for bad in ["CX 0", "QUBITS 1\nH 3", "FOO 1"]:
        with pytest.raises(SimError):
            Circuit.parse(bad)


def test_pauli_conventions():
    """pauli/mod.rs phase conventions: Y = i X Z, Hermiticity, dense and sparse forms."""
    y = Pauli.parse("Y0")
    assert (y.x, y.z, y.phase) == (1, 1, 1) and y.is_hermitian()
    assert Pauli.parse("-i*X0*Z1") == Pauli(2, 1, 2, 3)
    assert Pauli.parse("+XZ_Y") == Pauli.parse("X0*Z1*Y3", 4)
    assert not Pauli(1, 1, 0, 1).is_hermitian()
    m = np.zeros((2, 2), dtype=complex)
    for v in range(2):
        m[:, v] = y.apply_dense(np.eye(2, dtype=complex)[:, v])
    assert np.abs(m - np.array([[0, -1j], [1j, 0]])).max() == 0


def test_noise_sampling_tables():
    """Noise::errors / sample (circuit.rs:119-146)."""
    from oracle.circuit import noise_errors, noise_sample

    assert noise_errors("DEPOLARIZE1", (0.3,)) == [(0.3 / 3, 1, 0), (0.3 / 3, 1, 1), (0.3 / 3, 0, 1)]
    d2 = noise_errors("DEPOLARIZE2", (0.15,))
    assert len(d2) == 15 and d2[0] == (0.15 / 15, 1, 0) and d2[-1] == (0.15 / 15, 3, 3)
    assert [(x, z) for _, x, z in d2] == [(k & 3, k >> 2) for k in range(1, 16)]

    class U:
        def __init__(self, u):
            self.u = u

        def gen_f64(self):
            return self.u

    assert noise_sample("DEPOLARIZE1", (0.3,), U(0.05)) == (1, 0)
    assert noise_sample("DEPOLARIZE1", (0.3,), U(0.15)) == (1, 1)
    assert noise_sample("DEPOLARIZE1", (0.3,), U(0.25)) == (0, 1)
    assert noise_sample("DEPOLARIZE1", (0.3,), U(0.35)) is None


def test_errors_follow_the_reference():
    v = StateVector.zero_state(3)
    with pytest.raises(SimError) as e:
        v.apply_gate(G.named("CX"), [0, 0])
    assert e.value.kind == "DuplicateQubit"
    with pytest.raises(SimError) as e:
        v.apply_gate(G.named("H"), [3])
    assert e.value.kind == "QubitOutOfRange"
    with pytest.raises(SimError) as e:
        v.apply_gate(G.named("CX"), [0])
    assert e.value.kind == "Arity"
    with pytest.raises(SimError) as e:
        v.project(Pauli(3, 1, 0, 1), False)
    assert e.value.kind == "Invalid"
    with pytest.raises(SimError) as e:
        v.expectation(Pauli(2, 1, 0, 0))
    assert e.value.kind == "Invalid"
    w = v.clone()
    w.project(Pauli.parse("Z0", 3), True)
    with pytest.raises(SimError) as e:
        w.normalize()
    assert e.value.kind == "ZeroState"


# ------------------------------------------- inner, expectation_sum, probability
def dense_pauli(n, x, z, phase):
    """i^phase X^x Z^z as a 2^n x 2^n matrix built from 2x2 factors (qubit q
    is bit q of the index, so qubit 0 is the last Kronecker factor)."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Z = np.array([[1, 0], [0, -1]], dtype=complex)
    m = np.ones((1, 1), dtype=complex)
    for q in reversed(range(n)):
        f = np.eye(2, dtype=complex)
        if (x >> q) & 1:
            f = f @ X
        if (z >> q) & 1:
            f = f @ Z
        m = np.kron(m, f)
    return (1j ** phase) * m


def unnormalized_state(n, seed, scale):
    rng = np.random.default_rng(seed)
    a = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
    return scale * a


def test_inner_is_the_conjugate_linear_dot_product():
    """vector.rs:431-443: <a|b> = sum conj(a_i) b_i, conjugating the left
    state; inner(a, a) = norm_sq; a size mismatch is Invalid (432-434)."""
    for n, seed in [(1, 1), (5, 2), (11, 3), (15, 4)]:
        a = unnormalized_state(n, seed, 1.7)
        b = unnormalized_state(n, seed + 100, 0.3) + 0.2j
        sa, sb = StateVector(n, a.copy()), StateVector(n, b.copy())
        want = np.vdot(a, b)
        got = sa.inner(sb)
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))
        # the left argument is conjugated: swapping gives the conjugate
        assert abs(sb.inner(sa) - np.conj(want)) <= 1e-12 * max(1.0, abs(want))
        # a conjugated right argument would give sum a_i b_i... which differs here
        assert abs(np.sum(np.conj(a) * np.conj(b)) - want) > 1e-3
        nn = sa.norm_sq()
        assert abs(sa.inner(sa) - nn) <= 1e-12 * nn.real
        assert abs(nn - np.vdot(a, a).real) <= 1e-12 * nn.real and nn.imag == 0.0
    with pytest.raises(SimError) as e:
        StateVector.zero_state(3).inner(StateVector.zero_state(4))
    assert e.value.kind == "Invalid"


def test_expectation_sum_is_the_dense_quadratic_form():
    """mod.rs:93-101: sum_k c_k <psi|P_k|psi> with complex coefficients and
    Paulis of every phase (non-Hermitian terms included), against psi^H H psi
    with H built densely from 2x2 factors.  The state is unnormalized, so the
    result is not divided by the norm."""
    rng = np.random.default_rng(7)
    for n in (3, 6, 10):
        a = unnormalized_state(n, n, 2.5)
        s = StateVector(n, a.copy())
        terms = []
        H = np.zeros((1 << n, 1 << n), dtype=complex)
        for _ in range(11):
            x, z, ph = int(rng.integers(0, 1 << n)), int(rng.integers(0, 1 << n)), int(rng.integers(0, 4))
            c = complex(rng.normal(), rng.normal())
            terms.append((c, Pauli(n, x, z, ph)))
            H += c * dense_pauli(n, x, z, ph)
        want = np.vdot(a, H @ a)
        got = s.expectation_sum(terms)
        assert abs(got - want) <= 1e-11 * max(1.0, abs(want)), (n, got, want)
        # a conjugated coefficient gives something else
        wrong = sum(np.conj(c) * np.vdot(a, dense_pauli(n, p.x, p.z, p.phase) @ a) for c, p in terms)
        assert abs(wrong - want) > 1e-3
        # each term alone is the dense form too (expectation, vector.rs:412-429)
        for c, p in terms[:4]:
            e = s.expectation(p)
            assert abs(e - np.vdot(a, dense_pauli(n, p.x, p.z, p.phase) @ a)) <= 1e-11 * np.vdot(a, a).real


def test_probability_divides_by_the_norm():
    """mod.rs:104-108: |<x|psi>|^2 / <psi|psi> on unnormalized states."""
    for n, scale in [(4, 3.0), (9, 0.01), (13, 17.0)]:
        a = unnormalized_state(n, 30 + n, scale)
        s = StateVector(n, a.copy())
        norm = float(np.vdot(a, a).real)
        assert abs(norm - 1.0) > 1e-3
        ps = [s.probability(x) for x in range(0, 1 << n, max(1, (1 << n) // 37))]
        want = [abs(a[x]) ** 2 / norm for x in range(0, 1 << n, max(1, (1 << n) // 37))]
        assert np.allclose(ps, want, rtol=1e-13, atol=0)
    # over every index the probabilities sum to one
    a = unnormalized_state(6, 5, 4.0)
    s = StateVector(6, a.copy())
    assert abs(sum(s.probability(x) for x in range(64)) - 1.0) < 1e-13


def test_measuring_a_zero_state_draws_and_takes_outcome_zero():
    """mod.rs:120-128: norm 0 gives p_plus = NaN, which f64::clamp keeps, so
    neither certainty branch is taken, one uniform is drawn and `u >= NaN`
    is false: outcome false, weight 0."""

    class Draws:
        def __init__(self):
            self.count = 0

        def gen_f64(self):
            self.count += 1
            return 0.75

    s = StateVector.zero_state(3)
    assert s.project(Pauli.parse("Z0", 3), True) == 0
    d = Draws()
    assert s.measure(Pauli.parse("Z1", 3), d) is False
    assert d.count == 1
    assert s.norm_sq() == 0
