"""Parity of the HIP library (through its C ABI) with the CPU oracle on the
same seeded inputs.  Tolerances: 1e-12 max-abs for the complex128 form;
permutation gates bit-exact; sampling indices identical given the same
uniforms.  Each test restates a reference test where one exists."""
import math

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import workloads as W
from oracle import gates as G
from oracle.circuit import Circuit as OCircuit
from oracle.exact import ExactState
from oracle.pauli import Pauli, SimError
from oracle.statevector import SplitMix64, StateVector

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
    return a / np.linalg.norm(a)


def pair(n, seed):
    a = rand_state(n, seed)
    return StateVector(n, a.copy()), H.HipStateVector.from_amplitudes(n, a)


def maxerr(o, g):
    return float(np.abs(o.to_vector() - g.to_vector()).max())


NAMED = ["I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG", "T", "T_DAG",
         "CX", "CY", "CZ", "SWAP", "CS", "CS_DAG", "CCZ", "CCX"]
PERMS = {"I", "X", "Y", "Z", "S", "S_DAG", "CX", "CY", "CZ", "SWAP", "CCZ", "CCX"}  # exact phases 0, +-1, +-i


def gate_cases():
    out = [(nm, G.named(nm)) for nm in NAMED]
    out += [("PHASE", G.Gate("PHASE", k=3, bits=4)), ("PHASE33", G.Gate("PHASE", k=5, bits=33)),
            ("CPHASE", G.Gate("CPHASE", k=-1, bits=7)), ("RX", G.Gate("RX", theta=0.37)),
            ("RY", G.Gate("RY", theta=-2.1)), ("RZ", G.Gate("RZ", theta=4.4))]
    return out


@pytest.mark.parametrize("name,g", gate_cases(), ids=[c[0] for c in gate_cases()])
def test_every_gate_matches_the_oracle(name, g):
    n = 14
    rng = np.random.default_rng(abs(hash(name)) % 1000)
    for trial in range(3):
        o, d = pair(n, 100 + trial)
        qs = [int(q) for q in rng.permutation(n)[: g.arity()]]
        if trial == 0:
            qs = sorted(qs)
        elif trial == 1:
            qs = [n - 1 - i for i in range(g.arity())]  # high qubits, descending
        o.apply_gate(g, qs)
        d.apply_gate(g, qs)
        if name in PERMS:
            assert np.array_equal(o.to_vector(), d.to_vector()), (name, qs)
        else:
            assert maxerr(o, d) < TOL, (name, qs)


def test_fiber_kernel_matches_naive_on_high_and_low_qubits():
    """vector.rs:685-720 with k = 1..6 (k = 6 takes the out-of-place path)."""
    n = 15
    o, d = pair(n, 8)
    rng = np.random.default_rng(8)
    for trial in range(18):
        k = 1 + trial % 6
        qs = []
        while len(qs) < k:
            q = int(rng.integers(12, n)) if rng.random() < 0.5 else int(rng.integers(0, 12))
            if q not in qs:
                qs.append(q)
        m = rng.random((1 << k, 1 << k)) + 1j * rng.random((1 << k, 1 << k))
        o.apply_matrix(m, qs)
        d.apply_matrix(m, qs)
        err = maxerr(o, d)
        scale = np.abs(o.to_vector()).max()
        assert err < 1e-12 * max(1.0, scale), (trial, qs, err)
        s = 1.0 / math.sqrt(o.norm_sq().real)
        o.scale(s)
        d.scale(s)


def test_pauli_expectation_and_projection():
    """vector.rs:721-756 against the oracle (which is pinned to the dense Pauli action)."""
    n = 15
    rng = np.random.default_rng(3)
    o, d = pair(n, 9)
    for trial in range(10):
        kinds = [["I", "X", "Y", "Z"][int(rng.integers(0, 4))] for _ in range(n)]
        p = Pauli.from_kinds(kinds)
        if trial % 3 == 0:
            p = p.negated()
        assert abs(o.expectation(p) - d.expectation(p)) < TOL
        u, v = o.clone(), d.clone()
        u.apply_pauli(p)
        v.apply_pauli(p)
        assert np.array_equal(u.to_vector(), v.to_vector())
        b = bool(rng.integers(0, 2))
        w1, w2 = o.clone(), d.clone()
        a1 = w1.project(p, b)
        a2 = w2.project(p, b)
        assert abs(a1 - a2) < TOL
        assert np.array_equal(w1.to_vector(), w2.to_vector())
    # Z-type projections keep or zero entries exactly
    z = Pauli.parse("Z3*Z11", n)
    w1, w2 = o.clone(), d.clone()
    assert abs(w1.project(z, True) - w2.project(z, True)) < TOL
    assert np.array_equal(w1.to_vector(), w2.to_vector())


@pytest.mark.parametrize("fusion", [1, 2])
def test_batched_expectations_match_single(fusion):
    """Grouped expectations (LDS-partner kernel below 20 qubits, specialized
    kernels with fusion 2) against the oracle's one-at-a-time values."""
    n = 16
    o, d = pair(n, 4)
    d.set_fusion(fusion)
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 40, 5, max_weight=6)]
    got = d.expectations(ps)
    for p, g in zip(ps, got):
        assert abs(o.expectation(p) - g) < TOL
    terms = [(complex(0.3 * i, -0.1 * i), p) for i, p in enumerate(ps)]
    assert abs(o.expectation_sum(terms) - d.expectation_sum(terms)) < 1e-11


def test_reductions_scalars_and_readout():
    n = 13
    o, d = pair(n, 5)
    assert abs(o.norm_sq() - d.norm_sq()) < TOL
    o2, d2 = pair(n, 6)
    assert abs(o.inner(o2) - d.inner(d2)) < TOL
    for x in (0, 5, (1 << n) - 1):
        assert o.amplitude(x) == d.amplitude(x)
        assert abs(o.probability(x) - d.probability(x)) < TOL
    c = complex(0.3, -1.7)
    o.scale(c)
    d.scale(c)
    assert np.array_equal(o.to_vector(), d.to_vector())
    o.normalize()
    d.normalize()
    assert maxerr(o, d) < TOL
    # construction
    p = H.HipStateVector.plus_state(9)
    assert np.array_equal(p.to_vector(), StateVector.plus_state(9).to_vector())
    b = H.HipStateVector.basis_state(9, 77)
    assert np.array_equal(b.to_vector(), StateVector.basis_state(9, 77).to_vector())


def test_kron_convention():
    """tests/backends.rs:181-200."""
    ca = "QUBITS 3\nH 0 1\nT 0\nCX 0 2\nS 2\nH 2\nCCZ 0 1 2\n"
    cb = "QUBITS 2\nH 0\nCX 0 1\nT_DAG 1\nH 1\n"
    va, vb = StateVector.zero_state(3), StateVector.zero_state(2)
    va.run(OCircuit.parse(ca), SplitMix64(1))
    vb.run(OCircuit.parse(cb), SplitMix64(1))
    ga, gb = H.HipStateVector.zero_state(3), H.HipStateVector.zero_state(2)
    ga.run(H.Circuit.parse(ca))
    gb.run(H.Circuit.parse(cb))
    want = va.kron(vb).to_vector()
    got = ga.tensor(gb).to_vector()
    assert np.abs(want - got).max() < TOL
    t = H.HipStateVector.tensor_all([ga, gb, H.HipStateVector.plus_state(1)])
    assert t.num_qubits() == 6 and abs(t.norm_sq() - 1) < TOL


def test_sampling_matches_the_oracle_given_the_same_uniforms():
    """vector.rs:767-800 (Born rule, 200k shots, never a zero-weight index) and
    identical indices against the oracle for the same draws."""
    n = 14
    i = np.arange(1 << n)
    amp = np.where(i % 4099 == 7, 3.0, 0.01 * (i % 13)).astype(complex)
    o = StateVector(n, amp)
    d = H.HipStateVector.from_amplitudes(n, amp)
    us = np.random.default_rng(2).random(200_000)
    so = o.sample_with_uniforms(us)
    sd = d.sample_with_uniforms(us)
    assert np.array_equal(so, sd)
    total = np.sum(np.abs(amp) ** 2)
    counts = np.bincount(sd.astype(np.int64), minlength=1 << n)
    for j in range(7, 1 << n, 4099):
        p = abs(amp[j]) ** 2 / total
        assert abs(counts[j] / len(us) - p) < 6 * math.sqrt(p * (1 - p) / len(us))
    assert np.all((sd % 13 != 0) | (sd % 4099 == 7))
    # a vanished state is ZeroState
    z = H.HipStateVector.zero_state(3)
    z.project(Pauli.parse("Z0", 3), True)
    with pytest.raises(H.NcsvError) as e:
        z.sample_with_uniforms([0.5])
    assert e.value.kind == "ZeroState"


def boundary_uniforms(amp, blocks, per_block=6):
    """Draws u/total that land on, just below and just above the block
    prefixes acc_b + w_b and the in-block running sums of vector.rs:509-535,
    computed here in the reference's order (sequential sums of re*re + im*im)."""
    p = amp.real * amp.real + amp.imag * amp.imag
    bl = min(4096, len(amp))
    w = np.array([np.cumsum(p[b * bl:(b + 1) * bl])[-1] for b in range(len(amp) // bl)])
    hi = np.cumsum(w)
    total = hi[-1]
    targets = []
    rng = np.random.default_rng(len(amp))
    for b in blocks:
        targets.append(hi[b])
        if b:
            targets.append(hi[b - 1])
        run = hi[b - 1] if b else 0.0
        stops = set(int(j) for j in rng.integers(0, bl, per_block))
        for j in range(bl):
            run = run + p[b * bl + j]
            if j in stops:
                targets.append(run)
    us = []
    for t in targets:
        u = t / total
        for v in (u, np.nextafter(u, 0.0), np.nextafter(u, 1.0), np.nextafter(np.nextafter(u, 0.0), 0.0)):
            if 0.0 <= v < 1.0:
                us.append(v)
    return np.array(us + [0.0, np.nextafter(1.0, 0.0)])


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_sampling_is_bit_exact_at_block_and_prefix_boundaries(precision):
    """vector.rs:494-540 at 22 qubits (1024 blocks): identical indices for
    draws aimed at block prefixes and in-block running sums, on a state with
    runs of exact zeros (the zero-index back-off).  The complex64 state holds
    values exact in float32, so its FP64 weights equal the oracle's."""
    n = 22
    rng = np.random.default_rng(22)
    amp = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
    amp[rng.integers(0, 1 << n, 1 << 16)] = 0.0
    amp[4096 * 5: 4096 * 6] = 0.0  # an empty block
    amp[4096 * 9 + 4000: 4096 * 10 + 50] = 0.0  # zeros across a block end
    amp = (amp.real.astype(np.float32) + 1j * amp.imag.astype(np.float32)).astype(np.complex128)
    o = StateVector(n, amp.copy())
    d = H.HipStateVector.from_amplitudes(n, amp, precision)
    us = boundary_uniforms(amp, [0, 4, 5, 6, 9, 10, 511, 1023])
    us = np.concatenate([us, rng.random(50_000)])
    so = o.sample_with_uniforms(us)
    sd = d.sample_with_uniforms(us)
    bad = np.nonzero(so != sd)[0]
    assert len(bad) == 0, (len(bad), us[bad[:5]], so[bad[:5]], sd[bad[:5]])
    assert np.all(amp[sd.astype(np.int64)] != 0)
    # a second, smaller call reuses the handle's scratch
    assert np.array_equal(o.sample_with_uniforms(us[:100]), d.sample_with_uniforms(us[:100]))
    assert len(d.sample_with_uniforms([])) == 0


def test_ghz_and_clifford_t_against_the_exact_oracle():
    """vector.rs:580-633 and tests/backends.rs:18-75."""
    c = OCircuit(3)
    c.gate("H", [0]).gate("CX", [0, 1]).gate("CX", [1, 2])
    d = H.HipStateVector.zero_state(3)
    d.run(H.Circuit.parse(str(c)))
    e = ExactState(3).run(c)
    assert np.abs(np.array(e.to_vector()) - d.to_vector()).max() < 1e-15
    assert abs(d.expectation(Pauli.parse("X0*X1*X2")) - 1) < 1e-15
    assert d.expectation(Pauli.parse("Z0", 3)) == 0
    w = d.clone().project(Pauli.parse("Z0", 3), True)
    assert abs(w.real - 0.5) < 2e-16
    from tests.test_oracle import CLIFFORD_T

    oc = OCircuit.parse(CLIFFORD_T)
    d = H.HipStateVector.zero_state(4)
    d.run(H.Circuit.parse(CLIFFORD_T))
    e = ExactState(4).run(oc)
    assert np.abs(np.array(e.to_vector()) - d.to_vector()).max() < 1e-14
    for text in ["X0*Z1", "Y2", "Z0*Z1*Z2*Z3", "X3"]:
        p = Pauli.parse(text, 4)
        assert abs(d.expectation(p) - e.expectation(p)) < 1e-12


def run_both(text, seed, fusion=None):
    o = StateVector.zero_state(OCircuit.parse(text).n)
    ro = o.run(OCircuit.parse(text), SplitMix64(seed))
    d = H.HipStateVector.zero_state(o.n)
    if fusion is not None:
        d.set_fusion(fusion)
    rd = d.run(H.Circuit.parse(text), SplitMix64(seed))
    return o, d, ro, rd


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_large_copies_through_pinned_staging(precision, monkeypatch):
    """set_amplitudes and to_vector of 256 MiB go through the pinned staging
    buffers (chunked, threaded host copies) and equal one plain copy each."""
    n = 24
    rng = np.random.default_rng(3)
    a = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
    got = []
    for stage in ("1", "0"):
        monkeypatch.setenv("NCSV_STAGE", stage)
        d = H.HipStateVector.from_amplitudes(n, a, precision)
        got.append(d.to_vector())
    assert np.array_equal(got[0], got[1])
    if precision == H.C128:
        assert np.array_equal(got[0], a)
    else:
        assert np.abs(got[0] - a).max() < 1e-6


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_out_of_place_sweeps_equal_in_place(precision, monkeypatch):
    """NCSV_JIT_OOP=1: the specialized sweeps write a second buffer and the
    two swap after each sweep; states, expectations and a second run on the
    swapped buffers equal the in-place runs bit for bit."""
    n = 22
    circ = H.Circuit.parse(W.rotation_brickwork(n, 8, 11))
    ps = [H.Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 16, 12)]
    res = []
    for oop in ("0", "1"):
        monkeypatch.setenv("NCSV_JIT_OOP", oop)
        d = H.HipStateVector.zero_state(n, precision)
        d.run(circ)
        v1 = d.to_vector()
        d.run(circ)  # again, from the buffer the first run ended in
        res.append((v1, d.to_vector(), d.expectations(ps)))
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_deferred_basis_state(precision):
    """set_basis defers its fill: a run's first specialized sweep (22 qubits)
    starts from the basis state in registers, and any other reader writes the
    state first.  Every path against the oracle started from the same basis
    state."""
    n = 22
    text = W.rotation_brickwork(n, 6, 9)
    tol = 1e-12 if precision == H.C128 else 2e-5
    d = H.HipStateVector.zero_state(n, precision)
    for b in (0, 0x2A5F3, (1 << n) - 1):
        o = StateVector.basis_state(n, b)
        o.run(OCircuit.parse(text), SplitMix64(0))
        want = o.to_vector()
        # run straight after set_basis: the sweep from the basis state
        d.set_basis(b)
        d.run(H.Circuit.parse(text))
        assert np.abs(d.to_vector() - want).max() < tol, b
        # readers before a run write the basis state first
        d.set_basis(b)
        assert abs(d.amplitude(b) - 1.0) < 1e-15 and abs(d.norm_sq() - 1.0) < 1e-15
        d.run(H.Circuit.parse(text))
        assert np.abs(d.to_vector() - want).max() < tol, b
        d.set_basis(b)
        x = np.zeros(1 << n)
        x[b] = 1.0
        assert np.array_equal(np.abs(d.to_vector()), x)
        # a run whose first op is a measurement writes it first as well
        d.set_basis(b)
        rec = d.run(H.Circuit.parse(f"QUBITS {n}\nM 3\n" + text.split("\n", 1)[1]), SplitMix64(1))
        assert rec == [bool((b >> 3) & 1)]


def test_measurements_resets_and_noise_follow_the_same_draws():
    """mod.rs:113-165: the records and final states agree when both sides
    draw from the same generator."""
    text = ("QUBITS 5\nH 0 1 2 3 4\nT 1\nCX 0 1 2 3\nM 0\nMPP X1*X2 Z3*Z4\nR 2\nDEPOLARIZE1(0.3) 0 1 2 3 4\n"
            "RY(0.4) 3\nMY 3\nDEPOLARIZE2(0.5) 0 4 1 2\nX_ERROR(0.5) 3\nM 1 2 3 4\n")
    for seed in range(6):
        o, d, ro, rd = run_both(text, seed)
        assert ro == rd, seed
        assert maxerr(o, d) < TOL


def test_measured_clifford_circuit():
    """tests/backends.rs:77-107."""
    text = "QUBITS 3\nH 0\nCX 0 1\nCX 1 2\nS 2\nMPP Z0*Z1\nMPP X0*X1*X2\nM 0\n"
    o, d, ro, rd = run_both(text, 3)
    assert rd[0] is False and ro == rd
    z0 = Pauli.parse("Z0", 3)
    sign = -1.0 if rd[2] else 1.0
    assert abs(d.expectation(z0).real / d.norm_sq().real - sign) < 1e-12


@pytest.mark.parametrize("fusion", [0, 1])
def test_random_clifford_t_layers_exact_at_12_qubits(fusion):
    """The 20-qubit random-layer workload's family at 12 qubits and depth 30,
    against the exact Z[w] simulator."""
    text = W.random_layers(12, 30, 7, measure=False)
    e = ExactState(12).run(OCircuit.parse(text))
    d = H.HipStateVector.zero_state(12)
    d.set_fusion(fusion)
    d.run(H.Circuit.parse(text))
    assert np.abs(np.array(e.to_vector()) - d.to_vector()).max() < 1e-12


@pytest.mark.parametrize("fusion", [0, 1])
def test_random_layers_20_qubits_with_measurements(fusion):
    """20 qubits, 200 layers of {H,S,T,SQRT_X} + random CX matching, then M
    on all qubits and 1000 shots: amplitudes before measuring, the record and
    the shots, against the oracle."""
    text = W.random_layers(20, 200, 11, measure=False)
    o, d, _, _ = run_both(text, 1, fusion)
    assert maxerr(o, d) < TOL
    us = np.random.default_rng(0).random(1000)
    assert np.array_equal(o.sample_with_uniforms(us), d.sample_with_uniforms(us))
    meas = "QUBITS 20\nM " + " ".join(map(str, range(20))) + "\n"
    ro = o.run(OCircuit.parse(meas), SplitMix64(5))
    rd = d.run(H.Circuit.parse(meas), SplitMix64(5))
    assert ro == rd


@pytest.mark.parametrize("fusion", [0, 1])
def test_rotation_brickwork_22_qubits_and_64_paulis(fusion):
    """The 30-qubit brickwork's family (RX/RY/RZ + CZ, depth 24) at 22 qubits:
    amplitudes and 64 Pauli expectations within 1e-12 of the oracle."""
    n = 22
    text = W.rotation_brickwork(n, 24, 3)
    o, d, _, _ = run_both(text, 0, fusion)
    assert maxerr(o, d) < TOL
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 64, 4)]
    got = d.expectations(ps)
    for p, g in zip(ps, got):
        assert abs(o.expectation(p) - g) < TOL


def test_qft_phase_bits_beyond_32():
    """CPHASE(1, bits) up to bits = 24 here (the reference stops at 32 bits,
    phasepoly.rs:550; the library takes 62): against the oracle."""
    text = W.qft_with_rz(24, 50, 2)
    o, d, _, _ = run_both(text, 0)
    assert maxerr(o, d) < TOL
    # the closed form of one 33-bit phase
    s = H.HipStateVector.zero_state(2)
    s.apply_gate("X", [0])
    s.apply_gate("X", [1])
    s.apply_gate(H.Gate("CPHASE", k=1, bits=33), [0, 1])
    a = s.amplitude(3)
    assert abs(a - complex(math.cos(2 * math.pi / 2 ** 33), math.sin(2 * math.pi / 2 ** 33))) < 1e-16


def test_errors_follow_the_reference():
    d = H.HipStateVector.zero_state(3)
    for g, qs, kind in [("CX", [0, 0], "DuplicateQubit"), ("H", [3], "QubitOutOfRange"), ("CX", [0], "Arity")]:
        with pytest.raises(H.NcsvError) as e:
            d.apply_gate(g, qs)
        assert e.value.kind == kind
    with pytest.raises(H.NcsvError) as e:
        d.project(Pauli(3, 1, 0, 1), False)
    assert e.value.kind == "Invalid"
    with pytest.raises(H.NcsvError) as e:
        d.expectation(Pauli(2, 1, 0, 0))
    assert e.value.kind == "Invalid"
    w = d.clone()
    w.project(Pauli.parse("Z0", 3), True)
    with pytest.raises(H.NcsvError) as e:
        w.normalize()
    assert e.value.kind == "ZeroState"
    with pytest.raises(H.NcsvError) as e:
        d.run(H.Circuit.parse("QUBITS 4\nH 0\n"))
    assert e.value.kind == "Invalid"


def test_complex64_form_tracks_complex128():
    """The complex64 form (FP32 amplitudes, FP64 reductions and scale): its
    fidelity with the complex128 form on a 20-qubit brickwork and after
    measurements."""
    n = 20
    text = W.rotation_brickwork(n, 12, 9) + "M 0 1 2\n"
    a = H.HipStateVector.zero_state(n, H.C128)
    b = H.HipStateVector.zero_state(n, H.C64)
    ra = a.run(H.Circuit.parse(text), SplitMix64(4))
    rb = b.run(H.Circuit.parse(text), SplitMix64(4))
    assert ra == rb
    va, vb = a.to_vector(), b.to_vector()
    fid = abs(np.vdot(va, vb)) ** 2 / (np.vdot(va, va).real * np.vdot(vb, vb).real)
    assert 1 - fid < 1e-5
    assert np.abs(va - vb).max() < 1e-5
    assert abs(a.norm_sq() - b.norm_sq()) < 1e-6


def _mixed():
    from tests.test_plan import MIXED

    return MIXED + MIXED.split("\n", 1)[1] * 2


@pytest.mark.parametrize("fusion", [1, 2])
@pytest.mark.parametrize("name,text", [
    ("mixed", None),
    ("brickwork", W.rotation_brickwork(15, 12, 3)),
    ("layers", W.random_layers(14, 20, 5, measure=False)),
    ("qft", W.qft_with_rz(14, 40, 1)),
], ids=["mixed", "brickwork", "layers", "qft"])
def test_fused_sweeps_match_the_oracle(name, text, fusion):
    """The fused tile sweeps — op-list interpreter (fusion 1 below 20 qubits)
    and specialized kernels (fusion 2) — against the oracle, every op kind."""
    text = text or _mixed()
    n = OCircuit.parse(text).n
    a = rand_state(n, 21)
    o = StateVector(n, a.copy())
    o.run(OCircuit.parse(text), SplitMix64(0))
    d = H.HipStateVector.from_amplitudes(n, a)
    d.set_fusion(fusion)
    d.run(H.Circuit.parse(text))
    assert maxerr(o, d) < TOL
    # complex64 form through the same sweeps
    f = H.HipStateVector.from_amplitudes(n, a, H.C64)
    f.set_fusion(fusion)
    f.run(H.Circuit.parse(text))
    assert np.abs(f.to_vector() - o.to_vector()).max() < 2e-5


def test_large_state_round_trip():
    """32 qubits, complex64 (2^32 amplitudes: launches need grid-stride
    loops): H on every qubit gate by gate gives the uniform state (closed
    form 2^-16); the fused sweeps undo it back to |0...0>."""
    n = 32
    s = H.HipStateVector.zero_state(n, H.C64)
    s.set_fusion(0)
    text = f"QUBITS {n}\nH " + " ".join(map(str, range(n))) + "\n"
    s.run(H.Circuit.parse(text))
    for x in (0, 1, (1 << n) - 1, 0x9E3779B9):
        assert abs(s.amplitude(x) - 2.0 ** -16) < 1e-6 * 2.0 ** -16
    assert abs(s.norm_sq() - 1) < 1e-6
    s.set_fusion(1)
    s.run(H.Circuit.parse(text))
    assert abs(s.amplitude(0) - 1) < 1e-5
    assert abs(s.amplitude(12345)) < 1e-6


def test_apply_monomial_diagonal_phase_polynomial_and_permutations():
    """StateVector::apply_monomial (vector.rs:96-139) through the C ABI with
    caller-computed tables: the Diagonal(PhasePoly) of tests/backends.rs:109-136
    gives <0|H D_f H|0> = 2^-3 sum_x zeta8^f(x); 3- and 4-qubit permutations
    with unit phases (1, -1, i, -i) move amplitudes bit-exactly; a permutation
    with general phases matches the oracle within 1e-12."""
    f = G.PhasePoly(3, 3)
    f.add_monomial(1, [0])
    f.add_monomial(3, [1, 2])
    f.add_monomial(4, [0, 1, 2])
    mono = G.Gate("DIAG", poly=f).monomial()
    table = mono.images()
    for prec, tol in ((H.C128, 1e-15), (H.C64, 5e-7)):
        d = H.HipStateVector.zero_state(3, prec)
        for q in range(3):
            d.apply_gate("H", [q])
        d.apply_monomial([0, 1, 2], [y for _, y in table], [c for c, _ in table])
        for q in range(3):
            d.apply_gate("H", [q])
        want = sum(np.exp(2j * np.pi * (((x & 1) + 3 * ((x >> 1) & 1) * ((x >> 2) & 1) + 4 * (x == 7)) % 8) / 8)
                   for x in range(8)) / 8
        assert abs(d.amplitude(0) - want) < tol
    rng = np.random.default_rng(12)
    n = 13
    units = [1, -1, 1j, -1j]
    for k in (3, 4):
        for trial in range(3):
            o, d = pair(n, 300 + 10 * k + trial)
            qs = [int(q) for q in rng.permutation(n)[:k]]
            images = [int(v) for v in rng.permutation(1 << k)]
            phases = np.array([units[int(i)] for i in rng.integers(0, 4, 1 << k)], dtype=complex)
            ref = apply_perm_numpy(o.to_vector(), qs, images, phases)
            d.apply_monomial(qs, images, phases)
            # moves and unit phases are exact: the permuted, sign-flipped input
            assert np.array_equal(ref, d.to_vector()), (k, trial)
        o, d = pair(n, 400 + k)
        qs = [int(q) for q in rng.permutation(n)[:k]]
        images = [int(v) for v in rng.permutation(1 << k)]
        phases = np.exp(2j * np.pi * rng.random(1 << k))
        ref = apply_perm_numpy(o.to_vector(), qs, images, phases)
        d.apply_monomial(qs, images, phases)
        assert np.abs(ref - d.to_vector()).max() < TOL


def apply_perm_numpy(v, qs, images, phases):
    """out[y(i)] = phase(x(i)) v[i], where x(i) is the local index of i on
    qs and y(i) replaces it by images[x(i)]."""
    idx = np.arange(len(v))
    loc = np.zeros(len(v), dtype=np.int64)
    for j, q in enumerate(qs):
        loc |= ((idx >> q) & 1) << j
    img = np.asarray(images, dtype=np.int64)[loc]
    dst = idx.copy()
    for j, q in enumerate(qs):
        dst = (dst & ~(1 << q)) | (((img >> j) & 1) << q)
    out = np.zeros_like(v)
    ph = np.asarray(phases)[loc]
    unit = np.isin(ph, [1, -1, 1j, -1j])
    prod = v * ph
    # exact for unit phases: multiply by +-1 / +-i as sign flips and swaps
    exact = np.where(ph == 1, v, np.where(ph == -1, -v, np.where(ph == 1j, 1j * v.real - v.imag,
                                                                 np.where(ph == -1j, v.imag - 1j * v.real, prod))))
    out[dst] = np.where(unit, exact, prod)
    return out


def test_expectations_from_several_host_threads_on_one_handle():
    """expectation_sum calls expectation on one state from many threads at
    once (mod.rs:93-101, bridge.rs:459-464): concurrent calls on one handle
    give the same values as one thread."""
    import concurrent.futures as cf

    n = 20
    d = H.HipStateVector.zero_state(n)
    d.run(H.Circuit.parse(W.rotation_brickwork(n, 8, 5)))
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 48, 6)]
    want = [d.expectation(p) for p in ps]
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        got = list(ex.map(d.expectation, ps))
        sums = list(ex.map(lambda k: d.expectation_sum([(1.0, p) for p in ps[k:k + 6]]), range(0, 48, 6)))
    assert got == want
    for k, s in zip(range(0, 48, 6), sums):
        assert abs(s - sum(want[k:k + 6])) < 1e-12


@pytest.mark.parametrize("fusion", [1, 2])
def test_complex64_multi_pauli_expectations(fusion):
    """The grouped expectation kernels on complex64 tiles (13 Paulis per
    group) against the oracle: a 22-qubit brickwork with 64 random Paulis
    under fusion 1 and 2, within 2e-5 of the FP64 values."""
    n = 22
    text = W.rotation_brickwork(n, 10, 13)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    d = H.HipStateVector.zero_state(n, H.C64)
    d.set_fusion(fusion)
    d.run(H.Circuit.parse(text))
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 64, 14)]
    got = d.expectations(ps)
    assert max(abs(o.expectation(p) - g) for p, g in zip(ps, got)) < 2e-5


@pytest.mark.parametrize("precision,tol", [(H.C128, 1e-12), (H.C64, 2e-5)])
def test_multi_pauli_expectations_with_any_phase(precision, tol):
    """Paulis whose phase is not the Hermitian one (i^k P with k shifted by
    one or three) have purely imaginary expectations: the grouped kernels at
    22 qubits put those terms on the imaginary sums, Hermitian ones on the
    real sums, and both match the oracle (mod.rs:102-108)."""
    n = 22
    text = W.rotation_brickwork(n, 10, 21)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    d = H.HipStateVector.zero_state(n, precision)
    d.run(H.Circuit.parse(text))
    r = W.SplitMix64(22)
    ps = [Pauli(n, x, z, (ph + r.below(4)) & 3) for x, z, ph in W.random_paulis(n, 40, 23)]
    want = [o.expectation(p) for p in ps]
    got = d.expectations(ps)
    assert max(abs(w - g) for w, g in zip(want, got)) < tol
    # both sums carry values in this set
    assert max(abs(w.imag) for w in want) > 1e-3 and max(abs(w.real) for w in want) > 1e-3


def test_complex64_thirteen_pauli_group_on_a_small_state(monkeypatch):
    """NCSV_EXPECT_GMAX=13 under fusion 2 on 16 qubits: a 13-Pauli group on a
    complex64 state smaller than the default specialization size."""
    monkeypatch.setenv("NCSV_EXPECT_GMAX", "13")
    n = 16
    text = W.rotation_brickwork(n, 8, 15)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    d = H.HipStateVector.zero_state(n, H.C64)
    d.set_fusion(2)
    d.run(H.Circuit.parse(text))
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 26, 16)]
    got = d.expectations(ps)
    assert max(abs(o.expectation(p) - g) for p, g in zip(ps, got)) < 2e-5


@pytest.mark.parametrize("precision", [H.C128, H.C64])
def test_measuring_a_zero_state_follows_the_reference(precision):
    """mod.rs:120-128 on a zero state: p_plus = NaN survives f64::clamp, one
    uniform is drawn and the outcome is false, as in the oracle."""
    o = StateVector.zero_state(3)
    o.project(Pauli.parse("Z0", 3), True)
    d = H.HipStateVector.zero_state(3, precision)
    d.project(Pauli.parse("Z0", 3), True)
    ro, rd = SplitMix64(3), SplitMix64(3)
    assert o.measure(Pauli.parse("Z1", 3), ro) is False
    assert d.measure(Pauli.parse("Z1", 3), rd) is False
    assert ro.gen_f64() == rd.gen_f64()  # the same number of draws on both sides


def test_jit_cache_evicts_without_breaking_plans(monkeypatch):
    """With room for three compiled kernels, new expectation kernels evict
    old ones (unloading their modules) while a circuit's plan keeps its
    kernels: expectations of cycling Pauli sets and a re-run of the planned
    circuit stay equal to the oracle."""
    monkeypatch.setenv("NCSV_JIT_CACHE", "3")
    n = 20
    text = W.rotation_brickwork(n, 6, 21)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    circ = H.Circuit.parse(text)
    d = H.HipStateVector.zero_state(n)
    d.run(circ)
    sets = [[Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 6, 30 + k)] for k in range(5)]
    for rnd in range(2):
        for ps in sets:
            got = d.expectations(ps)
            assert max(abs(o.expectation(p) - g) for p, g in zip(ps, got)) < TOL
        d.set_basis(0)
        d.run(circ)
        assert maxerr(o, d) < TOL


def test_out_of_device_memory_cap_on_the_device(monkeypatch):
    """NCSV_DEVICE_MEMORY_LIMIT_GB: a state above the cap is a Device error
    with no handle; states below it, and the handles made before, work."""
    a = H.HipStateVector.zero_state(12)
    monkeypatch.setenv("NCSV_DEVICE_MEMORY_LIMIT_GB", "0.2")  # units of 2^30 bytes: 204.8 MiB
    with pytest.raises(H.NcsvError) as e:
        H.HipStateVector.zero_state(24)  # 256 MiB of complex128
    assert e.value.kind == "Device"
    with pytest.raises(H.NcsvError) as e:
        H.HipStateVector.zero_state(12).tensor(H.HipStateVector.zero_state(12))
    assert e.value.kind == "Device"
    b = H.HipStateVector.zero_state(23, H.C64)  # 64 MiB
    b.apply_gate("H", [22])
    assert abs(b.norm_sq() - 1) < 1e-6
    a.apply_gate("X", [0])
    assert abs(a.amplitude(1) - 1) == 0
