"""The benchmark configurations at (or near) their stated sizes on the GPU,
each against the oracle where the oracle finishes in seconds and through
size-independent properties where it does not:

- the rotation brickwork (RX/RY/RZ + CZ, depth 24) at 26 qubits against the
  oracle, and at 30 qubits fused against unfused and run back to |0...0>;
- QFT + 2000 RZ at 32 qubits, complex64 against complex128 (fidelity loss
  and max-abs), and the 33-qubit complex64 state;
- a slice of the 10-22-qubit circuit mix in one batched launch;
- 24-qubit noisy trajectories sharing one program.

Tolerances: complex128 1e-12 max-abs; complex64 2e-5 against FP64 results
and fidelity loss <= 1e-5."""
import math
import re

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import workloads as W
from oracle.circuit import Circuit as OCircuit
from oracle.pauli import Pauli
from oracle.statevector import SplitMix64, StateVector
from tests.test_batch import _draws, oracle_entry

pytestmark = pytest.mark.gpu

TOL = 1e-12


def inverse_of(text):
    """The inverse of a rotation/CZ circuit: gates in reverse order with
    negated angles (RX/RY/RZ(-t) undo RX/RY/RZ(t); CZ is its own inverse)."""
    lines = text.strip().splitlines()
    out = [lines[0]]
    for line in reversed(lines[1:]):
        m = re.match(r"(R[XYZ])\((.*)\) (.*)", line)
        out.append(f"{m.group(1)}({-float(m.group(2))!r}) {m.group(3)}" if m else line)
    return "\n".join(out) + "\n"


def test_brickwork_26_qubits_amplitudes_and_64_paulis():
    """The 30-qubit brickwork's stated check: the same family at 26 qubits,
    every amplitude and 64 Pauli expectations within 1e-12 of the oracle."""
    n = 26
    text = W.rotation_brickwork(n, 24, 3)
    print(f"brickwork {n} q: oracle", flush=True)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    d = H.HipStateVector.zero_state(n)
    d.run(H.Circuit.parse(text))
    err = float(np.abs(o.to_vector() - d.to_vector()).max())
    ps = [Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 64, 4)]
    got = d.expectations(ps)
    perr = max(abs(o.expectation(p) - g) for p, g in zip(ps, got))
    print(f"brickwork {n} q: max|dpsi| = {err:.2e}, max|d<P>| = {perr:.2e}", flush=True)
    assert err < TOL and perr < TOL


def test_brickwork_30_qubits_fused_equals_unfused_and_inverts():
    """At 30 qubits (no oracle at this size): the fused plan equals the
    gate-by-gate kernels (each checked against the oracle elsewhere) within
    1e-12, the 64 expectations of Hermitian Paulis are real and within
    [-1, 1], and the inverse circuit takes the state back to |0...0>."""
    n = 30
    text = W.rotation_brickwork(n, 24, 3)
    c = H.Circuit.parse(text)
    a = H.HipStateVector.zero_state(n)
    a.run(c)
    b = H.HipStateVector.zero_state(n)
    b.set_fusion(0)
    b.run(c)
    cmp = a.compare(b)
    print(f"30 q fused vs unfused: max-abs {cmp['max_abs']:.2e}, fidelity loss {1 - cmp['fidelity']:.2e}", flush=True)
    assert cmp["max_abs"] < TOL
    del b
    ps = [H.Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 64, 4)]
    ev = a.expectations(ps)
    assert max(abs(e.imag) for e in ev) < 1e-14
    assert max(abs(e.real) for e in ev) <= 1.0 + 1e-12
    assert abs(a.norm_sq() - 1.0) < TOL
    a.run(H.Circuit.parse(inverse_of(text)))
    assert abs(a.amplitude(0) - 1.0) < TOL
    assert abs(a.norm_sq() - 1.0) < TOL


def test_qft_32_qubits_complex64_against_complex128():
    """QFT + 2000 RZ (X prefix) at 32 qubits in both forms: fidelity loss at
    most 1e-5 and max-abs amplitude error reported (and below 1e-5)."""
    n = 32
    c = H.Circuit.parse(W.qft_with_rz(n, 2000, 7))
    a = H.HipStateVector.zero_state(n, H.C128)
    a.run(c)
    b = H.HipStateVector.zero_state(n, H.C64)
    b.run(c)
    cmp = a.compare(b)
    print(f"QFT 32 q: fidelity loss {1 - cmp['fidelity']:.3e}, max-abs {cmp['max_abs']:.3e}", flush=True)
    assert 1 - cmp["fidelity"] <= 1e-5
    assert cmp["max_abs"] < 1e-5
    assert abs(cmp["norm_self"] - 1) < TOL and abs(cmp["norm_other"] - 1) < 1e-5


def test_qft_33_qubits_complex64():
    """The 33-qubit complex64 state (64 GiB): QFT + 2000 RZ keeps the norm in
    FP64 to 1e-5; a 33-bit controlled phase on |1...1> equals the closed form
    exp(2 pi i / 2^33) (the reference's PhasePoly stops at 32 bits)."""
    n = 33
    s = H.HipStateVector.zero_state(n, H.C64)
    s.run(H.Circuit.parse(W.qft_with_rz(n, 2000, 7)))
    nn = s.norm_sq().real
    print(f"QFT 33 q complex64: norm_sq - 1 = {nn - 1:.3e}", flush=True)
    assert abs(nn - 1.0) < 1e-5
    idx = 1 | (1 << 32)
    s.set_basis(idx)
    s.apply_gate(H.Gate("CPHASE", k=1, bits=33), [32, 0])
    want = complex(math.cos(2 * math.pi / 2 ** 33), math.sin(2 * math.pi / 2 ** 33))
    assert abs(s.amplitude(idx) - want) < 1e-15
    assert abs(s.norm_sq() - 1.0) < 1e-12


def test_circuit_mix_slice_in_one_launch():
    """The first 40 circuits of the uneven 10-22-qubit mix (the benchmark's
    seed), all in one batched launch per precision: complex128 entries equal
    the oracle within 1e-12 (every circuit up to 17 qubits and four of
    20-22 qubits checked), complex64 entries equal the complex128 ones
    within 2e-5."""
    specs = W.batch_circuits(40, 11)
    ps = [[Pauli(n, x, z, ph) for x, z, ph in p] for n, _, p in specs]
    circs = [H.Circuit.parse(t) for _, t, _ in specs]
    res = {}
    for prec in (H.C128, H.C64):
        b = H.HipBatch(prec)
        for c, p in zip(circs, ps):
            b.add(c, p)
        b.run()
        assert b.stats()["launches"] == 1
        res[prec] = [b.result(i) for i in range(len(specs))]
    check = [i for i, (n, _, _) in enumerate(specs) if n <= 17] + [2, 4, 12, 21]
    print(f"mix slice: sizes {sorted(set(n for n, _, _ in specs))}, oracle on {len(check)}", flush=True)
    for i in check:
        n, text, _ = specs[i]
        rex, rnn, _ = oracle_entry(text, ps[i], [])
        ex, nn, _ = res[H.C128][i]
        assert max(abs(a - r) for a, r in zip(ex, rex)) < TOL, (i, n)
        assert abs(nn - rnn) < TOL
    worst = 0.0
    for i in range(len(specs)):
        e128, n128, _ = res[H.C128][i]
        e64, n64, _ = res[H.C64][i]
        worst = max(worst, max(abs(a - c) for a, c in zip(e128, e64)), abs(n128 - n64))
    print(f"mix slice: max |complex64 - complex128| = {worst:.2e}", flush=True)
    assert worst < 2e-5


@pytest.mark.parametrize("precision", [H.C64, H.C128])
def test_noisy_trajectories_24_qubits_share_a_program(precision, capfd, monkeypatch):
    """Sixteen trajectories of one 24-qubit noisy brickwork (Haar SU(2) +
    CX, DEPOLARIZE1 after every gate, M+R on four qubits after layer 5)
    share a program, so steps where exactly one error was drawn run the
    one-error op lists.  Every trajectory equals the oracle run with its own
    draws."""
    monkeypatch.setenv("NCSV_BATCH_DUMP", "1")
    n = 24
    text = W.noisy_brickwork(n, 6, 5, p=3e-3, every=5)
    ps = [Pauli(n, 0, 1 << i, 0) for i in range(0, n, 3)] + [Pauli(n, 0, 3 << i, 0) for i in range(0, n - 1, 5)]
    b = H.HipBatch(precision)
    c = H.Circuit.parse(text)
    for _ in range(16):
        b.add(c, ps)
    us = _draws(b, 29)
    capfd.readouterr()
    b.run(uniforms=us)
    stats = [line for line in capfd.readouterr().err.splitlines() if line.startswith("records=")]
    assert stats and int(stats[-1].split("one-error=")[1]) > 0, stats
    tol = TOL if precision == H.C128 else 2e-5
    for i in range(16):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(text, ps, us[i])
        assert oc == roc, i
        assert abs(nn - rnn) < tol
        assert max(abs(a - r) for a, r in zip(ex, rex)) < tol, i


def test_noisy_trajectories_full_size_complex64_against_complex128():
    """The 24-qubit depth-40 trajectories (M+R every 10 layers): 48 complex64
    trajectories from SplitMix64 seeds, then the same draws in complex128
    with the complex64 outcomes forced; the normalized <Z_i> and
    <Z_i Z_i+1> estimates agree within 1e-5."""
    n = 24
    c = H.Circuit.parse(W.noisy_brickwork(n, 40, 5))
    ps = [H.Pauli(n, 0, 1 << i, 0) for i in range(n)] + [H.Pauli(n, 0, 3 << i, 0) for i in range(n - 1)]
    T = 48
    seeds = np.arange(T, dtype=np.uint64) + 1000
    b64 = H.HipBatch(H.C64)
    b128 = H.HipBatch(H.C128)
    for _ in range(T):
        b64.add(c, ps)
        b128.add(c, ps)
    b64.run(seeds=seeds)
    forced = []
    for i in range(T):
        forced.append(np.array([0xff if o is None else o for o in b64.result(i)[2]], dtype=np.uint8))
    b128.run(seeds=seeds, forced=forced)
    diff = 0.0
    for i in range(T):
        e64, n64, o64 = b64.result(i)
        e128, n128, o128 = b128.result(i)
        assert o64 == o128
        diff = max(diff, max(abs(a.real / n64 - c_.real / n128) for a, c_ in zip(e64, e128)))
    print(f"trajectories: max |estimate64 - estimate128| = {diff:.2e}", flush=True)
    assert diff < 1e-5
