"""The frame sampler's C++ side on the CPU: its tableau's reference records
and the kernel's own per-word code run on the host (ncsv_frames_emulate_packed)
equal the oracle (oracle/frames.py) bit for bit, on random noisy Clifford
circuits with resets, Pauli-product measurements and qubits beyond 64; and a
non-Clifford gate is Unsupported (frame.rs:42-52)."""
import numpy as np
import pytest

import ncsim_amd as H
from oracle import frames as F
from oracle.circuit import Circuit
from ncsim_amd.workloads import repetition_code
from tests.test_frames_oracle import random_clifford_circuit


def noisy_clifford(n, depth, seed, wide=0):
    """random_clifford_circuit with noise channels and, optionally, `wide`
    extra qubits above 64 carrying gates and noise (measurements stay below
    64, as the text format requires)."""
    r = np.random.default_rng(seed + 1000)
    lines = random_clifford_circuit(n, depth, seed).strip().splitlines()
    out = [f"QUBITS {n + wide if wide == 0 else 64 + wide}"]
    for line in lines[1:]:
        if wide and line.startswith("MPP"):
            # a Pauli product on a register over 64 qubits is outside the text format
            line = "M " + line.split()[1].split("*")[0][1:]
        out.append(line)
        x = r.random()
        if x < 0.2:
            out.append(f"DEPOLARIZE1({r.uniform(0, 0.3)!r}) {r.integers(n)}")
        elif x < 0.3:
            a, b = r.choice(n, 2, replace=False)
            out.append(f"DEPOLARIZE2({r.uniform(0, 0.3)!r}) {a} {b}")
        elif x < 0.35:
            out.append(f"PAULI_CHANNEL_1(0.05,0.1,0.15) {r.integers(n)}")
        elif x < 0.4:
            out.append(f"X_ERROR(1.0) {r.integers(n)}")
        if wide and x < 0.5:
            q = 64 + int(r.integers(wide))
            out.append(f"CX {r.integers(n)} {q}")
            out.append(f"H {q}")
            out.append(f"Z_ERROR(0.2) {q}")
    if wide:
        out.append(f"CX {64} {0}")
    return "\n".join(out) + "\n"


CASES = [(noisy_clifford(n, d, s, w), s) for n, d, s, w in
         [(3, 6, 1, 0), (5, 10, 2, 0), (8, 12, 3, 0), (12, 8, 4, 0), (6, 8, 5, 40)]]


@pytest.mark.parametrize("text,seed", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_emulated_kernel_equals_the_oracle(text, seed):
    hs = H.HipFrameSampler(H.Circuit.parse(text), seed)
    fs = F.FrameSampler(Circuit.parse(text), seed)
    assert hs.reference() == fs.reference
    for shots in (1, 64, 77, 640):
        got = hs.emulate_packed(shots, 100 + seed)
        want = fs.sample_packed(shots, 100 + seed)
        assert len(got) == len(want)
        for k, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g, w), (shots, k)


def test_shape_and_reference_of_a_repetition_code():
    c = repetition_code(5, 3, 0.02)
    hs = H.HipFrameSampler(H.Circuit.parse(c), 5)
    assert (hs.num_measurements, hs.num_detectors, hs.num_observables) == (3 * 4 + 5, 12, 1)
    assert hs.reference() == F.FrameSampler(Circuit.parse(c), 5).reference


def test_non_clifford_gates_are_unsupported():
    for text in ("QUBITS 1\nT 0\nM 0\n", "QUBITS 2\nCS 0 1\n", "QUBITS 1\nRZ(0.5) 0\n", "QUBITS 1\nPHASE(1,3) 0\n",
                 "QUBITS 2\nCPHASE(1,2) 0 1\n", "QUBITS 3\nCCX 0 1 2\n"):
        with pytest.raises(H.NcsvError) as e:
            H.HipFrameSampler(H.Circuit.parse(text))
        assert e.value.kind == "Unsupported"
    H.HipFrameSampler(H.Circuit.parse("QUBITS 2\nPHASE(1,2) 0\nPHASE(4,3) 1\nCPHASE(1,1) 0 1\nM 0\n"))
