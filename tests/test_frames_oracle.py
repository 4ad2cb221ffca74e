"""Pins the Pauli-frame oracle (oracle/frames.py) on the CPU:

- its record histograms against trajectories of the dense state-vector
  oracle with sampled noise (the reference's frame tests, frame.rs:362-399,
  compare with stabilizer trajectories; the dense oracle is pinned
  independently by test_oracle.py);
- the tableau's reference records against the dense oracle: every recorded
  outcome has positive probability and deterministic ones match;
- closed forms of the reference's repetition-code tests (frame.rs:401-462):
  first-round detector rate 2p(1-p), observable rate (1-(1-2p)^r)/2,
  majority-vote failure 3p^2(1-p) + p^3, and silent detectors without noise;
- the counter-based draws: every site and shot independent and uniform
  enough that a Bernoulli(p) hit rate matches p.
"""
import math

import numpy as np
import pytest

from oracle import frames as F
from oracle.circuit import Circuit
from oracle.pauli import Pauli, SimError
from oracle.statevector import SplitMix64, StateVector, lib
from ncsim_amd.workloads import repetition_code


def histogram(records, m):
    h = np.zeros(1 << m)
    for r in records:
        h[sum(int(b) << k for k, b in enumerate(r))] += 1
    return h / len(records)


def trajectories(text, shots, seed):
    lib().orc_set_num_threads(1)
    c = Circuit.parse(text)
    rng = SplitMix64(seed)
    out = []
    for _ in range(shots):
        v = StateVector.zero_state(c.n)
        out.append(v.run(c, rng))
    return out


@pytest.mark.parametrize("text,shots,tol", [
    ("QUBITS 3\nH 0\nCX 0 1\nCX 1 2\nM 0 1 2\nMX 0\nM 2\n", 12000, 0.035),
    ("QUBITS 2\nH 0\nS 0\nCX 0 1\nMY 0\nMPP X0*X1 Z0*Z1\nR 0\nM 0\nH 1\nM 1\n", 12000, 0.035),
    ("QUBITS 3\nH 0\nDEPOLARIZE1(0.1) 0\nCX 0 1\nDEPOLARIZE2(0.2) 0 1\nCX 1 2\nX_ERROR(0.15) 2\nM 0 1 2\n"
     "MPP Z0*Z1\n", 20000, 0.04),
    ("QUBITS 2\nH 0\nCX 0 1\nPAULI_CHANNEL_1(0.05,0.1,0.15) 1\nSQRT_X 0\nCY 0 1\nMX 0 1\nZ_ERROR(0.3) 0\nR 1\n"
     "MX 0\nM 1\n", 20000, 0.04),
    ("QUBITS 3\nH 0 1\nSQRT_Y 2\nCZ 0 2\nSQRT_X_DAG 1\nS_DAG 0\nSWAP 1 2\nPHASE(3,2) 2\nCPHASE(1,1) 0 1\n"
     "SQRT_Y_DAG 0\nY_ERROR(0.1) 1\nMPP Y0*X1 Z2\nM 0 1\n", 12000, 0.04),
], ids=["ghz", "bell-mpp-reset", "noisy", "pauli-channel", "clifford-mix"])
def test_records_match_dense_trajectories(text, shots, tol):
    """frame.rs:362-399: total variation between the frame sampler's record
    histogram and the dense oracle's trajectories."""
    c = Circuit.parse(text)
    m = c.num_measurements()
    fs = F.FrameSampler(c, 7)
    hf = histogram(fs.sample(shots, 11), m)
    ht = histogram(trajectories(text, shots, 5), m)
    tv = 0.5 * np.abs(hf - ht).sum()
    assert tv < tol, (tv, hf, ht)


def random_clifford_circuit(n, depth, seed):
    r = np.random.default_rng(seed)
    one = ["H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG", "X", "Y", "Z", "I"]
    two = ["CX", "CY", "CZ", "SWAP"]
    lines = [f"QUBITS {n}"]
    for _ in range(depth):
        for q in range(n):
            lines.append(f"{one[r.integers(len(one))]} {q}")
        perm = r.permutation(n)
        for i in range(0, n - 1, 2):
            lines.append(f"{two[r.integers(len(two))]} {perm[i]} {perm[i + 1]}")
        k = r.random()
        if k < 0.3:
            lines.append(f"M {r.integers(n)}")
        elif k < 0.5:
            a, b = r.choice(n, 2, replace=False)
            lines.append(f"MPP {'XYZ'[r.integers(3)]}{a}*{'XYZ'[r.integers(3)]}{b}")
        elif k < 0.6:
            lines.append(f"R {r.integers(n)}")
    lines.append("M " + " ".join(str(q) for q in range(n)))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("seed", range(6))
def test_tableau_records_are_possible_outcomes(seed):
    """The tableau's record forced onto the dense oracle: every outcome has
    positive weight, and outcomes the dense state makes certain agree."""
    text = random_clifford_circuit(5, 8, seed)
    c = Circuit.parse(text)
    rec = F.reference_record(c, seed)
    v = StateVector.zero_state(c.n)
    k = 0
    for op in c.ops:
        if op[0] == "gate":
            v.apply_gate(op[1], op[2])
        elif op[0] in ("measure", "reset"):
            p = op[1] if op[0] == "measure" else Pauli.single(c.n, op[1], "Z")
            nn = v.norm_sq().real
            e = v.expectation(p).real / nn
            o = rec[k] if op[0] == "measure" else None
            if op[0] == "reset":
                # the tableau's reset measures Z and flips a 1: both branches possible here
                o = 0 if (1 + e) / 2 > 0.5 else 1
            w = v.project(p, bool(o)).real / nn
            assert w > 1e-9, (seed, k, e, o)
            if abs(e) > 1 - 1e-9:
                assert o == (0 if e > 0 else 1)
            if op[0] == "reset" and o:
                from oracle.gates import Gate
                v.apply_gate(Gate("X"), [op[1]])
            if op[0] == "measure":
                k += 1


def test_repetition_code_detectors_and_observable():
    """frame.rs:401-442 at d = 5, 3 rounds, p = 0.02."""
    d, rounds, p = 5, 3, 0.02
    c = Circuit.parse(repetition_code(d, rounds, p))
    quiet = F.FrameSampler(c.without_noise(), 5)
    dq, oq = quiet.sample_detectors(1000, 1)
    assert all(not b.any() for b in dq + oq)
    fs = F.FrameSampler(c, 5)
    shots = 50_000
    det, obs = fs.sample_detectors(shots, 2)
    pop = lambda b: sum(bin(int(w)).count("1") for w in b)  # noqa: E731
    rate = pop(det[0]) / shots
    want = 2 * p * (1 - p)
    assert abs(rate - want) < 5 * math.sqrt(want / shots)
    q = 1 - (1 - 2 * p) ** rounds
    orate = pop(obs[0]) / shots
    assert abs(orate - q / 2) < 5 * math.sqrt(q / 2 / shots)


def test_repetition_code_logical_error_rate():
    """frame.rs:444-462: majority vote of three bit-flipped qubits."""
    p = 0.1
    c = Circuit.parse(f"QUBITS 3\nX_ERROR({p}) 0 1 2\nM 0 1 2\n")
    fs = F.FrameSampler(c, 0)
    shots = 100_000
    packed = fs.sample_packed(shots, 3)
    bits = np.array([[(int(packed[k][s // 64]) >> (s % 64)) & 1 for s in range(shots)] for k in range(3)])
    fails = (bits.sum(axis=0) >= 2).mean()
    want = 3 * p * p * (1 - p) + p ** 3
    assert abs(fails - want) < 5 * math.sqrt(want / shots)
    par = F.parity(packed, [0, 1])
    rate = sum(bin(int(w)).count("1") for w in par) / shots
    assert abs(rate - 2 * p * (1 - p)) < 0.01


def test_counter_draws():
    """rnd(seed, a, b): a hit rate of p per shot, distinct sites and seeds
    independent (their coincidence rate is that of independent bits)."""
    s = np.arange(200_000, dtype=np.uint64)
    u = F.rnd(42, F.site(F.SITE_HIT, 3), s)
    for p in (0.5, 0.1, 1e-3):
        rate = (u < np.uint64(int(p * 2.0 ** 64))).mean()
        assert abs(rate - p) < 5 * math.sqrt(p / len(s))
    v = F.rnd(42, F.site(F.SITE_HIT, 4), s)
    w = F.rnd(43, F.site(F.SITE_HIT, 3), s)
    for a, b in ((u, v), (u, w)):
        same = ((a >> np.uint64(63)) == (b >> np.uint64(63))).mean()
        assert abs(same - 0.5) < 0.01
    assert F.mix(0) == 0xE220A8397B1DCDAF  # SplitMix64's first output of seed 0


def test_non_clifford_is_unsupported():
    """frame.rs:42-52 and gate.rs:658-666."""
    for text in ("QUBITS 1\nT 0\nM 0\n", "QUBITS 2\nCS 0 1\n", "QUBITS 1\nRX(0.5) 0\n", "QUBITS 1\nPHASE(1,3) 0\n",
                 "QUBITS 2\nCPHASE(1,2) 0 1\n", "QUBITS 3\nCCZ 0 1 2\n"):
        with pytest.raises(SimError) as e:
            F.FrameSampler(Circuit.parse(text), 0)
        assert e.value.kind == "Unsupported"
    for text in ("QUBITS 1\nPHASE(1,2) 0\nPHASE(2,2) 0\nPHASE(4,3) 0\nM 0\n", "QUBITS 2\nCPHASE(1,1) 0 1\nM 0\n"):
        F.FrameSampler(Circuit.parse(text), 0)
