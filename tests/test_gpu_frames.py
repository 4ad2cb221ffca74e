"""The frame sampler's kernels on the GPU against the oracle
(oracle/frames.py): packed records and detectors bit for bit on random
noisy Clifford circuits (shots not a multiple of 64, qubits beyond 64), and
at a million shots the reference's repetition-code statistics
(frame.rs:401-462) and silent detectors without noise."""
import math

import numpy as np
import pytest

import ncsim_amd as H
from oracle import frames as F
from oracle.circuit import Circuit
from tests.test_frames_host import CASES
from ncsim_amd.workloads import repetition_code

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("text,seed", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_packed_records_equal_the_oracle(text, seed):
    hs = H.HipFrameSampler(H.Circuit.parse(text), seed)
    fs = F.FrameSampler(Circuit.parse(text), seed)
    for shots in (1, 64, 1000, 4097):
        got = hs.sample_packed(shots, 7 + seed)
        want = fs.sample_packed(shots, 7 + seed)
        for k, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g, w), (shots, k)


def test_detectors_equal_the_oracle():
    text = repetition_code(5, 3, 0.05)
    hs = H.HipFrameSampler(H.Circuit.parse(text), 5)
    fs = F.FrameSampler(Circuit.parse(text), 5)
    for shots in (100, 5000):
        gd, go = hs.sample_detectors(shots, 3)
        wd, wo = fs.sample_detectors(shots, 3)
        assert all(np.array_equal(a, b) for a, b in zip(gd + go, wd + wo))


def test_large_downloads_through_pinned_staging(monkeypatch):
    """Records and detectors of 144-162 MiB come back through the pinned
    staging buffers bit-identical to one plain copy of each."""
    text = repetition_code(9, 9, 0.01)
    hs = H.HipFrameSampler(H.Circuit.parse(text), 3)
    shots = (1 << 24) - 5
    got = [hs.sample_packed(shots, 11), hs.sample_detectors(shots, 12)]
    monkeypatch.setenv("NCSV_FRAMES_STAGE", "0")
    want = [hs.sample_packed(shots, 11), hs.sample_detectors(shots, 12)]
    assert all(np.array_equal(a, b) for a, b in zip(got[0], want[0]))
    for g, w in zip(got[1], want[1]):
        assert all(np.array_equal(a, b) for a, b in zip(g, w))
    assert sum(len(a) for a in got[0]) * 8 >= 128 << 20


def popcount(words):
    return int(sum(bin(int(w)).count("1") for w in words))


def test_repetition_code_statistics_at_a_million_shots():
    """d = 5, 3 rounds, p = 0.02 (frame.rs:401-442) and the majority vote of
    three flipped qubits (444-462), at 2^20 shots."""
    shots = 1 << 20
    d, rounds, p = 5, 3, 0.02
    text = repetition_code(d, rounds, p)
    quiet = H.HipFrameSampler(H.Circuit.parse(
        "".join(line + "\n" for line in text.splitlines() if not line.startswith("X_ERROR"))), 5)
    dq, oq = quiet.sample_detectors(shots, 1)
    assert all(not b.any() for b in dq + oq)
    hs = H.HipFrameSampler(H.Circuit.parse(text), 5)
    det, obs = hs.sample_detectors(shots, 2)
    want = 2 * p * (1 - p)
    for k in range(d - 1):  # every first-round detector
        assert abs(popcount(det[k]) / shots - want) < 5 * math.sqrt(want / shots)
    q = 1 - (1 - 2 * p) ** rounds
    assert abs(popcount(obs[0]) / shots - q / 2) < 5 * math.sqrt(q / 2 / shots)
    p = 0.1
    maj = H.HipFrameSampler(H.Circuit.parse(f"QUBITS 3\nX_ERROR({p}) 0 1 2\nM 0 1 2\n"))
    a, b, c = maj.sample_packed(shots, 3)
    fails = popcount((a & b) | (a & c) | (b & c)) / shots
    want = 3 * p * p * (1 - p) + p ** 3
    assert abs(fails - want) < 5 * math.sqrt(want / shots)
