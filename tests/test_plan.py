"""CPU checks of the fusion planner: the fused tile schedule that ncsv_run
executes, emulated on the host with the device ops' semantics, equals the
oracle's gate-by-gate result; and it needs far fewer sweeps than gates."""
import ctypes

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import _lib as L
from ncsim_amd import workloads as W
from oracle.circuit import Circuit as OCircuit
from oracle.statevector import SplitMix64, StateVector


def emulate(text, prec, amp=None):
    c = H.Circuit.parse(text)
    stats = (ctypes.c_uint64 * 4)()
    ptr = None
    if amp is not None:
        ptr = amp.view(np.float64).ctypes.data_as(L.c_dp)
    H._check(L.lib().ncsv_plan_emulate(c._h, prec, ptr, stats))
    return list(stats)


def oracle_run(text, amp):
    o = StateVector(OCircuit.parse(text).n, amp.copy())
    o.run(OCircuit.parse(text), SplitMix64(0))
    return o.to_vector()


MIXED = """QUBITS 14
H 0 5 13 9
CX 0 13 12 1
CY 3 7
SWAP 2 11
CCX 4 8 12
CCZ 1 2 13
CS 6 10
CS_DAG 9 0
CPHASE(3,5) 13 2
PHASE(-7,6) 11
T 12
T_DAG 3
S 4
S_DAG 5
Z 6
X 7
Y 8
SQRT_X 9
SQRT_X_DAG 10
SQRT_Y 11
SQRT_Y_DAG 13
RX(0.3) 12
RY(1.2) 0
RZ(-2.2) 7
CZ 5 6
SWAP 0 1
CX 13 0
Y 13
CY 13 12
CCX 13 12 0
"""


@pytest.mark.parametrize("prec", [H.C128, H.C64])
@pytest.mark.parametrize("name,text", [
    ("mixed", MIXED + MIXED.split("\n", 1)[1] * 3),
    ("brickwork", W.rotation_brickwork(15, 12, 3)),
    ("layers", W.random_layers(14, 20, 5, measure=False)),
    ("qft", W.qft_with_rz(14, 40, 1)),
])
def test_fused_plan_equals_the_oracle(name, text, prec):
    n = OCircuit.parse(text).n
    rng = np.random.default_rng(1)
    amp = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
    amp /= np.linalg.norm(amp)
    want = oracle_run(text, amp)
    got = amp.copy()
    stats = emulate(text, prec, got)
    assert stats[3] == 0
    assert np.abs(got - want).max() < 1e-12, stats


def test_fusion_cuts_sweeps():
    text = W.rotation_brickwork(30, 24, 3)
    st = emulate(text, H.C128)
    gates = W.count_gates(text)
    assert gates == 1068
    assert st[0] <= 20, st
    st64 = emulate(text, H.C64)
    assert st64[0] <= st[0], (st, st64)


@pytest.mark.parametrize("prec", [H.C128, H.C64])
@pytest.mark.parametrize("name,text", [("mixed", MIXED), ("brickwork", W.rotation_brickwork(16, 6, 3))])
def test_specialized_sweep_kernels_compile(name, text, prec):
    """Every tile sweep of the plan, written out as a specialized kernel,
    compiles for gfx950 (hiprtc, no device needed)."""
    c = H.Circuit.parse(text)
    k = ctypes.c_uint32()
    H._check(L.lib().ncsv_jit_check(c._h, prec, b"gfx950", ctypes.byref(k)))
    assert k.value >= 1


@pytest.mark.parametrize("prec", [H.C128, H.C64])
def test_specialized_expectation_kernels_compile(prec):
    n = 30
    ps = W.random_paulis(n, 64, 4)
    k = len(ps)
    xs = (ctypes.c_uint64 * k)(*[p[0] for p in ps])
    zs = (ctypes.c_uint64 * k)(*[p[1] for p in ps])
    ph = (ctypes.c_uint32 * k)(*[p[2] for p in ps])
    g = ctypes.c_uint32()
    H._check(L.lib().ncsv_jit_check_expect(n, prec, k, xs, zs, ph, b"gfx950", ctypes.byref(g)))
    assert 1 <= g.value <= 12


@pytest.mark.parametrize("prec", [H.C128, H.C64])
@pytest.mark.parametrize("family", ["brickwork", "mixed"])
def test_multistart_plan_at_20_qubits_equals_the_oracle(prec, family):
    """From 20 qubits on the planner keeps, of the greedy tile choice and 23
    permuted variants, the plan with the fewest sweeps; that plan applied in
    FP64 on the host equals the oracle from a random state (30-qubit
    rotation brickwork depth 24: 12 sweeps instead of 13 in complex128, 11
    in complex64)."""
    n = 20
    if family == "brickwork":
        text = W.rotation_brickwork(n, 8, 5)
    else:
        rng = np.random.default_rng(6)
        lines = [f"QUBITS {n}"]
        for _ in range(6):
            for q in range(n):
                lines.append(f"{['H', 'T', 'SQRT_X', 'S'][int(rng.integers(4))]} {q}")
            perm = rng.permutation(n)
            for i in range(0, n - 1, 2):
                lines.append(f"CX {perm[i]} {perm[i + 1]}")
            a, b, c = rng.choice(n, 3, replace=False)
            lines.append(f"CCZ {a} {b} {c}")
        text = "\n".join(lines) + "\n"
    rng = np.random.default_rng(7)
    amp = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    amp /= np.linalg.norm(amp)
    want = oracle_run(text, amp)
    got = amp.copy()
    stats = emulate(text, prec, got)
    assert stats[3] == 0 and stats[0] >= 1
    assert np.abs(got - want).max() < 1e-12


def test_multistart_plan_saves_sweeps_on_the_30_qubit_brickwork():
    text = W.rotation_brickwork(30, 24, 3)
    assert emulate(text, H.C128)[0] == 12
    assert emulate(text, H.C64)[0] == 11


def test_balanced_plan_equals_the_oracle_and_keeps_its_sweeps(monkeypatch):
    """The planner evens out its sweeps' arithmetic by moving gates that
    commute with the rest of a heavy sweep into the next one (22-qubit
    rotation brickwork, depth 24: the heaviest sweep's 2x2s go from 65 to
    61): the same number of sweeps, and the balanced plan applied in FP64 on
    the host equals the oracle from a random state."""
    n = 22
    text = W.rotation_brickwork(n, 24, 3)
    monkeypatch.setenv("NCSV_PLAN_BALANCE", "0")
    plain = emulate(text, H.C128)[0]
    monkeypatch.setenv("NCSV_PLAN_BALANCE", "1")
    rng = np.random.default_rng(8)
    amp = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    amp /= np.linalg.norm(amp)
    want = oracle_run(text, amp)
    got = amp.copy()
    stats = emulate(text, H.C128, got)
    assert stats[0] == plain
    assert np.abs(got - want).max() < 1e-12
