"""FP32 MFMA dense blocks in the complex64 register sweeps, on the GPU.

A random two-qubit circuit (Haar 4x4 unitaries on alternating neighbour
pairs) takes the blocks by cost; forcing them (NCSV_JIT_MFMA=2) on the
rotation brickwork composes real-form 2x2s, conditioned phases and flips into
blocks.  Each is checked against the oracle (complex128, FP64) within the
complex64 tolerance 2e-5, and at 30 qubits the MFMA plan against the
vector-only plan and against complex128 (fidelity loss <= 1e-5)."""
import glob
import os
import tempfile

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import workloads as W
from oracle.circuit import Circuit as OCircuit
from oracle.gates import Gate as OGate
from oracle.statevector import SplitMix64, StateVector

pytestmark = pytest.mark.gpu

TOL64 = 2e-5


def _blocks(dump):
    return sum(open(f).read().count("MFMA block") for f in glob.glob(os.path.join(dump, "*.hip")))


def _run(n, gates_or_text, prec, mfma, monkeypatch):
    """Runs the circuit on a fresh state with NCSV_JIT_MFMA=mfma; returns the
    state and the MFMA blocks its kernels hold."""
    dump = tempfile.mkdtemp()
    monkeypatch.setenv("NCSV_JIT_MFMA", str(mfma))
    monkeypatch.setenv("NCSV_JIT_DUMP", dump)
    if isinstance(gates_or_text, str):
        c = H.Circuit.parse(gates_or_text)
    else:
        c = W.build_circuit(n, gates_or_text, H.Circuit, H.Gate)
    s = H.HipStateVector.zero_state(n, prec)
    s.run(c)
    s.synchronize()
    monkeypatch.delenv("NCSV_JIT_DUMP")
    blocks = _blocks(dump)
    for f in glob.glob(os.path.join(dump, "*")):
        os.remove(f)
    os.rmdir(dump)
    return s, blocks


def test_su4_brickwork_complex64_mfma_matches_the_oracle(monkeypatch):
    n, gates = 22, W.su4_brickwork(22, 8, 11)
    print(f"su4 brickwork {n} q: oracle", flush=True)
    o = StateVector.zero_state(n)
    o.run(W.build_circuit(n, gates, OCircuit, OGate), SplitMix64(0))
    d, blocks = _run(n, gates, H.C64, 1, monkeypatch)
    err = float(np.abs(o.to_vector() - d.to_vector()).max())
    print(f"su4 brickwork {n} q complex64: {blocks} MFMA blocks, max|dpsi| = {err:.2e}", flush=True)
    assert blocks > 0
    assert err < TOL64


def test_forced_mfma_blocks_on_the_rotation_brickwork(monkeypatch):
    n = 22
    text = W.rotation_brickwork(n, 6, 5)
    o = StateVector.zero_state(n)
    o.run(OCircuit.parse(text), SplitMix64(0))
    d, blocks = _run(n, text, H.C64, 2, monkeypatch)
    err = float(np.abs(o.to_vector() - d.to_vector()).max())
    print(f"rotation brickwork {n} q complex64, forced blocks: {blocks} MFMA blocks, max|dpsi| = {err:.2e}", flush=True)
    assert blocks > 0
    assert err < TOL64


def test_su4_brickwork_30_qubits_mfma_against_vector_and_complex128(monkeypatch):
    n, gates = 30, W.su4_brickwork(30, 12, 12)
    a, blocks = _run(n, gates, H.C64, 1, monkeypatch)
    b, none = _run(n, gates, H.C64, 0, monkeypatch)
    assert blocks > 0 and none == 0
    ab = a.compare(b)
    del b
    c, _ = _run(n, gates, H.C128, 1, monkeypatch)
    ac = a.compare(c)
    print(f"su4 brickwork {n} q: {blocks} MFMA blocks; against the vector plan fidelity loss "
          f"{1 - ab['fidelity']:.2e} max-abs {ab['max_abs']:.2e}; against complex128 fidelity loss "
          f"{1 - ac['fidelity']:.2e} max-abs {ac['max_abs']:.2e}", flush=True)
    assert 1 - ac["fidelity"] <= 1e-5 and ac["max_abs"] < 1e-5
    assert 1 - ab["fidelity"] <= 1e-5 and ab["max_abs"] < 1e-5
