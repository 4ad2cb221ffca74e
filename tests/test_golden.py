"""Known answers from closed forms (tests/golden/kat.json, written by
tests/golden/make_golden.py): the oracle, the exact oracle, the batched
form's emulator and the HIP paths reproduce them.  The 3-qubit circuits
also run embedded in 14-qubit registers on spread-out qubits, so the fused
tile sweeps (interpreter and specialized kernels) see them too."""
import json
import os
import re

import numpy as np
import pytest

import ncsim_amd as nc
from oracle.circuit import Circuit as OCircuit
from oracle.exact import ExactState
from oracle.pauli import Pauli as OPauli
from oracle.statevector import SplitMix64, StateVector

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kat.json")))
IDS = [k["name"] for k in KATS]


def amps(k):
    return np.array([complex(a, b) for a, b in k["amplitudes"]])


def embed(text, n, qmap):
    """The circuit text on n qubits, qubit q -> qmap[q]."""
    out = [f"QUBITS {n}"]
    for line in text.strip().splitlines()[1:]:
        head, *args = line.split()
        out.append(" ".join([head] + [str(qmap[int(a)]) for a in args]))
    return "\n".join(out) + "\n"


def embed_pauli(s, qmap):
    return re.sub(r"([XYZ])(\d+)", lambda m: f"{m.group(1)}{qmap[int(m.group(2))]}", s)


def index_map(n, qmap):
    return [sum(((x >> q) & 1) << qmap[q] for q in range(3)) for x in range(8)]


@pytest.mark.parametrize("k", KATS, ids=IDS)
def test_oracle_reproduces_known_answers(k):
    c = OCircuit.parse(k["text"])
    v = StateVector.zero_state(3)
    v.run(c, SplitMix64(0))
    assert np.abs(v.to_vector() - amps(k)).max() < 1e-12
    for p, re_, im_ in k["expectations"]:
        assert abs(v.expectation(OPauli.parse(p, 3)) - complex(re_, im_)) < 1e-12


@pytest.mark.parametrize("k", [k for k in KATS if "RX" not in k["text"]], ids=lambda k: k["name"])
def test_exact_oracle_reproduces_known_answers(k):
    e = ExactState(3).run(OCircuit.parse(k["text"]))
    assert np.abs(np.array(e.to_vector()) - amps(k)).max() < 1e-14


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_emulated_batch_reproduces_known_expectations(precision):
    b = nc.HipBatch(precision)
    ks = [k for k in KATS if k["expectations"]]
    for k in ks:
        ps = [OPauli.parse(p, 3) for p, _, _ in k["expectations"]]
        b.add(nc.Circuit.parse(k["text"]), [nc.Pauli(3, p.x, p.z, p.phase) for p in ps])
    b.emulate(seeds=list(range(len(ks))))
    for i, k in enumerate(ks):
        ex, nn, _ = b.result(i)
        assert abs(nn - 1.0) < 1e-12
        for got, (_, re_, im_) in zip(ex, k["expectations"]):
            assert abs(got - complex(re_, im_)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("fusion", [0, 1, 2])
@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
@pytest.mark.parametrize("k", KATS, ids=IDS)
def test_gpu_reproduces_known_answers(k, precision, fusion):
    tol = 1e-12 if precision == nc.C128 else 2e-6
    for n, qmap in [(3, [0, 1, 2]), (14, [0, 5, 13]), (14, [13, 2, 7])]:
        s = nc.HipStateVector.zero_state(n, precision)
        s.set_fusion(fusion)
        s.run(nc.Circuit.parse(embed(k["text"], n, qmap)))
        v = s.to_vector()
        want = np.zeros(1 << n, dtype=complex)
        want[index_map(n, qmap)] = amps(k)
        assert np.abs(v - want).max() < tol, (n, qmap)
        for p, re_, im_ in k["expectations"]:
            o = OPauli.parse(embed_pauli(p, qmap), n)
            assert abs(s.expectation(nc.Pauli(n, o.x, o.z, o.phase)) - complex(re_, im_)) < tol, (p, qmap)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_gpu_batch_reproduces_known_expectations(precision):
    tol = 1e-12 if precision == nc.C128 else 2e-6
    b = nc.HipBatch(precision)
    ks = [k for k in KATS if k["expectations"]]
    for k in ks:
        for n, qmap in [(3, [0, 1, 2]), (16, [2, 9, 15])]:
            ps = [OPauli.parse(embed_pauli(p, qmap), n) for p, _, _ in k["expectations"]]
            b.add(nc.Circuit.parse(embed(k["text"], n, qmap)), [nc.Pauli(n, p.x, p.z, p.phase) for p in ps])
    b.run(seeds=list(range(len(b))))
    i = 0
    for k in ks:
        for _ in range(2):
            ex, nn, _ = b.result(i)
            i += 1
            assert abs(nn - 1.0) < tol
            for got, (_, re_, im_) in zip(ex, k["expectations"]):
                assert abs(got - complex(re_, im_)) < tol
