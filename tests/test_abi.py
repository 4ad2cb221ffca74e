"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side logic (text parser, printer, gate and Pauli
validation) follows the reference — no device calls."""
import ctypes
import os
import re

import pytest

import ncsim_amd
from ncsim_amd import _lib as L
from ncsim_amd import workloads as W
from oracle.circuit import Circuit as OCircuit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ncsv.h")).read()
    return sorted(set(re.findall(r"^(?:ncsv_status|size_t|int|void|uint32_t|ncsv_batch \*)\s*(ncsv_[a-z0-9_]+)\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = header_functions()
    assert len(declared) >= 45
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == L.exported_symbols()


def test_no_device_is_an_error_not_a_fallback():
    if ncsim_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(ncsim_amd.NcsvError) as e:
        ncsim_amd.HipStateVector.zero_state(3)
    assert e.value.kind == "Device"


TEXTS = [
    "H 0\nCX 0 1 1 2\nT 2\nPHASE(3,4) 1\nM 0\nMX 1\nMPP X0*Z2\nR 2\nTICK\n",
    "QUBITS 5\nRX(0.25) 0 1\nCPHASE(1,33) 3 4\nCCX 0 1 2\nDEPOLARIZE1(0.001) 2 3\nMY 4\nDETECTOR rec[-1]\n"
    "OBSERVABLE_INCLUDE(0) rec[-1]\nDEPOLARIZE2(0.01) 0 1\nPAULI_CHANNEL_1(0.1,0,0.2) 3\n",
    W.random_layers(6, 3, 1),
    W.rotation_brickwork(5, 3, 2),
    W.qft_with_rz(5, 4, 3),
]


@pytest.mark.parametrize("text", TEXTS)
def test_parser_agrees_with_the_oracle(text):
    """circuit.rs:459-728: the C++ parser and printer against the oracle's."""
    c = ncsim_amd.Circuit.parse(text)
    o = OCircuit.parse(text)
    assert c.num_qubits() == o.n
    assert c.num_measurements() == o.num_measurements()
    assert c.num_gates() == o.num_gates() == W.count_gates(text)
    # the printed form parses back to the same circuit on both sides
    printed = str(c)
    assert OCircuit.parse(printed).num_gates() == o.num_gates()
    c2 = ncsim_amd.Circuit.parse(printed)
    assert str(c2) == printed


@pytest.mark.parametrize("bad,kind", [("CX 0", "Parse"), ("QUBITS 1\nH 3", "Parse"), ("FOO 1", "Parse"),
                                      ("CX 1 1", "Parse"), ("MPP X0*X0", "Parse"), ("PHASE(1,0) 0", "Parse"),
                                      ("DETECTOR rec[-1]", "Parse"), ("X_ERROR(1.5) 0", "Parse")])
def test_parser_errors(bad, kind):
    with pytest.raises(ncsim_amd.NcsvError) as e:
        ncsim_amd.Circuit.parse(bad)
    assert e.value.kind == kind


def test_circuit_builder_checks():
    """Circuit::push checks (circuit.rs:177-229)."""
    c = ncsim_amd.Circuit(3)
    c.gate("H", [0]).gate("CX", [0, 1]).gate(ncsim_amd.Gate("RZ", theta=0.5), [2])
    with pytest.raises(ncsim_amd.NcsvError) as e:
        c.gate("CX", [0])
    assert e.value.kind == "Arity"
    with pytest.raises(ncsim_amd.NcsvError) as e:
        c.gate("H", [3])
    assert e.value.kind == "QubitOutOfRange"
    with pytest.raises(ncsim_amd.NcsvError) as e:
        c.gate("CZ", [1, 1])
    assert e.value.kind == "DuplicateQubit"
    with pytest.raises(ncsim_amd.NcsvError) as e:
        c.measure(ncsim_amd.Pauli(3, 1, 0, 1))  # i X is not Hermitian
    assert e.value.kind == "Invalid"
    c.measure_z(2).reset(1).noise("DEPOLARIZE1", [0.01], [0]).tick()
    assert c.num_gates() == 3 and c.num_measurements() == 1
    assert "RZ(0.5) 2" in str(c)


def test_out_of_device_memory_is_a_device_error(monkeypatch):
    """A state larger than NCSV_DEVICE_MEMORY_LIMIT_GB fails with the Device
    status and no handle (the path a full device takes), with or without a
    GPU; clone and tensor allocate through the same check."""
    monkeypatch.setenv("NCSV_DEVICE_MEMORY_LIMIT_GB", "0.001")
    h = ctypes.c_void_p()
    st = L.lib().ncsv_create(17, 0, ctypes.byref(h))  # 2 MiB of complex128
    assert st == 9 and not h.value
    buf = ctypes.create_string_buffer(256)
    L.lib().ncsv_last_error(buf, 256)
    assert b"out of device memory" in buf.value
    with pytest.raises(ncsim_amd.NcsvError) as e:
        ncsim_amd.HipStateVector.zero_state(20, ncsim_amd.C64)
    assert e.value.kind == "Device"
