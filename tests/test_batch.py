"""The batched form (many circuits or trajectories per launch) against the
oracle.  Each entry is the reference's run (mod.rs:168-183) from |0...0>
with one uniform per Measure/Reset/Noise op in program order, followed by
StateVector::expectation (vector.rs:412-429) of its Paulis.

CPU tests run the batch programs through the library's FP64 host emulator
(the same step/op lists the kernel runs), so the planner's tiles, reduce /
collapse steps and noise slots are checked without a device; GPU tests run
the persistent kernel."""
import math

import numpy as np
import pytest

import ncsim_amd as nc
from oracle import circuit as OC
from oracle import statevector as OS
from oracle.pauli import Pauli as OPauli

ONE_Q = ["H", "S", "T", "SQRT_X", "X", "Y", "Z", "S_DAG", "SQRT_Y"]
TWO_Q = ["CX", "CZ", "SWAP", "CY", "CS"]


def random_circuit(n, depth, seed, measure=True, noise=True):
    """Random circuit text: rotations, Clifford+T, CCZ/CCX, mid-circuit
    measurements (M, MPP of Z strings), resets and Pauli noise."""
    rng = np.random.default_rng(seed)
    out = [f"QUBITS {n}"]
    for _ in range(depth):
        for q in range(n):
            r = rng.random()
            if r < 0.25:
                out.append(f"{['RX', 'RY', 'RZ'][rng.integers(3)]}({rng.uniform(0, 2 * math.pi)!r}) {q}")
            else:
                out.append(f"{ONE_Q[rng.integers(len(ONE_Q))]} {q}")
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            if rng.random() < 0.6:
                out.append(f"{TWO_Q[rng.integers(len(TWO_Q))]} {perm[i]} {perm[i + 1]}")
        if n >= 3 and rng.random() < 0.3:
            a, b, c = rng.choice(n, 3, replace=False)
            out.append(f"{'CCZ' if rng.random() < 0.5 else 'CCX'} {a} {b} {c}")
        if noise and rng.random() < 0.5:
            q = int(rng.integers(n))
            out.append(f"DEPOLARIZE1(0.3) {q}")
            if rng.random() < 0.3:
                a, b = rng.choice(n, 2, replace=False)
                out.append(f"DEPOLARIZE2(0.2) {a} {b}")
        if measure and rng.random() < 0.3:
            r = rng.random()
            q = int(rng.integers(n))
            if r < 0.4:
                out.append(f"M {q}")
            elif r < 0.7:
                out.append(f"R {q}")
            else:
                a, b = rng.choice(n, 2, replace=False)
                out.append(f"MPP Z{a}*Z{b}")
    return "\n".join(out) + "\n"


def random_paulis(n, count, seed):
    rng = np.random.default_rng(seed + 7)
    ps = []
    for _ in range(count):
        w = int(rng.integers(1, min(4, n) + 1))
        sup = rng.choice(n, w, replace=False)
        x = z = 0
        for q in sup:
            k = int(rng.integers(1, 4))
            if k in (1, 2):
                x |= 1 << int(q)
            if k in (2, 3):
                z |= 1 << int(q)
        ps.append(OPauli(n, x, z, int(rng.integers(4))))
    return ps


class _One:
    def __init__(self, u):
        self.u = u

    def gen_f64(self):
        return self.u


def oracle_entry(text, paulis, us, forced=None):
    """The oracle run with the batched form's draw rule."""
    c = OC.Circuit.parse(text)
    sv = OS.StateVector.zero_state(c.n)
    k = 0
    outc = []
    for op in c.ops:
        if op[0] not in ("measure", "reset", "noise"):
            sv.apply_op(op, None)
            continue
        u = us[k]
        f = None if forced is None else forced[k]
        k += 1
        if op[0] == "noise":
            sv.apply_op(op, _One(u))
            outc.append(None)
            continue
        p = op[1] if op[0] == "measure" else OPauli.single(c.n, op[1], "Z")
        if f is not None and f != 0xff:
            sv.project(p, bool(f))
            b = bool(f)
        else:
            b = sv.measure(p, _One(u))
        if op[0] == "reset" and b:
            from oracle.gates import Gate
            sv.apply_gate(Gate("X"), [op[1]])
        outc.append(int(b))
    ex = [sv.expectation(p) for p in paulis]
    return ex, sv.norm_sq().real, outc


def _build(texts, paulis, precision, tile_bits):
    b = nc.HipBatch(precision)
    if tile_bits:
        b.set_tile_bits(tile_bits)
    circs = [nc.Circuit.parse(t) for t in texts]
    for c, ps in zip(circs, paulis):
        b.add(c, ps)
    return b


def _draws(b, seed):
    rng = np.random.default_rng(seed)
    return [rng.random(b.draws(i)) for i in range(len(b))]


CASES = [(6, 8, 1), (7, 10, 2), (9, 6, 3), (10, 9, 4), (11, 5, 5), (12, 4, 6)]


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
@pytest.mark.parametrize("tile_bits", [0, 6, 8])
def test_emulated_batch_matches_oracle(precision, tile_bits):
    """Whole-state tiles (tile_bits 0) and multi-tile programs (6, 8 tile
    qubits), with measurements, resets and noise, FP64 host emulation."""
    texts = [random_circuit(n, d, s) for n, d, s in CASES]
    paulis = [random_paulis(n, 5, s) for n, d, s in CASES]
    b = _build(texts, paulis, precision, tile_bits)
    us = _draws(b, 11)
    b.emulate(uniforms=us)
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc, f"entry {i}"
        assert abs(nn - rnn) < 1e-12
        for a, r in zip(ex, rex):
            assert abs(a - r) < 1e-12, (i, a, r)


def test_emulated_batch_seeds_and_forced():
    """Seeded draws equal the SplitMix64 stream; forced outcomes override."""
    n, d, s = 9, 8, 21
    text = random_circuit(n, d, s)
    ps = random_paulis(n, 3, s)
    b = _build([text, text], [ps, ps], nc.C128, 7)
    b.emulate(seeds=[5, 6])
    for i, seed in enumerate([5, 6]):
        g = OS.SplitMix64(seed)
        us = [g.gen_f64() for _ in range(b.draws(i))]
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(text, ps, us)
        assert oc == roc
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12
    # forced outcomes: other uniforms at the measurement slots, the outcomes
    # of the first run forced, give the first run's results again
    us = _draws(b, 3)
    b.emulate(uniforms=us)
    first = [b.result(i) for i in range(2)]
    rng = np.random.default_rng(8)
    us2, forced = [], []
    for i in range(2):
        oc = first[i][2]
        us2.append(np.array([u if o is None else rng.random() for u, o in zip(us[i], oc)]))
        forced.append(np.array([0xff if o is None else o for o in oc], dtype=np.uint8))
    assert any(o is not None for o in first[0][2])
    b.emulate(uniforms=us2, forced=forced)
    for i in range(2):
        ex, nn, oc = b.result(i)
        assert oc == first[i][2]
        assert ex == first[i][0] and nn == first[i][1]


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
@pytest.mark.parametrize("tile_bits", [0, 6])
def test_measure_reset_runs_match_oracle(precision, tile_bits):
    """Runs of M and R on several qubits (decided from one joint reduction)
    in noisy brickwork trajectories, emulated, against the oracle run with
    the same draws."""
    from ncsim_amd import workloads as W
    text = W.noisy_brickwork(8, 12, 4, p=0.05, every=4, mqubits=[0, 2, 5, 7])
    c = nc.Circuit.parse(text)
    ps = [OPauli(8, 0, 1 << i, 0) for i in range(8)] + [OPauli(8, 0, 3 << i, 0) for i in range(7)]
    b = nc.HipBatch(precision)
    if tile_bits:
        b.set_tile_bits(tile_bits)
    for _ in range(6):
        b.add(c, ps)
    b.emulate(seeds=list(range(6)))
    for i in range(6):
        g = OS.SplitMix64(i)
        us = [g.gen_f64() for _ in range(b.draws(i))]
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(text, ps, us)
        assert oc == roc, f"entry {i}"
        assert abs(nn - rnn) < 1e-12
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12


def test_batch_programs_share_and_count():
    """Entries running one circuit share its program; stats count them."""
    text = random_circuit(8, 5, 9, measure=True, noise=True)
    ps = random_paulis(8, 2, 9)
    b = nc.HipBatch(nc.C64)
    b.set_tile_bits(6)
    c = nc.Circuit.parse(text)
    for _ in range(5):
        b.add(c, ps)
    b.emulate(seeds=list(range(5)))
    st = b.stats()
    assert st["programs"] == 1
    assert st["amp_updates"] > 0
    for i in range(5):
        g = OS.SplitMix64(i)
        us = [g.gen_f64() for _ in range(b.draws(i))]
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(text, ps, us)
        assert oc == roc
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12


def _planned_kinds(capfd, b):
    capfd.readouterr()
    b.plan()
    err = capfd.readouterr().err
    return {k.split("=")[0] for line in err.splitlines() if line.startswith("step") for k in line.split(":")[1].split()}


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
@pytest.mark.parametrize("tile_bits", [0, 6, 8])
def test_dense_blocks_match_oracle(precision, tile_bits, capfd, monkeypatch):
    """Programs fuse ops on up to three tile positions into dense 8x8 blocks
    (U3, run on MFMA); the same programs without them (NCSV_BATCH_MFMA=0) and
    the oracle agree with them (FP64 emulation of the step lists)."""
    monkeypatch.setenv("NCSV_BATCH_DUMP", "1")
    texts = [random_circuit(n, d, s) for n, d, s in CASES]
    paulis = [random_paulis(n, 4, s) for n, d, s in CASES]
    b = _build(texts, paulis, precision, tile_bits)
    assert "U3" in _planned_kinds(capfd, b)
    us = _draws(b, 12)
    b.emulate(uniforms=us)
    monkeypatch.setenv("NCSV_BATCH_MFMA", "0")
    b0 = _build(texts, paulis, precision, tile_bits)
    assert "U3" not in _planned_kinds(capfd, b0)
    b0.emulate(uniforms=us)
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        ex0, nn0, oc0 = b0.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == oc0 == roc, f"entry {i}"
        assert abs(nn - rnn) < 1e-12
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12
        assert max(abs(a - r) for a, r in zip(ex0, rex)) < 1e-12


def test_dense_blocks_fuse_brickwork(capfd, monkeypatch):
    """A 16-qubit brickwork of SU(2)s and CXs on 13-qubit tiles: blocks cut
    the passes per step at least in half, and CXs controlled by qubits
    outside the tile join blocks as a second matrix for the tiles where the
    control is set."""
    from ncsim_amd import workloads as W
    monkeypatch.setenv("NCSV_BATCH_DUMP", "2")
    c = nc.Circuit.parse(W.noisy_brickwork(16, 12, 3, p=1e-3, every=100))

    def plan(env):
        monkeypatch.setenv("NCSV_BATCH_MFMA", env)
        b = nc.HipBatch(nc.C64)
        b.set_tile_bits(13)
        b.add(c, [])
        capfd.readouterr()
        b.plan()
        return capfd.readouterr().err.splitlines()

    def passes(lines):
        return sum(int(line.split("clean=")[1].split()[0]) for line in lines if line.startswith("step"))

    fused = plan("1")
    assert 2 * passes(fused) <= passes(plan("0"))
    assert any(line.strip().startswith("U3") and not line.endswith("o=0") for line in fused)


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_outer_phases_taken_at_the_store(precision, capfd, monkeypatch):
    """Phases on qubits outside the tile (here the next layer's first RZ on
    the outer qubits, after a measure-and-reset) become one phase per tile
    multiplied in at the store, without a pass of their own; a noisy 9-qubit
    brickwork on 6-qubit tiles, with errors drawn (p = 0.05), equals the
    oracle through the FP64 emulation and through the kernel's source on the
    host."""
    from ncsim_amd import workloads as W
    monkeypatch.setenv("NCSV_BATCH_DUMP", "2")
    text = W.noisy_brickwork(9, 8, 3, p=0.05, every=4)
    paulis = random_paulis(9, 6, 51)
    b = _build([text], [paulis], precision, 6)
    capfd.readouterr()
    b.plan()
    step, later = -1, 0
    for line in capfd.readouterr().err.splitlines():
        if line.startswith("step"):
            step = int(line.split()[1])
        elif line.strip().startswith("OPH") and step > 0:
            later += 1
    assert later >= 3
    tol = 1e-12 if precision == nc.C128 else 2e-5
    for seed in (52, 53):
        us = _draws(b, seed)
        rex, rnn, roc = oracle_entry(text, paulis, us[0])
        # the kernel's own source once (complex64), the emulation for every case
        runs = (b.emulate, b.simulate) if precision == nc.C64 and seed == 52 else (b.emulate,)
        for run in runs:
            run(uniforms=us)
            ex, nn, oc = b.result(0)
            assert oc == roc
            assert abs(nn - rnn) < tol
            assert max(abs(a - r) for a, r in zip(ex, rex)) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_gpu_outer_phases_taken_at_the_store(precision):
    """The per-tile phases of outer-only diagonals on the device: noisy
    brickworks on tiles smaller than the state against the oracle."""
    from ncsim_amd import workloads as W
    tol = 1e-12 if precision == nc.C128 else 2e-5
    for n, L in ((9, 6), (12, 8), (14, 10)):
        text = W.noisy_brickwork(n, 8, 3, p=0.05, every=4)
        paulis = random_paulis(n, 6, 51)
        b = _build([text], [paulis], precision, L)
        for seed in (52, 53, 54):
            us = _draws(b, seed)
            b.run(uniforms=us)
            ex, nn, oc = b.result(0)
            rex, rnn, roc = oracle_entry(text, paulis, us[0])
            assert oc == roc, (n, L, seed)
            assert abs(nn - rnn) < tol
            assert max(abs(a - r) for a, r in zip(ex, rex)) < tol, (n, L, seed)


@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_conditioned_one_position_runs(precision, capfd, monkeypatch):
    """Ops on one tile position under one condition on outer qubits (a CX
    whose control is outside the tile, then the target's next gate) become
    one 2x2 per outcome of the condition, in one pass: a 12-qubit circuit of
    the 10-22 qubit mix on 6-qubit tiles plans five of them, and the program
    equals the oracle through the FP64 emulation (the kernel runs them in
    test_gpu_conditioned_one_position_runs)."""
    from ncsim_amd import workloads as W
    monkeypatch.setenv("NCSV_BATCH_DUMP", "2")
    n, text, _ = W.batch_circuits(64, 11)[35]
    paulis = [nc.Pauli(n, 1 << q, z, 0) for q, z in ((0, 0), (3, 8), (5, 0), (7, 1), (9, 0), (11, 4))]
    paulis += [nc.Pauli(n, (1 << q) | (1 << (q + 1)), 1 << q, 1) for q in (1, 4, 6, 8, 10)]
    b = _build([text], [paulis], precision, 6)
    capfd.readouterr()
    b.plan()
    cond = [line for line in capfd.readouterr().err.splitlines()
            if line.startswith("  U1 ") and not line.endswith("o=0")]
    assert len(cond) >= 5
    us = _draws(b, 61)
    rex, rnn, roc = oracle_entry(text, paulis, us[0])
    assert max(abs(r) for r in rex) > 1e-3  # expectations that see the state
    tol = 1e-12 if precision == nc.C128 else 2e-5
    b.emulate(uniforms=us)
    ex, nn, oc = b.result(0)
    assert abs(nn - rnn) < tol
    assert max(abs(a - r) for a, r in zip(ex, rex)) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
def test_gpu_conditioned_one_position_runs(precision):
    """One-position runs under an outer condition on the device: circuits of
    the 10-22 qubit mix that plan them, on tiles smaller than the state,
    against the oracle."""
    from ncsim_amd import workloads as W
    specs = W.batch_circuits(64, 11)
    tol = 1e-12 if precision == nc.C128 else 2e-5
    for i, L in ((35, 6), (33, 7), (59, 6), (30, 6)):
        n, text, _ = specs[i]
        paulis = [nc.Pauli(n, 1 << q, 0, 0) for q in range(n)] + random_paulis(n, 6, 62)
        b = _build([text], [paulis], precision, L)
        us = _draws(b, 63)
        b.run(uniforms=us)
        ex, nn, oc = b.result(0)
        rex, rnn, roc = oracle_entry(text, paulis, us[0])
        assert abs(nn - rnn) < tol, (i, L)
        assert max(abs(a - r) for a, r in zip(ex, rex)) < tol, (i, L)


def test_one_error_lists_for_shared_noisy_programs(capfd, monkeypatch):
    """Sixteen trajectories of one noisy 8-qubit circuit (7-qubit tiles) share
    a program, which then
    also plans, per step, the op lists keeping one noise slot each; a step
    where exactly one error was drawn runs that list.  Every trajectory
    equals the oracle with its own draws (the kernel's source on the host,
    and the FP64 emulation)."""
    monkeypatch.setenv("NCSV_BATCH_DUMP", "1")
    text = random_circuit(8, 4, 71, measure=True, noise=False)
    lines = text.strip().splitlines()
    noisy = [lines[0]]
    for k, line in enumerate(lines[1:]):
        noisy.append(line)
        if k % 3 == 0 and not line.startswith(("M", "R", "MPP")):
            noisy.append(f"DEPOLARIZE1(0.03) {k % 8}")
    text = "\n".join(noisy) + "\n"
    paulis = random_paulis(8, 3, 72)
    b = nc.HipBatch(nc.C64)
    b.set_tile_bits(7)
    c = nc.Circuit.parse(text)
    for _ in range(16):
        b.add(c, paulis)
    capfd.readouterr()
    b.plan()
    assert any("one-error-lists=" in line and "one-error-lists=0" not in line
               for line in capfd.readouterr().err.splitlines())
    us = _draws(b, 73)
    for run in (b.simulate, b.emulate):
        run(uniforms=us)
        stats = [line for line in capfd.readouterr().err.splitlines() if line.startswith("records=")]
        assert stats and int(stats[-1].split("one-error=")[1]) > 0
        for i in range(16):
            ex, nn, oc = b.result(i)
            rex, rnn, roc = oracle_entry(text, paulis, us[i])
            assert oc == roc, i
            assert abs(nn - rnn) < 2e-5
            assert max(abs(a - r) for a, r in zip(ex, rex)) < 2e-5, i


@pytest.mark.parametrize("precision,threads,tile_bits", [
    (nc.C128, "512", 7), (nc.C64, "512", 7), (nc.C128, "256", 7), (nc.C64, "256", 7),
    (nc.C128, "256", 11), (nc.C64, "256", 12), (nc.C128, "512", 12)])
def test_kernel_code_on_host_matches_oracle(precision, threads, tile_bits, monkeypatch):
    """The device kernel's own source run on the host (one workgroup as 512
    or 256 threads with real barriers): whole-state entries and multi-tile
    entries whose steps wait on each other, on tiles below the thread bits
    and on the widest tile each workgroup size runs (chunked tile loads,
    paired MFMA rounds)."""
    monkeypatch.setenv("NCSV_BATCH_THREADS", threads)
    cases = [(5, 4, 31), (8, 5, 32), (9, 4, 33)]
    if tile_bits > 7:
        cases += [(tile_bits + 1, 4, 34), (tile_bits, 3, 35)]
    texts = [random_circuit(n, d, s) for n, d, s in cases]
    paulis = [random_paulis(n, 3, s) for n, d, s in cases]
    b = _build(texts, paulis, precision, tile_bits)
    us = _draws(b, 21)
    b.simulate(uniforms=us)
    tol = 1e-12 if precision == nc.C128 else 2e-5
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc, f"entry {i}"
        assert abs(nn - rnn) < tol
        for a, r in zip(ex, rex):
            assert abs(a - r) < tol, (i, a, r)


@pytest.mark.parametrize("low_bits,tiles", [(3, 1), (0, 2), (2, 32)])
def test_emulated_shapes_match_oracle(low_bits, tiles):
    """Other tile shapes (fewer fixed low qubits, queue item sizes) plan
    programs that still equal the oracle."""
    cases = CASES[2:]
    texts = [random_circuit(n, d, s) for n, d, s in cases]
    paulis = [random_paulis(n, 4, s) for n, d, s in cases]
    b = _build(texts, paulis, nc.C128, 7)
    b.set_shape(low_bits, tiles)
    us = _draws(b, 5)
    b.emulate(uniforms=us)
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12


def test_autotune_choice_breaks_near_ties_toward_longer_items():
    """The fastest candidate wins unless another within 3 % has longer queue
    items (the measured samples of the trajectory and circuit mixes)."""
    from ncsim_amd import autotune as AT
    cfg = lambda b, lo, it: {"tile_bits": b, "low_bits": lo, "tiles_per_item": it}
    traj = [(cfg(13, 5, 32), 293.56), (cfg(13, 5, 128), 299.62), (cfg(13, 3, 32), 295.47),
            (cfg(13, 3, 128), 302.07), (cfg(12, 5, 32), 333.75), (cfg(12, 3, 128), 319.68)]
    assert AT.choose(traj) == cfg(13, 5, 128)
    mix = [(cfg(13, 5, 32), 38.30), (cfg(13, 5, 128), 66.99), (cfg(12, 5, 32), 33.46),
           (cfg(12, 3, 32), 31.66), (cfg(12, 3, 128), 96.60)]
    assert AT.choose(mix) == cfg(12, 3, 32)
    assert AT.choose([(cfg(12, 3, 32), 10.0), (cfg(12, 3, 128), 10.5)]) == cfg(12, 3, 32)


def test_autotune_candidates_and_signature():
    from ncsim_amd import autotune as AT
    c64 = AT.candidates(nc.C64)
    assert {c["tile_bits"] for c in c64} == {13, 12}
    assert all(c["low_bits"] in (5, 3) for c in c64)
    b = _build([random_circuit(10, 3, 1), random_circuit(14, 3, 2)], [[], []], nc.C128, 0)
    sig = AT._signature(b)
    assert sig.startswith("c128|") and "draws" in sig


def test_batch_rejects_unsupported():
    b = nc.HipBatch(nc.C128)
    c = nc.Circuit.parse("QUBITS 3\nH 0\nMX 1\n")
    b.add(c, [])
    with pytest.raises(nc.NcsvError):
        b.emulate(uniforms=[np.array([0.5])])


# ----------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("precision", [nc.C128, nc.C64])
@pytest.mark.parametrize("tile_bits", [0, 7])
def test_gpu_batch_matches_oracle(precision, tile_bits):
    texts = [random_circuit(n, d, s) for n, d, s in CASES]
    paulis = [random_paulis(n, 5, s) for n, d, s in CASES]
    b = _build(texts, paulis, precision, tile_bits)
    us = _draws(b, 13)
    b.run(uniforms=us)
    tol = 1e-12 if precision == nc.C128 else 2e-5
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc, f"entry {i}"
        assert abs(nn - rnn) < tol * max(1.0, rnn)
        for a, r in zip(ex, rex):
            assert abs(a - r) < tol, (i, a, r)


@pytest.mark.gpu
def test_gpu_batch_uneven_sizes_and_many_entries():
    """A few hundred entries of 4..16 qubits (whole-state and HBM-resident)
    in one launch, each equal to its own oracle run."""
    rng = np.random.default_rng(5)
    specs = [(int(rng.integers(4, 17)), int(rng.integers(2, 12)), 1000 + i) for i in range(160)]
    texts = [random_circuit(n, d, s, measure=True, noise=True) for n, d, s in specs]
    paulis = [random_paulis(n, 3, s) for n, d, s in specs]
    b = _build(texts, paulis, nc.C128, 0)
    us = _draws(b, 17)
    b.run(uniforms=us)
    for i in range(0, len(specs), 7):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(texts[i], paulis[i], us[i])
        assert oc == roc, f"entry {i}"
        assert max([0.0] + [abs(a - r) for a, r in zip(ex, rex)]) < 1e-12
        assert abs(nn - rnn) < 1e-12


@pytest.mark.gpu
def test_gpu_graph_replays_match_eager(monkeypatch):
    """Repeated runs replay the captured launch; every replay, with other
    draws in between, equals an eager launch of the same draws."""
    texts = [random_circuit(n, d, s) for n, d, s in CASES[:4]] + [random_circuit(15, 6, 77)]
    paulis = [random_paulis(n, 3, s) for n, d, s in CASES[:4]] + [random_paulis(15, 3, 77)]
    b = _build(texts, paulis, nc.C128, 0)
    u1, u2 = _draws(b, 1), _draws(b, 2)
    res = {}
    for tag, us in (("a", u1), ("b", u2), ("a2", u1), ("b2", u2)):
        b.run(uniforms=us)
        res[tag] = [b.result(i) for i in range(len(b))]
    assert b.stats()["graph_replays"] >= 3
    monkeypatch.setenv("NCSV_BATCH_NO_GRAPH", "1")
    for tag, us in (("a", u1), ("b", u2)):
        b.run(uniforms=us)
        eager = [b.result(i) for i in range(len(b))]
        for got in (res[tag], res[tag + "2"]):
            for (e1, n1, o1), (e2, n2, o2) in zip(got, eager):
                assert o1 == o2 and n1 == n2 and e1 == e2


@pytest.mark.gpu
def test_gpu_queue_spread_and_item_sizes_keep_results(monkeypatch):
    """Item sizes change only the queue: the thin depths cut into shorter
    items (a low spread so that this small batch qualifies), the default
    queue and long items give bit-identical results, equal to the oracle."""
    specs = [(n, d, 900 + k) for k, (n, d) in enumerate([(14, 9), (15, 14), (16, 6), (13, 20), (15, 4), (16, 11)])]
    texts = [random_circuit(n, d, s) for n, d, s in specs]
    paulis = [random_paulis(n, 3, s) for n, d, s in specs]
    b = _build(texts, paulis, nc.C128, 8)
    us = _draws(b, 5)
    runs = []
    for spread, items in (("0", -1), ("16", -1), ("16", 128), ("0", 1)):
        monkeypatch.setenv("NCSV_BATCH_SPREAD", spread)
        b.set_shape(-1, items)
        b.run(uniforms=us)
        runs.append((b.stats()["items"], [b.result(i) for i in range(len(b))]))
    assert runs[1][0] > runs[0][0]  # the spread cut some depths finer
    for _, res in runs[1:]:
        assert res == runs[0][1]
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = runs[0][1][i]
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc
        assert max(abs(a - r) for a, r in zip(ex, rex)) < 1e-12


@pytest.mark.gpu
def test_gpu_autotune_keeps_results(tmp_path):
    """The autotuner times its candidates on a sample, then the queue item
    sizes on the whole batch, applies one, caches it; results stay those of
    the oracle."""
    from ncsim_amd import autotune as AT
    specs = [(n, 6, 300 + n) for n in (10, 12, 14, 15, 16)]
    texts = [random_circuit(n, d, s, noise=False) for n, d, s in specs]
    paulis = [random_paulis(n, 4, s) for n, d, s in specs]
    b = _build(texts, paulis, nc.C128, 0)
    cache = str(tmp_path / "tune.json")
    r = b.autotune(sample=3, reps=1, cache=cache)
    assert not r["cached"] and r["config"] in [c for c, _ in r["timings"]]
    assert len(r["timings"]) == len(AT.candidates(nc.C128)) + len(AT.ITEMS)
    assert b.autotune(sample=5, reps=1, cache=cache)["cached"]
    us = _draws(b, 3)
    b.run(uniforms=us)
    for i, (t, ps) in enumerate(zip(texts, paulis)):
        ex, nn, oc = b.result(i)
        rex, rnn, roc = oracle_entry(t, ps, us[i])
        assert oc == roc
        assert max(abs(a - r_) for a, r_ in zip(ex, rex)) < 1e-12


@pytest.mark.gpu
def test_gpu_batch_equals_emulator_and_single_state():
    """The kernel equals the host emulator; a noiseless unitary entry equals
    the single-state fused run of the same circuit."""
    texts = [random_circuit(14, 12, 40, measure=False, noise=False), random_circuit(16, 10, 41, measure=True)]
    paulis = [random_paulis(14, 8, 40), random_paulis(16, 8, 41)]
    b = _build(texts, paulis, nc.C128, 0)
    us = _draws(b, 19)
    b.run(uniforms=us)
    dev = [b.result(i) for i in range(2)]
    b.emulate(uniforms=us)
    emu = [b.result(i) for i in range(2)]
    for (e1, n1, o1), (e2, n2, o2) in zip(dev, emu):
        assert o1 == o2
        assert abs(n1 - n2) < 1e-12
        assert max(abs(a - c) for a, c in zip(e1, e2)) < 1e-12
    s = nc.HipStateVector.zero_state(14)
    s.run(nc.Circuit.parse(texts[0]))
    ref = s.expectations([nc.Pauli(14, p.x, p.z, p.phase) for p in paulis[0]])
    assert max(abs(a - c) for a, c in zip(dev[0][0], ref)) < 1e-12
