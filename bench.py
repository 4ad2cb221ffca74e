"""Benchmark: amplitude updates per second of the dense complex128 state
vector on one MI355X, on the 30-qubit rotation brickwork (depth 24: RX/RY/RZ
on every qubit, uniform axis and angle, then CZ on alternating neighbour
pairs) followed by 64 random Pauli expectations (weight 1-4).

One step = reset to |0...0>, run the circuit, compute the 64 expectations.
amp-updates = (unfused gate count) x 2^n per step.  Prints one JSON line with
the roofline of the dominant kernel (HIP-event timed on the state's stream),
the precision form measured in the same run (QFT + 2000 RZ: 33 qubits
complex64, 32 qubits complex128 and complex64, fidelity and max-abs error of
complex64 against complex128), the complex64 sweeps' FP32 MFMA blocks against
the vector-only plan on a 30-qubit random SU(4) brickwork, and a CPU
baseline (the C oracle on a bounded sample of the 26-qubit member of the
brickwork family, on all of this host's cores and on one thread).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(n, depth, fusion, kernel):
    """HBM bytes per launch of the dominant kernel from the committed counter
    runs of this exact workload (profiles/traffic_*.json), else None."""
    path = os.path.join(ROOT, "profiles", "traffic_sweep_c128_brickwork30.json")
    if (n, depth, fusion) != (30, 24, 1) or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t["bytes_per_launch"] if t.get("kernel") == kernel else None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(seconds=12.0, seconds_one=8.0):
    """The oracle (C restatement of StateVector<C64>, OpenMP) applying the
    26-qubit member of the same circuit family gate by gate until `seconds`
    have passed, on every thread OpenMP is given (OMP_NUM_THREADS, else all
    cores) and again on one thread."""
    from oracle.circuit import Circuit as OCircuit
    from oracle.statevector import StateVector, lib
    from ncsim_amd import workloads as W

    n = 26
    circ = OCircuit.parse(W.rotation_brickwork(n, 24, 3))
    v = StateVector.zero_state(n)

    def timed(limit):
        t0 = time.perf_counter()
        done = 0
        for op in circ.ops:
            v.apply_gate(op[1], op[2])
            done += 1
            if time.perf_counter() - t0 > limit:
                break
        return done, time.perf_counter() - t0

    threads = lib().orc_num_threads()
    done, dt = timed(seconds)
    lib().orc_set_num_threads(1)
    done1, dt1 = timed(seconds_one)
    lib().orc_set_num_threads(threads)
    del v
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"value": done * float(1 << n) / dt, "unit": "amp-updates/s", "cores": threads, "kind": "port",
            "sample": f"first {done} gates of the 26-qubit depth-24 rotation brickwork, "
                      f"{dt:.1f} s, {threads} OpenMP threads",
            "single_thread": {"value": done1 * float(1 << n) / dt1, "cores": 1,
                              "sample": f"first {done1} gates of the same circuit, {dt1:.1f} s, 1 thread"},
            "host": {"nproc": os.cpu_count(), "affinity": affinity, "cpu_model": cpu_model(),
                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
            "what": "C restatement of ncsim StateVector<C64> (oracle/sv_oracle.c), -O3, OpenMP"}


def precision_form(n=33, steps=3, fusion=1, only_n=False):
    """The precision form on QFT (H + CPHASE(1, 2..n)) with 2000 RZ after a
    seeded X layer: complex128 and complex64 at n-1 qubits (fidelity and
    max-abs error of complex64 against complex128, on the device, and the
    complex64 / complex128 speed-up at equal size), then complex64 at n
    qubits with its dominant kernel's roofline.  One step = set_basis(0) +
    the circuit; `steps` timed after one untimed step (plans, compiles)."""
    import ncsim_amd as H
    from ncsim_amd import workloads as W

    def timed(m, prec):
        text = W.qft_with_rz(m, 2000, 7)
        gates = W.count_gates(text)
        circ = H.Circuit.parse(text)
        s = H.HipStateVector.zero_state(m, prec)
        s.set_fusion(fusion)
        s.set_basis(0)
        s.run(circ)  # warm-up (plans, kernels), the same calls as a timed step
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.set_basis(0)
            s.run(circ)
        s.synchronize()
        ms = (time.perf_counter() - t0) * 1000 / steps
        s.set_basis(0)
        s.profile_reset()
        s.profile(True)
        s.run(circ)
        s.profile(False)
        st = s.profile_read()
        sweeps = s.sweep_count()
        log(f"qft: n={m} prec={'c128' if prec == H.C128 else 'c64'} gates={gates} {ms:.1f} ms/step  " +
            ", ".join(f"{k}: {v[0]} x {v[1] / v[0]:.3f} ms" for k, v in sorted(st.items())))
        return s, ms, gates, st, sweeps

    if only_n:  # the n-qubit complex64 part alone (counter runs)
        c, ms64, g2, st64, sweeps = timed(n, H.C64)
        return None
    a, ms128, g1, _, _ = timed(n - 1, H.C128)
    b, ms64s, _, _, _ = timed(n - 1, H.C64)
    cmp = a.compare(b)
    del a, b
    c, ms64, g2, st64, sweeps = timed(n, H.C64)
    del c
    name, (launches, tot_ms, byts) = max(st64.items(), key=lambda kv: kv[1][1])
    achieved = byts / launches / (tot_ms / launches / 1000.0)
    return {"workload": f"QFT + 2000 RZ (seeded X prefix): {n} qubits complex64; {n - 1} qubits complex128 "
                        f"and complex64", "qubits": n, "gates": g2, "steps": steps,
            "ms_per_step": ms64, "value": g2 * float(1 << n) / (ms64 / 1000.0), "unit": "amp-updates/s",
            "dtype": "c64 (f32, f64 reductions)",
            "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK,
                         "traffic": measured_traffic_c2(n, name),
                         "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                         "algorithmic_bytes": byts / launches, "launches": launches, "avg_ms": tot_ms / launches,
                         "sweeps_per_step": sweeps},
            "check": {"qubits": n - 1, "fidelity": cmp["fidelity"], "fidelity_loss": 1 - cmp["fidelity"],
                      "max_abs": cmp["max_abs"]},
            "c128": {"qubits": n - 1, "ms_per_step": ms128, "value": g1 * float(1 << (n - 1)) / (ms128 / 1000.0)},
            "c64_same_size": {"qubits": n - 1, "ms_per_step": ms64s,
                              "value": g1 * float(1 << (n - 1)) / (ms64s / 1000.0)},
            "speedup_c64_over_c128": ms128 / ms64s,
            "gate_layers_per_s_ratio_n_c64_vs_n1_c128": (ms128 / ms64)}


def measured_traffic_c2(n, kernel):
    """HBM bytes per launch of the complex64 QFT's dominant kernel from the
    committed counter runs (profiles/traffic_sweep_c64_qft33.json), else None."""
    path = os.path.join(ROOT, "profiles", "traffic_sweep_c64_qft33.json")
    if n != 33 or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t["bytes_per_launch"] if t.get("kernel") == kernel else None


def run_qft(args):
    """The precision form alone; value = complex64 amp-updates/s at n qubits."""
    n = args.qubits if args.qubits != 30 else 33
    pf = precision_form(n, args.steps, args.fusion, args.only_n)
    if pf is None:
        return
    out = {"metric": "amp-updates/s", "value": pf["value"], "unit": "amp-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 1, "ms_per_step": pf["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": pf["dtype"], "data": "synthetic (seeded circuit)",
           "config": {"workload": pf["workload"], "gates": pf["gates"]}, "roofline": pf["roofline"],
           "precision_check": pf["check"], "c128": pf["c128"], "c64_same_size": pf["c64_same_size"],
           "speedup_c64_over_c128": pf["speedup_c64_over_c128"]}
    print(json.dumps(out), flush=True)


def run_batch(args):
    """Batched form: `--circuits` uneven circuits (10-22 qubits, depth
    10-120, H/S/T/SQRT_X/RZ layers, CX matchings, CCZ with p = 0.05) with 8
    Paulis each, all in one persistent launch per run, complex64 by default.
    Compared with per-circuit launches in the same precision (each circuit
    on its own state through the fused single-state path) over every
    `--stride`-th circuit; the same for the subset of 10-14 qubits.  In
    complex64 each sampled circuit is also run in complex128 and its
    expectations compared with the batch's.  value = circuits/s of the
    batched form on the whole set."""
    import numpy as np

    import ncsim_amd as H
    from ncsim_amd import workloads as W

    prec = H.C128 if args.precision == "c128" else H.C64
    specs = W.batch_circuits(args.circuits, 11)
    if args.qmin or args.qmax:  # one size band of the mix
        specs = [sp for sp in specs if (args.qmin or 0) <= sp[0] <= (args.qmax or 64)]
    circs = [H.Circuit.parse(t) for _, t, _ in specs]
    paulis = [[H.Pauli(n, x, z, ph) for x, z, ph in ps] for n, _, ps in specs]
    gates = np.array([W.count_gates(t) for _, t, _ in specs], dtype=np.float64)
    ns = np.array([n for n, _, _ in specs])

    tuned = {}

    def batched(idx):
        b = H.HipBatch(prec)
        for i in idx:
            b.add(circs[i], paulis[i])
        if args.tile_bits:
            b.set_tile_bits(args.tile_bits)
            b.set_shape(args.low_bits, args.items)
        elif not args.no_autotune:
            t0 = time.perf_counter()
            r = b.autotune(log=log)
            tuned[len(idx)] = {"config": r["config"], "cached": r["cached"],
                               "seconds": time.perf_counter() - t0}
            log(f"autotune: {r['config']} ({'cached' if r['cached'] else 'measured'})")
        t0 = time.perf_counter()
        b.run()  # plans every circuit and uploads the programs
        cold = time.perf_counter() - t0
        for _ in range(max(0, args.warmup - 1)):
            b.run()
        t0 = time.perf_counter()
        kms = 0.0
        for _ in range(args.steps):
            b.run()
            kms += b.stats()["kernel_ms"]
        wall = (time.perf_counter() - t0) / args.steps
        st = b.stats()
        log(f"batch of {len(idx)}: {wall * 1000:.1f} ms per run (kernel {kms / args.steps:.1f} ms, first run "
            f"{cold:.2f} s, planning {st['plan_ms']:.0f} ms), {st['steps']:.0f} steps, {st['items']:.0f} queue items")
        return b, wall, kms / args.steps, cold

    os.environ["NCSV_JIT_MIN_QUBITS"] = "99"  # per-circuit runs: no per-circuit kernel compiles

    def per_circuit(idx, b, pos):
        states = {}
        for i in idx:  # warm-up: plans cached on each circuit
            n = specs[i][0]
            s = states.setdefault(n, H.HipStateVector.zero_state(n, prec))
            s.set_basis(0)
            s.run(circs[i])
            s.expectations(paulis[i])
        for s in states.values():
            s.synchronize()
        t0 = time.perf_counter()
        evs = []
        for i in idx:
            s = states[specs[i][0]]
            s.set_basis(0)
            s.run(circs[i])
            evs.append(s.expectations(paulis[i]))
        per = (time.perf_counter() - t0) / len(idx)
        maxdiff = 0.0
        for i, ev in zip(idx, evs):
            bv = b.result(pos[i])[0]
            maxdiff = max(maxdiff, max(abs(a - c) for a, c in zip(ev, bv)))
        diff128 = None
        if prec == H.C64:  # each sampled circuit against its own complex128 run
            states = {}
            diff128 = 0.0
            for i in idx:
                s = states.setdefault(specs[i][0], H.HipStateVector.zero_state(specs[i][0], H.C128))
                s.set_basis(0)
                s.run(circs[i])
                ev = s.expectations(paulis[i])
                diff128 = max(diff128, max(abs(a - c) for a, c in zip(ev, b.result(pos[i])[0])))
            del states
        log(f"per-circuit: {per * 1000:.2f} ms per circuit over {len(idx)} circuits; max |batch - single| = "
            f"{maxdiff:.2e}" + ("" if diff128 is None else f"; max |batch - complex128 single| = {diff128:.2e}"))
        return per, maxdiff, diff128

    allidx = list(range(len(specs)))
    b, wall, kms, cold = batched(allidx)
    per, maxdiff, d128 = per_circuit(allidx[::args.stride], b, {i: i for i in allidx})
    del b
    small = [i for i in allidx if ns[i] <= 14]
    if small:
        bs, wall_s, kms_s, _ = batched(small)
        per_s, maxdiff_s, d128_s = per_circuit(small[::max(1, args.stride // 2)], bs,
                                               {i: k for k, i in enumerate(small)})
        del bs
    else:
        wall_s = kms_s = per_s = 1.0
        maxdiff_s, d128_s = 0.0, (None if d128 is None else 0.0)
    amp_updates = float((gates * np.exp2(ns)).sum())
    cps = len(specs) / wall
    out = {"metric": "circuits/s", "value": cps, "unit": "circuits/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1000, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64 (f32, f64 reductions)" if prec == H.C64 else "c128 (f64)",
           "data": "synthetic (seeded circuits)",
           "config": {"workload": f"{len(specs)} circuits, {int(ns.min())}-{int(ns.max())} qubits, depth 10-120, "
                                  f"8 Paulis each",
                      "gates": int(gates.sum()), "tile_bits": args.tile_bits or None},
           "amp_updates_per_s": amp_updates / wall, "kernel_ms": kms, "first_run_s": cold,
           "per_circuit": {"circuits_per_s": 1.0 / per, "sample": len(allidx[::args.stride]),
                           "ms_per_circuit": per * 1000,
                           "path": "single-state fused interpreter, one state per qubit count"},
           "speedup_vs_per_circuit": cps * per, "max_abs_batch_vs_single": max(maxdiff, maxdiff_s),
           "max_abs_vs_complex128_single": None if d128 is None else max(d128, d128_s),
           "autotune": tuned,
           "small_circuits": {"qubits": "10-14", "count": len(small), "circuits_per_s": len(small) / wall_s,
                              "kernel_ms": kms_s, "per_circuit_circuits_per_s": 1.0 / per_s,
                              "speedup_vs_per_circuit": len(small) / wall_s * per_s}}
    print(json.dumps(out), flush=True)


def run_traj(args):
    """Noisy trajectories: 24-qubit depth-40 brickwork of Haar SU(2) and CX
    with DEPOLARIZE1(1e-3) after every gate and M+R on 4 qubits every 10
    layers; `--trajectories` complex64 trajectories (SplitMix64 seeds) in
    the batched form; the first 256 again in complex128 with the same draws
    and the complex64 run's measurement outcomes forced.  Estimators: <Z_i>
    and <Z_i Z_i+1> per trajectory (normalized).  value = trajectories/s."""
    import numpy as np

    import ncsim_amd as H
    from ncsim_amd import workloads as W

    n = 24 if args.qubits == 30 else args.qubits
    text = W.noisy_brickwork(n, 40, 5)
    circ = H.Circuit.parse(text)
    ps = [H.Pauli(n, 0, 1 << i, 0) for i in range(n)] + [H.Pauli(n, 0, 3 << i, 0) for i in range(n - 1)]
    T = args.trajectories

    def make(prec, count):
        b = H.HipBatch(prec)
        for _ in range(count):
            b.add(circ, ps)
        return b

    b64 = make(H.C64, T)
    tuned = {}
    if args.tile_bits:
        b64.set_tile_bits(args.tile_bits)
        b64.set_shape(args.low_bits, args.items)
    elif not args.no_autotune:
        t0 = time.perf_counter()
        r = b64.autotune(log=log, sample=48)
        tuned["c64"] = {"config": r["config"], "cached": r["cached"], "seconds": time.perf_counter() - t0}
        log(f"autotune complex64: {r['config']} ({'cached' if r['cached'] else 'measured'})")
    seeds = np.arange(T, dtype=np.uint64) + 1000
    b64.run(seeds=seeds)  # warm-up (planning, programs)
    t0 = time.perf_counter()
    b64.run(seeds=seeds)
    wall = time.perf_counter() - t0
    st = b64.stats()
    est64 = []
    for i in range(T):
        ev, nn, oc = b64.result(i)
        est64.append([e.real / nn for e in ev])
    est64 = np.array(est64)
    log(f"trajectories: {T} x {n} qubits complex64 in {wall:.2f} s (kernel {st['kernel_ms']:.0f} ms, "
        f"{st['launches']:.0f} launches)")
    m = min(256, T)
    b128 = make(H.C128, m)
    if not args.tile_bits and not args.no_autotune:
        r = b128.autotune(log=log, sample=48)
        tuned["c128"] = {"config": r["config"], "cached": r["cached"]}
        log(f"autotune complex128: {r['config']} ({'cached' if r['cached'] else 'measured'})")
    forced = []
    for i in range(m):
        oc = b64.result(i)[2]
        forced.append(np.array([0xff if o is None else o for o in oc], dtype=np.uint8))
    t1 = time.perf_counter()
    b128.run(seeds=seeds[:m], forced=forced)
    w128 = time.perf_counter() - t1
    est128 = np.array([[e.real / b128.result(i)[1] for e in b128.result(i)[0]] for i in range(m)])
    diff = float(np.abs(est64[:m] - est128).max())
    log(f"check: {m} complex128 trajectories in {w128:.2f} s, max |est64 - est128| = {diff:.2e}")
    # per-trajectory launches: the same circuit through the single-state run
    # (noise ops split its fused segments), its own SplitMix64 draws
    single = min(args.single, T)
    # per-trajectory runs through the interpreter kernels: each trajectory's noise
    # splits the circuit differently, so specialized kernels would compile per run
    os.environ["NCSV_JIT_MIN_QUBITS"] = "99"
    s = H.HipStateVector.zero_state(n, H.C64)
    ms_single = None
    if single:
        s.run(circ, rng=W.SplitMix64(1))  # warm-up: plans
        s.synchronize()
        t2 = time.perf_counter()
        for k in range(single):
            s.set_basis(0)
            s.run(circ, rng=W.SplitMix64(100 + k))
            s.expectations(ps)
        s.synchronize()
        ms_single = (time.perf_counter() - t2) * 1000 / single
        log(f"per-trajectory launches: {ms_single:.1f} ms per trajectory over {single}")
    del s
    gates = W.count_gates(text)
    out = {"metric": "trajectories/s", "value": T / wall, "unit": "trajectories/s", "n_gpus": 1, "steps": 1,
           "warmup": 1, "ms_per_step": wall * 1000, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "c64 (f32, f64 reductions)", "data": "synthetic (seeded circuit and draws)",
           "config": {"workload": f"{T} noisy trajectories, {n} qubits, depth 40, DEPOLARIZE1(1e-3), M+R every 10 layers",
                      "gates": gates},
           "amp_updates_per_s": gates * float(1 << n) * T / wall,
           "estimates": {"mean_Z": est64[:, :n].mean(axis=0).tolist()[:4],
                         "mean_ZZ": est64[:, n:].mean(axis=0).tolist()[:4]},
           "fp64_check": {"trajectories": m, "max_abs_estimate_diff": diff, "seconds_c128": w128},
           "autotune": tuned,
           "per_trajectory": {"ms_per_trajectory": ms_single, "trajectories": single,
                              "speedup_batched": (ms_single / (wall * 1000 / T)) if ms_single else None}}
    print(json.dumps(out), flush=True)


def run_copy(args):
    """Counter calibration: device-to-device copies of an n-qubit state
    (ncsv_clone), each reading and writing 2^n amplitudes once."""
    import ncsim_amd as H
    prec = H.C64 if args.precision == "c64" else H.C128
    s = H.HipStateVector.plus_state(args.qubits, prec)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        c = s.clone()
        c.synchronize()
        del c
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    b = (1 << args.qubits) * (16 if prec == H.C128 else 8)
    print(json.dumps({"workload": f"copy of a {args.qubits}-qubit state", "bytes_read": b, "bytes_written": b,
                      "copies": args.steps, "ms_per_copy": ms}), flush=True)


def run_frames(args):
    """The Pauli-frame sampler: a distance-d repetition code (d data qubits,
    d - 1 ancillas measured and reset every round, X_ERROR(p) on the data
    before each round, detectors comparing consecutive rounds), sampled for
    detectors and observables; shots/s including the records' download."""
    import ncsim_amd as H
    from ncsim_amd import workloads as W
    d, rounds, p = args.distance, args.rounds, 1e-3
    shots = args.shots
    text = W.repetition_code(d, rounds, p)
    f = H.HipFrameSampler(H.Circuit.parse(text), 1)
    for i in range(args.warmup):
        f.sample_detectors(shots, 1000 + i)
    t0 = time.perf_counter()
    for i in range(args.steps):
        f.sample_detectors(shots, i)
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    print(json.dumps({"metric": "shots/s", "value": shots / (ms / 1000.0), "unit": "shots/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "u64 (bit-packed frames)",
                      "data": "synthetic (seeded draws)",
                      "config": {"workload": f"repetition code d={d}, {rounds} rounds, X_ERROR({p})",
                                 "qubits": 2 * d - 1, "measurements": f.num_measurements,
                                 "detectors": f.num_detectors, "shots": shots}}), flush=True)


def run_stream(args):
    """Streaming ceilings of the library's own elementwise kernels on an
    n-qubit state (default 30, complex128): the FP64 norm reduction
    (read-only), ncsv_scale (in-place read and write) and the plus-state fill
    (write-only), timed by HIP events on the state's stream."""
    import ncsim_amd as H
    prec = H.C64 if args.precision == "c64" else H.C128
    s = H.HipStateVector.plus_state(args.qubits, prec)
    s.synchronize()
    s.norm_sq()
    s.scale(1.0)
    s.profile_reset()
    s.profile(True)
    for _ in range(args.steps):
        s.norm_sq()
        s.scale(1.0)
        H._lib.lib().ncsv_set_plus(s._h)
    s.profile(False)
    st = s.profile_read()
    out = {}
    for k, (launches, ms, byts) in sorted(st.items()):
        out[k] = {"launches": launches, "avg_ms": ms / launches, "bytes": byts / launches,
                  "tb_per_s": byts / (ms / 1000.0) / 1e12, "frac_of_8tbs": byts / (ms / 1000.0) / HBM_PEAK}
    print(json.dumps({"workload": f"streaming kernels on a {args.qubits}-qubit "
                                  f"{'complex64' if prec == H.C64 else 'complex128'} state",
                      "kernels": out}), flush=True)


def su4_measure(n, depth, prec, mfma, steps=3, warmup=1, fusion=1):
    """The random two-qubit circuit (a line of Haar-random 4x4 unitaries on
    alternating neighbour pairs, `depth` layers) on an n-qubit state: one step
    = set_basis(0) + the circuit, with the sweeps' FP32 MFMA blocks off
    (mfma=0) or chosen by cost (mfma=1).  Returns ms per step, amp-updates/s
    and the dominant kernel's roofline."""
    import ncsim_amd as H
    from ncsim_amd import workloads as W
    old = os.environ.get("NCSV_JIT_MFMA")
    os.environ["NCSV_JIT_MFMA"] = str(mfma)
    try:
        gates = W.su4_brickwork(n, depth, 3)
        circ = W.build_circuit(n, gates, H.Circuit, H.Gate)
        s = H.HipStateVector.zero_state(n, prec)
        s.set_fusion(fusion)
        for _ in range(max(1, warmup)):
            s.set_basis(0)
            s.run(circ)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.set_basis(0)
            s.run(circ)
        s.synchronize()
        ms = (time.perf_counter() - t0) * 1000 / steps
        s.set_basis(0)
        s.profile_reset()
        s.profile(True)
        s.run(circ)
        s.profile(False)
        st = s.profile_read()
        sweeps = s.sweep_count()
        del s
    finally:
        if old is None:
            os.environ.pop("NCSV_JIT_MFMA", None)
        else:
            os.environ["NCSV_JIT_MFMA"] = old
    name, (launches, tot_ms, byts) = max(st.items(), key=lambda kv: kv[1][1])
    achieved = byts / launches / (tot_ms / launches / 1000.0)
    log(f"su4: n={n} depth={depth} mfma={mfma} {ms:.1f} ms/step  " +
        ", ".join(f"{k}: {v[0]} x {v[1] / v[0]:.3f} ms" for k, v in sorted(st.items())))
    return {"ms_per_step": ms, "value": len(gates) * float(1 << n) / (ms / 1000.0), "gates": len(gates),
            "sweeps_per_step": sweeps,
            "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK, "launches": launches,
                         "avg_ms": tot_ms / launches}}


def mfma_blocks(n=30, depth=12, steps=2):
    """FP32 MFMA dense blocks against the vector-only sweeps on the random
    SU(4) brickwork, complex64, measured in the same run."""
    import ncsim_amd as H
    on = su4_measure(n, depth, H.C64, 1, steps)
    off = su4_measure(n, depth, H.C64, 0, steps)
    return {"workload": f"random SU(4) brickwork {n} qubits depth {depth}, complex64", "gates": on["gates"],
            "mfma_by_cost": on, "vector_only": off, "mfma_over_vector_time": on["ms_per_step"] / off["ms_per_step"],
            "default": "vector only (NCSV_JIT_MFMA=0)"}


def run_su4(args):
    """The random SU(4) brickwork alone (complex64 unless --precision c128);
    MFMA blocks as NCSV_JIT_MFMA says (default 0: none)."""
    import ncsim_amd as H
    prec = H.C128 if args.precision == "c128" else H.C64
    r = su4_measure(args.qubits, args.depth, prec, int(os.environ.get("NCSV_JIT_MFMA", "0")), args.steps,
                    args.warmup, args.fusion)
    print(json.dumps({"metric": "amp-updates/s", "value": r["value"], "unit": "amp-updates/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": max(1, args.warmup), "ms_per_step": r["ms_per_step"],
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "c64 (f32, f64 reductions)" if prec == H.C64 else "c128 (f64)",
                      "data": "synthetic (seeded circuit)",
                      "config": {"workload": f"random SU(4) brickwork {args.qubits} qubits depth {args.depth}",
                                 "gates": r["gates"], "mfma_blocks": os.environ.get("NCSV_JIT_MFMA", "0")},
                      "roofline": r["roofline"], "sweeps_per_step": r["sweeps_per_step"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=24)
    ap.add_argument("--fusion", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision-form", action="store_true")
    ap.add_argument("--only-n", action="store_true", help="qft workload: only the n-qubit complex64 runs")
    ap.add_argument("--workload", default="brickwork",
                    choices=["brickwork", "qft", "batch", "trajectories", "frames", "copy", "su4", "stream"])
    ap.add_argument("--distance", type=int, default=25, help="frames workload: repetition-code distance")
    ap.add_argument("--rounds", type=int, default=25, help="frames workload: rounds")
    ap.add_argument("--shots", type=int, default=1 << 24, help="frames workload: shots per step")
    ap.add_argument("--circuits", type=int, default=4096)
    ap.add_argument("--stride", type=int, default=8)
    ap.add_argument("--trajectories", type=int, default=2000)
    ap.add_argument("--single", type=int, default=3, help="trajectories timed through per-trajectory launches")
    ap.add_argument("--precision", default=None, choices=["c128", "c64"],
                    help="default: complex128 for the brickwork, complex64 for the batch")
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=-1, help="batch workload with --tile-bits: fixed low tile qubits")
    ap.add_argument("--items", type=int, default=-1, help="batch workload with --tile-bits: tiles per queue item")
    ap.add_argument("--qmin", type=int, default=0, help="batch workload: only circuits of at least this many qubits")
    ap.add_argument("--qmax", type=int, default=0, help="batch workload: only circuits of at most this many qubits")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only: the state-vector path runs on a single "
                                   "MI355X; sharding across GPUs is not built"}))
        sys.exit(2)

    if args.workload == "qft":
        run_qft(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "trajectories":
        run_traj(args)
        return
    if args.workload == "frames":
        run_frames(args)
        return
    if args.workload == "copy":
        run_copy(args)
        return
    if args.workload == "su4":
        run_su4(args)
        return
    if args.workload == "stream":
        run_stream(args)
        return

    import ncsim_amd as H
    from ncsim_amd import workloads as W

    n, depth = args.qubits, args.depth
    text = W.rotation_brickwork(n, depth, 3)
    gates = W.count_gates(text)
    paulis = [H.Pauli(n, x, z, ph) for x, z, ph in W.random_paulis(n, 64, 4)]
    circ = H.Circuit.parse(text)
    prec = H.C64 if args.precision == "c64" else H.C128
    log(f"bench: n={n} depth={depth} gates={gates} paulis=64 fusion={args.fusion} precision={args.precision or 'c128'}")
    s = H.HipStateVector.zero_state(n, prec)
    s.set_fusion(args.fusion)

    def step():
        s.run(circ)
        return s.expectations(paulis)

    for i in range(args.warmup):
        H._lib.lib().ncsv_set_basis(s._h, 0)
        step()
        log(f"bench: warmup {i} done")
    s.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        H._lib.lib().ncsv_set_basis(s._h, 0)
        step()
        log(f"bench: step {i} done")
    s.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1000.0 / args.steps
    value = gates * float(1 << n) / (ms / 1000.0)

    # one profiled step: per-kernel durations by HIP events on the state's stream
    H._lib.lib().ncsv_set_basis(s._h, 0)
    s.profile_reset()
    s.profile(True)
    step()
    s.profile(False)
    stats = s.profile_read()
    sweeps = s.sweep_count()
    name, (launches, tot_ms, byts) = max(stats.items(), key=lambda kv: kv[1][1])
    avg_s = tot_ms / launches / 1000.0
    achieved = byts / launches / avg_s
    log("bench: kernels " + ", ".join(f"{k}: {v[0]} x {v[1] / v[0]:.3f} ms" for k, v in sorted(stats.items())))
    roofline = {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK,
                "traffic": measured_traffic(n, depth, args.fusion, name) if prec == H.C128 else None,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                "algorithmic_bytes": byts / launches, "launches": launches,
                "avg_ms": tot_ms / launches, "sweeps_per_step": sweeps}
    del s
    pf = mb = None
    if not args.no_precision_form and n == 30:
        pf = precision_form(33, max(1, min(args.steps, 3)), args.fusion)
        mb = mfma_blocks()
    cpu = None if args.no_cpu else cpu_baseline()
    out = {
        "metric": "amp-updates/s", "value": value, "unit": "amp-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (f64)" if prec == H.C128 else "c64 (f32, f64 reductions)", "data": "synthetic (seeded circuit)",
        "config": {"workload": f"rotation brickwork {n} qubits depth {depth} + 64 Pauli expectations",
                   "qubits": n, "depth": depth, "gates": gates, "fusion": args.fusion,
                   "precision": "complex128" if prec == H.C128 else "complex64"},
        "roofline": roofline, "precision_form": pf, "mfma_blocks": mb, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
