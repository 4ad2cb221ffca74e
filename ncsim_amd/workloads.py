"""Seeded synthetic circuits in ncsim's text format (circuit.rs:4-27), so the
same file runs on this library and on the reference.

Every generator draws from SplitMix64 and prints angles with repr(), which
round-trips exactly through the parser.
"""
import math


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def gen_f64(self):
        """[0, 1) with 53 random bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, m):
        return int(self.gen_f64() * m)

    def shuffle(self, xs):
        for i in range(len(xs) - 1, 0, -1):
            j = self.below(i + 1)
            xs[i], xs[j] = xs[j], xs[i]
        return xs


def random_layers(n, layers, seed, measure=True):
    """Random Clifford+T layers on n qubits: per layer every qubit gets one
    of H, S, T, SQRT_X; then CX on a uniformly random perfect matching, each
    pair kept with probability 1/2; finally M on every qubit."""
    r = SplitMix64(seed)
    out = [f"QUBITS {n}"]
    for _ in range(layers):
        for q in range(n):
            out.append(f"{['H', 'S', 'T', 'SQRT_X'][r.below(4)]} {q}")
        perm = r.shuffle(list(range(n)))
        for i in range(0, n - 1, 2):
            if r.gen_f64() < 0.5:
                out.append(f"CX {perm[i]} {perm[i + 1]}")
    if measure:
        out.append("M " + " ".join(str(q) for q in range(n)))
    return "\n".join(out) + "\n"


def rotation_brickwork(n, depth, seed):
    """Per layer: RX/RY/RZ on every qubit (axis uniform, angle uniform in
    [0, 2pi)), then CZ on (2i, 2i+1) in even layers and (2i+1, 2i+2) in odd
    layers."""
    r = SplitMix64(seed)
    out = [f"QUBITS {n}"]
    for layer in range(depth):
        for q in range(n):
            ax = ["RX", "RY", "RZ"][r.below(3)]
            out.append(f"{ax}({r.gen_f64() * 2.0 * math.pi!r}) {q}")
        for a in range(layer % 2, n - 1, 2):
            out.append(f"CZ {a} {a + 1}")
    return "\n".join(out) + "\n"


def random_paulis(n, count, seed, max_weight=4):
    """(x, z, phase) Hermitian Pauli strings: weight uniform 1..max_weight,
    support a uniform subset, letters uniform over X, Y, Z (Y = i X Z, so the
    stored phase counts the Ys)."""
    r = SplitMix64(seed)
    out = []
    for _ in range(count):
        w = 1 + r.below(max_weight)
        sup = r.shuffle(list(range(n)))[:w]
        x = z = ph = 0
        for q in sup:
            letter = r.below(3)
            if letter in (0, 1):
                x |= 1 << q
            if letter in (1, 2):
                z |= 1 << q
            if letter == 1:
                ph += 1
        out.append((x, z, ph & 3))
    return out


def qft_with_rz(n, rz_count, seed, prefix_x=True):
    """A seeded X layer (so the controlled phases act), the QFT without the
    final swaps (H(j), then CPHASE(1, m-j+1) on (m, j) for m > j), and
    rz_count RZ(theta ~ U[0, 2pi)) gates at uniformly random places."""
    r = SplitMix64(seed)
    body = []
    for j in range(n):
        body.append(f"H {j}")
        for m in range(j + 1, n):
            body.append(f"CPHASE(1,{m - j + 1}) {m} {j}")
    for _ in range(rz_count):
        pos = r.below(len(body) + 1)
        body.insert(pos, f"RZ({r.gen_f64() * 2.0 * math.pi!r}) {r.below(n)}")
    out = [f"QUBITS {n}"]
    if prefix_x:
        for q in range(n):
            if r.gen_f64() < 0.5:
                out.append(f"X {q}")
    return "\n".join(out + body) + "\n"


def batch_circuits(count, seed, qmin=10, qmax=22, dmin=10, dmax=120, npauli=8):
    """Uneven small circuits: qubits uniform in [qmin, qmax], depth uniform in
    [dmin, dmax]; per layer every qubit gets one of H, S, T, SQRT_X or
    RZ(theta ~ U[0, 2pi)), then CX on a random perfect matching with each pair
    kept with probability 1/2, then with probability 0.05 a CCZ on a random
    triple.  Each comes with `npauli` Pauli strings of weight 1-4.
    Returns [(n, text, [(x, z, phase), ...])]."""
    r = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = qmin + r.below(qmax - qmin + 1)
        depth = dmin + r.below(dmax - dmin + 1)
        lines = [f"QUBITS {n}"]
        for _ in range(depth):
            for q in range(n):
                k = r.below(5)
                if k == 4:
                    lines.append(f"RZ({r.gen_f64() * 2.0 * math.pi!r}) {q}")
                else:
                    lines.append(f"{['H', 'S', 'T', 'SQRT_X'][k]} {q}")
            perm = r.shuffle(list(range(n)))
            for i in range(0, n - 1, 2):
                if r.gen_f64() < 0.5:
                    lines.append(f"CX {perm[i]} {perm[i + 1]}")
            if r.gen_f64() < 0.05:
                a, b, c = r.shuffle(list(range(n)))[:3]
                lines.append(f"CCZ {a} {b} {c}")
        out.append((n, "\n".join(lines) + "\n", random_paulis(n, npauli, r.next_u64())))
    return out


def noisy_brickwork(n, depth, seed, p=1e-3, every=10, mqubits=None):
    """Noisy trajectories' circuit: per layer a Haar-random SU(2) on every
    qubit as RZ(gamma) RY(beta) RZ(alpha) (alpha, gamma ~ U[0, 2pi), beta =
    arccos(1 - 2u)) followed by DEPOLARIZE1(p), then CX on alternating pairs
    ((2i, 2i+1) in even layers, (2i+1, 2i+2) in odd ones) each followed by
    DEPOLARIZE1(p) on both qubits; after every `every` layers M then R on four
    fixed qubits."""
    r = SplitMix64(seed)
    mq = mqubits if mqubits is not None else [0, n // 4, n // 2, (3 * n) // 4]
    lines = [f"QUBITS {n}"]
    for layer in range(depth):
        for q in range(n):
            a, u, g = r.gen_f64() * 2.0 * math.pi, r.gen_f64(), r.gen_f64() * 2.0 * math.pi
            b = math.acos(1.0 - 2.0 * u)
            lines.append(f"RZ({g!r}) {q}")
            lines.append(f"RY({b!r}) {q}")
            lines.append(f"RZ({a!r}) {q}")
            lines.append(f"DEPOLARIZE1({p!r}) {q}")
        for a in range(layer % 2, n - 1, 2):
            lines.append(f"CX {a} {a + 1}")
            lines.append(f"DEPOLARIZE1({p!r}) {a} {a + 1}")
        if (layer + 1) % every == 0:
            lines.append("M " + " ".join(str(q) for q in mq))
            lines.append("R " + " ".join(str(q) for q in mq))
    return "\n".join(lines) + "\n"


def su4_brickwork(n, depth, seed):
    """Random two-qubit circuit on a line (quantum-volume style): per layer a
    Haar-random 4x4 unitary on each pair (2i, 2i+1) in even layers and
    (2i+1, 2i+2) in odd ones (complex Gaussian matrix from SplitMix64 draws by
    Box-Muller, QR, phases of R's diagonal divided out).  Returns a list of
    (qubits, matrix); bit k of the matrix index is qubits[k]."""
    import numpy as np

    r = SplitMix64(seed)

    def gauss():
        u1, u2 = 1.0 - r.gen_f64(), r.gen_f64()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    out = []
    for layer in range(depth):
        for a in range(layer % 2, n - 1, 2):
            z = np.array([[complex(gauss(), gauss()) for _ in range(4)] for _ in range(4)]) / math.sqrt(2.0)
            q, rr = np.linalg.qr(z)
            d = np.diagonal(rr)
            out.append(((a, a + 1), q * (d / np.abs(d))))
    return out


def build_circuit(n, gates, circuit_cls, gate_cls):
    """A circuit of (qubits, matrix) UNITARY gates through an API with
    `circuit_cls(n)` and `.gate(gate_cls("UNITARY", matrix=m), qubits)`."""
    c = circuit_cls(n)
    for qs, m in gates:
        c.gate(gate_cls("UNITARY", matrix=m), list(qs))
    return c


def repetition_code(d, rounds, p):
    """The reference's repetition-code memory experiment (frame.rs:401-442):
    data qubits 0..d-1, ancilla d + a measures Z_a Z_{a+1} with two CXs and
    is reset; X_ERROR(p) on every data qubit before each round; a round's
    detectors compare each ancilla with the previous round (the first round
    with nothing); the final data measurement's last bit is observable 0."""
    text = f"QUBITS {2 * d - 1}\n"
    for r in range(rounds):
        text += f"X_ERROR({p})" + "".join(f" {q}" for q in range(d)) + "\n"
        for a in range(d - 1):
            anc = d + a
            text += f"CX {a} {anc}\nCX {a + 1} {anc}\nM {anc}\nR {anc}\n"
        for a in range(d - 1):
            if r == 0:
                text += f"DETECTOR rec[-{d - 1 - a}]\n"
            else:
                text += f"DETECTOR rec[-{d - 1 - a}] rec[-{2 * (d - 1) - a}]\n"
    text += "M" + "".join(f" {q}" for q in range(d)) + "\nOBSERVABLE_INCLUDE(0) rec[-1]\n"
    return text


def count_gates(text):
    """Unfused gate applications of a text circuit (targets grouped by arity)."""
    ar = {"CX": 2, "CY": 2, "CZ": 2, "SWAP": 2, "CS": 2, "CS_DAG": 2, "CPHASE": 2, "CNOT": 2, "CCZ": 3, "CCX": 3,
          "TOFFOLI": 3}
    skip = {"QUBITS", "M", "MX", "MY", "MZ", "MPP", "R", "TICK", "DETECTOR", "OBSERVABLE_INCLUDE", "X_ERROR",
            "Y_ERROR", "Z_ERROR", "DEPOLARIZE1", "DEPOLARIZE2", "PAULI_CHANNEL_1"}
    g = 0
    for line in text.splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        toks = line.split()
        name = toks[0].split("(")[0].upper()
        if name in skip:
            continue
        g += (len(toks) - 1) // ar.get(name, 1)
    return g
