"""ORACLE (test infrastructure only): ncsim circuits and their text format,
restated from src/circuit.rs (Op 37-79, Noise 82-147, push checks 177-229,
parse 459-691, Display 693-728).

Ops are tuples: ("gate", Gate, qubits), ("measure", Pauli), ("reset", q),
("noise", (name, probs), qubits), ("detector", lookback),
("observable", index, lookback), ("tick",).
"""
from . import gates as G
from .pauli import Pauli, SimError


def noise_errors(name, probs):
    """Noise::errors (circuit.rs:119-128): (probability, x bits, z bits)."""
    if name == "X_ERROR":
        return [(probs[0], 1, 0)]
    if name == "Y_ERROR":
        return [(probs[0], 1, 1)]
    if name == "Z_ERROR":
        return [(probs[0], 0, 1)]
    if name == "PAULI_CHANNEL_1":
        a, b, c = probs
        return [(a, 1, 0), (b, 1, 1), (c, 0, 1)]
    if name == "DEPOLARIZE1":
        p = probs[0]
        return [(p / 3.0, 1, 0), (p / 3.0, 1, 1), (p / 3.0, 0, 1)]
    if name == "DEPOLARIZE2":
        p = probs[0]
        return [(p / 15.0, k & 3, k >> 2) for k in range(1, 16)]
    raise SimError("Parse", f"unknown channel {name}")


def noise_arity(name):
    return 2 if name == "DEPOLARIZE2" else 1


def noise_sample(name, probs, rng):
    """Noise::sample (circuit.rs:136-146): one uniform per application."""
    u = rng.gen_f64()
    for p, x, z in noise_errors(name, probs):
        if u < p:
            return x, z
        u -= p
    return None


def noise_name(name, probs):
    return f"{name}({','.join(repr(p) for p in probs)})"


def check_qubits(qubits, n):
    """error.rs:69-82."""
    for i, q in enumerate(qubits):
        if q >= n:
            raise SimError("QubitOutOfRange", f"qubit {q} out of range for {n} qubits")
        if q in qubits[:i]:
            raise SimError("DuplicateQubit", f"qubit {q} listed twice")


class Circuit:
    def __init__(self, n):
        self.n, self.ops = n, []

    def num_qubits(self):
        return self.n

    def num_measurements(self):
        return sum(1 for o in self.ops if o[0] == "measure")

    def num_gates(self):
        return sum(1 for o in self.ops if o[0] == "gate")

    def push(self, op):
        """Circuit::push (circuit.rs:177-229)."""
        kind = op[0]
        if kind == "gate":
            check_qubits(op[2], self.n)
            if len(op[2]) != op[1].arity():
                raise SimError("Arity", f"gate {op[1]!r} acts on {op[1].arity()} qubits, {len(op[2])} given")
        elif kind == "measure":
            if op[1].n != self.n:
                raise SimError("Invalid", "measurement on the wrong register")
            if not op[1].is_hermitian():
                raise SimError("Invalid", f"measured operator {op[1]!r} is not Hermitian")
        elif kind == "reset":
            check_qubits([op[1]], self.n)
        elif kind == "noise":
            check_qubits(op[2], self.n)
            if len(op[2]) != noise_arity(op[1][0]):
                raise SimError("Arity", "noise arity")
            errs = noise_errors(*op[1])
            if not (all(0.0 <= p <= 1.0 for p, _, _ in errs) and sum(p for p, _, _ in errs) <= 1.0 + 1e-12):
                raise SimError("Invalid", "not a probability distribution")
        elif kind in ("detector", "observable"):
            lb = op[-1]
            m = self.num_measurements()
            if any(k == 0 or k > m for k in lb):
                raise SimError("Invalid", f"rec[-k] with {m} measurements so far")
        self.ops.append(op)
        return self

    def gate(self, g, qubits):
        if isinstance(g, str):
            g = G.named(g)
        return self.push(("gate", g, list(qubits)))

    def measure_z(self, q):
        return self.push(("measure", Pauli.single(self.n, q, "Z")))

    def reset(self, q):
        return self.push(("reset", q))

    def detectors(self):
        """Circuit::detectors (circuit.rs:297-308): absolute measurement indices."""
        m, out = 0, []
        for o in self.ops:
            if o[0] == "measure":
                m += 1
            elif o[0] == "detector":
                out.append([m - k for k in o[1]])
        return out

    def observables(self):
        """Circuit::observables (circuit.rs:311-327): OBSERVABLE_INCLUDE(i) lists
        collected per index i."""
        m, out = 0, []
        for o in self.ops:
            if o[0] == "measure":
                m += 1
            elif o[0] == "observable":
                while len(out) <= o[1]:
                    out.append([])
                out[o[1]].extend(m - k for k in o[2])
        return out

    def without_noise(self):
        c = Circuit(self.n)
        c.ops = [o for o in self.ops if o[0] != "noise"]
        return c

    @staticmethod
    def parse(text):
        """Circuit::parse (circuit.rs:459-691)."""
        lines, declared = [], None
        for ln, raw in enumerate(text.splitlines()):
            line = raw.split("#")[0].strip()
            if not line:
                continue
            toks = line.split()
            if toks[0].upper() == "QUBITS":
                try:
                    declared = int(toks[1])
                except (IndexError, ValueError):
                    raise SimError("Parse", f"line {ln + 1}: QUBITS n")
                continue
            lines.append((ln, toks))
        parsed, max_q = [], 0
        for ln, toks in lines:
            try:
                p = _parse_line(toks)
            except SimError as e:
                raise SimError("Parse", f"line {ln + 1}: {e}")
            if p[0] in ("gate", "mbasis", "reset", "noise"):
                for q in p[-1]:
                    max_q = max(max_q, q + 1)
            elif p[0] == "mpp":
                for pp in p[1]:
                    max_q = max(max_q, pp.n)
            parsed.append((ln, p))
        n = declared if declared is not None else max_q
        if n < max_q:
            raise SimError("Parse", f"QUBITS {n} but qubit {max_q - 1} is used")
        c = Circuit(n)
        for ln, p in parsed:
            try:
                if p[0] == "gate":
                    g, qs = p[1], p[2]
                    k = g.arity()
                    if k == 0 or len(qs) % k:
                        raise SimError("Parse", f"{g!r} takes groups of {k} qubits")
                    for i in range(0, len(qs), k):
                        c.push(("gate", g, qs[i:i + k]))
                elif p[0] == "mbasis":
                    for q in p[2]:
                        c.push(("measure", Pauli.single(n, q, p[1])))
                elif p[0] == "mpp":
                    for pp in p[1]:
                        c.push(("measure", Pauli.parse(pp.to_sparse_string(), n)))
                elif p[0] == "reset":
                    for q in p[1]:
                        c.push(("reset", q))
                elif p[0] == "noise":
                    ch, qs = p[1], p[2]
                    k = noise_arity(ch[0])
                    if len(qs) % k:
                        raise SimError("Parse", f"{ch[0]} takes groups of {k} qubits")
                    for i in range(0, len(qs), k):
                        c.push(("noise", ch, qs[i:i + k]))
                elif p[0] == "detector":
                    c.push(("detector", p[1]))
                elif p[0] == "observable":
                    c.push(("observable", p[1], p[2]))
                else:
                    c.push(("tick",))
            except SimError as e:
                if e.kind == "Parse":
                    raise
                raise SimError("Parse", f"line {ln + 1}: {e}")
        return c

    def __str__(self):
        """Display (circuit.rs:693-728)."""
        out = [f"QUBITS {self.n}"]
        for o in self.ops:
            if o[0] == "gate":
                out.append(f"{o[1]!r} {' '.join(map(str, o[2]))}")
            elif o[0] == "measure":
                p = o[1]
                if p.weight() == 1 and p.display_phase() == 0:
                    q = p.support()[0]
                    out.append(f"M{p.kind(q)} {q}")
                else:
                    out.append(f"MPP {p.to_sparse_string()}")
            elif o[0] == "reset":
                out.append(f"R {o[1]}")
            elif o[0] == "noise":
                out.append(f"{noise_name(*o[1])} {' '.join(map(str, o[2]))}")
            elif o[0] == "detector":
                out.append("DETECTOR " + " ".join(f"rec[-{k}]" for k in o[1]))
            elif o[0] == "observable":
                out.append(f"OBSERVABLE_INCLUDE({o[1]}) " + " ".join(f"rec[-{k}]" for k in o[2]))
            else:
                out.append("TICK")
        return "\n".join(out) + "\n"


def _parse_args(s):
    i = s.find("(")
    if i < 0:
        return s.upper(), []
    rest = s[i + 1:]
    close = rest.rfind(")")
    if close < 0:
        raise SimError("Parse", f"missing ')' in {s}")
    return s[:i].upper(), [a.strip() for a in rest[:close].split(",")]


def _parse_line(toks):
    """parse_line (circuit.rs:557-691)."""
    name, args = _parse_args(toks[0])

    def qubits():
        try:
            return [int(t) for t in toks[1:]]
        except ValueError:
            raise SimError("Parse", f"bad qubit in {toks[1:]}")

    def fl(i):
        try:
            return float(args[i])
        except (IndexError, ValueError):
            raise SimError("Parse", f"{name} needs a numeric argument")

    def it(i):
        try:
            return int(args[i])
        except (IndexError, ValueError):
            raise SimError("Parse", f"{name} needs integer argument {i + 1}")

    simple = {"I", "X", "Y", "Z", "H", "S", "S_DAG", "SQRT_X", "SQRT_X_DAG", "SQRT_Y", "SQRT_Y_DAG",
              "T", "T_DAG", "CX", "CNOT", "CY", "CZ", "SWAP", "CS", "CS_DAG", "CCZ", "CCX", "TOFFOLI"}
    if name in simple:
        return ("gate", G.named(name), qubits())
    if name in ("PHASE", "CPHASE"):
        return ("gate", G.Gate(name, k=it(0), bits=it(1)), qubits())
    if name in ("RX", "RY", "RZ"):
        return ("gate", G.Gate(name, theta=fl(0)), qubits())
    if name in ("M", "MZ"):
        return ("mbasis", "Z", qubits())
    if name == "MX":
        return ("mbasis", "X", qubits())
    if name == "MY":
        return ("mbasis", "Y", qubits())
    if name == "MPP":
        return ("mpp", [Pauli.parse(t) for t in toks[1:]])
    if name == "R":
        return ("reset", qubits())
    if name in ("DETECTOR", "OBSERVABLE_INCLUDE"):
        recs = []
        for t in toks[1:]:
            if not (t.startswith("rec[-") and t.endswith("]")):
                raise SimError("Parse", f"bad target '{t}' (expected rec[-k])")
            recs.append(int(t[5:-1]))
        if name == "DETECTOR":
            return ("detector", recs)
        return ("observable", it(0), recs)
    if name in ("X_ERROR", "Y_ERROR", "Z_ERROR", "DEPOLARIZE1", "DEPOLARIZE2"):
        return ("noise", (name, (fl(0),)), qubits())
    if name == "PAULI_CHANNEL_1":
        return ("noise", (name, (fl(0), fl(1), fl(2))), qubits())
    if name == "TICK":
        return ("tick",)
    raise SimError("Parse", f"unknown instruction '{name}'")
