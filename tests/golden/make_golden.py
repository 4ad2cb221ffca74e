"""Writes kat.json: known answers of small circuits, each computed here from
its closed form (no simulator involved), restating the reference's own
known-answer tests for the state-vector path:

- ghz3: vector.rs:580-633 (GHZ amplitudes 1/sqrt2 on |000> and |111>, Pauli
  expectations of the GHZ stabilizers);
- t_ccz: pathsum.rs:1256-1288 (T on |+>^3 then CCZ: zeta8^(|x| + 4[x=7]) / 2^(3/2));
- diag_sum: tests/backends.rs:109-136 (H^3 D_f H^3 |0>, f = x0 + 3 x1 x2 + 4 x0 x1 x2
  mod 8 as T, CPHASE(3,3) and CCZ; amplitude y = 2^-3 sum_x (-1)^(x.y) zeta8^f(x));
- rot_product: the rotation matrices of gate.rs:706-730 (RX(t) = [[c, -is], [-is, c]],
  RY(t) = [[c, -s], [s, c]], RZ(t) = diag(e^(-it/2), e^(it/2)), c = cos(t/2),
  s = sin(t/2)) on |000>: a product state with closed-form Pauli expectations.

    python tests/golden/make_golden.py
"""
import cmath
import json
import math
import os

Z8 = [cmath.exp(2j * math.pi * k / 8) for k in range(8)]


def ghz3():
    a = [0j] * 8
    a[0] = a[7] = 1 / math.sqrt(2)
    # <Z0 Z1> = 1, <X0 X1 X2> = 1, <Z0> = 0, <Y0 Y1 X2> = -1 (Y|0> = i|1>, Y|1> = -i|0>)
    ex = [("Z0*Z1", 1.0), ("X0*X1*X2", 1.0), ("Z0", 0.0), ("Y0*Y1*X2", -1.0)]
    return "QUBITS 3\nH 0\nCX 0 1\nCX 1 2\n", a, ex


def t_ccz():
    a = [Z8[(bin(x).count("1") + (4 if x == 7 else 0)) % 8] / 2 ** 1.5 for x in range(8)]
    return "QUBITS 3\nH 0 1 2\nT 0 1 2\nCCZ 0 1 2\n", a, []


def diag_sum():
    def f(x):
        x0, x1, x2 = x & 1, (x >> 1) & 1, (x >> 2) & 1
        return (x0 + 3 * x1 * x2 + 4 * x0 * x1 * x2) % 8

    a = [sum((-1) ** bin(x & y).count("1") * Z8[f(x)] for x in range(8)) / 8 for y in range(8)]
    return "QUBITS 3\nH 0 1 2\nT 0\nCPHASE(3,3) 1 2\nCCZ 0 1 2\nH 0 1 2\n", a, []


def rot_product():
    t0, t1, t2 = 0.3, 1.1, 2.5
    v0 = [math.cos(t0 / 2), -1j * math.sin(t0 / 2)]
    v1 = [math.cos(t1 / 2), math.sin(t1 / 2)]
    v2 = [cmath.exp(-0.5j * t2), 0j]
    a = [v0[x & 1] * v1[(x >> 1) & 1] * v2[(x >> 2) & 1] for x in range(8)]
    ex = [("Z0", math.cos(t0)), ("Y0", -math.sin(t0)), ("X1", math.sin(t1)), ("Z2", 1.0), ("X0*X1", 0.0),
          ("Z0*X1", math.cos(t0) * math.sin(t1))]
    return f"QUBITS 3\nRX({t0!r}) 0\nRY({t1!r}) 1\nRZ({t2!r}) 2\n", a, ex


def main():
    out = []
    for name, ref, fn in [("ghz3", "vector.rs:580-633", ghz3), ("t_ccz", "pathsum.rs:1256-1288", t_ccz),
                          ("diag_sum", "tests/backends.rs:109-136", diag_sum),
                          ("rot_product", "gate.rs:706-730", rot_product)]:
        text, amps, ex = fn()
        out.append({"name": name, "ref": ref, "text": text, "n": 3,
                    "amplitudes": [[z.real, z.imag] for z in amps],
                    "expectations": [[p, v, 0.0] for p, v in ex]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
