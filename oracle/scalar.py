"""ORACLE (test infrastructure only): ncsim's C64 scalar constants and exact
phasors, restated.

Follows src/scalar.rs: `C64::root_of_unity` (112-132), `C64::sqrt2_pow`
(133-140), `Amplitude::from_phasor` (39-52), `Amplitude::i_pow` (54-61) and
`Phasor::{new, mul}` (234-295).  Only tests, the smoke check and the CPU
baseline import this package.
"""
import math

FRAC_1_SQRT_2 = 0.7071067811865476  # std::f64::consts::FRAC_1_SQRT_2
SQRT_2 = 1.4142135623730951  # std::f64::consts::SQRT_2


def cmul(a, b):
    """The reference's complex product (textbook form, no FMA)."""
    return complex(a.real * b.real - a.imag * b.imag, a.real * b.imag + a.imag * b.real)


def quarter_turn(j):
    """scalar.rs:101-108."""
    return [complex(1.0, 0.0), complex(0.0, 1.0), complex(-1.0, 0.0), complex(0.0, -1.0)][j % 4]


def root_of_unity(k, bits):
    """C64::root_of_unity (scalar.rs:112-132): exact at multiples of pi/4,
    else from_polar(1, 2 pi k / 2^bits).  bits <= 62."""
    assert bits <= 62, "phase bits must be at most 62"
    m = 1 << bits
    k = k % m
    if (4 * k) % m == 0:
        return quarter_turn(4 * k // m)
    if (8 * k) % m == 0:
        h = FRAC_1_SQRT_2
        return {1: complex(h, h), 3: complex(-h, h), 5: complex(-h, -h)}.get(8 * k // m, complex(h, -h))
    theta = 2.0 * math.pi * float(k) / float(m)
    return complex(math.cos(theta), math.sin(theta))


def sqrt2_pow(k):
    """C64::sqrt2_pow (scalar.rs:133-140)."""
    half = k // 2
    v = 2.0 ** half
    if k % 2 == 1:
        v *= SQRT_2
    return complex(v, 0.0)


def i_pow(k):
    """Amplitude::i_pow (scalar.rs:54-61), with the reference's negated zeros."""
    return [complex(1.0, 0.0), complex(0.0, 1.0), complex(-1.0, -0.0), complex(-0.0, -1.0)][k % 4]


class Phasor:
    """sqrt2^p * zeta_{2^bits}^phase, normalized (scalar.rs:210-295); `None`
    stands for Phasor::Zero where needed."""

    __slots__ = ("p", "phase", "bits")

    def __init__(self, p, phase, bits):
        assert bits <= 62, "phase bits must be at most 62"
        m = 1 << bits
        phase %= m
        while bits > 0 and phase % 2 == 0:
            phase //= 2
            bits -= 1
        self.p, self.phase, self.bits = p, phase, bits

    @staticmethod
    def one():
        return Phasor(0, 0, 0)

    @staticmethod
    def root(phase, bits):
        return Phasor(0, phase, bits)

    @staticmethod
    def ipow(k):
        return Phasor(0, k, 2)

    def mul(self, o):
        b = max(self.bits, o.bits)
        f = (self.phase << (b - self.bits)) + (o.phase << (b - o.bits))
        return Phasor(self.p + o.p, f, b)

    def __eq__(self, o):
        return (self.p, self.phase, self.bits) == (o.p, o.phase, o.bits)

    def to_c64(self):
        """Amplitude::from_phasor for C64 (scalar.rs:39-52)."""
        return cmul(root_of_unity(self.phase, self.bits), sqrt2_pow(self.p))

    def is_one(self):
        return self.p == 0 and self.bits == 0
