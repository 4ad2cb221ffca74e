"""ORACLE (test infrastructure only): Pauli-frame sampling of noisy Clifford
circuits, restated from src/frame.rs (FrameSampler: Clifford check 42-52,
reference record 54-56, frame updates 104-147, noise injection 307-343,
detectors 150-211) and src/pauli/columns.rs (column layout, gate updates
338-405, anticommutation 222-232, masked Pauli products 147-159).

The reference draws its randomness from rand's StdRng; the sampler here, and
the HIP kernels it checks, draw from a documented counter-based generator
instead, so that a GPU run and this restatement give the same bits:

    mix(v)          = SplitMix64's output step on v + 0x9E3779B97F4A7C15
    rnd(seed, a, b) = mix(mix(seed ^ mix(a)) ^ b)

with `a` naming the draw site (high byte: 1 initial Z randomization of qubit
q, 2 measurement k, 3 reset r, 4 noise application j: hit, 5 noise
application j: which error) and `b` the shot word (64 shots) or the shot.
A noise application hits shot s when rnd(seed, 4<<56 | j, s) is below
floor(total * 2^64) (always when total >= 1), and picks the first error whose
cumulative probability, divided by the total and scaled by 2^64, exceeds
rnd(seed, 5<<56 | j, s).  Those are the reference's Bernoulli(total) hit
and `u * total` walk (frame.rs:307-343) with integer draws.

The noiseless reference record comes from a stabilizer tableau (Aaronson and
Gottesman's CHP with general Pauli measurements; the reference uses its own
stabilizer state, src/state/stabilizer.rs), its random outcomes taking the top
bit of successive SplitMix64 outputs of `reference_seed`.

Rows and frames follow the reference's convention P = i^phase X^x Z^z.
"""
import numpy as np

from .circuit import noise_errors
from .pauli import Pauli, SimError

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15
C1 = 0xBF58476D1CE4E5B9
C2 = 0x94D049BB133111EB
SITE_INIT, SITE_MEAS, SITE_RESET, SITE_HIT, SITE_WHICH = 1, 2, 3, 4, 5


def mix(v):
    """SplitMix64's output step on v + gamma (Python ints)."""
    z = (v + GAMMA) & M64
    z = ((z ^ (z >> 30)) * C1) & M64
    z = ((z ^ (z >> 27)) * C2) & M64
    return z ^ (z >> 31)


def mix_np(v):
    """The same on a uint64 array."""
    with np.errstate(over="ignore"):
        z = v + np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(C1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(C2)
    return z ^ (z >> np.uint64(31))


def rnd(seed, a, b):
    """rnd(seed, a, b) for an array of b (uint64)."""
    k = np.uint64(mix((seed ^ mix(a)) & M64))
    return mix_np(k ^ np.asarray(b, dtype=np.uint64))


def site(kind, index):
    return (kind << 56) | index


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & M64

    def next_u64(self):
        self.s = (self.s + GAMMA) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * C1) & M64
        z = ((z ^ (z >> 27)) * C2) & M64
        return z ^ (z >> 31)


# ------------------------------------------------------------ Clifford views
# Images of X and Z of each local qubit under conjugation, as signed Pauli
# strings (gate.rs:587-604), for the named Clifford gates.
NAMED_IMAGES = {
    "I": (["X0"], ["Z0"]), "X": (["X0"], ["-Z0"]), "Y": (["-X0"], ["-Z0"]), "Z": (["-X0"], ["Z0"]),
    "H": (["Z0"], ["X0"]), "S": (["Y0"], ["Z0"]), "S_DAG": (["-Y0"], ["Z0"]),
    "SQRT_X": (["X0"], ["-Y0"]), "SQRT_X_DAG": (["X0"], ["Y0"]),
    "SQRT_Y": (["-Z0"], ["X0"]), "SQRT_Y_DAG": (["Z0"], ["-X0"]),
    "CX": (["X0*X1", "X1"], ["Z0", "Z0*Z1"]), "CY": (["X0*Y1", "Z0*X1"], ["Z0", "Z0*Z1"]),
    "CZ": (["X0*Z1", "Z0*X1"], ["Z0", "Z1"]), "SWAP": (["X1", "X0"], ["Z1", "Z0"]),
}


def parse_signed(s, k):
    """A signed Pauli string like "-X0*Y1" on k qubits as (x, z, phase)."""
    ph = 0
    if s.startswith("-"):
        ph, s = 2, s[1:]
    x = z = 0
    for f in s.split("*"):
        letter, q = f[0], int(f[1:])
        if letter in "XY":
            x |= 1 << q
        if letter in "YZ":
            z |= 1 << q
        if letter == "Y":
            ph += 1  # Y = i X Z
    return x, z, ph & 3


def pmul(a, b):
    """(i^pa X^xa Z^za)(i^pb X^xb Z^zb) = i^(pa+pb+2 popcount(za & xb)) X^(xa^xb) Z^(za^zb)."""
    xa, za, pa = a
    xb, zb, pb = b
    return xa ^ xb, za ^ zb, (pa + pb + 2 * bin(za & xb).count("1")) & 3


def clifford_images(g):
    """Images of X_j, Z_j (each (x, z, phase) on the gate's qubits) for a
    Clifford gate, or None (T, CS, CCZ, rotations, dense unitaries and
    diagonal phases that are not multiples of i are not Clifford)."""
    name = g.name
    if name in NAMED_IMAGES:
        xs, zs = NAMED_IMAGES[name]
        k = len(xs)
        return [parse_signed(s, k) for s in xs], [parse_signed(s, k) for s in zs]
    if name == "PAULI":
        p = g.pauli
        xs, zs = [], []
        for q in range(p.n):
            # conjugation by P flips the sign of the generators P anticommutes with
            xs.append((1 << q, 0, 2 if (p.z >> q) & 1 else 0))
            zs.append((0, 1 << q, 2 if (p.x >> q) & 1 else 0))
        return xs, zs
    if name in ("PHASE", "CPHASE"):
        # diag phase zeta_{2^bits}^k on |1> (|11>): a power of i when 4k = 0 mod 2^bits
        m = 1 << g.bits
        if (4 * g.k) % m:
            return None
        quarter = ((g.k % m) * 4) // m
        if name == "PHASE":
            return clifford_images_named({0: "I", 1: "S", 2: "Z", 3: "S_DAG"}[quarter])
        if quarter == 0:
            return [(1, 0, 0), (2, 0, 0)], [(0, 1, 0), (0, 2, 0)]
        if quarter == 2:
            return clifford_images_named("CZ")
        return None  # a controlled +-i phase (CS) is not Clifford
    return None


def clifford_images_named(name):
    xs, zs = NAMED_IMAGES[name]
    k = len(xs)
    return [parse_signed(s, k) for s in xs], [parse_signed(s, k) for s in zs]


def local_image(images, x, z):
    """Image of X^x Z^z (local bits) under the Clifford: the product of the
    generators' images, X's first, in qubit order."""
    xs, zs = images
    acc = (0, 0, 0)
    for j in range(len(xs)):
        if (x >> j) & 1:
            acc = pmul(acc, xs[j])
    for j in range(len(zs)):
        if (z >> j) & 1:
            acc = pmul(acc, zs[j])
    return acc


# ------------------------------------------------------------ tableau (CHP)
class Tableau:
    """A stabilizer state on n qubits as 2n rows (destabilizers, then
    stabilizers), each (x, z, phase) with Python-int masks."""

    def __init__(self, n):
        self.n = n
        self.rows = [(1 << q, 0, 0) for q in range(n)] + [(0, 1 << q, 0) for q in range(n)]

    def apply(self, images, qubits):
        out = []
        for (x, z, ph) in self.rows:
            lx = lz = 0
            for j, q in enumerate(qubits):
                lx |= ((x >> q) & 1) << j
                lz |= ((z >> q) & 1) << j
            if lx == 0 and lz == 0:
                out.append((x, z, ph))
                continue
            ix, iz, c = local_image(images, lx, lz)
            for j, q in enumerate(qubits):
                x = (x & ~(1 << q)) | (((ix >> j) & 1) << q)
                z = (z & ~(1 << q)) | (((iz >> j) & 1) << q)
            out.append((x, z, (ph + c) & 3))
        self.rows = out

    @staticmethod
    def anticommutes(r, p):
        return (bin((r[0] & p[1]) ^ (r[1] & p[0])).count("1") & 1) == 1

    def measure(self, p, draw):
        """Measures the Hermitian Pauli p = (x, z, phase); `draw()` gives the
        outcome bit of a random measurement.  Returns the outcome (1 = -1)."""
        n = self.n
        rows = self.rows
        piv = next((i for i in range(n, 2 * n) if self.anticommutes(rows[i], p)), None)
        if piv is not None:
            for j in range(2 * n):
                if j != piv and self.anticommutes(rows[j], p):
                    rows[j] = pmul(rows[j], rows[piv])
            rows[piv - n] = rows[piv]
            o = draw()
            rows[piv] = (p[0], p[1], (p[2] + 2 * o) & 3)
            return o
        acc = (0, 0, 0)
        for i in range(n):
            if self.anticommutes(rows[i], p):
                acc = pmul(acc, rows[i + n])
        assert acc[0] == p[0] and acc[1] == p[1], "tableau: deterministic outcome is not +-P"
        d = (acc[2] - p[2]) & 3
        assert d in (0, 2)
        return 1 if d == 2 else 0


def check_clifford(circuit):
    for op in circuit.ops:
        if op[0] == "gate" and clifford_images(op[1]) is None:
            raise SimError("Unsupported", f"non-Clifford gate {op[1]!r}")


def reference_record(circuit, reference_seed):
    """The noiseless run on a tableau (frame.rs:54-56)."""
    t = Tableau(circuit.n)
    rng = SplitMix64(reference_seed)
    draw = lambda: rng.next_u64() >> 63  # noqa: E731
    rec = []
    for op in circuit.ops:
        if op[0] == "gate":
            t.apply(clifford_images(op[1]), op[2])
        elif op[0] == "measure":
            p = op[1]
            rec.append(t.measure((p.x, p.z, p.phase), draw))
        elif op[0] == "reset":
            q = op[1]
            if t.measure((0, 1 << q, 0), draw):
                t.apply(clifford_images_named("X"), [q])
    return rec


# ------------------------------------------------------------ frames
def frame_matrix(images):
    """The GF(2) map of the gate on frame bits (signs dropped): out bit list
    per input generator."""
    xs, zs = images
    return [(ix, iz) for ix, iz, _ in xs], [(ix, iz) for ix, iz, _ in zs]


class FrameSampler:
    """FrameSampler (frame.rs:31-150) with the counter-based draws above."""

    def __init__(self, circuit, reference_seed):
        check_clifford(circuit)
        self.circuit = circuit
        self.reference = reference_record(circuit, reference_seed)

    def sample_packed(self, shots, seed):
        """Records packed by measurement: out[k] is a uint64 array of
        ceil(shots/64) words, bit s of word w = shot 64 w + s."""
        c = self.circuit
        n = c.n
        W = (shots + 63) // 64
        words = np.arange(W, dtype=np.uint64)
        last = np.full(W, M64, dtype=np.uint64)
        if shots % 64:
            last[-1] = np.uint64((1 << (shots % 64)) - 1)
        x = np.zeros((n, W), dtype=np.uint64)
        z = np.zeros((n, W), dtype=np.uint64)
        for q in range(n):
            z[q] = rnd(seed, site(SITE_INIT, q), words) & last
        out = []
        k = r = j = 0
        shot_ids = np.arange(W * 64, dtype=np.uint64)
        for op in c.ops:
            kind = op[0]
            if kind == "gate":
                xs, zs = frame_matrix(clifford_images(op[1]))
                qs = op[2]
                nx = [np.zeros(W, dtype=np.uint64) for _ in qs]
                nz = [np.zeros(W, dtype=np.uint64) for _ in qs]
                for jj, q in enumerate(qs):
                    for src, (ix, iz) in ((x[q], xs[jj]), (z[q], zs[jj])):
                        for t in range(len(qs)):
                            if (ix >> t) & 1:
                                nx[t] ^= src
                            if (iz >> t) & 1:
                                nz[t] ^= src
                for t, q in enumerate(qs):
                    x[q], z[q] = nx[t], nz[t]
            elif kind == "measure":
                p = op[1]
                rec = np.zeros(W, dtype=np.uint64)
                for q in range(n):
                    if (p.z >> q) & 1:
                        rec ^= x[q]
                    if (p.x >> q) & 1:
                        rec ^= z[q]
                if self.reference[k]:
                    rec = ~rec
                out.append(rec & last)
                m = rnd(seed, site(SITE_MEAS, k), words) & last
                for q in range(n):
                    if (p.x >> q) & 1:
                        x[q] ^= m
                    if (p.z >> q) & 1:
                        z[q] ^= m
                k += 1
            elif kind == "reset":
                q = op[1]
                x[q] = 0
                z[q] = rnd(seed, site(SITE_RESET, r), words) & last
                r += 1
            elif kind == "noise":
                name, probs = op[1]
                errs = noise_errors(name, probs)
                total = 0.0
                for pe, _, _ in errs:
                    total += pe
                if total > 0.0:
                    u = rnd(seed, site(SITE_HIT, j), shot_ids)
                    hit = np.ones(W * 64, dtype=bool) if total >= 1.0 else u < np.uint64(int(total * 2.0 ** 64))
                    hit[shots:] = False
                    thr = []
                    acc = 0.0
                    for pe, _, _ in errs:
                        acc += pe
                        f = acc / total
                        thr.append(M64 if f >= 1.0 else int(f * 2.0 ** 64))
                    v = rnd(seed, site(SITE_WHICH, j), shot_ids)
                    pick = np.full(W * 64, len(errs) - 1)
                    for i in reversed(range(len(errs))):
                        pick = np.where(v < np.uint64(thr[i]), i, pick)
                    for s in np.nonzero(hit)[0]:
                        _, ex, ez = errs[pick[s]]
                        w, b = s // 64, np.uint64(1 << (s % 64))
                        for jj, q in enumerate(op[2]):
                            if (ex >> jj) & 1:
                                x[q, w] ^= b
                            if (ez >> jj) & 1:
                                z[q, w] ^= b
                j += 1
        return out

    def sample(self, shots, seed):
        """One record per shot (lists of bools)."""
        packed = self.sample_packed(shots, seed)
        return [[bool((int(packed[k][s // 64]) >> (s % 64)) & 1) for k in range(len(packed))] for s in range(shots)]

    def sample_detectors(self, shots, seed):
        """Detector and observable flips (frame.rs:150-211), packed."""
        c = self.circuit
        packed = self.sample_packed(shots, seed)
        W = (shots + 63) // 64
        last = np.full(W, M64, dtype=np.uint64)
        if shots % 64:
            last[-1] = np.uint64((1 << (shots % 64)) - 1)

        def form(indices):
            b = np.zeros(W, dtype=np.uint64)
            flip = False
            for kk in indices:
                b ^= packed[kk]
                flip ^= bool(self.reference[kk])
            if flip:
                b = ~b
            return b & last

        return [form(d) for d in c.detectors()], [form(o) for o in c.observables()]


def parity(packed, indices):
    """frame.rs:346-357."""
    out = np.zeros_like(packed[0])
    for k in indices:
        out ^= packed[k]
    return out
