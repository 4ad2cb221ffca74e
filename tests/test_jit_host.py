"""CPU check of the specialized sweep kernels: the HIP source the generator
writes for every sweep of a plan is compiled for the host with g++ (HIP's
thread and block indices, __syncthreads and __shared__ mapped onto host
threads and a barrier), run one workgroup at a time on 256 host threads,
and compared with the plan's FP64 host emulation (ncsv_plan_emulate).  This
exercises the generator's register renaming, every LDS exchange form
(half-tile passes on a shared register position, wavefront turns,
intermediate layouts, whole-tile), the lazily applied diagonal factors and
their thread and block-index tables, without a GPU.  The grouped expectation
kernels run the same way, with __shfl_down emulated per wavefront."""
import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import ncsim_amd as H
from ncsim_amd import _lib as L
from ncsim_amd import workloads as W

PRE = r'''
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
struct double2 { double x, y; };
struct float2 { float x, y; };
struct TI { unsigned x; };
struct BI { unsigned long long x; };
thread_local TI threadIdx;
thread_local BI blockIdx;
static std::barrier<> *g_bar;
#define __syncthreads() g_bar->arrive_and_wait()
#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__
#define __shared__ static
#define __popcll __builtin_popcountll
#define __builtin_amdgcn_sched_barrier(x)
inline int __double2hiint(double d) { long long b; memcpy(&b, &d, 8); return (int)(b >> 32); }
inline int __double2loint(double d) { long long b; memcpy(&b, &d, 8); return (int)(unsigned)(b & 0xffffffffu); }
inline double __hiloint2double(int hi, int lo) {
    long long b = ((long long)hi << 32) | (unsigned)lo; double d; memcpy(&d, &b, 8); return d; }
inline int __float_as_int(float f) { int i; memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; memcpy(&f, &i, 4); return f; }
// the wavefront operations of the MFMA blocks on host threads (lane = tid & 63,
// wavefront = tid >> 6); every thread of the workgroup makes the same calls in
// the same order, so the workgroup barrier orders each exchange
struct f16v { float v[16]; float &operator[](int i) { return v[i]; } };
static float g_wa[4][64], g_wb[4][64];
inline f16v mfma32(float a, float b, f16v c) {
    // v_mfma_f32_32x32x2f32: A[m][k] in lane m + 32k, B[k][n] in lane n + 32k,
    // D register i of lane l = row 8(i/4) + 4(l/32) + i%4, column l%32
    const unsigned w = threadIdx.x >> 6, l = threadIdx.x & 63;
    g_wa[w][l] = a;
    g_wb[w][l] = b;
    g_bar->arrive_and_wait();
    f16v d = c;
    for (int i = 0; i < 16; i++) {
        const unsigned M = 8 * (i / 4) + 4 * (l / 32) + i % 4, n = l % 32;
        d[i] = g_wa[w][M] * g_wb[w][n] + g_wa[w][M + 32] * g_wb[w][n + 32] + d[i];
    }
    g_bar->arrive_and_wait();
    return d;
}
inline void swap32(float &x, float &y) {
    // v_permlane32_swap: (x lanes 0-31 | y lanes 0-31), (x lanes 32-63 | y lanes 32-63)
    const unsigned w = threadIdx.x >> 6, l = threadIdx.x & 63;
    g_wa[w][l] = x;
    g_wb[w][l] = y;
    g_bar->arrive_and_wait();
    x = l < 32 ? g_wa[w][l] : g_wb[w][l - 32];
    y = l < 32 ? g_wa[w][l + 32] : g_wb[w][l];
    g_bar->arrive_and_wait();
}
'''

POST = r'''
extern "C" void run_host(V *amp, unsigned long long nblocks, unsigned nthreads, const T *coef) {
    for (unsigned long long b = 0; b < nblocks; b++) {
        std::barrier<> bar(nthreads);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; t++)
            th.emplace_back([=] { threadIdx.x = t; blockIdx.x = b; ncsv_sweep(amp, coef, amp); });
        for (auto &x : th) x.join();
    }
}
extern "C" void run_host_oop(V *in, V *out, unsigned long long nblocks, unsigned nthreads, const T *coef) {
    for (unsigned long long b = 0; b < nblocks; b++) {
        std::barrier<> bar(nthreads);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; t++)
            th.emplace_back([=] { threadIdx.x = t; blockIdx.x = b; ncsv_sweep(in, coef, out); });
        for (auto &x : th) x.join();
    }
}
'''


HOST_NT = "inline V ld_nt(const V *p) { return *p; }\ninline void st_nt(V *p, V v) { *p = v; }"


def host_kernel(src, workdir, name, post=None):
    lines = []
    for line in src.splitlines():
        if ("asm volatile" in line or "ext_vector_type" in line or "__builtin_nontemporal" in line
                or "// @mfma" in line):
            continue  # the nontemporal and MFMA helpers become the host versions
        lines.append(line)
        if line.startswith("typedef") and line.endswith(" V;") and "ld_nt(" in src:
            lines.append(HOST_NT)
    src = "\n".join(lines)
    src = src.replace('extern "C" ', "")
    cpp = os.path.join(workdir, name + ".cpp")
    with open(cpp, "w") as f:
        f.write(PRE + src + (POST if post is None else post))
    so = os.path.join(workdir, name + ".so")
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-shared", "-fPIC", "-o", so, cpp, "-lpthread"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[:4000]
    return ctypes.CDLL(so)


def run_plan_on_host(text, precision, monkeypatch, half, env=None, oop=False):
    """Every sweep kernel of the circuit's plan, in order, on the host;
    returns (host result, emulated result, plan stats).  oop: each sweep
    reads one buffer and writes another that starts as NaN."""
    work = tempfile.mkdtemp()
    try:
        monkeypatch.setenv("NCSV_JIT_DUMP", work)
        monkeypatch.setenv("NCSV_JIT_HALF", str(half))
        for k, v in (env or {}).items():
            monkeypatch.setenv(k, v)
        c = text if isinstance(text, H.Circuit) else H.Circuit.parse(text)
        n = c.num_qubits()
        k = ctypes.c_uint32()
        assert L.lib().ncsv_jit_check(c._h, int(precision), b"gfx950", ctypes.byref(k)) == 0
        order = [line.split() for line in open(os.path.join(work, "order.txt")).read().splitlines()]
        hashes = [h for h, from_basis in order if from_basis == "0"]
        assert len(hashes) == k.value
        rng = np.random.default_rng(n)
        a = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
        want = a.copy()
        stats = (ctypes.c_uint64 * 4)()
        assert L.lib().ncsv_plan_emulate(c._h, int(precision), want.view(np.float64).ctypes.data_as(L.c_dp),
                                         stats) == 0
        buf = a.copy() if precision == H.C128 else a.astype(np.complex64)
        tile = 12 if precision == H.C128 else 13
        for i, h in enumerate(hashes):
            src = open(os.path.join(work, f"ncsv_{h}.hip")).read()
            coef = np.fromfile(os.path.join(work, f"ncsv_{h}.coef"), dtype=np.float64)
            coef = np.ascontiguousarray(coef if precision == H.C128 else coef.astype(np.float32))
            lib = host_kernel(src, work, f"k{i}")
            if oop:
                out = np.full_like(buf, np.nan)
                lib.run_host_oop(buf.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_ulonglong(1 << (n - tile)), ctypes.c_uint(256),
                                 coef.ctypes.data_as(ctypes.c_void_p))
                buf = out
            else:
                lib.run_host(buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_ulonglong(1 << (n - tile)),
                             ctypes.c_uint(256), coef.ctypes.data_as(ctypes.c_void_p))
        return buf.astype(np.complex128), want, list(stats)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def outer_heavy(n, tile, seed):
    """Gates on the low `tile` qubits plus phases and CZs reaching the
    others: one sweep on a multi-tile state whose diagonal factors depend on
    block-index bits."""
    r = np.random.default_rng(seed)
    lines = [f"QUBITS {n}"]
    for _ in range(4):
        for q in range(tile):
            lines.append(f"{['H', 'RX(0.3)', 'RY(1.1)', 'SQRT_X', 'T'][r.integers(5)]} {q}")
        for q in range(0, tile - 1, 2):
            lines.append(f"{'CX' if r.random() < 0.5 else 'CZ'} {q} {q + 1}")
        for q in range(tile, n):
            lines.append(f"RZ({r.random() * 6!r}) {q}")
            lines.append(f"CZ {q} {int(r.integers(tile))}")
            lines.append(f"CPHASE(1,{int(r.integers(2, 30))}) {q} {int(r.integers(tile))}")
    return "\n".join(lines) + "\n"


CASES = [
    ("brickwork", H.C128, lambda: W.rotation_brickwork(12, 3, 0), 1),
    ("qft", H.C128, lambda: W.qft_with_rz(14, 12, 1), 1),
    ("qft-c64", H.C64, lambda: W.qft_with_rz(15, 12, 2), 1),
    ("brickwork-c64", H.C64, lambda: W.rotation_brickwork(15, 4, 3), 1),
    ("outer", H.C128, lambda: outer_heavy(21, 12, 4), 1),
    ("outer-c64", H.C64, lambda: outer_heavy(15, 13, 5), 1),
    ("wavefront-turns", H.C128, lambda: W.rotation_brickwork(14, 4, 6), 2),
    ("whole-tile", H.C64, lambda: W.qft_with_rz(13, 8, 7), 0),
]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with C++20")
@pytest.mark.parametrize("name,precision,make,half", CASES, ids=[c[0] for c in CASES])
def test_generated_sweeps_on_host_threads(name, precision, make, half, monkeypatch):
    got, want, stats = run_plan_on_host(make(), precision, monkeypatch, half)
    tol = 1e-13 if precision == H.C128 else 2e-6
    assert np.abs(got - want).max() < tol, (name, stats)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with C++20")
@pytest.mark.parametrize("precision,make", [(H.C128, lambda: outer_heavy(16, 12, 4)),
                                            (H.C64, lambda: W.rotation_brickwork(17, 4, 3))], ids=["c128", "c64"])
def test_sweeps_out_of_place_on_host_threads(precision, make, monkeypatch):
    """Each sweep reading one buffer and writing another (initialised to NaN)
    writes every amplitude and computes the in-place result."""
    got, want, stats = run_plan_on_host(make(), precision, monkeypatch, 1, oop=True)
    assert not np.isnan(got).any()
    tol = 1e-13 if precision == H.C128 else 2e-6
    assert np.abs(got - want).max() < tol, stats


# memory shapes: loads and stores in other register layouts (the candidates
# the layout timing chooses from, including different load and store
# layouts) and the XCD-chunked block order; every shape computes the same
# state.  The block order is a compile-time constant, so each case runs in a
# fresh process.
SHAPES = [
    ("outer-load2", H.C128, lambda: outer_heavy(16, 12, 4), {"NCSV_JIT_MEMSHAPE": "2"}),
    ("outer-load3-store5-border", H.C128, lambda: outer_heavy(16, 12, 4),
     {"NCSV_JIT_MEMSHAPE": "3", "NCSV_JIT_MEMSHAPE_STORE": "5", "NCSV_JIT_BORDER": "1"}),
    ("brickwork-c64-first-last", H.C64, lambda: W.rotation_brickwork(17, 4, 3),
     {"NCSV_JIT_MEMSHAPE": "5", "NCSV_JIT_MEMSHAPE_STORE": "6", "NCSV_JIT_BORDER": "1"}),
    ("qft-c64-load1-border", H.C64, lambda: W.qft_with_rz(17, 12, 2), {"NCSV_JIT_MEMSHAPE": "1", "NCSV_JIT_BORDER": "1"}),
    ("brickwork-load4-store1", H.C128, lambda: W.rotation_brickwork(16, 4, 6),
     {"NCSV_JIT_MEMSHAPE": "4", "NCSV_JIT_MEMSHAPE_STORE": "1"}),
]


def _shape_worker(name, q):
    try:
        case = {c[0]: c for c in SHAPES}[name]
        _, precision, make, env = case

        class MP:  # a minimal monkeypatch for the child process
            def setenv(self, k, v):
                os.environ[k] = v

        got, want, stats = run_plan_on_host(make(), precision, MP(), 1, env)
        q.put(float(np.abs(got - want).max()))
    except BaseException as e:  # noqa: BLE001
        q.put(repr(e))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with C++20")
@pytest.mark.parametrize("name", [c[0] for c in SHAPES])
def test_generated_sweeps_in_other_memory_shapes(name):
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=_shape_worker, args=(name, q))
    p.start()
    r = q.get(timeout=600)
    p.join(60)
    assert not isinstance(r, str), r
    precision = {c[0]: c[1] for c in SHAPES}[name]
    assert r < (1e-13 if precision == H.C128 else 2e-6), (name, r)


# FP32 MFMA dense blocks in complex64 sweeps: a random two-qubit circuit takes
# them by cost; forcing them (NCSV_JIT_MFMA=2) on the rotation brickwork and
# the QFT composes real-form 2x2s, controlled flips, swaps and phases into
# blocks.  The kernels' v_mfma_f32_32x32x2f32 and v_permlane32_swap run as
# wavefront exchanges on the host threads.
def su4_circuit(n, depth, seed):
    return W.build_circuit(n, W.su4_brickwork(n, depth, seed), H.Circuit, H.Gate)


MFMA_CASES = [
    ("su4-auto", lambda: su4_circuit(14, 6, 1), {"NCSV_JIT_MFMA": "1"}),
    ("su4-multi-tile", lambda: su4_circuit(15, 4, 2), {"NCSV_JIT_MFMA": "2"}),
    ("brickwork-forced", lambda: W.rotation_brickwork(14, 3, 3), {"NCSV_JIT_MFMA": "2"}),
    ("qft-forced", lambda: W.qft_with_rz(14, 8, 2), {"NCSV_JIT_MFMA": "2"}),
    ("outer-forced", lambda: outer_heavy(14, 13, 5), {"NCSV_JIT_MFMA": "2"}),
]


def _mfma_worker(name, q):
    try:
        _, make, env = {c[0]: c for c in MFMA_CASES}[name]
        os.environ.update(env)

        class MP:
            def setenv(self, k, v):
                os.environ[k] = v

        work = tempfile.mkdtemp()
        os.environ["NCSV_JIT_DUMP"] = work
        c = make()
        got, want, stats = run_plan_on_host(c, H.C64, MP(), 1)
        q.put(float(np.abs(got - want).max()))
    except BaseException as e:  # noqa: BLE001
        import traceback
        q.put(traceback.format_exc())


def _mfma_blocks_in(circuit_or_text, env):
    """MFMA blocks the generator writes for the circuit's complex64 plan."""
    import subprocess as sp
    import sys
    code = (
        "import os, ctypes, glob, sys\n"
        "import ncsim_amd as H\nfrom ncsim_amd import _lib as L, workloads as W\n"
        "c = W.build_circuit(15, W.su4_brickwork(15, 6, 1), H.Circuit, H.Gate) if sys.argv[1] == 'su4' "
        "else H.Circuit.parse(W.rotation_brickwork(15, 4, 3))\n"
        "k = ctypes.c_uint32()\n"
        "assert L.lib().ncsv_jit_check(c._h, int(H.C64), b'gfx950', ctypes.byref(k)) == 0\n"
        "print(sum(open(f).read().count('MFMA block') for f in glob.glob(os.environ['NCSV_JIT_DUMP'] + '/*.hip')))\n")
    work = tempfile.mkdtemp()
    try:
        e = dict(os.environ, NCSV_JIT_DUMP=work, **env)
        r = sp.run([sys.executable, "-c", code, circuit_or_text], env=e, capture_output=True, text=True,
                   cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        assert r.returncode == 0, r.stderr[-2000:]
        return int(r.stdout.strip().splitlines()[-1])
    finally:
        shutil.rmtree(work, ignore_errors=True)


def test_mfma_blocks_chosen_by_cost():
    # by cost: generic 4x4s get blocks, rotations (two instructions per
    # amplitude each) none; off by default; forced on every run
    assert _mfma_blocks_in("su4", {"NCSV_JIT_MFMA": "1"}) > 0
    assert _mfma_blocks_in("brickwork", {"NCSV_JIT_MFMA": "1"}) == 0
    assert _mfma_blocks_in("su4", {}) == 0
    assert _mfma_blocks_in("brickwork", {"NCSV_JIT_MFMA": "2"}) > 0


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with C++20")
@pytest.mark.parametrize("name", [c[0] for c in MFMA_CASES])
def test_mfma_blocks_on_host_threads(name):
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p = ctx.Process(target=_mfma_worker, args=(name, q))
    p.start()
    r = q.get(timeout=900)
    p.join(60)
    assert not isinstance(r, str), r
    assert r < 2e-6, (name, r)


POST_BASIS = r'''
extern "C" void run_host_basis(V *amp, unsigned long long nblocks, unsigned nthreads, const T *coef,
                               unsigned long long binit) {
    for (unsigned long long b = 0; b < nblocks; b++) {
        std::barrier<> bar(nthreads);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; t++)
            th.emplace_back([=] { threadIdx.x = t; blockIdx.x = b; ncsv_sweep_from_basis(amp, coef, binit); });
        for (auto &x : th) x.join();
    }
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with C++20")
@pytest.mark.parametrize("precision,index", [(H.C128, 0), (H.C128, 40961), (H.C64, 70001)])
def test_from_basis_sweep_on_host_threads(precision, index, monkeypatch):
    """A plan's first sweep in its from-basis form (the basis amplitude made
    in registers; tiles without the basis index store zeros and return)
    followed by the plan's other sweeps, on host threads, equals the plan's
    FP64 emulation from the same basis state (17 qubits: 32 or 16 tiles)."""
    n = 17
    text = W.rotation_brickwork(n, 4, 9)
    work = tempfile.mkdtemp()
    try:
        monkeypatch.setenv("NCSV_JIT_DUMP", work)
        c = H.Circuit.parse(text)
        k = ctypes.c_uint32()
        assert L.lib().ncsv_jit_check(c._h, int(precision), b"gfx950", ctypes.byref(k)) == 0
        order = [line.split() for line in open(os.path.join(work, "order.txt")).read().splitlines()]
        hashes = [h for h, fb in order if fb == "0"]
        want = np.zeros(1 << n, dtype=np.complex128)
        want[index] = 1.0
        stats = (ctypes.c_uint64 * 4)()
        assert L.lib().ncsv_plan_emulate(c._h, int(precision), want.view(np.float64).ctypes.data_as(L.c_dp),
                                         stats) == 0
        buf = np.full(1 << n, np.nan + 1j * np.nan, dtype=np.complex128 if precision == H.C128 else np.complex64)
        tile = 12 if precision == H.C128 else 13
        for i, h in enumerate(hashes):
            src = open(os.path.join(work, f"ncsv_{h}.hip")).read()
            coef = np.fromfile(os.path.join(work, f"ncsv_{h}.coef"), dtype=np.float64)
            coef = np.ascontiguousarray(coef if precision == H.C128 else coef.astype(np.float32))
            if i == 0:
                fb = [f for f, flag in order if flag == "1"]
                assert fb, "the check wrote no from-basis form"
                src_b = open(os.path.join(work, f"ncsv_{fb[0]}.hip")).read()
                lib2 = host_kernel(src_b, work, "kb", POST_BASIS)
                lib2.run_host_basis(buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_ulonglong(1 << (n - tile)),
                                    ctypes.c_uint(256), coef.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_ulonglong(index))
                continue
            lib = host_kernel(src, work, f"k{i}")
            lib.run_host(buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_ulonglong(1 << (n - tile)), ctypes.c_uint(256),
                         coef.ctypes.data_as(ctypes.c_void_p))
        got = buf.astype(np.complex128)
        assert not np.isnan(got).any()
        assert np.abs(got - want).max() < (1e-13 if precision == H.C128 else 2e-6)
    finally:
        shutil.rmtree(work, ignore_errors=True)


EXPECT_PRE = r'''
struct GD { unsigned x; };
static GD gridDim;
// __shfl_down over a wavefront of host threads: every thread of the
// workgroup makes the same calls in the same order (wsum after the tile loop)
static double g_sd[4][64];
inline double __shfl_down(double v, int k, int w) {
    const unsigned wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    g_sd[wv][l] = v;
    g_bar->arrive_and_wait();
    const double r = (int)l + k < w ? g_sd[wv][l + k] : v;
    g_bar->arrive_and_wait();
    return r;
}
inline void lds_bar() { g_bar->arrive_and_wait(); }
'''

POST_EXPECT = r'''
extern "C" void run_host_expect(const V *amp, double *part, unsigned long long ntiles, unsigned grid, unsigned nthreads) {
    gridDim.x = grid;
    for (unsigned b = 0; b < grid; b++) {
        std::barrier<> bar(nthreads);
        g_bar = &bar;
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nthreads; t++)
            th.emplace_back([=] { threadIdx.x = t; blockIdx.x = b; ncsv_expect(amp, part, ntiles); });
        for (auto &x : th) x.join();
    }
}
'''


@pytest.mark.parametrize("precision,prefetch", [(H.C128, "1"), (H.C128, "0"), (H.C64, "0")])
def test_expectation_kernels_on_host_threads(precision, prefetch, monkeypatch):
    """The grouped expectation kernels the generator writes for 40 Paulis on
    16 qubits (phases shifted off the Hermitian ones, so both the real and
    the imaginary sums carry values), run on host threads: each group's
    per-Pauli sums equal the oracle's expectations (mod.rs:102-108), each
    matched to a distinct Pauli."""
    from oracle.pauli import Pauli
    from oracle.statevector import StateVector

    n = 16
    work = tempfile.mkdtemp()
    try:
        monkeypatch.setenv("NCSV_JIT_DUMP", work)
        monkeypatch.setenv("NCSV_JIT_NT", "0")
        monkeypatch.setenv("NCSV_EXPECT_PREFETCH", prefetch)
        r = W.SplitMix64(41)
        ps = [(x, z, (ph + r.below(4)) & 3) for x, z, ph in W.random_paulis(n, 40, 42)]
        k = len(ps)
        xs = (ctypes.c_uint64 * k)(*[p[0] for p in ps])
        zs = (ctypes.c_uint64 * k)(*[p[1] for p in ps])
        ph = (ctypes.c_uint32 * k)(*[p[2] for p in ps])
        g = ctypes.c_uint32()
        assert L.lib().ncsv_jit_check_expect(n, int(precision), k, xs, zs, ph, b"gfx950", ctypes.byref(g)) == 0
        srcs = [open(os.path.join(work, f)).read() for f in sorted(os.listdir(work)) if f.endswith(".hip")]
        srcs = [s for s in srcs if "ncsv_expect" in s]
        assert len(srcs) == g.value >= 2
        rng = np.random.default_rng(7)
        a = (rng.random(1 << n) - 0.5) + 1j * (rng.random(1 << n) - 0.5)
        a /= np.linalg.norm(a)
        buf = a.copy() if precision == H.C128 else a.astype(np.complex64)
        o = StateVector.from_amplitudes(n, buf.astype(np.complex128))
        want = [o.expectation(Pauli(n, x, z, p)) for x, z, p in ps]
        assert max(abs(w.imag) for w in want) > 1e-3 and max(abs(w.real) for w in want) > 1e-3
        tile = 12 if precision == H.C128 else 13
        ntiles = 1 << (n - tile)
        got = []
        for i, src in enumerate(srcs):
            two_cnt = int(re.search(r"red\[\d+\]\[(\d+)\]", src).group(1))
            lib = host_kernel(EXPECT_PRE + src, work, f"e{i}", POST_EXPECT)
            part = np.zeros(two_cnt * ntiles)
            lib.run_host_expect(buf.ctypes.data_as(ctypes.c_void_p), part.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_ulonglong(ntiles), ctypes.c_uint(ntiles), ctypes.c_uint(256))
            sums = part.reshape(two_cnt, ntiles).sum(axis=1)
            got += [complex(sums[2 * j], sums[2 * j + 1]) for j in range(two_cnt // 2)]
        assert len(got) == len(want)
        tol = 1e-13 if precision == H.C128 else 1e-6
        free = list(range(len(want)))
        for v in got:
            j = min(free, key=lambda j: abs(want[j] - v))
            assert abs(want[j] - v) < tol, (v, want[j])
            free.remove(j)
    finally:
        shutil.rmtree(work, ignore_errors=True)
