// Gate segments of a circuit (the gates between two measurements, resets or
// noise ops) executed on the device: one sweep per gate, or fused tile
// sweeps that apply many gates per pass over the amplitudes.
//
// A tile sweep reads every amplitude once and writes it once.  A tile is the
// 2^L amplitudes that agree on the n-L "outer" qubits; its L "tile" qubits
// always include qubits 0..c-1 (contiguous in memory).  A workgroup of 2^LT
// threads holds a tile in registers, 2^R amplitudes per thread (L = LT + R);
// the gates are applied in stages, each with R "register" tile positions
// (the qubits a non-diagonal gate may act on), and the tile is re-laid out
// through LDS between stages.  Diagonal gates and controls may sit on any
// qubit, tile or outer.
#pragma once

#include <vector>

#include "ncsv_internal.hpp"

namespace ncsv {

enum TileKind : uint8_t {
    TK_U1 = 0,     // 2x2 on register bit j0
    TK_U2 = 1,     // 4x4 on register bits (j0 = local bit 0, j1 = local bit 1), j0 < j1
    TK_D1R = 2,    // diag(d0, d1) on register bit j0
    TK_D1T = 3,    // diag(d0, d1) on the non-register qubit gmask
    TK_DM = 4,     // phase m[0] where (g & gmask) == gmask and (r & rmask) == rmask
    TK_PERM1 = 5,  // controlled flip of register bit j0: new1 = i^k0 v0, new0 = i^k1 v1
    TK_SWAPR = 6,  // swap register bits j0 < j1
    TK_U1R = 7,    // 2x2 with real entries (H, RY, ...)
    TK_U1X = 8,    // 2x2 with a real diagonal and an imaginary off-diagonal (RX, ...)
    TK_NSLOT = 9,  // a noise site on register bit j0: the Pauli error drawn for slot `aux` (batched form)
};

struct TileOpH {
    uint8_t kind = 0, j0 = 0, j1 = 0, k0 = 0, k1 = 0;
    uint32_t rmask = 0;
    uint64_t gmask = 0;
    uint32_t aux = 0;
    cd m[16];
};

struct StageH {
    uint32_t rmask_pos = 0;  // the R register tile positions (bit p = position p)
    uint32_t op_begin = 0, op_end = 0;
};

struct SweepH {
    std::vector<uint32_t> tq;  // global qubit of each tile position (size L)
    std::vector<StageH> stages;
    std::vector<TileOpH> ops;
    uint32_t gates = 0;  // unfused gates applied by this sweep
};

struct SegItem {
    bool tile = false;
    SweepH sw;
    LOp op;  // when !tile: a gate applied by its own sweep
};

struct TileGeom {
    uint32_t LT = 8, R = 4, c = 4;  // threads = 2^LT, registers per thread = 2^R, fixed low qubits
    bool pipe = false;              // load the next op while applying the current one
    bool merge_diag = false;        // multiply one-qubit diagonals into neighbouring 2x2s (saves passes
                                    // where ops cost a pass each; costs VALU where 2x2s have cheap forms)
    uint32_t floor_ops = 117;       // tile ops whose arithmetic takes as long as a sweep's memory
                                    // traffic (complex128 on this part; about half for complex64)
    uint32_t L() const { return LT + R; }
    int key() const { return (int)((LT << 12) | (R << 8) | (c << 4) | (pipe ? 1 : 0)); }
};

// The fused schedule of a segment of lowered gates on n qubits.
std::vector<SegItem> plan_segment(const std::vector<LOp> &ops, uint32_t n, const TileGeom &geo);

// Tile sweeps for a run of lowered gates when every tile position is
// addressable (the batched form's LDS-resident tiles): L tile qubits of which
// the low c are qubits 0..c-1; `first_fixed` qubits must be tile qubits of
// the first sweep.  Fails with NCSV_UNSUPPORTED on gates no tile op covers.
std::vector<SweepH> plan_lds_sweeps(const std::vector<LOp> &ops, uint32_t n, uint32_t L, uint32_t c,
                                    uint64_t first_fixed);

// plan_segment; with `specialized`, also with one and two more fixed low
// qubits (longer contiguous runs), keeping the plan of least estimated cost
// and setting *geo to its shape.
std::vector<SegItem> plan_segment_tuned(const std::vector<LOp> &ops, uint32_t n, TileGeom *geo, bool specialized);

// The memory layouts (register-position masks) a specialized sweep may load
// and store in, the default first.
std::vector<uint32_t> mem_candidates(const SweepH &sw, const TileGeom &geo);
namespace jit {
struct MemShape;
}
// The memory shape of a specialized sweep: loads in the first register
// stage's layout and stores in the last one's where those leave the fixed low
// positions to the threads (an LDS exchange fewer each), else the default
// (top R positions).  NCSV_JIT_MEMSHAPE=k (and NCSV_JIT_MEMSHAPE_STORE=k)
// force candidate k of mem_candidates instead (tests), "default" the default.
jit::MemShape forced_memory(const SweepH &sw, const TileGeom &geo);

// Whether the state runs kernels specialized per plan (fusion 2, or fusion 1
// from NCSV_JIT_MIN_QUBITS = 20 qubits on; never with NCSV_JIT=0).
bool use_specialized(const ncsv_state *s);
// Applies the gate ops c->ops[i..j) (ticks and annotations skipped).
void run_gate_segment(ncsv_state *s, ncsv_circuit *c, size_t i, size_t j);
// Frees the fused-engine buffers a state holds.
void fuse_release(ncsv_state *s);

namespace dev {
// The tile shape used for a state of this precision (tile.hip).
TileGeom tile_geom(ncsv_precision prec);
// Bytes of a sweep's device program (stages, then ops) and its packing.
size_t tile_prog_bytes(const SweepH &sw, ncsv_precision prec);
void tile_prog_pack(const SweepH &sw, ncsv_precision prec, void *dst);
// Sums each of `rows` contiguous rows of `cols` doubles; results to host.
void rowsum(ncsv_state *s, const double *part, uint32_t rows, uint32_t cols, double *host_out);
// Runs one tile sweep whose program sits at `dprog` in device memory.
void run_tile_sweep(ncsv_state *s, const SweepH &sw, const void *dprog, const TileGeom &geo);
}  // namespace dev

}  // namespace ncsv
