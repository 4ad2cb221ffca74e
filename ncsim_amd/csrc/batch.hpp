// Batched form: many circuits or noise trajectories of uneven sizes run by
// one persistent kernel launch (no counterpart in the reference, which runs
// one state at a time: mod.rs:168-183 per circuit).
//
// Each entry (a circuit started from |0...0>, with Paulis to evaluate at its
// end) follows a program of steps.  A step is one pass over the entry's
// state in tiles of 2^L amplitudes (the amplitudes that agree on the n-L
// outer qubits), each tile held in LDS while the step's ops run on it:
//   gate ops   — the fusion planner's tile ops, on tile positions,
//   NSLOT      — the Pauli error drawn for a noise site of this trajectory,
//   MEASURE    — whole-state tile: reduce, decide, collapse, reset,
//   REDUCE     — per-tile FP64 partials (norm, <Z-string>) for a measurement,
//   COLLAPSE   — sums a REDUCE's partials (every tile the same way), decides
//                the outcome, collapses the tile, optional reset,
//   EXPECT     — per-tile FP64 partials of up to 9 Pauli expectations.
// An entry of n <= L qubits is one step on one tile: its state never leaves
// LDS.  Larger entries keep their state in HBM between steps.
//
// The launch claims items (a step's run of 2^lgt tiles) from one counter in
// depth-major order (step d of every entry before step d+1 of any), and an
// item waits for all tiles of its entry's previous step through that step's
// completion counter; the lowest unfinished item never waits, so the queue
// drains.  Only steps that a later step waits on publish their stores with
// a device-scope release (once per item).
#pragma once

#include <cstdint>

namespace ncsv {
namespace batch {

enum BOpKind : uint8_t {
    BK_U1 = 0,      // 2x2 on position p0 (coef: 4; when omask != 0, 4 more for tiles meeting omask, and
                    // k0 = 1: tiles that do not meet it are left as they are)
    BK_U1R = 1,     // 2x2 with real entries
    BK_U1X = 2,     // real diagonal, imaginary off-diagonal
    BK_U2 = 3,      // 4x4 on positions p0 < p1 (coef: 16)
    BK_DM = 4,      // phase coef[0] where (x & lmask) == lmask and outer bits omask set
    BK_PERM1 = 5,   // flip p0 where the conditions hold: new1 = i^k0 v0, new0 = i^k1 v1
    BK_SWAP = 6,    // swap positions p0 < p1
    BK_NSLOT = 7,   // noise slot a on position p0: error code bit0 = X, bit1 = Z (Z first)
    BK_MEASURE = 8, // whole tile: measurement group c (reduce, decide, collapse, resets)
    BK_REDUCE = 9,  // group c's cell weights to part[a + 2^D tile]
    BK_COLLAPSE = 10,  // group c decided from b tiles of cell weights at a; collapse, resets
    BK_EXPECT = 11, // Paulis [c, c+b) to part[a + 2b tile + 2k]
    BK_G3 = 12,     // k1 (2 or 3) unconditional 2x2s on distinct positions p0, p1, p2 in one pass
    BK_U3 = 13,     // dense 8x8 on distinct positions p0, p1, p2 (coef: 64, row-major, then 64 more for tiles
                    // meeting omask when omask != 0; a: the FP32 copies in complex64)
    BK_OPHASE = 14, // phase coef[0] on the whole tile where the outer bits omask are set, multiplied in at
                    // the store (a phase on outer qubits only commutes with every op of the step)
};

struct BOpD {
    uint8_t kind, p0, p1, k0, k1, p2, pad1, pad2;
    uint32_t lmask;  // condition / Z mask on tile positions
    uint32_t a, b, c;
    uint32_t coef;   // first complex coefficient
    uint64_t omask;  // condition / Z mask on outer qubits (global qubit bits)
};

struct BPauliD {
    uint32_t xl, zl;  // on tile positions
    uint64_t zo;      // on outer qubits
    uint32_t ph, pad;
};

// A run of consecutive Z-type measurements and resets decided together: the
// weights of the 2^D cells (parities of the D distinct Z masks) are reduced
// once; the events then take their outcomes in order from those weights,
// each restricting the kept cells, and resets flip their qubit at the end
// (no later event of the run involves a reset qubit).
struct BMGroupD {
    uint32_t D, E;
    uint32_t lm[4];  // masks on tile positions
    uint64_t om[4];  // masks on outer qubits
    struct Ev {
        uint8_t d, sign, rpos, pad;  // mask index, sign (ph / 2), reset position or 0xff
        uint32_t ord;                // measurement ordinal
    } ev[16];
};

enum BStepFlags : uint32_t { BF_INIT = 1, BF_LOAD = 2, BF_STORE = 4 };

struct BStepD {
    uint32_t L, n;
    uint32_t op0, nops;    // the ops
    uint32_t op0c, nopsc;  // the same without noise slots (entries with no error drawn in this step)
    uint32_t flags;
    uint32_t qoff;   // qubit table: tile qubit of positions 0..L-1, then outer qubits ascending
    uint32_t omask;  // the outer qubits (tile number bit b <-> the b-th lowest)
    uint32_t slot_lo, slot_n;  // the noise slots of this step lie in [slot_lo, slot_lo + slot_n)
    uint32_t var0;             // ops[var0 + k]: header (a = first op, b = op count) of the op list
                               // keeping only the step's k-th noise slot (records with one error)
    uint8_t tpos[16];          // the qubits of tile positions 0..min(L, 16)-1
};

struct BEntryD {
    uint64_t amp_off;  // first amplitude in the pool (multi-tile entries)
    uint64_t part0;    // first measurement partial (doubles)
    uint64_t exp0;     // first expectation partial (doubles)
    uint32_t m0;       // first measurement ordinal slot (uniforms, outcomes, forced)
    uint32_t err0;     // first noise slot
};

struct BRecD {
    uint64_t first;       // first item number of this (entry, step)
    uint32_t entry, step;  // step: index into the step table
    uint32_t prev;        // record of the entry's previous step, or ~0u
    uint32_t prev_tiles;  // its tile count
    uint32_t flags;       // RF_ERR: a noise error was drawn in this step; RF_RELEASE: a later step waits on it
    uint32_t lgt;         // log2 of the tiles per item
    uint32_t var;         // RF_ONE: the step's noise slot (index in its slot list) holding the one error
};

// One expectation sum of an entry (a Pauli's or the norm's real or imaginary
// part): its per-tile partials sit at off + stride * t for t < tiles.
struct BSumD {
    uint64_t off;
    uint32_t stride, tiles;
};

// RF_ONE (with RF_ERR): exactly one error was drawn in this step, so the op
// list keeping only its slot runs instead of the list with every slot
enum BRecFlags : uint32_t { RF_ERR = 1, RF_RELEASE = 2, RF_ONE = 4 };

struct BArgs {
    const BStepD *steps;
    const BOpD *ops;
    const double *coef;  // complex coefficients as (re, im)
    const BPauliD *paulis;
    const BMGroupD *mgroups;
    const uint8_t *qtab;
    const BEntryD *entries;
    const BRecD *recs;        // + sentinel
    const uint32_t *rec_blk;  // record holding item 16k
    uint64_t total;           // items
    void *amp;
    double *part;             // measurement partials
    double *pexp;             // expectation partials
    double *lsc;              // per entry: logical scale (complex64 renormalization)
    const double *uni;        // per measurement ordinal
    const uint8_t *forced;    // per ordinal: 0, 1 or 0xff (free); may be null
    uint8_t *outc;            // per ordinal
    const uint8_t *errs;      // per noise slot
    uint32_t *done;           // per record
    unsigned long long *next;  // queue counter
    uint32_t *flag;           // bit 0: a wait ran out of spins
    uint32_t spin_limit;
};

}  // namespace batch
}  // namespace ncsv
