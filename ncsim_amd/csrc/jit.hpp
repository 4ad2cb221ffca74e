// Sweep kernels specialized per plan and compiled at run time (jit.cpp).
#pragma once

#include <string>
#include <vector>

#include "fuse.hpp"

namespace ncsv {
namespace jit {
// How a sweep kernel meets memory: the register positions of its loads and
// of its stores (R-subsets of the tile positions c..L-1; 0 = the top R
// positions) and the order of its workgroups over the tiles (0 ascending
// block index, 1 XCD-chunked: XCD x takes a contiguous run of tiles; -1
// NCSV_JIT_BORDER, default 0).
struct MemShape {
    uint32_t load = 0, store = 0;
    int border = -1;
};
uint32_t default_memr(const TileGeom &geo);
// HIP source of one tile sweep; appends the sweep's coefficients to *coef.
// from_basis: the sweep starts from the basis state |binit> (a kernel
// argument) computed in registers instead of loaded, with the same
// coefficients in the same order.
std::string generate(const SweepH &sw, uint32_t n, const TileGeom &geo, bool dbl, std::vector<double> *coef,
                     bool from_basis = false, const MemShape *ms = nullptr);
// hiprtc compile for `arch`; empty on failure (message in *log).
std::vector<char> compile_to_code(const std::string &src, const std::string &arch, std::string *log);
// Starts (or finds) the compile of `src`; returns an opaque handle.
void *prepare(const std::string &src);
// The same for a kernel named fname (sweeps from a basis state: ncsv_sweep_from_basis).
void *prepare_named(const std::string &src, const char *fname);
void release(void *handle);
// Launches a prepared sweep on the state's stream (waits for its compile).
void launch(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, const void *dcoef);
// Launches a prepared from_basis sweep: the state becomes its result applied
// to |binit> (no amplitude is read).
void launch_from_basis(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, const void *dcoef, uint64_t binit);
// false when NCSV_JIT=0.
bool available();

// A group of Paulis measured by one specialized kernel over one tile shape:
// the tile's qubits (positions 0..L-1) and stages of register positions,
// each with the Paulis whose X part lies on its registers.
struct ExpPauli {
    uint64_t x, z;
    uint32_t ph;
};
struct ExpStage {
    uint32_t rpos;
    std::vector<uint32_t> paulis;
};
struct ExpGroup {
    std::vector<uint32_t> tq;
    std::vector<ExpPauli> paulis;
    std::vector<ExpStage> stages;
    std::vector<uint32_t> index;  // position of each Pauli in the caller's list
};
std::string generate_expect(const ExpGroup &eg, uint32_t n, const TileGeom &geo, bool dbl);
void *prepare_named(const std::string &src, const char *fname);
// part: 2 x paulis rows of `grid` partial sums
void launch_expect(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, uint32_t grid, double *dpart);
// <P_k> for all k through specialized group kernels; false when some X part
// does not fit a tile's register stage (nothing done then).
bool plan_expect(uint32_t n, const TileGeom &geo, uint32_t count, const uint64_t *x, const uint64_t *z,
                 const uint32_t *ph, std::vector<ExpGroup> *out_groups);
// the expectation kernels' tile shape
TileGeom expect_geom(ncsv_precision prec);
bool expectations(ncsv_state *s, uint32_t count, const uint64_t *x, const uint64_t *z, const uint32_t *ph,
                  double *out);
}  // namespace jit
}  // namespace ncsv
