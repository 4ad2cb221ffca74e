// Sweep kernels specialized per plan and compiled at run time (jit.cpp).
#pragma once

#include <string>
#include <vector>

#include "fuse.hpp"

namespace ncsv {
namespace jit {
// HIP source of one tile sweep; appends the sweep's coefficients to *coef.
std::string generate(const SweepH &sw, uint32_t n, const TileGeom &geo, bool dbl, std::vector<double> *coef);
// hiprtc compile for `arch`; empty on failure (message in *log).
std::vector<char> compile_to_code(const std::string &src, const std::string &arch, std::string *log);
// Starts (or finds) the compile of `src`; returns an opaque handle.
void *prepare(const std::string &src);
void release(void *handle);
// Launches a prepared sweep on the state's stream (waits for its compile).
void launch(ncsv_state *s, void *handle, uint32_t L, uint32_t LT, const void *dcoef);
// false when NCSV_JIT=0.
bool available();
}  // namespace jit
}  // namespace ncsv
