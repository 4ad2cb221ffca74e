"""ORACLE — test infrastructure only.

A CPU restatement of ncsim's dense `StateVector<C64>` (src/state/vector.rs,
src/state/mod.rs) and the gate/Pauli/circuit views it uses, plus an exact
Z[w, 1/sqrt2] simulator for Clifford+T circuits.  Only the tests, the smoke
check and the benchmark's CPU-baseline leg use it, as the checker; the
product (ncsim_amd) never imports or links it.

Pinned by: the reference's own known answers restated in tests/ (GHZ,
T^3 CCZ amplitudes, the diagonal exponential sum, unitarity/dagger/composition
identities, naive O(4^k 2^n) application, dense Pauli action, Born-rule
sampling statistics) and tests/golden fixtures.
"""
