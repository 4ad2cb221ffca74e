// Runs the library's CPU-only paths under AddressSanitizer and
// UndefinedBehaviorSanitizer (no device): circuit text parsing and printing
// (valid and invalid texts), the circuit builder's checks, the fused-schedule
// planner with its FP64 host emulation, and the batched form's planner, FP64
// emulator and host run of its kernel code.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ncsv.h"

static int fails = 0;
#define CHECK(cond)                                                  \
    do {                                                             \
        if (!(cond)) {                                               \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond); \
            fails++;                                                 \
        }                                                            \
    } while (0)

static std::string brickwork(int n, int depth) {
    std::string t = "QUBITS " + std::to_string(n) + "\n";
    const char *ax[3] = {"RX", "RY", "RZ"};
    for (int l = 0; l < depth; l++) {
        for (int q = 0; q < n; q++) t += std::string(ax[(q * 7 + l) % 3]) + "(" + std::to_string(0.1 + q * 0.37 + l) + ") " + std::to_string(q) + "\n";
        for (int a = l % 2; a + 1 < n; a += 2) t += "CZ " + std::to_string(a) + " " + std::to_string(a + 1) + "\n";
    }
    return t;
}

static std::string qft(int n) {
    std::string t = "QUBITS " + std::to_string(n) + "\nX 1\nX 3\n";
    for (int j = 0; j < n; j++) {
        t += "H " + std::to_string(j) + "\n";
        for (int m = j + 1; m < n; m++)
            t += "CPHASE(1," + std::to_string(m - j + 1) + ") " + std::to_string(m) + " " + std::to_string(j) + "\n";
        t += "RZ(" + std::to_string(0.3 * j) + ") " + std::to_string((j * 3) % n) + "\n";
    }
    return t;
}

int main() {
    // parsing and printing
    const char *good[] = {
        "H 0\nCX 0 1 1 2\nT 2\nPHASE(3,4) 1\nM 0\nMX 1\nMPP X0*Z2\nR 2\nTICK\n",
        "QUBITS 5\nRX(0.25) 0 1\nCPHASE(1,33) 3 4\nCCX 0 1 2\nDEPOLARIZE1(0.001) 2 3\nMY 4\nDETECTOR rec[-1]\n"
        "OBSERVABLE_INCLUDE(0) rec[-1]\nDEPOLARIZE2(0.01) 0 1\nPAULI_CHANNEL_1(0.1,0,0.2) 3\n",
        "QUBITS 3\nH 0 # comment\n\nSQRT_X_DAG 1\nCY 0 2\nSWAP 1 2\nCS_DAG 0 1\nCCZ 0 1 2\nX_ERROR(0.5) 1\n"};
    for (const char *t : good) {
        ncsv_circuit *c = nullptr;
        CHECK(ncsv_circuit_parse(t, &c) == NCSV_OK);
        if (!c) continue;
        const size_t len = ncsv_circuit_to_text(c, nullptr, 0);
        std::vector<char> buf(len + 1);
        ncsv_circuit_to_text(c, buf.data(), buf.size());
        ncsv_circuit *c2 = nullptr;
        CHECK(ncsv_circuit_parse(buf.data(), &c2) == NCSV_OK);
        std::vector<char> buf2(len + 1);
        ncsv_circuit_to_text(c2, buf2.data(), buf2.size());
        CHECK(strcmp(buf.data(), buf2.data()) == 0);
        char small[7];
        ncsv_circuit_to_text(c, small, sizeof(small));
        CHECK(strlen(small) == 6);
        ncsv_circuit_destroy(c2);
        ncsv_circuit_destroy(c);
    }
    const char *bad[] = {"CX 0", "QUBITS 1\nH 3", "FOO 1", "CX 1 1", "MPP X0*X0", "PHASE(1,0) 0", "DETECTOR rec[-1]",
                         "X_ERROR(1.5) 0", "RX( 0", "H -1", "QUBITS x", "CPHASE(1,63) 0 1",
                         "OBSERVABLE_INCLUDE(-1) rec[-1]", "M 0\nDETECTOR rec[-2]"};
    for (const char *t : bad) {
        ncsv_circuit *c = nullptr;
        const ncsv_status st = ncsv_circuit_parse(t, &c);
        CHECK(st != NCSV_OK && c == nullptr);
        char msg[256];
        ncsv_last_error(msg, sizeof(msg));
    }
    // builder checks
    ncsv_circuit *c = nullptr;
    CHECK(ncsv_circuit_new(3, &c) == NCSV_OK);
    const uint32_t q01[2] = {0, 1}, q00[2] = {0, 0}, q5[1] = {5};
    CHECK(ncsv_circuit_push_gate(c, NCSV_GATE_CX, nullptr, nullptr, q01, 2) == NCSV_OK);
    CHECK(ncsv_circuit_push_gate(c, NCSV_GATE_CX, nullptr, nullptr, q00, 2) == NCSV_DUPLICATE_QUBIT);
    CHECK(ncsv_circuit_push_gate(c, NCSV_GATE_H, nullptr, nullptr, q5, 1) == NCSV_QUBIT_OUT_OF_RANGE);
    CHECK(ncsv_circuit_push_gate(c, NCSV_GATE_CX, nullptr, nullptr, q01, 1) == NCSV_ARITY);
    CHECK(ncsv_circuit_push_measure(c, 1, 0, 1) == NCSV_INVALID);
    const double p[1] = {0.01};
    CHECK(ncsv_circuit_push_noise(c, NCSV_NOISE_DEPOLARIZE1, p, q01, 1) == NCSV_OK);
    CHECK(ncsv_circuit_push_noise(c, NCSV_NOISE_DEPOLARIZE1, p, q01, 2) == NCSV_ARITY);
    ncsv_circuit_destroy(c);

    // planner and its emulation on both precisions
    for (int prec = 0; prec < 2; prec++) {
        const std::string texts[] = {brickwork(14, 6), qft(15), brickwork(21, 3)};
        for (const std::string &t : texts) {
            ncsv_circuit *cc = nullptr;
            CHECK(ncsv_circuit_parse(t.c_str(), &cc) == NCSV_OK);
            uint32_t n = 0;
            ncsv_circuit_num_qubits(cc, &n);
            std::vector<double> amp(2ull << n, 0.0);
            amp[0] = 1.0;
            uint64_t stats[4];
            CHECK(ncsv_plan_emulate(cc, (ncsv_precision)prec, amp.data(), stats) == NCSV_OK);
            double nn = 0;
            for (double v : amp) nn += v * v;
            CHECK(std::fabs(nn - 1.0) < 1e-9);
            ncsv_circuit_destroy(cc);
        }
    }

    // batched form: planning, FP64 emulation and the kernel code on host threads
    for (int prec = 0; prec < 2; prec++) {
        ncsv_batch *b = ncsv_batch_new((ncsv_precision)prec);
        const std::string texts[] = {
            "QUBITS 7\nH 0 1 2\nCX 0 3\nT 4\nCCZ 1 2 5\nDEPOLARIZE1(0.3) 6\nM 0\nR 2\nRY(0.5) 3\nMPP Z1*Z4\n",
            brickwork(9, 4) + "M 0 4\nX_ERROR(0.2) 3\n", qft(8)};
        const uint64_t zx[2] = {0, 0}, zz[2] = {1, 6};
        const uint32_t zp[2] = {0, 0};
        std::vector<ncsv_circuit *> keep;
        uint32_t draws = 0;
        for (const std::string &t : texts) {
            ncsv_circuit *cc = nullptr;
            CHECK(ncsv_circuit_parse(t.c_str(), &cc) == NCSV_OK);
            uint32_t idx = 0, d = 0;
            CHECK(ncsv_batch_add(b, cc, 2, zx, zz, zp, &idx) == NCSV_OK);
            ncsv_batch_draws(b, idx, &d);
            draws += d;
            keep.push_back(cc);
        }
        ncsv_batch_set_tile_bits(b, 6);
        std::vector<double> us(draws + 1);
        for (uint32_t i = 0; i < draws; i++) us[i] = (i * 0.37 + 0.11) - (int)(i * 0.37 + 0.11);
        CHECK(ncsv_batch_plan(b) == NCSV_OK);
        CHECK(ncsv_batch_emulate(b, us.data(), nullptr, nullptr) == NCSV_OK);
        double ex[4], nn;
        uint8_t oc[16];
        for (uint32_t e = 0; e < 3; e++) CHECK(ncsv_batch_result(b, e, ex, &nn, oc) == NCSV_OK);
        CHECK(ncsv_batch_simulate(b, us.data(), nullptr, nullptr) == NCSV_OK);
        ncsv_batch_destroy(b);
        for (ncsv_circuit *cc : keep) ncsv_circuit_destroy(cc);
    }
    // no device: an error, not a crash
    ncsv_state *s = nullptr;
    if (ncsv_device_count() == 0) CHECK(ncsv_create(4, NCSV_C128, &s) == NCSV_DEVICE && s == nullptr);
    printf("library driver: %s (%d failed checks)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
