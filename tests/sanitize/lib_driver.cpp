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
    // frame sampler: Clifford check, tableau reference record, op compiler
    // and the kernel's per-word code on the host (a register over 64 qubits,
    // and Pauli-product measurements on a small one)
    {
        const char *ts[2] = {
            "QUBITS 70\nH 0\nCX 0 1\nS 1\nSQRT_Y 2\nCZ 2 0\nDEPOLARIZE1(0.1) 0 2\nDEPOLARIZE2(0.2) 1 2\n"
            "M 0 1\nR 1\nCX 2 69\nPAULI_CHANNEL_1(0.05,0.1,0.15) 69\nX_ERROR(1.0) 3\nM 2 3\n"
            "DETECTOR rec[-1] rec[-2]\nOBSERVABLE_INCLUDE(0) rec[-3]\n",
            "QUBITS 3\nH 0\nCX 0 1\nSQRT_X 2\nMPP X0*Z1*Y2 Z0*Z1\nDEPOLARIZE1(0.3) 1\nM 0 1 2\nMX 0\nMY 1\n"
            "DETECTOR rec[-6] rec[-5]\nOBSERVABLE_INCLUDE(1) rec[-1]\n"};
        const uint64_t want[2][3] = {{4, 1, 1}, {7, 1, 2}};
        for (int i = 0; i < 2; i++) {
            ncsv_circuit *cc = nullptr;
            CHECK(ncsv_circuit_parse(ts[i], &cc) == NCSV_OK);
            ncsv_frames *f = nullptr;
            CHECK(ncsv_frames_new(cc, 7, &f) == NCSV_OK);
            uint64_t m = 0, d = 0, o = 0;
            CHECK(ncsv_frames_shape(f, &m, &d, &o) == NCSV_OK && m == want[i][0] && d == want[i][1] &&
                  o == want[i][2]);
            std::vector<uint8_t> rec(m + 1);
            CHECK(ncsv_frames_reference(f, rec.data(), m) == NCSV_OK);
            for (uint64_t shots : {1ull, 64ull, 77ull, 300ull}) {
                std::vector<uint64_t> out(m * ((shots + 63) / 64));
                CHECK(ncsv_frames_emulate_packed(f, shots, 11, out.data()) == NCSV_OK);
            }
            ncsv_frames_destroy(f);
            ncsv_circuit_destroy(cc);
        }
        ncsv_circuit *cc = nullptr;
        CHECK(ncsv_circuit_parse("QUBITS 1\nT 0\nM 0\n", &cc) == NCSV_OK);
        ncsv_frames *f = nullptr;
        CHECK(ncsv_frames_new(cc, 0, &f) == NCSV_UNSUPPORTED && f == nullptr);
        ncsv_circuit_destroy(cc);
    }
    // no device: an error, not a crash
    ncsv_state *s = nullptr;
    if (ncsv_device_count() == 0) CHECK(ncsv_create(4, NCSV_C128, &s) == NCSV_DEVICE && s == nullptr);
    printf("library driver: %s (%d failed checks)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
