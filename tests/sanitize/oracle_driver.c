/* Runs every kernel of the oracle (oracle/sv_oracle.c) on small states under
 * AddressSanitizer and UndefinedBehaviorSanitizer: dense k = 1..5 on qubits
 * below and above the 4096-amplitude block, monomials, Paulis with and
 * without X parts, projections, reductions, tensor products and sampling
 * with edge draws (0 and the largest double below 1). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_apply_matrix(double *amp, int n, const double *m, const int *qubits, int k);
void orc_apply_monomial(double *amp, int n, const int *qubits, int k, const int *y, const double *c);
void orc_apply_pauli(double *amp, int n, uint64_t x, uint64_t z, int ph);
void orc_norm_sq(const double *amp, int n, double *out);
void orc_expectation(const double *amp, int n, uint64_t x, uint64_t z, int ph, double *out);
void orc_inner(const double *a, const double *b, int n, double *out);
void orc_project(double *amp, int n, uint64_t x, uint64_t z, int ph, int outcome, double *out);
void orc_scale(double *amp, int n, double cre, double cim);
void orc_kron(const double *a, int na, const double *b, int nb, double *out);
int orc_sample(const double *amp, int n, const double *uniforms, int64_t shots, uint64_t *out);

static uint64_t s = 12345;
static double rnd(void) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

int main(void) {
    for (int n = 1; n <= 14; n++) {
        const size_t N = (size_t)1 << n;
        double *a = malloc(16 * N), *b = malloc(16 * N), out[2];
        for (size_t i = 0; i < 2 * N; i++) {
            a[i] = rnd();
            b[i] = rnd();
        }
        for (int k = 1; k <= (n < 5 ? n : 5); k++) {
            const size_t d = (size_t)1 << k;
            double *m = malloc(16 * d * d);
            for (size_t i = 0; i < 2 * d * d; i++) m[i] = rnd();
            int q[5];
            for (int j = 0; j < k; j++) q[j] = (n - 1 - 2 * j + 3 * n) % n;
            int ok = 1;
            for (int j = 0; j < k; j++)
                for (int l = 0; l < j; l++) ok &= q[j] != q[l];
            if (!ok)
                for (int j = 0; j < k; j++) q[j] = j;
            orc_apply_matrix(a, n, m, q, k);
            int y[32];
            for (size_t x = 0; x < d; x++) y[x] = (int)((x * 5 + 3) % d);
            if (d == 2) y[0] = 1, y[1] = 0;
            orc_apply_monomial(a, n, q, k, y, m);
            free(m);
        }
        const uint64_t mask = N - 1;
        orc_apply_pauli(a, n, 0, 5 & mask, 1);
        orc_apply_pauli(a, n, 6 & mask, 3 & mask, 2);
        orc_norm_sq(a, n, out);
        orc_expectation(a, n, 0, 1, 0, out);
        orc_expectation(a, n, 3 & mask, 6 & mask, (int)(__builtin_popcountll(3 & 6 & mask) & 1), out);
        orc_inner(a, b, n, out);
        orc_project(a, n, 0, 1, 0, 1, out);
        orc_project(a, n, 1, 0, 0, 0, out);
        orc_scale(a, n, 0.5, -0.25);
        if (n <= 7) {
            double *k2 = malloc(16 * N * N);
            orc_kron(a, n, b, n, k2);
            free(k2);
        }
        double us[5] = {0.0, 0.25, 0.5, 0.9999, 1.0 - 1.0 / 9007199254740992.0};
        uint64_t picks[5];
        if (orc_sample(a, n, us, 5, picks) != 0) return 2;
        memset(a, 0, 16 * N);
        if (orc_sample(a, n, us, 5, picks) != 6) return 3;  /* ZeroState */
        free(a);
        free(b);
    }
    printf("oracle driver: ok\n");
    return 0;
}
