/* Calls every function of the oracle's C restatement (oracle/oracle.c) on ragged
 * shapes, both layouts and 1..4 reference threads, for an AddressSanitizer /
 * UndefinedBehaviorSanitizer build (tests/test_sanitize.py). Prints a checksum per
 * call; the sanitizers abort on the first error. Test infrastructure only. */
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static uint64_t state = 88172645463325252ull;
static uint64_t next(void)
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return state;
}
static float unit(void) { return (float)((next() >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f; }

int main(void)
{
    const long Ns[] = {1, 7, 33, 96, 130};
    const long Ms[] = {12, 24, 36};
    const long K1 = 40;
    const int n = 3;
    for (size_t t = 0; t < sizeof Ns / sizeof Ns[0]; t++) {
        const long N = Ns[t];
        long M[3], K[3];
        for (int i = 0; i < n; i++) {
            M[i] = Ms[i];
            K[i] = i ? Ms[i - 1] : K1;
        }
        float *Bf = malloc(sizeof(float) * K1 * N), *Af[3], *Cf = malloc(sizeof(float) * M[2] * N);
        uint16_t *Bb = malloc(2 * K1 * N), *Ab[3];
        uint8_t *Bi = malloc(K1 * N);
        int8_t *Ai[3];
        int32_t *Ci = malloc(4 * M[2] * N);
        double *Bd = malloc(8 * K1 * N), *Ad[3], *Cd = malloc(8 * M[2] * N);
        for (long e = 0; e < K1 * N; e++) {
            Bf[e] = unit();
            Bb[e] = oracle_bf16_rne(Bf[e]);
            Bi[e] = (uint8_t)next();
            Bd[e] = Bf[e];
        }
        for (int i = 0; i < n; i++) {
            Af[i] = malloc(sizeof(float) * M[i] * K[i]);
            Ab[i] = malloc(2 * M[i] * K[i]);
            Ai[i] = malloc(M[i] * K[i]);
            Ad[i] = malloc(8 * M[i] * K[i]);
            for (long e = 0; e < M[i] * K[i]; e++) {
                Af[i][e] = unit();
                Ab[i][e] = oracle_bf16_trunc(Af[i][e]);
                Ai[i][e] = (int8_t)(next() % 7) - 3;
                Ad[i][e] = Af[i][e];
            }
        }
        double sum = 0;
        oracle_chain_f32(Cf, Bf, N, n, (const float *const *)Af, M, K);
        for (long e = 0; e < M[2] * N; e++) sum += Cf[e];
        oracle_chain_f64(Cd, Bd, N, n, (const double *const *)Ad, M, K);
        for (long e = 0; e < M[2] * N; e++) sum += Cd[e];
        for (int layout = 0; layout < 2; layout++)
            for (int th = 1; th <= 4; th++) {
                if (oracle_chain_bf16(Cf, Bb, N, n, (const uint16_t *const *)Ab, M, K, layout, th) == 0)
                    for (long e = 0; e < M[2] * N; e++) sum += Cf[e];
                if (oracle_chain_int8(Ci, Bi, N, n, (const int8_t *const *)Ai, M, K, layout, th) == 0)
                    for (long e = 0; e < M[2] * N; e++) sum += Ci[e];
            }
        long nb[16], ne[16], num;
        for (int th = 1; th <= 16; th++) oracle_partition(N, th, nb, ne, &num);
        float *Cs = malloc(sizeof(float) * M[2] * N);
        oracle_sgemm(Cs, Af[2], Bf, M[2], N, K1 < K[2] ? K1 : K[2]);
        for (long e = 0; e < M[2] * N; e++) sum += Cs[e];
        printf("N=%ld checksum %.6e\n", N, sum);
        free(Cs);
        for (int i = 0; i < n; i++) {
            free(Af[i]);
            free(Ab[i]);
            free(Ai[i]);
            free(Ad[i]);
        }
        free(Bf), free(Cf), free(Bb), free(Bi), free(Ci), free(Bd), free(Cd);
    }
    /* a K tail past one 2048-deep block of the singlegemm restatement */
    const long M = 20, N = 16, K = 2336;
    float *A = malloc(sizeof(float) * M * K), *B = malloc(sizeof(float) * K * N), *C = malloc(sizeof(float) * M * N);
    for (long e = 0; e < M * K; e++) A[e] = unit();
    for (long e = 0; e < K * N; e++) B[e] = unit();
    oracle_sgemm(C, A, B, M, N, K);
    printf("sgemm tail %.6e\n", (double)C[0] + C[M * N - 1]);
    free(A), free(B), free(C);
    return 0;
}
