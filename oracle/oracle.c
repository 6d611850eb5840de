/*
 * oracle.c - CPU restatement of the reference flashgemm arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg. The product library never links or calls it.
 *
 * Parity pins (tests/test_oracle.py):
 *   - golden fixtures written by the compiled reference (tests/golden/), bit-exact
 *     for every function below except oracle_chain_f64;
 *   - brute-force Python loops on tiny shapes (INT8, FP32 sequential fma);
 *   - closed forms: ones-chains, saturation boundaries, textbook RNE cases;
 *   - oracle_chain_f64 against numpy float64 matmul.
 *
 * Built with -ffp-contract=off: every fused multiply-add below is an explicit fmaf.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ bf16 -- */

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* Denormal operand read as a signed zero (the "DAZ" behaviour of the AVX512-BF16
 * instructions, kernel_bf16.h:38 vdpbf16ps and :357 vcvtne2ps2bf16). */
static inline float daz(float x)
{
    uint32_t u = f2u(x);
    if ((u & 0x7f800000u) == 0) return u2f(u & 0x80000000u);
    return x;
}

uint16_t oracle_bf16_rne(float x)
{
    uint32_t u = f2u(daz(x));
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u); /* quiet NaN */
    uint32_t lsb = (u >> 16) & 1u;
    return (uint16_t)((u + 0x7fffu + lsb) >> 16);
}

uint16_t oracle_bf16_trunc(float x) { return (uint16_t)(f2u(x) >> 16); }

float oracle_bf16_to_f32(uint16_t b) { return u2f((uint32_t)b << 16); }

/* ------------------------------------------------------------- partition -- */

void oracle_partition(long N, int nthreads, long *nb, long *ne, long *num_n)
{
    /* multigemm/src/flashgemm.c:42-51 */
    long groups = (N + 15) / 16;
    long b = (groups % nthreads == 0) ? (groups / nthreads) * 16
                                      : ((groups + nthreads - 1) / nthreads) * 16;
    long e = N % b;
    *nb = b;
    *ne = e;
    *num_n = N / b + (e == 0 ? 0 : 1);
}

/* ------------------------------------------------- reference intermediate -- */

/*
 * The reference's GEMM i epilogue writes the requantized panel Z (M_i rows x w
 * columns) into the buffer GEMM i+1 reads as packed B, in an order that differs
 * from the packed-B order for INT8 and BF16 (FP32 agrees). Reading it back as
 * packed B gives X[k][c] = Z[src_row(k,c)][panel + src_col(k,c)] with, for
 * panel-local column c of a panel of width w:
 *
 *   INT8, w == 32 (kernel_int8.h:329-385 vs :6-50):
 *       src_row = (k & ~3) + (c >> 3),     src_col = 4*(c & 7) + (k & 3)
 *   INT8, w == 16 (kernel_int8.h:1332-1386):
 *       src_row = (k & ~3) + 3 - (k & 3),  src_col = c
 *   BF16, w == 32 (kernel_bf16.h:356-390 vs :6-32):
 *       src_row = (k & ~1) + (c >> 4),     src_col = 2*(c & 15) + (k & 1)
 *   BF16, w == 16 (kernel_bf16.h:1332-1356):
 *       src_row = b(k >> 1) + (c >> 3),    src_col = 2*(c & 7) + (k & 1),
 *       b(g) = g + (g + 1) / 2   (the 16-wide epilogue's row pairs advance by
 *       2, 1, 2, 1, ... rows, so every third row is read twice and the last
 *       rows of Z are never read; measured on the compiled reference)
 *
 * Panels start at id*nb + 32*j inside each thread's column range
 * (flashgemm.c:109-127); widths other than 16 and 32 are outside the reference's
 * domain and rejected here.
 */
static int ref_src(int is_int8, long w, long k, long c, long *row, long *col)
{
    if (is_int8) {
        if (w == 32) { *row = (k & ~3L) + (c >> 3); *col = 4 * (c & 7) + (k & 3); return 0; }
        if (w == 16) { *row = (k & ~3L) + 3 - (k & 3); *col = c; return 0; }
    } else {
        if (w == 32) { *row = (k & ~1L) + (c >> 4); *col = 2 * (c & 15) + (k & 1); return 0; }
        if (w == 16) {
            long g = k >> 1;
            *row = g + (g + 1) / 2 + (c >> 3);
            *col = 2 * (c & 7) + (k & 1);
            return 0;
        }
    }
    return -1;
}

/* Apply the reference re-interpretation to a whole intermediate Z (rows x N, elem
 * bytes each) producing X. Returns -2 when a panel width is outside {16, 32} or the
 * row count is not a multiple of the k-group. */
static int ref_permute(const void *Zv, void *Xv, long rows, long N, size_t elem,
                       int is_int8, int nthreads)
{
    const unsigned char *Z = (const unsigned char *)Zv;
    unsigned char *X = (unsigned char *)Xv;
    long grp = is_int8 ? 4 : 2;
    if (rows % grp != 0) return -2;
    long nb, ne, num_n;
    oracle_partition(N, nthreads, &nb, &ne, &num_n);
    for (long id = 0; id < num_n; id++) {
        long NB = (id < num_n - 1) ? nb : (ne > 0 ? ne : nb);
        for (long j = 0; j < NB; j += 32) {
            long w = (NB - j < 32) ? NB - j : 32;
            long p0 = id * nb + j;
            for (long k = 0; k < rows; k++)
                for (long c = 0; c < w; c++) {
                    long r, cc;
                    if (ref_src(is_int8, w, k, c, &r, &cc) != 0) return -2;
                    memcpy(X + (k * N + p0 + c) * elem, Z + (r * N + p0 + cc) * elem, elem);
                }
        }
    }
    return 0;
}

/* ------------------------------------------------------------ GEMM cores -- */

/* Y(M x N) = A(M x K) . X(K x N); Y[m][n] = fmaf chain over k = 0..K-1
 * (kernel_f32.h:917-942: one vfmadd231ps per k into a register accumulator). */
static void gemm_f32_seq(const float *A, const float *X, float *Y, long M, long K, long N)
{
#pragma omp parallel for schedule(static)
    for (long m = 0; m < M; m++) {
        float *y = Y + m * N;
        for (long n = 0; n < N; n++) y[n] = 0.0f;
        for (long k = 0; k < K; k++) {
            float a = A[m * K + k];
            const float *x = X + k * N;
            for (long n = 0; n < N; n++) y[n] = fmaf(a, x[n], y[n]);
        }
    }
}

/* One vdpbf16ps step: acc + a*b with the exact bf16 product, fp32 RNE, denormal
 * inputs read as zero and a denormal result flushed to a signed zero. */
static inline float dp_step(float acc, float a, float b)
{
    float r = fmaf(daz(a), daz(b), daz(acc));
    return daz(r);
}

/* BF16 GEMM with the vdpbf16ps pair order (kernel_bf16.h:38, :797-822). K even. */
static void gemm_bf16_pairs(const uint16_t *A, const uint16_t *X, float *Y, long M, long K, long N)
{
#pragma omp parallel for schedule(static)
    for (long m = 0; m < M; m++) {
        float *y = Y + m * N;
        for (long n = 0; n < N; n++) y[n] = 0.0f;
        for (long g = 0; g < K / 2; g++) {
            float a1 = oracle_bf16_to_f32(A[m * K + 2 * g + 1]);
            float a0 = oracle_bf16_to_f32(A[m * K + 2 * g]);
            const uint16_t *x1 = X + (2 * g + 1) * N;
            const uint16_t *x0 = X + (2 * g) * N;
            for (long n = 0; n < N; n++) {
                float acc = dp_step(y[n], a1, oracle_bf16_to_f32(x1[n]));
                y[n] = dp_step(acc, a0, oracle_bf16_to_f32(x0[n]));
            }
        }
    }
}

/* INT8 GEMM: u8(A) . s8(X) in wrapping int32 (vpdpbusd, kernel_int8.h:56-57,113). */
static void gemm_u8s8(const uint8_t *A, const int8_t *X, int32_t *Y, long M, long K, long N)
{
#pragma omp parallel for schedule(static)
    for (long m = 0; m < M; m++) {
        uint32_t *y = (uint32_t *)(Y + m * N);
        for (long n = 0; n < N; n++) y[n] = 0u;
        for (long k = 0; k < K; k++) {
            int32_t a = (int32_t)A[m * K + k];
            const int8_t *x = X + k * N;
            for (long n = 0; n < N; n++) y[n] += (uint32_t)(a * (int32_t)x[n]);
        }
    }
}

static int check_chain(long N, int num_gemm, const long *M, const long *K)
{
    if (N <= 0 || num_gemm <= 0) return -1;
    for (int i = 0; i < num_gemm; i++) {
        if (M[i] <= 0 || K[i] <= 0) return -1;
        if (i > 0 && K[i] != M[i - 1]) return -1;
    }
    return 0;
}

/* ---------------------------------------------------------------- chains -- */

int oracle_chain_f32(float *C, const float *B, long N, int num_gemm,
                     const float *const *A, const long *M, const long *K)
{
    if (check_chain(N, num_gemm, M, K)) return -1;
    const float *X = B;
    float *tmp = NULL;
    for (int i = 0; i < num_gemm; i++) {
        float *Y = (i == num_gemm - 1) ? C : (float *)malloc(sizeof(float) * M[i] * N);
        gemm_f32_seq(A[i], X, Y, M[i], K[i], N);
        free(tmp);
        tmp = (i == num_gemm - 1) ? NULL : Y;
        X = Y;
    }
    return 0;
}

int oracle_chain_bf16(float *C, const uint16_t *B, long N, int num_gemm,
                      const uint16_t *const *A, const long *M, const long *K,
                      int layout, int nthreads)
{
    if (check_chain(N, num_gemm, M, K)) return -1;
    for (int i = 0; i < num_gemm; i++) if (K[i] % 2) return -1; /* flashgemm.c:16 */
    const uint16_t *X = B;
    uint16_t *xbuf = NULL;
    int rc = 0;
    for (int i = 0; i < num_gemm && rc == 0; i++) {
        if (i == num_gemm - 1) { gemm_bf16_pairs(A[i], X, C, M[i], K[i], N); break; }
        float *Y = (float *)malloc(sizeof(float) * M[i] * N);
        uint16_t *Z = (uint16_t *)malloc(sizeof(uint16_t) * M[i] * N);
        gemm_bf16_pairs(A[i], X, Y, M[i], K[i], N);
        for (long e = 0; e < M[i] * N; e++) Z[e] = oracle_bf16_rne(Y[e]);
        free(Y);
        free(xbuf);
        if (layout == ORACLE_LAYOUT_REFERENCE) {
            uint16_t *P = (uint16_t *)malloc(sizeof(uint16_t) * M[i] * N);
            rc = ref_permute(Z, P, M[i], N, sizeof(uint16_t), 0, nthreads);
            free(Z);
            Z = P;
        }
        xbuf = Z;
        X = Z;
    }
    free(xbuf);
    return rc;
}

int oracle_chain_int8(int32_t *C, const uint8_t *B, long N, int num_gemm,
                      const int8_t *const *A, const long *M, const long *K,
                      int layout, int nthreads)
{
    if (check_chain(N, num_gemm, M, K)) return -1;
    const int8_t *X = (const int8_t *)B;
    int8_t *xbuf = NULL;
    int rc = 0;
    for (int i = 0; i < num_gemm && rc == 0; i++) {
        if (i == num_gemm - 1) {
            gemm_u8s8((const uint8_t *)A[i], X, C, M[i], K[i], N);
            break;
        }
        int32_t *Y = (int32_t *)malloc(sizeof(int32_t) * M[i] * N);
        int8_t *Z = (int8_t *)malloc(M[i] * N);
        gemm_u8s8((const uint8_t *)A[i], X, Y, M[i], K[i], N);
        for (long e = 0; e < M[i] * N; e++) {
            int32_t v = Y[e];
            Z[e] = (int8_t)(v > 127 ? 127 : (v < -128 ? -128 : v)); /* vpmovsdb */
        }
        free(Y);
        free(xbuf);
        if (layout == ORACLE_LAYOUT_REFERENCE) {
            int8_t *P = (int8_t *)malloc(M[i] * N);
            rc = ref_permute(Z, P, M[i], N, 1, 1, nthreads);
            free(Z);
            Z = P;
        }
        xbuf = Z;
        X = Z;
    }
    free(xbuf);
    return rc;
}

/* ----------------------------------------------------------- single GEMM -- */

int oracle_sgemm(float *C, const float *A, const float *B, long M, long N, long K)
{
    if (M <= 0 || N <= 0 || K <= 0) return -1;
    const long KB = 2048; /* GEMM_K, singlegemm/src/NN_mp.h:13,247 */
#pragma omp parallel for schedule(static)
    for (long m = 0; m < M; m++) {
        float *acc = (float *)malloc(sizeof(float) * N);
        float *c = C + m * N;
        for (long k0 = 0; k0 < K; k0 += KB) {
            long kc = (K - k0 < KB) ? K - k0 : KB;
            for (long n = 0; n < N; n++) acc[n] = 0.0f;
            for (long k = k0; k < k0 + kc; k++) {
                float a = A[m * K + k];
                const float *b = B + k * N;
                for (long n = 0; n < N; n++) acc[n] = fmaf(a, b[n], acc[n]);
            }
            if (k0 == 0)
                for (long n = 0; n < N; n++) c[n] = acc[n];
            else
                for (long n = 0; n < N; n++) c[n] = c[n] + acc[n];
        }
        free(acc);
    }
    return 0;
}

/* ------------------------------------------------------------------ FP64 -- */

int oracle_chain_f64(double *C, const double *B, long N, int num_gemm,
                     const double *const *A, const long *M, const long *K)
{
    if (check_chain(N, num_gemm, M, K)) return -1;
    const double *X = B;
    double *tmp = NULL;
    for (int i = 0; i < num_gemm; i++) {
        double *Y = (i == num_gemm - 1) ? C : (double *)malloc(sizeof(double) * M[i] * N);
        const double *Ai = A[i];
        long Mi = M[i], Ki = K[i];
#pragma omp parallel for schedule(static)
        for (long m = 0; m < Mi; m++) {
            double *y = Y + m * N;
            for (long n = 0; n < N; n++) y[n] = 0.0;
            for (long k = 0; k < Ki; k++) {
                double a = Ai[m * Ki + k];
                const double *x = X + k * N;
                for (long n = 0; n < N; n++) y[n] += a * x[n];
            }
        }
        free(tmp);
        tmp = (i == num_gemm - 1) ? NULL : Y;
        X = Y;
    }
    return 0;
}
