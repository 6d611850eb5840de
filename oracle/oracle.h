/*
 * oracle.h - CPU restatement of the reference flashgemm arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. The product library (flashgemm_amd/) never links,
 * loads or calls anything declared here. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Every routine restates one piece of the reference (file:line cited beside
 * each declaration) in plain scalar C. Each is pinned by the checks listed
 * beside it; the pinning tests live in tests/test_oracle.py.
 */
#ifndef FLASHGEMM_ORACLE_H
#define FLASHGEMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout of the intermediate that GEMM i+1 reads (see oracle.c for the maps). */
enum {
    ORACLE_LAYOUT_MATH = 0,      /* X_{i+1} = requant(Y_i): the mathematical chain */
    ORACLE_LAYOUT_REFERENCE = 1  /* the reference's packed-B re-interpretation     */
};

/* bf16 helpers.
 * oracle_bf16_rne   : vcvtne2ps2bf16 (multigemm/src/kernel_bf16.h:357), RNE, denormal
 *                     inputs read as signed zero, NaN quieted.
 *                     Pinned: textbook RNE cases + golden BF16 chain fixtures.
 * oracle_bf16_trunc : float_to_bf16 (multigemm/src/utils.h:20-33), truncation.
 * oracle_bf16_to_f32: bf16_to_float (multigemm/src/utils.h:5-18). */
uint16_t oracle_bf16_rne(float x);
uint16_t oracle_bf16_trunc(float x);
float oracle_bf16_to_f32(uint16_t b);

/* Thread partition of N (multigemm/src/flashgemm.c:40-51, :109-120).
 * Returns nb (columns per thread, multiple of 16 unless exact), ne (= N % nb) and
 * num_n (threads that own columns). Pinned: golden chain fixtures at 1, 2, 3 threads. */
void oracle_partition(long N, int nthreads, long *nb, long *ne, long *num_n);

/* FP32 chain C = A_n ... A_1 B (multigemm/src/flashgemm.c:24-156, kernel_f32.h:917-942):
 * each output element is a sequential fmaf chain over k = 0..K-1; the fp32 intermediate
 * is consumed unrounded. Pinned: golden fixtures (bit-exact), ones-chain closed form.
 * Returns 0, or -1 for an invalid shape. */
int oracle_chain_f32(float *C, const float *B, long N, int num_gemm,
                     const float *const *A, const long *M, const long *K);

/* BF16 chain (multigemm/src/flashgemm.c:158-290, kernel_bf16.h): per k-pair g,
 * acc += a[2g+1]*b[2g+1] then acc += a[2g]*b[2g] (vdpbf16ps order, fp32 RNE, denormal
 * flush); intermediate = oracle_bf16_rne(Y); final C fp32.
 * layout = ORACLE_LAYOUT_REFERENCE applies the reference's intermediate re-interpretation
 * for the panel decomposition of `nthreads` threads. Pinned: golden fixtures (bit-exact). */
int oracle_chain_bf16(float *C, const uint16_t *B, long N, int num_gemm,
                      const uint16_t *const *A, const long *M, const long *K,
                      int layout, int nthreads);

/* INT8 chain (multigemm/src/flashgemm.c:292-424, kernel_int8.h): Y = u8(A) . s8(X) in
 * wrapping int32; intermediate = clamp(Y, -128, 127) read as s8 by the next GEMM
 * (vpmovsdb, kernel_int8.h:330-333); final C int32. Layout as for BF16.
 * Pinned: golden fixtures, brute force on tiny shapes, saturation closed forms. */
int oracle_chain_int8(int32_t *C, const uint8_t *B, long N, int num_gemm,
                      const int8_t *const *A, const long *M, const long *K,
                      int layout, int nthreads);

/* FP32 single GEMM C = A.B (singlegemm/src/NN_mp.h:141-229, kernel.h:384-454,1113-1116):
 * K is cut into blocks of 2048; within a block each element is a sequential fmaf chain;
 * the first block stores, later blocks add (C = C + acc). Pinned: golden fixtures from
 * the compiled reference (NGM and MGN paths), FP64 bound. */
int oracle_sgemm(float *C, const float *A, const float *B, long M, long N, long K);

/* FP64 chain in the mathematical layout, used for error norms only. Pinned against
 * numpy float64 matmul. */
int oracle_chain_f64(double *C, const double *B, long N, int num_gemm,
                     const double *const *A, const long *M, const long *K);

#ifdef __cplusplus
}
#endif
#endif
