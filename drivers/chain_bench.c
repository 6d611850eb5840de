/* Chain benchmark driver: the structure of the reference's multigemm drivers
 * (multigemm/test/test_{f32,bf16,int8}{,_3}.c) on libflashgemm.
 *
 *   chain_bench <f32|bf16|int8> <1|3|conv|x3> [threads]
 *
 * Mode 1 runs each of the reference's 20 conv-layer GEMM shapes as a one-GEMM
 * chain (multigemm/test/test_f32.c:14-24); mode 3 runs its 10 three-GEMM chains
 * (multigemm/test/test_f32_3.c:14-48). The calls go through the reference's own
 * entry points (host pointers, blocking), so times include host<->device copies.
 * Mode conv runs the same 10 chains as the convolutions they stand for: GEMM 1 is
 * a k x k convolution (k = 3 when 9 divides K1, else 5; pad k/2, stride 1) of a
 * K1/k^2-channel sqrt(N) x sqrt(N) image, through fg_chain_conv_host, which
 * gathers the patches on the GPU instead of taking a formed B. Mode x3 runs the
 * benchmark's 3-GEMM chain (N = 16384, 512 -> 512 -> 1024 -> 512).
 * `threads` sets flashgemm_thread_num (the reference layout's panel count). */
#include "drivers_common.h"
#include "flashgemm.h"

/* (M, N, K) of the first and second GEMMs of the single-GEMM table */
static const long SINGLE[60] = {
    32, 65536, 288, 24, 65536, 32,     144, 16384, 1296, 32, 16384, 144,
    192, 16384, 1728, 32, 16384, 192,  192, 4096, 4800, 56, 4096, 192,
    336, 4096, 8400, 56, 4096, 336,    32, 50176, 288, 16, 50176, 32,
    96, 12544, 864, 24, 12544, 96,     144, 12544, 1296, 24, 12544, 144,
    144, 3136, 1296, 32, 3136, 144,    192, 3136, 1728, 32, 3136, 192};

/* three-GEMM chains: (M1, N, K1), (M2, N, K2), (M3, N, K3) */
static const long CH1[30] = {32, 65536, 288, 144, 16384, 1296, 192, 16384, 1728, 192, 4096, 4800, 336, 4096, 8400,
                             32, 50176, 288, 96, 12544, 864, 144, 12544, 1296, 144, 3136, 1296, 192, 3136, 1728};
static const long CH2[30] = {24, 65536, 32, 32, 16384, 144, 32, 16384, 192, 56, 4096, 192, 56, 4096, 336,
                             16, 50176, 32, 24, 12544, 96, 24, 12544, 144, 32, 3136, 144, 32, 3136, 192};
static const long CH3[30] = {144, 65536, 24, 192, 16384, 32, 192, 16384, 32, 336, 4096, 56, 336, 4096, 56,
                             96, 50176, 16, 144, 12544, 24, 144, 12544, 24, 192, 3136, 32, 192, 3136, 32};

static void run(const char *dt, long N, int n, const long *M, const long *K)
{
    size_t e = dt[0] == 'f' ? 4 : dt[0] == 'b' ? 2 : 1;
    void *A[3];
    void *B = malloc((size_t)K[0] * N * e);
    void *C = malloc((size_t)M[n - 1] * N * 4);
    for (int i = 0; i < n; i++) A[i] = malloc((size_t)M[i] * K[i] * e);
    for (int i = 0; i <= n; i++) {
        void *p = i < n ? A[i] : B;
        long cnt = i < n ? M[i] * K[i] : K[0] * N;
        if (e == 4) fg_fill_f32((float *)p, cnt, 17 + i);
        else if (e == 2) fg_fill_bf16((uint16_t *)p, cnt, 17 + i);
        else fg_fill_i8((int8_t *)p, cnt, -127, 127, 17 + i);
    }
    double ops = 0;
    for (int i = 0; i < n; i++) ops += 2.0 * M[i] * K[i] * N;
    const int warm = 5, loop = 10;
    double t0 = 0;
    for (int it = 0; it < warm + loop; it++) {
        if (it == warm) t0 = fg_now();
#define CALL(fn, T) \
        (n == 1 ? fn((T *)C, B, N, 1, A[0], M[0], K[0]) \
                : fn((T *)C, B, N, 3, A[0], M[0], K[0], A[1], M[1], K[1], A[2], M[2], K[2]))
        if (e == 4) CALL(flashgemm_multi_f32f32f32, float);
        else if (e == 2) CALL(flashgemm_multi_bf16bf16f32, float);
        else CALL(flashgemm_multi_int8uint8int32, int);
#undef CALL
    }
    double cost = (fg_now() - t0) / loop;
    double peak = e == 4 ? PEAK_FP32_GFLOPS : e == 2 ? PEAK_BF16_GFLOPS : PEAK_INT8_GOPS;
    printf("%s: N=%-6ld", dt, N);
    for (int i = 0; i < n; i++) printf(" M%d=%-5ld K%d=%-5ld", i + 1, M[i], i + 1, K[i]);
    printf(" G%s=%-10.1f effic=%.3f %%\n", e == 1 ? "OPS" : "FLOPS", ops / cost * 1e-9, ops / cost * 1e-9 / peak * 100);
    for (int i = 0; i < n; i++) free(A[i]);
    free(B);
    free(C);
}

static void run_conv(const char *dt, long N, const long *M, const long *K)
{
    size_t e = dt[0] == 'f' ? 4 : dt[0] == 'b' ? 2 : 1;
    int dtype = e == 4 ? FG_DTYPE_F32 : e == 2 ? FG_DTYPE_BF16 : FG_DTYPE_INT8;
    long k = K[0] % 9 == 0 ? 3 : 5, side = (long)(sqrt((double)N) + 0.5);
    fg_conv_geom g = {1, K[0] / (k * k), side, side, k, k, 1, 1, k / 2, k / 2, 1, 1};
    long K1, Nc;
    if (fg_conv_dims(&g, &K1, &Nc) != FG_OK || K1 != K[0] || Nc != N) {
        printf("%s conv: no k x k geometry for K1=%ld N=%ld\n", dt, K[0], N);
        return;
    }
    void *A[3];
    void *X = malloc((size_t)K1 / (k * k) * N * e);
    void *C = malloc((size_t)M[2] * N * 4);
    for (int i = 0; i < 3; i++) A[i] = malloc((size_t)M[i] * K[i] * e);
    for (int i = 0; i <= 3; i++) {
        void *p = i < 3 ? A[i] : X;
        long cnt = i < 3 ? M[i] * K[i] : K1 / (k * k) * N;
        if (e == 4) fg_fill_f32((float *)p, cnt, 17 + i);
        else if (e == 2) fg_fill_bf16((uint16_t *)p, cnt, 17 + i);
        else fg_fill_i8((int8_t *)p, cnt, -127, 127, 17 + i);
    }
    double ops = 0;
    for (int i = 0; i < 3; i++) ops += 2.0 * M[i] * K[i] * N;
    const int warm = 5, loop = 10;
    double t0 = 0;
    int rc = FG_OK;
    for (int it = 0; it < warm + loop && rc == FG_OK; it++) {
        if (it == warm) t0 = fg_now();
        rc = fg_chain_conv_host(dtype, C, X, &g, 3, (const void *const *)A, M, K, NULL);
    }
    if (rc != FG_OK) {
        printf("%s conv: %s: %s\n", dt, fg_status_string(rc), fg_last_error());
    } else {
        double cost = (fg_now() - t0) / loop;
        double peak = e == 4 ? PEAK_FP32_GFLOPS : e == 2 ? PEAK_BF16_GFLOPS : PEAK_INT8_GOPS;
        printf("%s conv %ldx%ld, %ld channels, %ldx%ld image: M1=%-4ld M2=%-4ld M3=%-4ld G%s=%-10.1f effic=%.3f %%\n",
               dt, k, k, K1 / (k * k), side, side, M[0], M[1], M[2], e == 1 ? "OPS" : "FLOPS", ops / cost * 1e-9,
               ops / cost * 1e-9 / peak * 100);
    }
    for (int i = 0; i < 3; i++) free(A[i]);
    free(X);
    free(C);
}

int main(int argc, char **argv)
{
    const char *dt = argc > 1 ? argv[1] : "f32";
    const int conv = argc > 2 && strcmp(argv[2], "conv") == 0;
    int mode = argc > 2 && !conv ? atoi(argv[2]) : 3;
    flashgemm_set_thread_num(argc > 3 ? atoi(argv[3]) : 1);
    if (argc > 2 && strcmp(argv[2], "x3") == 0) {
        /* the benchmark's three-GEMM chain X(16384 x 512).W1(512 x 512).W2(512 x 1024).W3(1024 x 512),
           reference orientation: N = 16384, 512 -> 512 -> 1024 -> 512 */
        long M[3] = {512, 1024, 512}, K[3] = {512, 512, 1024};
        run(dt, 16384, 3, M, K);
        return 0;
    }
    for (int j = 0; j < 10; j++) {
        if (conv) {
            long M[3] = {CH1[3 * j], CH2[3 * j], CH3[3 * j]};
            long K[3] = {CH1[3 * j + 2], CH2[3 * j + 2], CH3[3 * j + 2]};
            run_conv(dt, CH1[3 * j + 1], M, K);
        } else if (mode == 1) {
            for (int h = 0; h < 2; h++) {
                long M = SINGLE[6 * j + 3 * h], N = SINGLE[6 * j + 3 * h + 1], K = SINGLE[6 * j + 3 * h + 2];
                run(dt, N, 1, &M, &K);
            }
        } else {
            long M[3] = {CH1[3 * j], CH2[3 * j], CH3[3 * j]};
            long K[3] = {CH1[3 * j + 2], CH2[3 * j + 2], CH3[3 * j + 2]};
            run(dt, CH1[3 * j + 1], 3, M, K);
        }
    }
    return 0;
}
