/* FP32 singlegemm benchmark driver: the structure of singlegemm/FP32_ID23.c on
 * libflashgemm (its 23 shapes, FP32_ID23.c:11-33; 10 warm-up and 10 timed calls of
 * LibShalom_sgemm_mp through the reference's entry point; host pointers, so times
 * include host<->device copies). The reference checks against OpenBLAS with
 * |C - C1| <= 1e-3; this driver checks a sample of 64 entries against a double
 * precision dot product with the same bound. */
#include "drivers_common.h"
#include "flashgemm.h"

static const long MNK[69] = {128, 2048, 4096, 320, 3072, 4096, 1632, 36548, 1024, 2048, 4096, 32,
                             31999, 1024, 80, 84, 1024, 4096, 256, 256, 2048, 1024, 16, 500000,
                             35, 8457, 2560, 64, 18496, 128, 128, 4624, 256, 256, 1156, 512,
                             512, 289, 1024, 256, 3136, 64, 64, 3136, 64, 64, 3136, 256,
                             512, 784, 128, 128, 784, 512, 32, 16384, 288, 144, 4096, 1296,
                             144, 1024, 1296, 192, 1024, 1728, 96, 4096, 864};

int main(void)
{
    LibShalom_set_thread_nums(1);
    for (int j = 0; j < 23; j++) {
        long M = MNK[3 * j], N = MNK[3 * j + 1], K = MNK[3 * j + 2];
        float *A = malloc(sizeof(float) * M * K), *B = malloc(sizeof(float) * K * N);
        float *C = malloc(sizeof(float) * M * N);
        fg_fill_f32(A, M * K, 3 * j + 1);
        fg_fill_f32(B, K * N, 3 * j + 2);
        for (long i = 0; i < M * K; i++) A[i] *= 0.5f; /* drand48()-0.5 range */
        for (long i = 0; i < K * N; i++) B[i] *= 0.5f;
        double t0 = 0, ops = 2.0 * M * N * K;
        for (int it = 0; it < 20; it++) {
            if (it == 10) t0 = fg_now();
            LibShalom_sgemm_mp(0, 0, C, A, B, M, N, K);
        }
        double cost = (fg_now() - t0) / 10;
        int ok = 1;
        uint64_t s = 99 + j;
        for (int t = 0; t < 64 && ok; t++) {
            long m = (long)(fg_splitmix64(&s) % (uint64_t)M), n = (long)(fg_splitmix64(&s) % (uint64_t)N);
            double ref = 0;
            for (long k = 0; k < K; k++) ref += (double)A[m * K + k] * B[k * N + n];
            if (fabs(ref - C[m * N + n]) > 1e-3) ok = 0;
        }
        printf("gpu: id=%-2d M= %-10ld N=%-10ld K=%-10ld GFLOPS= %-10.1f effic= %.3f %%%s\n", j + 1, M, N, K,
               ops / cost * 1e-9, ops / cost * 1e-9 / PEAK_FP32_GFLOPS * 100, ok ? "" : " error!");
        free(A);
        free(B);
        free(C);
    }
    return 0;
}
