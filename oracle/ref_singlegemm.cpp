// Test-side harness: exposes the reference's FP32 singlegemm through a C symbol so
// the golden-fixture script and the CPU baseline can call it. It is compiled by
// oracle/Makefile against the reference's own headers, read in place
// (-I <reference>/singlegemm/src); no reference source is copied here.
//
// The reference defines its dispatcher inside its benchmark driver
// (singlegemm/FP32_ID23.c:35-62), so this harness restates that dispatch:
// M/N >= 8 -> Small_MGN_NN_SGEMM with Tm = threads, Tn = 1; otherwise
// Dete_grad_N_threads_nums then Small_NGM_NN_SGEMM. NN mode only.
#include <stdio.h>
#include <stdlib.h>
#include "NN_mp.h"
#include "set_TmTn.h"

static int g_threads = 1;
static bool g_ready = false;

extern "C" void ref_singlegemm_set_threads(int threads)
{
    // LibShalom_set_thread_nums appends the divisors of `threads` to a global list
    // (set_TmTn.h:61-74); clear it first so a second call does not mix two counts.
    vec.clear();
    LibShalom_set_thread_nums(threads);
    g_threads = threads;
    g_ready = true;
}

extern "C" int ref_singlegemm_threads(void) { return g_threads; }

extern "C" void ref_sgemm(float *C, float *A, float *B, long M, long N, long K)
{
    if (!g_ready) ref_singlegemm_set_threads(1);
    if ((M / N) >= 8) {
        Tm = g_threads;
        Tn = g_threads / Tm;
        Small_MGN_NN_SGEMM(C, A, B, M, N, K, g_threads, Tm, Tn);
    } else {
        Dete_grad_N_threads_nums(g_threads, M, N, 0, 0);
        Small_NGM_NN_SGEMM(C, A, B, M, N, K, g_threads, Tm, Tn);
    }
}
