/* Shared helpers for the benchmark drivers (host side, plain C).
 *
 * The drivers keep the structure of the reference's drivers (shape tables, warm-up
 * calls, a timed loop, GFLOPS and efficiency printout; multigemm/test/test_*.c,
 * singlegemm/FP32_ID23.c) and call the library only through include/flashgemm.h.
 * Differences: the timed loop divides by the number of calls it makes (the
 * reference runs loop+1 calls and divides by loop, multigemm/test/test_f32.c:77-82),
 * efficiency is against MI355X peaks, and inputs come from a seeded splitmix64
 * generator instead of rand(). */
#ifndef FG_DRIVERS_COMMON_H
#define FG_DRIVERS_COMMON_H

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

/* MI355X dense peaks (256 CUs x 2.4 GHz), in GFLOPS / GOPS. */
#define PEAK_FP32_GFLOPS 157286.4
#define PEAK_BF16_GFLOPS 2516582.4
#define PEAK_INT8_GOPS 5033164.8

static double fg_now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + tv.tv_usec * 1.0e-6;
}

static uint64_t fg_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform [-1, 1) */
static void fg_fill_f32(float *a, long n, uint64_t seed)
{
    uint64_t s = seed;
    for (long i = 0; i < n; i++) a[i] = (float)((fg_splitmix64(&s) >> 40) * (2.0 / 16777216.0) - 1.0);
}

/* uniform [-1, 1) rounded to bf16 (nearest even) */
static void fg_fill_bf16(uint16_t *a, long n, uint64_t seed)
{
    uint64_t s = seed;
    for (long i = 0; i < n; i++) {
        float x = (float)((fg_splitmix64(&s) >> 40) * (2.0 / 16777216.0) - 1.0);
        uint32_t u;
        memcpy(&u, &x, 4);
        a[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
}

/* uniform integers in [lo, hi] */
static void fg_fill_i8(int8_t *a, long n, int lo, int hi, uint64_t seed)
{
    uint64_t s = seed;
    for (long i = 0; i < n; i++) a[i] = (int8_t)(lo + (int)(fg_splitmix64(&s) % (uint64_t)(hi - lo + 1)));
}

#endif
