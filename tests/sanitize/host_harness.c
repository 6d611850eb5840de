/* Host-side C-ABI calls that need no GPU, for an AddressSanitizer /
 * UndefinedBehaviorSanitizer build of the library's host sources
 * (tests/test_sanitize.py): the planner over many chain shapes and options, the
 * autotune table's save / load / clear with good and malformed files, status strings,
 * packed sizes, convolution geometry and the serving config defaults. Exit status 0
 * when every call returned what it should; the sanitizers abort on the first error.
 * Test infrastructure only. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "flashgemm.h"

static int fails = 0;
#define CHECK(c)                                                  \
    do {                                                          \
        if (!(c)) {                                               \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            fails++;                                              \
        }                                                         \
    } while (0)

static void write_file(const char *path, const char *text)
{
    FILE *f = fopen(path, "w");
    fputs(text, f);
    fclose(f);
}

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";
    char good[512], bad[512], saved[512];
    snprintf(good, sizeof good, "%s/good.txt", dir);
    snprintf(bad, sizeof bad, "%s/bad.txt", dir);
    snprintf(saved, sizeof saved, "%s/saved.txt", dir);

    /* planner: every dtype and mode, chain lengths 1..8, narrow to very wide GEMMs */
    const long widths[] = {1, 12, 32, 96, 200, 512, 2048, 8192, 65536};
    const int nw = sizeof widths / sizeof widths[0];
    unsigned seed = 7;
    for (int it = 0; it < 3000; it++) {
        seed = seed * 1103515245u + 12345u;
        int dtype = (int)(seed >> 8) % 3, mode = (int)(seed >> 12) % 3, n = 1 + (int)(seed >> 16) % 8;
        long M[8], K[8], N = widths[(seed >> 20) % nw];
        for (int i = 0; i < n; i++) {
            seed = seed * 1103515245u + 12345u;
            M[i] = widths[(seed >> 10) % nw];
            K[i] = i ? M[i - 1] : widths[(seed >> 18) % nw];
        }
        fg_chain_opts o;
        fg_chain_opts_default(&o);
        o.f32_mode = dtype == FG_DTYPE_F32 ? mode : FG_F32_MFMA;
        o.layout = (int)(seed >> 5) % 2;
        o.ref_threads = 1 + (int)(seed >> 7) % 16;
        o.fuse = (int)(seed >> 9) % 3 - 1;
        int fused = -1, tile = -1, ns = 0, ends[8], tiles[8];
        int rc = fg_chain_plan(dtype, N, n, M, K, &o, &fused, &tile);
        CHECK(rc == FG_OK || rc == FG_ERR_UNSUPPORTED || rc == FG_ERR_DOMAIN || rc == FG_ERR_INVALID);
        if (rc == FG_OK) CHECK(fused == 0 || fused == 1);
        rc = fg_chain_plan_segments(dtype, N, n, M, K, &o, &ns, ends, tiles);
        if (rc == FG_OK) CHECK(ns >= 1 && ns <= n && ends[ns - 1] == n - 1);
    }

    /* autotune table: round trip, malformed files refused whole */
    fg_autotune_clear();
    write_file(good, "flashgemm-autotune v2 gfx950\n"
                     "2 3 14 9 10 9 9 9 10 1 2 64 0 0.033\n"
                     "0 4 14 10 12 10 12 10 10 12 10 3 1 32 2 0 3 0 0 1.5\n");
    CHECK(fg_autotune_load(good) == FG_OK);
    CHECK(fg_autotune_size() == 2);
    CHECK(fg_autotune_save(saved) == FG_OK);
    fg_autotune_clear();
    CHECK(fg_autotune_load(saved) == FG_OK && fg_autotune_size() == 2);
    const char *bads[] = {
        "flashgemm-autotune v2 gfx950\n0 2 10 6 6\n",
        "flashgemm-autotune v2 gfx950\n0 99 10 6 6 6 6 1 1 64 0 0.1\n",
        "flashgemm-autotune v2 gfx950\n0 2 10 6 6 6 6 9999 1 64 0 0.1\n",
        "flashgemm-autotune v2 gfx950\n0 2 10 6 6 6 6 -3 1 64 0 0.1\n",
        "flashgemm-autotune v2 gfx950\n0 2 10 6 6 6 6 1 1 64 0 nan-ish\n",
        "flashgemm-autotune v9 gfx950\n",
        "",
        "flashgemm-autotune v2 gfx950\n0 2 10 6 6 6 6 1 1 64 0 0.1 trailing tokens 1 2 3 4 5 6 7 8 9\n",
    };
    for (size_t i = 0; i < sizeof bads / sizeof bads[0]; i++) {
        write_file(bad, bads[i]);
        int rc = fg_autotune_load(bad);
        CHECK(rc != FG_OK || fg_autotune_size() >= 2);
        CHECK(fg_autotune_size() >= 2);
    }
    CHECK(fg_autotune_load("/nonexistent/dir/table.txt") == FG_ERR_INVALID);
    fg_autotune_clear();

    /* small pure functions */
    for (int s = 1; s >= -8; s--) CHECK(fg_status_string(s) != NULL && strlen(fg_status_string(s)) > 0);
    for (int d = -1; d <= 3; d++)
        for (int m = 0; m <= 2; m++) (void)fg_packed_weight_bytes(d, m, 100, 300);
    fg_conv_geom g = {2, 3, 17, 19, 3, 3, 2, 2, 1, 1, 1, 1};
    long K1 = 0, N = 0;
    CHECK(fg_conv_dims(&g, &K1, &N) == FG_OK && K1 == 27);
    g.stride_h = 0;
    CHECK(fg_conv_dims(&g, &K1, &N) == FG_ERR_INVALID);
    fg_server_config c;
    fg_server_config_default(&c);
    CHECK(c.num_streams >= 1 && c.max_batch >= 1);
    printf("host harness: %d failed checks\n", fails);
    return fails != 0;
}
