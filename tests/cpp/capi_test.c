/* End-to-end check of the C interface, written in C so include/compute_c.h
 * is compiled as a C header: library-owned buffers (clc_buffer_*), then the
 * algorithms on them through clc_buffer_data(), following the scenarios of
 * the reference's C test (lib/test/Compute/capi.c:52-160: buffer copies and
 * bounds, sum / max reduce, scan, transform, copy_if, sort, sequence, fill,
 * a build error carrying its log, an fp64 reduce).
 *
 * Needs a GPU. Prints "capi_test: N checks, 0 failures" and exits 0 on
 * success. */
#include "compute_c.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int n_checks;
static int n_failed;

static void expect_true(int ok, const char* what, int line)
{
    ++n_checks;
    if (!ok)
    {
        ++n_failed;
        fprintf(stderr, "capi_test:%d: %s failed (last error: %s)\n", line, what,
                clc_last_error());
    }
}

#define EXPECT(cond) expect_true((cond) != 0, #cond, __LINE__)
#define EXPECT_OK(call) expect_true((call) == CLC_SUCCESS, #call, __LINE__)

static clc_buffer upload(clc_context ctx, const void* host, size_t bytes)
{
    clc_buffer b = NULL;
    EXPECT_OK(clc_buffer_alloc(ctx, bytes, &b));
    EXPECT_OK(clc_buffer_write(ctx, b, 0, bytes, host));
    return b;
}

static void check_buffers(clc_context ctx)
{
    enum { K = 128 };
    int32_t src[K], got[K];
    clc_buffer a, b = NULL, none = NULL;
    int i;
    for (i = 0; i < K; ++i) src[i] = 7 * i - 300;
    a = upload(ctx, src, sizeof src);
    EXPECT_OK(clc_buffer_alloc(ctx, sizeof src, &b));
    EXPECT_OK(clc_buffer_alloc(ctx, 0, &none));
    EXPECT(clc_buffer_size(a) == sizeof src);
    EXPECT(clc_buffer_size(none) == 0 && clc_buffer_data(none) == NULL);
    EXPECT(clc_buffer_data(a) != NULL);
    /* elements 32..111 of a -> the front of b */
    EXPECT_OK(clc_buffer_copy(ctx, a, 32 * sizeof(int32_t), b, 0, 80 * sizeof(int32_t)));
    memset(got, 0, sizeof got);
    EXPECT_OK(clc_buffer_read(ctx, b, 0, 80 * sizeof(int32_t), got));
    EXPECT(memcmp(got, src + 32, 80 * sizeof(int32_t)) == 0);
    EXPECT(clc_buffer_read(ctx, a, 8, sizeof src, got) == CLC_ERROR_INVALID_ARGUMENT);
    EXPECT(clc_buffer_write(ctx, a, sizeof src + 4, 0, src) == CLC_ERROR_INVALID_ARGUMENT);
    EXPECT(clc_buffer_copy(ctx, a, 0, b, 4, sizeof src) == CLC_ERROR_INVALID_ARGUMENT);
    EXPECT_OK(clc_buffer_write(ctx, none, 0, 0, NULL));
    clc_buffer_release(b);
    EXPECT(clc_buffer_alloc(NULL, 16, &b) == CLC_ERROR_INVALID_ARGUMENT && b == NULL);
    clc_buffer_release(a);
    clc_buffer_release(none);
    clc_buffer_release(NULL);
    EXPECT(clc_buffer_size(NULL) == 0);
}

static void check_algorithms(clc_context ctx)
{
    enum { N = 12011 };
    int32_t* x = malloc(N * sizeof(int32_t));
    int32_t* y = malloc(N * sizeof(int32_t));
    float* f = malloc(N * sizeof(float));
    double* d = malloc(N * sizeof(double));
    uint32_t hist[8], want_hist[8];
    clc_buffer in, out = NULL, fb = NULL, db, hb = NULL;
    int64_t sum = 0;
    int32_t mx = INT32_MIN, r = 0, init = 0, lo = INT32_MIN, start = -4, step = 3, zero = 0;
    size_t kept = 0, want_kept = 0, i;
    double dsum = 0.0, dinit = 0.0, dwant = 0.0;

    for (i = 0; i < N; ++i)
    {
        x[i] = (int32_t)((i * 2654435761u) % 2001u) - 1000;
        sum += x[i];
        if (x[i] > mx) mx = x[i];
        if (x[i] % 4 == 0) ++want_kept;
        d[i] = x[i] / 8.0;
        dwant += d[i];
    }
    in = upload(ctx, x, N * sizeof(int32_t));
    EXPECT_OK(clc_buffer_alloc(ctx, N * sizeof(int32_t), &out));
    EXPECT_OK(clc_buffer_alloc(ctx, N * sizeof(float), &fb));

    EXPECT_OK(clc_reduce(ctx, clc_buffer_data(in), N, CLC_TYPE_I32, "(a + b)", "0", &init, &r));
    EXPECT(r == (int32_t)sum);
    EXPECT_OK(clc_reduce(ctx, clc_buffer_data(in), N, CLC_TYPE_I32, "max(a, b)", "INT_MIN",
                         &lo, &r));
    EXPECT(r == mx);

    EXPECT_OK(clc_scan(ctx, clc_buffer_data(in), clc_buffer_data(out), N, CLC_TYPE_I32,
                       "(a + b)", "0", 0));
    EXPECT_OK(clc_buffer_read(ctx, out, 0, N * sizeof(int32_t), y));
    EXPECT(y[0] == x[0] && y[N - 1] == (int32_t)sum);

    EXPECT_OK(clc_transform(ctx, clc_buffer_data(in), CLC_TYPE_I32, clc_buffer_data(fb),
                            CLC_TYPE_F32, N, "(float)x * 0.25f"));
    EXPECT_OK(clc_buffer_read(ctx, fb, 0, N * sizeof(float), f));
    EXPECT(f[5] == (float)x[5] * 0.25f && f[N - 1] == (float)x[N - 1] * 0.25f);

    EXPECT_OK(clc_copy_if(ctx, clc_buffer_data(in), clc_buffer_data(out), N, CLC_TYPE_I32,
                          "(x & 3) == 0", &kept));
    EXPECT(kept == want_kept);

    EXPECT_OK(clc_sort(ctx, clc_buffer_data(in), CLC_TYPE_I32, NULL, N, 0));
    EXPECT_OK(clc_buffer_read(ctx, in, 0, N * sizeof(int32_t), y));
    for (i = 1; i < N && y[i - 1] <= y[i]; ++i)
        ;
    EXPECT(i == N);

    EXPECT_OK(clc_sequence(ctx, clc_buffer_data(out), N, CLC_TYPE_I32, &start, &step));
    EXPECT_OK(clc_buffer_read(ctx, out, 0, N * sizeof(int32_t), y));
    EXPECT(y[0] == -4 && y[N - 1] == -4 + 3 * (N - 1));
    EXPECT_OK(clc_fill(ctx, clc_buffer_data(out), N, CLC_TYPE_I32, &zero));
    EXPECT_OK(clc_buffer_read(ctx, out, N * sizeof(int32_t) - 4, 4, y));
    EXPECT(y[0] == 0);

    /* 8 bins over [-1000, 1000): bin (x + 1000) / 250. */
    memset(want_hist, 0, sizeof want_hist);
    for (i = 0; i < N; ++i)
        if (x[i] < 1000) ++want_hist[(x[i] + 1000) / 250];
    EXPECT_OK(clc_buffer_alloc(ctx, sizeof hist, &hb));
    EXPECT_OK(clc_histogram_even(ctx, clc_buffer_data(in), N, CLC_TYPE_I32,
                                 (uint32_t*)clc_buffer_data(hb), 8, -1000.0, 1000.0));
    EXPECT_OK(clc_buffer_read(ctx, hb, 0, sizeof hist, hist));
    EXPECT(memcmp(hist, want_hist, sizeof hist) == 0);

    /* An expression that does not compile: a build error with its log. */
    EXPECT(clc_transform(ctx, clc_buffer_data(in), CLC_TYPE_I32, clc_buffer_data(out),
                         CLC_TYPE_I32, N, "x */ nonsense") == CLC_ERROR_BUILD);
    EXPECT(strstr(clc_last_error(), "build log") != NULL);

    db = upload(ctx, d, N * sizeof(double));
    EXPECT_OK(clc_reduce(ctx, clc_buffer_data(db), N, CLC_TYPE_F64, "(a + b)", "0.0", &dinit,
                         &dsum));
    EXPECT(fabs(dsum - dwant) <= 1e-9 * fabs(dwant) + 1e-9);

    EXPECT_OK(clc_finish(ctx));
    clc_buffer_release(in);
    clc_buffer_release(out);
    clc_buffer_release(fb);
    clc_buffer_release(db);
    clc_buffer_release(hb);
    free(x);
    free(y);
    free(f);
    free(d);
}

int main(void)
{
    clc_context ctx = NULL;
    size_t count = 0;
    clc_device_info info;
    EXPECT_OK(clc_device_count(&count));
    EXPECT(count > 0);
    EXPECT_OK(clc_device_get_info(0, &info));
    EXPECT(strlen(info.name) > 0);
    EXPECT(clc_device_get_info(count, &info) == CLC_ERROR_INVALID_ARGUMENT);
    if (clc_context_create(CLC_DEFAULT_DEVICE, &ctx) != CLC_SUCCESS)
    {
        fprintf(stderr, "capi_test: no context: %s\n", clc_last_error());
        return 2;
    }
    check_buffers(ctx);
    check_algorithms(ctx);
    /* A buffer may outlive its context. */
    {
        clc_buffer keep = NULL;
        EXPECT_OK(clc_buffer_alloc(ctx, 4096, &keep));
        clc_context_release(ctx);
        EXPECT(clc_buffer_size(keep) == 4096);
        clc_buffer_release(keep);
    }
    printf("capi_test: %d checks, %d failures\n", n_checks, n_failed);
    return n_failed == 0 ? 0 : 1;
}
