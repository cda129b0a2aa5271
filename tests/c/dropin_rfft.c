/*
 * dropin_rfft.c -- a plain C11 program written against the reference
 * FastAndFourier API only (include/fastandfourier.h): malloc'd host
 * buffers, create -> execute -> inverse -> destroy, error strings.  It is
 * linked against the MI355X library unchanged to show the drop-in.
 * Exit status 0 when every check holds.
 */
#include <fastandfourier.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define PI 3.14159265358979323846

static int fail(const char *what) {
    fprintf(stderr, "FAIL %s: %s\n", what, faf_get_error());
    return 1;
}

static double rel_err(const float *a, const float *b, size_t n) {
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < n; ++i) {
        num += ((double)a[i] - b[i]) * ((double)a[i] - b[i]);
        den += (double)b[i] * b[i];
    }
    return sqrt(num / (den > 0 ? den : 1.0));
}

int main(void) {
    if (faf_init() != 0) return fail("init");
    /* rfft N=4096 of two tones: the packed spectrum peaks at bins 40 and 300 */
    const size_t n = 4096, nb = n / 2 + 1;
    faf_config cfg = faf_config_init(n);
    faf_transform *fwd = faf_create_rfft(&cfg);
    if (!fwd) return fail("create rfft");
    faf_transform *inv = faf_create_inverse(fwd);
    if (!inv) return fail("create irfft");
    if (faf_spectrum_len(fwd) != nb) return fail("spectrum_len");
    float *x = malloc(n * sizeof(float)), *y = malloc(n * sizeof(float));
    float *re = malloc(nb * sizeof(float)), *im = malloc(nb * sizeof(float));
    for (size_t i = 0; i < n; ++i)
        x[i] = (float)(sin(2 * PI * 40 * (double)i / n) + 0.5 * cos(2 * PI * 300 * (double)i / n));
    faf_buffer bx = faf_buffer_real(x, n), by = faf_buffer_real(y, n);
    faf_buffer bs = faf_buffer_hermitian(re, im, nb);
    if (faf_execute(fwd, &bs, &bx) != 0) return fail("execute rfft");
    /* sin at bin 40 -> -i n/2, cos at bin 300 -> n/4 */
    if (fabs(im[40] + n / 2.0) > 1e-2 * n || fabs(re[300] - n / 4.0) > 1e-2 * n ||
        fabs(re[41]) + fabs(im[41]) > 1e-2 * n)
        return fail("rfft peaks");
    if (faf_execute(inv, &by, &bs) != 0) return fail("execute irfft");
    if (rel_err(y, x, n) > 1e-5) return fail("rfft round trip");
    /* wrong layout is an error, as in the reference */
    faf_buffer bad = faf_buffer_split(re, im, nb);
    if (faf_execute(fwd, &bad, &bx) == 0) return fail("layout check");

    /* mixed-radix C2C N=1000 round trip, NumPy norm */
    const size_t m = 1000;
    faf_config c2 = faf_config_init(m);
    faf_transform *f2 = faf_create_fft(&c2);
    faf_transform *i2 = f2 ? faf_create_inverse(f2) : NULL;
    if (!f2 || !i2) return fail("create fft 1000");
    float *ar = calloc(m, sizeof(float)), *ai = calloc(m, sizeof(float));
    float *br = calloc(m, sizeof(float)), *bi = calloc(m, sizeof(float));
    float *cr = calloc(m, sizeof(float)), *ci = calloc(m, sizeof(float));
    unsigned s = 12345u;
    for (size_t i = 0; i < m; ++i) {
        s = s * 1664525u + 1013904223u;
        ar[i] = (float)((s >> 8) / 16777216.0 - 0.5);
        s = s * 1664525u + 1013904223u;
        ai[i] = (float)((s >> 8) / 16777216.0 - 0.5);
    }
    faf_buffer ba = faf_buffer_split(ar, ai, m), bb = faf_buffer_split(br, bi, m);
    faf_buffer bc = faf_buffer_split(cr, ci, m);
    if (faf_execute(f2, &bb, &ba) != 0 || faf_execute(i2, &bc, &bb) != 0)
        return fail("execute fft 1000");
    if (rel_err(cr, ar, m) > 1e-5 || rel_err(ci, ai, m) > 1e-5) return fail("fft round trip");

    faf_destroy_transform(i2);
    faf_destroy_transform(f2);
    faf_destroy_transform(inv);
    faf_destroy_transform(fwd);
    free(x); free(y); free(re); free(im);
    free(ar); free(ai); free(br); free(bi); free(cr); free(ci);
    faf_cleanup();
    printf("dropin OK\n");
    return 0;
}
