/*
 * faf_oracle.c -- CPU oracle for the transform execution path.
 *
 * TEST INFRASTRUCTURE ONLY (see faf_oracle.h).  This file is never linked
 * into the product library.  It restates, in plain C11, the algorithms of
 * the reference's CPU path so that the GPU results can be compared with
 * (a) FP64 ground truth and (b) what the reference itself computes.
 */

#include "faf_oracle.h"
#include "fastandfourier.h"   /* the reference's ABI (identical header), for orc_time_ref_rfft_batch */
#include <dlfcn.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* integer size policy -- reference src/faf_fft_stage.c:25-135,       */
/* src/faf_fft_kernels.c:196-205, src/faf_core.c:580-593               */
/* ------------------------------------------------------------------ */

static size_t strip(size_t n, size_t p) { while (n % p == 0) n /= p; return n; }

int orc_is_5_smooth(size_t n) {
    return n != 0 && strip(strip(strip(n, 2), 3), 5) == 1;
}

int orc_is_7_smooth(size_t n) {
    return n != 0 && strip(strip(strip(strip(n, 2), 3), 5), 7) == 1;
}

int orc_is_prime(size_t n) {
    if (n < 2) return 0;
    if (n % 2 == 0) return n == 2;
    for (size_t d = 3; d * d <= n; d += 2)
        if (n % d == 0) return 0;
    return 1;
}

int orc_rader_eligible(size_t n) {          /* faf_fft_stage.c:49-51 */
    return n >= 11 && orc_is_prime(n) && orc_is_7_smooth(n - 1);
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    for (; e; e >>= 1, b = b * b % m)
        if (e & 1) r = r * b % m;
    return r;
}

size_t orc_primitive_root(size_t p) {       /* faf_fft_stage.c:64-93 */
    if (p < 3 || !orc_is_prime(p)) return 0;
    size_t phi = p - 1, m = phi, f[32];
    int nf = 0;
    for (size_t d = 2; d * d <= m; d++)
        if (m % d == 0) { f[nf++] = d; m = strip(m, d); }
    if (m > 1) f[nf++] = m;
    for (size_t g = 2; g < p; g++) {
        int ok = 1;
        for (int i = 0; i < nf && ok; i++)
            if (powmod(g, phi / f[i], p) == 1) ok = 0;
        if (ok) return g;
    }
    return 0;
}

int orc_factor_7smooth(size_t n, int *fac, int *nfac) {  /* faf_fft_stage.c:95-105 */
    static const int order[5] = {4, 2, 3, 5, 7};
    int k = 0;
    if (n == 0) return -1;
    for (int o = 0; o < 5; o++)
        while (n % (size_t)order[o] == 0 && k < 16) { fac[k++] = order[o]; n /= (size_t)order[o]; }
    *nfac = k;
    return n == 1 ? 0 : -1;
}

size_t orc_digit_reverse(size_t i, const int *fac, int nfac) {  /* faf_fft_stage.c:115-124 */
    size_t r = 0;
    for (int k = 0; k < nfac; k++) {
        r = r * (size_t)fac[k] + i % (size_t)fac[k];
        i /= (size_t)fac[k];
    }
    return r;
}

size_t orc_next_5_smooth(size_t n) {        /* faf_fft_stage.c:126-135 */
    if (n <= 1) return 1;
    while (!orc_is_5_smooth(n)) n++;
    return n;
}

int orc_is_codelet_size(size_t n) {         /* faf_fft_kernels.c:196-205 */
    return (n >= 2 && n <= 10) || n == 12;
}

int orc_fft_size_legal(size_t n, int bluestein_flag) {   /* faf_core.c:580-593 */
    if (orc_is_5_smooth(n) || orc_is_codelet_size(n)) return 1;
    if (orc_rader_eligible(n)) return 2;
    if (bluestein_flag && n >= 1) return 3;
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 ground truth                                                   */
/* ------------------------------------------------------------------ */

/* exp(sign*2*pi*i*m/n) for m in [0,n): exact integer phase index. */
static void phase_table(size_t n, int inverse, double *c, double *s) {
    double sg = inverse ? 1.0 : -1.0;
    for (size_t m = 0; m < n; m++) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        c[m] = cos(a);
        s[m] = sg * sin(a);
    }
}

void orc_dft_f64(size_t n, int inverse, const double *xr, const double *xi,
                 double *yr, double *yi) {
    double *c = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    phase_table(n, inverse, c, s);
    for (size_t k = 0; k < n; k++) {
        double ar = 0.0, ai = 0.0;
        size_t m = 0;                       /* m = j*k mod n, updated exactly */
        for (size_t j = 0; j < n; j++) {
            double xre = xr[j], xim = xi ? xi[j] : 0.0;
            ar += xre * c[m] - xim * s[m];
            ai += xre * s[m] + xim * c[m];
            m += k;
            if (m >= n) m -= n;
        }
        yr[k] = ar;
        yi[k] = ai;
    }
    free(c);
    free(s);
}

void orc_dft_bin_f64(size_t n, int inverse, const double *xr, const double *xi,
                     size_t k, double *yr, double *yi) {
    double sg = inverse ? 1.0 : -1.0, ar = 0.0, ai = 0.0;
    size_t m = 0;
    for (size_t j = 0; j < n; j++) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        double c = cos(a), s = sg * sin(a);
        double xre = xr[j], xim = xi ? xi[j] : 0.0;
        ar += xre * c - xim * s;
        ai += xre * s + xim * c;
        m = (m + k) % n;
    }
    *yr = ar;
    *yi = ai;
}

static size_t smallest_factor(size_t n) {
    for (size_t p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

/* Recursive decimation-in-time over the smallest prime factor; every
 * twiddle is looked up in an n-entry FP64 table at an exact index. */
static void fft64_rec(size_t n, size_t stride, const double *xr, const double *xi,
                      double *yr, double *yi, const double *c, const double *s,
                      size_t tstep, double *tmp) {
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi ? xi[0] : 0.0; return; }
    size_t p = smallest_factor(n), m = n / p;
    if (p == n) {                          /* prime: direct DFT */
        for (size_t k = 0; k < n; k++) {
            double ar = 0.0, ai = 0.0;
            size_t e = 0;
            for (size_t j = 0; j < n; j++) {
                double xre = xr[j * stride], xim = xi ? xi[j * stride] : 0.0;
                size_t ix = e * tstep;
                ar += xre * c[ix] - xim * s[ix];
                ai += xre * s[ix] + xim * c[ix];
                e += k;
                if (e >= n) e -= n;
            }
            yr[k] = ar;
            yi[k] = ai;
        }
        return;
    }
    /* sub-transforms of the p decimated sequences into y[r*m .. r*m+m) */
    for (size_t r = 0; r < p; r++)
        fft64_rec(m, stride * p, xr + r * stride, xi ? xi + r * stride : NULL,
                  yr + r * m, yi + r * m, c, s, tstep * p, tmp);
    double *tr = tmp, *ti = tmp + n;
    for (size_t q = 0; q < p; q++)
        for (size_t k = 0; k < m; k++) {
            size_t kk = k + q * m;
            double ar = 0.0, ai = 0.0;
            for (size_t r = 0; r < p; r++) {
                size_t ix = ((r * kk) % n) * tstep;
                double vr = yr[r * m + k], vi = yi[r * m + k];
                ar += vr * c[ix] - vi * s[ix];
                ai += vr * s[ix] + vi * c[ix];
            }
            tr[kk] = ar;
            ti[kk] = ai;
        }
    memcpy(yr, tr, n * sizeof(double));
    memcpy(yi, ti, n * sizeof(double));
}

int orc_fft_f64(size_t n, int inverse, const double *xr, const double *xi,
                double *yr, double *yi) {
    if (n == 0) return -1;
    double *c = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    double *tmp = malloc(2 * n * sizeof(double) * 2);
    phase_table(n, inverse, c, s);
    /* recursion uses tmp levels sequentially; each level needs 2n scratch,
     * and a level finishes its children before touching tmp itself */
    fft64_rec(n, 1, xr, xi, yr, yi, c, s, 1, tmp);
    free(c); free(s); free(tmp);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP32 restatement: plans                                             */
/* ------------------------------------------------------------------ */

enum { K_ONE, K_CODELET, K_SPLITRADIX, K_MIXED, K_RADER, K_BLUESTEIN };

typedef struct oplan {
    size_t n;
    int inverse, kind;
    float *tw;            /* SR level tables, or full mixed table (2n) */
    float *work;          /* 2n */
    size_t *gather;       /* mixed: digit-reversed load order */
    int fac[16], nf;
    struct oplan *inner, *inner_inv;
    size_t *rin, *rout;
    float *bre, *bim;     /* Rader B or Bluestein B (length m or M) */
    float *phr, *phi;     /* Bluestein chirp (n) */
    float *ar, *ai;       /* work (m or M) */
    float *cr, *ci;       /* Bluestein second work (M) */
} oplan;

static oplan *plan_create(size_t n, int inverse, int allow7, int blue_flag);
static void plan_exec(const oplan *p, const float *xr, const float *xi,
                      float *yr, float *yi);

static void plan_free(oplan *p) {
    if (!p) return;
    plan_free(p->inner);
    plan_free(p->inner_inv);
    free(p->tw); free(p->work); free(p->gather); free(p->rin); free(p->rout);
    free(p->bre); free(p->bim); free(p->phr); free(p->phi); free(p->ar);
    free(p->ai); free(p->cr); free(p->ci);
    free(p);
}

/* --- codelets: reference src/faf_fft_kernels.c:19-292 ------------- */

static void codelet_direct(float *re, float *im, size_t n, int inverse) {
    /* src/faf_fft_kernels.c:206-226: direct DFT, twiddles from double */
    float xr[12], xi[12], wr[12], wi[12];
    double sg = inverse ? 1.0 : -1.0;
    for (size_t t = 0; t < n; t++) {
        xr[t] = re[t]; xi[t] = im[t];
        double a = sg * 2.0 * M_PI * (double)t / (double)n;
        wr[t] = (float)cos(a); wi[t] = (float)sin(a);
    }
    for (size_t k = 0; k < n; k++) {
        float sr = 0.0f, si = 0.0f;
        for (size_t t = 0; t < n; t++) {
            size_t e = (k * t) % n;
            sr += xr[t] * wr[e] - xi[t] * wi[e];
            si += xr[t] * wi[e] + xi[t] * wr[e];
        }
        re[k] = sr; im[k] = si;
    }
}

static void codelet4(float *re, float *im, int inverse) {   /* :37-59 */
    float s0r = re[0] + re[2], s0i = im[0] + im[2];
    float d0r = re[0] - re[2], d0i = im[0] - im[2];
    float s1r = re[1] + re[3], s1i = im[1] + im[3];
    float d1r = re[1] - re[3], d1i = im[1] - im[3];
    re[0] = s0r + s1r; im[0] = s0i + s1i;
    re[2] = s0r - s1r; im[2] = s0i - s1i;
    /* forward: y1 = d0 - i d1, y3 = d0 + i d1; inverse swaps */
    float g = inverse ? -1.0f : 1.0f;
    re[1] = d0r + g * d1i; im[1] = d0i - g * d1r;
    re[3] = d0r - g * d1i; im[3] = d0i + g * d1r;
}

static void codelet8(float *re, float *im, int inverse) {   /* :97-189, radix-2 DIT */
    static const int br[8] = {0, 4, 2, 6, 1, 5, 3, 7};
    float r[8], i[8];
    for (int k = 0; k < 8; k++) { r[k] = re[br[k]]; i[k] = im[br[k]]; }
    const float h = 0.70710678118654752440f, sg = inverse ? 1.0f : -1.0f;
    for (int len = 2; len <= 8; len <<= 1) {
        for (int b = 0; b < 8; b += len)
            for (int j = 0; j < len / 2; j++) {
                float wr, wi;
                int e = j * (8 / len);       /* W_8^e */
                if (e == 0) { wr = 1.0f; wi = 0.0f; }
                else if (e == 1) { wr = h; wi = sg * h; }
                else if (e == 2) { wr = 0.0f; wi = sg; }
                else { wr = -h; wi = sg * h; }
                int a = b + j, c = a + len / 2;
                float tr = r[c] * wr - i[c] * wi, ti = r[c] * wi + i[c] * wr;
                r[c] = r[a] - tr; i[c] = i[a] - ti;
                r[a] += tr; i[a] += ti;
            }
    }
    memcpy(re, r, sizeof r); memcpy(im, i, sizeof i);
}

static void codelet_exec(size_t n, int inverse, float *re, float *im) {
    if (n == 2) {
        float a = re[0], b = re[1], c = im[0], d = im[1];
        re[0] = a + b; re[1] = a - b; im[0] = c + d; im[1] = c - d;
    } else if (n == 4) codelet4(re, im, inverse);
    else if (n == 8) codelet8(re, im, inverse);
    else codelet_direct(re, im, n, inverse);
}

/* --- split-radix DIF: reference src/faf_fft_splitradix.c:23-369 --- */

static size_t sr_table_floats(size_t n) {
    size_t c = 0;
    for (size_t m = n; m >= 16; m >>= 1) c += m;
    return c;
}

static void sr_tables(float *tw, size_t n, int inverse) {   /* :30-43 */
    double sg = inverse ? 1.0 : -1.0;
    size_t o = 0;
    for (size_t m = n; m >= 16; m >>= 1)
        for (size_t k = 0; k < m / 4; k++) {
            double a = sg * 2.0 * M_PI * (double)k / (double)m;
            tw[o++] = (float)cos(a);     tw[o++] = (float)sin(a);
            tw[o++] = (float)cos(3 * a); tw[o++] = (float)sin(3 * a);
        }
}

static void sr_rec(size_t n, const float *xr, const float *xi, float *tr, float *ti,
                   float *yr, float *yi, size_t sy, const float *tw, int inverse) {
    if (n <= 8) {                               /* leaves :244-282 */
        float r[8], i[8];
        for (size_t k = 0; k < n; k++) { r[k] = xr[k]; i[k] = xi[k]; }
        if (n > 1) codelet_exec(n, inverse, r, i);
        for (size_t k = 0; k < n; k++) { yr[k * sy] = r[k]; yi[k * sy] = i[k]; }
        return;
    }
    size_t q = n / 4, h = n / 2;
    for (size_t k = 0; k < q; k++) {            /* L-butterfly :323-359 */
        float ar = xr[k], ai = xi[k], br = xr[k + h], bi = xi[k + h];
        float cr = xr[k + q], ci = xi[k + q], dr = xr[k + 3 * q], di = xi[k + 3 * q];
        float er = cr - dr, ei = ci - di;       /* t3 = +-j*(c-d) */
        float jr = inverse ? ei : -ei, ji = inverse ? -er : er;
        float ur = (ar - br) - jr, ui = (ai - bi) - ji;
        float vr = (ar - br) + jr, vi = (ai - bi) + ji;
        const float *w = tw + 4 * k;
        tr[k] = ar + br; ti[k] = ai + bi;
        tr[k + q] = cr + dr; ti[k + q] = ci + di;
        tr[k + h] = ur * w[0] - ui * w[1]; ti[k + h] = ur * w[1] + ui * w[0];
        tr[k + 3 * q] = vr * w[2] - vi * w[3]; ti[k + 3 * q] = vr * w[3] + vi * w[2];
    }
    const float *nx = tw + n;
    sr_rec(h, tr, ti, tr, ti, yr, yi, 2 * sy, nx, inverse);
    sr_rec(q, tr + h, ti + h, tr + h, ti + h, yr + sy, yi + sy, 4 * sy, nx + h, inverse);
    sr_rec(q, tr + 3 * q, ti + 3 * q, tr + 3 * q, ti + 3 * q, yr + 3 * sy, yi + 3 * sy,
           4 * sy, nx + h, inverse);
}

/* --- mixed-radix DIT bytecode: reference src/faf_transforms.c:147-187,
 *     stage kernels src/faf_fft_stage.c:163-407 ---------------------- */

static void tw_full_f32(float *tw, size_t n, int inverse) {  /* faf_twiddles.c:43-51 */
    const float sg = inverse ? 1.0f : -1.0f;
    const float step = 2.0f * (float)M_PI / (float)n;          /* float angle */
    for (size_t k = 0; k < n; k++) {
        float a = sg * step * (float)k;
        tw[2 * k] = cosf(a);
        tw[2 * k + 1] = sinf(a);
    }
}

static void dft_small(float *vr, float *vi, int r, int inverse) {
    /* radix-3/5/7 leg DFT with table constants (faf_fft_stage.c:203-345) */
    float yr[7], yi[7], wr[7], wi[7];
    double sg = inverse ? 1.0 : -1.0;
    for (int e = 0; e < r; e++) {
        wr[e] = (float)cos(2.0 * M_PI * e / r);
        wi[e] = (float)(sg * sin(2.0 * M_PI * e / r));
    }
    for (int k = 0; k < r; k++) {
        float sr = 0.0f, si = 0.0f;
        for (int j = 0; j < r; j++) {
            int e = (j * k) % r;
            sr += vr[j] * wr[e] - vi[j] * wi[e];
            si += vr[j] * wi[e] + vi[j] * wr[e];
        }
        yr[k] = sr; yi[k] = si;
    }
    memcpy(vr, yr, (size_t)r * sizeof(float));
    memcpy(vi, yi, (size_t)r * sizeof(float));
}

static void mixed_stage(float *re, float *im, size_t n, size_t group, int r,
                        size_t tws, const float *tw, int inverse) {
    size_t m = group / (size_t)r;
    for (size_t base = 0; base < n; base += group)
        for (size_t j = 0; j < m; j++) {
            float vr[7], vi[7];
            for (int p = 0; p < r; p++) {
                size_t ix = base + j + (size_t)p * m;
                float a = re[ix], b = im[ix];
                if (p > 0) {
                    const float *w = tw + 2 * (j * (size_t)p * tws);
                    float t = a * w[0] - b * w[1];
                    b = a * w[1] + b * w[0];
                    a = t;
                }
                vr[p] = a; vi[p] = b;
            }
            if (r == 2) {
                float a = vr[0], b = vi[0];
                vr[0] = a + vr[1]; vi[0] = b + vi[1];
                vr[1] = a - vr[1]; vi[1] = b - vi[1];
            } else if (r == 4) {
                codelet4(vr, vi, inverse);
            } else {
                dft_small(vr, vi, r, inverse);
            }
            for (int p = 0; p < r; p++) {
                size_t ix = base + j + (size_t)p * m;
                re[ix] = vr[p]; im[ix] = vi[p];
            }
        }
}

/* --- planner: reference src/faf_core.c:572-719 --------------------- */

static oplan *plan_create(size_t n, int inverse, int allow7, int blue_flag) {
    oplan *p = calloc(1, sizeof(oplan));
    p->n = n;
    p->inverse = inverse;
    if (n == 1) { p->kind = K_ONE; return p; }
    int legal = orc_is_5_smooth(n) || orc_is_codelet_size(n) ||
                (allow7 && orc_is_7_smooth(n));
    if (!legal && orc_rader_eligible(n)) {
        /* reference src/faf_fft_rader.c:19-134 */
        size_t m = n - 1, g = orc_primitive_root(n);
        size_t ginv = (size_t)powmod(g, n - 2, n);
        p->kind = K_RADER;
        p->inner = plan_create(m, 0, 1, 0);
        p->inner_inv = plan_create(m, 1, 1, 0);
        p->rin = malloc(m * sizeof(size_t));
        p->rout = malloc(m * sizeof(size_t));
        size_t a = 1, b = 1;
        for (size_t j = 0; j < m; j++) {
            p->rin[j] = a; p->rout[j] = b;
            a = a * g % n; b = b * ginv % n;
        }
        p->bre = malloc(m * sizeof(float)); p->bim = malloc(m * sizeof(float));
        p->ar = malloc(m * sizeof(float)); p->ai = malloc(m * sizeof(float));
        double sg = inverse ? 1.0 : -1.0;
        for (size_t j = 0; j < m; j++) {
            double ang = sg * 2.0 * M_PI * (double)p->rout[j] / (double)n;
            p->ar[j] = (float)cos(ang); p->ai[j] = (float)sin(ang);
        }
        plan_exec(p->inner, p->ar, p->ai, p->bre, p->bim);   /* B = FFT_m(w) */
        return p;
    }
    if (!legal) {
        if (!blue_flag) { free(p); return NULL; }
        /* reference src/faf_bluestein.c:19-163 */
        size_t M = orc_next_5_smooth(2 * n - 1);
        p->kind = K_BLUESTEIN;
        p->inner = plan_create(M, 0, 0, 0);
        p->inner_inv = plan_create(M, 1, 0, 0);
        p->phr = malloc(n * sizeof(float)); p->phi = malloc(n * sizeof(float));
        double sg = inverse ? 1.0 : -1.0;
        for (size_t k = 0; k < n; k++) {
            double ang = sg * M_PI * ((double)k * (double)k) / (double)n;
            p->phr[k] = (float)cos(ang); p->phi[k] = (float)sin(ang);
        }
        p->ar = calloc(M, sizeof(float)); p->ai = calloc(M, sizeof(float));
        p->cr = calloc(M, sizeof(float)); p->ci = calloc(M, sizeof(float));
        p->bre = malloc(M * sizeof(float)); p->bim = malloc(M * sizeof(float));
        p->ar[0] = 1.0f;
        for (size_t l = 1; l < n; l++) {
            p->ar[l] = p->phr[l]; p->ai[l] = -p->phi[l];
            p->ar[M - l] = p->phr[l]; p->ai[M - l] = -p->phi[l];
        }
        plan_exec(p->inner, p->ar, p->ai, p->bre, p->bim);
        memset(p->ar, 0, M * sizeof(float)); memset(p->ai, 0, M * sizeof(float));
        return p;
    }
    p->work = malloc(2 * n * sizeof(float));
    if (orc_is_codelet_size(n)) { p->kind = K_CODELET; return p; }
    if ((n & (n - 1)) == 0 && n >= 16) {
        p->kind = K_SPLITRADIX;
        p->tw = malloc(sr_table_floats(n) * sizeof(float));
        sr_tables(p->tw, n, inverse);
        return p;
    }
    p->kind = K_MIXED;
    orc_factor_7smooth(n, p->fac, &p->nf);
    p->tw = malloc(2 * n * sizeof(float));
    tw_full_f32(p->tw, n, inverse);
    p->gather = malloc(n * sizeof(size_t));
    for (size_t i = 0; i < n; i++) p->gather[i] = orc_digit_reverse(i, p->fac, p->nf);
    return p;
}

static void plan_exec(const oplan *p, const float *xr, const float *xi,
                      float *yr, float *yi) {
    size_t n = p->n;
    switch (p->kind) {
    case K_ONE:
        yr[0] = xr[0]; yi[0] = xi[0];
        return;
    case K_CODELET:
        if (yr != xr) memcpy(yr, xr, n * sizeof(float));
        if (yi != xi) memcpy(yi, xi, n * sizeof(float));
        codelet_exec(n, p->inverse, yr, yi);
        return;
    case K_SPLITRADIX: {
        float *wr = p->work, *wi = p->work + n;
        if (yr == xr) {           /* in place: reference copies through scratch */
            float *cr = malloc(2 * n * sizeof(float));
            memcpy(cr, xr, n * sizeof(float)); memcpy(cr + n, xi, n * sizeof(float));
            sr_rec(n, cr, cr + n, wr, wi, yr, yi, 1, p->tw, p->inverse);
            free(cr);
        } else {
            sr_rec(n, xr, xi, wr, wi, yr, yi, 1, p->tw, p->inverse);
        }
        return;
    }
    case K_MIXED: {
        /* gather into the work planes then stage in place (faf_vm.c:1426-1500) */
        float *wr = p->work, *wi = p->work + n;
        for (size_t i = 0; i < n; i++) { wr[i] = xr[p->gather[i]]; wi[i] = xi[p->gather[i]]; }
        size_t group = 1;
        for (int k = 0; k < p->nf; k++) {
            group *= (size_t)p->fac[k];
            mixed_stage(wr, wi, n, group, p->fac[k], n / group, p->tw, p->inverse);
        }
        memcpy(yr, wr, n * sizeof(float)); memcpy(yi, wi, n * sizeof(float));
        return;
    }
    case K_RADER: {               /* reference src/faf_fft_rader.c:192-231 */
        size_t m = n - 1;
        float x0r = xr[0], x0i = xi[0], dr = 0.0f, di = 0.0f;
        for (size_t i = 0; i < n; i++) { dr += xr[i]; di += xi[i]; }
        for (size_t j = 0; j < m; j++) { p->ar[j] = xr[p->rin[j]]; p->ai[j] = xi[p->rin[j]]; }
        plan_exec(p->inner, p->ar, p->ai, p->ar, p->ai);
        for (size_t j = 0; j < m; j++) {
            float a = p->ar[j], b = p->ai[j];
            p->ar[j] = a * p->bre[j] - b * p->bim[j];
            p->ai[j] = a * p->bim[j] + b * p->bre[j];
        }
        plan_exec(p->inner_inv, p->ar, p->ai, p->ar, p->ai);
        float sc = 1.0f / (float)m;
        yr[0] = dr; yi[0] = di;
        for (size_t j = 0; j < m; j++) {
            yr[p->rout[j]] = x0r + p->ar[j] * sc;
            yi[p->rout[j]] = x0i + p->ai[j] * sc;
        }
        return;
    }
    case K_BLUESTEIN: {           /* reference src/faf_bluestein.c:165-223 */
        size_t M = p->inner->n;
        for (size_t k = 0; k < n; k++) {
            p->ar[k] = xr[k] * p->phr[k] - xi[k] * p->phi[k];
            p->ai[k] = xr[k] * p->phi[k] + xi[k] * p->phr[k];
        }
        for (size_t k = n; k < M; k++) p->ar[k] = p->ai[k] = 0.0f;
        plan_exec(p->inner, p->ar, p->ai, p->cr, p->ci);
        for (size_t k = 0; k < M; k++) {
            float a = p->cr[k], b = p->ci[k];
            p->cr[k] = a * p->bre[k] - b * p->bim[k];
            p->ci[k] = a * p->bim[k] + b * p->bre[k];
        }
        plan_exec(p->inner_inv, p->cr, p->ci, p->ar, p->ai);
        float sc = (float)(1.0 / (double)M);   /* inner inverse NumPy norm */
        for (size_t k = 0; k < n; k++) {
            float a = p->ar[k] * sc, b = p->ai[k] * sc;
            yr[k] = a * p->phr[k] - b * p->phi[k];
            yi[k] = a * p->phi[k] + b * p->phr[k];
        }
        return;
    }
    }
}

/* norm factor: reference src/faf_core.c:1251-1279 (double, cast to float) */
static float norm_scale(size_t n, int norm, int inverse) {
    if (norm == 2) return (float)(1.0 / sqrt((double)n));
    if (norm == 3) return inverse ? 1.0f : (float)(1.0 / (double)n);
    return inverse ? (float)(1.0 / (double)n) : 1.0f;   /* NONE / default */
}

static void scale_f32(float *a, size_t n, float s) {
    if (s != 1.0f)
        for (size_t i = 0; i < n; i++) a[i] *= s;
}

int orc_fft_f32(size_t n, int inverse, int norm, int blue, const float *xr,
                const float *xi, float *yr, float *yi) {
    if (n == 0 || !orc_fft_size_legal(n, blue)) return -1;
    oplan *p = plan_create(n, inverse, 0, blue);
    if (!p) return -1;
    plan_exec(p, xr, xi, yr, yi);
    float s = norm_scale(n, norm, inverse);
    scale_f32(yr, n, s);
    scale_f32(yi, n, s);
    plan_free(p);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Kiss R2C / C2R: reference src/faf_rfft.c:37-172, 337-395            */
/* ------------------------------------------------------------------ */

typedef struct {
    size_t n;
    oplan *inner;
    float *stw;                 /* super-twiddles, nc/2 complex */
    float *pr, *pi, *fr, *fi;   /* scratch nc each */
} rplan;

static rplan *rplan_create(size_t n, int inverse) {
    if (n < 2 || (n & 1) || !orc_is_5_smooth(n)) return NULL;
    rplan *r = calloc(1, sizeof(rplan));
    size_t nc = n / 2;
    r->n = n;
    r->inner = plan_create(nc, inverse, 0, 0);
    r->stw = malloc((nc / 2 + 1) * 2 * sizeof(float));
    for (size_t i = 0; i < nc / 2; i++) {     /* :37-45 */
        double ph = -M_PI * ((double)(i + 1) / (double)nc + 0.5);
        if (inverse) ph = -ph;
        r->stw[2 * i] = (float)cos(ph);
        r->stw[2 * i + 1] = (float)sin(ph);
    }
    r->pr = malloc(nc * sizeof(float)); r->pi = malloc(nc * sizeof(float));
    r->fr = malloc(nc * sizeof(float)); r->fi = malloc(nc * sizeof(float));
    return r;
}

static void rplan_free(rplan *r) {
    if (!r) return;
    plan_free(r->inner);
    free(r->stw); free(r->pr); free(r->pi); free(r->fr); free(r->fi);
    free(r);
}

static void rplan_fwd(const rplan *r, const float *x, float *xr, float *xi) {
    size_t nc = r->n / 2;
    for (size_t k = 0; k < nc; k++) { r->pr[k] = x[2 * k]; r->pi[k] = x[2 * k + 1]; }
    plan_exec(r->inner, r->pr, r->pi, r->fr, r->fi);
    const float *fr = r->fr, *fi = r->fi;
    xr[0] = fr[0] + fi[0]; xi[0] = 0.0f;              /* :61-89 */
    xr[nc] = fr[0] - fi[0]; xi[nc] = 0.0f;
    for (size_t k = 1; k <= nc / 2; k++) {
        float ar = fr[k], ai = fi[k], br = fr[nc - k], bi = -fi[nc - k];
        float sr = ar + br, si = ai + bi, dr = ar - br, di = ai - bi;
        float tr = r->stw[2 * (k - 1)], ti = r->stw[2 * (k - 1) + 1];
        float wr = dr * tr - di * ti, wi = dr * ti + di * tr;
        xr[k] = 0.5f * (sr + wr); xi[k] = 0.5f * (si + wi);
        xr[nc - k] = 0.5f * (sr - wr); xi[nc - k] = 0.5f * (wi - si);
    }
}

static void rplan_inv(const rplan *r, const float *xr, const float *xi, float *y) {
    size_t nc = r->n / 2;
    float *fr = r->fr, *fi = r->fi;
    fr[0] = xr[0] + xr[nc]; fi[0] = xr[0] - xr[nc];     /* :127-149 */
    for (size_t k = 1; k <= nc / 2; k++) {
        float ar = xr[k], ai = xi[k], br = xr[nc - k], bi = -xi[nc - k];
        float er = ar + br, ei = ai + bi, dr = ar - br, di = ai - bi;
        float tr = r->stw[2 * (k - 1)], ti = r->stw[2 * (k - 1) + 1];
        float orr = dr * tr - di * ti, oi = dr * ti + di * tr;
        fr[k] = er + orr; fi[k] = ei + oi;
        fr[nc - k] = er - orr; fi[nc - k] = -(ei - oi);
    }
    plan_exec(r->inner, fr, fi, r->pr, r->pi);
    for (size_t k = 0; k < nc; k++) { y[2 * k] = r->pr[k]; y[2 * k + 1] = r->pi[k]; }
}

int orc_rfft_f32(size_t n, int norm, const float *x, float *re, float *im) {
    rplan *r = rplan_create(n, 0);
    if (!r) return -1;
    rplan_fwd(r, x, re, im);
    float s = norm_scale(n, norm, 0);
    scale_f32(re, n / 2 + 1, s);
    scale_f32(im, n / 2 + 1, s);
    rplan_free(r);
    return 0;
}

int orc_irfft_f32(size_t n, int norm, const float *re, const float *im, float *y) {
    rplan *r = rplan_create(n, 1);
    if (!r) return -1;
    rplan_inv(r, re, im, y);
    scale_f32(y, n, norm_scale(n, norm, 1));
    rplan_free(r);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Hermitian multiply and pipeline ops                                 */
/* ------------------------------------------------------------------ */

void orc_herm_mul_f32(float *yr, float *yi, const float *xr, const float *xi,
                      const float *hr, const float *hi, size_t nb, int conj_h) {
    /* reference src/faf_core.c:150-166 (conj_h=0) and :186-202 (conj_h=1) */
    if (nb == 0) return;
    float sg = conj_h ? -1.0f : 1.0f;
    yr[0] = xr[0] * hr[0]; yi[0] = 0.0f;
    for (size_t k = 1; k + 1 < nb; k++) {
        float ar = xr[k], ai = xi[k], br = hr[k], bi = sg * hi[k];
        yr[k] = ar * br - ai * bi;
        yi[k] = ar * bi + ai * br;
    }
    if (nb > 1) { yr[nb - 1] = xr[nb - 1] * hr[nb - 1]; yi[nb - 1] = 0.0f; }
}

int orc_pipeline_f32(size_t n, int nops, const int *ops, const uint32_t *a1,
                     const uint32_t *a2, const float *hr, const float *hi,
                     size_t h_bins, const float *x, float *y) {
    /* reference src/faf_pipeline.c:53-90 with src/chirp.c:421-471 */
    size_t nb = n / 2 + 1;
    float *sr = malloc(nb * sizeof(float)), *si = malloc(nb * sizeof(float));
    if (orc_rfft_f32(n, 1, x, sr, si) != 0) { free(sr); free(si); return -1; }
    for (int o = 0; o < nops; o++) {
        if (ops[o] == 1) {
            for (size_t k = 0; k < nb; k++) si[k] = -si[k];
        } else if (ops[o] == 2) {
            if (hr && hi) {
                size_t m = h_bins < nb ? h_bins : nb;
                orc_herm_mul_f32(sr, si, sr, si, hr, hi, m, 0);
            }
        } else if (ops[o] == 3) {
            for (size_t k = 0; k < nb; k++)
                if (k < a1[o] || k > a2[o]) { sr[k] = 0.0f; si[k] = 0.0f; }
        }
    }
    int rc = orc_irfft_f32(n, 1, sr, si, y);
    free(sr); free(si);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Hann window and framed STFT                                         */
/* ------------------------------------------------------------------ */

void orc_hann_f32(float *w, size_t n) {       /* faf_twiddles.c:159-163 */
    for (size_t i = 0; i < n; i++)
        w[i] = 0.5f * (1.0f - cosf(2.0f * (float)M_PI * (float)i / (float)(n - 1)));
}

size_t orc_stft_frames(size_t len, size_t win, size_t hop) {
    if (len <= win) return 1;               /* a short record is one zero-padded frame */
    return (len - win) / hop + 1;
}

int orc_stft_f32(size_t len, size_t win, size_t hop, const float *x, float *re,
                 float *im) {
    size_t nf = orc_stft_frames(len, win, hop), nb = win / 2 + 1;
    float *w = malloc(win * sizeof(float)), *fr = malloc(win * sizeof(float));
    rplan *r = rplan_create(win, 0);
    if (!r) { free(w); free(fr); return -1; }
    orc_hann_f32(w, win);
    for (size_t f = 0; f < nf; f++) {
        for (size_t i = 0; i < win; i++) {
            size_t s = f * hop + i;
            fr[i] = (s < len ? x[s] : 0.0f) * w[i];
        }
        rplan_fwd(r, fr, re + f * nb, im + f * nb);
    }
    rplan_free(r); free(w); free(fr);
    return 0;
}

int orc_stft_f64(size_t len, size_t win, size_t hop, const float *x, double *re,
                 double *im) {
    size_t nf = orc_stft_frames(len, win, hop), nb = win / 2 + 1;
    float *w = malloc(win * sizeof(float));
    double *fr = malloc(win * sizeof(double)), *yr = malloc(win * sizeof(double));
    double *yi = malloc(win * sizeof(double));
    orc_hann_f32(w, win);
    for (size_t f = 0; f < nf; f++) {
        for (size_t i = 0; i < win; i++) {
            size_t s = f * hop + i;
            fr[i] = (s < len ? (double)x[s] : 0.0) * (double)w[i];
        }
        orc_fft_f64(win, 0, fr, NULL, yr, yi);
        memcpy(re + f * nb, yr, nb * sizeof(double));
        memcpy(im + f * nb, yi, nb * sizeof(double));
    }
    free(w); free(fr); free(yr); free(yi);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CWT execution given a filter table                                  */
/* ------------------------------------------------------------------ */

int orc_cwt_real_f32(size_t n, size_t rows, const float *psi, size_t stride,
                     const float *x, float *w) {
    /* reference src/faf_cwt.c:769-843, src/faf_cwt_kernels.c:24-35 */
    size_t nb = n / 2 + 1;
    rplan *f = rplan_create(n, 0), *b = rplan_create(n, 1);
    if (!f || !b) { rplan_free(f); rplan_free(b); return -1; }
    float *xr = malloc(nb * sizeof(float)), *xi = malloc(nb * sizeof(float));
    float *yr = malloc(nb * sizeof(float)), *yi = malloc(nb * sizeof(float));
    rplan_fwd(f, x, xr, xi);
    float s = (float)(1.0 / (double)n);
    for (size_t r = 0; r < rows; r++) {
        const float *p = psi + r * stride;
        for (size_t k = 0; k < nb; k++) { yr[k] = xr[k] * p[k]; yi[k] = xi[k] * p[k]; }
        yi[0] = 0.0f;
        yi[nb - 1] = 0.0f;
        rplan_inv(b, yr, yi, w + r * n);
        scale_f32(w + r * n, n, s);
    }
    free(xr); free(xi); free(yr); free(yi);
    rplan_free(f); rplan_free(b);
    return 0;
}

int orc_icwt_l1_f32(size_t n, size_t n_scales, int has_lp, unsigned voices, double c,
                    const float *w, float *y) {
    /* reference src/faf_cwt.c:878-915 (the FP32 branch) */
    if (!voices) return -1;
    if (c <= 0.0) c = 1.0;
    const float dlog = (float)(0.693147180559945309417 / (double)voices), cf = (float)c;
    const float *rows = w + (has_lp ? n : 0);
    for (size_t i = 0; i < n; i++) y[i] = 0.0f;
    for (size_t j = 0; j < n_scales; j++)
        for (size_t i = 0; i < n; i++) y[i] += rows[j * n + i] * dlog;
    for (size_t i = 0; i < n; i++) y[i] /= cf;
    if (has_lp)
        for (size_t i = 0; i < n; i++) y[i] += w[i];
    return 0;
}

int orc_cwt_complex_f64(size_t n, size_t rows, const float *psi, size_t stride,
                        const float *x, double *zr, double *zi) {
    size_t nb = n / 2 + 1;
    double *xd = malloc(n * sizeof(double)), *Xr = malloc(n * sizeof(double));
    double *Xi = malloc(n * sizeof(double)), *Zr = malloc(n * sizeof(double));
    double *Zi = malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) xd[i] = x[i];
    orc_fft_f64(n, 0, xd, NULL, Xr, Xi);
    for (size_t r = 0; r < rows; r++) {
        const float *p = psi + r * stride;
        /* one-sided spectrum: DC and Nyquist real (imag zeroed as in the
         * reference multiply), interior bins doubled, negative bins zero */
        memset(Zr, 0, n * sizeof(double)); memset(Zi, 0, n * sizeof(double));
        Zr[0] = Xr[0] * p[0];
        for (size_t k = 1; k + 1 < nb; k++) {
            Zr[k] = 2.0 * Xr[k] * p[k];
            Zi[k] = 2.0 * Xi[k] * p[k];
        }
        Zr[nb - 1] = Xr[nb - 1] * p[nb - 1];
        orc_fft_f64(n, 1, Zr, Zi, zr + r * n, zi + r * n);
        for (size_t i = 0; i < n; i++) { zr[r * n + i] /= (double)n; zi[r * n + i] /= (double)n; }
    }
    free(xd); free(Xr); free(Xi); free(Zr); free(Zi);
    return 0;
}

/* ------------------------------------------------------------------ */
/* CPU baseline timing: one plan per worker thread, like the reference  */
/* executing one signal per call                                        */
/* ------------------------------------------------------------------ */

typedef struct {
    size_t n, lo, hi;
    const float *x;
    float *re, *im;
} job_t;

static void *rfft_worker(void *arg) {
    job_t *j = arg;
    size_t nb = j->n / 2 + 1;
    rplan *r = rplan_create(j->n, 0);
    for (size_t s = j->lo; s < j->hi; s++)
        rplan_fwd(r, j->x + s * j->n, j->re + s * nb, j->im + s * nb);
    rplan_free(r);
    return NULL;
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

double orc_time_rfft_batch(size_t n, size_t count, int threads, const float *x,
                           float *re, float *im) {
    if (threads < 1) threads = 1;
    pthread_t *th = malloc((size_t)threads * sizeof(pthread_t));
    job_t *jb = malloc((size_t)threads * sizeof(job_t));
    double t0 = now_s();
    for (int t = 0; t < threads; t++) {
        jb[t].n = n; jb[t].x = x; jb[t].re = re; jb[t].im = im;
        jb[t].lo = count * (size_t)t / (size_t)threads;
        jb[t].hi = count * (size_t)(t + 1) / (size_t)threads;
        pthread_create(&th[t], NULL, rfft_worker, &jb[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    double dt = now_s() - t0;
    free(th); free(jb);
    return dt;
}

/* The same measurement through the reference library itself, compiled from
 * its own sources into oracle/_ref (see the Makefile): one plan per thread,
 * created before the clock starts, and one faf_execute per signal, as a
 * reference program calls it (src/faf_rfft.c:268-335).  backend is the
 * reference's faf_backend (0 = AUTO: JIT above its threshold, else VM). */
typedef faf_transform *(*ref_create_fn)(const faf_config *);
typedef int (*ref_exec_fn)(const faf_transform *, faf_buffer *, const faf_buffer *);
typedef void (*ref_destroy_fn)(faf_transform *);

typedef struct {
    size_t n, lo, hi;
    const float *x;
    float *re, *im;
    int backend, failed;
    ref_create_fn create;
    ref_exec_fn exec;
    ref_destroy_fn destroy;
    pthread_barrier_t *ready;
} rjob_t;

static void *ref_worker(void *arg) {
    rjob_t *j = arg;
    const size_t nb = j->n / 2 + 1;
    faf_config cfg = faf_config_init(j->n);
    cfg.backend = (faf_backend)j->backend;
    faf_transform *t = j->create(&cfg);
    pthread_barrier_wait(j->ready);
    if (!t) { j->failed = 1; return NULL; }
    for (size_t s = j->lo; s < j->hi; s++) {
        faf_buffer in = faf_buffer_real((float *)(j->x + s * j->n), j->n);
        faf_buffer out = faf_buffer_hermitian(j->re + s * nb, j->im + s * nb, nb);
        if (j->exec(t, &out, &in)) { j->failed = 1; break; }
    }
    j->destroy(t);
    return NULL;
}

double orc_time_ref_rfft_batch(const char *so, size_t n, size_t count, int threads,
                               int backend, const float *x, float *re, float *im) {
    void *h = dlopen(so, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1.0;
    ref_create_fn create = (ref_create_fn)dlsym(h, "faf_create_rfft");
    ref_exec_fn exec = (ref_exec_fn)dlsym(h, "faf_execute");
    ref_destroy_fn destroy = (ref_destroy_fn)dlsym(h, "faf_destroy_transform");
    if (!create || !exec || !destroy) return -1.0;
    if (threads < 1) threads = 1;
    pthread_barrier_t ready;
    pthread_barrier_init(&ready, NULL, (unsigned)threads + 1);
    pthread_t *th = malloc((size_t)threads * sizeof(pthread_t));
    rjob_t *jb = calloc((size_t)threads, sizeof(rjob_t));
    for (int t = 0; t < threads; t++) {
        jb[t].n = n; jb[t].x = x; jb[t].re = re; jb[t].im = im;
        jb[t].lo = count * (size_t)t / (size_t)threads;
        jb[t].hi = count * (size_t)(t + 1) / (size_t)threads;
        jb[t].backend = backend;
        jb[t].create = create; jb[t].exec = exec; jb[t].destroy = destroy;
        jb[t].ready = &ready;
        pthread_create(&th[t], NULL, ref_worker, &jb[t]);
    }
    pthread_barrier_wait(&ready);              /* every plan exists: start the clock */
    double t0 = now_s();
    int failed = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(th[t], NULL);
        failed |= jb[t].failed;
    }
    double dt = now_s() - t0;
    pthread_barrier_destroy(&ready);
    free(th); free(jb);
    return failed ? -1.0 : dt;
}
