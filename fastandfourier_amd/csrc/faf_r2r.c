/*
 * faf_r2r.c -- DCT I-IV, DST I-IV and MDCT / IMDCT on the device (C11 host
 * side), each riding on one complex FFT.
 *
 * The reference declares these transforms (include/fastandfourier.h,
 * README "DCT I-IV, DST I-IV, MDCT") but its generators emit a plain FFT in
 * their place (src/faf_transforms.c:195-222), so there is no reference
 * output to match.  What is computed here are the transforms those names
 * mean, with scipy.fft's definitions and norms: NONE is scipy's "backward"
 * (forward unscaled), ORTHO its "ortho" (orthogonalised: the first / last
 * terms of types I-III weighted as there), FORWARD its "forward".  The type
 * names the matrix computed and dir the side of the normalisation, so that
 * faf_create_inverse -- which, as in the reference (src/faf_core.c:1104-
 * 1118), swaps types II and III and flips dir -- yields the true inverse:
 * type T with dir INVERSE is scipy's idct / idst of the partner type (II <->
 * III; I and IV are their own partners), e.g. DCT-III inverse = dct3/(2N) =
 * idct(type=2).  The MDCT of
 * 2M real samples is X[k] = sum_n x[n] cos(pi/M (n + 1/2 + M/2)(k + 1/2)),
 * k < M, and the IMDCT y[n] = (2/M) sum_k X[k] cos(...) (n < 2M), so that
 * windowed overlap-add with a Princen-Bradley window reconstructs the signal
 * (ORTHO: both scaled by sqrt(2/M); FORWARD: the 2/M on the MDCT).
 *
 * Every case is one of a single pattern, three launches per chunk:
 *   u[j] = pre[j] * x[j - shift]     (zero where j - shift is outside x)
 *   U    = FFT_L(u)                  (forward; IMDCT: unnormalised inverse)
 *   y[k] = Re or Im of post[k] * U[k + off]
 * with L = 2N (DCT/DST II-IV), 2(N-1) (DCT-I), 2(N+1) (DST-I) or 2M (MDCT),
 * and pre / post tables built here in FP64 from exact integer phases,
 * carrying the norm weights.  Any L the FFT planner takes works (5-smooth
 * directly; otherwise Rader or Bluestein), so sizes are unrestricted.
 */
#include "faf_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* a complex table as float2 (FP32 plans) or double2 (FP64 plans) */
static void *upload_c(const double *re, const double *im, size_t n, int fp64) {
    const size_t es = fp64 ? sizeof(double) : sizeof(float);
    void *h = malloc(2 * n * es);
    if (!h) { fafi_err("r2r: alloc failed"); return NULL; }
    for (size_t i = 0; i < n; ++i) {
        if (fp64) { ((double *)h)[2 * i] = re[i]; ((double *)h)[2 * i + 1] = im[i]; }
        else { ((float *)h)[2 * i] = (float)re[i]; ((float *)h)[2 * i + 1] = (float)im[i]; }
    }
    void *d = fafk_malloc(2 * n * es);
    if (!d || fafk_h2d(d, h, 2 * n * es, NULL) || fafk_sync(NULL)) {
        fafi_err("r2r: %s", fafk_last_error());
        fafk_free(d);
        d = NULL;
    }
    free(h);
    return d;
}

/* e^{-i pi m / q} (sign -1) or e^{+i pi m / q} as exp(-+2 pi i m / (2q)), exact phase */
static void phase(long m, size_t q, int plus, double *c, double *s) {
    const long per = (long)(2 * q);
    long r = m % per;
    if (r < 0) r += per;
    fafi_unit_root((size_t)r, (size_t)per, plus, c, s);
}

int fafi_plan_r2r(fafh_plan *p, int kind, int type, size_t n, int inverse, int norm, int fp64) {
    const int ortho = norm == FAF_NORM_ORTHO, fwdn = norm == FAF_NORM_FORWARD;
    size_t N = n, L, nin, nout;
    int shift = 0, off = 0, im = 0, inner_inv = 0;
    /* the matrix to compute (type T), and its overall scale */
    const int T = type;
    double sc = 1.0;
    if (kind == 2) {
        if (n < 2 || (n & 1)) { fafi_err("MDCT size must be even, got %zu", n); return -1; }
    } else if (type == 1 && kind == 0 && n < 2) {
        fafi_err("DCT-I needs n >= 2, got %zu", n);
        return -1;
    } else if (n < 1) {
        fafi_err("r2r size must be > 0");
        return -1;
    }
    if (kind != 2) {
        /* NONE scales the inverse side, FORWARD the forward side, by 1/d */
        const double d = T == 1 ? (kind == 0 ? 2.0 * (double)(N - 1) : 2.0 * (double)(N + 1))
                                : 2.0 * (double)N;
        if ((inverse && !ortho && !fwdn) || (!inverse && fwdn)) sc = 1.0 / d;
    }
    double *pr, *pi, *qr, *qi, *al, *be;
    if (kind == 2) {
        const size_t M = n / 2;
        L = n;
        nin = inverse ? M : n;
        nout = inverse ? n : M;
        inner_inv = inverse;
    } else {
        nin = nout = N;
        L = (T == 1) ? (kind == 0 ? 2 * (N - 1) : 2 * (N + 1)) : 2 * N;
    }
    pr = calloc(L, sizeof(double)); pi = calloc(L, sizeof(double));
    qr = calloc(nout, sizeof(double)); qi = calloc(nout, sizeof(double));
    al = malloc(nin * sizeof(double)); be = malloc(nout * sizeof(double));
    if (!pr || !pi || !qr || !qi || !al || !be) {
        free(pr); free(pi); free(qr); free(qi); free(al); free(be);
        fafi_err("r2r: alloc failed");
        return -1;
    }
    for (size_t i = 0; i < nin; ++i) al[i] = 1.0;
    for (size_t i = 0; i < nout; ++i) be[i] = sc;
    double c, s;
    if (kind == 2) {
        const size_t M = n / 2;
        const double w = ortho ? sqrt(2.0 / (double)M) : 1.0;
        if (!inverse) {
            /* u[j] = x[j] e^{-i pi j / (2M)}; X[k] = Re(e^{-i pi (M+1)(2k+1)/(4M)} U[k]) */
            const double f = fwdn ? 2.0 / (double)M : w;
            for (size_t j = 0; j < L; ++j) { phase((long)j, 2 * M, 0, &c, &s); pr[j] = c; pi[j] = s; }
            for (size_t k = 0; k < nout; ++k) {
                phase((long)((M + 1) * (2 * k + 1)), 4 * M, 0, &c, &s);
                qr[k] = f * c; qi[k] = f * s;
            }
        } else {
            /* u[j] = X[j] e^{+i pi (M+1) j / (2M)}, j < M; U = unnormalised
             * inverse FFT_2M; y[n] = scale Re(e^{+i pi (2n+M+1)/(4M)} U[n]) */
            const double f = fwdn ? 1.0 : (ortho ? w : 2.0 / (double)M);
            for (size_t j = 0; j < M; ++j) {
                phase((long)((M + 1) * j), 2 * M, 1, &c, &s);
                pr[j] = c; pi[j] = s;
            }
            for (size_t k = 0; k < nout; ++k) {
                phase((long)(2 * k + M + 1), 4 * M, 1, &c, &s);
                qr[k] = f * c; qi[k] = f * s;
            }
        }
    } else {
        /* orthogonalising weights (scipy, orthogonalize=True): input side
         * for the types whose sums carry a single-weight end term, output
         * side for the others; the 1/sqrt(...) overall factor on the output */
        if (ortho) {
            double g;
            if (T == 1) g = 1.0 / sqrt(kind == 0 ? 2.0 * (double)(N - 1) : 2.0 * (double)(N + 1));
            else g = 1.0 / sqrt(2.0 * (double)N);
            for (size_t k = 0; k < nout; ++k) be[k] = g;
            if (kind == 0 && T == 1) {
                al[0] *= sqrt(2.0); al[N - 1] *= sqrt(2.0);
                be[0] /= sqrt(2.0); be[N - 1] /= sqrt(2.0);
            } else if (kind == 0 && T == 2) {
                be[0] /= sqrt(2.0);
            } else if (kind == 0 && T == 3) {
                al[0] *= sqrt(2.0);
            } else if (kind == 1 && T == 2) {
                be[N - 1] /= sqrt(2.0);
            } else if (kind == 1 && T == 3) {
                al[N - 1] *= sqrt(2.0);
            }
        }
        if (kind == 0) {
            switch (T) {
            case 1:    /* u = e_n x (e = 1 at the ends, else 2), y = Re U */
                for (size_t j = 0; j < N; ++j)
                    pr[j] = al[j] * ((j == 0 || j == N - 1) ? 1.0 : 2.0);
                for (size_t k = 0; k < nout; ++k) qr[k] = be[k];
                break;
            case 2:    /* u = x, y = Re(2 e^{-i pi k/(2N)} U[k]) */
                for (size_t j = 0; j < N; ++j) pr[j] = al[j];
                for (size_t k = 0; k < nout; ++k) {
                    phase((long)k, 2 * N, 0, &c, &s);
                    qr[k] = 2.0 * be[k] * c; qi[k] = 2.0 * be[k] * s;
                }
                break;
            case 3:    /* u = c_n x e^{-i pi n/(2N)} (c_0 = 1, else 2), y = Re U */
                for (size_t j = 0; j < N; ++j) {
                    phase((long)j, 2 * N, 0, &c, &s);
                    const double f = al[j] * (j == 0 ? 1.0 : 2.0);
                    pr[j] = f * c; pi[j] = f * s;
                }
                for (size_t k = 0; k < nout; ++k) qr[k] = be[k];
                break;
            default:   /* 4: u = x e^{-i pi n/(2N)}, y = Re(2 e^{-i pi (2k+1)/(4N)} U[k]) */
                for (size_t j = 0; j < N; ++j) {
                    phase((long)j, 2 * N, 0, &c, &s);
                    pr[j] = al[j] * c; pi[j] = al[j] * s;
                }
                for (size_t k = 0; k < nout; ++k) {
                    phase((long)(2 * k + 1), 4 * N, 0, &c, &s);
                    qr[k] = 2.0 * be[k] * c; qi[k] = 2.0 * be[k] * s;
                }
            }
        } else {
            im = 1;
            switch (T) {
            case 1:    /* u[j] = x[j-1] (j = 1..N), y = Im(-2 U[k+1]) */
                shift = 1;
                off = 1;
                for (size_t j = 1; j <= N; ++j) pr[j] = al[j - 1];
                for (size_t k = 0; k < nout; ++k) qr[k] = -2.0 * be[k];
                break;
            case 2:    /* u = x, y = Im(-2 e^{-i pi (k+1)/(2N)} U[k+1]) */
                off = 1;
                for (size_t j = 0; j < N; ++j) pr[j] = al[j];
                for (size_t k = 0; k < nout; ++k) {
                    phase((long)(k + 1), 2 * N, 0, &c, &s);
                    qr[k] = -2.0 * be[k] * c; qi[k] = -2.0 * be[k] * s;
                }
                break;
            case 3:    /* u[j] = d x[j-1] e^{-i pi j/(2N)} (d = 1 at j = N, else 2), y = Im(-U) */
                shift = 1;
                for (size_t j = 1; j <= N; ++j) {
                    phase((long)j, 2 * N, 0, &c, &s);
                    const double f = al[j - 1] * (j == N ? 1.0 : 2.0);
                    pr[j] = f * c; pi[j] = f * s;
                }
                for (size_t k = 0; k < nout; ++k) qr[k] = -be[k];
                break;
            default:   /* 4: u = x e^{-i pi n/(2N)}, y = Im(-2 e^{-i pi (2k+1)/(4N)} U[k]) */
                for (size_t j = 0; j < N; ++j) {
                    phase((long)j, 2 * N, 0, &c, &s);
                    pr[j] = al[j] * c; pi[j] = al[j] * s;
                }
                for (size_t k = 0; k < nout; ++k) {
                    phase((long)(2 * k + 1), 4 * N, 0, &c, &s);
                    qr[k] = -2.0 * be[k] * c; qi[k] = -2.0 * be[k] * s;
                }
            }
        }
    }
    p->r2r_pre = upload_c(pr, pi, L, fp64);
    p->r2r_post = upload_c(qr, qi, nout, fp64);
    free(pr); free(pi); free(qr); free(qi); free(al); free(be);
    if (!p->r2r_pre || !p->r2r_post) return -1;
    /* the inner C2C: any length (Rader / Bluestein where not 5-smooth) */
    fafh_plan *q = fafi_plan_new();
    if (!q) { fafi_err("r2r: alloc failed"); return -1; }
    if (L == 1 || (fp64 ? fafi_plan_c2c64(q, L, inner_inv, 1, 0)
                        : fafi_plan_c2c(q, L, inner_inv, 1))) {
        if (L == 1) fafi_err("r2r: inner length 1");
        fafi_plan_free(q);
        return -1;
    }
    q->scale = 1.0f;
    q->scale64 = 1.0;
    p->fp64 = fp64;
    p->inner_f = q;
    p->kind = PK_R2R;
    p->n = n;
    p->inverse = inverse;
    p->r2r_L = (long)L;
    p->r2r_nin = (long)nin;
    p->r2r_nout = (long)nout;
    p->r2r_shift = shift;
    p->r2r_off = off;
    p->r2r_im = im;
    return 0;
}

/* chunks of signals: two plane pairs of L per signal, then the inner
 * plan's own workspace */
static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t r2r_bytes(const fafh_plan *p, long cs) {
    return al256((size_t)cs * (size_t)p->r2r_L * 16) + fafi_run_ws_bytes(p->inner_f, cs);
}

int fafi_run_r2r(fafh_plan *p, const float *x, long idist, float *y, long odist, long count,
                 void *stream) {
    const long L = p->r2r_L;
    long cs = (long)(((size_t)64 << 20) / ((size_t)L * 16));
    if (cs < 1) cs = 1;
    if (cs > count) cs = count;
    if (p->ws_ext)
        while (cs > 1 && r2r_bytes(p, cs) > p->ws_ext_bytes) cs = (cs + 1) / 2;
    const size_t need = r2r_bytes(p, cs);
    char *ws = fafi_workspace(p, need);
    if (!ws) return -1;
    float *ur = (float *)ws, *ui = ur + (size_t)cs * L;
    float *vr = ui + (size_t)cs * L, *vi = vr + (size_t)cs * L;
    char *sub = ws + al256((size_t)cs * (size_t)L * 16);
    const size_t subb = need - al256((size_t)cs * (size_t)L * 16);
    fafh_plan *q = p->inner_f;
    for (long s0 = 0; s0 < count; s0 += cs) {
        const long c = count - s0 < cs ? count - s0 : cs;
        if (fafk_r2r_pre(x + s0 * idist, idist, p->r2r_nin, p->r2r_pre, p->r2r_shift, L, ur, ui,
                         L, c, stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
        q->ws_ext = sub;
        q->ws_ext_bytes = subb;
        int rc = fafi_run_c2c(q, ur, ui, L, vr, vi, L, c, stream);
        q->ws_ext = NULL;
        q->ws_ext_bytes = 0;
        if (rc) return -1;
        if (fafk_r2r_post(vr, vi, L, p->r2r_post, p->r2r_off, p->r2r_im, p->r2r_nout,
                          y + s0 * odist, odist, c, stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
    }
    return 0;
}
