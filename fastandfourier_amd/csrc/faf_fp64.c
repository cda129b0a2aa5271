/*
 * faf_fp64.c -- the FP64 execution variant (C11 host side).
 *
 * The reference plans FFT, RFFT and the DCT family in FP64 as well as FP32
 * (precision in faf_config; its FP64 rfft round trip is held to 1e-12,
 * tests/test_rfft.cpp:239-261).  Here an FP64 config gets a device plan
 * that computes in double throughout: the generic double-precision
 * Stockham kernel of faf_kernels64.hip over runtime radix plans (lengths
 * up to 4096 in one workgroup, four-step beyond), FP64 twiddle tables from
 * exact integer phases, the Kiss post/pre passes for R2C/C2R (the real
 * signal is read or written directly as packed complex pairs), Rader and
 * Bluestein as gather / chirp, inner FFTs, kernel product and scatter /
 * chirp, and DCT/DST/MDCT around one complex FFT as in FP32.  Buffers hold
 * doubles; layouts, norms and the buffer contract are the FP32 ones.
 */
#include "faf_internal.h"
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CHUNK64 ((size_t)64 << 20)
#define MAX1 4096                 /* longest transform of one workgroup in FP64 */

static int k64(int rc) {
    if (rc) fafi_err("%s", fafk64_last_error());
    return rc;
}

/* device double2 table of exp(-+2 pi i m / n), m < n, cached per (n, inv)
 * for the life of the process (one upload per table) */
typedef struct tab64 { size_t n; int inv; int kind; void *dev; struct tab64 *next; } tab64;
static tab64 *g_t64;
static pthread_mutex_t g_t64_mu = PTHREAD_MUTEX_INITIALIZER;

static void *upload64(const double *h, size_t doubles) {
    void *d = fafk_malloc(doubles * sizeof(double));
    if (!d) { fafi_err("%s", fafk_last_error()); return NULL; }
    if (fafk_h2d(d, h, doubles * sizeof(double), NULL) || fafk_sync(NULL)) {
        fafi_err("%s", fafk_last_error());
        fafk_free(d);
        return NULL;
    }
    return d;
}

/* kind 0: W_n^m (m < n); 1: Kiss post -i W_n^k (k <= n/2); 2: Kiss pre i W_n^-k (k < n/2) */
static const void *table64(int kind, size_t n, int inv) {
    pthread_mutex_lock(&g_t64_mu);
    for (tab64 *e = g_t64; e; e = e->next)
        if (e->n == n && e->inv == inv && e->kind == kind) {
            pthread_mutex_unlock(&g_t64_mu);
            return e->dev;
        }
    pthread_mutex_unlock(&g_t64_mu);
    size_t cnt = kind == 0 ? n : kind == 1 ? n / 2 + 1 : n / 2;
    double *h = malloc(2 * cnt * sizeof(double));
    if (!h) { fafi_err("fp64: alloc failed"); return NULL; }
    for (size_t m = 0; m < cnt; ++m) {
        double c, s;
        if (kind == 0) {
            fafi_unit_root(m, n, inv, &c, &s);
            h[2 * m] = c; h[2 * m + 1] = s;
        } else if (kind == 1) {
            fafi_unit_root(m, n, 0, &c, &s);     /* W = (c, s); -i W = (s, -c) */
            h[2 * m] = s; h[2 * m + 1] = -c;
        } else {
            fafi_unit_root(m, n, 1, &c, &s);     /* W^-1 = (c, s); i W^-1 = (-s, c) */
            h[2 * m] = -s; h[2 * m + 1] = c;
        }
    }
    void *d = upload64(h, 2 * cnt);
    free(h);
    if (!d) return NULL;
    pthread_mutex_lock(&g_t64_mu);
    tab64 *e = malloc(sizeof *e);
    e->n = n; e->inv = inv; e->kind = kind; e->dev = d; e->next = g_t64;
    g_t64 = e;
    pthread_mutex_unlock(&g_t64_mu);
    return d;
}

/* the largest n1 <= sqrt(n) with n1 and n / n1 both runnable in one FP64
 * workgroup */
static int split64(size_t n, int allow7, int *n1, int *n2) {
    size_t best = 0;
    for (size_t a = 2; a * a <= n; ++a) {
        if (n % a) continue;
        size_t b = n / a;
        if (a <= MAX1 && b <= MAX1 && fafi_radix_fits(a, allow7) && fafi_radix_fits(b, allow7))
            best = a;
    }
    if (!best) return -1;
    *n1 = (int)best;
    *n2 = (int)(n / best);
    return 0;
}

static int powmod(size_t b, size_t e, size_t m) {
    unsigned long long r = 1, x = b % m;
    while (e) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
        e >>= 1;
    }
    return (int)r;
}

static fafh_plan *inner64(size_t m, int inv, int allow7);

int fafi_plan_c2c64(fafh_plan *p, size_t n, int inv, int bluestein, int allow7) {
    p->n = n;
    p->inverse = inv;
    p->fp64 = 1;
    int legal = fafi_is_5_smooth(n) || fafi_is_codelet_size(n) ||
                (allow7 && fafi_is_7_smooth(n));
    if (!legal && fafi_rader_eligible(n)) {
        /* Rader: a[j] = x[g^j], B = FFT_m(w^(g^-j)) / m in FP64 */
        const size_t m = n - 1, g = fafi_primitive_root(n), gi = (size_t)powmod(g, n - 2, n);
        int32_t *rin = malloc(m * sizeof(int32_t)), *rout = malloc(m * sizeof(int32_t));
        double *br = malloc(m * sizeof(double)), *bi = malloc(m * sizeof(double));
        double *bd = malloc(2 * m * sizeof(double));
        size_t a = 1, b = 1;
        for (size_t j = 0; j < m; ++j) {
            rin[j] = (int32_t)a;
            rout[j] = (int32_t)b;
            fafi_unit_root(b, n, inv, &br[j], &bi[j]);
            a = a * g % n;
            b = b * gi % n;
        }
        int rc = fafi_fft64(m, 0, br, bi);
        for (size_t j = 0; j < m && !rc; ++j) {
            bd[2 * j] = br[j] / (double)m;
            bd[2 * j + 1] = bi[j] / (double)m;
        }
        if (!rc) {
            p->rader_mem[0] = fafk_malloc(m * sizeof(int32_t));
            p->rader_mem[1] = fafk_malloc(m * sizeof(int32_t));
            p->rader_mem[2] = upload64(bd, 2 * m);
            rc = !p->rader_mem[0] || !p->rader_mem[1] || !p->rader_mem[2] ||
                 fafk_h2d(p->rader_mem[0], rin, m * sizeof(int32_t), NULL) ||
                 fafk_h2d(p->rader_mem[1], rout, m * sizeof(int32_t), NULL) || fafk_sync(NULL);
            if (rc) fafi_err("fp64 Rader: %s", fafk_last_error());
        }
        free(rin); free(rout); free(br); free(bi); free(bd);
        if (rc) return -1;
        p->rader.p = (int32_t)n;
        p->rader.rin = p->rader_mem[0];
        p->rader.rout = p->rader_mem[1];
        p->rader.b = p->rader_mem[2];
        p->inner_f = inner64(m, 0, 1);
        p->inner_i = p->inner_f ? inner64(m, 1, 1) : NULL;
        if (!p->inner_i) return -1;
        p->kind = PK_RADER64;
        return 0;
    }
    if (!legal) {
        if (!bluestein) {
            fafi_err(fafi_is_7_smooth(n) ? "size %zu is not a public FFT length"
                                         : "size %zu needs FAF_FLAG_BLUESTEIN", n);
            return -1;
        }
        /* Bluestein: M = next 5-smooth >= 2n-1, chirp exp(-+i pi k^2 / n)
         * from k^2 mod 2n, kernel spectrum / M, all FP64 */
        const size_t M = fafi_next_5_smooth(2 * n - 1);
        double *ch = malloc(2 * n * sizeof(double));
        double *br = calloc(M, sizeof(double)), *bi = calloc(M, sizeof(double));
        double *bd = malloc(2 * M * sizeof(double));
        for (size_t k = 0; k < n; ++k) {
            unsigned long long q = ((unsigned long long)k * k) % (unsigned long long)(2 * n);
            double c, s;
            fafi_unit_root((size_t)q, 2 * n, inv, &c, &s);
            ch[2 * k] = c; ch[2 * k + 1] = s;
            br[k] = c; bi[k] = -s;
            if (k) { br[M - k] = c; bi[M - k] = -s; }
        }
        int rc = fafi_fft64(M, 0, br, bi);
        for (size_t j = 0; j < M && !rc; ++j) {
            bd[2 * j] = br[j] / (double)M;
            bd[2 * j + 1] = bi[j] / (double)M;
        }
        if (!rc) {
            p->blue_mem[0] = upload64(ch, 2 * n);
            p->blue_mem[1] = upload64(bd, 2 * M);
            rc = (p->blue_mem[0] && p->blue_mem[1]) ? 0 : -1;
        }
        free(ch); free(br); free(bi); free(bd);
        if (rc) return -1;
        p->blue.n = (int32_t)n;
        p->blue.m = (int32_t)M;
        p->blue.chirp = p->blue_mem[0];
        p->blue.b = p->blue_mem[1];
        p->inner_f = inner64(M, 0, 0);
        p->inner_i = p->inner_f ? inner64(M, 1, 0) : NULL;
        if (!p->inner_i) return -1;
        p->kind = PK_BLUE64;
        return 0;
    }
    if (n == 1) { fafi_err("length 1 is not a transform"); return -1; }
    if (n <= MAX1 && fafi_radix_fits(n, 1)) {
        if (fafi_make_fft(n, inv, 1, NULL, 0, &p->a)) return -1;
        p->tw64 = table64(0, n, inv);
        if (!p->tw64) return -1;
        p->kind = PK_C2C64;
        return 0;
    }
    int n1, n2;
    if (split64(n, allow7, &n1, &n2)) {
        fafi_err("fp64: length %zu has no four-step split", n);
        return -1;
    }
    if (fafi_make_fft((size_t)n1, inv, allow7, NULL, 0, &p->a) ||
        fafi_make_fft((size_t)n2, inv, allow7, NULL, 0, &p->b))
        return -1;
    p->n1 = n1;
    p->n2 = n2;
    p->tw64 = table64(0, (size_t)n1, inv);
    p->tw64b = table64(0, (size_t)n2, inv);
    p->twn64 = table64(0, n, inv);
    if (!p->tw64 || !p->tw64b || !p->twn64) return -1;
    p->kind = PK_C2C64_4S;
    return 0;
}

static fafh_plan *inner64(size_t m, int inv, int allow7) {
    fafh_plan *q = fafi_plan_new();
    if (!q) { fafi_err("alloc failed"); return NULL; }
    q->scale64 = 1.0;
    if (fafi_plan_c2c64(q, m, inv, 1, allow7)) { fafi_plan_free(q); return NULL; }
    return q;
}

int fafi_plan_real64(fafh_plan *p, size_t n, int inv) {
    const size_t nc = n / 2;
    p->n = n;
    p->inverse = inv;
    p->fp64 = 1;
    if (nc < 2) { fafi_err("fp64 rfft needs n >= 4"); return -1; }
    p->inner_f = inner64(nc, inv, 0);
    if (!p->inner_f) return -1;
    p->inner_f->interleaved = 1;        /* packed pairs / Z in workspace */
    p->tw64 = table64(inv ? 2 : 1, n, 0);
    if (!p->tw64) return -1;
    p->kind = inv ? PK_C2R64 : PK_R2C64;
    return 0;
}

/* ------------------------------------------------------------------ */

static size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

static fafk_acc64 acc(const void *re, const void *im, long G, long gdist, long tdist, long es) {
    fafk_acc64 a;
    a.re = re; a.im = im; a.G = G; a.gdist = gdist; a.tdist = tdist; a.estride = es;
    return a;
}

/* workspace bytes one run of chunk c signals takes, by plan kind */
static size_t need64(const fafh_plan *p, long c) {
    switch (p->kind) {
    case PK_C2C64_4S:
        return (size_t)c * p->n * 16;
    case PK_RADER64:
    case PK_BLUE64: {
        const size_t L = p->kind == PK_RADER64 ? p->n - 1 : (size_t)p->blue.m;
        size_t w1 = fafi_ws64_bytes(p->inner_f, c), w2 = fafi_ws64_bytes(p->inner_i, c);
        return a256((size_t)c * L * 32) + a256((size_t)c * 16) + (w1 > w2 ? w1 : w2);
    }
    case PK_R2C64:
    case PK_C2R64:
        return a256((size_t)c * (p->n / 2) * 16) + fafi_ws64_bytes(p->inner_f, c);
    case PK_R2R:
        return a256((size_t)c * (size_t)p->r2r_L * 32) + fafi_ws64_bytes(p->inner_f, c);
    default:
        return 0;
    }
}

/* signals per chunk: about 64 MB of workspace, and within a lent
 * workspace when there is one */
static long chunk64(const fafh_plan *p, long count) {
    const size_t one = need64(p, 1);
    long c = (long)(CHUNK64 / (one ? one : 1));
    if (c < 1) c = 1;
    if (c > count) c = count;
    if (p->ws_ext)
        while (c > 1 && need64(p, c) > p->ws_ext_bytes) c = (c + 1) / 2;
    return c;
}

size_t fafi_ws64_bytes(const fafh_plan *p, long count) {
    if (!p || count <= 0) return 0;
    return need64(p, chunk64(p, count));
}

/* runs q (FP64 C2C) on lent workspace */
static int run_lent(fafh_plan *q, void *ws, size_t wb, const double *ir, const double *ii,
                    long idist, double *orr, double *oi, long odist, long count, void *stream) {
    q->ws_ext = ws;
    q->ws_ext_bytes = wb;
    int rc = fafi_run_c2c64(q, ir, ii, idist, orr, oi, odist, count, stream);
    q->ws_ext = NULL;
    q->ws_ext_bytes = 0;
    return rc;
}

/* double2 loads and stores need 16-byte aligned complex elements */
static int aligned16(const void *a, const void *b) {
    return !(((uintptr_t)a | (uintptr_t)b) & 15);
}

int fafi_run_c2c64(fafh_plan *p, const double *ir, const double *ii, long idist, double *orr,
                   double *oi, long odist, long count, void *stream) {
    const int inter = p->interleaved;
    if (inter && !aligned16(ir, orr)) {
        fafi_err("FP64 interleaved buffers must be 16-byte aligned");
        return -1;
    }
    switch (p->kind) {
    case PK_C2C64: {
        fafk_acc64 in = acc(ir, inter ? NULL : ii, 1, idist, 0, 1);
        fafk_acc64 out = acc(orr, inter ? NULL : oi, 1, odist, 0, 1);
        return k64(fafk64_fft(fafk_plan_record(&p->a), (int)p->n, p->tw64, p->inverse, count, 0,
                              &in, &out, NULL, 0, p->scale64, stream));
    }
    case PK_C2C64_4S: {
        const long n = (long)p->n, n1 = p->n1, n2 = p->n2;
        const long cs = chunk64(p, count);
        double *ws = fafi_workspace(p, (size_t)cs * (size_t)n * 16);
        if (!ws) return -1;
        for (long b0 = 0; b0 < count; b0 += cs) {
            const long c = count - b0 < cs ? count - b0 : cs;
            const long fi = inter ? 2 : 1;
            /* columns: c * n2 transforms of n1 points, W_n^(k1 c) on output */
            fafk_acc64 in = acc(ir + b0 * idist * fi, inter ? NULL : ii + b0 * idist, n2, idist, 1,
                                n2);
            fafk_acc64 w = acc(ws, NULL, n2, n, 1, n2);
            if (k64(fafk64_fft(fafk_plan_record(&p->a), (int)n1, p->tw64, p->inverse, c * n2, 1,
                               &in, &w, p->twn64, n, 1.0, stream)))
                return -1;
            /* rows: c * n1 transforms of n2 points, written transposed */
            fafk_acc64 r = acc(ws, NULL, n1, n, n2, 1);
            fafk_acc64 out = acc(orr + b0 * odist * fi, inter ? NULL : oi + b0 * odist, n1, odist,
                                 1, n1);
            if (k64(fafk64_fft(fafk_plan_record(&p->b), (int)n2, p->tw64b, p->inverse, c * n1, 1,
                               &r, &out, NULL, 0, p->scale64, stream)))
                return -1;
        }
        return 0;
    }
    case PK_RADER64:
    case PK_BLUE64: {
        if (inter) { fafi_err("Rader/Bluestein lengths need SPLIT layout on the device path"); return -1; }
        const int rader = p->kind == PK_RADER64;
        const size_t n = p->n, L = rader ? n - 1 : (size_t)p->blue.m;
        const long cs = chunk64(p, count);
        const size_t ab = a256((size_t)cs * L * 32), zb = a256((size_t)cs * 16);
        const size_t subb = need64(p, cs) - ab - zb;
        char *ws = fafi_workspace(p, need64(p, cs));
        if (!ws) return -1;
        double *ar = (double *)ws, *ai = ar + (size_t)cs * L;
        double *br = ai + (size_t)cs * L, *bi = br + (size_t)cs * L;
        void *z = ws + ab;
        char *sub = ws + ab + zb;
        const long l = (long)L;
        for (long s0 = 0; s0 < count; s0 += cs) {
            const long c = count - s0 < cs ? count - s0 : cs;
            const double *xr = ir + s0 * idist, *xi = ii + s0 * idist;
            double *yr = orr + s0 * odist, *yi = oi + s0 * odist;
            if (k64(fafk64_in(xr, xi, idist, rader ? p->rader.rin : NULL,
                              rader ? NULL : p->blue.chirp, (long)n, l, ar, ai, l, c, stream)))
                return -1;
            if (run_lent(p->inner_f, sub, subb, ar, ai, l, br, bi, l, c, stream)) return -1;
            if (k64(fafk64_mul(br, bi, l, rader ? p->rader.b : p->blue.b, l, rader ? z : NULL, c,
                               stream)))
                return -1;
            if (run_lent(p->inner_i, sub, subb, br, bi, l, ar, ai, l, c, stream)) return -1;
            if (k64(fafk64_out(ar, ai, l, rader ? p->rader.rout : NULL,
                               rader ? NULL : p->blue.chirp, rader ? l : (long)n, xr, xi, idist,
                               z, yr, yi, odist, p->scale64, c, stream)))
                return -1;
        }
        return 0;
    }
    default:
        fafi_err("plan is not an FP64 C2C transform");
        return -1;
    }
}

int fafi_run_r2c64(fafh_plan *p, const double *x, long idist, double *orr, double *oi,
                   long odist, long count, void *stream) {
    const long n = (long)p->n, nc = n / 2;
    if ((idist & 1) || !aligned16(x, x)) {
        fafi_err("FP64 rfft reads signals as complex pairs: the input must be 16-byte aligned "
                 "with an even row distance");
        return -1;
    }
    const long cs = chunk64(p, count);
    const size_t zb = a256((size_t)cs * (size_t)nc * 16);
    const size_t subb = need64(p, cs) - zb;
    char *ws = fafi_workspace(p, zb + subb);
    if (!ws) return -1;
    double *Z = (double *)ws;
    for (long s0 = 0; s0 < count; s0 += cs) {
        const long c = count - s0 < cs ? count - s0 : cs;
        /* z[k] = x[2k] + i x[2k+1]: the real row read as nc complex pairs */
        if (run_lent(p->inner_f, ws + zb, subb, x + s0 * idist, NULL, idist / 2, Z, NULL, nc, c,
                     stream))
            return -1;
        const long fo = p->interleaved ? 2 : 1;
        if (k64(fafk64_post(Z, nc, p->tw64, nc, orr + s0 * odist * fo,
                            p->interleaved ? NULL : oi + s0 * odist, odist, p->scale64, c,
                            stream)))
            return -1;
    }
    return 0;
}

int fafi_run_c2r64(fafh_plan *p, const double *ir, const double *ii, long idist, double *y,
                   long odist, long count, void *stream) {
    const long n = (long)p->n, nc = n / 2;
    if ((odist & 1) || !aligned16(y, y)) {
        fafi_err("FP64 irfft writes signals as complex pairs: the output must be 16-byte "
                 "aligned with an even row distance");
        return -1;
    }
    const long cs = chunk64(p, count);
    const size_t zb = a256((size_t)cs * (size_t)nc * 16);
    const size_t subb = need64(p, cs) - zb;
    char *ws = fafi_workspace(p, zb + subb);
    if (!ws) return -1;
    double *Z = (double *)ws;
    const long fi = p->interleaved ? 2 : 1;
    for (long s0 = 0; s0 < count; s0 += cs) {
        const long c = count - s0 < cs ? count - s0 : cs;
        if (k64(fafk64_pre(ir + s0 * idist * fi, p->interleaved ? NULL : ii + s0 * idist, idist,
                           p->tw64, nc, Z, nc, c, stream)))
            return -1;
        fafh_plan *q = p->inner_f;
        const double keep = q->scale64;
        q->scale64 = p->scale64;         /* the norm rides on the inverse passes */
        int rc = run_lent(q, ws + zb, subb, Z, NULL, nc, y + s0 * odist, NULL, odist / 2, c,
                          stream);
        q->scale64 = keep;
        if (rc) return -1;
    }
    return 0;
}

/* DCT / DST / MDCT in FP64: the FP32 plan's pre/post tables rebuilt in
 * double around an FP64 inner FFT */
int fafi_run_r2r64(fafh_plan *p, const double *x, long idist, double *y, long odist, long count,
                   void *stream) {
    const long L = p->r2r_L;
    fafh_plan *q = p->inner_f;
    const long cs = chunk64(p, count);
    const size_t ub = a256((size_t)cs * (size_t)L * 32), subb = need64(p, cs) - ub;
    char *ws = fafi_workspace(p, ub + subb);
    if (!ws) return -1;
    double *ur = (double *)ws, *ui = ur + (size_t)cs * L;
    double *vr = ui + (size_t)cs * L, *vi = vr + (size_t)cs * L;
    for (long s0 = 0; s0 < count; s0 += cs) {
        const long c = count - s0 < cs ? count - s0 : cs;
        if (k64(fafk64_r2r_pre(x + s0 * idist, idist, p->r2r_nin, p->r2r_pre, p->r2r_shift, L, ur,
                               ui, L, c, stream)))
            return -1;
        if (run_lent(q, ws + ub, subb, ur, ui, L, vr, vi, L, c, stream)) return -1;
        if (k64(fafk64_r2r_post(vr, vi, L, p->r2r_post, p->r2r_off, p->r2r_im, p->r2r_nout,
                                y + s0 * odist, odist, c, stream)))
            return -1;
    }
    return 0;
}
