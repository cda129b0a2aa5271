/*
 * faf_plan.c -- create-time planning for the device kernels (C11).
 *
 * Chooses radix passes and launch shapes, and builds the device tables:
 * twiddles W_n^m, Kiss super-twiddles, Rader and Bluestein tables.  Every
 * table value is computed in FP64 from an exact integer phase index and
 * rounded once to FP32 (the reference computes its mixed-radix twiddles
 * from an FP32 angle, src/faf_twiddles.c:43-51; the device path does not).
 */
#define _GNU_SOURCE
#include "faf_internal.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* size policy -- reference src/faf_fft_stage.c:25-135                 */
/* ------------------------------------------------------------------ */

static size_t strip(size_t n, size_t p) { while (n % p == 0) n /= p; return n; }

int fafi_is_5_smooth(size_t n) { return n && strip(strip(strip(n, 2), 3), 5) == 1; }
int fafi_is_7_smooth(size_t n) { return n && strip(strip(strip(strip(n, 2), 3), 5), 7) == 1; }

int fafi_is_prime(size_t n) {
    if (n < 2) return 0;
    if (n % 2 == 0) return n == 2;
    for (size_t d = 3; d * d <= n; d += 2)
        if (n % d == 0) return 0;
    return 1;
}

int fafi_rader_eligible(size_t n) {
    return n >= 11 && fafi_is_prime(n) && fafi_is_7_smooth(n - 1);
}

static uint64_t powmod_u64(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    b %= m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return r;
}

size_t fafi_primitive_root(size_t p) {
    if (p < 3 || !fafi_is_prime(p)) return 0;
    size_t q[24];
    int nq = 0;
    size_t m = p - 1;
    for (size_t d = 2; d * d <= m; ++d)
        if (m % d == 0) { q[nq++] = d; m = strip(m, d); }
    if (m > 1) q[nq++] = m;
    for (size_t g = 2; g < p; ++g) {
        int ok = 1;
        for (int i = 0; i < nq && ok; ++i) ok = powmod_u64(g, (p - 1) / q[i], p) != 1;
        if (ok) return g;
    }
    return 0;
}

size_t fafi_next_5_smooth(size_t n) {
    if (n <= 1) return 1;
    while (!fafi_is_5_smooth(n)) ++n;
    return n;
}

int fafi_is_codelet_size(size_t n) { return (n >= 2 && n <= 10) || n == 12; }

/* ------------------------------------------------------------------ */
/* radix passes                                                        */
/* ------------------------------------------------------------------ */

static int per_thread(int r) { return FAFD_MAXE / r; }

/* default pass order: radix 16 while possible, then the leftover power of
 * two as one pass, then 3s, 5s, 7s */
static int default_radices(size_t n, int allow7, int *rad) {
    int k = 0;
    size_t m = n;
    int a = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    while (a >= 4) { rad[k++] = 16; a -= 4; }
    if (a == 3) rad[k++] = 8;
    else if (a == 2) rad[k++] = 4;
    else if (a == 1) rad[k++] = 2;
    while (m % 3 == 0) { rad[k++] = 3; m /= 3; }
    while (m % 5 == 0) { rad[k++] = 5; m /= 5; }
    while (allow7 && m % 7 == 0) { rad[k++] = 7; m /= 7; }
    if (m != 1 || k > FAFD_MAX_STAGES) return -1;
    return k;
}

static int threads_for(size_t n, const int *rad, int k) {
    int T = 1;
    for (int s = 0; s < k; ++s) {
        size_t nb = n / (size_t)rad[s];
        int need = (int)((nb + (size_t)per_thread(rad[s]) - 1) / (size_t)per_thread(rad[s]));
        if (need > T) T = need;
    }
    return T;
}

static int plane_for(size_t n) { return (int)(n + (n >> 4) + 1); }

int fafi_radix_fits(size_t n, int allow7) {
    int rad[FAFD_MAX_STAGES];
    if (n < 2) return 0;
    int k = default_radices(n, allow7, rad);
    if (k <= 0) return 0;
    return threads_for(n, rad, k) <= 1024 &&
           (size_t)plane_for(n) * 8 <= FAFD_MAX_LDS_BYTES;
}

int fafi_make_fft(size_t n, int inv, int allow7, const int *order, int norder, fafd_fft *P) {
    int rad[FAFD_MAX_STAGES];
    int k;
    memset(P, 0, sizeof *P);
    if (order && norder > 0) {
        size_t prod = 1;
        for (int s = 0; s < norder; ++s) { rad[s] = order[s]; prod *= (size_t)order[s]; }
        if (prod != n) { fafi_err("radix order does not multiply to %zu", n); return -1; }
        k = norder;
    } else {
        k = default_radices(n, allow7, rad);
    }
    if (k <= 0) {
        fafi_err("length %zu has no radix-2/3/5%s decomposition", n, allow7 ? "/7" : "");
        return -1;
    }
    P->n = (int32_t)n;
    P->nstages = k;
    P->threads = threads_for(n, rad, k);
    P->plane = plane_for(n);
    size_t ns = 1;
    for (int s = 0; s < k; ++s) {
        P->st[s].radix = rad[s];
        P->st[s].ns = (int32_t)ns;
        P->st[s].ns_magic = ns > 1 ? (uint32_t)(((uint64_t)1 << 32) / ns +
                                                ((((uint64_t)1 << 32) % ns) ? 1 : 0))
                                   : 0u;
        P->st[s].tw_stride = (int32_t)(n / (ns * (size_t)rad[s]));
        ns *= (size_t)rad[s];
    }
    if (P->threads > 1024) {
        fafi_err("length %zu needs %d threads in one workgroup", n, P->threads);
        return -1;
    }
    P->tw = fafi_twiddles(n, inv);
    if (!P->tw) return -1;
    /* the kernels read the record from device memory: upload it now rather
     * than at the first launch */
    if (!fafk_plan_record(P)) {
        fafi_err("%s", fafk_last_error());
        return -1;
    }
    return 0;
}

int fafi_default_spb(const fafd_fft *P, size_t extra_per_signal) {
    int spb = 256 / P->threads;
    if (spb < 1) spb = 1;
    size_t per = ((size_t)P->plane + extra_per_signal) * 8;
    size_t budget = 80 * 1024;                 /* leave room for two workgroups per CU */
    if (per > budget) budget = FAFD_MAX_LDS_BYTES;
    while (spb > 1 && (size_t)spb * per > budget) --spb;
    if ((size_t)spb * per > FAFD_MAX_LDS_BYTES) return 0;
    return spb;
}

/* n = n1 * n2 with n1 <= n2, n1 as close to sqrt(n) as possible and both
 * lengths runnable in one workgroup (radix 7 allowed when allow7) */
static int split_4step(size_t n, int allow7, int *n1, int *n2) {
    size_t best = 0;
    for (size_t a = 2; a * a <= n; ++a) {
        if (n % a) continue;
        size_t b = n / a;
        if (fafi_radix_fits(a, allow7) && fafi_radix_fits(b, allow7)) best = a;
    }
    if (!best) return -1;
    *n1 = (int)best;
    *n2 = (int)(n / best);
    return 0;
}

int fafi_split_4step(size_t n, int *n1, int *n2) { return split_4step(n, 0, n1, n2); }

/* ------------------------------------------------------------------ */
/* table cache: device tables keyed by (kind, n, direction), kept for   */
/* the life of the process                                              */
/* ------------------------------------------------------------------ */

typedef struct tab_ent {
    int kind;
    size_t n;
    int inv;
    void *dev;
    struct tab_ent *next;
} tab_ent;

static tab_ent *g_tabs;
static pthread_mutex_t g_tab_mu = PTHREAD_MUTEX_INITIALIZER;

static void *tab_get(int kind, size_t n, int inv) {
    void *r = NULL;
    pthread_mutex_lock(&g_tab_mu);
    for (tab_ent *e = g_tabs; e; e = e->next)
        if (e->kind == kind && e->n == n && e->inv == inv) { r = e->dev; break; }
    pthread_mutex_unlock(&g_tab_mu);
    return r;
}

static void *tab_put(int kind, size_t n, int inv, const float *host, size_t floats) {
    void *d = fafk_malloc(floats * sizeof(float));
    if (!d) { fafi_err("%s", fafk_last_error()); return NULL; }
    if (fafk_h2d(d, host, floats * sizeof(float), NULL) || fafk_sync(NULL)) {
        fafi_err("%s", fafk_last_error());
        fafk_free(d);
        return NULL;
    }
    pthread_mutex_lock(&g_tab_mu);
    for (tab_ent *e = g_tabs; e; e = e->next)
        if (e->kind == kind && e->n == n && e->inv == inv) {   /* lost a race */
            pthread_mutex_unlock(&g_tab_mu);
            fafk_free(d);
            return e->dev;
        }
    tab_ent *e = malloc(sizeof *e);
    e->kind = kind; e->n = n; e->inv = inv; e->dev = d; e->next = g_tabs;
    g_tabs = e;
    pthread_mutex_unlock(&g_tab_mu);
    return d;
}

/* exp(sign * 2 pi i * m / n) with the phase reduced by symmetry so the FP64
 * argument stays within [0, pi/4] */
static void unit_root(size_t m, size_t n, int inv, double *c, double *s) {
    /* angle = 2 pi m / n = (pi / 4n) A with A = 8 (m mod n) in [0, 8n):
     * folded in exact integers into [0, pi/4] (sin and cos signs and a
     * swap), so the FP64 sin/cos see a small argument with one rounding */
    uint64_t A = 8 * (uint64_t)(m % n), N = n;
    int neg_s = 0, neg_c = 0, swap = 0;
    if (A > 4 * N) { A = 8 * N - A; neg_s = 1; }       /* (pi, 2pi): sin < 0 */
    if (A > 2 * N) { A = 4 * N - A; neg_c = 1; }       /* (pi/2, pi]: cos < 0 */
    if (A > N) { A = 2 * N - A; swap = 1; }            /* (pi/4, pi/2]: cos <-> sin */
    const double x = (M_PI / 4.0) * ((double)A / (double)N);
    double cc = cos(x), ss = sin(x);
    if (swap) { double t = cc; cc = ss; ss = t; }
    if (neg_c) cc = -cc;
    if (neg_s) ss = -ss;
    *c = cc;
    *s = inv ? ss : -ss;
}

void fafi_unit_root(size_t m, size_t n, int inv, double *c, double *s) {
    unit_root(m, n, inv, c, s);
}

const void *fafi_twiddles(size_t n, int inv) {
    void *d = tab_get(1, n, inv);
    if (d) return d;
    float *h = malloc(2 * n * sizeof(float));
    for (size_t m = 0; m < n; ++m) {
        double c, s;
        unit_root(m, n, inv, &c, &s);
        h[2 * m] = (float)c;
        h[2 * m + 1] = (float)s;
    }
    d = tab_put(1, n, inv, h, 2 * n);
    free(h);
    return d;
}

/* Kiss super-twiddles indexed by k = 1..nc/2 (slot 0 unused):
 * forward exp(-i pi (k/nc + 1/2)), inverse the conjugate
 * (reference src/faf_rfft.c:37-45, which indexes them by k-1) */
const void *fafi_kiss_table(size_t nc, int inv) {
    void *d = tab_get(2, nc, inv);
    if (d) return d;
    size_t cnt = nc / 2 + 1;
    float *h = calloc(2 * cnt, sizeof(float));
    for (size_t k = 1; k < cnt; ++k) {
        /* -i * exp(-i pi k / nc) = -i * W_{2nc}^k */
        double c, s;
        unit_root(k, 2 * nc, 0, &c, &s);         /* c + i s = exp(-i pi k/nc) */
        double re = s, im = -c;                  /* times -i */
        h[2 * k] = (float)re;
        h[2 * k + 1] = (float)(inv ? -im : im);
    }
    d = tab_put(2, nc, inv, h, 2 * cnt);
    free(h);
    return d;
}

/* ------------------------------------------------------------------ */
/* host FP64 FFT for create-time tables (Rader and Bluestein kernels)   */
/* ------------------------------------------------------------------ */

static size_t least_factor(size_t n) {
    for (size_t p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

static void fft64_rec(size_t n, size_t s, const double *xr, const double *xi, double *yr,
                      double *yi, const double *cr, const double *ci, size_t N,
                      double *scratch) {
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    size_t p = least_factor(n), m = n / p;
    if (p == n) {
        for (size_t k = 0; k < n; ++k) {
            double ar = 0, ai = 0;
            for (size_t j = 0; j < n; ++j) {
                size_t e = (j * k % n) * (N / n);
                ar += xr[j * s] * cr[e] - xi[j * s] * ci[e];
                ai += xr[j * s] * ci[e] + xi[j * s] * cr[e];
            }
            yr[k] = ar; yi[k] = ai;
        }
        return;
    }
    for (size_t r = 0; r < p; ++r)
        fft64_rec(m, s * p, xr + r * s, xi + r * s, yr + r * m, yi + r * m, cr, ci, N,
                  scratch);
    double *tr = scratch, *ti = scratch + n;
    for (size_t q = 0; q < p; ++q)
        for (size_t k = 0; k < m; ++k) {
            size_t kk = q * m + k;
            double ar = 0, ai = 0;
            for (size_t r = 0; r < p; ++r) {
                size_t e = (r * kk % n) * (N / n);
                double vr = yr[r * m + k], vi = yi[r * m + k];
                ar += vr * cr[e] - vi * ci[e];
                ai += vr * ci[e] + vi * cr[e];
            }
            tr[kk] = ar; ti[kk] = ai;
        }
    memcpy(yr, tr, n * sizeof(double));
    memcpy(yi, ti, n * sizeof(double));
}

int fafi_fft64(size_t n, int inv, double *re, double *im) {
    double *cr = malloc(n * sizeof(double)), *ci = malloc(n * sizeof(double));
    double *xr = malloc(n * sizeof(double)), *xi = malloc(n * sizeof(double));
    double *sc = malloc(2 * n * sizeof(double));
    if (!cr || !ci || !xr || !xi || !sc) {
        free(cr); free(ci); free(xr); free(xi); free(sc);
        fafi_err("out of host memory");
        return -1;
    }
    for (size_t m = 0; m < n; ++m) unit_root(m, n, inv, &cr[m], &ci[m]);
    memcpy(xr, re, n * sizeof(double));
    memcpy(xi, im, n * sizeof(double));
    fft64_rec(n, 1, xr, xi, re, im, cr, ci, n, sc);
    free(cr); free(ci); free(xr); free(xi); free(sc);
    return 0;
}

/* ------------------------------------------------------------------ */
/* plan objects                                                        */
/* ------------------------------------------------------------------ */

fafh_plan *fafi_plan_new(void) {
    fafh_plan *p = calloc(1, sizeof(fafh_plan));
    if (p) p->scale = 1.0f;
    return p;
}

void fafi_plan_free(fafh_plan *p) {
    if (!p) return;
    fafi_cwt_free(p);
    fafi_plan_free(p->inner_f);
    fafi_plan_free(p->inner_i);
    fafk_free(p->r2r_pre);
    fafk_free(p->r2r_post);
    fafi_plan_free(p->sub_r2c);
    fafi_plan_free(p->sub_c2c);
    fafi_plan_free(p->sub_c2r);
    for (int i = 0; i < 3; ++i) fafk_free(p->rader_mem[i]);
    for (int i = 0; i < 2; ++i) fafk_free(p->blue_mem[i]);
    fafk_free(p->ws);
    for (int i = 0; i < p->n_retired; ++i) fafk_free(p->ws_retired[i]);
    free(p->ws_retired);
    fafk_free(p->stg);
    fafk_free(p->h_dev);
    fafk_free(p->win_dev);
    free(p);
}

void *fafi_workspace(fafh_plan *p, size_t bytes) {
    if (p->ws_ext) {
        if (p->ws_ext_bytes >= bytes) return p->ws_ext;
        if (fafi_noalloc())
            fafi_err("faf_execute_batch_noalloc: the run needs %zu bytes of scratch, more than "
                     "the %zu bytes of the static scratch; use a smaller count",
                     bytes, p->ws_ext_bytes);
        else
            fafi_err("lent workspace holds %zu bytes, the run needs %zu", p->ws_ext_bytes,
                     bytes);
        return NULL;
    }
    if (p->ws_bytes >= bytes) {
        /* a graph captured now keeps this buffer's address */
        if (fafi_capturing()) p->ws_captured = 1;
        return p->ws;
    }
    if (fafi_capturing()) {
        fafi_err("the workspace must grow to %zu bytes, which a graph capture cannot do; "
                 "execute once before capturing", bytes);
        return NULL;
    }
    fafk_sync(fafi_stream());
    if (p->ws_captured) {
        /* graphs may still replay on the old buffer: keep it until the plan
         * is destroyed instead of freeing it under them */
        void **rt = realloc(p->ws_retired, (size_t)(p->n_retired + 1) * sizeof(void *));
        if (!rt) { fafi_err("workspace: out of host memory"); return NULL; }
        p->ws_retired = rt;
        p->ws_retired[p->n_retired++] = p->ws;
        p->ws_captured = 0;
    } else {
        fafk_free(p->ws);
    }
    p->ws = fafk_malloc(bytes);
    p->ws_bytes = p->ws ? bytes : 0;
    if (!p->ws) fafi_err("%s", fafk_last_error());
    return p->ws;
}

void *fafi_staging(fafh_plan *p, size_t bytes) {
    if (p->stg_bytes >= bytes) return p->stg;
    if (fafi_capturing()) { fafi_err("staging cannot grow inside a graph capture"); return NULL; }
    fafk_sync(fafi_stream());
    fafk_free(p->stg);
    p->stg = fafk_malloc(bytes);
    p->stg_bytes = p->stg ? bytes : 0;
    if (!p->stg) fafi_err("%s", fafk_last_error());
    return p->stg;
}

static void *upload(const void *host, size_t bytes) {
    void *d = fafk_malloc(bytes);
    if (!d) { fafi_err("%s", fafk_last_error()); return NULL; }
    if (fafk_h2d(d, host, bytes, NULL) || fafk_sync(NULL)) {
        fafi_err("%s", fafk_last_error());
        fafk_free(d);
        return NULL;
    }
    return d;
}

static fafh_plan *plan_inner(size_t m, int inv);

static int plan_rader(fafh_plan *p, size_t n, int inv) {
    /* reference src/faf_fft_rader.c:19-134: a[j] = x[g^j], B = FFT_m(w^(g^-j));
     * here B is built in FP64 and carries the 1/m of the inner inverse */
    size_t m = n - 1, g = fafi_primitive_root(n);
    size_t ginv = (size_t)powmod_u64(g, n - 2, n);
    const int one_wg = fafi_radix_fits(m, 1);
    if (one_wg &&
        (fafi_make_fft(m, 0, 1, NULL, 0, &p->a) || fafi_make_fft(m, 1, 1, NULL, 0, &p->b)))
        return -1;
    int32_t *rin = malloc(m * sizeof(int32_t)), *rout = malloc(m * sizeof(int32_t));
    double *br = malloc(m * sizeof(double)), *bi = malloc(m * sizeof(double));
    float *bf = malloc(2 * m * sizeof(float));
    size_t a = 1, b = 1;
    for (size_t j = 0; j < m; ++j) {
        rin[j] = (int32_t)a;
        rout[j] = (int32_t)b;
        unit_root(b, n, inv, &br[j], &bi[j]);
        a = a * g % n;
        b = b * ginv % n;
    }
    int rc = fafi_fft64(m, 0, br, bi);
    for (size_t j = 0; j < m && !rc; ++j) {
        bf[2 * j] = (float)(br[j] / (double)m);
        bf[2 * j + 1] = (float)(bi[j] / (double)m);
    }
    if (!rc) {
        p->rader_mem[0] = upload(rin, m * sizeof(int32_t));
        p->rader_mem[1] = upload(rout, m * sizeof(int32_t));
        p->rader_mem[2] = upload(bf, 2 * m * sizeof(float));
        rc = (p->rader_mem[0] && p->rader_mem[1] && p->rader_mem[2]) ? 0 : -1;
    }
    free(rin); free(rout); free(br); free(bi); free(bf);
    if (rc) return -1;
    p->rader.p = (int32_t)n;
    p->rader.rin = p->rader_mem[0];
    p->rader.rout = p->rader_mem[1];
    p->rader.b = p->rader_mem[2];
    p->kind = PK_RADER;
    p->spb = one_wg ? fafi_default_spb(&p->a, (size_t)(n + (n >> 4) + 1)) : 0;
    if (p->spb >= 1) return 0;
    /* beyond one workgroup: gather, inner FFTs through HBM, scatter */
    p->inner_f = plan_inner(m, 0);
    p->inner_i = p->inner_f ? plan_inner(m, 1) : NULL;
    if (!p->inner_i) return -1;
    p->kind = PK_RADER_L;
    p->spb = 1;
    return 0;
}

static int plan_blue(fafh_plan *p, size_t n, int inv) {
    /* reference src/faf_bluestein.c:19-163 with the chirp phase k^2 reduced
     * mod 2n in integers before the FP64 sin/cos */
    size_t M = fafi_next_5_smooth(2 * n - 1);
    const int one_wg = fafi_radix_fits(M, 0);
    if (one_wg &&
        (fafi_make_fft(M, 0, 0, NULL, 0, &p->a) || fafi_make_fft(M, 1, 0, NULL, 0, &p->b)))
        return -1;
    float *ch = malloc(2 * n * sizeof(float));
    double *br = calloc(M, sizeof(double)), *bi = calloc(M, sizeof(double));
    float *bf = malloc(2 * M * sizeof(float));
    for (size_t k = 0; k < n; ++k) {
        uint64_t q = ((uint64_t)k * (uint64_t)k) % (uint64_t)(2 * n);
        double c, s;
        unit_root((size_t)q, 2 * n, inv, &c, &s);   /* exp(-+ i pi q / n) */
        ch[2 * k] = (float)c;
        ch[2 * k + 1] = (float)s;
        /* kernel = conj(chirp) at l and M-l */
        br[k] = c; bi[k] = -s;
        if (k) { br[M - k] = c; bi[M - k] = -s; }
    }
    int rc = fafi_fft64(M, 0, br, bi);
    for (size_t j = 0; j < M && !rc; ++j) {
        bf[2 * j] = (float)(br[j] / (double)M);
        bf[2 * j + 1] = (float)(bi[j] / (double)M);
    }
    if (!rc) {
        p->blue_mem[0] = upload(ch, 2 * n * sizeof(float));
        p->blue_mem[1] = upload(bf, 2 * M * sizeof(float));
        rc = (p->blue_mem[0] && p->blue_mem[1]) ? 0 : -1;
    }
    free(ch); free(br); free(bi); free(bf);
    if (rc) return -1;
    p->blue.n = (int32_t)n;
    p->blue.m = (int32_t)M;
    p->blue.chirp = p->blue_mem[0];
    p->blue.b = p->blue_mem[1];
    p->kind = PK_BLUE;
    p->spb = one_wg ? fafi_default_spb(&p->a, 0) : 0;
    if (p->spb >= 1) return 0;
    /* beyond one workgroup: chirp, inner FFTs through HBM, chirp */
    p->inner_f = plan_inner(M, 0);
    p->inner_i = p->inner_f ? plan_inner(M, 1) : NULL;
    if (!p->inner_i) return -1;
    p->kind = PK_BLUE_L;
    p->spb = 1;
    return 0;
}

static int plan_4step_ex(fafh_plan *p, size_t n, int inv, int allow7) {
    int n1, n2;
    if (split_4step(n, allow7, &n1, &n2)) {
        fafi_err("length %zu has no four-step split", n);
        return -1;
    }
    if (fafi_make_fft((size_t)n1, inv, allow7, NULL, 0, &p->a) ||
        fafi_make_fft((size_t)n2, inv, allow7, NULL, 0, &p->b))
        return -1;
    p->n1 = n1;
    p->n2 = n2;
    p->twn = fafi_twiddles(n, inv);
    if (!p->twn) return -1;
    /* columns per workgroup for pass A, rows per workgroup for pass B */
    int c = 32;
    while (c > 1 && (c * p->a.threads > 1024 || (size_t)c * p->a.plane * 8 > FAFD_MAX_LDS_BYTES))
        c >>= 1;
    int r = 32;
    while (r > 1 && (r * p->b.threads > 1024 || (size_t)r * p->b.plane * 8 > FAFD_MAX_LDS_BYTES))
        r >>= 1;
    p->cols = c;
    p->rows = r;
    p->ct4 = !allow7 && fafi_ct_enabled() && fafk_ct4_supported(n1, n2);
    p->rs = !allow7 && fafi_ct_enabled() && n == 32768;
    return 0;
}

static int plan_4step(fafh_plan *p, size_t n, int inv) { return plan_4step_ex(p, n, inv, 0); }

/* The inner C2C of a Rader / Bluestein transform too long for one
 * workgroup: split planes, unnormalised.  5-smooth lengths get the full C2C
 * planner (specialised kernels included); lengths with a factor 7 (Rader
 * primes with p-1 7-smooth) get runtime radix-7 passes, in one workgroup
 * or four-step. */
static fafh_plan *plan_inner(size_t m, int inv) {
    fafh_plan *q = fafi_plan_new();
    if (!q) { fafi_err("alloc failed"); return NULL; }
    q->scale = 1.0f;
    int rc;
    if (fafi_is_5_smooth(m)) {
        rc = fafi_plan_c2c(q, m, inv, 0);
    } else if (fafi_radix_fits(m, 1)) {
        q->n = m;
        q->inverse = inv;
        rc = fafi_make_fft(m, inv, 1, NULL, 0, &q->a);
        q->kind = PK_C2C;
        q->spb = rc ? 0 : fafi_default_spb(&q->a, 0);
        if (!rc && q->spb < 1) rc = -1;
    } else {
        q->n = m;
        q->inverse = inv;
        rc = plan_4step_ex(q, m, inv, 1);
        q->kind = PK_C2C_4S;
    }
    if (rc) { fafi_plan_free(q); return NULL; }
    return q;
}

int fafi_plan_c2c(fafh_plan *p, size_t n, int inv, int bluestein) {
    p->n = n;
    p->inverse = inv;
    int legal = fafi_is_5_smooth(n) || fafi_is_codelet_size(n);
    if (!legal && fafi_rader_eligible(n)) return plan_rader(p, n, inv);
    if (!legal && fafi_is_7_smooth(n) && !bluestein) {
        fafi_err("size %zu is not a public FFT length", n);
        return -1;
    }
    if (!legal) {
        if (!bluestein) { fafi_err("size %zu needs FAF_FLAG_BLUESTEIN", n); return -1; }
        return plan_blue(p, n, inv);
    }
    if (n == 1) {
        fafi_err("length 1 is not a transform");
        return -1;
    }
    /* a legal length carries a factor 7 only as the codelet n = 7 */
    if (fafi_radix_fits(n, 1)) {
        if (fafi_make_fft(n, inv, 1, NULL, 0, &p->a)) return -1;
        p->kind = PK_C2C;
        p->ct = fafi_ct_enabled() && fafk_ct_supports(0, (int)n);
        p->spb = fafi_default_spb(&p->a, 0);
        if (p->spb < 1) return -1;
        return fafi_apply_tuned(p);
    }
    if (plan_4step(p, n, inv)) return -1;
    p->kind = PK_C2C_4S;
    return 0;
}

int fafi_plan_real(fafh_plan *p, size_t n, int inv) {
    size_t nc = n / 2;
    p->n = n;
    p->inverse = inv;
    if (nc == 1) {
        fafi_err("rfft length 2 is handled by the codelet path only on the host");
        return -1;
    }
    p->post = fafi_kiss_table(nc, 0);
    p->pre = fafi_kiss_table(nc, 1);
    if (!p->post || !p->pre) return -1;
    if (fafi_radix_fits(nc, 0)) {
        if (fafi_make_fft(nc, inv, 0, NULL, 0, &p->a)) return -1;
        p->kind = inv ? PK_C2R : PK_R2C;
        p->ct = fafi_ct_enabled() && fafk_ct_supports(1, (int)nc);
        /* the single-pass register-resident kernels: rfft / irfft 16384 and
         * irfft 4096 (the rfft 4096 one-signal-per-workgroup kernel is faster
         * than the single pass) */
        p->rs = fafi_ct_enabled() && (nc == 8192 || (nc == 2048 && inv));
        p->spb = fafi_default_spb(&p->a, 0);
        if (p->spb < 1) return -1;
        return fafi_apply_tuned(p);
    }
    if (plan_4step(p, nc, inv)) return -1;
    p->kind = inv ? PK_C2R_4S : PK_R2C_4S;
    return 0;
}

int fafi_plan_pipe(fafh_plan *p, size_t n) {
    size_t nc = n / 2;
    p->n = n;
    if (nc < 2 || !fafi_radix_fits(nc, 0)) {
        fafi_err("fused pipeline length %zu must be even with n/2 in one workgroup", n);
        return -1;
    }
    p->post = fafi_kiss_table(nc, 0);
    p->pre = fafi_kiss_table(nc, 1);
    if (!p->post || !p->pre) return -1;
    if (fafi_make_fft(nc, 0, 0, NULL, 0, &p->a) || fafi_make_fft(nc, 1, 0, NULL, 0, &p->b))
        return -1;
    p->kind = PK_PIPE;
    p->ct = fafi_ct_enabled() && fafk_ct_supports(3, (int)nc);
    p->spb = fafi_default_spb(&p->a, 0);
    if (p->spb < 1) return -1;
    return fafi_apply_tuned(p);
}

int fafi_plan_stft(fafh_plan *p, size_t win, size_t hop) {
    size_t nc = win / 2;
    if (win < 4 || (win & 1) || !fafi_is_5_smooth(win) || !fafi_radix_fits(nc, 0)) {
        fafi_err("STFT window %zu must be even, 5-smooth and fit one workgroup", win);
        return -1;
    }
    if (hop == 0) { fafi_err("STFT hop must be > 0"); return -1; }
    p->post = fafi_kiss_table(nc, 0);
    if (!p->post || fafi_make_fft(nc, 0, 0, NULL, 0, &p->a)) return -1;
    /* symmetric Hann, reference src/faf_twiddles.c:159-163, from FP64 */
    float *w = malloc(win * sizeof(float));
    for (size_t i = 0; i < win; ++i) {
        double c, s;
        unit_root(i, win - 1, 1, &c, &s);
        w[i] = (float)(0.5 * (1.0 - c));
    }
    p->win_dev = upload(w, win * sizeof(float));
    free(w);
    if (!p->win_dev) return -1;
    p->win = win;
    p->hop = hop;
    p->kind = PK_STFT;
    p->ct = fafi_ct_enabled() && fafk_ct_supports(4, (int)nc);
    p->spb = fafi_default_spb(&p->a, 0);
    return p->spb >= 1 ? 0 : -1;
}
