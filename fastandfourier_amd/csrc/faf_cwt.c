/*
 * faf_cwt.c -- CWT filter bank (C11 host side) and its device launches.
 *
 * The bank is built on the host in FP64 at create time, following the
 * reference's construction (src/faf_cwt.c:104-677: scale geometry, L2
 * shadow with plateau rescale, residual lowpass, Littlewood-Paley report,
 * live FP32 table, dual-frame table, wrap check), then uploaded once.  The
 * forward transform runs as one fused kernel per segment: one rfft, then
 * per row a one-sided multiply and an inverse FFT.
 */
#define _GNU_SOURCE
#include "faf_internal.h"
#include <pthread.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define LN2 0.693147180559945309417
#define MAX_SCALES 512

typedef struct fafi_cwt_bank {
    faf_cwt_config cfg;      /* resolved config: first member (faf_create_inverse) */
    size_t n, n_bins, stride, n_scales, n_rows;
    int has_lp;
    double *scales, *freqs, *l2, *A, *dual;
    float *psi32, *dual32;
    faf_cwt_lp_report report;
    int inv_kind;
} fafi_cwt_bank;

/* ------------------------------------------------------------------ */
/* config defaults (reference src/faf_cwt.c:27-51)                      */
/* ------------------------------------------------------------------ */

faf_cwt_config faf_cwt_config_init(size_t n) {
    faf_cwt_config c;
    memset(&c, 0, sizeof c);
    c.n = n;
    c.fs = 1.0;
    c.precision = FAF_PREC_FP32;
    c.backend = FAF_BACKEND_AUTO;
    c.wavelet = FAF_CWT_WAVELET_MORSE;
    c.norm = FAF_CWT_NORM_L1;
    c.scale_kind = FAF_CWT_SCALE_GEOMETRIC;
    c.center_kind = FAF_CWT_CENTER_PEAK;
    c.voices = 10;
    c.morlet_mu = 6.0;
    c.morse_gamma = 3.0;
    c.morse_beta = 20.0;     /* time-bandwidth beta*gamma = 60 */
    c.bump_center = 5.0;
    c.bump_width = 1.0;
    c.bandpass_peak = 2.0;
    c.lp_alpha = 0.05;
    c.lp_beta = 0.05;
    c.lp_floor = 1e-4;
    c.flags = FAF_CWT_FLAG_INCLUDE_LOWPASS | FAF_CWT_FLAG_VALIDATE_STRICT;
    return c;
}

/* ------------------------------------------------------------------ */
/* analytic prototypes in the Fourier domain (reference                 */
/* src/faf_cwt_wavelets.c:13-80); all vanish for omega <= 0             */
/* ------------------------------------------------------------------ */

static double meyer_nu(double x) {
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    return x * x * x * x * (35.0 - 84.0 * x + 70.0 * x * x - 20.0 * x * x * x);
}

/* peak normalisation of the generalized Morse prototype: lp = b log(wp) -
 * wp^g with wp = (b/g)^(1/g), the log of its value at the peak.  The last
 * pair (b, g) is kept per thread, since a bank evaluates one prototype
 * hundreds of thousands of times (same bits as computing it every call). */
static double morse_lp(double b, double g) {
    static _Thread_local double kb = -1.0, kg = -1.0, klp;
    if (b != kb || g != kg) {
        const double wp = pow(b / g, 1.0 / g);
        klp = b * log(wp) - pow(wp, g);
        kb = b;
        kg = g;
    }
    return klp;
}

static double proto(double w, const faf_cwt_config *c) {
    if (w <= 0.0) return 0.0;
    switch (c->wavelet) {
    case FAF_CWT_WAVELET_MORLET: {
        double mu = c->morlet_mu;
        return exp(-0.5 * (w - mu) * (w - mu)) - exp(-0.5 * mu * mu) * exp(-0.5 * w * w);
    }
    case FAF_CWT_WAVELET_MORSE: {
        /* peak-normalised generalized Morse, evaluated in the log domain */
        double g = c->morse_gamma, b = c->morse_beta;
        return exp(b * log(w) - pow(w, g) - morse_lp(b, g));
    }
    case FAF_CWT_WAVELET_BUMP: {
        double u = (w - c->bump_center) / c->bump_width;
        if (fabs(u) >= 1.0) return 0.0;
        return exp(1.0 - 1.0 / (1.0 - u * u));
    }
    case FAF_CWT_WAVELET_SHANNON:
        return (w >= M_PI / 2.0 && w <= M_PI) ? 1.0 : 0.0;
    case FAF_CWT_WAVELET_MEYER:
        if (w < 2.0 * M_PI / 3.0) return 0.0;
        if (w <= 4.0 * M_PI / 3.0) return sin(0.5 * M_PI * meyer_nu(3.0 * w / (2.0 * M_PI) - 1.0));
        if (w <= 8.0 * M_PI / 3.0) return cos(0.5 * M_PI * meyer_nu(3.0 * w / (4.0 * M_PI) - 1.0));
        return 0.0;
    default:
        return 0.0;
    }
}

static double peak_omega(const faf_cwt_config *c) {
    switch (c->wavelet) {
    case FAF_CWT_WAVELET_MORLET: return c->morlet_mu;
    case FAF_CWT_WAVELET_MORSE: return pow(c->morse_beta / c->morse_gamma, 1.0 / c->morse_gamma);
    case FAF_CWT_WAVELET_BUMP: return c->bump_center;
    case FAF_CWT_WAVELET_SHANNON: return 0.75 * M_PI;
    case FAF_CWT_WAVELET_MEYER: return 4.0 * M_PI / 3.0;
    default: return 1.0;
    }
}

/* energy centroid of |psi|^2 and the admissibility integral: Riemann sums
 * with step 1e-4 accumulated the same way as reference src/faf_cwt.c:75-98 */
static double energy_omega(const faf_cwt_config *c) {
    double dw = 1e-4, num = 0.0, den = 0.0;
    for (double w = dw; w <= 80.0; w += dw) {
        double e = proto(w, c);
        e *= e;
        num += w * e;
        den += e;
    }
    return den > 0.0 ? num / den : peak_omega(c);
}

static double admissibility(const faf_cwt_config *c, double xi0) {
    double dw = 1e-4, sum = 0.0;
    for (double w = dw; w <= xi0 * 20.0; w += dw) sum += proto(w, c) / w;
    return sum * dw;
}

/* The two Riemann sums above depend on the prototype alone (family and
 * shape parameters), not on n or the scales, and take 1.3M prototype
 * evaluations: a small process-wide table keeps them per prototype. */
typedef struct {
    int wavelet, center_kind;
    double mu, gamma, beta, bc, bw;
    double xi0, adm;
} proto_sums;
static proto_sums g_sums[32];
static int g_nsums;
static pthread_mutex_t g_sums_mu = PTHREAD_MUTEX_INITIALIZER;

static int sums_match(const proto_sums *e, const faf_cwt_config *c) {
    return e->wavelet == (int)c->wavelet && e->center_kind == (int)c->center_kind &&
           e->mu == c->morlet_mu && e->gamma == c->morse_gamma && e->beta == c->morse_beta &&
           e->bc == c->bump_center && e->bw == c->bump_width;
}

/* xi0 (the energy centroid or the peak, as c asks) and the admissibility
 * constant of c's prototype */
static void prototype_sums(const faf_cwt_config *c, double *xi0, double *adm) {
    pthread_mutex_lock(&g_sums_mu);
    for (int i = 0; i < g_nsums; ++i)
        if (sums_match(&g_sums[i], c)) {
            *xi0 = g_sums[i].xi0;
            *adm = g_sums[i].adm;
            pthread_mutex_unlock(&g_sums_mu);
            return;
        }
    pthread_mutex_unlock(&g_sums_mu);
    const double x = c->center_kind == FAF_CWT_CENTER_ENERGY ? energy_omega(c) : peak_omega(c);
    const double a = x > 0.0 && x == x ? admissibility(c, x) : 0.0;
    pthread_mutex_lock(&g_sums_mu);
    proto_sums e = {(int)c->wavelet, (int)c->center_kind, c->morlet_mu, c->morse_gamma,
                    c->morse_beta, c->bump_center, c->bump_width, x, a};
    g_sums[g_nsums < 32 ? g_nsums++ : 31] = e;
    pthread_mutex_unlock(&g_sums_mu);
    *xi0 = x;
    *adm = a;
}

/* rows[j][k] = proto(scales[j] xi[k]) with DC and Nyquist zeroed, for
 * j < J.  Rows are independent, so a large bank fills them on up to eight
 * threads (each row is computed by the same code either way). */
typedef struct {
    const faf_cwt_config *c;
    const double *scales, *xi;
    size_t J, nb, stride, j0, step;
    double *rows;
} proto_job;

static void *proto_rows_run(void *arg) {
    const proto_job *q = arg;
    for (size_t j = q->j0; j < q->J; j += q->step) {
        double *row = q->rows + j * q->stride;
        for (size_t k = 0; k < q->nb; ++k) row[k] = proto(q->scales[j] * q->xi[k], q->c);
        row[0] = 0.0;
        if (q->nb > 1) row[q->nb - 1] = 0.0;
    }
    return NULL;
}

static void proto_rows(const faf_cwt_config *c, const double *scales, size_t J,
                       const double *xi, size_t nb, double *rows, size_t stride) {
    enum { MAXT = 8 };
    size_t nt = J * nb >= 65536 ? (J < MAXT ? J : MAXT) : 1;
    proto_job jobs[MAXT];
    pthread_t th[MAXT];
    int started[MAXT] = {0};
    for (size_t t = 0; t < nt; ++t)
        jobs[t] = (proto_job){c, scales, xi, J, nb, stride, t, nt, rows};
    for (size_t t = 1; t < nt; ++t)
        started[t] = pthread_create(&th[t], NULL, proto_rows_run, &jobs[t]) == 0;
    proto_rows_run(&jobs[0]);
    for (size_t t = 1; t < nt; ++t) {
        if (started[t]) pthread_join(th[t], NULL);
        else proto_rows_run(&jobs[t]);       /* no thread: fill its rows here */
    }
}

static void bank_free(fafi_cwt_bank *b) {
    if (!b) return;
    free(b->scales); free(b->freqs); free(b->l2); free(b->A); free(b->dual);
    free(b->psi32); free(b->dual32);
    free(b);
}

/* ------------------------------------------------------------------ */
/* bank construction (reference src/faf_cwt.c:104-677)                  */
/* ------------------------------------------------------------------ */

static fafi_cwt_bank *bank_build(const faf_cwt_config *cfg) {
    faf_cwt_config c = *cfg;
    if (c.n < 2 || (c.n & 1) || !fafi_is_5_smooth(c.n)) {
        fafi_err("cwt: n=%zu not even 5-smooth", c.n);
        return NULL;
    }
    if (c.precision != FAF_PREC_FP32) {
        fafi_err(c.precision == FAF_PREC_FP64
                     ? "cwt: this build computes in FP32 only"
                     : "cwt: precision %d unsupported", (int)c.precision);
        return NULL;
    }
    if (c.voices == 0) { fafi_err("cwt: voices must be >= 1"); return NULL; }
    if (c.fs <= 0.0) c.fs = 1.0;
    if (c.wavelet == FAF_CWT_WAVELET_MORSE && (c.morse_gamma <= 0.0 || c.morse_beta <= 0.0)) {
        fafi_err("cwt: morse gamma/beta must be > 0");
        return NULL;
    }
    if (c.wavelet == FAF_CWT_WAVELET_MORLET && c.morlet_mu <= 0.0) {
        fafi_err("cwt: morlet mu must be > 0");
        return NULL;
    }
    if (c.wavelet == FAF_CWT_WAVELET_BUMP && c.bump_width <= 0.0) {
        fafi_err("cwt: bump width must be > 0");
        return NULL;
    }
    if (c.flags & FAF_CWT_FLAG_FIXED_LOWPASS) {
        fafi_err("cwt: FAF_CWT_FLAG_FIXED_LOWPASS is reserved");
        return NULL;
    }
    double xi0, adm_c;
    prototype_sums(&c, &xi0, &adm_c);
    if (!(xi0 > 0.0) || xi0 != xi0) {
        fafi_err("cwt: prototype center frequency is not finite");
        return NULL;
    }
    const size_t n = c.n, nb = n / 2 + 1, stride = (nb + 15) & ~(size_t)15;
    const double nyq = c.fs / 2.0;
    if (c.fmax <= 0.0) c.fmax = 0.45 * nyq;
    if (c.fmin <= 0.0) c.fmin = 16.0 * c.fs / (double)n;
    if (c.fmin >= c.fmax) { fafi_err("cwt: fmin >= fmax"); return NULL; }
    if (c.fmax >= nyq) { fafi_err("cwt: fmax at or above Nyquist"); return NULL; }

    size_t J;
    if (c.scale_kind == FAF_CWT_SCALE_GEOMETRIC) {
        double oct = log(c.fmax / c.fmin) / LN2;
        if (!(oct >= 0.0)) { fafi_err("cwt: fmin/fmax do not form a valid octave span"); return NULL; }
        J = (size_t)floor((double)c.voices * oct + 1.0 + 1e-12);
        if (J < 1) J = 1;
    } else {
        J = c.voices < 2 ? 2 : (size_t)c.voices;
    }
    if (J > MAX_SCALES) { fafi_err("cwt: too many scales (%zu > %d)", J, MAX_SCALES); return NULL; }

    fafi_cwt_bank *b = calloc(1, sizeof *b);
    const int lp = (c.flags & FAF_CWT_FLAG_INCLUDE_LOWPASS) ? 1 : 0;
    const size_t rows = J + (size_t)lp, off = (size_t)lp;
    b->cfg = c;
    b->n = n; b->n_bins = nb; b->stride = stride; b->n_scales = J; b->n_rows = rows;
    b->has_lp = lp;
    b->scales = calloc(J, sizeof(double));
    b->freqs = calloc(J, sizeof(double));
    b->l2 = calloc(rows * stride, sizeof(double));
    b->A = calloc(nb, sizeof(double));
    b->dual = calloc(rows * stride, sizeof(double));
    b->psi32 = calloc(rows * stride, sizeof(float));
    b->dual32 = calloc(rows * stride, sizeof(float));
    double *xi = calloc(nb, sizeof(double));
    for (size_t k = 0; k < nb; ++k) xi[k] = 2.0 * M_PI * (double)k / (double)n;

    for (size_t j = 0; j < J; ++j) {
        double f = c.scale_kind == FAF_CWT_SCALE_GEOMETRIC
                       ? c.fmin * pow(2.0, (double)j / (double)c.voices)
                       : c.fmin + (double)j * (c.fmax - c.fmin) / (double)(J > 1 ? J - 1 : 1);
        b->freqs[j] = f;
        b->scales[j] = xi0 / (2.0 * M_PI * f / c.fs);
    }
    /* L2 shadow: rows coarse -> fine after the lowpass slot */
    proto_rows(&c, b->scales, J, xi, nb, b->l2 + off * stride, stride);
    if (c.scale_kind == FAF_CWT_SCALE_GEOMETRIC) {
        /* rescale so the plateau of sum_j psi^2 (bins within 5% of its peak) is 1 */
        double amax = 0.0, mean = 0.0;
        double *ap = calloc(nb, sizeof(double));
        size_t cnt = 0;
        for (size_t k = 0; k < nb; ++k) {
            double s = 0.0;
            for (size_t j = 0; j < J; ++j) {
                double v = b->l2[(off + j) * stride + k];
                s += v * v;
            }
            ap[k] = s;
            if (s > amax) amax = s;
        }
        for (size_t k = 0; k < nb; ++k)
            if (ap[k] >= 0.95 * amax) { mean += ap[k]; ++cnt; }
        if (cnt) mean /= (double)cnt;
        if (mean <= 0.0) mean = 1.0;
        double r = 1.0 / sqrt(mean);
        for (size_t j = 0; j < J; ++j)
            for (size_t k = 0; k < nb; ++k) b->l2[(off + j) * stride + k] *= r;
        free(ap);
    }
    if (lp) {
        for (size_t k = 0; k < nb; ++k) {
            double s = 0.0;
            for (size_t j = 0; j < J; ++j) {
                double v = b->l2[(off + j) * stride + k];
                s += v * v;
            }
            b->l2[k] = 1.0 - s > 0.0 ? sqrt(1.0 - s) : 0.0;
        }
    }
    for (size_t k = 0; k < nb; ++k) {
        double s = 0.0;
        for (size_t r = 0; r < rows; ++r) s += b->l2[r * stride + k] * b->l2[r * stride + k];
        b->A[k] = s;
    }
    /* live FP32 analysis table */
    for (size_t r = 0; r < rows; ++r) {
        float *live = b->psi32 + r * stride;
        const double *l2 = b->l2 + r * stride;
        if ((int)r < lp) {
            for (size_t k = 0; k < nb; ++k) live[k] = (float)l2[k];
            continue;
        }
        double s = b->scales[r - off];
        if (c.norm == FAF_CWT_NORM_L2) {
            for (size_t k = 0; k < nb; ++k) live[k] = (float)l2[k];
        } else {
            float peak = 0.0f;
            for (size_t k = 0; k < nb; ++k) {
                live[k] = (float)proto(s * xi[k], &c);
                if (live[k] > peak) peak = live[k];
            }
            live[0] = 0.0f;
            if (c.norm == FAF_CWT_NORM_BANDPASS && peak > 0.0f) {
                float sc = (float)(c.bandpass_peak / (double)peak);
                for (size_t k = 0; k < nb; ++k) live[k] *= sc;
            }
        }
        if (nb > 1) live[nb - 1] = 0.0f;
    }
    /* dual frame from the live table: dual = psi / (sum_r psi^2 + eps) */
    for (size_t k = 0; k < nb; ++k) {
        double s = 0.0;
        for (size_t r = 0; r < rows; ++r) {
            double v = b->psi32[r * stride + k];
            s += v * v;
        }
        for (size_t r = 0; r < rows; ++r) {
            double v = b->psi32[r * stride + k];
            double d = s > 0.0 ? v / (s + 1e-6) : 0.0;
            b->dual[r * stride + k] = d;
            b->dual32[r * stride + k] = (float)d;
        }
    }
    /* Littlewood-Paley report */
    faf_cwt_lp_report *rep = &b->report;
    rep->n = n; rep->n_bins = nb; rep->n_scales = J; rep->has_lowpass = lp;
    rep->alpha = c.lp_alpha; rep->beta = c.lp_beta; rep->A = b->A;
    size_t klo = 1;
    if (!lp) {
        const double *lo = b->l2 + off * stride;
        double pk = 0.0;
        for (size_t k = 0; k < nb; ++k) if (lo[k] > pk) pk = lo[k];
        for (klo = 1; klo < nb; ++klo) if (lo[klo] >= pk * c.lp_floor) break;
    }
    size_t khi;
    {
        const double *hi = b->l2 + (off + J - 1) * stride;
        double pk = 0.0;
        for (size_t k = 0; k < nb; ++k) if (hi[k] > pk) pk = hi[k];
        for (khi = nb - 2; khi > klo; --khi) if (hi[khi] >= pk * c.lp_floor) break;
    }
    rep->k_lo = klo; rep->k_hi = khi;
    rep->hole_bins_dc = klo;
    rep->hole_bins_nyq = (nb - 1) - khi;
    double mx = 0.0, sum = 0.0, amin = 1e30, amx = -1e30;
    size_t cnt = 0;
    for (size_t k = klo; k <= khi; ++k) {
        double a = b->A[k], d = fabs(a - 1.0);
        if (d > mx) mx = d;
        sum += d;
        if (a < amin) amin = a;
        if (a > amx) amx = a;
        ++cnt;
    }
    if (!cnt) { amin = 0.0; amx = 0.0; }
    rep->max_abs_dev = mx;
    rep->mean_abs_dev = cnt ? sum / (double)cnt : 0.0;
    rep->min_A = amin; rep->max_A = amx;
    rep->frame_cond = amin > 0.0 ? amx / amin : INFINITY;
    double mdc = 0.0;
    for (size_t j = 0; j < J; ++j) {
        double v = fabs(b->l2[(off + j) * stride]);
        if (v > mdc) mdc = v;
    }
    rep->max_dc_wavelet = mdc;
    rep->admissibility_C = adm_c;
    rep->passed = (mx <= c.lp_alpha && rep->mean_abs_dev <= c.lp_beta && cnt > 0) ? 1 : 0;
    free(xi);
    const int strict = (c.flags & FAF_CWT_FLAG_VALIDATE_STRICT) &&
                       !(c.flags & FAF_CWT_FLAG_ALLOW_UNTILED);
    if (strict && !rep->passed) {
        fafi_err("cwt: LP bound violated max|A-1|=%g on [%zu,%zu]", mx, klo, khi);
        bank_free(b);
        return NULL;
    }
    /* wrap check: energy of the coarsest L2 kernel near t = n/2, from an
     * FP64 inverse DFT of its (real, even) spectrum */
    {
        double *re = calloc(n, sizeof(double)), *im = calloc(n, sizeof(double));
        const double *psi = b->l2 + off * stride;
        for (size_t k = 0; k < nb; ++k) {
            re[k] = psi[k];
            if (k && k < n - k) re[n - k] = psi[k];
        }
        fafi_fft64(n, 1, re, im);
        double tot = 0.0, wrap = 0.0;
        size_t half = n / 2, win = n / 32 ? n / 32 : 1;
        for (size_t i = 0; i < n; ++i) tot += re[i] * re[i];
        for (size_t i = half - win; i < half + win && i < n; ++i) wrap += re[i] * re[i];
        rep->max_wrap_energy = tot > 0.0 ? wrap / tot : 0.0;
        free(re); free(im);
        if (strict && rep->max_wrap_energy > 1e-3) {
            fafi_err("cwt: coarsest scale wraps (energy=%g)", rep->max_wrap_energy);
            bank_free(b);
            return NULL;
        }
    }
    return b;
}

/* ------------------------------------------------------------------ */
/* transforms                                                          */
/* ------------------------------------------------------------------ */

static void *upload_f32(const float *h, size_t count) {
    void *d = fafk_malloc(count * sizeof(float));
    if (!d || fafk_h2d(d, h, count * sizeof(float), NULL) || fafk_sync(NULL)) {
        fafi_err("%s", fafk_last_error());
        fafk_free(d);
        return NULL;
    }
    return d;
}

/* the rfft of a segment and one length-n inverse row side by side in LDS */
static int cwt_fits_one_workgroup(size_t n) {
    size_t nc = n / 2;
    if (!fafi_radix_fits(nc, 0) || !fafi_radix_fits(n, 0)) return 0;
    size_t planes = (nc + (nc >> 4) + 1) + (n + (n >> 4) + 1);
    return planes * 8 <= FAFD_MAX_LDS_BYTES;
}

/* sub-plan of a long-segment CWT: interleaved device buffers, given scale */
static fafh_plan *cwt_sub(size_t n, int real, int inv, float scale) {
    fafh_plan *q = fafi_plan_new();
    if (!q) { fafi_err("cwt: alloc failed"); return NULL; }
    int rc = real ? fafi_plan_real(q, n, inv) : fafi_plan_c2c(q, n, inv, 0);
    if (rc) { fafi_plan_free(q); return NULL; }
    q->interleaved = 1;
    q->scale = scale;
    return q;
}

int fafi_cwt_build(fafh_plan *p, const faf_cwt_config *cfg, int inverse_kind) {
    fafi_cwt_bank *b = bank_build(cfg);
    if (!b) return -1;
    p->bank = b;
    b->inv_kind = inverse_kind;
    size_t n = b->n, nc = n / 2;
    p->n = n;
    p->psi_stride = (long)b->stride;
    p->rows_total = (int)b->n_rows;
    p->post = fafi_kiss_table(nc, 0);
    p->pre = fafi_kiss_table(nc, 1);
    if (!p->post || !p->pre) return -1;
    if (n % 2) { fafi_err("cwt: length %zu must be even", n); return -1; }
    if (inverse_kind < 0) {
        p->kind = PK_CWT;
        p->scale = (float)(1.0 / (double)n);
        p->psi_dev = upload_f32(b->psi32, b->n_rows * b->stride);
        if (!p->psi_dev) return -1;
        if (cwt_fits_one_workgroup(n)) {
            /* one kernel: rfft in LDS, then every row's length-n inverse */
            if (fafi_make_fft(nc, 0, 0, NULL, 0, &p->a) ||
                fafi_make_fft(n, 1, 0, NULL, 0, &p->full_inv))
                return -1;
            p->ct = fafi_ct_enabled() && fafk_ct_supports(5, (int)n);
            return 0;
        }
        /* longer segments: rfft(n), spread the filtered rows, then one
         * batched inverse over all rows (four-step where n needs it) */
        p->sub_r2c = cwt_sub(n, 1, 0, 1.0f);
        p->sub_c2c = p->sub_r2c ? cwt_sub(n, 0, 1, p->scale) : NULL;
        p->sub_c2r = p->sub_c2c ? cwt_sub(n, 1, 1, p->scale) : NULL;
        return p->sub_c2r ? 0 : -1;
    }
    p->kind = PK_ICWT;
    p->scale = (float)(1.0 / (double)n);
    p->dual_dev = upload_f32(b->dual32, b->n_rows * b->stride);
    if (!p->dual_dev) return -1;
    if (inverse_kind == FAF_CWT_INV_L1) return 0;
    /* dual inverse: one kernel when a row spectrum and the accumulator fit
     * one workgroup's LDS side by side */
    if (fafi_radix_fits(nc, 0) && 2 * (nc + (nc >> 4) + 1) * 8 <= FAFD_MAX_LDS_BYTES &&
        !fafi_make_fft(nc, 0, 0, NULL, 0, &p->a))
        return fafi_make_fft(nc, 1, 0, NULL, 0, &p->b);
    /* longer segments: one batched rfft over every row, the dual
     * accumulation across rows, then a batched irfft (reference
     * src/faf_cwt.c:958-980 does the same per row) */
    p->sub_r2c = cwt_sub(n, 1, 0, 1.0f);
    p->sub_c2r = p->sub_r2c ? cwt_sub(n, 1, 1, p->scale) : NULL;
    return p->sub_c2r ? 0 : -1;
}

void fafi_cwt_free(fafh_plan *p) {
    if (!p) return;
    bank_free(p->bank);
    p->bank = NULL;
    fafk_free(p->psi_dev);
    fafk_free(p->dual_dev);
    p->psi_dev = p->dual_dev = NULL;
}

/* long-segment CWT: segments per chunk, so the filtered rows stay <= 256 MB */
static long cwt_chunk(const fafh_plan *p, long count) {
    size_t per = (size_t)p->rows_total * (p->n + 2) * 8;
    long c = (long)(((size_t)256 << 20) / per);
    if (c < 1) c = 1;
    return c < count ? c : count;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t fafi_cwt_ws_bytes(const fafh_plan *p, long count, int complex_out) {
    if (!p->sub_r2c || count <= 0) return 0;
    const size_t nc = p->n / 2, rows = (size_t)p->rows_total;
    const long c = cwt_chunk(p, count);
    size_t z = align256((size_t)c * (nc + 1) * 8);
    size_t sp = complex_out ? 0 : align256((size_t)c * rows * (nc + 1) * 8);
    size_t s1 = fafi_run_ws_bytes(p->sub_r2c, c);
    size_t s2 = fafi_run_ws_bytes(complex_out ? p->sub_c2c : p->sub_c2r, c * (long)rows);
    return z + sp + (s1 > s2 ? s1 : s2);
}

/* run a sub-plan on lent workspace */
#define LENT(q, w, wb, call)              \
    ((q)->ws_ext = (w), (q)->ws_ext_bytes = (wb), rc = (call), (q)->ws_ext = NULL, \
     (q)->ws_ext_bytes = 0, rc)

static int run_cwt_long(fafh_plan *p, const float *x, long idist, float *out, long odist,
                        long count, int complex_out, void *stream) {
    const long n = (long)p->n, nc = n / 2, rows = p->rows_total;
    const long c = cwt_chunk(p, count);
    const size_t need = fafi_cwt_ws_bytes(p, count, complex_out);
    char *ws = fafi_workspace(p, need);
    if (!ws) return -1;
    float *Z = (float *)ws;
    size_t zb = align256((size_t)c * (size_t)(nc + 1) * 8);
    size_t sb = complex_out ? 0 : align256((size_t)c * (size_t)rows * (size_t)(nc + 1) * 8);
    float *S = (float *)(ws + zb);
    char *sub = ws + zb + sb;
    const size_t subb = need - zb - sb;
    int rc = 0;
    for (long b0 = 0; b0 < count; b0 += c) {
        const long cc = count - b0 < c ? count - b0 : c;
        if (LENT(p->sub_r2c, sub, subb,
                 fafi_run_r2c(p->sub_r2c, x + b0 * idist, idist, Z, NULL, nc + 1, cc, stream)))
            return -1;
        if (complex_out) {
            /* analytic rows: doubled one-sided spectrum, inverse C2C in place */
            float *o = out + 2 * b0 * odist;
            if (fafk_cwt_spread(Z, nc + 1, p->psi_dev, p->psi_stride, (int)nc, 1, o, n, odist,
                                (int)rows, cc, stream)) {
                fafi_err("%s", fafk_last_error());
                return -1;
            }
            if (odist == rows * n) {
                if (LENT(p->sub_c2c, sub, subb,
                         fafi_run_c2c(p->sub_c2c, o, NULL, n, o, NULL, n, cc * rows, stream)))
                    return -1;
            } else {
                for (long s = 0; s < cc; ++s) {
                    float *os = o + 2 * s * odist;
                    if (LENT(p->sub_c2c, sub, subb,
                             fafi_run_c2c(p->sub_c2c, os, NULL, n, os, NULL, n, rows, stream)))
                        return -1;
                }
            }
        } else {
            /* real rows = irfft(X psi_r): spread into S, batched C2R */
            if (fafk_cwt_spread(Z, nc + 1, p->psi_dev, p->psi_stride, (int)nc, 0, S, nc + 1,
                                rows * (nc + 1), (int)rows, cc, stream)) {
                fafi_err("%s", fafk_last_error());
                return -1;
            }
            float *o = out + b0 * odist;
            if (odist == rows * n) {
                if (LENT(p->sub_c2r, sub, subb,
                         fafi_run_c2r(p->sub_c2r, S, NULL, nc + 1, o, n, cc * rows, stream)))
                    return -1;
            } else {
                for (long s = 0; s < cc; ++s)
                    if (LENT(p->sub_c2r, sub, subb,
                             fafi_run_c2r(p->sub_c2r, S + 2 * s * rows * (nc + 1), NULL, nc + 1,
                                          o + s * odist, n, rows, stream)))
                        return -1;
            }
        }
    }
    return 0;
}

int fafi_run_cwt(fafh_plan *p, const float *x, long idist, float *out, long odist,
                 long count, int complex_out, void *stream) {
    if (p->sub_r2c) return run_cwt_long(p, x, idist, out, odist, count, complex_out, stream);
    int rows = p->rows_total;
    /* enough workgroups to fill the device: split the rows of a segment
     * across workgroups when there are few segments; and at most 11 rows per
     * workgroup, so the last round of workgroups is short (each group
     * recomputes the segment's rfft, a small cost next to 11 inverse rows).
     * Measured on 1024 segments of 4096 with 65 rows: 1 group 0.70 ms,
     * 3 groups 0.60, 6 groups (11 rows) 0.56, 11 groups 0.61. */
    long want = 2 * (long)fafi_cus();
    int groups = (int)((want + count - 1) / count);
    if (groups < (rows + 10) / 11) groups = (rows + 10) / 11;
    if (groups < 1) groups = 1;
    if (groups > rows) groups = rows;
    int rpw = (rows + groups - 1) / groups;
    int rc;
    if (p->ct) {
        rc = fafk_ct_cwt(&p->a, &p->full_inv, !complex_out, p->post, p->psi_dev, p->psi_stride,
                         rows, rpw, x, idist, out, odist, count, p->scale, stream);
        if (rc) fafi_err("%s", fafk_ct_last_error());
        return rc;
    }
    rc = fafk_cwt(&p->a, &p->full_inv, !complex_out, p->post, p->psi_dev, p->psi_stride, rows,
                  rpw, x, idist, out, odist, count, p->scale, stream);
    if (rc) fafi_err("%s", fafk_last_error());
    return rc;
}

/* long-segment dual inverse: segments per chunk, so the row spectra stay
 * <= 256 MB; workspace = spectra, accumulators, then the sub-plans' own */
static long icwt_chunk(const fafh_plan *p, long count) { return cwt_chunk(p, count); }

static size_t icwt_ws_bytes(const fafh_plan *p, long count) {
    const size_t nb = p->n / 2 + 1, rows = (size_t)p->rows_total;
    const long c = icwt_chunk(p, count);
    size_t yb = align256((size_t)c * rows * nb * 8), ab = align256((size_t)c * nb * 8);
    size_t s1 = fafi_run_ws_bytes(p->sub_r2c, c * (long)rows);
    size_t s2 = fafi_run_ws_bytes(p->sub_c2r, c);
    return yb + ab + (s1 > s2 ? s1 : s2);
}

static int run_icwt_long(fafh_plan *p, const float *w, long idist, float *y, long odist,
                         long count, void *stream) {
    const long n = (long)p->n, nb = n / 2 + 1, rows = p->rows_total;
    const long c = icwt_chunk(p, count);
    const size_t need = icwt_ws_bytes(p, count);
    char *ws = fafi_workspace(p, need);
    if (!ws) return -1;
    float *Y = (float *)ws;
    const size_t yb = align256((size_t)c * (size_t)rows * (size_t)nb * 8);
    float *A = (float *)(ws + yb);
    const size_t ab = align256((size_t)c * (size_t)nb * 8);
    char *sub = ws + yb + ab;
    const size_t subb = need - yb - ab;
    int rc = 0;
    for (long s0 = 0; s0 < count; s0 += c) {
        const long cc = count - s0 < c ? count - s0 : c;
        if (idist == rows * n) {       /* the chunk's rows are one contiguous batch */
            if (LENT(p->sub_r2c, sub, subb,
                     fafi_run_r2c(p->sub_r2c, w + s0 * idist, n, Y, NULL, nb, cc * rows,
                                  stream)))
                return -1;
        } else {
            for (long s = 0; s < cc; ++s)
                if (LENT(p->sub_r2c, sub, subb,
                         fafi_run_r2c(p->sub_r2c, w + (s0 + s) * idist, n,
                                      Y + 2 * s * rows * nb, NULL, nb, rows, stream)))
                    return -1;
        }
        if (fafk_icwt_mulacc(Y, nb, (int)rows, p->dual_dev, p->psi_stride, A, cc, stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
        if (LENT(p->sub_c2r, sub, subb,
                 fafi_run_c2r(p->sub_c2r, A, NULL, nb, y + s0 * odist, odist, cc, stream)))
            return -1;
    }
    return rc;
}

int fafi_run_icwt(fafh_plan *p, const float *w, long idist, float *y, long odist, long count,
                  void *stream) {
    fafi_cwt_bank *b = p->bank;
    if (b->inv_kind == FAF_CWT_INV_L1) {
        /* reference src/faf_cwt.c:878-915: weight ln2/V, C from the LP report
         * (1 when it is not positive), both rounded to FP32 as there */
        double c = b->report.admissibility_C;
        if (c <= 0.0) c = 1.0;
        int rc = fafk_icwt_l1(w, idist, y, odist, (long)b->n, (int)b->n_scales, b->has_lp,
                              (float)(LN2 / (double)b->cfg.voices), (float)c, count, stream);
        if (rc) fafi_err("%s", fafk_last_error());
        return rc;
    }
    if (p->sub_r2c) return run_icwt_long(p, w, idist, y, odist, count, stream);
    int rc = fafk_icwt(&p->a, &p->b, p->post, p->pre, p->dual_dev, p->psi_stride,
                       p->rows_total, w, idist, y, odist, count, p->scale, stream);
    if (rc) fafi_err("%s", fafk_last_error());
    return rc;
}

static faf_transform *cwt_transform(const faf_cwt_config *cfg, int inverse_kind) {
    if (!cfg) { fafi_err("cwt: null argument"); return NULL; }
    if (fafi_create_ready()) return NULL;
    faf_transform *t = calloc(1, sizeof(faf_transform));
    fafh_plan *p = fafi_plan_new();
    if (!t || !p) { free(t); free(p); fafi_err("cwt: alloc failed"); return NULL; }
    t->amd = p;
    if (fafi_cwt_build(p, cfg, inverse_kind)) {
        fafi_plan_free(p);
        free(t);
        return NULL;
    }
    fafi_cwt_bank *b = p->bank;
    t->type = inverse_kind < 0 ? FAF_TRANSFORM_CWT : FAF_TRANSFORM_ICWT;
    t->n = b->n;
    t->precision = b->cfg.precision;
    t->cfg = faf_config_init(b->n);
    t->cfg.precision = b->cfg.precision;
    t->cfg.layout = FAF_LAYOUT_REAL;
    t->cfg.norm = FAF_NORM_NONE;
    t->cfg.backend = cfg->backend;
    if (inverse_kind >= 0) t->cfg.dir = FAF_DIR_INVERSE;
    t->user_aux_n = b->n_rows;
    return t;
}

faf_transform *faf_create_cwt(const faf_cwt_config *cfg) { return cwt_transform(cfg, -1); }

faf_transform *faf_create_icwt(const faf_cwt_config *cfg, faf_cwt_inverse_kind kind) {
    if (!cfg) { fafi_err("cwt: null argument"); return NULL; }
    if (kind == FAF_CWT_INV_L1 && cfg->norm != FAF_CWT_NORM_L1) {
        fafi_err("cwt: L1 inverse requires L1 bank");
        return NULL;
    }
    return cwt_transform(cfg, (int)kind);
}

/* ------------------------------------------------------------------ */
/* queries (reference src/faf_cwt.c:995-1079)                           */
/* ------------------------------------------------------------------ */

static const fafi_cwt_bank *bank_of(const faf_transform *t) {
    if (!t || (t->type != FAF_TRANSFORM_CWT && t->type != FAF_TRANSFORM_ICWT) || !t->amd)
        return NULL;
    return ((const fafh_plan *)t->amd)->bank;
}

const faf_cwt_lp_report *faf_cwt_bank_report(const faf_transform *t) {
    const fafi_cwt_bank *b = bank_of(t);
    return b ? &b->report : NULL;
}
size_t faf_cwt_n_scales(const faf_transform *t) { const fafi_cwt_bank *b = bank_of(t); return b ? b->n_scales : 0; }
size_t faf_cwt_n_rows(const faf_transform *t) { const fafi_cwt_bank *b = bank_of(t); return b ? b->n_rows : 0; }
size_t faf_cwt_n_bins(const faf_transform *t) { const fafi_cwt_bank *b = bank_of(t); return b ? b->n_bins : 0; }
int faf_cwt_has_lowpass(const faf_transform *t) { const fafi_cwt_bank *b = bank_of(t); return b ? b->has_lp : 0; }

int faf_cwt_freqs(const faf_transform *t, double *hz, size_t cap) {
    const fafi_cwt_bank *b = bank_of(t);
    if (!b || !hz) return -1;
    size_t m = b->n_scales < cap ? b->n_scales : cap;
    memcpy(hz, b->freqs, m * sizeof(double));
    return (int)m;
}

int faf_cwt_scales(const faf_transform *t, double *s, size_t cap) {
    const fafi_cwt_bank *b = bank_of(t);
    if (!b || !s) return -1;
    size_t m = b->n_scales < cap ? b->n_scales : cap;
    memcpy(s, b->scales, m * sizeof(double));
    return (int)m;
}

void faf_cwt_report_fprint(FILE *fp, const faf_cwt_lp_report *r) {
    if (!fp || !r) return;
    fprintf(fp, "CWT LP Report:\n");
    fprintf(fp, "  n=%zu  n_bins=%zu  n_scales=%zu  lowpass=%d\n", r->n, r->n_bins,
            r->n_scales, r->has_lowpass);
    fprintf(fp, "  passed=%d  max|A-1|=%.6g  mean|A-1|=%.6g\n", r->passed, r->max_abs_dev,
            r->mean_abs_dev);
    fprintf(fp, "  A range: [%.6g, %.6g]  cond=%.4g\n", r->min_A, r->max_A, r->frame_cond);
    fprintf(fp, "  certified band: [%zu, %zu]\n", r->k_lo, r->k_hi);
    fprintf(fp, "  holes: DC=%zu bins, Nyquist=%zu bins\n", r->hole_bins_dc, r->hole_bins_nyq);
    fprintf(fp, "  max DC wavelet=%.3e  wrap energy=%.3e\n", r->max_dc_wavelet,
            r->max_wrap_energy);
    fprintf(fp, "  admissibility C=%.6g\n", r->admissibility_C);
}

const float *faf_cwt_psi_f32(const faf_transform *t, size_t idx) {
    const fafi_cwt_bank *b = bank_of(t);
    if (!b) return NULL;
    size_t row = idx + (size_t)b->has_lp;
    return row < b->n_rows ? b->psi32 + row * b->stride : NULL;
}

const double *faf_cwt_psi_f64(const faf_transform *t, size_t idx) {
    (void)t; (void)idx;
    return NULL;   /* FP64 banks are not built by this FP32 device library */
}

const float *faf_cwt_phi_f32(const faf_transform *t) {
    const fafi_cwt_bank *b = bank_of(t);
    return (b && b->has_lp) ? b->psi32 : NULL;
}

const double *faf_cwt_phi_f64(const faf_transform *t) { (void)t; return NULL; }

/* dual table accessor for tests and callers that build their own inverse */
const float *faf_amd_cwt_dual_f32(const faf_transform *t, size_t row) {
    const fafi_cwt_bank *b = bank_of(t);
    return (b && row < b->n_rows) ? b->dual32 + row * b->stride : NULL;
}
