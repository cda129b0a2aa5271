/*
 * faf_api.c -- the drop-in C API (C11): config, errors, create, execute,
 * destroy, size policy and edge helpers.  Entry points and their error
 * behaviour follow the reference's include/fastandfourier.h; the work of
 * faf_execute runs in the gfx950 kernels behind faf_kern.h.
 */
#define _GNU_SOURCE
#include "faf_internal.h"
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* errors (reference src/faf_core.c:23, 70-90: thread-local 256 bytes)  */
/* ------------------------------------------------------------------ */

static _Thread_local char g_err[256];

const char *faf_get_error(void) { return g_err; }
void faf_clear_error(void) { g_err[0] = '\0'; }

void faf_set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

void fafi_err(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

/* ------------------------------------------------------------------ */
/* runtime state                                                       */
/* ------------------------------------------------------------------ */

static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_rt_rc = -1;
static char g_rt_msg[256];
static int g_cus = 256;

static void rt_init(void) {
    g_rt_rc = fafk_init();
    if (!g_rt_rc) g_rt_rc = fafk_ct_init();
    if (!g_rt_rc) g_rt_rc = fafk_rs_init();
    if (!g_rt_rc) g_rt_rc = fafk_rg_init();
    if (!g_rt_rc) g_rt_rc = fafk64_init();
    if (g_rt_rc) {
        snprintf(g_rt_msg, sizeof g_rt_msg, "%s", fafk_last_error());
        return;
    }
    int cus = 0;
    if (fafk_device_info(&cus, NULL, NULL) == 0 && cus > 0) g_cus = cus;
}

int fafi_runtime_ready(void) {
    pthread_once(&g_once, rt_init);
    if (g_rt_rc) {
        fafi_err("GPU runtime unavailable: %s", g_rt_msg);
        return -1;
    }
    return 0;
}

size_t fafi_cus(void) { return (size_t)g_cus; }

static _Thread_local void *t_stream;
static _Thread_local int t_async;
static _Thread_local int t_timing;
static _Thread_local float t_last_ms = -1.0f;

void *fafi_stream(void) { return t_stream; }
int fafi_sync_mode(void) { return !t_async; }

int faf_amd_set_stream(void *s) { t_stream = s; return 0; }
void *faf_amd_get_stream(void) { return t_stream; }
int faf_amd_set_sync(int on) { t_async = !on; return 0; }
int faf_amd_set_timing(int on) { t_timing = on; return 0; }

static _Thread_local int t_no_ct;
int faf_amd_set_specialised(int on) { t_no_ct = !on; return 0; }
int fafi_ct_enabled(void) { return !t_no_ct; }
float faf_amd_last_kernel_ms(void) { return t_last_ms; }

/* ---- graph capture of executes (thread-local) ---- */
static _Thread_local int t_capture;        /* 1 while faf_graph_begin is open */
static _Thread_local int t_cap_failed;     /* an execute failed inside the capture */
static _Thread_local void *t_cap_prev;     /* caller's stream before the capture */
static _Thread_local int t_cap_prev_async;
static _Thread_local void *t_cap_own;      /* stream made for a capture on the default stream */

struct faf_graph {
    void *exec;
    int nodes;
};

int fafi_capturing(void) { return t_capture; }

/* 1 while faf_execute_batch_noalloc runs on this thread */
static _Thread_local int t_noalloc;
int fafi_noalloc(void) { return t_noalloc; }

/* create-time entry: plans upload their tables synchronously, which a
 * capture cannot hold */
int fafi_create_ready(void) {
    if (t_capture) {
        fafi_err("plans cannot be created while a graph capture is open");
        return -1;
    }
    return fafi_runtime_ready();
}

int faf_graph_begin(void) {
    if (fafi_runtime_ready()) return -1;
    if (t_capture) { fafi_err("faf_graph_begin: a capture is already open"); return -1; }
    void *s = t_stream;
    if (!s) {
        /* the legacy default stream cannot be captured; capture on a private one */
        t_cap_own = fafk_stream_create();
        if (!t_cap_own) { fafi_err("faf_graph_begin: no stream"); return -1; }
        s = t_cap_own;
    }
    if (fafk_capture_begin(s)) {
        fafi_err("%s", fafk_last_error());
        fafk_stream_destroy(t_cap_own);
        t_cap_own = NULL;
        return -1;
    }
    t_cap_prev = t_stream;
    t_cap_prev_async = t_async;
    t_stream = s;
    t_async = 1;
    t_capture = 1;
    t_cap_failed = 0;
    return 0;
}

faf_graph *faf_graph_end(void) {
    if (!t_capture) { fafi_err("faf_graph_end: no capture is open"); return NULL; }
    int nodes = 0;
    void *x = fafk_capture_end(t_stream, &nodes);
    t_stream = t_cap_prev;
    t_async = t_cap_prev_async;
    t_capture = 0;
    fafk_stream_destroy(t_cap_own);
    t_cap_own = NULL;
    if (!x) { fafi_err("faf_graph_end: %s", fafk_last_error()); return NULL; }
    if (t_cap_failed) {
        fafk_graph_destroy(x);
        fafi_err("faf_graph_end: an execute failed inside the capture");
        return NULL;
    }
    faf_graph *g = calloc(1, sizeof *g);
    if (!g) { fafk_graph_destroy(x); fafi_err("faf_graph_end: out of memory"); return NULL; }
    g->exec = x;
    g->nodes = nodes;
    return g;
}

int faf_graph_launch(const faf_graph *g) {
    if (!g) { fafi_err("faf_graph_launch: null graph"); return -1; }
    if (t_capture) { fafi_err("faf_graph_launch: inside an open capture"); return -1; }
    if (fafk_graph_launch(g->exec, t_stream) || (!t_async && fafk_sync(t_stream))) {
        fafi_err("%s", fafk_last_error());
        return -1;
    }
    return 0;
}

int faf_graph_nodes(const faf_graph *g) { return g ? g->nodes : 0; }

void faf_graph_destroy(faf_graph *g) {
    if (!g) return;
    fafk_graph_destroy(g->exec);
    free(g);
}

int faf_amd_synchronize(void) {
    if (t_capture) { fafi_err("faf_amd_synchronize: a graph capture is open"); return -1; }
    if (fafi_runtime_ready()) return -1;
    if (fafk_sync(t_stream)) { fafi_err("%s", fafk_last_error()); return -1; }
    return 0;
}

int faf_amd_device_ready(void) { return fafi_runtime_ready(); }

int faf_amd_copy_ceiling_ex(void *dst, const void *src, size_t bytes, int grid,
                            int streaming) {
    if (!dst || !src) { fafi_err("copy: null buffer"); return -1; }
    if (fafi_runtime_ready()) return -1;
    if (grid <= 0) grid = (int)(8 * fafi_cus());
    void *stream = fafi_stream();
    int rc = fafk_copy16(dst, src, bytes, grid, streaming, stream);
    if (!rc && fafi_sync_mode()) rc = fafk_sync(stream);
    if (rc) fafi_err("%s", fafk_last_error());
    return rc;
}

int faf_amd_copy_ceiling(void *dst, const void *src, size_t bytes, int grid) {
    return faf_amd_copy_ceiling_ex(dst, src, bytes, grid, 0);
}

/* ------------------------------------------------------------------ */
/* lifecycle and names (reference src/faf_core.c:27-68, 222-307)        */
/* ------------------------------------------------------------------ */

int faf_init(void) {
    faf_clear_error();
    return 0;
}

void faf_cleanup(void) {}

const char *faf_version(void) { return FASTANDFOURIER_VERSION_STRING; }
const char *faf_arch_name(void) { return "gfx950"; }
size_t faf_get_alignment(void) { return FAF_ALIGN; }

size_t faf_precision_size(faf_precision p) {
    switch (p) {
    case FAF_PREC_FP8: case FAF_PREC_INT8: return 1;
    case FAF_PREC_FP16: case FAF_PREC_BF16: case FAF_PREC_INT16: return 2;
    case FAF_PREC_FP64: return 8;
    default: return 4;
    }
}

const char *faf_precision_name(faf_precision p) {
    static const char *nm[] = {"fp8", "fp16", "bf16", "fp32", "fp64", "int8", "int16", "int32"};
    return (unsigned)p < 8 ? nm[p] : "unknown";
}

const char *faf_layout_name(faf_layout l) {
    static const char *nm[] = {"default", "split", "hermitian", "real", "interleaved"};
    return (unsigned)l < 5 ? nm[l] : "unknown";
}

const char *faf_norm_name(faf_norm m) {
    static const char *nm[] = {"default", "none", "ortho", "forward", "lazy", "jpeg2000"};
    return (unsigned)m < 6 ? nm[m] : "unknown";
}

const char *faf_transform_name(faf_transform_type t) {
    static const char *nm[] = {"fft", "ifft", "rfft", "irfft", "dct_i", "dct_ii", "dct_iii",
                               "dct_iv", "dst_i", "dst_ii", "dst_iii", "dst_iv", "stft",
                               "mdct", "imdct", "haar", "daubechies4", "cdf53", "cdf97",
                               "sym4", "pipeline", "cwt", "icwt"};
    return (unsigned)t < 23 ? nm[t] : "unknown";
}

/* ------------------------------------------------------------------ */
/* size policy (reference src/faf_core.c:374-429)                       */
/* ------------------------------------------------------------------ */

size_t faf_spectrum_len(const faf_transform *t) {
    if (!t) return 0;
    if (t->type == FAF_TRANSFORM_RFFT || t->type == FAF_TRANSFORM_IRFFT ||
        t->type == FAF_TRANSFORM_PIPELINE)
        return t->n / 2 + 1;
    return t->n;
}

static int is_pow2(size_t n) { return n && !(n & (n - 1)); }

bool faf_is_size_supported(faf_transform_type type, size_t n) {
    if (n == 0) return false;
    switch (type) {
    case FAF_TRANSFORM_FFT: case FAF_TRANSFORM_IFFT:
        return fafi_is_5_smooth(n) || fafi_is_codelet_size(n) || fafi_rader_eligible(n);
    case FAF_TRANSFORM_RFFT: case FAF_TRANSFORM_IRFFT:
    case FAF_TRANSFORM_CWT: case FAF_TRANSFORM_ICWT:
        return fafi_is_5_smooth(n) && n % 2 == 0 && n >= 2;
    case FAF_TRANSFORM_HAAR: case FAF_TRANSFORM_DAUBECHIES4: case FAF_TRANSFORM_CDF53:
    case FAF_TRANSFORM_CDF97: case FAF_TRANSFORM_SYM4:
        return is_pow2(n);
    case FAF_TRANSFORM_DCT_I: return n >= 2;
    case FAF_TRANSFORM_MDCT: case FAF_TRANSFORM_IMDCT: return n % 2 == 0;
    default: return true;
    }
}

size_t faf_get_recommended_size(faf_transform_type type, size_t min_size) {
    int fft = type == FAF_TRANSFORM_FFT || type == FAF_TRANSFORM_IFFT;
    int real = type == FAF_TRANSFORM_RFFT || type == FAF_TRANSFORM_IRFFT ||
               type == FAF_TRANSFORM_CWT || type == FAF_TRANSFORM_ICWT;
    if (fft || real) {
        if (fft && faf_is_size_supported(type, min_size)) return min_size;
        size_t n = fafi_next_5_smooth(min_size);
        if (real) {
            if (n < 2) n = 2;
            if (n % 2) n = fafi_next_5_smooth(n + 1);
        }
        return n;
    }
    size_t p = 1;
    while (p < min_size) p <<= 1;
    return p;
}

/* ------------------------------------------------------------------ */
/* creation                                                            */
/* ------------------------------------------------------------------ */

static float norm_scale(size_t n, faf_norm norm, int inverse) {
    /* reference src/faf_core.c:1251-1279; FP64 value rounded once */
    switch (norm) {
    case FAF_NORM_ORTHO: return (float)(1.0 / sqrt((double)n));
    case FAF_NORM_FORWARD: return inverse ? 1.0f : (float)(1.0 / (double)n);
    default: return inverse ? (float)(1.0 / (double)n) : 1.0f;
    }
}

static faf_transform *new_transform(void) {
    faf_transform *t = calloc(1, sizeof(faf_transform));
    if (!t) { fafi_err("Failed to allocate transform"); return NULL; }
    t->amd = fafi_plan_new();
    if (!t->amd) { free(t); fafi_err("Failed to allocate transform"); return NULL; }
    return t;
}

static void resolve_into(faf_transform *t, faf_config c) {
    t->cfg = c;
    t->n = c.n;
    t->precision = c.precision;
    t->flags = c.flags;
    if (c.dir == FAF_DIR_INVERSE) t->flags |= FAF_FLAG_INVERSE;
    t->hop_length = c.hop_length;
    t->win_length = c.win_length;
    t->family = c.family;
    t->levels = c.levels;
}

/* FP32 everywhere; FP64 for the transforms that have the FP64 variant
 * (FFT, RFFT, DCT/DST/MDCT: fp64_ok) */
static int check_precision_ex(faf_precision p, const char *what, int fp64_ok) {
    if (p == FAF_PREC_FP32) return 0;
    if (p == FAF_PREC_FP64) {
        if (fp64_ok) return 0;
        fafi_err("%s: FP64 execution covers FFT, RFFT and DCT/DST/MDCT; this transform "
                 "computes in FP32", what);
        return -1;
    }
    fafi_err("%s supports FP32 and FP64 only", what);
    return -1;
}

static int check_precision(faf_precision p, const char *what) {
    return check_precision_ex(p, what, 0);
}

static double norm_scale64(size_t n, faf_norm norm, int inverse) {
    switch (norm) {
    case FAF_NORM_ORTHO: return 1.0 / sqrt((double)n);
    case FAF_NORM_FORWARD: return inverse ? 1.0 : 1.0 / (double)n;
    default: return inverse ? 1.0 / (double)n : 1.0;
    }
}

/* FAF_FLAG_MEASURE (reference src/faf_core.c:539-565, called at :694):
 * time the plan's launch shapes now when its key has no recorded shape (a
 * recorded one was already applied while planning); the batch is 64 MB of
 * complex signals */
static int measure_at_create(faf_transform *t) {
    fafh_plan *p = t->amd;
    if (!(t->flags & FAF_FLAG_MEASURE) || !p) return 0;
    if (p->kind != PK_C2C && p->kind != PK_R2C && p->kind != PK_C2R) return 0;
    if (fafi_tune_find(p->kind, p->n, p->inverse, p->interleaved, NULL)) return 0;
    size_t count = ((size_t)64 << 20) / (p->n * 8);
    if (count < 16) count = 16;
    if (count > 65536) count = 65536;
    return fafi_autotune_plan(p, count) < 0 ? -1 : 0;
}

faf_transform *faf_create_fft(const faf_config *cfg) {
    /* planner order of reference src/faf_core.c:572-719 */
    if (!cfg) { fafi_err("faf_create_fft: cfg is NULL"); return NULL; }
    faf_config c = *cfg;
    size_t n = c.n;
    int legal = fafi_is_5_smooth(n) || fafi_is_codelet_size(n);
    int rader = !legal && fafi_rader_eligible(n);
    int blue = !legal && !rader && (c.flags & FAF_FLAG_BLUESTEIN);
    if (!legal && !rader && !blue) {
        fafi_err("FFT size must be 5-smooth (2^a 3^b 5^c), a Rader prime (n-1 7-smooth), "
                 "or a small codelet; got %zu; nearest 5-smooth is %zu (or set "
                 "FAF_FLAG_BLUESTEIN)", n, faf_get_recommended_size(FAF_TRANSFORM_FFT, n));
        return NULL;
    }
    if (c.layout == FAF_LAYOUT_DEFAULT) c.layout = FAF_LAYOUT_SPLIT;
    if (c.norm == FAF_NORM_DEFAULT) c.norm = FAF_NORM_NONE;
    if (c.layout == FAF_LAYOUT_REAL || c.layout == FAF_LAYOUT_HERMITIAN) {
        fafi_err("FFT C2C does not support layout '%s'", faf_layout_name(c.layout));
        return NULL;
    }
    if (c.norm == FAF_NORM_LAZY || c.norm == FAF_NORM_JPEG2000) {
        fafi_err("FFT does not support norm '%s'", faf_norm_name(c.norm));
        return NULL;
    }
    if (check_precision_ex(c.precision, rader ? "Rader" : blue ? "Bluestein" : "FFT", 1))
        return NULL;
    if (fafi_create_ready()) return NULL;
    faf_transform *t = new_transform();
    if (!t) return NULL;
    int inv = c.dir == FAF_DIR_INVERSE;
    t->type = inv ? FAF_TRANSFORM_IFFT : FAF_TRANSFORM_FFT;
    if (blue) c.flags |= FAF_FLAG_BLUESTEIN;
    resolve_into(t, c);
    fafh_plan *p = t->amd;
    p->interleaved = c.layout == FAF_LAYOUT_INTERLEAVED;
    if (c.precision == FAF_PREC_FP64) {
        if (fafi_plan_c2c64(p, n, inv, blue, 0)) {
            faf_destroy_transform(t);
            return NULL;
        }
        p->scale64 = norm_scale64(n, c.norm, inv);
        return t;
    }
    if (fafi_plan_c2c(p, n, inv, blue)) {
        faf_destroy_transform(t);
        return NULL;
    }
    p->scale = norm_scale(n, c.norm, inv);
    if (measure_at_create(t)) {
        faf_destroy_transform(t);
        return NULL;
    }
    return t;
}

faf_transform *faf_create_rfft(const faf_config *cfg) {
    /* reference src/faf_rfft.c:174-266 */
    if (!cfg) { fafi_err("faf_create_rfft: cfg is NULL"); return NULL; }
    faf_config c = *cfg;
    if (c.flags & FAF_FLAG_BLUESTEIN) {
        fafi_err("R2C does not support Bluestein; use C2C with FAF_FLAG_BLUESTEIN on a real "
                 "buffer");
        return NULL;
    }
    if (c.n < 2 || (c.n & 1)) { fafi_err("R2C size must be even and >= 2, got %zu", c.n); return NULL; }
    if (!fafi_is_5_smooth(c.n)) {
        fafi_err("R2C size must be even and 5-smooth, got %zu; nearest is %zu", c.n,
                 faf_get_recommended_size(FAF_TRANSFORM_RFFT, c.n));
        return NULL;
    }
    if (check_precision_ex(c.precision, "R2C", 1)) return NULL;
    if (c.layout == FAF_LAYOUT_DEFAULT) c.layout = FAF_LAYOUT_HERMITIAN;
    if (c.norm == FAF_NORM_DEFAULT) c.norm = FAF_NORM_NONE;
    if (c.layout != FAF_LAYOUT_HERMITIAN && c.layout != FAF_LAYOUT_INTERLEAVED) {
        fafi_err("R2C layout must be hermitian or interleaved, got '%s'",
                 faf_layout_name(c.layout));
        return NULL;
    }
    if (c.norm == FAF_NORM_LAZY || c.norm == FAF_NORM_JPEG2000) {
        fafi_err("R2C does not support norm '%s'", faf_norm_name(c.norm));
        return NULL;
    }
    if (c.n == 2) {
        fafi_err("R2C size 2 has a one-point inner transform; use n >= 4");
        return NULL;
    }
    if (fafi_create_ready()) return NULL;
    faf_transform *t = new_transform();
    if (!t) return NULL;
    int inv = c.dir == FAF_DIR_INVERSE;
    t->type = inv ? FAF_TRANSFORM_IRFFT : FAF_TRANSFORM_RFFT;
    resolve_into(t, c);
    t->flags |= FAF_FLAG_REAL;
    fafh_plan *p = t->amd;
    p->interleaved = c.layout == FAF_LAYOUT_INTERLEAVED;
    if (c.precision == FAF_PREC_FP64) {
        if (fafi_plan_real64(p, c.n, inv)) {
            faf_destroy_transform(t);
            return NULL;
        }
        p->scale64 = norm_scale64(c.n, c.norm, inv);
        return t;
    }
    if (fafi_plan_real(p, c.n, inv)) {
        faf_destroy_transform(t);
        return NULL;
    }
    p->scale = norm_scale(c.n, c.norm, inv);
    if (measure_at_create(t)) {
        faf_destroy_transform(t);
        return NULL;
    }
    return t;
}

faf_transform *faf_create_stft(const faf_config *cfg) {
    /* reference src/faf_core.c:994-1032 builds a single-frame Hann-windowed
     * C2C FFT; that meaning is kept for SPLIT/INTERLEAVED.  layout REAL (an
     * error in the reference) selects the framed real STFT of faf_amd.h. */
    if (!cfg) { fafi_err("faf_create_stft: cfg is NULL"); return NULL; }
    faf_config c = *cfg;
    if (check_precision(c.precision, "STFT")) return NULL;
    if (c.win_length == 0) c.win_length = c.n;
    if (c.hop_length == 0) c.hop_length = c.win_length / 4;
    if (c.win_length > c.n) {
        fafi_err("STFT window length %zu exceeds FFT size %zu", c.win_length, c.n);
        return NULL;
    }
    if (c.layout == FAF_LAYOUT_REAL) {
        if (check_precision(c.precision, "STFT")) return NULL;
        if (fafi_create_ready()) return NULL;
        faf_transform *t = new_transform();
        if (!t) return NULL;
        c.dir = FAF_DIR_FORWARD;
        if (c.norm == FAF_NORM_DEFAULT) c.norm = FAF_NORM_NONE;
        resolve_into(t, c);
        t->type = FAF_TRANSFORM_STFT;
        if (fafi_plan_stft(t->amd, c.win_length, c.hop_length)) {
            faf_destroy_transform(t);
            return NULL;
        }
        return t;
    }
    faf_config fc = c;
    fc.dir = FAF_DIR_FORWARD;
    faf_transform *t = faf_create_fft(&fc);
    if (!t) return NULL;
    t->type = FAF_TRANSFORM_STFT;
    t->hop_length = t->cfg.hop_length = c.hop_length;
    t->win_length = t->cfg.win_length = c.win_length;
    /* window for the single-frame form (applied on the INTERLEAVED path,
     * reference src/faf_core.c:1369-1383) */
    fafh_plan *p = t->amd;
    float *w = calloc(c.n, sizeof(float));
    faf_gen_hann_window_f32(w, c.win_length);
    p->win_dev = fafk_malloc(c.n * sizeof(float));
    if (!p->win_dev || fafk_h2d(p->win_dev, w, c.n * sizeof(float), NULL) || fafk_sync(NULL)) {
        free(w);
        fafi_err("%s", fafk_last_error());
        faf_destroy_transform(t);
        return NULL;
    }
    free(w);
    p->win = c.win_length;
    p->hop = c.hop_length;
    return t;
}

static faf_transform *unsupported(const char *what) {
    fafi_err("%s is not part of the device transform path in this build", what);
    return NULL;
}

/* DCT / DST / MDCT: creation and validation as reference
 * src/faf_core.c:721-824 (type 0 means II, REAL layout, NONE norm by
 * default, MDCT sizes even); the transforms themselves are the ones the
 * reference names but leaves as FFT placeholders (faf_r2r.c) */
static faf_transform *create_r2r(const faf_config *cfg, int kind) {
    static const char *nm[3] = {"DCT", "DST", "MDCT"};
    if (!cfg) {
        fafi_err("faf_create_%s: cfg is NULL", kind == 0 ? "dct" : kind == 1 ? "dst" : "mdct");
        return NULL;
    }
    faf_config c = *cfg;
    int type = 2;
    if (kind != 2) {
        if (c.n == 0) { fafi_err("%s size must be > 0", nm[kind]); return NULL; }
        type = c.dct_type ? c.dct_type : 2;
        if (type < 1 || type > 4) {
            fafi_err("%s type must be 1-4, got %d", nm[kind], type);
            return NULL;
        }
        c.dct_type = type;
    } else if (c.n % 2 != 0 || c.n == 0) {
        fafi_err("MDCT size must be even, got %zu", c.n);
        return NULL;
    }
    if (c.layout == FAF_LAYOUT_DEFAULT) c.layout = FAF_LAYOUT_REAL;
    if (c.norm == FAF_NORM_DEFAULT) c.norm = FAF_NORM_NONE;
    if (c.layout != FAF_LAYOUT_REAL) {
        fafi_err("%s needs layout REAL, got '%s'", nm[kind], faf_layout_name(c.layout));
        return NULL;
    }
    if (c.norm != FAF_NORM_NONE && c.norm != FAF_NORM_ORTHO && c.norm != FAF_NORM_FORWARD) {
        fafi_err("%s does not support norm '%s'", nm[kind], faf_norm_name(c.norm));
        return NULL;
    }
    if (check_precision_ex(c.precision, nm[kind], 1)) return NULL;
    if (fafi_create_ready()) return NULL;
    faf_transform *t = new_transform();
    if (!t) return NULL;
    const int inv = c.dir == FAF_DIR_INVERSE;
    if (kind == 0) t->type = (faf_transform_type)(FAF_TRANSFORM_DCT_I + (type - 1));
    else if (kind == 1) t->type = (faf_transform_type)(FAF_TRANSFORM_DST_I + (type - 1));
    else t->type = inv ? FAF_TRANSFORM_IMDCT : FAF_TRANSFORM_MDCT;
    resolve_into(t, c);
    if (fafi_plan_r2r(t->amd, kind, type, c.n, inv, c.norm, c.precision == FAF_PREC_FP64)) {
        faf_destroy_transform(t);
        return NULL;
    }
    return t;
}

faf_transform *faf_create_dct(const faf_config *cfg) { return create_r2r(cfg, 0); }
faf_transform *faf_create_dst(const faf_config *cfg) { return create_r2r(cfg, 1); }
faf_transform *faf_create_mdct(const faf_config *cfg) { return create_r2r(cfg, 2); }
faf_transform *faf_create_dwt(const faf_config *cfg) { (void)cfg; return unsupported("DWT"); }

faf_transform *faf_create(faf_transform_type type, const faf_config *cfg) {
    /* reference src/faf_core.c:1034-1092 */
    if (!cfg) { fafi_err("faf_create: cfg is NULL"); return NULL; }
    faf_config c = *cfg;
    switch (type) {
    case FAF_TRANSFORM_IFFT: c.dir = FAF_DIR_INVERSE; return faf_create_fft(&c);
    case FAF_TRANSFORM_FFT: return faf_create_fft(&c);
    case FAF_TRANSFORM_IRFFT: c.dir = FAF_DIR_INVERSE; return faf_create_rfft(&c);
    case FAF_TRANSFORM_RFFT: return faf_create_rfft(&c);
    case FAF_TRANSFORM_STFT: return faf_create_stft(&c);
    case FAF_TRANSFORM_CWT: case FAF_TRANSFORM_ICWT:
        fafi_err("CWT: use faf_create_cwt / faf_create_icwt");
        return NULL;
    case FAF_TRANSFORM_DCT_I: case FAF_TRANSFORM_DCT_II: case FAF_TRANSFORM_DCT_III:
    case FAF_TRANSFORM_DCT_IV:
        c.dct_type = 1 + (int)(type - FAF_TRANSFORM_DCT_I);
        return faf_create_dct(&c);
    case FAF_TRANSFORM_DST_I: case FAF_TRANSFORM_DST_II: case FAF_TRANSFORM_DST_III:
    case FAF_TRANSFORM_DST_IV:
        c.dct_type = 1 + (int)(type - FAF_TRANSFORM_DST_I);
        return faf_create_dst(&c);
    case FAF_TRANSFORM_IMDCT: c.dir = FAF_DIR_INVERSE; return faf_create_mdct(&c);
    case FAF_TRANSFORM_MDCT: return faf_create_mdct(&c);
    case FAF_TRANSFORM_HAAR: case FAF_TRANSFORM_DAUBECHIES4: case FAF_TRANSFORM_CDF53:
    case FAF_TRANSFORM_CDF97: case FAF_TRANSFORM_SYM4: return faf_create_dwt(&c);
    default:
        fafi_err("Unknown transform type %d", (int)type);
        return NULL;
    }
}

faf_transform *faf_create_inverse(const faf_transform *fwd) {
    /* reference src/faf_core.c:1094-1158 */
    if (!fwd) { fafi_err("faf_create_inverse: transform is NULL"); return NULL; }
    faf_config c = faf_config_inverse(fwd->cfg);
    switch (fwd->type) {
    case FAF_TRANSFORM_FFT: case FAF_TRANSFORM_IFFT: return faf_create_fft(&c);
    case FAF_TRANSFORM_RFFT: case FAF_TRANSFORM_IRFFT: return faf_create_rfft(&c);
    case FAF_TRANSFORM_STFT: return faf_create_stft(&c);
    case FAF_TRANSFORM_DCT_I: case FAF_TRANSFORM_DCT_II: case FAF_TRANSFORM_DCT_III:
    case FAF_TRANSFORM_DCT_IV:
        /* II <-> III; I and IV pair with themselves (dir flips the scale side) */
        if (fwd->type == FAF_TRANSFORM_DCT_II) c.dct_type = 3;
        else if (fwd->type == FAF_TRANSFORM_DCT_III) c.dct_type = 2;
        return faf_create_dct(&c);
    case FAF_TRANSFORM_DST_I: case FAF_TRANSFORM_DST_II: case FAF_TRANSFORM_DST_III:
    case FAF_TRANSFORM_DST_IV:
        if (fwd->type == FAF_TRANSFORM_DST_II) c.dct_type = 3;
        else if (fwd->type == FAF_TRANSFORM_DST_III) c.dct_type = 2;
        return faf_create_dst(&c);
    case FAF_TRANSFORM_MDCT: case FAF_TRANSFORM_IMDCT: return faf_create_mdct(&c);
    case FAF_TRANSFORM_CWT: {
        const fafh_plan *p = fwd->amd;
        if (!p || !p->bank) { fafi_err("faf_create_inverse: CWT has no bank"); return NULL; }
        return faf_create_icwt((const faf_cwt_config *)p->bank, FAF_CWT_INV_DUAL);
    }
    default:
        fafi_err("faf_create_inverse: unsupported type %s", faf_transform_name(fwd->type));
        return NULL;
    }
}

void faf_destroy_transform(faf_transform *t) {
    if (!t) return;
    if (t->amd) {
        fafh_plan *p = t->amd;
        if (p->kind != PK_NONE && !t_capture) fafk_sync(fafi_stream());
        fafi_plan_free(p);
    }
    free(t->code);
    free(t);
}

/* ------------------------------------------------------------------ */
/* execution                                                           */
/* ------------------------------------------------------------------ */

/* A set of planes moved between caller memory and device memory. */
typedef struct {
    int n;
    const void *host[4];    /* caller pointers */
    size_t bytes[4];
    int is_out[4];
    void *dev[4];           /* device pointers the kernels use */
} planes_t;

static void planes_add(planes_t *ps, const void *ptr, size_t bytes, int is_out) {
    ps->host[ps->n] = ptr;
    ps->bytes[ps->n] = bytes;
    ps->is_out[ps->n] = is_out;
    ps->dev[ps->n] = (void *)ptr;
    ps->n++;
}

/* Give every host-resident plane a device twin; inputs are copied in. */
static int planes_stage_in(fafh_plan *p, planes_t *ps, void *stream) {
    size_t need = 0;
    int host[4];
    for (int i = 0; i < ps->n; ++i) {
        host[i] = ps->host[i] && !fafk_is_device_ptr(ps->host[i]);
        if (host[i]) need += (ps->bytes[i] + 255) & ~(size_t)255;
    }
    if (!need) return 0;
    if (t_capture) {
        fafi_err("host buffers cannot be captured into a graph; pass device pointers");
        return -1;
    }
    if (t_noalloc) {
        fafi_err("faf_execute_batch_noalloc: every buffer must be device memory "
                 "(host buffers need staging memory)");
        return -1;
    }
    char *base = fafi_staging(p, need);
    if (!base) return -1;
    size_t off = 0;
    for (int i = 0; i < ps->n; ++i) {
        if (!host[i]) continue;
        /* an output that aliases an input host pointer shares its twin */
        int twin = -1;
        for (int j = 0; j < i; ++j)
            if (host[j] && ps->host[j] == ps->host[i]) twin = j;
        if (twin >= 0) { ps->dev[i] = ps->dev[twin]; host[i] = 2; continue; }
        ps->dev[i] = base + off;
        off += (ps->bytes[i] + 255) & ~(size_t)255;
        if (!ps->is_out[i] && fafk_h2d(ps->dev[i], ps->host[i], ps->bytes[i], stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
    }
    return 0;
}

static int planes_stage_out(planes_t *ps, void *stream) {
    int any = 0;
    for (int i = 0; i < ps->n; ++i) {
        if (!ps->is_out[i] || ps->dev[i] == ps->host[i]) continue;
        if (fafk_d2h((void *)ps->host[i], ps->dev[i], ps->bytes[i], stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
        any = 1;
    }
    return any;
}

static int exec_planes(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                       int (*run)(const faf_transform *, planes_t *, long, long, long, void *)) {
    void *stream = fafi_stream();
    fafh_plan *p = t->amd;
    g_err[0] = '\0';
    if (planes_stage_in(p, ps, stream)) {
        t_cap_failed |= t_capture;
        return -1;
    }
    void *ev0 = NULL, *ev1 = NULL;
    if (t_timing && !t_capture && !t_noalloc) {
        ev0 = fafk_event_create();
        ev1 = fafk_event_create();
        if (ev0) fafk_event_record(ev0, stream);
    }
    int rc = run(t, ps, count, idist, odist, stream);
    if (!rc && ev1) fafk_event_record(ev1, stream);
    int staged = rc ? 0 : planes_stage_out(ps, stream);
    if (staged < 0) rc = -1;
    if (!rc && (staged || fafi_sync_mode() || ev1) && fafk_sync(stream)) {
        fafi_err("%s", fafk_last_error());
        rc = -1;
    }
    if (ev0 && ev1 && !rc) t_last_ms = fafk_event_ms(ev0, ev1);
    fafk_event_destroy(ev0);
    fafk_event_destroy(ev1);
    if (rc && !g_err[0]) fafi_err("%s", fafk_last_error());
    if (rc) t_cap_failed |= t_capture;
    return rc;
}

/* runners: planes layout fixed per kind */
static int run_c2c(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                   void *stream) {
    fafh_plan *p = t->amd;
    if (p->fp64) {
        if (p->interleaved)
            return fafi_run_c2c64(p, ps->dev[0], NULL, idist, ps->dev[1], NULL, odist, count,
                                  stream);
        return fafi_run_c2c64(p, ps->dev[0], ps->dev[1], idist, ps->dev[2], ps->dev[3], odist,
                              count, stream);
    }
    if (p->interleaved)
        return fafi_run_c2c(p, ps->dev[0], NULL, idist, ps->dev[1], NULL, odist, count, stream);
    return fafi_run_c2c(p, ps->dev[0], ps->dev[1], idist, ps->dev[2], ps->dev[3], odist, count,
                        stream);
}

static int run_r2c(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                   void *stream) {
    fafh_plan *p = t->amd;
    if (p->fp64)
        return fafi_run_r2c64(p, ps->dev[0], idist, ps->dev[1],
                              p->interleaved ? NULL : ps->dev[2], odist, count, stream);
    return fafi_run_r2c(p, ps->dev[0], idist, ps->dev[1], p->interleaved ? NULL : ps->dev[2],
                        odist, count, stream);
}

static int run_c2r(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                   void *stream) {
    fafh_plan *p = t->amd;
    if (p->fp64) {
        if (p->interleaved)
            return fafi_run_c2r64(p, ps->dev[0], NULL, idist, ps->dev[1], odist, count, stream);
        return fafi_run_c2r64(p, ps->dev[0], ps->dev[1], idist, ps->dev[2], odist, count,
                              stream);
    }
    if (p->interleaved)
        return fafi_run_c2r(p, ps->dev[0], NULL, idist, ps->dev[1], odist, count, stream);
    return fafi_run_c2r(p, ps->dev[0], ps->dev[1], idist, ps->dev[2], odist, count, stream);
}

static int run_pipe(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                    void *stream) {
    return fafi_run_pipe(t->amd, ps->dev[0], idist, ps->dev[1], odist, count, stream);
}

static int run_cwt(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                   void *stream) {
    return fafi_run_cwt(t->amd, ps->dev[0], idist, ps->dev[1], odist, count, 0, stream);
}

static int run_cwt_complex(const faf_transform *t, planes_t *ps, long count, long idist,
                           long odist, void *stream) {
    return fafi_run_cwt(t->amd, ps->dev[0], idist, ps->dev[1], odist, count, 1, stream);
}

static int run_icwt(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                    void *stream) {
    return fafi_run_icwt(t->amd, ps->dev[0], idist, ps->dev[1], odist, count, stream);
}

static int run_stft_single(const faf_transform *t, planes_t *ps, long count, long idist,
                           long odist, void *stream);

static int run_r2r(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                   void *stream) {
    fafh_plan *p = t->amd;
    if (p->fp64) return fafi_run_r2r64(p, ps->dev[0], idist, ps->dev[1], odist, count, stream);
    return fafi_run_r2r(p, ps->dev[0], idist, ps->dev[1], odist, count, stream);
}

/* the body shared by faf_execute and faf_execute_batch */
static int execute_n(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                     size_t count, size_t in_dist, size_t out_dist) {
    if (!t || !out || !in) return -1;
    fafh_plan *p = t->amd;
    if (!p) { fafi_err("transform has no device plan"); return -1; }
    if (fafi_runtime_ready()) return -1;
    if (count == 0) return 0;
    planes_t ps;
    memset(&ps, 0, sizeof ps);
    size_t n = t->n;
    const size_t f = p->fp64 ? sizeof(double) : sizeof(float);

    if (t->type == FAF_TRANSFORM_PIPELINE) {
        if (in->layout != FAF_LAYOUT_REAL || in->n != n || !in->re) {
            fafi_err("pipeline input must be REAL of length %zu", n);
            return -1;
        }
        if (out->layout != FAF_LAYOUT_REAL || out->n != n || !out->re) {
            fafi_err("pipeline output must be REAL of length %zu", n);
            return -1;
        }
        size_t id = in_dist ? in_dist : n, od = out_dist ? out_dist : n;
        planes_add(&ps, in->re, ((count - 1) * id + n) * f, 0);
        planes_add(&ps, out->re, ((count - 1) * od + n) * f, 1);
        return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_pipe);
    }

    if (t->type == FAF_TRANSFORM_RFFT || t->type == FAF_TRANSFORM_IRFFT) {
        /* buffer contract of reference src/faf_rfft.c:268-335 */
        if (!in->re || !out->re) { fafi_err("execute buffers must provide re"); return -1; }
        size_t nb = n / 2 + 1;
        faf_layout sl = t->cfg.layout;
        int inter = sl == FAF_LAYOUT_INTERLEAVED;
        if (t->type == FAF_TRANSFORM_RFFT) {
            if (in->layout != FAF_LAYOUT_REAL) {
                fafi_err("rfft input must be REAL, got %s", faf_layout_name(in->layout));
                return -1;
            }
            if (in->n != n) { fafi_err("rfft input n=%zu, expected %zu", in->n, n); return -1; }
            if (out->layout != sl) {
                fafi_err("rfft output layout (%s) does not match transform (%s)",
                         faf_layout_name(out->layout), faf_layout_name(sl));
                return -1;
            }
            if (out->n != nb) {
                fafi_err("rfft output n=%zu, expected %zu bins", out->n, nb);
                return -1;
            }
            if (!inter && !out->im) { fafi_err("hermitian rfft output requires im"); return -1; }
            size_t id = in_dist ? in_dist : n, od = out_dist ? out_dist : nb;
            planes_add(&ps, in->re, ((count - 1) * id + n) * f, 0);
            if (inter) {
                planes_add(&ps, out->re, ((count - 1) * od + nb) * 2 * f, 1);
            } else {
                planes_add(&ps, out->re, ((count - 1) * od + nb) * f, 1);
                planes_add(&ps, out->im, ((count - 1) * od + nb) * f, 1);
            }
            return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_r2c);
        }
        if (in->layout != sl) {
            fafi_err("irfft input layout (%s) does not match transform (%s)",
                     faf_layout_name(in->layout), faf_layout_name(sl));
            return -1;
        }
        if (in->n != nb) { fafi_err("irfft input n=%zu, expected %zu bins", in->n, nb); return -1; }
        if (!inter && !in->im) { fafi_err("hermitian irfft input requires im"); return -1; }
        if (out->layout != FAF_LAYOUT_REAL) {
            fafi_err("irfft output must be REAL, got %s", faf_layout_name(out->layout));
            return -1;
        }
        if (out->n != n) { fafi_err("irfft output n=%zu, expected %zu", out->n, n); return -1; }
        size_t id = in_dist ? in_dist : nb, od = out_dist ? out_dist : n;
        if (inter) {
            planes_add(&ps, in->re, ((count - 1) * id + nb) * 2 * f, 0);
        } else {
            planes_add(&ps, in->re, ((count - 1) * id + nb) * f, 0);
            planes_add(&ps, in->im, ((count - 1) * id + nb) * f, 0);
        }
        planes_add(&ps, out->re, ((count - 1) * od + n) * f, 1);
        return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_c2r);
    }

    if (p->kind == PK_R2R) {
        /* REAL in, REAL out: n -> n (DCT / DST), n -> n/2 (MDCT), n/2 -> n (IMDCT) */
        const size_t nin = (size_t)p->r2r_nin, nout = (size_t)p->r2r_nout;
        if (!in->re || !out->re) { fafi_err("execute buffers must provide re"); return -1; }
        if (in->layout != FAF_LAYOUT_REAL || out->layout != FAF_LAYOUT_REAL) {
            fafi_err("%s buffers must be REAL", faf_transform_name(t->type));
            return -1;
        }
        if (in->n != nin || out->n != nout) {
            fafi_err("%s buffer n (%zu/%zu), expected %zu in and %zu out",
                     faf_transform_name(t->type), in->n, out->n, nin, nout);
            return -1;
        }
        size_t id = in_dist ? in_dist : nin, od = out_dist ? out_dist : nout;
        planes_add(&ps, in->re, ((count - 1) * id + nin) * f, 0);
        planes_add(&ps, out->re, ((count - 1) * od + nout) * f, 1);
        return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_r2r);
    }

    if (t->type == FAF_TRANSFORM_CWT || t->type == FAF_TRANSFORM_ICWT) {
        size_t rows = faf_cwt_n_rows(t);
        if (!in->re || !out->re) { fafi_err("cwt: buffers must provide re"); return -1; }
        if (in->layout != FAF_LAYOUT_REAL || out->layout != FAF_LAYOUT_REAL) {
            fafi_err("cwt: layout mismatch");
            return -1;
        }
        if (t->type == FAF_TRANSFORM_CWT) {
            if (in->n != n) { fafi_err("cwt: buffer n mismatch"); return -1; }
            if (in->re == out->re) { fafi_err("cwt: in and out alias"); return -1; }
            size_t id = in_dist ? in_dist : n, od = out_dist ? out_dist : rows * n;
            planes_add(&ps, in->re, ((count - 1) * id + n) * f, 0);
            planes_add(&ps, out->re, ((count - 1) * od + rows * n) * f, 1);
            return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_cwt);
        }
        if (out->n != n) { fafi_err("cwt: buffer n mismatch"); return -1; }
        size_t id = in_dist ? in_dist : rows * n, od = out_dist ? out_dist : n;
        planes_add(&ps, in->re, ((count - 1) * id + rows * n) * f, 0);
        planes_add(&ps, out->re, ((count - 1) * od + n) * f, 1);
        return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_icwt);
    }

    if (t->type == FAF_TRANSFORM_STFT && p->kind == PK_STFT) {
        fafi_err("framed STFT: use faf_execute_stft");
        return -1;
    }

    /* C2C (and the single-frame STFT): reference src/faf_core.c:1322-1349 */
    if ((t->flags & FAF_FLAG_BLUESTEIN) && (in->n != n || out->n != n)) {
        fafi_err("Bluestein buffer n (%zu/%zu) must equal %zu", in->n, out->n, n);
        return -1;
    }
    if (in->layout != t->cfg.layout || out->layout != t->cfg.layout) {
        fafi_err("buffer layout (%s/%s) does not match transform (%s)",
                 faf_layout_name(in->layout), faf_layout_name(out->layout),
                 faf_layout_name(t->cfg.layout));
        return -1;
    }
    if (!in->re || !out->re) { fafi_err("execute buffers must provide re"); return -1; }
    size_t id = in_dist ? in_dist : n, od = out_dist ? out_dist : n;
    if (t->cfg.layout == FAF_LAYOUT_SPLIT) {
        if (!in->im || !out->im) { fafi_err("split/hermitian layout requires im planes"); return -1; }
        planes_add(&ps, in->re, ((count - 1) * id + n) * f, 0);
        planes_add(&ps, in->im, ((count - 1) * id + n) * f, 0);
        planes_add(&ps, out->re, ((count - 1) * od + n) * f, 1);
        planes_add(&ps, out->im, ((count - 1) * od + n) * f, 1);
    } else {
        planes_add(&ps, in->re, ((count - 1) * id + n) * 2 * f, 0);
        planes_add(&ps, out->re, ((count - 1) * od + n) * 2 * f, 1);
    }
    if (t->type == FAF_TRANSFORM_STFT && p->interleaved)
        return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_stft_single);
    return exec_planes(t, &ps, (long)count, (long)id, (long)od, run_c2c);
}

int faf_execute(const faf_transform *t, faf_buffer *out, const faf_buffer *in) {
    return execute_n(t, out, in, 1, 0, 0);
}

int faf_execute_batch(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                      size_t count, size_t in_dist, size_t out_dist) {
    return execute_n(t, out, in, count, in_dist, out_dist);
}

_Static_assert(FAFK_SCRATCH_BYTES == FAF_NOALLOC_SCRATCH_BYTES,
               "the kernel layer's scratch is the size the header states");

int faf_execute_batch_noalloc(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                              size_t count, size_t in_dist, size_t out_dist) {
    if (!t || !out || !in) { fafi_err("faf_execute_batch_noalloc: null argument"); return -1; }
    fafh_plan *p = t->amd;
    if (!p) { fafi_err("transform has no device plan"); return -1; }
    if (fafi_runtime_ready()) return -1;
    size_t cap = 0;
    void *scr = fafk_scratch(&cap);   /* this device's copy of the static scratch */
    if (!scr) { fafi_err("faf_execute_batch_noalloc: %s", fafk_last_error()); return -1; }
    /* the plan's own workspace is replaced by the scratch for this call:
     * four-step runs cut their batch into chunks that fit it, and a run
     * that cannot fit (one chunk larger than the scratch) fails in
     * fafi_workspace with the sizes */
    void *save = p->ws_ext;
    size_t save_b = p->ws_ext_bytes;
    p->ws_ext = scr;
    p->ws_ext_bytes = cap;
    t_noalloc = 1;
    fafk_set_noalloc(1);
    int rc = execute_n(t, out, in, count, in_dist, out_dist);
    fafk_set_noalloc(0);
    t_noalloc = 0;
    p->ws_ext = save;
    p->ws_ext_bytes = save_b;
    return rc;
}

int faf_execute_f32(const faf_transform *t, float *out, const float *in) {
    /* interleaved convenience wrapper (reference src/faf_vm.c:976-1083) */
    if (!t || !out || !in) return -1;
    if (t->amd && ((const fafh_plan *)t->amd)->fp64) {
        fafi_err("faf_execute_f32: the transform is FP64; use faf_execute with double buffers");
        return -1;
    }
    if (t->type == FAF_TRANSFORM_FFT || t->type == FAF_TRANSFORM_IFFT ||
        t->type == FAF_TRANSFORM_STFT) {
        if (t->cfg.layout != FAF_LAYOUT_INTERLEAVED) {
            fafi_err("faf_execute_f32 needs an INTERLEAVED transform");
            return -1;
        }
        faf_buffer i = faf_buffer_interleaved((void *)in, t->n);
        faf_buffer o = faf_buffer_interleaved(out, t->n);
        return faf_execute(t, &o, &i);
    }
    if (t->type == FAF_TRANSFORM_RFFT) {
        faf_buffer i = faf_buffer_real((void *)in, t->n);
        faf_buffer o = faf_buffer_interleaved(out, t->n / 2 + 1);
        return faf_execute(t, &o, &i);
    }
    if (t->type == FAF_TRANSFORM_IRFFT) {
        faf_buffer i = faf_buffer_interleaved((void *)in, t->n / 2 + 1);
        faf_buffer o = faf_buffer_real(out, t->n);
        return faf_execute(t, &o, &i);
    }
    if (t->amd && ((const fafh_plan *)t->amd)->kind == PK_R2R) {
        const fafh_plan *p = t->amd;
        faf_buffer i = faf_buffer_real((void *)in, (size_t)p->r2r_nin);
        faf_buffer o = faf_buffer_real(out, (size_t)p->r2r_nout);
        return faf_execute(t, &o, &i);
    }
    fafi_err("faf_execute_f32: unsupported transform type %s", faf_transform_name(t->type));
    return -1;
}

int faf_execute_cwt_complex(const faf_transform *t, float *out, const float *x, size_t count,
                            size_t in_dist) {
    if (!t || !out || !x || t->type != FAF_TRANSFORM_CWT) {
        fafi_err("faf_execute_cwt_complex: needs a CWT transform and buffers");
        return -1;
    }
    if (fafi_runtime_ready()) return -1;
    size_t n = t->n, rows = faf_cwt_n_rows(t);
    size_t id = in_dist ? in_dist : n;
    planes_t ps;
    memset(&ps, 0, sizeof ps);
    planes_add(&ps, x, ((count - 1) * id + n) * sizeof(float), 0);
    planes_add(&ps, out, count * rows * n * 2 * sizeof(float), 1);
    return exec_planes(t, &ps, (long)count, (long)id, (long)(rows * n), run_cwt_complex);
}

/* ---- framed STFT ---- */

size_t faf_stft_frames(const faf_transform *t, size_t len) {
    if (!t || !t->amd) return 0;
    const fafh_plan *p = t->amd;
    size_t win = p->win, hop = p->hop;
    if (!win || !hop) return 0;
    if (len <= win) return 1;
    return (len - win) / hop + 1;
}

typedef struct { size_t len, records, rec_dist, frames; } stft_args;
static _Thread_local stft_args t_stft;

static int run_stft(const faf_transform *t, planes_t *ps, long count, long idist, long odist,
                    void *stream) {
    (void)idist; (void)odist;
    fafh_plan *p = t->amd;
    size_t nb = p->win / 2 + 1;
    if (p->ct) {
        int rc = fafk_ct_stft(&p->a, p->post, p->win_dev, ps->dev[0], (long)t_stft.len,
                              (long)t_stft.rec_dist, (long)p->hop, (long)t_stft.frames,
                              ps->dev[1], ps->dev[2], (long)nb, (long)(t_stft.frames * nb),
                              count, stream);
        if (rc) fafi_err("%s", fafk_ct_last_error());
        return rc;
    }
    int rc = fafk_stft(&p->a, p->spb, p->post, p->win_dev, ps->dev[0], (long)t_stft.len,
                       (long)t_stft.rec_dist, (long)p->hop, (long)t_stft.frames, ps->dev[1],
                       ps->dev[2], (long)nb, (long)(t_stft.frames * nb), count, stream);
    if (rc) fafi_err("%s", fafk_last_error());
    return rc;
}

int faf_execute_stft(const faf_transform *t, float *out_re, float *out_im, const float *x,
                     size_t len, size_t records, size_t rec_dist) {
    if (!t || !t->amd || ((fafh_plan *)t->amd)->kind != PK_STFT) {
        fafi_err("faf_execute_stft: needs a framed STFT (faf_create_stft with layout REAL)");
        return -1;
    }
    if (!out_re || !out_im || !x || len == 0 || records == 0) {
        fafi_err("faf_execute_stft: null buffer or empty record");
        return -1;
    }
    if (fafi_runtime_ready()) return -1;
    fafh_plan *p = t->amd;
    size_t frames = faf_stft_frames(t, len), nb = p->win / 2 + 1;
    size_t rd = rec_dist ? rec_dist : len;
    t_stft.len = len; t_stft.records = records; t_stft.rec_dist = rd; t_stft.frames = frames;
    planes_t ps;
    memset(&ps, 0, sizeof ps);
    planes_add(&ps, x, ((records - 1) * rd + len) * sizeof(float), 0);
    planes_add(&ps, out_re, records * frames * nb * sizeof(float), 1);
    planes_add(&ps, out_im, records * frames * nb * sizeof(float), 1);
    return exec_planes(t, &ps, (long)(records * frames), 0, 0, run_stft);
}

/* single-frame STFT on INTERLEAVED: window then FFT (reference
 * src/faf_core.c:1369-1383); the window multiply runs as a pipeline op of
 * the C2C kernel's source */
static int run_stft_single(const faf_transform *t, planes_t *ps, long count, long idist,
                           long odist, void *stream) {
    fafh_plan *p = t->amd;
    size_t n = t->n;
    /* windowed copy in the workspace, then the plain C2C */
    float *ws = fafi_workspace(p, (size_t)count * n * 2 * sizeof(float));
    if (!ws) return -1;
    if (fafk_window_inter(ps->dev[0], idist, p->win_dev, (long)n, ws, count, stream)) {
        fafi_err("%s", fafk_last_error());
        return -1;
    }
    return fafi_run_c2c(p, ws, NULL, (long)n, ps->dev[1], NULL, odist, count, stream);
}

/* ------------------------------------------------------------------ */
/* edge helpers operating on caller host memory (reference              */
/* src/faf_core.c:96-220).  These are layout utilities, not transforms. */
/* ------------------------------------------------------------------ */

int faf_deinterleave_f32(float *FAF_RESTRICT re, float *FAF_RESTRICT im,
                         const float *FAF_RESTRICT iv, size_t n) {
    if (!re || !im || !iv) { fafi_err("deinterleave: null buffer"); return -1; }
    for (size_t i = 0; i < n; ++i) { re[i] = iv[2 * i]; im[i] = iv[2 * i + 1]; }
    return 0;
}

int faf_interleave_f32(float *FAF_RESTRICT iv, const float *FAF_RESTRICT re,
                       const float *FAF_RESTRICT im, size_t n) {
    if (!iv || !re || !im) { fafi_err("interleave: null buffer"); return -1; }
    for (size_t i = 0; i < n; ++i) { iv[2 * i] = re[i]; iv[2 * i + 1] = im[i]; }
    return 0;
}

int faf_deinterleave_f64(double *FAF_RESTRICT re, double *FAF_RESTRICT im,
                         const double *FAF_RESTRICT iv, size_t n) {
    if (!re || !im || !iv) { fafi_err("deinterleave: null buffer"); return -1; }
    for (size_t i = 0; i < n; ++i) { re[i] = iv[2 * i]; im[i] = iv[2 * i + 1]; }
    return 0;
}

int faf_interleave_f64(double *FAF_RESTRICT iv, const double *FAF_RESTRICT re,
                       const double *FAF_RESTRICT im, size_t n) {
    if (!iv || !re || !im) { fafi_err("interleave: null buffer"); return -1; }
    for (size_t i = 0; i < n; ++i) { iv[2 * i] = re[i]; iv[2 * i + 1] = im[i]; }
    return 0;
}

#define HERM_MUL(NAME, T, SG)                                                         \
    void NAME(T *yr, T *yi, const T *xr, const T *xi, const T *hr, const T *hi,       \
              size_t nb) {                                                            \
        if (!yr || !yi || !xr || !xi || !hr || !hi || nb == 0) return;                \
        yr[0] = xr[0] * hr[0];                                                        \
        yi[0] = 0;                                                                    \
        for (size_t k = 1; k + 1 < nb; ++k) {                                         \
            T ar = xr[k], ai = xi[k], br = hr[k], bi = (SG) * hi[k];                  \
            yr[k] = ar * br - ai * bi;                                                \
            yi[k] = ar * bi + ai * br;                                                \
        }                                                                             \
        if (nb > 1) {                                                                 \
            yr[nb - 1] = xr[nb - 1] * hr[nb - 1];                                     \
            yi[nb - 1] = 0;                                                           \
        }                                                                             \
    }
HERM_MUL(faf_herm_mul_f32, float, 1.0f)
HERM_MUL(faf_herm_mul_f64, double, 1.0)
HERM_MUL(faf_herm_mul_conj_f32, float, -1.0f)
HERM_MUL(faf_herm_mul_conj_f64, double, -1.0)

void faf_gen_hann_window_f32(float *w, size_t n) {
    /* symmetric Hann, reference src/faf_twiddles.c:159-163 */
    if (!w || n == 0) return;
    if (n == 1) { w[0] = 0.0f; return; }
    for (size_t i = 0; i < n; ++i)
        w[i] = (float)(0.5 * (1.0 - cos(2.0 * 3.14159265358979323846 * (double)i / (double)(n - 1))));
}
