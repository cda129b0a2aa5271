/*
 * faf_chirp.c -- the spectral-pipeline part of the Chirp front end (C11).
 *
 * chirp_compile accepts the fused pipeline form of the reference
 * (src/chirp.c:1922-2055):
 *     (pipeline (rfft :size N [:norm ortho|forward|none]) OP ... (irfft))
 * with OP one of  conj | (conj) | mul-spectrum | (mul-spectrum) |
 * (bandpass :lo A :hi B).  The result runs as ONE device kernel
 * (rfft -> ops -> irfft in LDS).  Host callbacks registered with
 * chirp_register_spectral are kept in the registry for API compatibility,
 * but a pipeline that names one is rejected: device kernels cannot call
 * host code.  The rest of the Chirp language is not part of this path.
 */
#define _GNU_SOURCE
#include "faf_internal.h"
#include "chirp.h"
#include <ctype.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* registry of named host callbacks (reference src/chirp.c:314-364)     */
/* ------------------------------------------------------------------ */

typedef struct { char *name; void *fn; void *fn64; void *ctx; int kind; } reg_ent;
static reg_ent *g_reg;
static int g_nreg, g_capreg;
static pthread_mutex_t g_reg_mu = PTHREAD_MUTEX_INITIALIZER;

static int reg_add(const char *name, void *fn, void *fn64, void *ctx, int kind) {
    if (!name || !fn) return -1;
    pthread_mutex_lock(&g_reg_mu);
    for (int i = 0; i < g_nreg; ++i)
        if (!strcmp(g_reg[i].name, name)) {
            g_reg[i].fn = fn; g_reg[i].fn64 = fn64; g_reg[i].ctx = ctx; g_reg[i].kind = kind;
            pthread_mutex_unlock(&g_reg_mu);
            return i;
        }
    if (g_nreg == g_capreg) {
        g_capreg = g_capreg ? 2 * g_capreg : 16;
        g_reg = realloc(g_reg, (size_t)g_capreg * sizeof(reg_ent));
    }
    g_reg[g_nreg].name = strdup(name);
    g_reg[g_nreg].fn = fn; g_reg[g_nreg].fn64 = fn64;
    g_reg[g_nreg].ctx = ctx; g_reg[g_nreg].kind = kind;
    int idx = g_nreg++;
    pthread_mutex_unlock(&g_reg_mu);
    return idx;
}

int chirp_register_spectral(const char *name, chirp_spectral_fn fn, void *ctx) {
    return reg_add(name, (void *)fn, NULL, ctx, CHIRP_KIND_SPECTRAL);
}

int chirp_register_spectral_ex(const char *name, chirp_spectral_fn f32, chirp_spectral_fn_f64 f64,
                               void *ctx) {
    return reg_add(name, (void *)f32, (void *)f64, ctx, CHIRP_KIND_SPECTRAL);
}

void chirp_cleanup(void) {
    pthread_mutex_lock(&g_reg_mu);
    for (int i = 0; i < g_nreg; ++i) free(g_reg[i].name);
    free(g_reg);
    g_reg = NULL;
    g_nreg = g_capreg = 0;
    pthread_mutex_unlock(&g_reg_mu);
}

int chirp_count(void) { return g_nreg; }

/* ------------------------------------------------------------------ */
/* a small S-expression reader                                         */
/* ------------------------------------------------------------------ */

typedef struct node {
    int is_list;
    char *atom;
    struct node **kid;
    int nkid;
} node;

static void node_free(node *n) {
    if (!n) return;
    for (int i = 0; i < n->nkid; ++i) node_free(n->kid[i]);
    free(n->kid);
    free(n->atom);
    free(n);
}

static void skip_ws(const char **s) {
    for (;;) {
        while (**s && isspace((unsigned char)**s)) ++*s;
        if (**s == ';') { while (**s && **s != '\n') ++*s; continue; }
        break;
    }
}

static node *parse(const char **s) {
    skip_ws(s);
    if (!**s || **s == ')') return NULL;
    node *n = calloc(1, sizeof *n);
    if (**s == '(') {
        ++*s;
        n->is_list = 1;
        for (;;) {
            skip_ws(s);
            if (**s == ')') { ++*s; return n; }
            if (!**s) { node_free(n); return NULL; }
            node *k = parse(s);
            if (!k) { node_free(n); return NULL; }
            n->kid = realloc(n->kid, (size_t)(n->nkid + 1) * sizeof(node *));
            n->kid[n->nkid++] = k;
        }
    }
    const char *b = *s;
    while (**s && !isspace((unsigned char)**s) && **s != '(' && **s != ')') ++*s;
    n->atom = strndup(b, (size_t)(*s - b));
    return n;
}

static const char *head(const node *n) {
    if (!n) return NULL;
    if (!n->is_list) return n->atom;
    return (n->nkid && !n->kid[0]->is_list) ? n->kid[0]->atom : NULL;
}

static const node *kwarg(const node *n, const char *key) {
    if (!n || !n->is_list) return NULL;
    for (int i = 1; i + 1 < n->nkid; ++i)
        if (!n->kid[i]->is_list && n->kid[i]->atom[0] == ':' && !strcmp(n->kid[i]->atom + 1, key))
            return n->kid[i + 1];
    return NULL;
}

static int kw_long(const node *n, const char *key, long *out) {
    const node *v = kwarg(n, key);
    if (!v || v->is_list) return -1;
    char *e = NULL;
    long x = strtol(v->atom, &e, 10);
    if (!e || *e) return -1;
    *out = x;
    return 0;
}

/* ------------------------------------------------------------------ */
/* pipeline transforms                                                 */
/* ------------------------------------------------------------------ */

faf_transform *faf_create_pipeline(const faf_config *cfg, int nops, const int *ops,
                                   const uint32_t *a1, const uint32_t *a2) {
    if (!cfg) { fafi_err("pipeline: cfg is NULL"); return NULL; }
    if (nops < 0 || nops > FAFD_MAX_OPS) {
        fafi_err("pipeline: at most %d spectral ops", FAFD_MAX_OPS);
        return NULL;
    }
    for (int i = 0; i < nops; ++i)
        if (ops[i] < FAFD_OP_CONJ || ops[i] > FAFD_OP_BANDPASS) {
            fafi_err("pipeline: unknown op code %d", ops[i]);
            return NULL;
        }
    faf_config fc = *cfg;
    fc.dir = FAF_DIR_FORWARD;
    fc.layout = FAF_LAYOUT_HERMITIAN;
    faf_transform *fwd = faf_create_rfft(&fc);
    if (!fwd) return NULL;
    faf_config ic = faf_config_inverse(fwd->cfg);
    faf_transform *inv = faf_create_rfft(&ic);
    if (!inv) { faf_destroy_transform(fwd); return NULL; }
    faf_transform *t = calloc(1, sizeof(faf_transform));
    fafh_plan *p = fafi_plan_new();
    if (!t || !p) {
        free(t); free(p);
        faf_destroy_transform(fwd); faf_destroy_transform(inv);
        fafi_err("pipeline: alloc failed");
        return NULL;
    }
    t->amd = p;
    t->type = FAF_TRANSFORM_PIPELINE;
    t->n = fwd->n;
    t->precision = fwd->precision;
    t->inner = fwd;
    t->inner_inv = inv;
    t->cfg = fwd->cfg;
    t->cfg.layout = FAF_LAYOUT_REAL;
    t->cfg.dir = FAF_DIR_INVERSE;
    t->flags = FAF_FLAG_REAL;
    if (fafi_plan_pipe(p, t->n)) { faf_destroy_transform(t); return NULL; }
    p->ops.nops = nops;
    for (int i = 0; i < nops; ++i) {
        p->ops.code[i] = ops[i];
        p->ops.a1[i] = a1 ? a1[i] : 0;
        p->ops.a2[i] = a2 ? a2[i] : 0;
    }
    /* the forward and inverse norms come from one config, so their product
     * is 1/n for NONE (1 * 1/n), ORTHO (1/sqrt(n) twice) and FORWARD
     * (1/n * 1); every op is linear, so the product is applied once */
    p->scale = (float)(1.0 / (double)t->n);
    /* keep an op list in t->code like the reference (CALL_BUILTIN, a0 = op) */
    if (nops) {
        t->code = calloc((size_t)nops, sizeof(faf_inst));
        t->n_inst = (size_t)nops;
        for (int i = 0; i < nops; ++i) {
            t->code[i].packed = 30;   /* FAF_CALL_BUILTIN */
            t->code[i].a0 = ops[i] == FAFD_OP_CONJ ? CHIRP_OP_CONJ
                            : ops[i] == FAFD_OP_MUL ? CHIRP_OP_MUL : CHIRP_OP_BANDPASS;
            t->code[i].a1 = p->ops.a1[i];
            t->code[i].a2 = p->ops.a2[i];
        }
    }
    return t;
}

int chirp_bind(faf_transform *t, const char *name, void *re, void *im, size_t n_bins) {
    /* reference src/chirp.c:371-386: the bound spectrum is borrowed, so it is
     * read again at every execute */
    if (!t || !name || !re || !im || !t->amd) { fafi_err("chirp_bind: bad argument"); return -1; }
    if (t->type != FAF_TRANSFORM_PIPELINE) { fafi_err("chirp_bind: not a pipeline"); return -1; }
    (void)name;
    fafh_plan *p = t->amd;
    t->user_aux = re;
    t->user_aux_im = im;
    t->user_aux_n = n_bins;
    p->h_host_re = re;
    p->h_host_im = im;
    p->h_host_bins = n_bins;
    return 0;
}

/* refresh the device copy of the bound spectrum (called before a launch) */
int fafi_pipe_refresh(fafh_plan *p, void *stream) {
    p->ops.h_re = p->ops.h_im = NULL;
    p->ops.h_bins = 0;
    if (!p->h_host_re) return 0;
    size_t nb = p->h_host_bins, cap = nb * 2 * sizeof(float);
    if (fafk_is_device_ptr(p->h_host_re) && fafk_is_device_ptr(p->h_host_im)) {
        p->ops.h_re = p->h_host_re;
        p->ops.h_im = p->h_host_im;
    } else {
        if (fafi_capturing()) {
            fafi_err("a graph capture needs the bound spectrum in device memory");
            return -1;
        }
        if (fafi_noalloc()) {
            fafi_err("faf_execute_batch_noalloc: the bound spectrum must be device memory");
            return -1;
        }
        if (p->h_dev_cap < cap) {
            fafk_free(p->h_dev);
            p->h_dev = fafk_malloc(cap);
            p->h_dev_cap = p->h_dev ? cap : 0;
            if (!p->h_dev) { fafi_err("%s", fafk_last_error()); return -1; }
        }
        float *d = p->h_dev;
        if (fafk_h2d(d, p->h_host_re, nb * sizeof(float), stream) ||
            fafk_h2d(d + nb, p->h_host_im, nb * sizeof(float), stream)) {
            fafi_err("%s", fafk_last_error());
            return -1;
        }
        p->ops.h_re = d;
        p->ops.h_im = d + nb;
    }
    p->ops.h_bins = (int64_t)nb;
    return 0;
}

faf_transform *chirp_compile(const char *src) {
    if (!src) return NULL;
    const char *s = src;
    node *ast = parse(&s);
    faf_transform *t = NULL;
    if (!ast || !ast->is_list || !head(ast) || strcmp(head(ast), "pipeline")) {
        fafi_err("Chirp: only (pipeline (rfft ...) ops ... (irfft)) is supported on the "
                 "device path");
        goto done;
    }
    if (ast->nkid < 3) { fafi_err("Chirp: pipeline needs (rfft ...) and (irfft)"); goto done; }
    const node *rf = ast->kid[1];
    if (!rf->is_list || strcmp(head(rf) ? head(rf) : "", "rfft")) {
        fafi_err("Chirp: pipeline must start with (rfft :size N)");
        goto done;
    }
    long n = 0;
    if (kw_long(rf, "size", &n) || n < 4) { fafi_err("Chirp: (rfft) needs :size >= 4"); goto done; }
    faf_config c = faf_config_init((size_t)n);
    const node *nv = kwarg(rf, "norm");
    if (nv && !nv->is_list) {
        if (!strcmp(nv->atom, "ortho")) c.norm = FAF_NORM_ORTHO;
        else if (!strcmp(nv->atom, "forward")) c.norm = FAF_NORM_FORWARD;
        else if (!strcmp(nv->atom, "none") || !strcmp(nv->atom, "backward")) c.norm = FAF_NORM_NONE;
        else { fafi_err("Chirp: unknown :norm %s", nv->atom); goto done; }
    }
    const node *tail = ast->kid[ast->nkid - 1];
    const char *th = head(tail);
    if (!th || (strcmp(th, "irfft") && strcmp(th, "inverse"))) {
        fafi_err("Chirp: the device pipeline must end with (irfft)");
        goto done;
    }
    int ops[FAFD_MAX_OPS];
    uint32_t a1[FAFD_MAX_OPS], a2[FAFD_MAX_OPS];
    int nops = 0;
    for (int i = 2; i < ast->nkid - 1; ++i) {
        const node *op = ast->kid[i];
        const char *h = head(op);
        if (nops == FAFD_MAX_OPS) { fafi_err("Chirp: too many spectral ops"); goto done; }
        a1[nops] = a2[nops] = 0;
        if (h && !strcmp(h, "conj")) ops[nops++] = FAFD_OP_CONJ;
        else if (h && !strcmp(h, "mul-spectrum")) ops[nops++] = FAFD_OP_MUL;
        else if (h && !strcmp(h, "bandpass")) {
            long lo = 0, hi = 0;
            kw_long(op, "lo", &lo);
            kw_long(op, "hi", &hi);
            a1[nops] = (uint32_t)lo;
            a2[nops] = (uint32_t)hi;
            ops[nops++] = FAFD_OP_BANDPASS;
        } else {
            fafi_err("Chirp: spectral op '%s' runs host code and cannot be fused into the "
                     "device pipeline", h ? h : "?");
            goto done;
        }
    }
    t = faf_create_pipeline(&c, nops, ops, a1, a2);
done:
    node_free(ast);
    return t;
}
