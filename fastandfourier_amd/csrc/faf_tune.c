/*
 * faf_tune.c -- the plan autotuner (C11).
 *
 * For a plan it times, on synthetic device data, every candidate launch
 * shape: the specialised kernel (compile-time passes or the single-pass
 * register-resident core) when the length has one, and the runtime radix
 * kernels in every candidate radix order (the default order, its rotations,
 * each radix-16 pass split into 4 x 4) with 1..64 signals per workgroup.
 * Each candidate is warmed once and timed as the median of 5 runs.  The
 * fastest shape -- specialised or not, radix order, signals per workgroup --
 * is applied to the plan and recorded for the process under the plan's key
 * (kind, length, direction, interleaved layout), and every plan of that key
 * created later starts from it.  This is the MI355X counterpart of the
 * reference's FAF_FLAG_MEASURE (src/faf_core.c:504-565, called at :694),
 * which times split-radix against iterative DIT at create time and keeps
 * the faster; faf_create_fft / faf_create_rfft honour that flag by running
 * this tuner when the key has no recorded shape yet.
 */
#include "faf_internal.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MAX_CAND_ORDERS 8
#define MAX_SHAPES 512

/* ---- the process-wide record ---- */

static fafi_shape g_shapes[MAX_SHAPES];
static int g_nshapes;
static pthread_mutex_t g_shape_mu = PTHREAD_MUTEX_INITIALIZER;

static int same_key(const fafi_shape *s, int kind, size_t n, int inverse, int interleaved) {
    return s->kind == kind && s->n == n && s->inverse == inverse && s->interleaved == interleaved;
}

int fafi_tune_find(int kind, size_t n, int inverse, int interleaved, fafi_shape *out) {
    int hit = 0;
    pthread_mutex_lock(&g_shape_mu);
    for (int i = 0; i < g_nshapes; ++i)
        if (same_key(&g_shapes[i], kind, n, inverse, interleaved)) {
            if (out) *out = g_shapes[i];
            hit = 1;
            break;
        }
    pthread_mutex_unlock(&g_shape_mu);
    return hit;
}

static void tune_store(const fafi_shape *s) {
    pthread_mutex_lock(&g_shape_mu);
    int i;
    for (i = 0; i < g_nshapes; ++i)
        if (same_key(&g_shapes[i], s->kind, s->n, s->inverse, s->interleaved)) break;
    if (i == g_nshapes && g_nshapes < MAX_SHAPES) ++g_nshapes;
    if (i < MAX_SHAPES) g_shapes[i] = *s;
    pthread_mutex_unlock(&g_shape_mu);
}

static int tunable(const fafh_plan *p) {
    return p->kind == PK_C2C || p->kind == PK_R2C || p->kind == PK_C2R || p->kind == PK_PIPE;
}

/* inner FFT length of a plan */
static size_t fft_len(const fafh_plan *p) { return p->kind == PK_C2C ? p->n : p->n / 2; }

/* builds the runtime-kernel records of p for a radix order */
static int shape_plans(const fafh_plan *p, const int *order, int norder, fafd_fft *a,
                       fafd_fft *b) {
    const size_t m = fft_len(p);
    /* the inner transform of C2C / C2R runs in the plan's direction, the
     * rfft and the pipeline's first half forward */
    if (fafi_make_fft(m, p->inverse, 1, order, norder, a)) return -1;
    *b = p->b;
    if (p->kind == PK_PIPE && fafi_make_fft(m, 1, 1, order, norder, b)) return -1;
    return 0;
}

int fafi_apply_tuned(fafh_plan *p) {
    if (!tunable(p)) return 0;
    fafi_shape s;
    if (!fafi_tune_find(p->kind, p->n, p->inverse, p->interleaved, &s)) return 0;
    if (s.specialised) {
        /* only meaningful while the specialised kernel is enabled */
        if (p->ct || p->rs) p->spb = s.spb > 0 ? s.spb : p->spb;
        return 0;
    }
    fafd_fft a, b;
    if (shape_plans(p, s.order, s.norder, &a, &b)) return -1;
    p->a = a;
    p->b = b;
    p->ct = 0;
    p->rs = 0;
    p->spb = s.spb;
    return 0;
}

/* ---- candidates ---- */

static int cand_orders(const fafd_fft *P, int orders[][FAFD_MAX_STAGES], int *lens) {
    int base[FAFD_MAX_STAGES], k = P->nstages, nc = 0;
    for (int s = 0; s < k; ++s) base[s] = P->st[s].radix;
    /* rotations of the default order */
    for (int r = 0; r < k && nc < MAX_CAND_ORDERS - 1; ++r) {
        for (int s = 0; s < k; ++s) orders[nc][s] = base[(s + r) % k];
        lens[nc++] = k;
    }
    /* every 16 split into 4 x 4 */
    int m = 0, split[FAFD_MAX_STAGES];
    for (int s = 0; s < k && m < FAFD_MAX_STAGES - 1; ++s) {
        if (base[s] == 16) { split[m++] = 4; split[m++] = 4; }
        else split[m++] = base[s];
    }
    if (m != k && m <= FAFD_MAX_STAGES) {
        memcpy(orders[nc], split, (size_t)m * sizeof(int));
        lens[nc++] = m;
    }
    return nc;
}

/* one run of the tuning batch: 3 back-to-back launches, ms, or -1 */
static float run_once(fafh_plan *p, float *x, float *y, long count, void *stream) {
    void *e0 = fafk_event_create(), *e1 = fafk_event_create();
    const size_t n = p->n, nb = n / 2 + 1;
    const int inter = p->interleaved;
    int rc = 0;
    fafk_event_record(e0, stream);
    for (int rep = 0; rep < 3 && !rc; ++rep) {
        switch (p->kind) {
        case PK_C2C:
            rc = fafi_run_c2c(p, x, inter ? NULL : x + count * n, (long)n, y,
                              inter ? NULL : y + count * n, (long)n, count, stream);
            break;
        case PK_R2C:
            rc = fafi_run_r2c(p, x, (long)n, y, inter ? NULL : y + count * nb, (long)nb, count,
                              stream);
            break;
        case PK_C2R:
            rc = fafi_run_c2r(p, x, inter ? NULL : x + count * nb, (long)nb, y, (long)n, count,
                              stream);
            break;
        case PK_PIPE:
            rc = fafi_run_pipe(p, x, (long)n, y, (long)n, count, stream);
            break;
        default:
            rc = -1;
        }
    }
    fafk_event_record(e1, stream);
    float ms = rc ? -1.0f : (fafk_event_sync(e1), fafk_event_ms(e0, e1));
    fafk_event_destroy(e0);
    fafk_event_destroy(e1);
    return ms;
}

/* median of 5 runs after one warm run */
static float time_median(fafh_plan *p, float *x, float *y, long count, void *stream) {
    if (run_once(p, x, y, count, stream) < 0.0f) return -1.0f;
    float t[5];
    for (int k = 0; k < 5; ++k) {
        t[k] = run_once(p, x, y, count, stream);
        if (t[k] < 0.0f) return -1.0f;
    }
    for (int i = 1; i < 5; ++i)
        for (int j = i; j > 0 && t[j] < t[j - 1]; --j) { float s = t[j]; t[j] = t[j - 1]; t[j - 1] = s; }
    return t[2];
}

int fafi_autotune_plan(fafh_plan *p, size_t count) {
    if (!tunable(p)) {
        fafi_err("autotune: plan kind has a fixed launch shape");
        return -1;
    }
    if (count == 0) count = 1024;
    void *stream = fafi_stream();
    const size_t n = p->n, bytes = count * (n + 2) * 2 * sizeof(float);
    float *x = fafk_malloc(bytes), *y = fafk_malloc(bytes);
    if (!x || !y) {
        fafk_free(x); fafk_free(y);
        fafi_err("autotune: %s", fafk_last_error());
        return -1;
    }
    /* zeros are fine for timing: the kernels have no data-dependent paths */
    float *zero = calloc(1, bytes);
    if (!zero || fafk_h2d(x, zero, bytes, stream) || fafk_sync(stream)) {
        free(zero);
        fafk_free(x); fafk_free(y);
        fafi_err("autotune: %s", fafk_last_error());
        return -1;
    }
    free(zero);

    fafi_shape best;
    memset(&best, 0, sizeof best);
    best.kind = p->kind;
    best.n = n;
    best.inverse = p->inverse;
    best.interleaved = p->interleaved;
    float tbest = 1e30f;
    fafd_fft best_a = p->a, best_b = p->b;
    if (p->ct || p->rs) {          /* the specialised kernel is one candidate */
        fafh_plan trial = *p;
        float ms = time_median(&trial, x, y, (long)count, stream);
        if (ms > 0.0f) {
            tbest = ms;
            best.specialised = 1;
            best.spb = p->spb;
        }
    }
    /* the runtime kernels: every candidate order from the default one */
    fafd_fft def, defb;
    int orders[MAX_CAND_ORDERS][FAFD_MAX_STAGES], lens[MAX_CAND_ORDERS];
    int no = 0;
    if (!shape_plans(p, NULL, 0, &def, &defb)) no = cand_orders(&def, orders, lens);
    for (int o = 0; o < no; ++o) {
        fafd_fft a, b;
        if (shape_plans(p, orders[o], lens[o], &a, &b)) continue;
        for (int spb = 1; spb <= 64; spb *= 2) {
            if (spb * a.threads > 1024) break;
            if ((size_t)spb * a.plane * 8 > FAFD_MAX_LDS_BYTES) break;
            fafh_plan trial = *p;
            trial.ct = 0;
            trial.rs = 0;
            trial.a = a;
            trial.b = b;
            trial.spb = spb;
            float ms = time_median(&trial, x, y, (long)count, stream);
            if (ms > 0.0f && ms < tbest) {
                tbest = ms;
                best.specialised = 0;
                best.norder = lens[o];
                memcpy(best.order, orders[o], (size_t)lens[o] * sizeof(int));
                best.spb = spb;
                best_a = a;
                best_b = b;
            }
        }
    }
    fafk_free(x);
    fafk_free(y);
    if (tbest >= 1e30f) { fafi_err("autotune: no candidate ran"); return -1; }
    best.ms = tbest;
    if (!best.specialised) {
        p->a = best_a;
        p->b = best_b;
        p->ct = 0;
        p->rs = 0;
        p->spb = best.spb;
    }
    tune_store(&best);
    return best.spb;
}

int faf_amd_autotune(faf_transform *t, size_t count) {
    if (!t || !t->amd) { fafi_err("autotune: no transform"); return -1; }
    if (fafi_create_ready()) return -1;
    return fafi_autotune_plan(t->amd, count);
}

int faf_amd_plan_shape(const faf_transform *t, int *specialised, int *order, int *norder,
                       int *spb) {
    if (!t || !t->amd) { fafi_err("plan shape: no transform"); return -1; }
    const fafh_plan *p = t->amd;
    if (!tunable(p)) { fafi_err("plan shape: plan kind has a fixed launch shape"); return -1; }
    if (specialised) *specialised = p->ct || p->rs;
    if (norder) *norder = p->a.nstages;
    if (order)
        for (int s = 0; s < p->a.nstages; ++s) order[s] = p->a.st[s].radix;
    if (spb) *spb = p->spb;
    return 0;
}
