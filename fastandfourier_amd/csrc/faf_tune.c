/*
 * faf_tune.c -- the autotuner (C11).
 *
 * For a plan it times, on synthetic device data, every candidate launch
 * shape: the order of the radix passes (the default order and its
 * rotations, plus splitting each radix-16 pass into 4 x 4) and the number
 * of signals per workgroup.  The fastest shape is kept in the plan and in a
 * process-wide table keyed by (kind, length), the MI355X counterpart of the
 * reference's FAF_FLAG_MEASURE (src/faf_core.c:504-565), which times
 * split-radix against iterative DIT and keeps the faster.
 */
#include "faf_internal.h"
#include <stdlib.h>
#include <string.h>

#define MAX_CAND_ORDERS 8

static int order_of(const fafd_fft *P, int *rad) {
    for (int s = 0; s < P->nstages; ++s) rad[s] = P->st[s].radix;
    return P->nstages;
}

static int cand_orders(const fafd_fft *P, int orders[][FAFD_MAX_STAGES], int *lens) {
    int base[FAFD_MAX_STAGES], k = order_of(P, base), nc = 0;
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

static float time_once(fafh_plan *p, void *x, void *y, long count, void *stream) {
    void *e0 = fafk_event_create(), *e1 = fafk_event_create();
    float *xf = x, *yf = y;
    size_t n = p->n, nb = n / 2 + 1;
    int rc = 0;
    fafk_event_record(e0, stream);
    for (int rep = 0; rep < 3 && !rc; ++rep) {
        switch (p->kind) {
        case PK_C2C:
            rc = fafi_run_c2c(p, xf, xf + count * n, (long)n, yf, yf + count * n, (long)n,
                              count, stream);
            break;
        case PK_R2C:
            rc = fafi_run_r2c(p, xf, (long)n, yf, yf + count * nb, (long)nb, count, stream);
            break;
        case PK_C2R:
            rc = fafi_run_c2r(p, xf, xf + count * nb, (long)nb, yf, (long)n, count, stream);
            break;
        case PK_PIPE:
            rc = fafi_run_pipe(p, xf, (long)n, yf, (long)n, count, stream);
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

int faf_amd_autotune(faf_transform *t, size_t count) {
    if (!t || !t->amd) { fafi_err("autotune: no transform"); return -1; }
    if (fafi_create_ready()) return -1;
    fafh_plan *p = t->amd;
    if (p->kind != PK_C2C && p->kind != PK_R2C && p->kind != PK_C2R && p->kind != PK_PIPE) {
        fafi_err("autotune: plan kind has a fixed launch shape");
        return -1;
    }
    if (count == 0) count = 1024;
    void *stream = fafi_stream();
    size_t n = p->n, bytes = count * (n + 2) * 2 * sizeof(float);
    void *x = fafk_malloc(bytes), *y = fafk_malloc(bytes);
    if (!x || !y) {
        fafk_free(x); fafk_free(y);
        fafi_err("autotune: %s", fafk_last_error());
        return -1;
    }
    /* zeros are fine for timing: the kernels have no data-dependent paths */
    float *zero = calloc(1, bytes);
    fafk_h2d(x, zero, bytes, stream);
    fafk_sync(stream);
    free(zero);

    int inv = p->kind == PK_C2R ? 1 : p->inverse;
    size_t fft_n = (p->kind == PK_C2C) ? n : n / 2;
    fafd_fft best_a = p->a, best_b = p->b;
    int best_spb = p->spb, best_ct = 0;
    float best = 1e30f;
    if (p->ct) {               /* the specialised kernel is one more candidate */
        fafh_plan trial = *p;
        time_once(&trial, x, y, (long)count, stream);
        best = time_once(&trial, x, y, (long)count, stream);
        if (best <= 0.0f) best = 1e30f;
        else best_ct = 1;
    }
    int orders[MAX_CAND_ORDERS][FAFD_MAX_STAGES], lens[MAX_CAND_ORDERS];
    int no = cand_orders(&p->a, orders, lens);
    for (int o = 0; o < no; ++o) {
        fafd_fft a, b;
        if (fafi_make_fft(fft_n, inv, 1, orders[o], lens[o], &a)) continue;
        b = p->b;
        if (p->kind == PK_PIPE && fafi_make_fft(fft_n, 1, 1, orders[o], lens[o], &b)) continue;
        for (int spb = 1; spb <= 64; spb *= 2) {
            if (spb * a.threads > 1024) break;
            if ((size_t)spb * a.plane * 8 > FAFD_MAX_LDS_BYTES) break;
            fafh_plan trial = *p;
            trial.ct = 0;
            trial.a = a;
            trial.b = b;
            trial.spb = spb;
            time_once(&trial, x, y, (long)count, stream);          /* warm */
            float ms = time_once(&trial, x, y, (long)count, stream);
            if (ms > 0.0f && ms < best) {
                best = ms;
                best_a = a;
                best_b = b;
                best_spb = spb;
                best_ct = 0;
            }
        }
    }
    fafk_free(x);
    fafk_free(y);
    if (best >= 1e30f) { fafi_err("autotune: no candidate ran"); return -1; }
    if (!best_ct) {
        p->a = best_a;
        p->b = best_b;
        p->spb = best_spb;
    }
    p->ct = best_ct;
    double gbps = (double)count * (double)n * 8.0 * 3.0 / (best * 1e-3) / 1e9;
    fafi_tune_record(p->kind, n, best_spb, gbps);
    return best_spb;
}
