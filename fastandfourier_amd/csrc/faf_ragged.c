/*
 * faf_ragged.c -- many C2C signals of uneven lengths in one persistent
 * launch (C11 host side).
 *
 * Signals are bucketed by length; each bucket has one device plan
 * (Stockham, Rader or Bluestein) in a device-resident plan table.  Work is
 * cut into units of up to `spb` same-length signals (one workgroup's worth)
 * and the units are ordered largest-first; persistent workgroups pull units
 * from an atomic counter until the queue is empty, so lengths from 11 to
 * 8192 share one launch without wave divergence inside a unit.
 */
#include "faf_internal.h"
#include <stdlib.h>
#include <string.h>

#define RAGGED_BLOCK 512
#define RAGGED_LDS_CAP (96 * 1024)
/* with compile-time paths only, two workgroups share a CU's 160 KB */
#define RAGGED_LDS_CAP_CT (79 * 1024)

struct faf_ragged {
    size_t count, total;
    int nplans;
    fafh_plan **hp;           /* host plans own the device tables */
    void *d_plans, *d_units, *d_order, *d_off, *d_counter;
    int nunits;
    int generic;              /* some length needs the runtime paths */
    int grid;
    size_t lds;
};

static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : x > y;
}

typedef struct { fafd_unit u; double cost; } unit_c;

static int cmp_cost(const void *a, const void *b) {
    double x = ((const unit_c *)a)->cost, y = ((const unit_c *)b)->cost;
    return x > y ? -1 : x < y;
}

static void *up(const void *h, size_t bytes) {
    void *d = fafk_malloc(bytes ? bytes : 1);
    if (!d) return NULL;
    if (bytes && (fafk_h2d(d, h, bytes, NULL) || fafk_sync(NULL))) { fafk_free(d); return NULL; }
    return d;
}

void faf_ragged_destroy(faf_ragged *r) {
    if (!r) return;
    for (int i = 0; i < r->nplans; ++i) fafi_plan_free(r->hp[i]);
    free(r->hp);
    fafk_free(r->d_plans); fafk_free(r->d_units); fafk_free(r->d_order);
    fafk_free(r->d_off); fafk_free(r->d_counter);
    free(r);
}

size_t faf_ragged_total(const faf_ragged *r) { return r ? r->total : 0; }

faf_ragged *faf_ragged_create(const size_t *lengths, size_t count, uint32_t flags) {
    if (!lengths || count == 0 || count > 0x7fffffff) {
        fafi_err("ragged: need 1..2^31-1 lengths");
        return NULL;
    }
    if (fafi_create_ready()) return NULL;
    faf_ragged *r = calloc(1, sizeof *r);
    r->count = count;
    /* distinct lengths */
    size_t *uniq = malloc(count * sizeof(size_t));
    memcpy(uniq, lengths, count * sizeof(size_t));
    qsort(uniq, count, sizeof(size_t), cmp_size);
    int nu = 0;
    for (size_t i = 0; i < count; ++i)
        if (i == 0 || uniq[i] != uniq[i - 1]) uniq[nu++] = uniq[i];
    r->hp = calloc((size_t)nu, sizeof(fafh_plan *));
    fafd_rplan *tab = calloc((size_t)nu, sizeof(fafd_rplan));
    size_t *pers = calloc((size_t)nu, sizeof(size_t));
    int ok = 1;
    size_t lds_max = 0;
    for (int k = 0; k < nu && ok; ++k) {
        fafh_plan *p = fafi_plan_new();
        r->hp[r->nplans++] = p;
        if (fafi_plan_c2c(p, uniq[k], 0, (flags & FAF_FLAG_BLUESTEIN) != 0)) { ok = 0; break; }
        fafd_rplan *d = &tab[k];
        d->len = (int32_t)uniq[k];
        size_t per;
        if (p->kind == PK_C2C) {
            /* lengths with compile-time passes in the ragged kernel; the
             * 5-smooth ones share the runtime plan's thread count */
            const size_t n = uniq[k];
            const int ct_t = fafi_ct_enabled() && n <= 8192 ? fafk_rg_ct_threads((int)n) : 0;
            const int pow2 = !(n & (n - 1));
            const int ct = ct_t > 0 && (pow2 || ct_t == p->a.threads);
            d->kind = ct ? FAFD_KIND_CT : FAFD_KIND_STOCKHAM;
            d->fwd = p->a;
            if (ct) d->fwd.threads = ct_t;
            per = (size_t)p->a.plane;
            if (ct && fafk_rg_ct_slots((int)n) > 0) per = (size_t)fafk_rg_ct_slots((int)n);
        } else if (p->kind == PK_RADER) {
            const int rt = fafi_ct_enabled() ? fafk_rg_rader_threads((int)uniq[k]) : 0;
            d->kind = rt > 0 ? FAFD_KIND_RADER_CT : FAFD_KIND_RADER;
            d->fwd = p->a; d->inv = p->b; d->rader = p->rader;
            if (rt > 0) d->fwd.threads = rt;
            per = (size_t)p->a.plane + uniq[k] + (uniq[k] >> 4) + 1;
        } else if (p->kind == PK_BLUE) {
            d->kind = FAFD_KIND_BLUESTEIN;
            d->fwd = p->a; d->inv = p->b; d->blue = p->blue;
            per = (size_t)p->a.plane;
        } else {
            fafi_err("ragged: length %zu exceeds one workgroup", uniq[k]);
            ok = 0;
            break;
        }
        d->threads = d->fwd.threads;
        if (d->threads > RAGGED_BLOCK) {
            fafi_err("ragged: length %zu needs %d threads", uniq[k], d->threads);
            ok = 0;
            break;
        }
        if (d->kind != FAFD_KIND_CT && d->kind != FAFD_KIND_RADER_CT) r->generic = 1;
        pers[k] = per;
    }
    /* signals per unit: as many as the threads and the LDS cap allow */
    const size_t cap = r->generic ? RAGGED_LDS_CAP : RAGGED_LDS_CAP_CT;
    for (int k = 0; k < nu && ok; ++k) {
        fafd_rplan *d = &tab[k];
        const size_t per = pers[k];
        int spb = RAGGED_BLOCK / d->threads;
        while (spb > 1 && (size_t)spb * per * 8 > cap) --spb;
        if ((size_t)spb * per * 8 > FAFD_MAX_LDS_BYTES) {
            fafi_err("ragged: length %zu does not fit LDS", uniq[k]);
            ok = 0;
            break;
        }
        d->spb = spb;
        if ((size_t)spb * per * 8 > lds_max) lds_max = (size_t)spb * per * 8;
    }
    free(pers);
    if (!ok) { free(uniq); free(tab); faf_ragged_destroy(r); return NULL; }

    /* order signals by plan, offsets as exclusive prefix sums */
    int32_t *order = malloc(count * sizeof(int32_t));
    long *off = malloc(count * sizeof(long));
    int *pid = malloc(count * sizeof(int));
    int *nper = calloc((size_t)nu, sizeof(int)), *start = calloc((size_t)nu + 1, sizeof(int));
    size_t acc = 0;
    for (size_t i = 0; i < count; ++i) {
        off[i] = (long)acc;
        acc += lengths[i];
        int lo = 0, hi = nu - 1;
        while (lo < hi) {            /* binary search in the sorted distinct list */
            int mid = (lo + hi) / 2;
            if (uniq[mid] < lengths[i]) lo = mid + 1; else hi = mid;
        }
        pid[i] = lo;
        nper[lo]++;
    }
    r->total = acc;
    for (int k = 0; k < nu; ++k) start[k + 1] = start[k] + nper[k];
    int *fill = calloc((size_t)nu, sizeof(int));
    for (size_t i = 0; i < count; ++i) order[start[pid[i]] + fill[pid[i]]++] = (int32_t)i;

    /* units, largest estimated cost first */
    int maxu = 0;
    for (int k = 0; k < nu; ++k) maxu += (nper[k] + tab[k].spb - 1) / tab[k].spb;
    unit_c *uc = malloc((size_t)maxu * sizeof(unit_c));
    int nuu = 0;
    for (int k = 0; k < nu; ++k) {
        double lg = 1.0;
        for (size_t m = uniq[k]; m > 1; m >>= 1) lg += 1.0;
        const int stockham = tab[k].kind == FAFD_KIND_STOCKHAM || tab[k].kind == FAFD_KIND_CT;
        double per_sig = (double)uniq[k] * lg * (stockham ? 1.0 : 2.5);
        for (int s = 0; s < nper[k]; s += tab[k].spb) {
            int c = nper[k] - s < tab[k].spb ? nper[k] - s : tab[k].spb;
            uc[nuu].u.plan = k;
            uc[nuu].u.count = c;
            uc[nuu].u.first = start[k] + s;
            uc[nuu].u.pad = 0;
            uc[nuu].cost = per_sig * c;
            ++nuu;
        }
    }
    qsort(uc, (size_t)nuu, sizeof(unit_c), cmp_cost);
    fafd_unit *units = malloc((size_t)nuu * sizeof(fafd_unit));
    for (int i = 0; i < nuu; ++i) units[i] = uc[i].u;
    r->nunits = nuu;
    r->lds = lds_max;
    long g = 4 * (long)fafi_cus();
    r->grid = (int)(nuu < g ? nuu : g);
    if (r->grid < 1) r->grid = 1;

    r->d_plans = up(tab, (size_t)nu * sizeof(fafd_rplan));
    r->d_units = up(units, (size_t)nuu * sizeof(fafd_unit));
    r->d_order = up(order, count * sizeof(int32_t));
    r->d_off = up(off, count * sizeof(long));
    r->d_counter = fafk_malloc(256);
    free(uniq); free(tab); free(order); free(off); free(pid); free(nper); free(start);
    free(fill); free(uc); free(units);
    if (!r->d_plans || !r->d_units || !r->d_order || !r->d_off || !r->d_counter) {
        fafi_err("ragged: %s", fafk_last_error());
        faf_ragged_destroy(r);
        return NULL;
    }
    return r;
}

int faf_ragged_execute(faf_ragged *r, float *out_re, float *out_im, const float *in_re,
                       const float *in_im) {
    if (!r || !out_re || !out_im || !in_re || !in_im) {
        fafi_err("ragged: null argument");
        return -1;
    }
    if (fafi_runtime_ready()) return -1;
    void *stream = fafi_stream();
    int dev = fafk_is_device_ptr(in_re) && fafk_is_device_ptr(in_im) &&
              fafk_is_device_ptr(out_re) && fafk_is_device_ptr(out_im);
    if (!dev) {
        fafi_err("ragged: buffers must be device memory");
        return -1;
    }
    int rc = fafk_ragged(r->d_plans, r->d_units, r->nunits, r->d_order, r->d_off, in_re, in_im,
                         out_re, out_im, r->d_counter, r->grid, RAGGED_BLOCK, r->lds, r->generic,
                         stream);
    if (!rc && fafi_sync_mode()) rc = fafk_sync(stream);
    if (rc) fafi_err("ragged: %s", fafk_rg_last_error());
    return rc;
}
