/*
 * faf_ragged.c -- many C2C signals of uneven lengths in one launch (C11
 * host side), and the autotuner of that launch.
 *
 * Signals are bucketed by length; each bucket has one device plan
 * (Stockham, Rader or Bluestein) in a device-resident plan table.  Work is
 * cut into units of up to `spb` same-length signals (one workgroup's worth)
 * and the units are ordered largest-first, so lengths from 11 to 8192 share
 * one launch without wave divergence inside a unit.  The launch either
 * gives every unit a workgroup of its own (dispatched in that order), or
 * runs persistent workgroups that pull units from an atomic counter until
 * the queue is empty; the autotuner times both.
 *
 * How each length runs -- compile-time passes or a runtime radix plan (and
 * which radix order), and how many signals a unit carries -- is a per-length
 * choice.  By default it is the planner's; faf_ragged_autotune times the
 * candidates of every length of a batch on the device, keeps the fastest
 * per length (and whether the whole launch is faster without the runtime
 * paths compiled in), and records the winners process-wide so that later
 * ragged plans containing those lengths use them.
 */
#include "faf_internal.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* workgroup sizes of the launch: 256 threads when every length of the
 * batch runs on that many (one 8192-point signal per workgroup instead of
 * two tied to the barriers of one; each length class measured as fast or
 * faster, the C3 mix 170 -> 154 us), else 512 */
#define RAGGED_BLOCK 512
#define RAGGED_BLOCK_SMALL 256
/* a unit's signals lie within this many elements of its first */
#define RAGGED_SPAN (1L << 30)
#define RAGGED_LDS_CAP (96 * 1024)
/* with compile-time paths only, two workgroups share a CU's 160 KB */
#define RAGGED_LDS_CAP_CT (80 * 1024 - 256)

/* how one length runs in the launch */
typedef struct rg_choice {
    int kind;                 /* FAFD_KIND_*, or -1: the planner's default */
    int norder;               /* runtime Stockham: radix order (0 = default) */
    int order[FAFD_MAX_STAGES];
    int spb;                  /* signals per unit (0 = as many as fit) */
    int variant;              /* compile-time kinds: pass sequence (-1: the default) */
} rg_choice;

struct faf_ragged {
    size_t count, total;
    size_t *lengths;          /* the caller's lengths, kept for the autotuner */
    uint32_t flags;
    int nplans;
    fafh_plan **hp;           /* host plans own the device tables */
    void *d_plans, *d_units, *d_uoff, *d_counter;
    int nunits;
    int generic;              /* some length needs the runtime paths */
    int persist;              /* 1: persistent workgroups and a queue; 0: one per unit */
    int block_want;           /* 0: 256 threads when every length fits, else 512 */
    int block;                /* threads per workgroup of the launch */
    int grid;
    size_t lds;
};

/* ---- per-length choices kept for the process (autotuner results) ---- */

typedef struct rg_tuned { size_t len; rg_choice c; } rg_tuned;
static rg_tuned g_rg_tuned[256];
static int g_rg_ntuned;
static pthread_mutex_t g_rg_mu = PTHREAD_MUTEX_INITIALIZER;

static int rg_tuned_find(size_t len, rg_choice *out) {
    int hit = 0;
    pthread_mutex_lock(&g_rg_mu);
    for (int i = 0; i < g_rg_ntuned; ++i)
        if (g_rg_tuned[i].len == len) { *out = g_rg_tuned[i].c; hit = 1; break; }
    pthread_mutex_unlock(&g_rg_mu);
    return hit;
}

static void rg_tuned_store(size_t len, const rg_choice *c) {
    pthread_mutex_lock(&g_rg_mu);
    int i;
    for (i = 0; i < g_rg_ntuned; ++i)
        if (g_rg_tuned[i].len == len) break;
    if (i == g_rg_ntuned && g_rg_ntuned < 256) ++g_rg_ntuned;
    if (i < 256) { g_rg_tuned[i].len = len; g_rg_tuned[i].c = *c; }
    pthread_mutex_unlock(&g_rg_mu);
}

static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : x > y;
}

typedef struct { fafd_unit u; double cost; } unit_c;

/* largest cost first; equal costs keep the signal order (units of one
 * length then walk the signals in memory order) */
static int cmp_cost(const void *a, const void *b) {
    const unit_c *p = a, *q = b;
    if (p->cost != q->cost) return p->cost > q->cost ? -1 : 1;
    return (p->u.first > q->u.first) - (p->u.first < q->u.first);
}

static void *up(const void *h, size_t bytes) {
    void *d = fafk_malloc(bytes ? bytes : 1);
    if (!d) return NULL;
    if (bytes && (fafk_h2d(d, h, bytes, NULL) || fafk_sync(NULL))) { fafk_free(d); return NULL; }
    return d;
}

static void ragged_free_device(faf_ragged *r) {
    for (int i = 0; i < r->nplans; ++i) fafi_plan_free(r->hp[i]);
    free(r->hp);
    r->hp = NULL;
    r->nplans = 0;
    fafk_free(r->d_plans); fafk_free(r->d_units); fafk_free(r->d_uoff);
    fafk_free(r->d_counter);
    r->d_plans = r->d_units = r->d_uoff = r->d_counter = NULL;
}

void faf_ragged_destroy(faf_ragged *r) {
    if (!r) return;
    ragged_free_device(r);
    free(r->lengths);
    free(r);
}

size_t faf_ragged_total(const faf_ragged *r) { return r ? r->total : 0; }

/* The device plan of one length under choice c (NULL or kind -1: the
 * planner's default).  Returns 0 and fills d / *per (LDS float2 per signal),
 * or -1 with an error.  A choice that the length cannot take (compile-time
 * passes it has none of) fails, so the autotuner skips it. */
/* The pass sequence a compile-time kind of length n (rader: the n - 1
 * inner points of prime n) runs when no choice names one: the composite
 * radices (variant 1) where the length has them, else the planner's order */
static int rg_default_variant(size_t n, int rader) {
    if (n > 8192) return 0;
    return (rader ? fafk_rg_rader_seq((int)n, 1, NULL, NULL)
                  : fafk_rg_ct_seq((int)n, 1, NULL, NULL)) > 0 ? 1 : 0;
}

/* records radices ord[0..nr) as the stage list of P (for reporting: the
 * compile-time kinds take their passes from the kernel's own sequence) */
static void set_order(fafd_fft *P, const int *ord, int nr) {
    int32_t ns = 1;
    P->nstages = nr;
    for (int s = 0; s < nr; ++s) {
        P->st[s].radix = ord[s];
        P->st[s].ns = ns;
        P->st[s].tw_stride = P->n / (ns * ord[s]);
        ns *= ord[s];
    }
}

static int plan_length(fafh_plan *p, size_t n, uint32_t flags, const rg_choice *c,
                       fafd_rplan *d, size_t *per) {
    if (fafi_plan_c2c(p, n, 0, (flags & FAF_FLAG_BLUESTEIN) != 0)) return -1;
    memset(d, 0, sizeof *d);
    d->len = (int32_t)n;
    d->path = -1;
    const int want = c ? c->kind : -1;
    if (p->kind == PK_C2C) {
        /* the planner's own radix order: variant 0 of the compile-time passes
         * of the 5-smooth lengths follows it (a plan-level autotune may have
         * changed p->a, so it is rebuilt here) */
        fafd_fft def;
        if (fafi_make_fft(n, 0, 1, NULL, 0, &def)) return -1;
        const int var = want == FAFD_KIND_CT && c->variant >= 0 ? c->variant : rg_default_variant(n, 0);
        int ord[FAFD_MAX_STAGES], nr = 0;
        const int ct_t = fafi_ct_enabled() && n <= 8192 ? fafk_rg_ct_seq((int)n, var, ord, &nr) : 0;
        const int pow2 = !(n & (n - 1));
        const int ct_ok = ct_t > 0 && (pow2 || var > 0 || ct_t == def.threads);
        int kind = want >= 0 ? want : (ct_ok ? FAFD_KIND_CT : FAFD_KIND_STOCKHAM);
        if (kind == FAFD_KIND_CT) {
            if (!ct_ok) { fafi_err("ragged: length %zu has no compile-time passes", n); return -1; }
            d->kind = FAFD_KIND_CT;
            d->variant = var;
            d->path = fafk_rg_path(FAFD_KIND_CT, (int)n, var);
            d->fwd = def;
            d->fwd.threads = ct_t;
            if (nr > 0) set_order(&d->fwd, ord, nr);
            *per = (size_t)def.plane;
            if (fafk_rg_ct_slots((int)n) > 0) *per = (size_t)fafk_rg_ct_slots((int)n);
        } else if (kind == FAFD_KIND_STOCKHAM) {
            d->kind = FAFD_KIND_STOCKHAM;
            if (c && c->norder > 0) {
                if (fafi_make_fft(n, 0, 1, c->order, c->norder, &d->fwd)) return -1;
            } else {
                d->fwd = def;
            }
            *per = (size_t)d->fwd.plane;
        } else {
            fafi_err("ragged: length %zu cannot run as kind %d", n, kind);
            return -1;
        }
    } else if (p->kind == PK_RADER) {
        const int var = want == FAFD_KIND_RADER_CT && c->variant >= 0 ? c->variant : rg_default_variant(n, 1);
        int ord[FAFD_MAX_STAGES], nr = 0;
        const int rt = fafi_ct_enabled() ? fafk_rg_rader_seq((int)n, var, ord, &nr) : 0;
        int kind = want >= 0 ? want : (rt > 0 ? FAFD_KIND_RADER_CT : FAFD_KIND_RADER);
        if (kind == FAFD_KIND_RADER_CT && rt <= 0) {
            fafi_err("ragged: prime %zu has no compile-time inner passes", n);
            return -1;
        }
        if (kind != FAFD_KIND_RADER_CT && kind != FAFD_KIND_RADER) {
            fafi_err("ragged: prime %zu cannot run as kind %d", n, kind);
            return -1;
        }
        d->kind = kind;
        d->fwd = p->a; d->inv = p->b; d->rader = p->rader;
        /* the runtime Rader path stages the signal in LDS beside the plane;
         * the compile-time one stages, gathers and scatters in the plane */
        *per = (size_t)p->a.plane + (kind == FAFD_KIND_RADER ? n + (n >> 4) + 1 : 0);
        if (kind == FAFD_KIND_RADER_CT) {
            d->variant = var;
            d->path = fafk_rg_path(FAFD_KIND_RADER_CT, (int)n, var);
            d->fwd.threads = rt;
            if (nr > 0) set_order(&d->fwd, ord, nr);
        }
    } else if (p->kind == PK_BLUE) {
        if (want >= 0 && want != FAFD_KIND_BLUESTEIN) {
            fafi_err("ragged: length %zu runs through Bluestein only", n);
            return -1;
        }
        d->kind = FAFD_KIND_BLUESTEIN;
        d->fwd = p->a; d->inv = p->b; d->blue = p->blue;
        *per = (size_t)p->a.plane;
    } else {
        fafi_err("ragged: length %zu exceeds one workgroup", n);
        return -1;
    }
    d->threads = d->fwd.threads;
    if ((d->kind == FAFD_KIND_CT || d->kind == FAFD_KIND_RADER_CT) && d->path < 0) {
        fafi_err("ragged: length %zu has no compiled path", n);
        return -1;
    }
    if (d->threads > RAGGED_BLOCK) {
        fafi_err("ragged: length %zu needs %d threads", n, d->threads);
        return -1;
    }
    return 0;
}

/* choice of length n: the caller's override table first, then the
 * process-wide autotuned table, else the default */
typedef struct { const size_t *len; const rg_choice *c; int n; } rg_overrides;

static const rg_choice *choice_for(size_t n, const rg_overrides *ov, rg_choice *tmp) {
    if (ov)
        for (int i = 0; i < ov->n; ++i)
            if (ov->len[i] == n) return &ov->c[i];
    if (rg_tuned_find(n, tmp)) return tmp;
    return NULL;
}

/* Workgroups of the ragged kernel a CU holds at once: the lean kernel's
 * 128 VGPRs allow 4 waves per SIMD (16 per CU), the one with the runtime
 * paths 2 (8 per CU); and the CU's 160 KB of LDS, in 512-byte blocks. */
static int ragged_per_cu(int generic, size_t lds, int block) {
    const int by_waves = (generic ? 8 : 16) / (block / 64);
    const size_t blk = (lds + 64 + 511) / 512 * 512;
    const int by_lds = (int)(FAFD_MAX_LDS_BYTES / blk);
    const int n = by_waves < by_lds ? by_waves : by_lds;
    return n < 1 ? 1 : n;
}

/* Signals in the next unit of one length: up to spb of the `left` signals
 * at sig[0..], as long as each ends within RAGGED_SPAN elements of the
 * first one's offset. */
static int unit_len(const long *off, const int32_t *sig, int left, int spb, size_t len) {
    int c = left < spb ? left : spb;
    while (c > 1 && off[sig[c - 1]] + (long)len - off[sig[0]] > RAGGED_SPAN) --c;
    return c;
}

/* Builds the device tables of a ragged batch into r (r->lengths/count set). */
static int ragged_build(faf_ragged *r, const rg_overrides *ov) {
    const size_t count = r->count;
    const size_t *lengths = r->lengths;
    size_t *uniq = malloc(count * sizeof(size_t));
    memcpy(uniq, lengths, count * sizeof(size_t));
    qsort(uniq, count, sizeof(size_t), cmp_size);
    int nu = 0;
    for (size_t i = 0; i < count; ++i)
        if (i == 0 || uniq[i] != uniq[i - 1]) uniq[nu++] = uniq[i];
    r->hp = calloc((size_t)nu, sizeof(fafh_plan *));
    fafd_rplan *tab = calloc((size_t)nu, sizeof(fafd_rplan));
    size_t *pers = calloc((size_t)nu, sizeof(size_t));
    int *want_spb = calloc((size_t)nu, sizeof(int));
    int ok = 1;
    size_t lds_max = 0;
    r->generic = 0;
    for (int k = 0; k < nu && ok; ++k) {
        fafh_plan *p = fafi_plan_new();
        r->hp[r->nplans++] = p;
        rg_choice tmp;
        const rg_choice *c = choice_for(uniq[k], ov, &tmp);
        if (plan_length(p, uniq[k], r->flags, c, &tab[k], &pers[k])) {
            /* a recorded choice this length cannot take falls back to the
             * default; an override from the autotuner's trial must hold */
            if (!c || (ov && c != &tmp)) { ok = 0; break; }
            fafi_plan_free(p);
            r->hp[r->nplans - 1] = p = fafi_plan_new();
            c = NULL;
            if (plan_length(p, uniq[k], r->flags, NULL, &tab[k], &pers[k])) { ok = 0; break; }
        }
        want_spb[k] = c ? c->spb : 0;
        if (tab[k].kind != FAFD_KIND_CT && tab[k].kind != FAFD_KIND_RADER_CT) r->generic = 1;
    }
    /* the workgroup size of the launch */
    int maxt = 0;
    for (int k = 0; k < nu && ok; ++k)
        if (tab[k].threads > maxt) maxt = tab[k].threads;
    const int block = r->block_want ? r->block_want
                                    : (maxt <= RAGGED_BLOCK_SMALL ? RAGGED_BLOCK_SMALL : RAGGED_BLOCK);
    if (ok && maxt > block) {
        fafi_err("ragged: a length needs %d threads, more than a %d-thread workgroup", maxt, block);
        ok = 0;
    }
    r->block = block;
    /* signals per unit: as many as the threads and the LDS cap allow, or
     * fewer when the choice asks */
    const size_t cap = r->generic ? RAGGED_LDS_CAP : RAGGED_LDS_CAP_CT;
    for (int k = 0; k < nu && ok; ++k) {
        fafd_rplan *d = &tab[k];
        const size_t per = pers[k];
        int spb = block / d->threads;
        while (spb > 1 && (size_t)spb * per * 8 > cap) --spb;
        if (want_spb[k] > 0 && want_spb[k] < spb) spb = want_spb[k];
        if ((size_t)spb * per * 8 > FAFD_MAX_LDS_BYTES) {
            fafi_err("ragged: length %zu does not fit LDS", uniq[k]);
            ok = 0;
            break;
        }
        d->spb = spb;
        /* the 8192-point single pass exchanges through LDS in every signal
         * slot of the workgroup, active or not, so its units reserve all */
        const size_t slots = fafk_rg_ct_slots(d->len) > 0 && d->kind == FAFD_KIND_CT
                                 ? (size_t)(block / d->threads) : (size_t)spb;
        if (slots * per * 8 > lds_max) lds_max = slots * per * 8;
    }
    free(pers);
    free(want_spb);
    if (!ok) { free(uniq); free(tab); return -1; }

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

    /* units, largest estimated cost first.  A unit's signals lie within
     * RAGGED_SPAN elements of its first (the kernel addresses them by 32-bit
     * byte offsets from it), so a unit ends early where the next signal of
     * its length lies further away. */
    int maxu = 0;
    for (int k = 0; k < nu; ++k)
        for (int s = 0; s < nper[k]; ++maxu)
            s += unit_len(off, order + start[k] + s, nper[k] - s, tab[k].spb, uniq[k]);
    unit_c *uc = malloc((size_t)maxu * sizeof(unit_c));
    int nuu = 0;
    for (int k = 0; k < nu; ++k) {
        double lg = 1.0;
        for (size_t m = uniq[k]; m > 1; m >>= 1) lg += 1.0;
        const int stockham = tab[k].kind == FAFD_KIND_STOCKHAM || tab[k].kind == FAFD_KIND_CT;
        double per_sig = (double)uniq[k] * lg * (stockham ? 1.0 : 2.5);
        for (int s = 0; s < nper[k];) {
            const long o0 = off[order[start[k] + s]];
            const int c = unit_len(off, order + start[k] + s, nper[k] - s, tab[k].spb, uniq[k]);
            uc[nuu].u.plan = k;
            uc[nuu].u.count = c;
            uc[nuu].u.first = start[k] + s;
            uc[nuu].u.path = tab[k].path;
            uc[nuu].u.threads = tab[k].threads;
            uc[nuu].u.pad = 0;
            uc[nuu].u.off0 = o0;
            uc[nuu].cost = per_sig * c;
            ++nuu;
            s += c;
        }
    }
    qsort(uc, (size_t)nuu, sizeof(unit_c), cmp_cost);
    fafd_unit *units = malloc((size_t)nuu * sizeof(fafd_unit));
    for (int i = 0; i < nuu; ++i) units[i] = uc[i].u;
    r->nunits = nuu;
    r->lds = lds_max;
    /* persistent grid: the workgroups resident at once (each starts on its
     * own unit without a queue round trip), never more than the units */
    int per_cu = ragged_per_cu(r->generic, r->lds, r->block);
    long g = (long)per_cu * (long)fafi_cus();
    r->grid = (int)(nuu < g ? nuu : g);
    if (r->grid < 1) r->grid = 1;

    /* offset of the signal in each slot of the unit list */
    long *uoff = malloc(count * sizeof(long));
    for (size_t i = 0; i < count; ++i) uoff[i] = off[order[i]];
    r->d_plans = up(tab, (size_t)nu * sizeof(fafd_rplan));
    r->d_units = up(units, (size_t)nuu * sizeof(fafd_unit));
    r->d_uoff = up(uoff, count * sizeof(long));
    r->d_counter = fafk_malloc(256);
    free(uniq); free(tab); free(order); free(off); free(pid); free(nper); free(start);
    free(fill); free(uc); free(units); free(uoff);
    if (!r->d_plans || !r->d_units || !r->d_uoff || !r->d_counter) {
        fafi_err("ragged: %s", fafk_last_error());
        return -1;
    }
    if (fafk_ragged_counter_init(r->d_counter, r->grid, fafi_stream())) {
        fafi_err("ragged: %s", fafk_rg_last_error());
        return -1;
    }
    if (fafk_sync(fafi_stream())) {
        fafi_err("ragged: %s", fafk_last_error());
        return -1;
    }
    return 0;
}

static faf_ragged *ragged_new(const size_t *lengths, size_t count, uint32_t flags,
                              const rg_overrides *ov, int block_want) {
    faf_ragged *r = calloc(1, sizeof *r);
    if (!r) { fafi_err("ragged: alloc failed"); return NULL; }
    r->count = count;
    r->flags = flags;
    r->block_want = block_want;
    r->lengths = malloc(count * sizeof(size_t));
    if (!r->lengths) { free(r); fafi_err("ragged: alloc failed"); return NULL; }
    memcpy(r->lengths, lengths, count * sizeof(size_t));
    if (ragged_build(r, ov)) { faf_ragged_destroy(r); return NULL; }
    return r;
}

faf_ragged *faf_ragged_create(const size_t *lengths, size_t count, uint32_t flags) {
    if (!lengths || count == 0 || count > 0x7fffffff) {
        fafi_err("ragged: need 1..2^31-1 lengths");
        return NULL;
    }
    for (size_t i = 0; i < count; ++i)
        if (lengths[i] < 2) { fafi_err("ragged: length %zu is not a transform", lengths[i]); return NULL; }
    if (fafi_create_ready()) return NULL;
    faf_ragged *r = ragged_new(lengths, count, flags, NULL, 0);
    /* the reference's MEASURE flag (src/faf_core.c:539-565, called at :694)
     * asks for timing at create: tune the batch's lengths that have no
     * recorded choice yet */
    if (r && (flags & FAF_FLAG_MEASURE) && faf_ragged_autotune(r) < 0) {
        faf_ragged_destroy(r);
        return NULL;
    }
    return r;
}

static int launch(faf_ragged *r, float *out_re, float *out_im, const float *in_re,
                  const float *in_im, void *stream) {
    int rc = fafk_ragged(r->d_plans, r->d_units, r->nunits, r->d_uoff, in_re, in_im,
                         out_re, out_im, r->d_counter, r->grid, r->block, r->lds,
                         r->generic, r->persist, stream);
    if (rc) fafi_err("ragged: %s", fafk_rg_last_error());
    return rc;
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
    int rc = launch(r, out_re, out_im, in_re, in_im, stream);
    if (!rc && fafi_sync_mode()) rc = fafk_sync(stream);
    return rc;
}

/* ------------------------------------------------------------------ */
/* autotuner of the ragged launch                                       */
/* ------------------------------------------------------------------ */

#define RG_MAX_CANDS 48

/* median launch time (ms) of r over `reps` (odd, at most 9) timed launches
 * after 2 warm ones */
static float time_ragged(faf_ragged *r, float *buf, void *stream, int reps) {
    float *ir = buf, *ii = buf + r->total, *orr = buf + 2 * r->total, *oi = buf + 3 * r->total;
    for (int w = 0; w < 2; ++w)
        if (launch(r, orr, oi, ir, ii, stream)) return -1.0f;
    float t[9];
    if (reps > 9) reps = 9;
    for (int k = 0; k < reps; ++k) {
        void *e0 = fafk_event_create(), *e1 = fafk_event_create();
        fafk_event_record(e0, stream);
        int rc = launch(r, orr, oi, ir, ii, stream);
        fafk_event_record(e1, stream);
        t[k] = rc ? -1.0f : (fafk_event_sync(e1), fafk_event_ms(e0, e1));
        fafk_event_destroy(e0);
        fafk_event_destroy(e1);
        if (t[k] < 0.0f) return -1.0f;
    }
    for (int i = 1; i < reps; ++i)      /* insertion sort, then the median */
        for (int j = i; j > 0 && t[j] < t[j - 1]; --j) { float s = t[j]; t[j] = t[j - 1]; t[j - 1] = s; }
    return t[reps / 2];
}

/* candidate choices of one length */
static int candidates(size_t n, uint32_t flags, rg_choice *out) {
    int nc = 0;
    fafh_plan *p = fafi_plan_new();
    if (!p || fafi_plan_c2c(p, n, 0, (flags & FAF_FLAG_BLUESTEIN) != 0)) {
        fafi_plan_free(p);
        return 0;
    }
    const int kind = p->kind;
    fafd_fft def = p->a;
    fafi_plan_free(p);
    const int spbs[3] = {0, -2, -4};      /* all that fit, half of that, a quarter */
    int kinds[2], nk = 0;
    if (kind == PK_C2C) { kinds[nk++] = FAFD_KIND_CT; kinds[nk++] = FAFD_KIND_STOCKHAM; }
    else if (kind == PK_RADER) { kinds[nk++] = FAFD_KIND_RADER_CT; kinds[nk++] = FAFD_KIND_RADER; }
    else kinds[nk++] = FAFD_KIND_BLUESTEIN;
    int orders[8][FAFD_MAX_STAGES], lens[8], no = 0;
    if (kind == PK_C2C) {
        int rad[FAFD_MAX_STAGES], k;
        if (fafi_make_fft(n, 0, 1, NULL, 0, &def) == 0) {
            k = def.nstages;
            for (int s = 0; s < k; ++s) rad[s] = def.st[s].radix;
            for (int rot = 0; rot < k && no < 7; ++rot) {      /* rotations of the default */
                for (int s = 0; s < k; ++s) orders[no][s] = rad[(s + rot) % k];
                lens[no++] = k;
            }
            int m = 0, split[FAFD_MAX_STAGES];                  /* every 16 as 4 x 4 */
            for (int s = 0; s < k && m < FAFD_MAX_STAGES - 1; ++s) {
                if (rad[s] == 16) { split[m++] = 4; split[m++] = 4; }
                else split[m++] = rad[s];
            }
            if (m != k && m <= FAFD_MAX_STAGES) {
                memcpy(orders[no], split, (size_t)m * sizeof(int));
                lens[no++] = m;
            }
        }
    }
    for (int a = 0; a < nk; ++a) {
        /* runtime passes: every radix order; compile-time kinds: every pass
         * sequence the kernel carries for this length */
        int nord = 1;
        if (kinds[a] == FAFD_KIND_STOCKHAM) nord = no;
        else if (kinds[a] == FAFD_KIND_CT || kinds[a] == FAFD_KIND_RADER_CT)
            while (nord < 4 && (kinds[a] == FAFD_KIND_CT
                                    ? fafk_rg_ct_seq((int)n, nord, NULL, NULL)
                                    : fafk_rg_rader_seq((int)n, nord, NULL, NULL)) > 0)
                ++nord;
        for (int o = 0; o < nord; ++o)
            for (int s = 0; s < 3 && nc < RG_MAX_CANDS; ++s) {
                rg_choice c;
                memset(&c, 0, sizeof c);
                c.kind = kinds[a];
                c.variant = -1;
                if (kinds[a] == FAFD_KIND_STOCKHAM) {
                    c.norder = lens[o];
                    memcpy(c.order, orders[o], (size_t)lens[o] * sizeof(int));
                } else if (kinds[a] != FAFD_KIND_BLUESTEIN && kinds[a] != FAFD_KIND_RADER) {
                    c.variant = o;
                }
                c.spb = spbs[s];          /* negative: a fraction, resolved below */
                out[nc++] = c;
            }
    }
    return nc;
}

/* the most signals of length n a unit can take under choice c, in a
 * batch of that length alone (which gets the workgroup size it fits) */
static int max_spb(size_t n, uint32_t flags, const rg_choice *c, int generic) {
    fafh_plan *p = fafi_plan_new();
    fafd_rplan d;
    size_t per = 0;
    int rc = p ? plan_length(p, n, flags, c, &d, &per) : -1;
    fafi_plan_free(p);
    if (rc) return 0;
    const size_t cap = generic ? RAGGED_LDS_CAP : RAGGED_LDS_CAP_CT;
    int spb = (d.threads <= RAGGED_BLOCK_SMALL ? RAGGED_BLOCK_SMALL : RAGGED_BLOCK) / d.threads;
    while (spb > 1 && (size_t)spb * per * 8 > cap) --spb;
    return spb;
}

int faf_ragged_autotune(faf_ragged *r) {
    if (!r) { fafi_err("ragged autotune: null plan"); return -1; }
    if (fafi_create_ready()) return -1;
    void *stream = fafi_stream();
    /* scratch for inputs and outputs of the largest trial: the whole batch
     * (zeros: no kernel path depends on the data) */
    const size_t bytes = 4 * r->total * sizeof(float);
    float *buf = fafk_malloc(bytes);
    if (!buf) { fafi_err("ragged autotune: %s", fafk_last_error()); return -1; }
    float *zero = calloc(1, bytes);
    if (!zero || fafk_h2d(buf, zero, bytes, stream) || fafk_sync(stream)) {
        free(zero);
        fafk_free(buf);
        fafi_err("ragged autotune: %s", fafk_last_error());
        return -1;
    }
    free(zero);

    /* distinct lengths and how many signals each has */
    size_t *uniq = malloc(r->count * sizeof(size_t));
    memcpy(uniq, r->lengths, r->count * sizeof(size_t));
    qsort(uniq, r->count, sizeof(size_t), cmp_size);
    int nu = 0;
    size_t *cnt = calloc(r->count, sizeof(size_t));
    for (size_t i = 0; i < r->count; ++i) {
        if (i == 0 || uniq[i] != uniq[i - 1]) uniq[nu++] = uniq[i];
        cnt[nu - 1]++;
    }
    rg_choice *best = calloc((size_t)nu, sizeof(rg_choice));
    rg_choice *best_ct = calloc((size_t)nu, sizeof(rg_choice));   /* best without runtime paths */
    int *has_ct = calloc((size_t)nu, sizeof(int));
    int tuned = 0, rc = 0;

    for (int k = 0; k < nu && !rc; ++k) {
        rg_choice known;
        if (rg_tuned_find(uniq[k], &known)) {      /* already measured in this process */
            best[k] = known;
            best_ct[k] = known;
            has_ct[k] = known.kind == FAFD_KIND_CT || known.kind == FAFD_KIND_RADER_CT;
            continue;
        }
        /* a batch of this length alone: as many signals as the caller's
         * batch holds (at least 64, so a unit count is realistic) */
        size_t m = cnt[k] < 64 ? 64 : cnt[k];
        while (m * uniq[k] > r->total && m > 1) --m;
        size_t *one = malloc(m * sizeof(size_t));
        for (size_t i = 0; i < m; ++i) one[i] = uniq[k];
        rg_choice cands[RG_MAX_CANDS];
        int nc = candidates(uniq[k], r->flags, cands);
        float tb = 1e30f, tct = 1e30f;
        for (int ci = 0; ci < nc; ++ci) {
            rg_choice c = cands[ci];
            const int gen = c.kind != FAFD_KIND_CT && c.kind != FAFD_KIND_RADER_CT;
            if (c.spb < 0) {
                const int mx = max_spb(uniq[k], r->flags, &c, gen);
                if (mx <= 0) continue;
                const int s = mx / -c.spb;
                if (s < 1 || (c.spb != 0 && s == mx)) continue;
                c.spb = s;
            }
            rg_overrides ov = {&uniq[k], &c, 1};
            faf_ragged *trial = ragged_new(one, m, r->flags, &ov, 0);
            if (!trial) continue;            /* this length cannot take that choice */
            float ms = time_ragged(trial, buf, stream, 5);
            const int lean = !trial->generic;
            faf_ragged_destroy(trial);
            if (ms <= 0.0f) continue;
            if (ms < tb) { tb = ms; best[k] = c; }
            if (lean && ms < tct) { tct = ms; best_ct[k] = c; has_ct[k] = 1; }
        }
        free(one);
        if (tb >= 1e30f) { fafi_err("ragged autotune: no candidate of length %zu ran", uniq[k]); rc = -1; }
        ++tuned;
    }
    if (!rc) {
        fafi_err("%s", "");     /* candidates a length cannot take left messages */
        /* whole-batch decision, timed on the batch itself: the per-length
         * winners (which may need the runtime paths, and with them the
         * larger kernel), the best compile-time choice of every length that
         * has one, and the planner's defaults -- so a batch never ends up
         * slower than untuned because single-length timings mislead */
        rg_overrides ov_best = {uniq, best, nu};
        rg_choice *lean = malloc((size_t)nu * sizeof(rg_choice));
        rg_choice *dflt = malloc((size_t)nu * sizeof(rg_choice));
        for (int k = 0; k < nu; ++k) {
            lean[k] = has_ct[k] ? best_ct[k] : best[k];
            memset(&dflt[k], 0, sizeof dflt[k]);
            dflt[k].kind = -1;
        }
        rg_overrides ov_lean = {uniq, lean, nu}, ov_dflt = {uniq, dflt, nu};
        const rg_choice *cand[3] = {best, lean, dflt};
        const rg_overrides *ovs[3] = {&ov_best, &ov_lean, &ov_dflt};
        const rg_choice *win = NULL;
        float twin = 1e30f;
        int win_persist = 0, win_block = 0;
        for (int c = 0; c < 3; ++c)
            for (int bw = 0, auto_block = 0; bw < 2; ++bw) {   /* the size that fits, then 512 */
                if (bw && auto_block == RAGGED_BLOCK) break;     /* that was 512 already */
                faf_ragged *a = ragged_new(r->lengths, r->count, r->flags, ovs[c],
                                           bw ? RAGGED_BLOCK : 0);
                if (a && !bw) auto_block = a->block;
                for (int ps = 0; a && ps < 2; ++ps) {  /* one workgroup per unit, persistent */
                    a->persist = ps;
                    /* the whole-batch candidates differ by a few percent:
                     * a median of 9 picks among them more reliably */
                    const float ta = time_ragged(a, buf, stream, 9);
                    if (ta > 0.0f && ta < twin) {
                        twin = ta;
                        win = cand[c];
                        win_persist = ps;
                        win_block = bw ? RAGGED_BLOCK : 0;
                    }
                }
                faf_ragged_destroy(a);
            }
        if (!win) {
            fafi_err("ragged autotune: the batch did not run");
            rc = -1;
        } else {
            for (int k = 0; k < nu; ++k) rg_tuned_store(uniq[k], &win[k]);
            /* rebuild this plan with the winners */
            ragged_free_device(r);
            rg_overrides ov = {uniq, win, nu};
            r->block_want = win_block;
            if (ragged_build(r, &ov)) rc = -1;
            r->persist = win_persist;
        }
        free(lean);
        free(dflt);
    }
    fafk_free(buf);
    free(uniq); free(cnt); free(best); free(best_ct); free(has_ct);
    return rc ? -1 : tuned;
}

int faf_ragged_choice(const faf_ragged *r, size_t len, int *kind, int *spb, int *order,
                      int *norder) {
    if (!r) return -1;
    for (size_t i = 0; i < r->count; ++i) {
        if (r->lengths[i] != len) continue;
        /* the plan table is on the device; read the entry of this length */
        int nu = r->nplans;
        fafd_rplan *tab = malloc((size_t)nu * sizeof(fafd_rplan));
        if (!tab || fafk_d2h(tab, r->d_plans, (size_t)nu * sizeof(fafd_rplan), NULL) ||
            fafk_sync(NULL)) {
            free(tab);
            fafi_err("ragged: %s", fafk_last_error());
            return -1;
        }
        int rc = -1;
        for (int k = 0; k < nu; ++k) {
            if ((size_t)tab[k].len != len) continue;
            if (kind) *kind = tab[k].kind;
            if (spb) *spb = tab[k].spb;
            if (norder) *norder = tab[k].fwd.nstages;
            if (order)
                for (int s = 0; s < tab[k].fwd.nstages; ++s) order[s] = tab[k].fwd.st[s].radix;
            rc = 0;
        }
        free(tab);
        return rc;
    }
    fafi_err("ragged: length %zu is not in this batch", len);
    return -1;
}

int faf_ragged_set_choice(size_t len, int kind, int spb, const int *order, int norder) {
    if (len < 2 || spb < 0 || norder < 0 || norder > FAFD_MAX_STAGES || (norder && !order)) {
        fafi_err("ragged choice: bad argument");
        return -1;
    }
    if (kind == -1) {                    /* forget the length: back to the defaults */
        pthread_mutex_lock(&g_rg_mu);
        for (int i = 0; i < g_rg_ntuned; ++i)
            if (g_rg_tuned[i].len == len) {
                g_rg_tuned[i] = g_rg_tuned[--g_rg_ntuned];
                break;
            }
        pthread_mutex_unlock(&g_rg_mu);
        return 0;
    }
    if (fafi_create_ready()) return -1;
    rg_choice c;
    memset(&c, 0, sizeof c);
    c.kind = kind;
    c.spb = spb;
    c.variant = -1;
    if (kind == FAFD_KIND_CT || kind == FAFD_KIND_RADER_CT) {
        /* the order names one of the kernel's compiled pass sequences */
        for (int v = 0; norder > 0 && c.variant < 0 && v < 4; ++v) {
            int ord[FAFD_MAX_STAGES], nr = 0;
            const int t = kind == FAFD_KIND_CT ? fafk_rg_ct_seq((int)len, v, ord, &nr)
                                               : fafk_rg_rader_seq((int)len, v, ord, &nr);
            if (t > 0 && nr == norder && !memcmp(ord, order, (size_t)nr * sizeof(int)))
                c.variant = v;
        }
        if (norder > 0 && c.variant < 0) {
            fafi_err("ragged choice: no compile-time pass sequence of length %zu in that order",
                     len);
            return -1;
        }
    } else if (kind == FAFD_KIND_STOCKHAM) {
        c.norder = norder;
        if (norder) memcpy(c.order, order, (size_t)norder * sizeof(int));
    } else if (kind != FAFD_KIND_RADER && kind != FAFD_KIND_BLUESTEIN) {
        fafi_err("ragged choice: unknown kind %d", kind);
        return -1;
    }
    /* the length must plan that way */
    fafh_plan *p = fafi_plan_new();
    fafd_rplan d;
    size_t per = 0;
    const int rc = p ? plan_length(p, len, kind == FAFD_KIND_BLUESTEIN ? FAF_FLAG_BLUESTEIN : 0,
                                   &c, &d, &per)
                     : -1;
    fafi_plan_free(p);
    if (rc) return -1;
    rg_tuned_store(len, &c);
    return 0;
}

int faf_ragged_persistent(const faf_ragged *r) { return r ? r->persist : -1; }

int faf_ragged_block(const faf_ragged *r) { return r ? r->block : -1; }

int faf_ragged_set_block(faf_ragged *r, int threads) {
    if (!r) { fafi_err("ragged: null plan"); return -1; }
    if (threads != 0 && threads != RAGGED_BLOCK_SMALL && threads != RAGGED_BLOCK) {
        fafi_err("ragged: workgroup size %d is not 0, 256 or 512", threads);
        return -1;
    }
    if (fafi_create_ready()) return -1;
    const int old = r->block_want;
    ragged_free_device(r);
    r->block_want = threads;
    if (ragged_build(r, NULL) == 0) return 0;
    char msg[256];
    snprintf(msg, sizeof msg, "%s", faf_get_error());
    ragged_free_device(r);
    r->block_want = old;
    if (ragged_build(r, NULL)) return -1;
    fafi_err("%s", msg);
    return -1;
}

int faf_ragged_set_persistent(faf_ragged *r, int on) {
    if (!r) { fafi_err("ragged: null plan"); return -1; }
    r->persist = on != 0;
    return 0;
}
