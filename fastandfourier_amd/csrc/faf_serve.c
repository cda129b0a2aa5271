/*
 * faf_serve.c -- continuous-batching request loop over one device (C11).
 *
 * Requests (rfft, forward C2C, framed STFT, CWT) are submitted from any
 * thread with device-resident input.  One worker thread
 *   1. admits requests in arrival order while their result (and, for long
 *      transforms, workspace) fits the HBM budget, carving it out of one
 *      device arena;
 *   2. gathers admitted requests into a launch when a stream slot is free and
 *      the device is idle, or the oldest admitted request has waited the
 *      batching window, or the batch is full;
 *   3. expands the batch into work units (one plan and several signals each,
 *      most expensive first) in the slot's pinned table and runs them all in
 *      one launch of the persistent serving kernel; requests longer than one
 *      workgroup holds (four-step lengths, long CWT segments) follow on the
 *      same stream through their cached plans on workspace lent from the
 *      arena;
 *   4. polls the slots' completion events and time-stamps each request.
 * Several slots, each with its own stream, let launches of successive
 * windows overlap on the device while new requests keep arriving.  Plans
 * (twiddles, Kiss tables, windows, CWT filter banks) are built once per
 * (kind, length) and stay resident in a device plan table.  A submission
 * whose plan does not exist yet hands the build to a small pool of builder
 * threads and returns at once; its request waits (parked) until the plan
 * is ready, while requests with ready plans go ahead of it.
 *
 * The reference library has no serving layer; this file is new.
 */
#define _POSIX_C_SOURCE 200809L
#include "faf_internal.h"
#include <math.h>
#include <unistd.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SV_MAX_ENTRIES 4096
#define SV_PAGE_BITS 16
#define SV_PAGE ((int64_t)1 << SV_PAGE_BITS)
#define SV_MAX_PAGES 4096
#define SV_MAX_SLOTS 8
#define SV_CWT_ROWS_PER_UNIT 4
#define SV_BLOCK_SMALL 512
#define SV_ALIGN 256
/* plan-building threads, at most; fewer on a host with few CPUs.  On a
 * cold stream of 836 new plans, 4 -> 8 -> 12 builders took the p99 latency
 * from 435 to 250 to 200 ms. */
#define SV_BUILDERS 12

enum { ST_QUEUED = 0, ST_ADMITTED, ST_LAUNCHED, ST_DONE, ST_FAILED };

/* buffers a phase reads or writes: the request's input, its result, or a
 * region of its workspace (which follows the result in its arena block) */
enum { SV_IN = -1, SV_OUT = -2 };
enum { SV_WS0 = 0, SV_WS1, SV_WSZ, SV_WSR, SV_NREG };
#define SV_MAX_PHASES 5

/* one phase of a request longer than one workgroup holds: `units` units
 * of the device plan `splan` (repeated for every CWT filter row when
 * each_row is set) */
typedef struct sphase {
    int splan;
    int units;               /* units per request (per row when each_row) */
    int per;                 /* columns / rows / bins per unit */
    int T;                   /* threads of one unit */
    int lds;                 /* LDS bytes of one unit */
    int each_row;
    int src, dst;            /* SV_IN, SV_OUT or a workspace region */
    int src_row, dst_row;    /* advance src / dst by one row per filter row */
    double cost;             /* relative work of one unit */
} sphase;

typedef struct sentry {
    int kind;                /* FAF_REQ_* */
    size_t n;                /* length (STFT: window) */
    int small;               /* one plan inside the serving kernel */
    int splan;               /* index in the device plan table (small) */
    int nph;                 /* phases of a long request */
    sphase ph[SV_MAX_PHASES];
    size_t reg[SV_NREG];     /* workspace region offsets */
    size_t row_bytes;        /* CWT: bytes of one analytic row */
    int T;                   /* threads per signal (CWT: per unit) */
    int lds_per;             /* LDS float2 per signal (CWT: per unit) */
    int spb;                 /* signals per unit */
    int rows;                /* CWT rows */
    size_t hop;              /* STFT */
    double cost;             /* relative work of one signal (unit ordering) */
    fafh_plan *plan;
    size_t ws_bytes;         /* long transforms: workspace of one request */
    uint64_t stamp;          /* worker: batch that last used it */
    int bucket;
    _Atomic int state;       /* 0 being built, 1 ready, -1 failed (err says why) */
    char err[160];
} sentry;

typedef struct srec {
    int kind;
    uint32_t flags;
    size_t n;
    const float *in;
    sentry *e;
    size_t out_bytes, blk_bytes, blk;
    char *out;
    uint64_t t_arr, t_done;
    double io_bytes;
    int waited;
    int freed;
    int sized;               /* out/blk/io bytes known (its plan was ready) */
    _Atomic int state;
} srec;

typedef struct slot {
    void *stream, *ev_small, *ev_done;
    int busy, nsmall, nreq, small_done, cap;
    uint64_t t_launched;
    int64_t *reqs;
    void *tab;               /* pinned: units then signals */
    size_t tab_bytes;
} slot;

typedef struct frange { size_t off, len; } frange;

typedef struct ring { int64_t *v; size_t head, n, cap; } ring;

struct faf_server {
    faf_server_config cfg;
    /* plans */
    pthread_mutex_t plan_mu;
    sentry *ent[SV_MAX_ENTRIES];
    int nent;
    int hash[2 * SV_MAX_ENTRIES];   /* open addressing: entry index + 1, 0 empty */
    fafd_splan *splans_dev;
    int nsplan;
    pthread_mutex_t splan_mu;        /* the device plan table's fill level */
    pthread_cond_t plan_cv;          /* an entry finished building (plan_mu) */
    /* plan builds handed to the builder threads (plan_mu) */
    pthread_cond_t build_cv;
    sentry **bq;
    int bq_head, bq_n, bq_cap;
    pthread_t builders[SV_BUILDERS];
    int nbuilders, build_stop;
    /* request records, in pages of SV_PAGE tickets; page number pg lives in
     * slot pg % SV_MAX_PAGES, and page_no[] says which page a slot holds
     * (-1 when empty).  A slot is recycled once every request of its page
     * has finished, its result is freed and the stats window has moved past
     * it, so the server runs indefinitely with a bounded record store.
     * next_ticket only advances under in_mu, after the record is filled. */
    srec *pages[SV_MAX_PAGES];
    int64_t page_no[SV_MAX_PAGES];
    _Atomic int64_t next_ticket;
    /* submissions and releases, handed to the worker */
    pthread_mutex_t in_mu;
    pthread_cond_t in_cv;
    ring inbox, rel;
    int sleeping;
    uint64_t plans_built;    /* builds finished (in_mu): wakes the worker */
    /* completion */
    pthread_mutex_t done_mu;
    pthread_cond_t done_cv;
    int64_t finished;        /* done + failed */
    /* worker-only state */
    pthread_t th;
    int started;
    _Atomic int stop;
    char *arena;
    size_t arena_bytes;
    _Atomic size_t in_use, peak;
    frange *fr;
    int nfr, frcap;
    ring pend, ready;
    ring parked;             /* requests whose plan is still being built */
    uint64_t plans_seen;     /* plans_built when parked was last rescanned */
    slot slots[SV_MAX_SLOTS];
    int nslots;
    _Atomic int plain;       /* one launch + synchronize per request */
    _Atomic double window_ns;
    uint64_t stamp;
    /* batch scratch */
    sentry **bk_e;
    int *bk_cnt, *bk_off, bk_cap;
    fafd_ssig *sigs;
    size_t sigs_cap;
    fafd_sunit *units;
    double *ucost;
    size_t units_cap;
    int *uorder, *ulds, *uthr, *uphase;
    /* stats */
    int64_t stats_first;
    _Atomic uint64_t launches, launched_reqs, admission_waits, long_reqs;
    _Atomic uint64_t host_ns, device_ns;     /* per batch: build + launch, launch -> done */
    char err[256];
};

static uint64_t now_ns(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

static inline void cpu_relax(void) {
#if defined(__x86_64__) || defined(__i386__)
    __builtin_ia32_pause();
#endif
}

static size_t al(size_t b) { return (b + SV_ALIGN - 1) & ~(size_t)(SV_ALIGN - 1); }

/* ------------------------------------------------------------------ */
/* rings of tickets                                                     */
/* ------------------------------------------------------------------ */

static int ring_push(ring *r, int64_t t) {
    if (r->n == r->cap) {
        size_t nc = r->cap ? 2 * r->cap : 1024;
        int64_t *v = malloc(nc * sizeof(int64_t));
        if (!v) return -1;
        for (size_t i = 0; i < r->n; ++i) v[i] = r->v[(r->head + i) % r->cap];
        free(r->v);
        r->v = v;
        r->head = 0;
        r->cap = nc;
    }
    r->v[(r->head + r->n) % r->cap] = t;
    r->n++;
    return 0;
}

static int64_t ring_at(const ring *r, size_t i) { return r->v[(r->head + i) % r->cap]; }

static int64_t ring_pop(ring *r) {
    int64_t t = r->v[r->head];
    r->head = (r->head + 1) % r->cap;
    r->n--;
    return t;
}

static srec *rec(faf_server *s, int64_t t) {
    return &s->pages[(t >> SV_PAGE_BITS) % SV_MAX_PAGES][t & (SV_PAGE - 1)];
}

/* 1 if ticket t's record is still stored (its page was not recycled) */
static int rec_live(faf_server *s, int64_t t) {
    const int64_t pg = t >> SV_PAGE_BITS;
    return s->pages[pg % SV_MAX_PAGES] && s->page_no[pg % SV_MAX_PAGES] == pg;
}

/* ------------------------------------------------------------------ */
/* arena: first fit over an address-ordered free list                   */
/* ------------------------------------------------------------------ */

static int arena_alloc(faf_server *s, size_t bytes, size_t *off) {
    for (int i = 0; i < s->nfr; ++i) {
        if (s->fr[i].len < bytes) continue;
        *off = s->fr[i].off;
        s->fr[i].off += bytes;
        s->fr[i].len -= bytes;
        if (!s->fr[i].len) {
            memmove(s->fr + i, s->fr + i + 1, (size_t)(s->nfr - i - 1) * sizeof(frange));
            s->nfr--;
        }
        size_t u = atomic_fetch_add(&s->in_use, bytes) + bytes;
        if (u > atomic_load(&s->peak)) atomic_store(&s->peak, u);
        return 0;
    }
    return -1;
}

static void arena_free(faf_server *s, size_t off, size_t bytes) {
    if (!bytes) return;
    int i = 0;
    while (i < s->nfr && s->fr[i].off < off) ++i;
    int merged = 0;
    if (i > 0 && s->fr[i - 1].off + s->fr[i - 1].len == off) {
        s->fr[i - 1].len += bytes;
        merged = 1;
        if (i < s->nfr && off + bytes == s->fr[i].off) {
            s->fr[i - 1].len += s->fr[i].len;
            memmove(s->fr + i, s->fr + i + 1, (size_t)(s->nfr - i - 1) * sizeof(frange));
            s->nfr--;
        }
    } else if (i < s->nfr && off + bytes == s->fr[i].off) {
        s->fr[i].off = off;
        s->fr[i].len += bytes;
        merged = 1;
    }
    if (!merged) {
        if (s->nfr == s->frcap) {
            int nc = s->frcap ? 2 * s->frcap : 64;
            frange *f = realloc(s->fr, (size_t)nc * sizeof(frange));
            if (!f) return;      /* the block stays lost; the budget shrinks */
            s->fr = f;
            s->frcap = nc;
        }
        memmove(s->fr + i + 1, s->fr + i, (size_t)(s->nfr - i) * sizeof(frange));
        s->fr[i].off = off;
        s->fr[i].len = bytes;
        s->nfr++;
    }
    atomic_fetch_sub(&s->in_use, bytes);
}

/* ------------------------------------------------------------------ */
/* plan cache                                                           */
/* ------------------------------------------------------------------ */

static size_t stft_frames(const faf_server *s, size_t n) {
    size_t w = s->cfg.stft_win, h = s->cfg.stft_hop;
    return n <= w ? 1 : (n - w) / h + 1;
}

static double nlogn(double n) { return n * log2(n > 2 ? n : 2); }

static int upload_splan(faf_server *s, const fafd_splan *sp) {
    pthread_mutex_lock(&s->splan_mu);
    if (s->nsplan >= SV_MAX_ENTRIES) {
        pthread_mutex_unlock(&s->splan_mu);
        fafi_err("serve: plan table full");
        return -1;
    }
    int idx = s->nsplan++;
    pthread_mutex_unlock(&s->splan_mu);
    if (fafk_h2d(s->splans_dev + idx, sp, sizeof *sp, NULL) || fafk_sync(NULL)) {
        fafi_err("serve: %s", fafk_last_error());
        return -1;
    }
    return idx;
}

static int shape_fits(sentry *e) {
    if (e->T < 1 || e->T > SV_BLOCK_SMALL) return 0;
    /* the serving kernel keeps a few bytes of static LDS beside the table */
    const size_t cap = FAFD_MAX_LDS_BYTES - SV_ALIGN;
    size_t per = (size_t)e->lds_per * 8;
    if (per > cap) return 0;
    /* units stay within 512 threads, the serving kernel's register budget
     * without spills */
    int by_t = SV_BLOCK_SMALL / e->T;
    int by_l = (int)(cap / per);
    e->spb = by_t < by_l ? by_t : by_l;
    return e->spb >= 1;
}

/* FFTs of P side by side in one unit: at most 16, within 512 threads and
 * the LDS */
static int fs_per_unit(const fafd_fft *P) {
    int per = 16;
    while (per > 1 && (per * P->threads > SV_BLOCK_SMALL ||
                       (size_t)per * (size_t)P->plane * 8 > FAFD_MAX_LDS_BYTES - SV_ALIGN))
        per >>= 1;
    return per;
}

static int add_phase(faf_server *s, sentry *e, const fafd_splan *sp, int units, int per, int T,
                     int lds, int src, int dst, double cost) {
    if (e->nph >= SV_MAX_PHASES || T > SV_BLOCK_SMALL) {
        fafi_err("serve: phase shape out of range");
        return -1;
    }
    int idx = upload_splan(s, sp);
    if (idx < 0) return -1;
    sphase *ph = &e->ph[e->nph++];
    memset(ph, 0, sizeof *ph);
    ph->splan = idx;
    ph->units = units;
    ph->per = per;
    ph->T = T;
    ph->lds = lds;
    ph->src = src;
    ph->dst = dst;
    ph->cost = cost;
    return 0;
}

/* four-step phases of a length-n C2C (direction inv) from src to dst;
 * fsin selects how the column pass gathers its input */
static int add_4step(faf_server *s, sentry *e, size_t n, int inv, int fsin, size_t nc_spread,
                     const float *psi, long psi_stride, float scale, int src, int mid, int dst) {
    int n1, n2;
    if (fafi_split_4step(n, &n1, &n2)) {
        fafi_err("serve: length %zu has no four-step split", n);
        return -1;
    }
    fafd_splan col, row;
    memset(&col, 0, sizeof col);
    memset(&row, 0, sizeof row);
    if (fafi_make_fft((size_t)n1, inv, 0, NULL, 0, &col.a) ||
        fafi_make_fft((size_t)n2, inv, 0, NULL, 0, &row.a))
        return -1;
    col.kind = FAFD_SK_FSCOL;
    col.inv = inv;
    col.fs_other = n2;
    col.fs_per = fs_per_unit(&col.a);
    col.fs_in = fsin;
    col.fs_nc = (int32_t)nc_spread;
    col.psi = psi;
    col.psi_stride = psi_stride;
    col.twn = fafi_twiddles(n, inv);
    row.kind = FAFD_SK_FSROW;
    row.inv = inv;
    row.fs_other = n1;
    row.fs_per = fs_per_unit(&row.a);
    row.scale = scale;
    if (!col.twn) return -1;
    const int cu = (n2 + col.fs_per - 1) / col.fs_per, ru = (n1 + row.fs_per - 1) / row.fs_per;
    if (add_phase(s, e, &col, cu, col.fs_per, col.fs_per * col.a.threads,
                  col.fs_per * col.a.plane * 8, src, mid, col.fs_per * nlogn(n1)) ||
        add_phase(s, e, &row, ru, row.fs_per, row.fs_per * row.a.threads,
                  row.fs_per * row.a.plane * 8, mid, dst, row.fs_per * nlogn(n2)))
        return -1;
    return 0;
}

/* Kiss post-pass over a packed length-nc spectrum, src -> dst */
static int add_post(faf_server *s, sentry *e, size_t nc, const void *post, int src, int dst) {
    fafd_splan sp;
    memset(&sp, 0, sizeof sp);
    sp.kind = FAFD_SK_POST;
    sp.fs_other = (int32_t)nc;
    sp.fs_per = 2048;
    sp.post = post;
    const int units = (int)((nc / 2 + 1 + 2047) / 2048);
    return add_phase(s, e, &sp, units, 2048, 256, 0, src, dst, 2048.0);
}

/* packed rfft of length 2 nc from src into the Hermitian spectrum at dst:
 * one serving-kernel signal when nc fits a 512-thread workgroup, else
 * four-step phases plus the post-pass */
static int add_rfft(faf_server *s, sentry *e, fafh_plan *q, size_t nc, int src, int dst) {
    fafd_fft a;
    if (fafi_radix_fits(nc, 0) && !fafi_make_fft(nc, 0, 0, NULL, 0, &a) &&
        a.threads <= SV_BLOCK_SMALL) {
        fafd_splan sp;
        memset(&sp, 0, sizeof sp);
        sp.kind = FAFD_SK_R2C;
        sp.a = a;
        sp.post = fafi_kiss_table(nc, 0);
        sp.len = (int32_t)(2 * nc);
        sp.threads = a.threads;
        sp.lds_per = a.plane;
        if (!sp.post) return -1;
        return add_phase(s, e, &sp, 1, 1, a.threads, a.plane * 8, src, dst, nlogn(nc));
    }
    (void)q;
    const void *post = fafi_kiss_table(nc, 0);
    if (!post) return -1;
    return add_4step(s, e, nc, 0, FAFD_FSIN_PAIR, 0, NULL, 0, 1.0f, src, SV_WS0, SV_WS1) ||
                   add_post(s, e, nc, post, SV_WS1, dst)
               ? -1
               : 0;
}

static int build_entry(faf_server *s, sentry *e) {
    fafh_plan *q = fafi_plan_new();
    if (!q) { fafi_err("serve: alloc failed"); return -1; }
    e->plan = q;
    fafd_splan sp;
    memset(&sp, 0, sizeof sp);
    size_t n = e->n;
    switch (e->kind) {
    case FAF_REQ_C2C:
        if (fafi_plan_c2c(q, n, 0, 0)) return -1;
        e->cost = nlogn((double)n);
        if (q->kind == PK_C2C && q->a.threads <= SV_BLOCK_SMALL) {
            sp.kind = FAFD_SK_C2C;
            sp.a = q->a;
            e->T = q->a.threads;
            e->lds_per = q->a.plane;
        } else if (q->kind == PK_RADER && q->a.threads <= SV_BLOCK_SMALL) {
            sp.kind = FAFD_SK_RADER;
            sp.a = q->a;
            sp.b = q->b;
            sp.rader = q->rader;
            e->T = q->a.threads;
            e->lds_per = q->a.plane + (int)(n + (n >> 4) + 1);
        } else if (q->kind == PK_BLUE && q->a.threads <= SV_BLOCK_SMALL) {
            sp.kind = FAFD_SK_BLUE;
            sp.a = q->a;
            sp.b = q->b;
            sp.blue = q->blue;
            e->T = q->a.threads;
            e->lds_per = q->a.plane;
        } else if (q->kind == PK_C2C || q->kind == PK_C2C_4S) {
            e->reg[SV_WS0] = 0;
            if (add_4step(s, e, n, 0, FAFD_FSIN_INTER, 0, NULL, 0, 1.0f, SV_IN, SV_WS0, SV_OUT))
                return -1;
            e->ws_bytes = al(n * 8);
            return 0;
        } else {
            fafi_err("serve: C2C length %zu is beyond the serving kernel", n);
            return -1;
        }
        break;
    case FAF_REQ_RFFT: {
        if (n < 4 || (n & 1)) { fafi_err("serve: rfft length %zu must be even and >= 4", n); return -1; }
        const size_t nc = n / 2;
        e->cost = nlogn((double)nc);
        fafd_fft a;
        if (fafi_radix_fits(nc, 0) && !fafi_make_fft(nc, 0, 0, NULL, 0, &a) &&
            a.threads <= SV_BLOCK_SMALL) {
            sp.kind = FAFD_SK_R2C;
            sp.a = a;
            sp.post = fafi_kiss_table(nc, 0);
            if (!sp.post) return -1;
            e->T = a.threads;
            e->lds_per = a.plane;
            break;
        }
        e->reg[SV_WS0] = 0;
        e->reg[SV_WS1] = al(nc * 8);
        if (add_rfft(s, e, q, nc, SV_IN, SV_OUT)) return -1;
        e->ws_bytes = 2 * al(nc * 8);
        return 0;
    }
    case FAF_REQ_STFT:
        if (fafi_plan_stft(q, n, s->cfg.stft_hop)) return -1;
        e->hop = s->cfg.stft_hop;
        e->cost = nlogn((double)n / 2);
        sp.kind = FAFD_SK_STFT;
        sp.a = q->a;
        sp.post = q->post;
        sp.win = q->win_dev;
        sp.hop = (int32_t)e->hop;
        e->T = q->a.threads;
        e->lds_per = q->a.plane;
        break;
    case FAF_REQ_CWT: {
        faf_cwt_config c = faf_cwt_config_init(n);
        c.fs = s->cfg.cwt_fs;
        c.voices = s->cfg.cwt_voices;
        c.fmin = 16.0 * c.fs / (double)n;
        c.fmax = c.fmin * pow(2.0, s->cfg.cwt_octaves);
        if (fafi_cwt_build(q, &c, -1)) return -1;
        e->rows = q->rows_total;
        e->cost = nlogn((double)n) * SV_CWT_ROWS_PER_UNIT;
        const int fits = !q->sub_r2c && q->a.threads <= SV_BLOCK_SMALL &&
                         q->full_inv.threads <= SV_BLOCK_SMALL;
        if (fits) {
            sp.kind = FAFD_SK_CWT;
            sp.a = q->a;
            sp.b = q->full_inv;
            sp.post = q->post;
            sp.psi = q->psi_dev;
            sp.psi_stride = q->psi_stride;
            sp.rows = q->rows_total;
            e->T = q->a.threads > q->full_inv.threads ? q->a.threads : q->full_inv.threads;
            e->lds_per = q->a.plane + q->full_inv.plane;
            break;
        }
        /* rfft into Z, then per filter row an inverse four-step over the
         * analytic spectrum Z psi_r gathered on the fly */
        const size_t nc = n / 2;
        e->reg[SV_WS0] = 0;
        e->reg[SV_WS1] = al(nc * 8);
        e->reg[SV_WSZ] = 2 * al(nc * 8);
        e->reg[SV_WSR] = e->reg[SV_WSZ] + al((nc + 1) * 8);
        e->row_bytes = n * 8;
        e->ws_bytes = e->reg[SV_WSR] + al((size_t)e->rows * n * 8);
        if (add_rfft(s, e, q, nc, SV_IN, SV_WSZ)) return -1;
        const int first = e->nph;
        if (add_4step(s, e, n, 1, FAFD_FSIN_SPREAD, nc, q->psi_dev, q->psi_stride,
                      (float)(1.0 / (double)n), SV_WSZ, SV_WSR, SV_OUT))
            return -1;
        for (int i = first; i < e->nph; ++i) e->ph[i].each_row = 1;
        e->ph[first].dst_row = 1;                     /* column pass: Z -> WSR row */
        e->ph[first + 1].src_row = 1;                 /* row pass: WSR row -> out row */
        e->ph[first + 1].dst_row = 1;
        return 0;
    }
    default:
        fafi_err("serve: unknown request kind %d", e->kind);
        return -1;
    }
    if (!shape_fits(e)) { fafi_err("serve: plan of length %zu exceeds one workgroup", n); return -1; }
    if (e->kind == FAF_REQ_CWT) e->spb = 1;
    sp.len = (int32_t)n;
    sp.threads = e->T;
    sp.lds_per = e->lds_per;
    int idx = upload_splan(s, &sp);
    if (idx < 0) return -1;
    e->splan = idx;
    e->small = 1;
    return 0;
}

static unsigned hslot(int kind, size_t key) {
    uint64_t h = ((uint64_t)key << 3 ^ (uint64_t)kind) * 0x9E3779B97F4A7C15ull;
    return (unsigned)(h >> 40) & (2 * SV_MAX_ENTRIES - 1);
}

/* the entry of (kind, key) in the table, or NULL; *slot receives the empty
 * hash slot the probe ended on.  plan_mu held. */
static sentry *entry_find(faf_server *s, int kind, size_t key, unsigned *slot) {
    unsigned h = hslot(kind, key);
    for (; s->hash[h]; h = (h + 1) & (2 * SV_MAX_ENTRIES - 1)) {
        sentry *e = s->ent[s->hash[h] - 1];
        if (e->kind == kind && e->n == key) return e;
    }
    *slot = h;
    return NULL;
}

/* the checks a request gets at once, before any plan is built */
static int request_ok(faf_server *s, int kind, size_t n) {
    if (kind < FAF_REQ_RFFT || kind > FAF_REQ_CWT) {
        fafi_err("serve: unknown request kind %d", kind);
        return 0;
    }
    if (n < 2 || n > ((size_t)1 << 16) * (kind == FAF_REQ_STFT ? 1024 : 1)) {
        fafi_err("serve: length %zu out of range", n);
        return 0;
    }
    if (kind == FAF_REQ_RFFT && (n < 4 || (n & 1))) {
        fafi_err("serve: rfft length %zu must be even and >= 4", n);
        return 0;
    }
    return 1;
}

/* builds e outside plan_mu and publishes the outcome */
static void build_publish(faf_server *s, sentry *e) {
    const int rc = build_entry(s, e);
    if (rc) snprintf(e->err, sizeof e->err, "%s", faf_get_error());
    pthread_mutex_lock(&s->plan_mu);
    atomic_store(&e->state, rc ? -1 : 1);
    pthread_cond_broadcast(&s->plan_cv);
    pthread_mutex_unlock(&s->plan_mu);
    pthread_mutex_lock(&s->in_mu);           /* the worker rescans its parked requests */
    s->plans_built++;
    if (s->sleeping) pthread_cond_signal(&s->in_cv);
    pthread_mutex_unlock(&s->in_mu);
}

/* builder threads to start: SV_BUILDERS, or the online CPUs less two */
static int builder_cap(void) {
    const long n = sysconf(_SC_NPROCESSORS_ONLN) - 2;
    return n < 1 ? 1 : n < SV_BUILDERS ? (int)n : SV_BUILDERS;
}

static void *builder_main(void *arg) {
    faf_server *s = arg;
    pthread_mutex_lock(&s->plan_mu);
    for (;;) {
        while (!s->bq_n && !s->build_stop) pthread_cond_wait(&s->build_cv, &s->plan_mu);
        /* stopping: builds still queued are dropped (the server is being
         * destroyed and its worker has already failed the requests parked
         * on them); unbuilt entries are freed with the rest */
        if (s->build_stop) break;
        sentry *e = s->bq[s->bq_head];
        s->bq_head = (s->bq_head + 1) % s->bq_cap;
        s->bq_n--;
        pthread_mutex_unlock(&s->plan_mu);
        build_publish(s, e);
        pthread_mutex_lock(&s->plan_mu);
    }
    pthread_mutex_unlock(&s->plan_mu);
    return NULL;
}

/* Entry of (kind, n), created when missing (state 0, being built); STFT
 * entries are keyed by the window.  With wait, a missing entry is built on
 * this thread and an entry being built elsewhere is waited for, so a ready
 * entry (or NULL with the build's error) comes back; without wait, a
 * missing entry goes to the builder threads and may come back unbuilt. */
static sentry *entry_acquire(faf_server *s, int kind, size_t n, int wait) {
    if (!request_ok(s, kind, n)) return NULL;
    size_t key = kind == FAF_REQ_STFT ? s->cfg.stft_win : n;
    unsigned h = 0;
    pthread_mutex_lock(&s->plan_mu);
    sentry *e = entry_find(s, kind, key, &h);
    int mine = 0;
    if (!e) {
        if (s->nent >= SV_MAX_ENTRIES) {
            pthread_mutex_unlock(&s->plan_mu);
            fafi_err("serve: too many distinct (kind, length) plans");
            return NULL;
        }
        e = calloc(1, sizeof *e);
        if (!e) {
            pthread_mutex_unlock(&s->plan_mu);
            fafi_err("serve: alloc failed");
            return NULL;
        }
        e->kind = kind;
        e->n = key;
        s->ent[s->nent++] = e;
        s->hash[h] = s->nent;
        mine = 1;
        if (!wait) {
            /* to the builder threads (started on the first such build) */
            int queued = 0;
            if (s->bq_n == s->bq_cap) {
                int nc = s->bq_cap ? 2 * s->bq_cap : 64;
                sentry **q = malloc((size_t)nc * sizeof *q);
                if (q) {
                    for (int i = 0; i < s->bq_n; ++i) q[i] = s->bq[(s->bq_head + i) % s->bq_cap];
                    free(s->bq);
                    s->bq = q;
                    s->bq_head = 0;
                    s->bq_cap = nc;
                }
            }
            if (s->bq_n < s->bq_cap) {
                s->bq[(s->bq_head + s->bq_n) % s->bq_cap] = e;
                s->bq_n++;
                while (s->nbuilders < builder_cap() &&
                       s->nbuilders < s->bq_n &&
                       pthread_create(&s->builders[s->nbuilders], NULL, builder_main, s) == 0)
                    s->nbuilders++;
                if (s->nbuilders > 0) {
                    pthread_cond_signal(&s->build_cv);
                    queued = 1;
                } else {
                    s->bq_n--;               /* no builder thread: build here */
                }
            }
            if (queued) {
                pthread_mutex_unlock(&s->plan_mu);
                return e;
            }
        }
    }
    if (mine) {
        pthread_mutex_unlock(&s->plan_mu);
        build_publish(s, e);
        pthread_mutex_lock(&s->plan_mu);
    }
    if (wait)
        while (atomic_load(&e->state) == 0) pthread_cond_wait(&s->plan_cv, &s->plan_mu);
    const int st = atomic_load(&e->state);
    pthread_mutex_unlock(&s->plan_mu);
    if (st < 0) {
        fafi_err("%s", e->err[0] ? e->err : "serve: plan build failed");
        return NULL;
    }
    return e;
}

static sentry *entry_get(faf_server *s, int kind, size_t n) { return entry_acquire(s, kind, n, 1); }

static size_t result_bytes(faf_server *s, const sentry *e, size_t n) {
    switch (e->kind) {
    case FAF_REQ_C2C: return n * 8;
    case FAF_REQ_RFFT: return (n / 2 + 1) * 8;
    case FAF_REQ_STFT: return stft_frames(s, n) * (e->n / 2 + 1) * 8;
    default: return (size_t)e->rows * n * 8;
    }
}

static double input_bytes(int kind, size_t n) { return (double)n * (kind == FAF_REQ_C2C ? 8 : 4); }

/* ------------------------------------------------------------------ */
/* worker                                                               */
/* ------------------------------------------------------------------ */

static void set_err(faf_server *s, const char *msg) {
    pthread_mutex_lock(&s->done_mu);
    snprintf(s->err, sizeof s->err, "%s", msg && msg[0] ? msg : "serve: launch failed");
    pthread_mutex_unlock(&s->done_mu);
}

static void finish(faf_server *s, int64_t *tk, int cnt, int state) {
    uint64_t t = now_ns();
    pthread_mutex_lock(&s->done_mu);
    for (int i = 0; i < cnt; ++i) {
        srec *r = rec(s, tk[i]);
        r->t_done = t;
        if (!r->freed && ((r->flags & FAF_REQ_DISCARD) || state == ST_FAILED)) {
            arena_free(s, r->blk, r->blk_bytes);
            r->freed = 1;
        }
        atomic_store(&r->state, state);
    }
    s->finished += cnt;
    pthread_cond_broadcast(&s->done_cv);
    pthread_mutex_unlock(&s->done_mu);
}

static int take_inbox(faf_server *s) {
    int did = 0;
    pthread_mutex_lock(&s->in_mu);
    while (s->inbox.n) {
        if (ring_push(&s->pend, ring_pop(&s->inbox))) break;
        did = 1;
    }
    while (s->rel.n) {
        const int64_t t = ring_pop(&s->rel);
        if (!rec_live(s, t)) continue;     /* recycled after an earlier release */
        srec *r = rec(s, t);
        if (!r->freed) {
            arena_free(s, r->blk, r->blk_bytes);
            r->freed = 1;
        }
        did = 1;
    }
    pthread_mutex_unlock(&s->in_mu);
    return did;
}

/* result, arena block and I/O bytes of r, from its (ready) plan */
static void size_request(faf_server *s, srec *r) {
    r->out_bytes = result_bytes(s, r->e, r->n);
    r->blk_bytes = al(r->out_bytes) + al(r->e->ws_bytes);
    r->io_bytes = input_bytes(r->kind, r->n) + (double)r->out_bytes;
    r->sized = 1;
}

/* Parked requests whose plan finished building (or failed) go back to the
 * admission queue; run when a build has finished since the last scan. */
static int unpark(faf_server *s) {
    pthread_mutex_lock(&s->in_mu);
    const uint64_t built = s->plans_built;
    pthread_mutex_unlock(&s->in_mu);
    if (!s->parked.n || built == s->plans_seen) return 0;
    s->plans_seen = built;
    int did = 0;
    for (size_t i = 0, m = s->parked.n; i < m; ++i) {
        int64_t t = ring_pop(&s->parked);
        if (atomic_load(&rec(s, t)->e->state) != 0) {
            if (!ring_push(&s->pend, t)) { did = 1; continue; }
        }
        ring_push(&s->parked, t);            /* still building (or no room to move) */
    }
    return did;
}

static int admit(faf_server *s) {
    int did = 0;
    while (s->pend.n) {
        int64_t t = ring_at(&s->pend, 0);
        srec *r = rec(s, t);
        if (!r->sized) {
            const int st = atomic_load(&r->e->state);
            if (st == 0) {                   /* its plan is being built: step aside */
                ring_pop(&s->pend);
                if (ring_push(&s->parked, t)) {
                    set_err(s, "serve: alloc failed");
                    r->freed = 1;
                    finish(s, &t, 1, ST_FAILED);
                }
                did = 1;
                continue;
            }
            if (st < 0) {
                ring_pop(&s->pend);
                set_err(s, r->e->err[0] ? r->e->err : "serve: plan build failed");
                r->freed = 1;
                finish(s, &t, 1, ST_FAILED);
                did = 1;
                continue;
            }
            size_request(s, r);
        }
        if (r->blk_bytes > s->arena_bytes) {
            ring_pop(&s->pend);
            set_err(s, "serve: request larger than the HBM budget");
            r->freed = 1;
            r->blk_bytes = 0;
            finish(s, &t, 1, ST_FAILED);
            did = 1;
            continue;
        }
        size_t off;
        if (arena_alloc(s, r->blk_bytes, &off)) {
            if (!r->waited) {
                r->waited = 1;
                atomic_fetch_add(&s->admission_waits, 1);
            }
            break;
        }
        r->blk = off;
        r->out = s->arena + off;
        atomic_store(&r->state, ST_ADMITTED);
        ring_pop(&s->pend);
        ring_push(&s->ready, t);
        did = 1;
    }
    return did;
}

static int grow(void **p, size_t *cap, size_t need, size_t elem) {
    if (need <= *cap) return 0;
    size_t nc = *cap ? *cap : 256;
    while (nc < need) nc *= 2;
    void *q = realloc(*p, nc * elem);
    if (!q) return -1;
    *p = q;
    *cap = nc;
    return 0;
}

static int signals_of(faf_server *s, const srec *r) {
    if (r->kind == FAF_REQ_STFT) return (int)stft_frames(s, r->n);
    if (r->kind == FAF_REQ_CWT)
        return (r->e->rows + SV_CWT_ROWS_PER_UNIT - 1) / SV_CWT_ROWS_PER_UNIT;
    return 1;
}

static faf_server *g_sort_srv;
static int by_cost_desc(const void *a, const void *b) {
    double ca = g_sort_srv->ucost[*(const int *)a], cb = g_sort_srv->ucost[*(const int *)b];
    return ca < cb ? 1 : (ca > cb ? -1 : 0);
}

/* where buffer `b` of request r starts, for filter row `row` */
static const float *buf_of(const srec *r, int b, int row, int advance) {
    const char *p;
    if (b == SV_IN) p = (const char *)r->in;
    else if (b == SV_OUT) p = r->out;
    else p = r->out + al(r->out_bytes) + r->e->reg[b];
    if (advance) p += (size_t)row * r->e->row_bytes;
    return (const float *)p;
}

/* Pops `take` admitted requests into the slot (short ones first) and
 * launches them: one serving-kernel launch per phase, the short requests
 * and the first phase of the long ones sharing the first launch.  Returns
 * 0, -1 when the popped requests must fail (after the slot's stream has
 * drained), or -2 when nothing was popped. */
static int launch(faf_server *s, slot *sl, int take) {
    const uint64_t stamp = ++s->stamp;
    const uint64_t t0 = now_ns();
    if (sl->cap < take) {
        int64_t *v = realloc(sl->reqs, (size_t)take * sizeof(int64_t));
        if (!v) return -2;
        sl->reqs = v;
        sl->cap = take;
    }
    int ns = 0, nl = 0, nb = 0, nph = 1;
    size_t nsig = 0, nunits = 0;
    for (int i = 0; i < take; ++i)
        if (rec(s, ring_at(&s->ready, (size_t)i))->e->small) ++ns;
    for (int i = 0; i < take; ++i) {
        int64_t t = ring_pop(&s->ready);
        srec *r = rec(s, t);
        atomic_store(&r->state, ST_LAUNCHED);
        if (r->e->small) sl->reqs[i - nl] = t;
        else sl->reqs[ns + nl++] = t;
    }
    sl->nsmall = ns;
    sl->nreq = take;
    sl->small_done = 0;
    for (int i = 0; i < ns; ++i) {
        srec *r = rec(s, sl->reqs[i]);
        sentry *e = r->e;
        if (e->stamp != stamp) {
            if (nb == s->bk_cap) {
                int nc = s->bk_cap ? 2 * s->bk_cap : 64;
                sentry **be = realloc(s->bk_e, (size_t)nc * sizeof(sentry *));
                if (be) s->bk_e = be;
                int *bc = be ? realloc(s->bk_cnt, (size_t)nc * sizeof(int)) : NULL;
                if (bc) s->bk_cnt = bc;
                int *bo = bc ? realloc(s->bk_off, (size_t)nc * sizeof(int)) : NULL;
                if (bo) s->bk_off = bo;
                if (!bo) return -1;
                s->bk_cap = nc;
            }
            e->stamp = stamp;
            e->bucket = nb;
            s->bk_e[nb] = e;
            s->bk_cnt[nb] = 0;
            ++nb;
        }
        int k = signals_of(s, r);
        s->bk_cnt[e->bucket] += k;
        nsig += (size_t)k;
    }
    const size_t nsig_small = nsig;
    for (int b = 0, off = 0; b < nb; ++b) {
        s->bk_off[b] = off;
        off += s->bk_cnt[b];
        nunits += (size_t)((s->bk_cnt[b] + s->bk_e[b]->spb - 1) / s->bk_e[b]->spb);
    }
    for (int i = ns; i < take; ++i) {
        const sentry *e = rec(s, sl->reqs[i])->e;
        for (int p = 0; p < e->nph; ++p)
            nunits += (size_t)e->ph[p].units * (size_t)(e->ph[p].each_row ? e->rows : 1);
        if (e->nph > nph) nph = e->nph;
    }
    for (int i = ns; i < take; ++i) {
        const sentry *e = rec(s, sl->reqs[i])->e;
        for (int p = 0; p < e->nph; ++p)
            nsig += (size_t)e->ph[p].units * (size_t)(e->ph[p].each_row ? e->rows : 1);
    }
    if (grow((void **)&s->sigs, &s->sigs_cap, nsig, sizeof(fafd_ssig))) return -1;
    if (grow((void **)&s->units, &s->units_cap, nunits, sizeof(fafd_sunit))) return -1;
    s->ucost = realloc(s->ucost, s->units_cap * sizeof(double));
    s->uorder = realloc(s->uorder, s->units_cap * sizeof(int));
    s->ulds = realloc(s->ulds, s->units_cap * sizeof(int));
    s->uthr = realloc(s->uthr, s->units_cap * sizeof(int));
    s->uphase = realloc(s->uphase, s->units_cap * sizeof(int));
    if (!s->ucost || !s->uorder || !s->ulds || !s->uthr || !s->uphase) return -1;

    /* short requests: signals grouped by plan, units of at most spb */
    for (int i = 0; i < ns; ++i) {
        srec *r = rec(s, sl->reqs[i]);
        sentry *e = r->e;
        fafd_ssig *g = s->sigs + s->bk_off[e->bucket];
        if (r->kind == FAF_REQ_STFT) {
            const size_t fr = stft_frames(s, r->n), bins = e->n / 2 + 1;
            for (size_t f = 0; f < fr; ++f) {
                g[f].in = r->in + f * e->hop;
                g[f].out = (float *)r->out + 2 * f * bins;
                g[f].avail = (int64_t)(r->n - f * e->hop);
                g[f].r0 = g[f].r1 = 0;
            }
            s->bk_off[e->bucket] += (int)fr;
        } else if (r->kind == FAF_REQ_CWT) {
            int k = 0;
            for (int r0 = 0; r0 < e->rows; r0 += SV_CWT_ROWS_PER_UNIT, ++k) {
                g[k].in = r->in;
                g[k].out = (float *)r->out;
                g[k].avail = (int64_t)r->n;
                g[k].r0 = r0;
                g[k].r1 = r0 + SV_CWT_ROWS_PER_UNIT < e->rows ? r0 + SV_CWT_ROWS_PER_UNIT
                                                               : e->rows;
            }
            s->bk_off[e->bucket] += k;
        } else {
            g[0].in = r->in;
            g[0].out = (float *)r->out;
            g[0].avail = (int64_t)r->n;
            g[0].r0 = g[0].r1 = 0;
            s->bk_off[e->bucket] += 1;
        }
    }
    size_t u = 0;
    for (int b = 0, first = 0; b < nb; ++b) {
        sentry *e = s->bk_e[b];
        for (int i = 0; i < s->bk_cnt[b]; i += e->spb, ++u) {
            int c = s->bk_cnt[b] - i < e->spb ? s->bk_cnt[b] - i : e->spb;
            s->units[u].plan = e->splan;
            s->units[u].count = c;
            s->units[u].first = first + i;
            s->units[u].pad = 0;
            s->ucost[u] = e->cost * (double)c;
            s->ulds[u] = c * e->lds_per * 8;
            s->uthr[u] = c * e->T;
            s->uphase[u] = 0;
        }
        first += s->bk_cnt[b];
    }
    /* long requests: one unit and one signal per column / row / bin block */
    size_t g = nsig_small;
    for (int i = ns; i < take; ++i) {
        const srec *r = rec(s, sl->reqs[i]);
        const sentry *e = r->e;
        for (int p = 0; p < e->nph; ++p) {
            const sphase *ph = &e->ph[p];
            const int reps = ph->each_row ? e->rows : 1;
            for (int row = 0; row < reps; ++row)
                for (int k = 0; k < ph->units; ++k, ++g, ++u) {
                    fafd_ssig *sg = &s->sigs[g];
                    sg->in = buf_of(r, ph->src, row, ph->src_row);
                    sg->out = (float *)buf_of(r, ph->dst, row, ph->dst_row);
                    sg->avail = row;
                    sg->r0 = k * ph->per;
                    sg->r1 = 0;
                    s->units[u].plan = ph->splan;
                    s->units[u].count = 1;
                    s->units[u].first = (int32_t)g;
                    s->units[u].pad = 0;
                    s->ucost[u] = ph->cost;
                    s->ulds[u] = ph->lds;
                    s->uthr[u] = ph->T;
                    s->uphase[u] = p;
                }
        }
    }
    /* per phase, most expensive units first */
    int pcount[SV_MAX_PHASES + 1] = {0}, pstart[SV_MAX_PHASES + 1] = {0};
    for (size_t k = 0; k < nunits; ++k) pcount[s->uphase[k]]++;
    for (int p = 1; p <= nph; ++p) pstart[p] = pstart[p - 1] + pcount[p - 1];
    {
        int fill[SV_MAX_PHASES];
        for (int p = 0; p < nph; ++p) fill[p] = pstart[p];
        for (size_t k = 0; k < nunits; ++k) s->uorder[fill[s->uphase[k]]++] = (int)k;
    }
    g_sort_srv = s;
    for (int p = 0; p < nph; ++p)
        qsort(s->uorder + pstart[p], (size_t)pcount[p], sizeof(int), by_cost_desc);
    size_t need = nunits * sizeof(fafd_sunit) + nsig * sizeof(fafd_ssig);
    if (need > sl->tab_bytes) {
        fafk_host_free(sl->tab);
        size_t nb2 = sl->tab_bytes ? sl->tab_bytes : 65536;
        while (nb2 < need) nb2 *= 2;
        sl->tab = fafk_host_alloc(nb2);
        sl->tab_bytes = sl->tab ? nb2 : 0;
        if (!sl->tab) { set_err(s, fafk_last_error()); return -1; }
    }
    fafd_sunit *tu = sl->tab;
    fafd_ssig *ts = (fafd_ssig *)(tu + nunits);
    for (size_t k = 0; k < nunits; ++k) tu[k] = s->units[s->uorder[k]];
    memcpy(ts, s->sigs, nsig * sizeof(fafd_ssig));
    void *st = sl->stream;
    for (int p = 0; p < nph; ++p) {
        if (!pcount[p]) continue;
        int blk = 64, lds = 0;
        for (int k = pstart[p]; k < pstart[p] + pcount[p]; ++k) {
            const int q = s->uorder[k];
            const int th = (s->uthr[q] + 63) & ~63;
            if (th > blk) blk = th;
            if (s->ulds[q] > lds) lds = s->ulds[q];
        }
        if (fafk_serve(s->splans_dev, tu + pstart[p], pcount[p], ts, NULL, 0, blk, (size_t)lds,
                       st)) {
            set_err(s, fafk_last_error());
            return -1;
        }
        if (p == 0 && ns && nl && fafk_event_record(sl->ev_small, st)) return -1;
    }
    if (fafk_event_record(sl->ev_done, st)) return -1;
    atomic_fetch_add(&s->launches, 1);
    atomic_fetch_add(&s->launched_reqs, (uint64_t)take);
    atomic_fetch_add(&s->long_reqs, (uint64_t)nl);
    sl->t_launched = now_ns();
    atomic_fetch_add(&s->host_ns, sl->t_launched - t0);
    sl->busy = 1;
    return 0;
}

static int poll_slots(faf_server *s) {
    int did = 0;
    for (int i = 0; i < s->nslots; ++i) {
        slot *sl = &s->slots[i];
        if (!sl->busy) continue;
        if (sl->nsmall && sl->nsmall < sl->nreq && !sl->small_done) {
            int q = fafk_event_query(sl->ev_small);
            if (q == 1) {
                finish(s, sl->reqs, sl->nsmall, ST_DONE);
                sl->small_done = 1;
                did = 1;
            }
        }
        int q = fafk_event_query(sl->ev_done);
        if (q == 0) continue;
        int from = sl->small_done ? sl->nsmall : 0;
        if (q < 0) set_err(s, fafk_last_error());
        atomic_fetch_add(&s->device_ns, now_ns() - sl->t_launched);
        finish(s, sl->reqs + from, sl->nreq - from, q == 1 ? ST_DONE : ST_FAILED);
        sl->busy = 0;
        did = 1;
    }
    return did;
}

static int maybe_launch(faf_server *s) {
    if (!s->ready.n) return 0;
    slot *free_sl = NULL;
    int busy = 0;
    for (int i = 0; i < s->nslots; ++i) {
        if (s->slots[i].busy) ++busy;
        else if (!free_sl) free_sl = &s->slots[i];
    }
    if (!free_sl) return 0;
    int take;
    const int plain = atomic_load(&s->plain);
    if (plain) {
        take = 1;
    } else {
        if (busy) {
            uint64_t age = now_ns() - rec(s, ring_at(&s->ready, 0))->t_arr;
            if ((double)age < atomic_load(&s->window_ns) && s->ready.n < (size_t)s->cfg.max_batch)
                return 0;
        }
        take = s->ready.n < (size_t)s->cfg.max_batch ? (int)s->ready.n : s->cfg.max_batch;
    }
    int rc = launch(s, free_sl, take);
    if (rc == -2) return 0;
    if (rc) {
        /* the popped requests fail once whatever was queued has finished
         * with their memory; the slot stays free */
        fafk_sync(free_sl->stream);
        finish(s, free_sl->reqs, take, ST_FAILED);
        free_sl->busy = 0;
        return 1;
    }
    if (plain) {
        /* plain per-call service: wait for this request before the next */
        rc = fafk_event_sync(free_sl->ev_done);
        if (rc) set_err(s, fafk_last_error());
        atomic_fetch_add(&s->device_ns, now_ns() - free_sl->t_launched);
        finish(s, free_sl->reqs, free_sl->nreq, rc ? ST_FAILED : ST_DONE);
        free_sl->busy = 0;
    }
    return 1;
}

static int has_busy(faf_server *s) {
    for (int i = 0; i < s->nslots; ++i)
        if (s->slots[i].busy) return 1;
    return 0;
}

static void *worker_main(void *arg) {
    faf_server *s = arg;
    for (;;) {
        int did = take_inbox(s);
        did |= unpark(s);
        did |= admit(s);
        did |= poll_slots(s);
        did |= maybe_launch(s);
        if (did) continue;
        if (atomic_load(&s->stop) && !s->ready.n && !has_busy(s)) {
            /* shutting down: what is still waiting for budget or a plan fails */
            while (s->pend.n) {
                int64_t t = ring_pop(&s->pend);
                rec(s, t)->freed = 1;
                finish(s, &t, 1, ST_FAILED);
            }
            while (s->parked.n) {
                int64_t t = ring_pop(&s->parked);
                rec(s, t)->freed = 1;
                finish(s, &t, 1, ST_FAILED);
            }
            break;
        }
        if (!s->pend.n && !s->ready.n && !has_busy(s)) {
            /* idle: sleep until a submission, release or stop */
            pthread_mutex_lock(&s->in_mu);
            if (!s->inbox.n && !s->rel.n && !atomic_load(&s->stop) &&
                (!s->parked.n || s->plans_built == s->plans_seen)) {
                s->sleeping = 1;
                pthread_cond_wait(&s->in_cv, &s->in_mu);
                s->sleeping = 0;
            }
            pthread_mutex_unlock(&s->in_mu);
        } else {
            for (int i = 0; i < 32; ++i) cpu_relax();
        }
    }
    return NULL;
}

/* ------------------------------------------------------------------ */
/* public API                                                           */
/* ------------------------------------------------------------------ */

faf_server_config faf_server_config_init(void) {
    faf_server_config c;
    memset(&c, 0, sizeof c);
    c.hbm_budget = (size_t)64 << 30;
    c.window_us = 60.0;
    c.streams = 3;
    c.max_batch = 1024;
    c.plain = 0;
    c.stft_win = 512;
    c.stft_hop = 128;
    c.cwt_fs = 16000.0;
    c.cwt_voices = 8;
    c.cwt_octaves = 31.5 / 8.0;
    return c;
}

void faf_server_destroy(faf_server *s) {
    if (!s) return;
    if (s->started) {
        pthread_mutex_lock(&s->in_mu);
        atomic_store(&s->stop, 1);
        pthread_cond_signal(&s->in_cv);
        pthread_mutex_unlock(&s->in_mu);
        pthread_join(s->th, NULL);
    }
    pthread_mutex_lock(&s->plan_mu);          /* builders finish the build they hold */
    s->build_stop = 1;
    pthread_cond_broadcast(&s->build_cv);
    pthread_mutex_unlock(&s->plan_mu);
    for (int i = 0; i < s->nbuilders; ++i) pthread_join(s->builders[i], NULL);
    free(s->bq);
    for (int i = 0; i < s->nslots; ++i) {
        slot *sl = &s->slots[i];
        if (sl->stream) fafk_sync(sl->stream);
        fafk_event_destroy(sl->ev_small);
        fafk_event_destroy(sl->ev_done);
        fafk_stream_destroy(sl->stream);
        fafk_host_free(sl->tab);
        free(sl->reqs);
    }
    for (int i = 0; i < s->nent; ++i) {
        fafi_plan_free(s->ent[i]->plan);
        free(s->ent[i]);
    }
    for (int i = 0; i < SV_MAX_PAGES; ++i) free(s->pages[i]);
    fafk_free(s->splans_dev);
    fafk_free(s->arena);
    free(s->fr);
    free(s->inbox.v); free(s->rel.v); free(s->pend.v); free(s->ready.v); free(s->parked.v);
    free(s->bk_e); free(s->bk_cnt); free(s->bk_off);
    free(s->sigs); free(s->units); free(s->ucost); free(s->uorder); free(s->ulds);
    free(s->uthr); free(s->uphase);
    pthread_mutex_destroy(&s->plan_mu);
    pthread_mutex_destroy(&s->splan_mu);
    pthread_cond_destroy(&s->plan_cv);
    pthread_cond_destroy(&s->build_cv);
    pthread_mutex_destroy(&s->in_mu);
    pthread_mutex_destroy(&s->done_mu);
    pthread_cond_destroy(&s->in_cv);
    pthread_cond_destroy(&s->done_cv);
    free(s);
}

faf_server *faf_server_create(const faf_server_config *cfg) {
    if (fafi_create_ready()) return NULL;
    faf_server_config c = cfg ? *cfg : faf_server_config_init();
    faf_server_config d = faf_server_config_init();
    if (!c.hbm_budget) c.hbm_budget = d.hbm_budget;
    if (c.window_us < 0) c.window_us = 0;
    if (c.streams <= 0) c.streams = d.streams;
    if (c.streams > SV_MAX_SLOTS) c.streams = SV_MAX_SLOTS;
    if (c.max_batch <= 0) c.max_batch = d.max_batch;
    if (!c.stft_win) c.stft_win = d.stft_win;
    if (!c.stft_hop) c.stft_hop = d.stft_hop;
    if (c.cwt_fs <= 0) c.cwt_fs = d.cwt_fs;
    if (!c.cwt_voices) c.cwt_voices = d.cwt_voices;
    if (c.cwt_octaves <= 0) c.cwt_octaves = d.cwt_octaves;
    faf_server *s = calloc(1, sizeof *s);
    if (!s) { fafi_err("serve: alloc failed"); return NULL; }
    s->cfg = c;
    pthread_mutex_init(&s->plan_mu, NULL);
    pthread_mutex_init(&s->splan_mu, NULL);
    pthread_cond_init(&s->plan_cv, NULL);
    pthread_cond_init(&s->build_cv, NULL);
    pthread_mutex_init(&s->in_mu, NULL);
    pthread_mutex_init(&s->done_mu, NULL);
    pthread_cond_init(&s->in_cv, NULL);
    pthread_cond_init(&s->done_cv, NULL);
    for (int i = 0; i < SV_MAX_PAGES; ++i) s->page_no[i] = -1;
    s->arena_bytes = c.hbm_budget & ~(size_t)(SV_ALIGN - 1);
    s->arena = fafk_malloc(s->arena_bytes);
    s->splans_dev = fafk_malloc(SV_MAX_ENTRIES * sizeof(fafd_splan));
    s->fr = malloc(64 * sizeof(frange));
    s->frcap = 64;
    if (!s->arena || !s->splans_dev || !s->fr) {
        fafi_err("serve: cannot reserve %zu bytes of device memory: %s", s->arena_bytes,
                 fafk_last_error());
        faf_server_destroy(s);
        return NULL;
    }
    s->fr[0].off = 0;
    s->fr[0].len = s->arena_bytes;
    s->nfr = 1;
    s->nslots = c.streams;
    atomic_store(&s->plain, c.plain ? 1 : 0);
    atomic_store(&s->window_ns, c.window_us * 1e3);
    for (int i = 0; i < s->nslots; ++i) {
        slot *sl = &s->slots[i];
        sl->stream = fafk_stream_create();
        sl->ev_small = fafk_event_create();
        sl->ev_done = fafk_event_create();
        if (!sl->stream || !sl->ev_small || !sl->ev_done) {
            fafi_err("serve: cannot create streams/events");
            faf_server_destroy(s);
            return NULL;
        }
    }
    if (pthread_create(&s->th, NULL, worker_main, s)) {
        fafi_err("serve: cannot start the worker thread");
        faf_server_destroy(s);
        return NULL;
    }
    s->started = 1;
    return s;
}

int faf_server_set_window(faf_server *s, double window_us) {
    if (!s || !(window_us >= 0.0)) { fafi_err("serve: bad window"); return -1; }
    faf_server_drain(s);
    atomic_store(&s->window_ns, window_us * 1e3);
    return 0;
}

int faf_server_set_plain(faf_server *s, int on) {
    if (!s) { fafi_err("serve: null server"); return -1; }
    faf_server_drain(s);
    atomic_store(&s->plain, on ? 1 : 0);
    return 0;
}

int faf_server_prepare(faf_server *s, faf_req_kind kind, size_t n) {
    if (!s) { fafi_err("serve: null server"); return -1; }
    return entry_get(s, (int)kind, n) ? 0 : -1;
}

int faf_server_prepare_many(faf_server *s, const int *kinds, const size_t *ns, size_t count) {
    if (!s || (count && (!kinds || !ns))) { fafi_err("serve: null argument"); return -1; }
    /* every new plan goes to the builder threads first, then each is waited
     * for; the first failure's message is the one reported */
    int rc = 0;
    char msg[192] = "";
    for (size_t i = 0; i < count; ++i)
        if (!entry_acquire(s, kinds[i], ns[i], 0) && !rc) {
            rc = -1;
            snprintf(msg, sizeof msg, "%s", faf_get_error());
        }
    for (size_t i = 0; i < count; ++i)
        if (!entry_acquire(s, kinds[i], ns[i], 1) && !rc) {
            rc = -1;
            snprintf(msg, sizeof msg, "%s", faf_get_error());
        }
    if (rc) fafi_err("%s", msg);
    return rc;
}

size_t faf_server_result_bytes(faf_server *s, faf_req_kind kind, size_t n) {
    if (!s) return 0;
    sentry *e = entry_get(s, (int)kind, n);
    return e ? result_bytes(s, e, n) : 0;
}

/* Frees the page in slot `sl` if every request in it has finished with its
 * result freed and the stats window starts past it.  Called with in_mu held
 * (no submit can touch the page); takes done_mu for the records. */
static int page_reclaim(faf_server *s, int sl) {
    const int64_t pg = s->page_no[sl];
    int ok = 1;
    pthread_mutex_lock(&s->done_mu);
    if ((pg + 1) * SV_PAGE > s->stats_first) ok = 0;
    for (int64_t i = 0; ok && i < SV_PAGE; ++i) {
        srec *r = &s->pages[sl][i];
        int st = atomic_load(&r->state);
        if ((st != ST_DONE && st != ST_FAILED) || !r->freed) ok = 0;
    }
    if (ok) {
        free(s->pages[sl]);
        s->pages[sl] = NULL;
        s->page_no[sl] = -1;
    }
    pthread_mutex_unlock(&s->done_mu);
    return ok;
}

static int64_t submit_at(faf_server *s, int kind, size_t n, const float *in, uint32_t flags,
                         uint64_t t_arr) {
    if (!s || !in) { fafi_err("serve: null argument"); return -1; }
    if (kind == FAF_REQ_STFT && n < 1) { fafi_err("serve: empty STFT record"); return -1; }
    sentry *e = entry_acquire(s, kind, n, 0);    /* may still be building */
    if (!e) return -1;
    /* the ticket is issued only once its record exists: a submission that
     * fails here leaves no ticket behind for drain or stats to wait on */
    pthread_mutex_lock(&s->in_mu);
    const int64_t t = atomic_load(&s->next_ticket);
    const int64_t pg = t >> SV_PAGE_BITS;
    const int sl = (int)(pg % SV_MAX_PAGES);
    if (s->page_no[sl] != pg) {
        if (s->pages[sl] && !page_reclaim(s, sl)) {
            pthread_mutex_unlock(&s->in_mu);
            fafi_err("serve: %lld requests are unreleased or inside the stats window; "
                     "release results (or submit with FAF_REQ_DISCARD) and reset the stats",
                     (long long)(SV_MAX_PAGES * SV_PAGE));
            return -1;
        }
        s->pages[sl] = calloc((size_t)SV_PAGE, sizeof(srec));
        if (!s->pages[sl]) {
            pthread_mutex_unlock(&s->in_mu);
            fafi_err("serve: alloc failed");
            return -1;
        }
        s->page_no[sl] = pg;
    }
    srec *r = rec(s, t);
    memset(r, 0, sizeof *r);
    r->kind = kind;
    r->flags = flags;
    r->n = n;
    r->in = in;
    r->e = e;
    if (atomic_load(&e->state) == 1) size_request(s, r);   /* else once its plan is built */
    r->t_arr = t_arr;
    atomic_store(&r->state, ST_QUEUED);
    if (ring_push(&s->inbox, t)) {
        /* not queued and no ticket issued: the record slot is reused */
        pthread_mutex_unlock(&s->in_mu);
        fafi_err("serve: alloc failed");
        return -1;
    }
    atomic_store(&s->next_ticket, t + 1);
    if (s->sleeping) pthread_cond_signal(&s->in_cv);
    pthread_mutex_unlock(&s->in_mu);
    return t;
}

int64_t faf_server_submit(faf_server *s, faf_req_kind kind, size_t n, const float *in,
                          uint32_t flags) {
    return submit_at(s, (int)kind, n, in, flags, now_ns());
}

static srec *valid(faf_server *s, int64_t t) {
    if (!s || t < 0 || t >= atomic_load(&s->next_ticket)) {
        fafi_err("serve: invalid ticket %lld", (long long)t);
        return NULL;
    }
    if (!rec_live(s, t)) {
        fafi_err("serve: ticket %lld has expired (released and recycled)", (long long)t);
        return NULL;
    }
    return rec(s, t);
}

int faf_server_poll(faf_server *s, int64_t t) {
    srec *r = valid(s, t);
    if (!r) return -1;
    int st = atomic_load(&r->state);
    return st == ST_DONE ? 1 : (st == ST_FAILED ? -1 : 0);
}

int faf_server_wait(faf_server *s, int64_t t) {
    srec *r = valid(s, t);
    if (!r) return -1;
    char msg[256];
    pthread_mutex_lock(&s->done_mu);
    while (atomic_load(&r->state) < ST_DONE) pthread_cond_wait(&s->done_cv, &s->done_mu);
    snprintf(msg, sizeof msg, "%s", s->err[0] ? s->err : "serve: request failed");
    pthread_mutex_unlock(&s->done_mu);
    if (atomic_load(&r->state) == ST_FAILED) {
        fafi_err("%s", msg);
        return -1;
    }
    return 0;
}

const float *faf_server_result(faf_server *s, int64_t t, size_t *bytes) {
    srec *r = valid(s, t);
    if (!r || atomic_load(&r->state) != ST_DONE || r->freed) return NULL;
    if (bytes) *bytes = r->out_bytes;
    return (const float *)r->out;
}

int faf_server_release(faf_server *s, int64_t t) {
    srec *r = valid(s, t);
    if (!r) return -1;
    if (atomic_load(&r->state) < ST_DONE) { fafi_err("serve: request still in flight"); return -1; }
    pthread_mutex_lock(&s->in_mu);
    int rc = ring_push(&s->rel, t);
    if (s->sleeping) pthread_cond_signal(&s->in_cv);
    pthread_mutex_unlock(&s->in_mu);
    return rc;
}

int faf_server_drain(faf_server *s) {
    if (!s) return -1;
    pthread_mutex_lock(&s->done_mu);
    while (s->finished < atomic_load(&s->next_ticket))
        pthread_cond_wait(&s->done_cv, &s->done_mu);
    pthread_mutex_unlock(&s->done_mu);
    return 0;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int faf_server_stats_get(faf_server *s, faf_server_stats *st) {
    if (!s || !st) return -1;
    memset(st, 0, sizeof *st);
    int64_t hi = atomic_load(&s->next_ticket), lo = s->stats_first;
    uint64_t *lat = malloc((size_t)(hi - lo + 1) * sizeof(uint64_t));
    if (!lat) return -1;
    size_t m = 0;
    uint64_t t0 = UINT64_MAX, t1 = 0;
    double bytes = 0.0, sum = 0.0;
    pthread_mutex_lock(&s->done_mu);
    for (int64_t t = lo; t < hi; ++t) {
        if (!rec_live(s, t)) continue;
        srec *r = rec(s, t);
        int state = atomic_load(&r->state);
        st->submitted++;
        if (state == ST_FAILED) { st->failed++; continue; }
        if (state != ST_DONE) continue;
        uint64_t l = r->t_done > r->t_arr ? r->t_done - r->t_arr : 0;
        lat[m++] = l;
        sum += (double)l;
        bytes += r->io_bytes;
        if (r->t_arr < t0) t0 = r->t_arr;
        if (r->t_done > t1) t1 = r->t_done;
    }
    st->peak_hbm_bytes = atomic_load(&s->peak);
    pthread_mutex_unlock(&s->done_mu);
    st->completed = m;
    st->launches = atomic_load(&s->launches);
    uint64_t lr = atomic_load(&s->launched_reqs);
    st->mean_batch = st->launches ? (double)lr / (double)st->launches : 0.0;
    st->admission_waits = atomic_load(&s->admission_waits);
    st->long_requests = atomic_load(&s->long_reqs);
    if (st->launches) {
        st->host_us_per_batch = (double)atomic_load(&s->host_ns) / 1e3 / (double)st->launches;
        st->device_us_per_batch =
            (double)atomic_load(&s->device_ns) / 1e3 / (double)st->launches;
    }
    if (m) {
        qsort(lat, m, sizeof(uint64_t), cmp_u64);
        st->p50_us = (double)lat[(m - 1) / 2] / 1e3;
        st->p99_us = (double)lat[(size_t)ceil(0.99 * (double)m) - 1] / 1e3;
        st->max_us = (double)lat[m - 1] / 1e3;
        st->mean_us = sum / (double)m / 1e3;
        st->span_s = t1 > t0 ? (double)(t1 - t0) / 1e9 : 0.0;
        if (st->span_s > 0) {
            st->req_per_s = (double)m / st->span_s;
            st->gb_per_s = bytes / st->span_s / 1e9;
        }
    }
    free(lat);
    return 0;
}

void faf_server_stats_reset(faf_server *s) {
    if (!s) return;
    faf_server_drain(s);
    pthread_mutex_lock(&s->done_mu);
    s->stats_first = atomic_load(&s->next_ticket);
    atomic_store(&s->peak, atomic_load(&s->in_use));
    pthread_mutex_unlock(&s->done_mu);
    /* pages now behind the stats window whose results are all freed go */
    pthread_mutex_lock(&s->in_mu);
    for (int i = 0; i < SV_MAX_PAGES; ++i)
        if (s->pages[i]) (void)page_reclaim(s, i);
    pthread_mutex_unlock(&s->in_mu);
    atomic_store(&s->launches, 0);
    atomic_store(&s->launched_reqs, 0);
    atomic_store(&s->long_reqs, 0);
    atomic_store(&s->host_ns, 0);
    atomic_store(&s->device_ns, 0);
    atomic_store(&s->admission_waits, 0);
}

int faf_server_replay(faf_server *s, const int *kinds, const size_t *ns,
                      const float *const *ins, const uint64_t *arrival_ns, size_t count,
                      uint32_t flags, uint64_t *latency_ns) {
    if (!s || (count && (!kinds || !ns || !ins || !arrival_ns))) {
        fafi_err("serve: null argument");
        return -1;
    }
    int64_t *tk = malloc((count ? count : 1) * sizeof(int64_t));
    if (!tk) { fafi_err("serve: alloc failed"); return -1; }
    int rc = 0;
    const uint64_t t0 = now_ns();
    for (size_t i = 0; i < count; ++i) {
        const uint64_t at = t0 + arrival_ns[i];
        while (now_ns() < at) cpu_relax();
        tk[i] = submit_at(s, kinds[i], ns[i], ins[i], flags, at);
        if (tk[i] < 0) rc = -1;
    }
    faf_server_drain(s);
    for (size_t i = 0; i < count; ++i) {
        if (tk[i] < 0 || !rec_live(s, tk[i])) {
            if (latency_ns) latency_ns[i] = 0;
            continue;
        }
        srec *r = rec(s, tk[i]);
        if (atomic_load(&r->state) != ST_DONE) rc = -1;
        if (latency_ns) latency_ns[i] = r->t_done > r->t_arr ? r->t_done - r->t_arr : 0;
    }
    free(tk);
    if (rc) {
        char msg[256];
        pthread_mutex_lock(&s->done_mu);
        snprintf(msg, sizeof msg, "%s", s->err[0] ? s->err : "serve: request failed");
        pthread_mutex_unlock(&s->done_mu);
        if (!faf_get_error()[0]) fafi_err("%s", msg);
    }
    return rc;
}
