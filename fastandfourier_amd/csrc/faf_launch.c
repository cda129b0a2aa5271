/*
 * faf_launch.c -- run a planned transform over device buffers (C11).
 *
 * Lengths that fit one workgroup's LDS run as a single kernel launch per
 * batch.  Longer lengths use the four-step split n = n1 * n2 (pass A over
 * columns, pass B over rows) through a device workspace, processed in
 * chunks of signals so the intermediate stays small enough to be served
 * largely from the MI355X's 256 MB last-level cache.
 */
#include "faf_internal.h"
#include <string.h>

/* workspace per four-step chunk: small enough that the intermediate the
 * column pass writes is still in the 256 MB Infinity Cache when the row
 * pass reads it (measured: 64 MB beats 256 MB by 27% and 16 MB, whose
 * launches are too short, by 65%) */
#ifndef CHUNK_BYTES
#define CHUNK_BYTES ((size_t)64 << 20)
#endif

static int kerr(int rc) {
    if (rc) fafi_err("%s", fafk_last_error());
    return rc;
}

static int kcerr(int rc) {
    if (rc) fafi_err("%s", fafk_ct_last_error());
    return rc;
}

static int krerr(int rc) {
    if (rc) fafi_err("%s", fafk_rs_last_error());
    return rc;
}

static int rs_grid(void) { return (int)fafi_cus(); }

static size_t big_bytes(const fafh_plan *p, long cs, size_t L);
static long big_chunk(const fafh_plan *p, size_t L, long count);

/* signals per four-step chunk; per_sig = workspace bytes one signal takes.
 * A lent workspace (the serving arena, the static scratch of the
 * allocation-free call) caps the chunk at what it holds. */
static long chunk_signals(const fafh_plan *p, size_t n_complex, size_t per_sig, long count) {
    long c = (long)(CHUNK_BYTES / (n_complex * 2 * sizeof(float)));
    if (p->ws_ext) {
        const long cap = (long)(p->ws_ext_bytes / per_sig);
        if (cap < c) c = cap;
    }
    if (c < 1) c = 1;
    return c < count ? c : count;
}

size_t fafi_run_ws_bytes(const fafh_plan *p, long count) {
    if (count <= 0) return 0;
    switch (p->kind) {
    case PK_C2C_4S:
        if (p->rs) return 0;
        return (size_t)chunk_signals(p, p->n, p->n * 8, count) * p->n * 2 * sizeof(float);
    case PK_R2C_4S:
    case PK_C2R_4S:
        /* with the single pass the workspace is needed only for rows that are
         * not 8-byte aligned, which fall back to four-step: keep its size */
        if (p->ct4)
            return (size_t)chunk_signals(p, p->n / 2, p->n / 2 * 8, count) * (p->n / 2) * 2 *
                   sizeof(float);
        return (size_t)chunk_signals(p, p->n / 2, p->n / 2 * 16, count) * (p->n / 2) * 4 *
               sizeof(float);
    case PK_CWT:
        return fafi_cwt_ws_bytes(p, count, 1);
    case PK_RADER_L:
    case PK_BLUE_L: {
        const size_t L = p->kind == PK_RADER_L ? p->n - 1 : (size_t)p->blue.m;
        return big_bytes(p, big_chunk(p, L, count), L);
    }
    default:
        return 0;
    }
}

/* ---- Rader and Bluestein beyond one workgroup ----
 * Per chunk of signals: the gather (Rader: a[j] = x[g^j]) or the chirp
 * (Bluestein: a[k] = x[k] chirp[k], zero-padded to M) into workspace plane
 * pair A, the inner forward FFT into pair B, the product with the
 * precomputed kernel spectrum (Rader saves each signal's DC sum first), the
 * unnormalised inner inverse back into A, then the scatter (Rader:
 * y[g^-j] = x0 + a[j], y[0] = x0 + sum) or the chirp out, with the plan's
 * norm.  Reference src/faf_fft_rader.c:136-232, src/faf_bluestein.c:165-223.
 * Workspace: two plane pairs of L = p-1 or M per signal, the DC sums, then
 * the inner plans' own workspace. */
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t big_bytes(const fafh_plan *p, long cs, size_t L) {
    size_t a = align256((size_t)cs * L * 16), z = align256((size_t)cs * 8);
    size_t w1 = fafi_run_ws_bytes(p->inner_f, cs), w2 = fafi_run_ws_bytes(p->inner_i, cs);
    return a + z + (w1 > w2 ? w1 : w2);
}

static long big_chunk(const fafh_plan *p, size_t L, long count) {
    long cs = (long)(CHUNK_BYTES / (L * 16));
    if (cs < 1) cs = 1;
    if (cs > count) cs = count;
    if (p->ws_ext)
        while (cs > 1 && big_bytes(p, cs, L) > p->ws_ext_bytes) cs = (cs + 1) / 2;
    return cs;
}

/* runs an inner plan on lent workspace */
static int inner_run(fafh_plan *q, void *ws, size_t wb, const float *ir, const float *ii,
                     float *orr, float *oi, long dist, long count, void *stream) {
    q->ws_ext = ws;
    q->ws_ext_bytes = wb;
    int rc = fafi_run_c2c(q, ir, ii, dist, orr, oi, dist, count, stream);
    q->ws_ext = NULL;
    q->ws_ext_bytes = 0;
    return rc;
}

static int run_big(fafh_plan *p, const float *ir, const float *ii, long idist, float *orr,
                   float *oi, long odist, long count, void *stream) {
    const int rader = p->kind == PK_RADER_L;
    const size_t n = p->n, L = rader ? n - 1 : (size_t)p->blue.m;
    const long cs = big_chunk(p, L, count);
    const size_t need = big_bytes(p, cs, L);
    char *ws = fafi_workspace(p, need);
    if (!ws) return -1;
    float *ar = (float *)ws, *ai = ar + (size_t)cs * L;
    float *br = ai + (size_t)cs * L, *bi = br + (size_t)cs * L;
    const size_t ab = align256((size_t)cs * L * 16);
    void *z = ws + ab;
    char *sub = ws + ab + align256((size_t)cs * 8);
    const size_t subb = need - ab - align256((size_t)cs * 8);
    const long l = (long)L;
    for (long s0 = 0; s0 < count; s0 += cs) {
        const long c = count - s0 < cs ? count - s0 : cs;
        const float *xr = ir + s0 * idist, *xi = ii + s0 * idist;
        float *yr = orr + s0 * odist, *yi = oi + s0 * odist;
        int rc = rader ? fafk_big_gather(xr, xi, idist, p->rader.rin, l, ar, ai, l, c, stream)
                       : fafk_big_chirp_in(xr, xi, idist, p->blue.chirp, (long)n, l, ar, ai, l,
                                           c, stream);
        if (kerr(rc)) return -1;
        if (inner_run(p->inner_f, sub, subb, ar, ai, br, bi, l, c, stream)) return -1;
        if (kerr(fafk_big_mul(br, bi, l, rader ? p->rader.b : p->blue.b, l,
                              rader ? z : NULL, c, stream)))
            return -1;
        if (inner_run(p->inner_i, sub, subb, br, bi, ar, ai, l, c, stream)) return -1;
        rc = rader ? fafk_big_rader_out(ar, ai, l, p->rader.rout, l, xr, xi, idist, z, yr, yi,
                                        odist, p->scale, c, stream)
                   : fafk_big_chirp_out(ar, ai, l, p->blue.chirp, (long)n, yr, yi, odist,
                                        p->scale, c, stream);
        if (kerr(rc)) return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */

int fafi_run_c2c(fafh_plan *p, const float *ir, const float *ii, long idist, float *orr,
                 float *oi, long odist, long count, void *stream) {
    const int inter = p->interleaved;
    switch (p->kind) {
    case PK_C2C:
        if (p->ct && p->n == 8192)   /* one pass, every count and layout */
            return krerr(fafk_rs8_c2c(p->inverse, inter, p->a.tw, ir, ii, idist, orr, oi, odist,
                                      count, p->scale, stream));
        if (p->ct)
            return kcerr(fafk_ct_c2c(&p->a, p->inverse, inter, ir, ii, idist, orr, oi, odist,
                                     count, p->scale, stream));
        return kerr(fafk_c2c(&p->a, p->inverse, inter, p->spb, ir, ii, idist, orr, oi, odist,
                             count, p->scale, stream));
    case PK_RADER_L:
    case PK_BLUE_L:
        if (inter) {
            fafi_err("Rader/Bluestein lengths need SPLIT layout on the device path");
            return -1;
        }
        return run_big(p, ir, ii, idist, orr, oi, odist, count, stream);
    case PK_RADER:
    case PK_BLUE: {
        /* these kernels read/write split planes; interleaved buffers are
         * viewed as two strided planes is not supported: deinterleave */
        if (inter) {
            fafi_err("Rader/Bluestein lengths need SPLIT layout on the device path");
            return -1;
        }
        if (p->kind == PK_RADER)
            return kerr(fafk_rader(&p->a, &p->b, &p->rader, p->spb, ir, ii, idist, orr, oi,
                                   odist, count, p->scale, stream));
        return kerr(fafk_blue(&p->a, &p->b, &p->blue, p->spb, ir, ii, idist, orr, oi, odist,
                              count, p->scale, stream));
    }
    case PK_C2C_4S: {
        size_t n = p->n;
        if (p->rs)   /* one pass, no workspace */
            return krerr(fafk_rs_c2c(p->inverse, inter, p->twn, ir, ii, idist, orr, oi, odist,
                                     count, p->scale, rs_grid(), stream));
        long cs = chunk_signals(p, n, n * 8, count);
        float *ws = fafi_workspace(p, (size_t)cs * n * 2 * sizeof(float));
        if (!ws) return -1;
        if (p->ct4) {
            const long fi = idist * (inter ? 2 : 1), fo = odist * (inter ? 2 : 1);
            for (long b0 = 0; b0 < count; b0 += cs) {
                long c = count - b0 < cs ? count - b0 : cs;
                int rc = fafk_ct4(inter ? FAFK_F4_IN_INTER : FAFK_F4_IN_SPLIT,
                                  inter ? FAFK_F4_OUT_INTER : FAFK_F4_OUT_SPLIT, p->inverse,
                                  &p->a, &p->b, p->twn, NULL, NULL, ir + b0 * fi,
                                  inter ? NULL : ii + b0 * fi, fi, ws, (long)n, orr + b0 * fo,
                                  inter ? NULL : oi + b0 * fo, fo, c, p->scale, stream);
                if (kcerr(rc)) return -1;
            }
            return 0;
        }
        float *tr = ws, *ti = ws + (size_t)cs * n;
        for (long b0 = 0; b0 < count; b0 += cs) {
            long c = count - b0 < cs ? count - b0 : cs;
            const long fi = idist * (inter ? 2 : 1), fo = odist * (inter ? 2 : 1);
            int rc = fafk_fs_col(&p->a, p->inverse, inter ? FAFK_FS_IN_INTER : FAFK_FS_IN_SPLIT,
                                 p->n2, p->cols, p->twn, ir + b0 * fi,
                                 inter ? NULL : ii + b0 * fi, fi, tr, ti, (long)n, c, stream);
            if (!rc)
                rc = fafk_fs_row(&p->b, p->inverse,
                                 inter ? FAFK_FS_OUT_INTER : FAFK_FS_OUT_SPLIT, p->n1, p->rows,
                                 tr, ti, (long)n, orr + b0 * fo, inter ? NULL : oi + b0 * fo,
                                 fo, c, p->scale, stream);
            if (kerr(rc)) return -1;
        }
        return 0;
    }
    default:
        fafi_err("plan is not a C2C transform");
        return -1;
    }
}

int fafi_run_r2c(fafh_plan *p, const float *x, long idist, float *orr, float *oi, long odist,
                 long count, void *stream) {
    const int inter = p->interleaved;
    if (p->kind == PK_R2C && p->rs && fafk_rs_pair_ok(x, idist))   /* N = 16384, one pass */
        return krerr(fafk_rs8_r2c(inter, p->a.tw, p->post, x, idist, orr, oi, odist, count,
                                  p->scale, stream));
    if (p->kind == PK_R2C && p->ct)
        return kcerr(fafk_ct_r2c(&p->a, inter, p->post, x, idist, orr, oi, odist, count,
                                 p->scale, stream));
    if (p->kind == PK_R2C)
        return kerr(fafk_r2c(&p->a, inter, p->spb, p->post, x, idist, orr, oi, odist, count,
                             p->scale, stream));
    if (p->kind != PK_R2C_4S) { fafi_err("plan is not an R2C transform"); return -1; }
    size_t nc = p->n / 2;
    if (p->rs && fafk_rs_pair_ok(x, idist))   /* one pass, no workspace */
        return krerr(fafk_rs_r2c(inter, p->twn, p->post, x, idist, orr, oi, odist, count,
                                 p->scale, rs_grid(), stream));
    long cs = chunk_signals(p, nc, p->ct4 ? nc * 8 : nc * 16, count);
    if (p->ct4) {
        /* column pass, then the row pass with the Kiss post-pass fused */
        void *w = fafi_workspace(p, (size_t)cs * nc * 2 * sizeof(float));
        if (!w) return -1;
        const long fo = odist * (inter ? 2 : 1);
        for (long b0 = 0; b0 < count; b0 += cs) {
            long c = count - b0 < cs ? count - b0 : cs;
            int rc = fafk_ct4(FAFK_F4_IN_PAIR, inter ? FAFK_F4_POST_INTER : FAFK_F4_POST_SPLIT, 0,
                              &p->a, &p->b, p->twn, NULL, p->post, x + b0 * idist, NULL, idist,
                              w, (long)nc, orr + b0 * fo, inter ? NULL : oi + b0 * fo, fo, c,
                              p->scale, stream);
            if (kcerr(rc)) return -1;
        }
        return 0;
    }
    float *ws = fafi_workspace(p, (size_t)cs * nc * 4 * sizeof(float));
    if (!ws) return -1;
    float *tr = ws, *ti = ws + (size_t)cs * nc;
    float *zr = ws + 2 * (size_t)cs * nc, *zi = ws + 3 * (size_t)cs * nc;
    for (long b0 = 0; b0 < count; b0 += cs) {
        long c = count - b0 < cs ? count - b0 : cs;
        long ooff = b0 * odist * (inter ? 2 : 1);
        /* packed z[i] = x[2i] + i x[2i+1] is read as complex pairs */
        int rc = fafk_fs_col(&p->a, 0, FAFK_FS_IN_PAIR, p->n2, p->cols, p->twn,
                             x + b0 * idist, NULL, idist, tr, ti, (long)nc, c, stream);
        if (!rc)
            rc = fafk_fs_row(&p->b, 0, FAFK_FS_OUT_SPLIT, p->n1, p->rows, tr, ti, (long)nc, zr,
                             zi, (long)nc, c, 1.0f, stream);
        if (!rc)
            rc = fafk_r2c_post((int)nc, p->post, inter, zr, zi, (long)nc, orr + ooff,
                               inter ? NULL : oi + ooff, odist, c, p->scale, stream);
        if (kerr(rc)) return -1;
    }
    return 0;
}

int fafi_run_c2r(fafh_plan *p, const float *ir, const float *ii, long idist, float *y,
                 long odist, long count, void *stream) {
    const int inter = p->interleaved;
    if (p->kind == PK_C2R && p->rs && fafk_rs_pair_ok(y, odist)) {   /* one pass */
        if (p->n == 4096)
            return krerr(fafk_rs2_c2r(inter, p->a.tw, p->pre, ir, ii, idist, y, odist,
                                      count, p->scale, stream));
        return krerr(fafk_rs8_c2r(inter, p->a.tw, p->pre, ir, ii, idist, y, odist, count,
                                  p->scale, stream));
    }
    if (p->kind == PK_C2R && p->ct)
        return kcerr(fafk_ct_c2r(&p->a, inter, p->pre, ir, ii, idist, y, odist, count,
                                 p->scale, stream));
    if (p->kind == PK_C2R)
        return kerr(fafk_c2r(&p->a, inter, p->spb, p->pre, ir, ii, idist, y, odist, count,
                             p->scale, stream));
    if (p->kind != PK_C2R_4S) { fafi_err("plan is not a C2R transform"); return -1; }
    size_t nc = p->n / 2;
    if (p->rs && fafk_rs_pair_ok(y, odist))   /* one pass, no workspace */
        return krerr(fafk_rs_c2r(inter, p->twn, p->pre, ir, ii, idist, y, odist, count,
                                 p->scale, rs_grid(), stream));
    long cs = chunk_signals(p, nc, p->ct4 ? nc * 8 : nc * 16, count);
    if (p->ct4) {
        /* Kiss pre-pass fused into the column pass, real pairs out of the row pass */
        void *w = fafi_workspace(p, (size_t)cs * nc * 2 * sizeof(float));
        if (!w) return -1;
        const long fi = idist * (inter ? 2 : 1);
        for (long b0 = 0; b0 < count; b0 += cs) {
            long c = count - b0 < cs ? count - b0 : cs;
            int rc = fafk_ct4(inter ? FAFK_F4_PRE_INTER : FAFK_F4_PRE_SPLIT, FAFK_F4_OUT_PAIR, 1,
                              &p->a, &p->b, p->twn, p->pre, NULL, ir + b0 * fi,
                              inter ? NULL : ii + b0 * fi, fi, w, (long)nc, y + b0 * odist, NULL,
                              odist, c, p->scale, stream);
            if (kcerr(rc)) return -1;
        }
        return 0;
    }
    float *ws = fafi_workspace(p, (size_t)cs * nc * 4 * sizeof(float));
    if (!ws) return -1;
    float *fr = ws, *fi = ws + (size_t)cs * nc;
    float *tr = ws + 2 * (size_t)cs * nc, *ti = ws + 3 * (size_t)cs * nc;
    for (long b0 = 0; b0 < count; b0 += cs) {
        long c = count - b0 < cs ? count - b0 : cs;
        long ioff = b0 * idist * (inter ? 2 : 1);
        int rc = fafk_c2r_pre((int)nc, p->pre, inter, ir + ioff, inter ? NULL : ii + ioff,
                              idist, fr, fi, (long)nc, c, stream);
        if (!rc)
            rc = fafk_fs_col(&p->a, 1, FAFK_FS_IN_SPLIT, p->n2, p->cols, p->twn, fr, fi,
                             (long)nc, tr, ti, (long)nc, c, stream);
        if (!rc)
            rc = fafk_fs_row(&p->b, 1, FAFK_FS_OUT_PAIR, p->n1, p->rows, tr, ti, (long)nc,
                             y + b0 * odist, NULL, odist, c, p->scale, stream);
        if (kerr(rc)) return -1;
    }
    return 0;
}

int fafi_run_pipe(fafh_plan *p, const float *x, long idist, float *y, long odist, long count,
                  void *stream) {
    if (p->kind != PK_PIPE) { fafi_err("plan is not a fused pipeline"); return -1; }
    if (fafi_pipe_refresh(p, stream)) return -1;
    fafd_ops ops = p->ops;
    for (int i = 0; i < ops.nops; ++i)
        if (ops.code[i] == FAFD_OP_MUL && !ops.h_re) {
            ops.code[i] = FAFD_OP_NONE;   /* unbound spectrum: no-op, as the reference */
        }
    if (p->ct && p->n == 4096 && ops.nops == 1 && ops.code[0] == FAFD_OP_MUL &&
        fafk_rs_pair_ok(x, idist) && fafk_rs_pair_ok(y, odist))   /* one pass */
        return krerr(fafk_rs2_pipe(p->a.tw, p->b.tw, p->post, p->pre, &ops, x, idist, y, odist,
                                   count, p->scale, stream));
    if (p->ct)
        return kcerr(fafk_ct_pipe(&p->a, &p->b, p->post, p->pre, &ops, x, idist, y, odist,
                                  count, p->scale, stream));
    return kerr(fafk_pipe(&p->a, &p->b, p->spb, p->post, p->pre, &ops, x, idist, y, odist,
                          count, p->scale, stream));
}
