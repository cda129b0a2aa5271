/*
 * faf_kernels_ct.hip -- C-ABI launchers of the compile-time specialised
 * kernels (faf_engine_ct.h) for power-of-two lengths 256 ... 16384.
 * fafk_ct_supports() tells the host which (kind, length) pairs exist.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "faf_engine_ct.h"
#include "faf_kern.h"

static thread_local char g_ct_err[160];

static int ct_fail(const char *what, hipError_t e) {
    snprintf(g_ct_err, sizeof g_ct_err, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

extern "C" const char *fafk_ct_last_error(void) { return g_ct_err; }

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

static inline bool ct_len(int n) {
    return n == 256 || n == 512 || n == 1024 || n == 2048 || n == 4096 || n == 8192 ||
           n == 16384;
}

extern "C" int fafk_ct_supports(int kind, int n) {
    if (kind == 5) return n >= 512 && n <= 8192 && ct_len(n);   /* CWT segment length */
    return ct_len(n) ? 1 : 0;
}

template <class K> static void ct_attr(K k) {
    hipFuncAttributes fa;
    size_t stat = 0;
    if (hipFuncGetAttributes(&fa, (const void *)k) == hipSuccess) stat = fa.sharedSizeBytes;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(FAFD_MAX_LDS_BYTES - stat));
    (void)hipGetLastError();
}

template <int N> static void ct_attrs_n() {
    ct_attr(kc_c2c<false, false, N>); ct_attr(kc_c2c<true, false, N>);
    ct_attr(kc_c2c<false, true, N>); ct_attr(kc_c2c<true, true, N>);
    ct_attr(kc_r2c<false, N>); ct_attr(kc_r2c<true, N>);
    ct_attr(kc_c2r<false, N>); ct_attr(kc_c2r<true, N>);
    ct_attr(kc_pipe<N, false>); ct_attr(kc_pipe<N, true>);
    ct_attr(kc_stft<false, N>); ct_attr(kc_stft<true, N>);
}

template <int N> static void ct_attrs_cwt() {
    ct_attr(kc_cwt<false, N>); ct_attr(kc_cwt<true, N>);
}

template <int N1, int N2, int RW> static void ct_attrs_4r() {
    ct_attr(kc4_row<F4_POST_SPLIT, false, N1, N2, RW>);
    ct_attr(kc4_row<F4_POST_INTER, false, N1, N2, RW>);
    ct_attr(kc4_row<F4_OUT_SPLIT, false, N1, N2, RW>);
    ct_attr(kc4_row<F4_OUT_SPLIT, true, N1, N2, RW>);
    ct_attr(kc4_row<F4_OUT_INTER, false, N1, N2, RW>);
    ct_attr(kc4_row<F4_OUT_INTER, true, N1, N2, RW>);
    ct_attr(kc4_row<F4_OUT_PAIR, true, N1, N2, RW>);
}

template <int N1, int N2> static void ct_attrs_4() {
    ct_attr(kc4_col<F4_IN_PAIR, false, N1, N2>);
    ct_attr(kc4_col<F4_IN_SPLIT, false, N1, N2>); ct_attr(kc4_col<F4_IN_SPLIT, true, N1, N2>);
    ct_attr(kc4_col<F4_IN_INTER, false, N1, N2>); ct_attr(kc4_col<F4_IN_INTER, true, N1, N2>);
    ct_attr(kc4_col<F4_PRE_SPLIT, true, N1, N2>); ct_attr(kc4_col<F4_PRE_INTER, true, N1, N2>);
    ct_attrs_4r<N1, N2, F4_ROWS>();
}

template <int N> static void ct_attrs_r2cw() {
    ct_attr(kc_r2cw<false, N>); ct_attr(kc_r2cw<true, N>);
}
template <int N> static void ct_attrs_c2rw() {
    ct_attr(kc_c2rw<false, N>); ct_attr(kc_c2rw<true, N>);
}

extern "C" int fafk_ct_init(void) {
    ct_attrs_r2cw<1024>(); ct_attrs_r2cw<2048>(); ct_attrs_r2cw<8192>();
    ct_attrs_c2rw<512>(); ct_attrs_c2rw<1024>(); ct_attrs_c2rw<2048>();
    ct_attr(kc_pipew<2048>);
    ct_attrs_4<128, 256>(); ct_attrs_4<256, 256>();
    ct_attrs_cwt<512>(); ct_attrs_cwt<1024>(); ct_attrs_cwt<2048>(); ct_attrs_cwt<4096>();
    ct_attrs_cwt<8192>();
    ct_attrs_n<256>(); ct_attrs_n<512>(); ct_attrs_n<1024>(); ct_attrs_n<2048>();
    ct_attrs_n<4096>(); ct_attrs_n<8192>(); ct_attrs_n<16384>();
    ct_attr(kc_r2c<false, 2048, 4>);
    return 0;
}

template <int N> struct Shape {
    static constexpr int T = N / 16;
    static constexpr int SPB = ct_spb<T>();
};

#define CT_SWITCH(n, CALL)                                \
    switch (n) {                                          \
    case 256: CALL(256); break;                           \
    case 512: CALL(512); break;                           \
    case 1024: CALL(1024); break;                         \
    case 2048: CALL(2048); break;                         \
    case 4096: CALL(4096); break;                         \
    case 8192: CALL(8192); break;                         \
    case 16384: CALL(16384); break;                       \
    default:                                              \
        snprintf(g_ct_err, sizeof g_ct_err, "no specialised kernel for length %d", n); \
        return -1;                                        \
    }

#define GRID(N) dim3((unsigned)((count + Shape<N>::SPB - 1) / Shape<N>::SPB))
#define BLOCK(N) dim3(Shape<N>::T * Shape<N>::SPB)
#define LDS(N, plane) ((size_t)Shape<N>::SPB * (size_t)(plane) * sizeof(float2))

extern "C" int fafk_ct_c2c(const fafd_fft *P, int inv, int inter, const float *in_re,
                           const float *in_im, long in_dist, float *out_re, float *out_im,
                           long out_dist, long count, float scale, void *stream) {
#define C2C(N)                                                                           \
    do {                                                                                 \
        if (inv && inter) hipLaunchKernelGGL((kc_c2c<true, true, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale); \
        else if (inv) hipLaunchKernelGGL((kc_c2c<true, false, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale); \
        else if (inter) hipLaunchKernelGGL((kc_c2c<false, true, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale); \
        else hipLaunchKernelGGL((kc_c2c<false, false, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale); \
    } while (0)
    CT_SWITCH(P->n, C2C)
#undef C2C
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_c2c", e);
}

extern "C" int fafk_ct_r2c(const fafd_fft *P, int inter_out, const void *post, const float *x,
                           long in_dist, float *out_re, float *out_im, long out_dist,
                           long count, float scale, void *stream) {
    const float2 *po = (const float2 *)post;
    /* rfft 4096 into split rows: four signals per 512-thread workgroup.  On
     * packed rows with 16-byte aligned planes a whole workgroup's 4 x 2049
     * bins per plane leave through LDS as aligned float4 stores; otherwise
     * (tail workgroup, other distances or alignments) per bin.  Every split
     * call takes this kernel, whatever its count and layout, so a signal's
     * bits do not depend on the batch it is in. */
    if (P->n == 2048 && !inter_out) {
        hipLaunchKernelGGL((kc_r2c<false, 2048, 4>), dim3((unsigned)((count + 3) / 4)), dim3(512),
                           (size_t)4 * P->plane * sizeof(float2), S(stream), *P, po, x, in_dist,
                           out_re, out_im, out_dist, count, scale);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : ct_fail("kc_r2c (4 signals)", e);
    }
    /* one signal of NC/8 threads per workgroup where that measured faster
     * (rfft 2048, 4096, 16384; the staged kernel wins at 1024) */
    if (P->n == 1024 || P->n == 2048 || P->n == 8192) {
        const dim3 g((unsigned)count), b((unsigned)(P->n / 8));
        const size_t lds = (size_t)P->plane * sizeof(float2);
#define R2CW(N)                                                                          \
        do {                                                                             \
            if (inter_out) hipLaunchKernelGGL((kc_r2cw<true, N>), g, b, lds, S(stream), *P, po, x, in_dist, out_re, out_im, out_dist, count, scale); \
            else hipLaunchKernelGGL((kc_r2cw<false, N>), g, b, lds, S(stream), *P, po, x, in_dist, out_re, out_im, out_dist, count, scale); \
        } while (0)
        if (P->n == 1024) R2CW(1024);
        else if (P->n == 2048) R2CW(2048);
        else R2CW(8192);
#undef R2CW
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : ct_fail("kc_r2cw", e);
    }
#define R2C(N)                                                                           \
    do {                                                                                 \
        if (inter_out) hipLaunchKernelGGL((kc_r2c<true, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, po, x, in_dist, out_re, out_im, out_dist, count, scale); \
        else hipLaunchKernelGGL((kc_r2c<false, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, po, x, in_dist, out_re, out_im, out_dist, count, scale); \
    } while (0)
    CT_SWITCH(P->n, R2C)
#undef R2C
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_r2c", e);
}

extern "C" int fafk_ct_c2r(const fafd_fft *P, int inter_in, const void *pre, const float *in_re,
                           const float *in_im, long in_dist, float *y, long out_dist, long count,
                           float scale, void *stream) {
    const float2 *pr = (const float2 *)pre;
    /* one signal of NC/8 threads per workgroup where that measured faster
     * (irfft 1024, 2048, 4096; the staged kernel wins at 16384) */
    if (P->n == 512 || P->n == 1024 || P->n == 2048) {
        const dim3 g((unsigned)count), b((unsigned)(P->n / 8));
        const size_t lds = (size_t)P->plane * sizeof(float2);
#define C2RW(N)                                                                          \
        do {                                                                             \
            if (inter_in) hipLaunchKernelGGL((kc_c2rw<true, N>), g, b, lds, S(stream), *P, pr, in_re, in_im, in_dist, y, out_dist, count, scale); \
            else hipLaunchKernelGGL((kc_c2rw<false, N>), g, b, lds, S(stream), *P, pr, in_re, in_im, in_dist, y, out_dist, count, scale); \
        } while (0)
        if (P->n == 512) C2RW(512);
        else if (P->n == 1024) C2RW(1024);
        else C2RW(2048);
#undef C2RW
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : ct_fail("kc_c2rw", e);
    }
#define C2R(N)                                                                           \
    do {                                                                                 \
        if (inter_in) hipLaunchKernelGGL((kc_c2r<true, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, pr, in_re, in_im, in_dist, y, out_dist, count, scale); \
        else hipLaunchKernelGGL((kc_c2r<false, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, pr, in_re, in_im, in_dist, y, out_dist, count, scale); \
    } while (0)
    CT_SWITCH(P->n, C2R)
#undef C2R
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_c2r", e);
}

extern "C" int fafk_ct_pipe(const fafd_fft *Pf, const fafd_fft *Pi, const void *post,
                            const void *pre, const fafd_ops *ops, const float *x, long in_dist,
                            float *y, long out_dist, long count, float scale, void *stream) {
    const float2 *po = (const float2 *)post, *pr = (const float2 *)pre;
    const bool mul1 = ops->nops == 1 && ops->code[0] == FAFD_OP_MUL;
    if (mul1 && Pf->n == 2048) {               /* one signal of 256 threads per workgroup */
        hipLaunchKernelGGL((kc_pipew<2048>), dim3((unsigned)count), dim3(256),
                           (size_t)Pf->plane * sizeof(float2), S(stream), *Pf, *Pi, po, pr, *ops,
                           x, in_dist, y, out_dist, count, scale);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : ct_fail("kc_pipew", e);
    }
#define PIPE(N)                                                                          \
    do {                                                                                 \
        if (mul1) hipLaunchKernelGGL((kc_pipe<N, true>), GRID(N), BLOCK(N), LDS(N, Pf->plane), S(stream), *Pf, *Pi, po, pr, *ops, x, in_dist, y, out_dist, count, scale); \
        else hipLaunchKernelGGL((kc_pipe<N, false>), GRID(N), BLOCK(N), LDS(N, Pf->plane), S(stream), *Pf, *Pi, po, pr, *ops, x, in_dist, y, out_dist, count, scale); \
    } while (0)
    CT_SWITCH(Pf->n, PIPE)
#undef PIPE
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_pipe", e);
}

extern "C" int fafk_ct_stft(const fafd_fft *P, const void *post, const float *win,
                            const float *x, long rec_len, long rec_dist, long hop, long frames,
                            float *out_re, float *out_im, long frame_dist, long rec_out_dist,
                            long count, void *stream) {
    const float2 *po = (const float2 *)post;
    /* with a record at least one window long every frame is whole (the
     * 32768-sample window keeps the checked loader: the other spills) */
    const bool full = rec_len >= 2L * P->n && P->n <= 8192;
#define STFT(N)                                                                          \
    do {                                                                                 \
        if (full) hipLaunchKernelGGL((kc_stft<true, N>), GRID(N), BLOCK(N), LDS(N, P->plane) + (N <= 512 ? 8 * N : 0), S(stream), *P, po, win, x, rec_len, rec_dist, hop, frames, out_re, out_im, frame_dist, rec_out_dist, count); \
        else hipLaunchKernelGGL((kc_stft<false, N>), GRID(N), BLOCK(N), LDS(N, P->plane), S(stream), *P, po, win, x, rec_len, rec_dist, hop, frames, out_re, out_im, frame_dist, rec_out_dist, count); \
    } while (0)
    CT_SWITCH(P->n, STFT)
#undef STFT
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_stft", e);
}

/* CWT bank rows of power-of-two segments 512 ... 8192 (kind 5) */
extern "C" int fafk_ct_cwt(const fafd_fft *Ph, const fafd_fft *Pn, int real_out, const void *post,
                           const float *psi, long psi_stride, int rows, int rows_per_wg,
                           const float *x, long in_dist, float *out, long out_dist, long count,
                           float scale, void *stream) {
    const int groups = (rows + rows_per_wg - 1) / rows_per_wg;
    dim3 g((unsigned)(count * groups));
    const size_t lds = (size_t)(Ph->plane + Pn->plane) * sizeof(float2);
    const float2 *po = (const float2 *)post;
#define CWT(N)                                                                           \
    do {                                                                                 \
        if (real_out) hipLaunchKernelGGL((kc_cwt<true, N>), g, dim3(N / 16), lds, S(stream), *Ph, *Pn, po, psi, psi_stride, rows, rows_per_wg, x, in_dist, out, out_dist, count, scale); \
        else hipLaunchKernelGGL((kc_cwt<false, N>), g, dim3(N / 16), lds, S(stream), *Ph, *Pn, po, psi, psi_stride, rows, rows_per_wg, x, in_dist, out, out_dist, count, scale); \
    } while (0)
    switch (Pn->n) {
    case 512: CWT(512); break;
    case 1024: CWT(1024); break;
    case 2048: CWT(2048); break;
    case 4096: CWT(4096); break;
    case 8192: CWT(8192); break;
    default:
        snprintf(g_ct_err, sizeof g_ct_err, "no specialised CWT kernel for length %d", Pn->n);
        return -1;
    }
#undef CWT
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc_cwt", e);
}

/* ---------------- four-step, power-of-two n1 x n2 ---------------- */

template <int N1, int N2, int RW>
static int ct4_row(int out_kind, int inv, const fafd_fft *P2, float2 *w, long ws_dist,
                   const void *post, float *out_re, float *out_im, long out_dist, long count,
                   float scale, hipStream_t st);

extern "C" int fafk_ct4_supported(int n1, int n2) {
    return (n1 == 128 && n2 == 256) || (n1 == 256 && n2 == 256);
}

template <int N1, int N2>
static int ct4_go(int in_kind, int out_kind, int inv, const fafd_fft *P1, const fafd_fft *P2,
                  const void *twn, const void *pre, const void *post, const float *in_re,
                  const float *in_im, long in_dist, void *ws, long ws_dist, float *out_re,
                  float *out_im, long out_dist, long count, float scale, hipStream_t st) {
    const dim3 gc(N2 / F4_COLS, (unsigned)count), bc(F4_COLS * (N1 / 16));
    const size_t lc = (size_t)F4_COLS * (N1 + N1 / 16 + 1) * sizeof(float2);
    const float2 *tw = (const float2 *)twn;
    float2 *w = (float2 *)ws;
#define COL(I, V) hipLaunchKernelGGL((kc4_col<I, V, N1, N2>), gc, bc, lc, st, *P1, tw, pre, in_re, in_im, in_dist, w, ws_dist)
    switch (in_kind) {
    case F4_IN_PAIR: COL(F4_IN_PAIR, false); break;
    case F4_IN_SPLIT: if (inv) COL(F4_IN_SPLIT, true); else COL(F4_IN_SPLIT, false); break;
    case F4_IN_INTER: if (inv) COL(F4_IN_INTER, true); else COL(F4_IN_INTER, false); break;
    case F4_PRE_SPLIT: COL(F4_PRE_SPLIT, true); break;
    case F4_PRE_INTER: COL(F4_PRE_INTER, true); break;
    default: snprintf(g_ct_err, sizeof g_ct_err, "four-step input kind %d", in_kind); return -1;
    }
#undef COL
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ct_fail("kc4_col", e);
    return ct4_row<N1, N2, F4_ROWS>(out_kind, inv, P2, w, ws_dist, post, out_re, out_im,
                                    out_dist, count, scale, st);
}

template <int N1, int N2, int RW>
static int ct4_row(int out_kind, int inv, const fafd_fft *P2, float2 *w, long ws_dist,
                   const void *post, float *out_re, float *out_im, long out_dist, long count,
                   float scale, hipStream_t st) {
    /* post-pass blocks hold 2 RW rows (RW rows of [0, N1/2) and their
     * mirrors, N1/2 in block 0); plain blocks 2 RW consecutive rows */
    const int blocks = N1 / (2 * RW);
    const dim3 gr((unsigned)blocks, (unsigned)count), br(2 * RW * (N2 / 16));
    const size_t lr = (size_t)2 * RW * (N2 + N2 / 16 + 1) * sizeof(float2);
    const float2 *po = (const float2 *)post;
#define ROW(O, V) hipLaunchKernelGGL((kc4_row<O, V, N1, N2, RW>), gr, br, lr, st, *P2, (const float2 *)w, ws_dist, po, out_re, out_im, out_dist, scale)
    switch (out_kind) {
    case F4_POST_SPLIT: ROW(F4_POST_SPLIT, false); break;
    case F4_POST_INTER: ROW(F4_POST_INTER, false); break;
    case F4_OUT_SPLIT: if (inv) ROW(F4_OUT_SPLIT, true); else ROW(F4_OUT_SPLIT, false); break;
    case F4_OUT_INTER: if (inv) ROW(F4_OUT_INTER, true); else ROW(F4_OUT_INTER, false); break;
    case F4_OUT_PAIR: ROW(F4_OUT_PAIR, true); break;
    default: snprintf(g_ct_err, sizeof g_ct_err, "four-step output kind %d", out_kind); return -1;
    }
#undef ROW
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ct_fail("kc4_row", e);
}

/* column pass then row pass of `count` signals; dists in floats for the
 * input and output planes (interleaved rows count two floats a bin) and in
 * float2 for the workspace */
extern "C" int fafk_ct4(int in_kind, int out_kind, int inv, const fafd_fft *P1,
                        const fafd_fft *P2, const void *twn, const void *pre, const void *post,
                        const float *in_re, const float *in_im, long in_dist, void *ws,
                        long ws_dist, float *out_re, float *out_im, long out_dist, long count,
                        float scale, void *stream) {
    if (P1->n == 128 && P2->n == 256)
        return ct4_go<128, 256>(in_kind, out_kind, inv, P1, P2, twn, pre, post, in_re, in_im,
                                in_dist, ws, ws_dist, out_re, out_im, out_dist, count, scale,
                                S(stream));
    if (P1->n == 256 && P2->n == 256)
        return ct4_go<256, 256>(in_kind, out_kind, inv, P1, P2, twn, pre, post, in_re, in_im,
                                in_dist, ws, ws_dist, out_re, out_im, out_dist, count, scale,
                                S(stream));
    snprintf(g_ct_err, sizeof g_ct_err, "no specialised four-step for %d x %d", P1->n, P2->n);
    return -1;
}
