/*
 * faf_kernels.hip -- the thin C-ABI layer between the C11 host code and the
 * gfx950 kernels of faf_engine.h.  Every function takes plain pointers,
 * sizes and plan records by value, launches on the given stream and
 * returns 0 or a negative code (message in fafk_last_error()).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include "faf_engine.h"
#include "faf_kern.h"

static thread_local char g_err[256];

extern "C" const char *fafk_last_error(void) { return g_err; }

static int fail(const char *what, hipError_t e) {
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

#define LAUNCH_CHECK(name)                                                \
    do {                                                                  \
        hipError_t e_ = hipGetLastError();                                \
        if (e_ != hipSuccess) return fail(name, e_);                      \
    } while (0)

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

/* allow every kernel the LDS left beside its static LDS (160 KB in all);
 * a refusal is not fatal (launches then stay within the default), and the
 * error it records is cleared so it cannot surface in a later call */
template <class K> static void lds_attr(K k) {
    hipFuncAttributes fa;
    size_t stat = 0;
    if (hipFuncGetAttributes(&fa, (const void *)k) == hipSuccess) stat = fa.sharedSizeBytes;
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(FAFD_MAX_LDS_BYTES - stat));
    (void)hipGetLastError();
}

extern "C" int fafk_init(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        snprintf(g_err, sizeof g_err, "no HIP device available (%s)",
                 e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return -1;
    }
    lds_attr(k_c2c<false, false>); lds_attr(k_c2c<true, false>);
    lds_attr(k_c2c<false, true>);  lds_attr(k_c2c<true, true>);
    lds_attr(k_r2c<false>); lds_attr(k_r2c<true>);
    lds_attr(k_c2r<false>); lds_attr(k_c2r<true>);
    lds_attr(k_pipe); lds_attr(k_stft);
    lds_attr(k_cwt<false>); lds_attr(k_cwt<true>);
    lds_attr(k_fs_col<false, FS_IN_SPLIT>); lds_attr(k_fs_col<true, FS_IN_SPLIT>);
    lds_attr(k_fs_col<false, FS_IN_PAIR>);  lds_attr(k_fs_col<true, FS_IN_PAIR>);
    lds_attr(k_fs_col<false, FS_IN_INTER>); lds_attr(k_fs_col<true, FS_IN_INTER>);
    lds_attr(k_fs_row<false, FS_OUT_SPLIT>); lds_attr(k_fs_row<true, FS_OUT_SPLIT>);
    lds_attr(k_fs_row<false, FS_OUT_PAIR>);  lds_attr(k_fs_row<true, FS_OUT_PAIR>);
    lds_attr(k_fs_row<false, FS_OUT_INTER>); lds_attr(k_fs_row<true, FS_OUT_INTER>);
    lds_attr(k_rader); lds_attr(k_blue); lds_attr(k_ragged); lds_attr(k_icwt);
    lds_attr(k_serve);
    return 0;
}

/* ---------------- memory and streams ---------------- */

extern "C" void *fafk_malloc(size_t bytes) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) { fail("hipMalloc", e); return nullptr; }
    return p;
}

extern "C" void *fafk_host_alloc(size_t bytes) {
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) { fail("hipHostMalloc", e); return nullptr; }
    return p;
}

extern "C" void fafk_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

extern "C" void fafk_free(void *p) {
    if (p) (void)hipFree(p);
}

extern "C" int fafk_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream));
    return e == hipSuccess ? 0 : fail("hipMemcpyAsync H2D", e);
}

extern "C" int fafk_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream));
    return e == hipSuccess ? 0 : fail("hipMemcpyAsync D2H", e);
}

extern "C" int fafk_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream));
    return e == hipSuccess ? 0 : fail("hipMemcpyAsync D2D", e);
}

extern "C" int fafk_sync(void *stream) {
    hipError_t e = hipStreamSynchronize(S(stream));
    return e == hipSuccess ? 0 : fail("hipStreamSynchronize", e);
}

/* 1 if p is device (or managed) memory, 0 if plain host memory */
extern "C" int fafk_is_device_ptr(const void *p) {
    hipPointerAttribute_t a;
    memset(&a, 0, sizeof a);
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) { (void)hipGetLastError(); return 0; }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

extern "C" void *fafk_stream_create(void) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    return (void *)s;
}

extern "C" void fafk_stream_destroy(void *s) {
    if (s) (void)hipStreamDestroy(S(s));
}

extern "C" void *fafk_event_create(void) {
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return (void *)ev;
}

extern "C" void fafk_event_destroy(void *ev) {
    if (ev) (void)hipEventDestroy((hipEvent_t)ev);
}

extern "C" int fafk_event_record(void *ev, void *stream) {
    hipError_t e = hipEventRecord((hipEvent_t)ev, S(stream));
    return e == hipSuccess ? 0 : fail("hipEventRecord", e);
}

/* 1 done, 0 pending, <0 error */
extern "C" int fafk_event_query(void *ev) {
    hipError_t e = hipEventQuery((hipEvent_t)ev);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail("hipEventQuery", e);
}

extern "C" int fafk_event_sync(void *ev) {
    hipError_t e = hipEventSynchronize((hipEvent_t)ev);
    return e == hipSuccess ? 0 : fail("hipEventSynchronize", e);
}

extern "C" float fafk_event_ms(void *a, void *b) {
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return -1.0f;
    return ms;
}

extern "C" int fafk_device_info(int *cus, size_t *lds_per_block, size_t *total_mem) {
    hipDeviceProp_t p;
    hipError_t e = hipGetDeviceProperties(&p, 0);
    if (e != hipSuccess) return fail("hipGetDeviceProperties", e);
    if (cus) *cus = p.multiProcessorCount;
    if (lds_per_block) *lds_per_block = p.sharedMemPerBlock;
    if (total_mem) *total_mem = p.totalGlobalMem;
    return 0;
}

extern "C" int fafk_fill_u32(unsigned *p, int n, unsigned v, void *stream) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, S(stream), p, n, v);
    LAUNCH_CHECK("k_fill_u32");
    return 0;
}

extern "C" int fafk_window_inter(const float *x, long in_dist, const float *w, long n, float *y,
                                 long count, void *stream) {
    long tot = n * count;
    hipLaunchKernelGGL(k_window_inter, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       S(stream), (const float2 *)x, in_dist, w, n, (float2 *)y, count);
    LAUNCH_CHECK("k_window_inter");
    return 0;
}

/* ---------------- transform launches ---------------- */

static inline unsigned blocks_for(long count, int spb) {
    return (unsigned)((count + spb - 1) / spb);
}

extern "C" int fafk_c2c(const fafd_fft *P, int inv, int inter, int spb, const float *in_re,
                        const float *in_im, long in_dist, float *out_re, float *out_im,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    if (inter) {
        if (inv) hipLaunchKernelGGL((k_c2c<true, true>), g, b, lds, S(stream), *P, spb, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale);
        else hipLaunchKernelGGL((k_c2c<false, true>), g, b, lds, S(stream), *P, spb, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale);
    } else {
        if (inv) hipLaunchKernelGGL((k_c2c<true, false>), g, b, lds, S(stream), *P, spb, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale);
        else hipLaunchKernelGGL((k_c2c<false, false>), g, b, lds, S(stream), *P, spb, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale);
    }
    LAUNCH_CHECK("k_c2c");
    return 0;
}

extern "C" int fafk_r2c(const fafd_fft *P, int inter_out, int spb, const void *post,
                        const float *x, long in_dist, float *out_re, float *out_im,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    if (inter_out)
        hipLaunchKernelGGL((k_r2c<true>), g, b, lds, S(stream), *P, spb, (const float2 *)post, x, in_dist, out_re, out_im, out_dist, count, scale);
    else
        hipLaunchKernelGGL((k_r2c<false>), g, b, lds, S(stream), *P, spb, (const float2 *)post, x, in_dist, out_re, out_im, out_dist, count, scale);
    LAUNCH_CHECK("k_r2c");
    return 0;
}

extern "C" int fafk_c2r(const fafd_fft *P, int inter_in, int spb, const void *pre,
                        const float *in_re, const float *in_im, long in_dist, float *y,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    if (inter_in)
        hipLaunchKernelGGL((k_c2r<true>), g, b, lds, S(stream), *P, spb, (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, count, scale);
    else
        hipLaunchKernelGGL((k_c2r<false>), g, b, lds, S(stream), *P, spb, (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, count, scale);
    LAUNCH_CHECK("k_c2r");
    return 0;
}

extern "C" int fafk_pipe(const fafd_fft *Pf, const fafd_fft *Pi, int spb, const void *post,
                         const void *pre, const fafd_ops *ops, const float *x, long in_dist,
                         float *y, long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(Pf->threads * spb);
    size_t lds = (size_t)spb * Pf->plane * sizeof(float2);
    hipLaunchKernelGGL(k_pipe, g, b, lds, S(stream), *Pf, *Pi, spb, (const float2 *)post,
                       (const float2 *)pre, *ops, x, in_dist, y, out_dist, count, scale);
    LAUNCH_CHECK("k_pipe");
    return 0;
}

extern "C" int fafk_stft(const fafd_fft *P, int spb, const void *post, const float *win,
                         const float *x, long rec_len, long rec_dist, long hop, long frames,
                         float *out_re, float *out_im, long frame_dist, long rec_out_dist,
                         long count, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    hipLaunchKernelGGL(k_stft, g, b, lds, S(stream), *P, spb, (const float2 *)post, win, x,
                       rec_len, rec_dist, hop, frames, out_re, out_im, frame_dist,
                       rec_out_dist, count);
    LAUNCH_CHECK("k_stft");
    return 0;
}

extern "C" int fafk_cwt(const fafd_fft *Ph, const fafd_fft *Pn, int real_out, const void *post,
                        const float *psi, long psi_stride, int rows, int rows_per_wg,
                        const float *x, long in_dist, float *out, long out_dist, long count,
                        float scale, void *stream) {
    const int groups = (rows + rows_per_wg - 1) / rows_per_wg;
    dim3 g((unsigned)(count * groups));
    dim3 b(Ph->threads > Pn->threads ? Ph->threads : Pn->threads);
    size_t lds = (size_t)(Ph->plane + Pn->plane) * sizeof(float2);
    if (real_out)
        hipLaunchKernelGGL((k_cwt<true>), g, b, lds, S(stream), *Ph, *Pn, (const float2 *)post, psi, psi_stride, rows, rows_per_wg, x, in_dist, out, out_dist, count, scale);
    else
        hipLaunchKernelGGL((k_cwt<false>), g, b, lds, S(stream), *Ph, *Pn, (const float2 *)post, psi, psi_stride, rows, rows_per_wg, x, in_dist, out, out_dist, count, scale);
    LAUNCH_CHECK("k_cwt");
    return 0;
}

extern "C" int fafk_icwt(const fafd_fft *Pf, const fafd_fft *Pi, const void *post,
                         const void *pre, const float *dual, long dual_stride, int rows,
                         const float *w, long in_dist, float *y, long out_dist, long count,
                         float scale, void *stream) {
    dim3 g((unsigned)count), b(Pf->threads);
    size_t lds = (size_t)2 * Pf->plane * sizeof(float2);
    hipLaunchKernelGGL(k_icwt, g, b, lds, S(stream), *Pf, *Pi, (const float2 *)post,
                       (const float2 *)pre, dual, dual_stride, rows, w, in_dist, y, out_dist,
                       count, scale);
    LAUNCH_CHECK("k_icwt");
    return 0;
}

extern "C" int fafk_fs_col(const fafd_fft *P1, int inv, int in_kind, int n2, int cols,
                           const void *twn, const float *in_re, const float *in_im,
                           long in_dist, float *t_re, float *t_im, long t_dist, long count,
                           void *stream) {
    dim3 g((unsigned)((n2 + cols - 1) / cols), (unsigned)count), b(P1->threads * cols);
    size_t lds = (size_t)cols * P1->plane * sizeof(float2);
    const float2 *tw = (const float2 *)twn;
#define FSC(I, K) hipLaunchKernelGGL((k_fs_col<I, K>), g, b, lds, S(stream), *P1, n2, cols, tw, in_re, in_im, in_dist, t_re, t_im, t_dist)
    if (inv) {
        if (in_kind == FS_IN_SPLIT) FSC(true, FS_IN_SPLIT);
        else if (in_kind == FS_IN_PAIR) FSC(true, FS_IN_PAIR);
        else FSC(true, FS_IN_INTER);
    } else {
        if (in_kind == FS_IN_SPLIT) FSC(false, FS_IN_SPLIT);
        else if (in_kind == FS_IN_PAIR) FSC(false, FS_IN_PAIR);
        else FSC(false, FS_IN_INTER);
    }
#undef FSC
    LAUNCH_CHECK("k_fs_col");
    return 0;
}

extern "C" int fafk_fs_row(const fafd_fft *P2, int inv, int out_kind, int n1, int rows,
                           const float *t_re, const float *t_im, long t_dist, float *out_re,
                           float *out_im, long out_dist, long count, float scale,
                           void *stream) {
    dim3 g((unsigned)((n1 + rows - 1) / rows), (unsigned)count), b(P2->threads * rows);
    size_t lds = (size_t)rows * P2->plane * sizeof(float2);
#define FSR(I, K) hipLaunchKernelGGL((k_fs_row<I, K>), g, b, lds, S(stream), *P2, n1, rows, t_re, t_im, t_dist, out_re, out_im, out_dist, scale)
    if (inv) {
        if (out_kind == FS_OUT_SPLIT) FSR(true, FS_OUT_SPLIT);
        else if (out_kind == FS_OUT_PAIR) FSR(true, FS_OUT_PAIR);
        else FSR(true, FS_OUT_INTER);
    } else {
        if (out_kind == FS_OUT_SPLIT) FSR(false, FS_OUT_SPLIT);
        else if (out_kind == FS_OUT_PAIR) FSR(false, FS_OUT_PAIR);
        else FSR(false, FS_OUT_INTER);
    }
#undef FSR
    LAUNCH_CHECK("k_fs_row");
    return 0;
}

extern "C" int fafk_r2c_post(int nc, const void *post, int inter_out, const float *z_re,
                             const float *z_im, long z_dist, float *out_re, float *out_im,
                             long out_dist, long count, float scale, void *stream) {
    dim3 g((unsigned)((nc / 2 + 256) / 256), (unsigned)count), b(256);
    if (inter_out)
        hipLaunchKernelGGL((k_r2c_post<true>), g, b, 0, S(stream), nc, (const float2 *)post, z_re, z_im, z_dist, out_re, out_im, out_dist, scale);
    else
        hipLaunchKernelGGL((k_r2c_post<false>), g, b, 0, S(stream), nc, (const float2 *)post, z_re, z_im, z_dist, out_re, out_im, out_dist, scale);
    LAUNCH_CHECK("k_r2c_post");
    return 0;
}

extern "C" int fafk_c2r_pre(int nc, const void *pre, int inter_in, const float *in_re,
                            const float *in_im, long in_dist, float *f_re, float *f_im,
                            long f_dist, long count, void *stream) {
    dim3 g((unsigned)((nc / 2 + 256) / 256), (unsigned)count), b(256);
    if (inter_in)
        hipLaunchKernelGGL((k_c2r_pre<true>), g, b, 0, S(stream), nc, (const float2 *)pre, in_re, in_im, in_dist, f_re, f_im, f_dist);
    else
        hipLaunchKernelGGL((k_c2r_pre<false>), g, b, 0, S(stream), nc, (const float2 *)pre, in_re, in_im, in_dist, f_re, f_im, f_dist);
    LAUNCH_CHECK("k_c2r_pre");
    return 0;
}

extern "C" int fafk_rader(const fafd_fft *Pf, const fafd_fft *Pi, const fafd_rader *R, int spb,
                          const float *in_re, const float *in_im, long in_dist, float *out_re,
                          float *out_im, long out_dist, long count, float scale,
                          void *stream) {
    dim3 g(blocks_for(count, spb)), b(Pf->threads * spb);
    const int per = Pf->plane + R->p + (R->p >> 4) + 1;
    size_t lds = (size_t)spb * per * sizeof(float2);
    hipLaunchKernelGGL(k_rader, g, b, lds, S(stream), *Pf, *Pi, *R, spb, in_re, in_im,
                       in_dist, out_re, out_im, out_dist, count, scale);
    LAUNCH_CHECK("k_rader");
    return 0;
}

extern "C" int fafk_blue(const fafd_fft *Pf, const fafd_fft *Pi, const fafd_blue *Bq, int spb,
                         const float *in_re, const float *in_im, long in_dist, float *out_re,
                         float *out_im, long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(Pf->threads * spb);
    size_t lds = (size_t)spb * Pf->plane * sizeof(float2);
    hipLaunchKernelGGL(k_blue, g, b, lds, S(stream), *Pf, *Pi, *Bq, spb, in_re, in_im,
                       in_dist, out_re, out_im, out_dist, count, scale);
    LAUNCH_CHECK("k_blue");
    return 0;
}

extern "C" int fafk_ragged(const void *plans, const void *units, int nunits,
                           const int32_t *order, const long *offsets, const float *in_re,
                           const float *in_im, float *out_re, float *out_im,
                           unsigned *counter, int grid, int block, size_t lds,
                           void *stream) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, S(stream), counter, 1, 0u);
    LAUNCH_CHECK("k_fill_u32");
    hipLaunchKernelGGL(k_ragged, dim3(grid), dim3(block), lds, S(stream),
                       (const fafd_rplan *)plans, (const fafd_unit *)units, nunits, order,
                       offsets, in_re, in_im, out_re, out_im, counter);
    LAUNCH_CHECK("k_ragged");
    return 0;
}

extern "C" int fafk_serve(const void *plans, const void *units, int nunits, const void *sigs,
                          unsigned *counter, int grid, int block, size_t lds, void *stream) {
    if (nunits <= 0) return 0;
    if (counter) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, S(stream), counter, 1, 0u);
        LAUNCH_CHECK("k_fill_u32");
    } else {
        grid = nunits;
    }
    hipLaunchKernelGGL(k_serve, dim3(grid), dim3(block), lds, S(stream),
                       (const fafd_splan *)plans, (const fafd_sunit *)units, nunits,
                       (const fafd_ssig *)sigs, counter);
    LAUNCH_CHECK("k_serve");
    return 0;
}

extern "C" int fafk_cwt_spread(const float *X, long x_dist, const float *psi, long psi_stride,
                               int nc, int analytic, float *out, long row_len, long seg_dist,
                               int rows, long count, void *stream) {
    if (count <= 0 || rows <= 0) return 0;
    if (count > 65535 || rows > 65535) return fail("k_cwt_spread grid", hipErrorInvalidValue);
    dim3 grid((unsigned)((row_len + 255) / 256), (unsigned)rows, (unsigned)count);
    hipLaunchKernelGGL(k_cwt_spread, grid, dim3(256), 0, S(stream), (const float2 *)X, x_dist,
                       psi, psi_stride, nc, analytic, (float2 *)out, row_len, seg_dist, rows);
    LAUNCH_CHECK("k_cwt_spread");
    return 0;
}
