/*
 * faf_kernels_rs.hip -- C-ABI launchers of the single-pass 32768-point
 * kernels (faf_engine_rs.h): rfft / irfft N = 65536 and C2C 32768, rfft /
 * irfft N = 16384, and irfft N = 4096 with the single-MUL pipeline at 4096.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include "faf_engine_rs.h"
#include "faf_kern.h"

static thread_local char g_rs_err[160];

static int rs_fail(const char *what, hipError_t e) {
    snprintf(g_rs_err, sizeof g_rs_err, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

extern "C" const char *fafk_rs_last_error(void) { return g_rs_err; }

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

template <class K> static void rs_attr(K k) {
    (void)hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                              rs::LDS_BYTES);
    (void)hipGetLastError();
}

extern "C" int fafk_rs_init(void) {
    rs_attr(kr_r2c<false>); rs_attr(kr_r2c<true>);
    rs_attr(kr_c2r<false>); rs_attr(kr_c2r<true>);
    rs_attr(kr_c2c<false, false, false>); rs_attr(kr_c2c<true, false, false>);
    rs_attr(kr_c2c<false, true, true>); rs_attr(kr_c2c<true, true, true>);
    return 0;
}

/* rfft / irfft N = 16384 in one pass (rs8): one 256-thread workgroup per
 * signal, grid = count */
extern "C" int fafk_rs8_r2c(int inter_out, const void *tw, const void *post, const float *x,
                            long in_dist, float *out_re, float *out_im, long out_dist,
                            long count, float scale, void *stream) {
    if (count <= 0) return 0;
    if (count > 0x7fffffffL) return rs_fail("kr8_r2c", hipErrorInvalidValue);
    const dim3 g((unsigned)count), b(rs8::T);
    if (inter_out)
        hipLaunchKernelGGL(kr8_r2c<true>, g, b, rs8::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)post, x, in_dist, out_re, out_im, out_dist, scale);
    else
        hipLaunchKernelGGL(kr8_r2c<false>, g, b, rs8::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)post, x, in_dist, out_re, out_im, out_dist, scale);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rs_fail("kr8_r2c", e);
}

extern "C" int fafk_rs8_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                            const float *in_im, long in_dist, float *y, long out_dist,
                            long count, float scale, void *stream) {
    if (count <= 0) return 0;
    if (count > 0x7fffffffL) return rs_fail("kr8_c2r", hipErrorInvalidValue);
    const dim3 g((unsigned)count), b(rs8::T);
    if (inter_in)
        hipLaunchKernelGGL(kr8_c2r<true>, g, b, rs8::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, scale);
    else
        hipLaunchKernelGGL(kr8_c2r<false>, g, b, rs8::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, scale);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rs_fail("kr8_c2r", e);
}

/* irfft N = 4096 in one pass (rs2): one wave per signal */
extern "C" int fafk_rs2_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                            const float *in_im, long in_dist, float *y, long out_dist,
                            long count, float scale, void *stream) {
    if (count <= 0) return 0;
    if (count > 0x7fffffffL) return rs_fail("kr2_c2r", hipErrorInvalidValue);
    const dim3 g((unsigned)count), b(rs2::T);
    if (inter_in)
        hipLaunchKernelGGL(kr2_c2r<true>, g, b, rs2::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, scale);
    else
        hipLaunchKernelGGL(kr2_c2r<false>, g, b, rs2::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, scale);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rs_fail("kr2_c2r", e);
}

/* the single-MUL pipeline N = 4096 in one pass (rs2): one wave per signal */
extern "C" int fafk_rs2_pipe(const void *twf, const void *twi, const void *post, const void *pre,
                             const fafd_ops *ops, const float *x, long in_dist, float *y,
                             long out_dist, long count, float scale, void *stream) {
    if (count <= 0) return 0;
    if (count > 0x7fffffffL) return rs_fail("kr2_pipe", hipErrorInvalidValue);
    hipLaunchKernelGGL(kr2_pipe, dim3((unsigned)count), dim3(rs2::T), rs2::LDS_BYTES, S(stream),
                       (const float2 *)twf, (const float2 *)twi, (const float2 *)post,
                       (const float2 *)pre, *ops, x, in_dist, y, out_dist, scale);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rs_fail("kr2_pipe", e);
}

/* C2C 8192 in one pass (rs8): one 256-thread workgroup per signal */
extern "C" int fafk_rs8_c2c(int inv, int inter, const void *tw, const float *in_re,
                            const float *in_im, long in_dist, float *out_re, float *out_im,
                            long out_dist, long count, float scale, void *stream) {
    if (count <= 0) return 0;
    if (count > 0x7fffffffL) return rs_fail("kr8_c2c", hipErrorInvalidValue);
    const dim3 g((unsigned)count), b(rs8::T);
    const float2 *t = (const float2 *)tw;
#define RS8_C2C(I, L)                                                                       \
    hipLaunchKernelGGL((kr8_c2c<I, L, L>), g, b, rs8::LDS_BYTES, S(stream), t, in_re, in_im, \
                       in_dist, out_re, out_im, out_dist, scale)
    if (inv) {
        if (inter) RS8_C2C(true, true); else RS8_C2C(true, false);
    } else {
        if (inter) RS8_C2C(false, true); else RS8_C2C(false, false);
    }
#undef RS8_C2C
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rs_fail("kr8_c2c", e);
}

/* float2 row access needs 8-byte aligned rows */
extern "C" int fafk_rs_pair_ok(const void *p, long dist) {
    return ((uintptr_t)p & 7u) == 0 && (dist & 1) == 0;
}

static unsigned rs_grid(long count, int grid) {
    long g = grid > 0 ? grid : 256;
    return (unsigned)(count < g ? count : g);
}

#define RS_CHECK(name)                                                    \
    do {                                                                  \
        hipError_t e_ = hipGetLastError();                                \
        if (e_ != hipSuccess) return rs_fail(name, e_);                   \
    } while (0)

extern "C" int fafk_rs_r2c(int inter_out, const void *tw, const void *post, const float *x,
                           long in_dist, float *out_re, float *out_im, long out_dist,
                           long count, float scale, int grid, void *stream) {
    if (count <= 0) return 0;
    const dim3 g(rs_grid(count, grid)), b(rs::T);
    if (inter_out)
        hipLaunchKernelGGL(kr_r2c<true>, g, b, rs::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)post, x, in_dist, out_re, out_im, out_dist, count,
                           scale);
    else
        hipLaunchKernelGGL(kr_r2c<false>, g, b, rs::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)post, x, in_dist, out_re, out_im, out_dist, count,
                           scale);
    RS_CHECK("kr_r2c");
    return 0;
}

extern "C" int fafk_rs_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                           const float *in_im, long in_dist, float *y, long out_dist,
                           long count, float scale, int grid, void *stream) {
    if (count <= 0) return 0;
    const dim3 g(rs_grid(count, grid)), b(rs::T);
    if (inter_in)
        hipLaunchKernelGGL(kr_c2r<true>, g, b, rs::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, count,
                           scale);
    else
        hipLaunchKernelGGL(kr_c2r<false>, g, b, rs::LDS_BYTES, S(stream), (const float2 *)tw,
                           (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, count,
                           scale);
    RS_CHECK("kr_c2r");
    return 0;
}

extern "C" int fafk_rs_c2c(int inv, int inter, const void *tw, const float *in_re,
                           const float *in_im, long in_dist, float *out_re, float *out_im,
                           long out_dist, long count, float scale, int grid, void *stream) {
    if (count <= 0) return 0;
    const dim3 g(rs_grid(count, grid)), b(rs::T);
    const float2 *t = (const float2 *)tw;
#define RS_C2C(I, L)                                                                   \
    hipLaunchKernelGGL((kr_c2c<I, L, L>), g, b, rs::LDS_BYTES, S(stream), t, in_re, in_im, \
                       in_dist, out_re, out_im, out_dist, count, scale)
    if (inv) {
        if (inter) RS_C2C(true, true); else RS_C2C(true, false);
    } else {
        if (inter) RS_C2C(false, true); else RS_C2C(false, false);
    }
#undef RS_C2C
    RS_CHECK("kr_c2c");
    return 0;
}
