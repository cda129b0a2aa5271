/*
 * faf_kernels.hip -- the thin C-ABI layer between the C11 host code and the
 * gfx950 kernels of faf_engine.h.  Every function takes plain pointers,
 * sizes and host copies of the plan records, launches on the given stream
 * and returns 0 or a negative code (message in fafk_last_error()).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <string>
#include <unordered_map>
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

template <int M> static void init_attrs() {
    lds_attr(k_c2c<false, false, M>); lds_attr(k_c2c<true, false, M>);
    lds_attr(k_c2c<false, true, M>);  lds_attr(k_c2c<true, true, M>);
    lds_attr(k_r2c<false, M>); lds_attr(k_r2c<true, M>);
    lds_attr(k_c2r<false, M>); lds_attr(k_c2r<true, M>);
    lds_attr(k_pipe<M>); lds_attr(k_stft<M>);
    lds_attr(k_cwt<false, M>); lds_attr(k_cwt<true, M>);
    lds_attr(k_rader<M>); lds_attr(k_blue<M>); lds_attr(k_icwt<M>);
    lds_attr(k_serve<M>);
}

/* ---------------- static scratch of the allocation-free entry point ----
 * faf_execute_batch_noalloc lends this array to a plan as its workspace
 * (four-step intermediates and the like) instead of allocating one.  Every
 * GPU that loads the library holds its own copy, at its own address, until
 * the process ends.  Host code never passes the array's name to a kernel
 * (on the host it names a stand-in, not device memory): its device address
 * is looked up once per device with hipGetSymbolAddress and kept in
 * g_scratch_ptr[device]. */
static __device__ __align__(256) unsigned char scratch[FAFK_SCRATCH_BYTES];

#define FAFK_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[FAFK_MAX_DEVICES];

/* resolves the current device's scratch address (once per device) */
static int scratch_resolve(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail("hipGetDevice", e);
    if (dev < 0 || dev >= FAFK_MAX_DEVICES) {
        snprintf(g_err, sizeof g_err, "device %d beyond the scratch table", dev);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return fail("hipGetSymbolAddress (scratch)", e);
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" void *fafk_scratch(size_t *bytes) {
    void *p = nullptr;
    if (scratch_resolve(&p)) return nullptr;
    if (bytes) *bytes = FAFK_SCRATCH_BYTES;
    return p;
}

/* while set on a thread, device-memory allocations of this layer fail
 * instead of happening (the plan-record cache included) */
static thread_local int t_noalloc;
extern "C" void fafk_set_noalloc(int on) { t_noalloc = on; }

extern "C" int fafk_init(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        snprintf(g_err, sizeof g_err, "no HIP device available (%s)",
                 e == hipSuccess ? "count 0" : hipGetErrorString(e));
        return -1;
    }
    {   /* the scratch address of the device current at set-up */
        void *p = nullptr;
        if (scratch_resolve(&p)) return -1;
    }
    init_attrs<512>();
    init_attrs<1024>();
    constexpr int M = 1024;
    lds_attr(k_fs_col<false, FS_IN_SPLIT, M>); lds_attr(k_fs_col<true, FS_IN_SPLIT, M>);
    lds_attr(k_fs_col<false, FS_IN_PAIR, M>);  lds_attr(k_fs_col<true, FS_IN_PAIR, M>);
    lds_attr(k_fs_col<false, FS_IN_INTER, M>); lds_attr(k_fs_col<true, FS_IN_INTER, M>);
    lds_attr(k_fs_row<false, FS_OUT_SPLIT, M>); lds_attr(k_fs_row<true, FS_OUT_SPLIT, M>);
    lds_attr(k_fs_row<false, FS_OUT_PAIR, M>);  lds_attr(k_fs_row<true, FS_OUT_PAIR, M>);
    lds_attr(k_fs_row<false, FS_OUT_INTER, M>); lds_attr(k_fs_row<true, FS_OUT_INTER, M>);
    return 0;
}

/* ---------------- memory and streams ---------------- */

extern "C" void *fafk_malloc(size_t bytes) {
    void *p = nullptr;
    if (t_noalloc) {
        snprintf(g_err, sizeof g_err, "allocation of %zu bytes refused: the call may not "
                 "allocate device memory", bytes);
        return nullptr;
    }
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

/* ---- stream capture into a replayable graph ----
 * Relaxed mode: the host code itself refuses every synchronising or
 * allocating call while a capture is open, so the runtime need not. */
extern "C" int fafk_capture_begin(void *stream) {
    hipError_t e = hipStreamBeginCapture(S(stream), hipStreamCaptureModeRelaxed);
    return e == hipSuccess ? 0 : fail("hipStreamBeginCapture", e);
}

/* Ends the capture and instantiates it; NULL (error set) if the capture was
 * invalidated or holds nothing.  *nodes receives the graph's node count. */
extern "C" void *fafk_capture_end(void *stream, int *nodes) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(S(stream), &g);
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        fail("hipStreamEndCapture", e != hipSuccess ? e : hipErrorStreamCaptureInvalidated);
        return nullptr;
    }
    size_t n = 0;
    e = hipGraphGetNodes(g, nullptr, &n);
    if (e == hipSuccess && n == 0) e = hipErrorInvalidValue;
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        fail(n ? "hipGraphInstantiate" : "empty capture", e);
        return nullptr;
    }
    if (nodes) *nodes = (int)n;
    return (void *)x;
}

extern "C" int fafk_graph_launch(void *x, void *stream) {
    hipError_t e = hipGraphLaunch((hipGraphExec_t)x, S(stream));
    return e == hipSuccess ? 0 : fail("hipGraphLaunch", e);
}

extern "C" void fafk_graph_destroy(void *x) {
    if (x) (void)hipGraphExecDestroy((hipGraphExec_t)x);
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

/* Copy ceiling of the device: each thread moves four 16-byte words per
 * step (loads issued before stores), grid-stride over the buffer.  Used to
 * state what a pure read-once/write-once stream reaches on this board. */
/* NT: non-temporal (streaming) stores, as the transforms use for rows
 * written once */
template <bool NT>
__global__ __launch_bounds__(256) void k_copy16(const float4 *__restrict__ src,
                                                float4 *__restrict__ dst, long n16) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const long stride = (long)gridDim.x * 256 * 4;
    for (long i = (long)blockIdx.x * 1024 + threadIdx.x; i < n16; i += stride) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * 256 < n16) v[u] = src[i + u * 256];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * 256 < n16) {
                if (NT) __builtin_nontemporal_store(*(const f4 *)&v[u], (f4 *)&dst[i + u * 256]);
                else dst[i + u * 256] = v[u];
            }
    }
}

extern "C" int fafk_copy16(void *dst, const void *src, size_t bytes, int grid, int streaming,
                           void *stream) {
    if ((bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15) || grid < 1) {
        snprintf(g_err, sizeof g_err, "copy: 16-byte aligned buffers and sizes only");
        return -1;
    }
    if (streaming)
        hipLaunchKernelGGL(k_copy16<true>, dim3(grid), dim3(256), 0, S(stream),
                           (const float4 *)src, (float4 *)dst, (long)(bytes / 16));
    else
        hipLaunchKernelGGL(k_copy16<false>, dim3(grid), dim3(256), 0, S(stream),
                           (const float4 *)src, (float4 *)dst, (long)(bytes / 16));
    LAUNCH_CHECK("k_copy16");
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

/* Plan records live in device memory: a kernel indexes P.st[s] with a
 * runtime stage number, which on a by-value kernel argument makes the
 * compiler materialise the whole record in registers and scratch.  Records
 * are immutable, so each distinct one is uploaded once (keyed by its bytes)
 * into 64 KB slabs that live until the process ends. */
static std::mutex g_rec_mu;
static std::unordered_map<std::string, const void *> g_recs;
static char *g_slab;
static size_t g_slab_left;

static const void *dev_rec(const void *h, size_t bytes) {
    std::string key((const char *)h, bytes);
    std::lock_guard<std::mutex> lk(g_rec_mu);
    auto it = g_recs.find(key);
    if (it != g_recs.end()) return it->second;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (t_noalloc) {
        /* a record not yet uploaded would need a copy (and maybe a slab) */
        snprintf(g_err, sizeof g_err, "device plan record not resident: run the plan once "
                 "through faf_execute_batch before the allocation-free call");
        return nullptr;
    }
    if (g_slab_left < need) {
        const size_t sz = need > 65536 ? need : 65536;
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail("hipMalloc (plan records)", e);
            return nullptr;
        }
        g_slab = (char *)p;
        g_slab_left = sz;
    }
    char *d = g_slab;
    hipError_t e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail("hipMemcpy (plan records)", e);
        return nullptr;
    }
    g_slab += need;
    g_slab_left -= need;
    g_recs.emplace(std::move(key), d);
    return d;
}

static const fafd_fft *dev_fft(const fafd_fft *P) {
    return (const fafd_fft *)dev_rec(P, sizeof *P);
}

/* the device copy of a spectral-op list, built from its fields so padding
 * bytes never split the cache */
static const fafd_ops *dev_ops(const fafd_ops *o) {
    fafd_ops c;
    memset(&c, 0, sizeof c);
    c.nops = o->nops;
    for (int i = 0; i < FAFD_MAX_OPS; ++i) {
        c.code[i] = o->code[i];
        c.a1[i] = o->a1[i];
        c.a2[i] = o->a2[i];
    }
    c.h_re = o->h_re;
    c.h_im = o->h_im;
    c.h_bins = o->h_bins;
    return (const fafd_ops *)dev_rec(&c, sizeof c);
}

/* upload a plan record ahead of its first launch (plan creation time) */
extern "C" const void *fafk_plan_record(const fafd_fft *P) { return dev_fft(P); }

#define NEED(p) \
    do { if (!(p)) { snprintf(g_err, sizeof g_err, "device plan record upload failed"); return -1; } } while (0)

/* kernels compiled for at most 512 threads may use up to 256 VGPRs and so
 * never spill; 1024-thread launches (lengths whose plan needs more than 512
 * threads) take the 1024 variant */
#define BY_MT(bx, ...)                                                    \
    do {                                                                  \
        if ((bx) <= 512) { constexpr int MT = 512; __VA_ARGS__; }         \
        else { constexpr int MT = 1024; __VA_ARGS__; }                    \
    } while (0)

static inline unsigned blocks_for(long count, int spb) {
    return (unsigned)((count + spb - 1) / spb);
}

extern "C" int fafk_c2c(const fafd_fft *P, int inv, int inter, int spb, const float *in_re,
                        const float *in_im, long in_dist, float *out_re, float *out_im,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    const fafd_fft *Pd = dev_fft(P);
    NEED(Pd);
#define C2C(I, X) hipLaunchKernelGGL((k_c2c<I, X, MT>), g, b, lds, S(stream), Pd, spb, in_re, in_im, in_dist, out_re, out_im, out_dist, count, scale)
    BY_MT(b.x, if (inter) { if (inv) C2C(true, true); else C2C(false, true); }
               else { if (inv) C2C(true, false); else C2C(false, false); });
#undef C2C
    LAUNCH_CHECK("k_c2c");
    return 0;
}

extern "C" int fafk_r2c(const fafd_fft *P, int inter_out, int spb, const void *post,
                        const float *x, long in_dist, float *out_re, float *out_im,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    const fafd_fft *Pd = dev_fft(P);
    NEED(Pd);
#define R2C(X) hipLaunchKernelGGL((k_r2c<X, MT>), g, b, lds, S(stream), Pd, spb, (const float2 *)post, x, in_dist, out_re, out_im, out_dist, count, scale)
    BY_MT(b.x, if (inter_out) R2C(true); else R2C(false));
#undef R2C
    LAUNCH_CHECK("k_r2c");
    return 0;
}

extern "C" int fafk_c2r(const fafd_fft *P, int inter_in, int spb, const void *pre,
                        const float *in_re, const float *in_im, long in_dist, float *y,
                        long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    const fafd_fft *Pd = dev_fft(P);
    NEED(Pd);
#define C2R(X) hipLaunchKernelGGL((k_c2r<X, MT>), g, b, lds, S(stream), Pd, spb, (const float2 *)pre, in_re, in_im, in_dist, y, out_dist, count, scale)
    BY_MT(b.x, if (inter_in) C2R(true); else C2R(false));
#undef C2R
    LAUNCH_CHECK("k_c2r");
    return 0;
}

extern "C" int fafk_pipe(const fafd_fft *Pf, const fafd_fft *Pi, int spb, const void *post,
                         const void *pre, const fafd_ops *ops, const float *x, long in_dist,
                         float *y, long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(Pf->threads * spb);
    size_t lds = (size_t)spb * Pf->plane * sizeof(float2);
    const fafd_fft *Pfd = dev_fft(Pf), *Pid = dev_fft(Pi);
    const fafd_ops *od = dev_ops(ops);
    NEED(Pfd && Pid && od);
    BY_MT(b.x, hipLaunchKernelGGL(k_pipe<MT>, g, b, lds, S(stream), Pfd, Pid, spb,
                                  (const float2 *)post, (const float2 *)pre, od, x, in_dist, y,
                                  out_dist, count, scale));
    LAUNCH_CHECK("k_pipe");
    return 0;
}

extern "C" int fafk_stft(const fafd_fft *P, int spb, const void *post, const float *win,
                         const float *x, long rec_len, long rec_dist, long hop, long frames,
                         float *out_re, float *out_im, long frame_dist, long rec_out_dist,
                         long count, void *stream) {
    dim3 g(blocks_for(count, spb)), b(P->threads * spb);
    size_t lds = (size_t)spb * P->plane * sizeof(float2);
    const fafd_fft *Pd = dev_fft(P);
    NEED(Pd);
    BY_MT(b.x, hipLaunchKernelGGL(k_stft<MT>, g, b, lds, S(stream), Pd, spb,
                                  (const float2 *)post, win, x, rec_len, rec_dist, hop, frames,
                                  out_re, out_im, frame_dist, rec_out_dist, count));
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
    const fafd_fft *Phd = dev_fft(Ph), *Pnd = dev_fft(Pn);
    NEED(Phd && Pnd);
#define CWT(R) hipLaunchKernelGGL((k_cwt<R, MT>), g, b, lds, S(stream), Phd, Pnd, (const float2 *)post, psi, psi_stride, rows, rows_per_wg, x, in_dist, out, out_dist, count, scale)
    BY_MT(b.x, if (real_out) CWT(true); else CWT(false));
#undef CWT
    LAUNCH_CHECK("k_cwt");
    return 0;
}

extern "C" int fafk_icwt(const fafd_fft *Pf, const fafd_fft *Pi, const void *post,
                         const void *pre, const float *dual, long dual_stride, int rows,
                         const float *w, long in_dist, float *y, long out_dist, long count,
                         float scale, void *stream) {
    dim3 g((unsigned)count), b(Pf->threads);
    size_t lds = (size_t)2 * Pf->plane * sizeof(float2);
    const fafd_fft *Pfd = dev_fft(Pf), *Pid = dev_fft(Pi);
    NEED(Pfd && Pid);
    BY_MT(b.x, hipLaunchKernelGGL(k_icwt<MT>, g, b, lds, S(stream), Pfd, Pid,
                                  (const float2 *)post, (const float2 *)pre, dual, dual_stride,
                                  rows, w, in_dist, y, out_dist, count, scale));
    LAUNCH_CHECK("k_icwt");
    return 0;
}

/* L1 one-integral inverse CWT (reference src/faf_cwt.c:878-915): per
 * sample, the wavelet rows summed in row order with weight dlog = ln2/V,
 * divided by the admissibility constant C, plus the lowpass row (row 0 when
 * the bank has one).  One thread per sample; a row's loads are coalesced
 * and eight rows are in flight before they are summed in order. */
__global__ __launch_bounds__(256) void k_icwt_l1(const float *__restrict__ w, long in_dist,
                                                 float *__restrict__ y, long out_dist, long n,
                                                 int n_scales, int has_lp, float dlog, float c,
                                                 long bps) {
    const long seg = blockIdx.x / bps;
    const long i = (blockIdx.x - seg * bps) * 256 + threadIdx.x;
    if (i >= n) return;
    const float *ws = w + seg * in_dist + (has_lp ? n : 0) + i;
    float acc = 0.0f;
    int j = 0;
    for (; j + 8 <= n_scales; j += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ws[(long)(j + u) * n];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u] * dlog;
    }
    for (; j < n_scales; ++j) acc += ws[(long)j * n] * dlog;
    acc = acc / c;
    if (has_lp) acc += w[seg * in_dist + i];
    y[seg * out_dist + i] = acc;
}

extern "C" int fafk_icwt_l1(const float *w, long in_dist, float *y, long out_dist, long n,
                            int n_scales, int has_lp, float dlog, float c, long count,
                            void *stream) {
    const long bps = (n + 255) / 256;
    if (n <= 0 || count <= 0 || n_scales < 0 || bps * count > 0x7fffffffL) {
        snprintf(g_err, sizeof g_err, "k_icwt_l1: unsupported shape n=%ld count=%ld", n, count);
        return -1;
    }
    hipLaunchKernelGGL(k_icwt_l1, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream), w,
                       in_dist, y, out_dist, n, n_scales, has_lp, dlog, c, bps);
    LAUNCH_CHECK("k_icwt_l1");
    return 0;
}

/* ---- Rader and Bluestein beyond one workgroup: the element-wise steps
 * around the inner FFTs, which run through HBM.  Split planes; signal s of
 * a plane starts at s * dist; one thread per element, bps blocks of 256
 * threads per signal. ---- */

/* Rader gather: a[j] = x[rin[j]] (rin[j] = g^j mod p), j < m = p - 1 */
__global__ __launch_bounds__(256) void k_big_gather(const float *__restrict__ xr,
                                                    const float *__restrict__ xi, long xdist,
                                                    const int32_t *__restrict__ rin, long m,
                                                    float *__restrict__ ar, float *__restrict__ ai,
                                                    long adist, long bps) {
    const long s = blockIdx.x / bps;
    const long j = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (j >= m) return;
    const long src = s * xdist + rin[j];
    ar[s * adist + j] = xr[src];
    ai[s * adist + j] = xi[src];
}

/* Bluestein chirp in: a[k] = x[k] chirp[k] for k < n, zero for n <= k < M */
__global__ __launch_bounds__(256) void k_big_chirp_in(const float *__restrict__ xr,
                                                      const float *__restrict__ xi, long xdist,
                                                      const float2 *__restrict__ chirp, long n,
                                                      long M, float *__restrict__ ar,
                                                      float *__restrict__ ai, long adist,
                                                      long bps) {
    const long s = blockIdx.x / bps;
    const long k = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (k >= M) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (k < n) v = cmul(make_float2(xr[s * xdist + k], xi[s * xdist + k]), chirp[k]);
    ar[s * adist + k] = v.x;
    ai[s * adist + k] = v.y;
}

/* spectrum times the precomputed kernel: A[j] *= B[j]; save0, when given,
 * receives each signal's A[0] before the product (Rader's DC term) */
__global__ __launch_bounds__(256) void k_big_mul(float *__restrict__ ar, float *__restrict__ ai,
                                                 long adist, const float2 *__restrict__ B,
                                                 long m, float2 *__restrict__ save0, long bps) {
    const long s = blockIdx.x / bps;
    const long j = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (j >= m) return;
    const float2 a = make_float2(ar[s * adist + j], ai[s * adist + j]);
    if (save0 && j == 0) save0[s] = a;
    const float2 v = cmul(a, B[j]);
    ar[s * adist + j] = v.x;
    ai[s * adist + j] = v.y;
}

/* Rader scatter: y[rout[j]] = scale (x0 + a[j]); y[0] = scale (x0 + A0) */
__global__ __launch_bounds__(256) void k_big_rader_out(const float *__restrict__ ar,
                                                       const float *__restrict__ ai, long adist,
                                                       const int32_t *__restrict__ rout, long m,
                                                       const float *__restrict__ xr,
                                                       const float *__restrict__ xi, long xdist,
                                                       const float2 *__restrict__ save0,
                                                       float *__restrict__ yr,
                                                       float *__restrict__ yi, long ydist,
                                                       float scale, long bps) {
    const long s = blockIdx.x / bps;
    const long j = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (j >= m) return;
    const float2 x0 = make_float2(xr[s * xdist], xi[s * xdist]);
    const float2 v = cadd(x0, make_float2(ar[s * adist + j], ai[s * adist + j]));
    const long d = s * ydist + rout[j];
    yr[d] = v.x * scale;
    yi[d] = v.y * scale;
    if (j == 0) {
        const float2 dc = cadd(x0, save0[s]);
        yr[s * ydist] = dc.x * scale;
        yi[s * ydist] = dc.y * scale;
    }
}

/* Bluestein chirp out: y[k] = scale a[k] chirp[k], k < n */
__global__ __launch_bounds__(256) void k_big_chirp_out(const float *__restrict__ ar,
                                                       const float *__restrict__ ai, long adist,
                                                       const float2 *__restrict__ chirp, long n,
                                                       float *__restrict__ yr,
                                                       float *__restrict__ yi, long ydist,
                                                       float scale, long bps) {
    const long s = blockIdx.x / bps;
    const long k = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (k >= n) return;
    const float2 v = cmul(make_float2(ar[s * adist + k], ai[s * adist + k]), chirp[k]);
    yr[s * ydist + k] = v.x * scale;
    yi[s * ydist + k] = v.y * scale;
}

static int big_grid(long len, long count, long *bps, const char *what) {
    *bps = (len + 255) / 256;
    if (len <= 0 || count <= 0 || *bps * count > 0x7fffffffL) {
        snprintf(g_err, sizeof g_err, "%s: unsupported shape len=%ld count=%ld", what, len, count);
        return -1;
    }
    return 0;
}

extern "C" int fafk_big_gather(const float *xr, const float *xi, long xdist, const int32_t *rin,
                               long m, float *ar, float *ai, long adist, long count,
                               void *stream) {
    long bps;
    if (big_grid(m, count, &bps, "k_big_gather")) return -1;
    hipLaunchKernelGGL(k_big_gather, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream), xr,
                       xi, xdist, rin, m, ar, ai, adist, bps);
    LAUNCH_CHECK("k_big_gather");
    return 0;
}

extern "C" int fafk_big_chirp_in(const float *xr, const float *xi, long xdist,
                                 const void *chirp, long n, long M, float *ar, float *ai,
                                 long adist, long count, void *stream) {
    long bps;
    if (big_grid(M, count, &bps, "k_big_chirp_in")) return -1;
    hipLaunchKernelGGL(k_big_chirp_in, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream),
                       xr, xi, xdist, (const float2 *)chirp, n, M, ar, ai, adist, bps);
    LAUNCH_CHECK("k_big_chirp_in");
    return 0;
}

extern "C" int fafk_big_mul(float *ar, float *ai, long adist, const void *B, long m,
                            void *save0, long count, void *stream) {
    long bps;
    if (big_grid(m, count, &bps, "k_big_mul")) return -1;
    hipLaunchKernelGGL(k_big_mul, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream), ar, ai,
                       adist, (const float2 *)B, m, (float2 *)save0, bps);
    LAUNCH_CHECK("k_big_mul");
    return 0;
}

extern "C" int fafk_big_rader_out(const float *ar, const float *ai, long adist,
                                  const int32_t *rout, long m, const float *xr, const float *xi,
                                  long xdist, const void *save0, float *yr, float *yi,
                                  long ydist, float scale, long count, void *stream) {
    long bps;
    if (big_grid(m, count, &bps, "k_big_rader_out")) return -1;
    hipLaunchKernelGGL(k_big_rader_out, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream),
                       ar, ai, adist, rout, m, xr, xi, xdist, (const float2 *)save0, yr, yi,
                       ydist, scale, bps);
    LAUNCH_CHECK("k_big_rader_out");
    return 0;
}

extern "C" int fafk_big_chirp_out(const float *ar, const float *ai, long adist,
                                  const void *chirp, long n, float *yr, float *yi, long ydist,
                                  float scale, long count, void *stream) {
    long bps;
    if (big_grid(n, count, &bps, "k_big_chirp_out")) return -1;
    hipLaunchKernelGGL(k_big_chirp_out, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream),
                       ar, ai, adist, (const float2 *)chirp, n, yr, yi, ydist, scale, bps);
    LAUNCH_CHECK("k_big_chirp_out");
    return 0;
}

/* ---- DCT / DST / MDCT around one complex FFT ----
 * pre:  u[j] = pre[j] * x[j - shift] where 0 <= j - shift < nin, else 0
 *       (j < L; u split planes of distance udist)
 * post: y[k] = Re or Im of post[k] * U[k + off], k < nout */
__global__ __launch_bounds__(256) void k_r2r_pre(const float *__restrict__ x, long xdist,
                                                 long nin, const float2 *__restrict__ pre,
                                                 int shift, long L, float *__restrict__ ur,
                                                 float *__restrict__ ui, long udist, long bps) {
    const long s = blockIdx.x / bps;
    const long j = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (j >= L) return;
    const long i = j - shift;
    float2 v = make_float2(0.0f, 0.0f);
    if (i >= 0 && i < nin) {
        const float a = x[s * xdist + i];
        const float2 w = pre[j];
        v = make_float2(a * w.x, a * w.y);
    }
    ur[s * udist + j] = v.x;
    ui[s * udist + j] = v.y;
}

__global__ __launch_bounds__(256) void k_r2r_post(const float *__restrict__ ur,
                                                  const float *__restrict__ ui, long udist,
                                                  const float2 *__restrict__ post, int off,
                                                  int use_im, long nout,
                                                  float *__restrict__ y, long ydist, long bps) {
    const long s = blockIdx.x / bps;
    const long k = (blockIdx.x - s * bps) * 256 + threadIdx.x;
    if (k >= nout) return;
    const float2 u = make_float2(ur[s * udist + k + off], ui[s * udist + k + off]);
    const float2 v = cmul(post[k], u);
    y[s * ydist + k] = use_im ? v.y : v.x;
}

extern "C" int fafk_r2r_pre(const float *x, long xdist, long nin, const void *pre, int shift,
                            long L, float *ur, float *ui, long udist, long count, void *stream) {
    long bps;
    if (big_grid(L, count, &bps, "k_r2r_pre")) return -1;
    hipLaunchKernelGGL(k_r2r_pre, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream), x, xdist,
                       nin, (const float2 *)pre, shift, L, ur, ui, udist, bps);
    LAUNCH_CHECK("k_r2r_pre");
    return 0;
}

extern "C" int fafk_r2r_post(const float *ur, const float *ui, long udist, const void *post,
                             int off, int use_im, long nout, float *y, long ydist, long count,
                             void *stream) {
    long bps;
    if (big_grid(nout, count, &bps, "k_r2r_post")) return -1;
    hipLaunchKernelGGL(k_r2r_post, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream), ur, ui,
                       udist, (const float2 *)post, off, use_im, nout, y, ydist, bps);
    LAUNCH_CHECK("k_r2r_post");
    return 0;
}

/* Dual-frame accumulation of the long-segment inverse CWT (reference
 * src/faf_cwt.c:958-977): acc[k] = sum over rows r, in row order, of
 * Y_r[k] * dual_r[k], with the imaginary parts of DC and Nyquist cleared.
 * Y: interleaved row spectra [count][rows][nb]; acc: interleaved [count][nb].
 * One thread per bin; a row's bins are read contiguously by a wave. */
__global__ __launch_bounds__(256) void k_icwt_mulacc(const float2 *__restrict__ Y, long nb,
                                                     int rows, const float *__restrict__ dual,
                                                     long stride, float2 *__restrict__ acc,
                                                     long bps) {
    const long seg = blockIdx.x / bps;
    const long k = (blockIdx.x - seg * bps) * 256 + threadIdx.x;
    if (k >= nb) return;
    const float2 *y = Y + seg * rows * nb + k;
    float ar = 0.0f, ai = 0.0f;
    for (int r = 0; r < rows; ++r) {
        const float2 v = y[(long)r * nb];
        const float d = dual[(long)r * stride + k];
        ar = fmaf(v.x, d, ar);
        ai = fmaf(v.y, d, ai);
    }
    if (k == 0 || k == nb - 1) ai = 0.0f;
    acc[seg * nb + k] = make_float2(ar, ai);
}

extern "C" int fafk_icwt_mulacc(const float *Y, long nb, int rows, const float *dual,
                                long stride, float *acc, long count, void *stream) {
    const long bps = (nb + 255) / 256;
    if (nb <= 0 || rows <= 0 || count <= 0 || stride < nb || bps * count > 0x7fffffffL) {
        snprintf(g_err, sizeof g_err, "k_icwt_mulacc: unsupported shape nb=%ld rows=%d count=%ld",
                 nb, rows, count);
        return -1;
    }
    hipLaunchKernelGGL(k_icwt_mulacc, dim3((unsigned)(bps * count)), dim3(256), 0, S(stream),
                       (const float2 *)Y, nb, rows, dual, stride, (float2 *)acc, bps);
    LAUNCH_CHECK("k_icwt_mulacc");
    return 0;
}

extern "C" int fafk_fs_col(const fafd_fft *P1, int inv, int in_kind, int n2, int cols,
                           const void *twn, const float *in_re, const float *in_im,
                           long in_dist, float *t_re, float *t_im, long t_dist, long count,
                           void *stream) {
    dim3 g((unsigned)((n2 + cols - 1) / cols), (unsigned)count), b(P1->threads * cols);
    size_t lds = (size_t)cols * P1->plane * sizeof(float2);
    const float2 *tw = (const float2 *)twn;
    const fafd_fft *Pd = dev_fft(P1);
    NEED(Pd);
    /* the strided column and row passes hide HBM latency with occupancy:
     * the 128-register variant (four waves per SIMD) beats the spill-free
     * 512-thread one here */
    constexpr int MT = 1024;
#define FSC(I, K) hipLaunchKernelGGL((k_fs_col<I, K, MT>), g, b, lds, S(stream), Pd, n2, cols, tw, in_re, in_im, in_dist, t_re, t_im, t_dist)
    do {
          if (inv) {
              if (in_kind == FS_IN_SPLIT) FSC(true, FS_IN_SPLIT);
              else if (in_kind == FS_IN_PAIR) FSC(true, FS_IN_PAIR);
              else FSC(true, FS_IN_INTER);
          } else {
              if (in_kind == FS_IN_SPLIT) FSC(false, FS_IN_SPLIT);
              else if (in_kind == FS_IN_PAIR) FSC(false, FS_IN_PAIR);
              else FSC(false, FS_IN_INTER);
          }
    } while (0);
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
    const fafd_fft *Pd = dev_fft(P2);
    NEED(Pd);
    constexpr int MT = 1024;     /* as for the column pass */
#define FSR(I, K) hipLaunchKernelGGL((k_fs_row<I, K, MT>), g, b, lds, S(stream), Pd, n1, rows, t_re, t_im, t_dist, out_re, out_im, out_dist, scale)
    do {
          if (inv) {
              if (out_kind == FS_OUT_SPLIT) FSR(true, FS_OUT_SPLIT);
              else if (out_kind == FS_OUT_PAIR) FSR(true, FS_OUT_PAIR);
              else FSR(true, FS_OUT_INTER);
          } else {
              if (out_kind == FS_OUT_SPLIT) FSR(false, FS_OUT_SPLIT);
              else if (out_kind == FS_OUT_PAIR) FSR(false, FS_OUT_PAIR);
              else FSR(false, FS_OUT_INTER);
          }
    } while (0);
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
    const fafd_fft *Pfd = dev_fft(Pf), *Pid = dev_fft(Pi);
    NEED(Pfd && Pid);
    BY_MT(b.x, hipLaunchKernelGGL(k_rader<MT>, g, b, lds, S(stream), Pfd, Pid, *R, spb, in_re,
                                  in_im, in_dist, out_re, out_im, out_dist, count, scale));
    LAUNCH_CHECK("k_rader");
    return 0;
}

extern "C" int fafk_blue(const fafd_fft *Pf, const fafd_fft *Pi, const fafd_blue *Bq, int spb,
                         const float *in_re, const float *in_im, long in_dist, float *out_re,
                         float *out_im, long out_dist, long count, float scale, void *stream) {
    dim3 g(blocks_for(count, spb)), b(Pf->threads * spb);
    size_t lds = (size_t)spb * Pf->plane * sizeof(float2);
    const fafd_fft *Pfd = dev_fft(Pf), *Pid = dev_fft(Pi);
    NEED(Pfd && Pid);
    BY_MT(b.x, hipLaunchKernelGGL(k_blue<MT>, g, b, lds, S(stream), Pfd, Pid, *Bq, spb, in_re,
                                  in_im, in_dist, out_re, out_im, out_dist, count, scale));
    LAUNCH_CHECK("k_blue");
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
    BY_MT(block, hipLaunchKernelGGL(k_serve<MT>, dim3(grid), dim3(block), lds, S(stream),
                                    (const fafd_splan *)plans, (const fafd_sunit *)units,
                                    nunits, (const fafd_ssig *)sigs, counter));
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
