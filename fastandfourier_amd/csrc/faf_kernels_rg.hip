/*
 * faf_kernels_rg.hip -- the ragged persistent launch: C2C signals of uneven
 * lengths in one launch (host side: faf_ragged.c).
 *
 * Workgroups pull units (up to spb same-length signals) from an atomic
 * counter until the queue is empty.  A unit of power-of-two length
 * 256 ... 8192 (T = N/16 threads per signal) or of one of the 5-smooth
 * lengths of MixSeq runs the compile-time passes of that length
 * (faf_engine_ct.h), and a Rader prime of RaderSeq runs Rader with
 * compile-time inner passes; other lengths run the plan's runtime passes,
 * Rader or Bluestein.  The length switch is uniform across
 * the workgroup, so the barriers inside every path stay uniform too.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "faf_engine_ct.h"
#define FAF_RS_CORE_ONLY
#include "faf_engine_rs.h"
#include "faf_kern.h"

static thread_local char g_rg_err[160];

static int rg_fail(const char *what, hipError_t e) {
    snprintf(g_rg_err, sizeof g_rg_err, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

__global__ void kg_fill_u32(unsigned *p, int n, unsigned v) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = v;
}

/* Passes issue their twiddle loads lazily (CtChain::run_lazy): the kernel's
 * occupancy is set by its register count, not by latency inside one pass. */
template <int N>
FAF_DEV void rg_ct(const fafd_rplan *pl, const GSplitIn &src, const GSplitOut &dst, int sid,
                   int lt, int active) {
    Lds lds{faf_smem + sid * pl->fwd.plane};
    typename Ct<false, N>::Chain ch;
    ch.run_lazy(src, lds, dst, lt, active, (const float2 *)pl->fwd.tw);
}

template <int N>
FAF_DEV void rg_mix(const fafd_rplan *pl, const GSplitIn &src, const GSplitOut &dst, int sid,
                    int lt, int active) {
    Lds lds{faf_smem + sid * pl->fwd.plane};
    typename MixFFT<false, N, MixSeq<N>::T, typename MixSeq<N>::type>::Chain ch;
    ch.run_lazy(src, lds, dst, lt, active, (const float2 *)pl->fwd.tw);
}

/* Rader for prime P with compile-time passes over the P - 1 inner points;
 * the same steps as rader_body (faf_engine.h) */
template <int P>
FAF_DEV void rg_rader(const fafd_rplan *pl, const GSplitIn &src, const GSplitOut &dst,
                      int sid, int lt, int active) {
    constexpr int T = RaderSeq<P>::T;
    using Seq = typename RaderSeq<P>::type;
    const fafd_rader &R = pl->rader;
    const int per = pl->fwd.plane + lpad(P) + 1;
    Lds xa{faf_smem + sid * per};
    Lds aa{faf_smem + sid * per + lpad(P) + 1};
    if (active)
        for (int i = lt; i < P; i += T) xa.st(i, src.ld(i));
    __syncthreads();
    const float2 x0 = active ? xa.ld(0) : make_float2(0.0f, 0.0f);
    GatherLds g{xa, R.rin};
    typename MixFFT<false, P - 1, T, Seq>::Chain fw;
    fw.run_lazy(g, aa, aa, lt, active, (const float2 *)pl->fwd.tw);
    __syncthreads();
    if (active && lt == 0) xa.st(0, cadd(x0, aa.ld(0)));   /* DC = x0 + sum x[1..] */
    MulLds mb{aa, (const float2 *)R.b};
    ScatterLds sc{xa, R.rout, x0};
    typename MixFFT<true, P - 1, T, Seq>::Chain bw;
    bw.run_lazy(mb, aa, sc, lt, active, (const float2 *)pl->inv.tw);
    __syncthreads();
    if (active)
        for (int i = lt; i < P; i += T) dst.st(i, xa.ld(i));
}

/* GEN: also the runtime Stockham, Rader and Bluestein paths.  Without them
 * the kernel needs far fewer registers (more workgroups per CU); the host
 * picks it when every length of the batch has compile-time passes. */
/* a pointer that is the same in every lane of the wave, in scalar registers */
template <class P> FAF_DEV P *wave_uniform(P *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (P *)(((uint64_t)hi << 32) | lo);
}

/* 8192 points: the single-pass register-resident FFT, 256 threads and
 * 35 KB of LDS per signal, so a 512-thread workgroup runs two signals.  A
 * signal's 256 threads are four whole waves, so its plane pointers are
 * wave-uniform: in scalar registers the 64 loads and stores address by a
 * 32-bit lane offset instead of a 64-bit address per load. */
FAF_DEV void rg_rs8(const fafd_rplan *pl, const GSplitIn &src, const GSplitOut &dst, int sid,
                    int lt, int active) {
    float *lds = (float *)faf_smem + sid * (rs8::LDS_BYTES / 4);
    const float2 *tw = (const float2 *)pl->fwd.tw;
    const float *ir = wave_uniform(src.re), *ii = wave_uniform(src.im);
    float *orr = wave_uniform(dst.re), *oi = wave_uniform(dst.im);
    float2 v[32], b1[4][3];
#pragma unroll
    for (int m = 0; m < 32; ++m) {
        const int k = lt + rs8::T * m;
        v[m] = active ? make_float2(ir[k], ii[k]) : make_float2(0.0f, 0.0f);
    }
    rs8::tw1_bases(tw, lt, b1);
    rs8::fft8192<false>(v, lds, tw, b1, lt);
    if (active) {
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int k = lt + rs8::T * j;
            orr[k] = v[j].x;
            oi[k] = v[j].y;
        }
    }
}

template <int MT, bool GEN>
__global__ __launch_bounds__(MT, GEN ? 1 : 4) void kg_ragged(const fafd_rplan *plans, const fafd_unit *units,
                                                int nunits, const int32_t *order,
                                                const long *offsets, const float *in_re,
                                                const float *in_im, float *out_re,
                                                float *out_im, unsigned *counter) {
    __shared__ int s_unit;
    const int tid = (int)threadIdx.x;
    for (;;) {
        if (tid == 0) s_unit = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int u = s_unit;
        __syncthreads();
        if (u >= nunits) break;
        const fafd_unit un = units[u];
        const fafd_rplan *pl = plans + un.plan;
        const int T = pl->threads;
        const int sid = tid / T, lt = tid - sid * T;
        const int active = sid < un.count;
        const int s = active ? order[un.first + sid] : order[un.first];
        const long off = offsets[s];
        GSplitIn src{in_re + off, in_im + off};
        GSplitOut dst{out_re + off, out_im + off, 1.0f};
        if (pl->kind == FAFD_KIND_CT) {
            switch (pl->len) {
            case 256: rg_ct<256>(pl, src, dst, sid, lt, active); break;
            case 512: rg_ct<512>(pl, src, dst, sid, lt, active); break;
            case 1024: rg_ct<1024>(pl, src, dst, sid, lt, active); break;
            case 2048: rg_ct<2048>(pl, src, dst, sid, lt, active); break;
            case 4096: rg_ct<4096>(pl, src, dst, sid, lt, active); break;
            case 8192: rg_rs8(pl, src, dst, sid, lt, active); break;
            case 60: rg_mix<60>(pl, src, dst, sid, lt, active); break;
            case 360: rg_mix<360>(pl, src, dst, sid, lt, active); break;
            case 1000: rg_mix<1000>(pl, src, dst, sid, lt, active); break;
            case 3000: rg_mix<3000>(pl, src, dst, sid, lt, active); break;
            case 4800: rg_mix<4800>(pl, src, dst, sid, lt, active); break;
            default: break;             /* the host sets this kind for the lengths above only */
            }
        } else if (pl->kind == FAFD_KIND_RADER_CT) {
            switch (pl->len) {
            case 11: rg_rader<11>(pl, src, dst, sid, lt, active); break;
            case 17: rg_rader<17>(pl, src, dst, sid, lt, active); break;
            case 31: rg_rader<31>(pl, src, dst, sid, lt, active); break;
            case 97: rg_rader<97>(pl, src, dst, sid, lt, active); break;
            case 1009: rg_rader<1009>(pl, src, dst, sid, lt, active); break;
            default: break;
            }
        } else if constexpr (GEN) {
            if (pl->kind == FAFD_KIND_STOCKHAM) {
                Lds lds{faf_smem + sid * pl->fwd.plane};
                fft_run<false>(pl->fwd, src, lds, dst, lt, active);
            } else if (pl->kind == FAFD_KIND_RADER) {
                const int per = pl->fwd.plane + lpad(pl->rader.p) + 1;
                Lds xa{faf_smem + sid * per};
                Lds aa{faf_smem + sid * per + lpad(pl->rader.p) + 1};
                rader_body(pl->fwd, pl->inv, pl->rader, src, dst, xa, aa, lt, T, active);
            } else {
                Lds aa{faf_smem + sid * pl->fwd.plane};
                blue_body(pl->fwd, pl->inv, pl->blue, src, dst, aa, lt, active);
            }
        }
        __syncthreads();
    }
}

extern "C" int fafk_rg_init(void) {
    (void)hipFuncSetAttribute((const void *)kg_ragged<512, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              FAFD_MAX_LDS_BYTES - 64);
    (void)hipFuncSetAttribute((const void *)kg_ragged<512, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              FAFD_MAX_LDS_BYTES - 64);
    (void)hipGetLastError();
    return 0;
}

extern "C" int fafk_ragged(const void *plans, const void *units, int nunits,
                           const int32_t *order, const long *offsets, const float *in_re,
                           const float *in_im, float *out_re, float *out_im,
                           unsigned *counter, int grid, int block, size_t lds, int generic,
                           void *stream) {
    hipLaunchKernelGGL(kg_fill_u32, dim3(1), dim3(64), 0, S(stream), counter, 1, 0u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return rg_fail("kg_fill_u32", e);
    if (block > 512) {               /* the host plans every unit for 512 threads */
        snprintf(g_rg_err, sizeof g_rg_err, "ragged launch of %d threads per workgroup", block);
        return -1;
    }
    if (generic)
        hipLaunchKernelGGL((kg_ragged<512, true>), dim3(grid), dim3(block), lds, S(stream),
                           (const fafd_rplan *)plans, (const fafd_unit *)units, nunits, order,
                           offsets, in_re, in_im, out_re, out_im, counter);
    else
        hipLaunchKernelGGL((kg_ragged<512, false>), dim3(grid), dim3(block), lds, S(stream),
                           (const fafd_rplan *)plans, (const fafd_unit *)units, nunits, order,
                           offsets, in_re, in_im, out_re, out_im, counter);
    e = hipGetLastError();
    return e == hipSuccess ? 0 : rg_fail("kg_ragged", e);
}

extern "C" const char *fafk_rg_last_error(void) { return g_rg_err; }

/* LDS per signal (float2 slots) of the compile-time path of length n when it
 * is not the plan's plane, else 0 */
extern "C" int fafk_rg_ct_slots(int n) { return n == 8192 ? rs8::LDS_BYTES / 8 : 0; }

/* threads per signal of the Rader path for prime p with compile-time inner
 * passes, 0 if p has none */
extern "C" int fafk_rg_rader_threads(int p) {
    switch (p) {
    case 11: return RaderSeq<11>::T;
    case 17: return RaderSeq<17>::T;
    case 31: return RaderSeq<31>::T;
    case 97: return RaderSeq<97>::T;
    case 1009: return RaderSeq<1009>::T;
    default: return 0;
    }
}

/* threads per signal of the compile-time passes of length n in the ragged
 * launch, 0 if n has none */
extern "C" int fafk_rg_ct_threads(int n) {
    switch (n) {
    case 256: case 512: case 1024: case 2048: case 4096: return n / 16;
    case 8192: return rs8::T;
    case 60: return MixSeq<60>::T;
    case 360: return MixSeq<360>::T;
    case 1000: return MixSeq<1000>::T;
    case 3000: return MixSeq<3000>::T;
    case 4800: return MixSeq<4800>::T;
    default: return 0;
    }
}
