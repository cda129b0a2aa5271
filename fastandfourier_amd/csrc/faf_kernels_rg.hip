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

/* Table pointers (twiddles, Rader index and chirp tables) come out of the
 * plan records in device memory, so the compiler cannot tell their address
 * space: as generic pointers their loads are flat loads, which also count
 * against the LDS wait counter (every LDS wait of a pass then waits for the
 * pass's twiddles too).  Every table is device memory: these give the loads
 * global pointers. */
FAF_DEV const faf_gtw *gtw(const void *p) { return (const faf_gtw *)p; }

/* A unit's signals are addressed from its first signal: one base pointer per
 * plane, the same for the whole workgroup (scalar registers), and a per-lane
 * offset d in elements of the lane's signal from that base.  The host keeps
 * every unit within 2^30 elements of its first signal, so the byte offsets
 * fit 32 bits and each load and store takes the scalar-base form with a
 * 32-bit lane offset (a per-lane 64-bit pointer costs a 64-bit add per
 * access group). */
struct RgIn {
    const float *re, *im;
    uint32_t d;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const {
        const uint32_t b = (d + (uint32_t)i) * 4u;
        return make_float2(*(const float *)((const char *)re + b),
                           *(const float *)((const char *)im + b));
    }
};

struct RgOut {
    float *re, *im;
    uint32_t d;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const {
        const uint32_t b = (d + (uint32_t)i) * 4u;
        *(float *)((char *)re + b) = v.x;
        *(float *)((char *)im + b) = v.y;
    }
};

/* Passes issue their twiddle loads lazily (CtChain::run_lazy): the kernel's
 * occupancy is set by its register count, not by latency inside one pass.
 * Only the Rader paths read their tables through global pointers (1009
 * alone: 18.6 -> 17.3 us, the C3 mix 1-2% faster); the 5-smooth chains and
 * the 8192 pass keep generic twiddle pointers, as global ones measured no
 * faster there. */
template <int N>
FAF_DEV void rg_ct(const fafd_rplan *pl, const RgIn &src, const RgOut &dst, int sid,
                   int lt, int active) {
    Lds lds{faf_smem + sid * pl->fwd.plane};
    typename Ct<false, N>::Chain ch;
    ch.run_lazy(src, lds, dst, lt, active, (const float2 *)pl->fwd.tw);
}

template <int N, int V>
FAF_DEV void rg_mix(const fafd_rplan *pl, const RgIn &src, const RgOut &dst, int sid,
                    int lt, int active) {
    Lds lds{faf_smem + sid * pl->fwd.plane};
    typename MixFFT<false, N, MixSeq<N, V>::T, typename MixSeq<N, V>::type>::Chain ch;
    ch.run_lazy(src, lds, dst, lt, active, (const float2 *)pl->fwd.tw);
}

/* Rader for prime P with compile-time passes over the P - 1 inner points:
 * the steps of rader_body (faf_engine.h) in one LDS buffer per signal.  The
 * plane of P - 1 points has lpad(P - 1) + 1 slots, room for the P samples:
 * the signal is staged there, the first forward pass gathers from it and
 * (after its barrier) writes the pass in place, and the last inverse pass
 * scatters y[rout[j]] back into it, again after its barrier; y[0] goes
 * straight to the output. */
template <int P, int V>
FAF_DEV void rg_rader(const fafd_rplan *pl, const RgIn &src, const RgOut &dst,
                      int sid, int lt, int active) {
    constexpr int T = RaderSeq<P, V>::T;
    using Seq = typename RaderSeq<P, V>::type;
    const fafd_rader &R = pl->rader;
    Lds aa{faf_smem + sid * pl->fwd.plane};
    if (active)
        for (int i = lt; i < P; i += T) aa.st(i, src.ld(i));
    __syncthreads();
    const float2 x0 = active ? aa.ld(0) : make_float2(0.0f, 0.0f);
    GatherLds g{aa, (const faf_gi32 *)R.rin};
    typename MixFFT<false, P - 1, T, Seq>::Chain fw;
    fw.run_lazy(g, aa, aa, lt, active, gtw(pl->fwd.tw));
    __syncthreads();
    if (active && lt == 0) dst.st(0, cadd(x0, aa.ld(0)));   /* DC = x0 + sum x[1..] */
    MulLds mb{aa, gtw(R.b)};
    ScatterLds sc{aa, (const faf_gi32 *)R.rout, x0};
    typename MixFFT<true, P - 1, T, Seq>::Chain bw;
    bw.run_lazy(mb, aa, sc, lt, active, gtw(pl->inv.tw));
    __syncthreads();
    if (active)
        for (int i = 1 + lt; i < P; i += T) dst.st(i, aa.ld(i));
}

/* 8192 points: the single-pass register-resident FFT, 256 threads and
 * 35 KB of LDS per signal (two signals per 512-thread unit).  A signal's
 * 256 threads are four whole waves, so its plane pointers and its active
 * flag are wave-uniform.  The lane index is taken from threadIdx.x so that
 * the compiler knows its range: with the pointers in scalar registers every
 * load and store then addresses by a 32-bit lane offset plus an immediate.
 * An inactive slot (a 512-thread unit holding one signal) reads the signal
 * of slot 0, which its source points at, and stores nothing, so the loads
 * need no per-load branch (with one the compiler formed a branch and a
 * 64-bit address per load pair: 13% more VALU instructions than the
 * batched kernel). */
FAF_DEV void rg_rs8(const fafd_rplan *pl, const RgIn &src, const RgOut &dst, int sid,
                    int, int active) {
    const int lt = (int)threadIdx.x & (rs8::T - 1);
    float *lds = (float *)faf_smem + __builtin_amdgcn_readfirstlane(sid) * (rs8::LDS_BYTES / 4);
    const float2 *tw = (const float2 *)pl->fwd.tw;
    /* global (not generic) pointers: loads and stores stay global_* */
    typedef __attribute__((address_space(1))) float gf;
    const uint32_t du = __builtin_amdgcn_readfirstlane(src.d);   /* whole waves per slot */
    const gf *ir = (const gf *)(src.re + du), *ii = (const gf *)(src.im + du);
    gf *orr = (gf *)(dst.re + du), *oi = (gf *)(dst.im + du);
    const int act = __builtin_amdgcn_readfirstlane(active);
    float2 v[32], b1[4][3];
    /* the twiddle bases first: issued after the data loads they keep more
     * registers live (129 against 125 VGPRs, over the kernel's 128) */
    rs8::tw1_bases(tw, lt, b1);
    if (act) {
#pragma unroll
        for (int m = 0; m < 32; ++m) {
            const int k = lt + rs8::T * m;
            v[m] = make_float2(ir[k], ii[k]);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 32; ++m) v[m] = make_float2(0.0f, 0.0f);
    }
    rs8::fft8192<false>(v, lds, tw, b1, lt);
    if (act) {
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int k = lt + rs8::T * j;
            orr[k] = v[j].x;
            oi[k] = v[j].y;
        }
    }
}

/* The compile-time code paths of the ragged kernel: (id, kind, length,
 * variant, body).  The host stores the id of each plan's path in the plan
 * record (fafk_rg_path), and the kernel dispatches on it with one flat
 * switch: nested tests of kind, length and variant made the register
 * allocator spill.  The 8192-point single pass, the path with the most
 * registers live (112 alone), comes first: placed after the others it
 * spilled 28-36 bytes per lane in the combined kernel, first it does not. */
#define RG_PATHS(X)                                                          \
    X(0, FAFD_KIND_CT, 8192, 0, (rg_rs8))                                    \
    X(1, FAFD_KIND_CT, 256, 0, (rg_ct<256>))                                 \
    X(2, FAFD_KIND_CT, 512, 0, (rg_ct<512>))                                 \
    X(3, FAFD_KIND_CT, 1024, 0, (rg_ct<1024>))                               \
    X(4, FAFD_KIND_CT, 2048, 0, (rg_ct<2048>))                               \
    X(5, FAFD_KIND_CT, 4096, 0, (rg_ct<4096>))                               \
    X(6, FAFD_KIND_CT, 60, 0, (rg_mix<60, 0>))                               \
    X(7, FAFD_KIND_CT, 60, 1, (rg_mix<60, 1>))                               \
    X(8, FAFD_KIND_CT, 360, 0, (rg_mix<360, 0>))                             \
    X(9, FAFD_KIND_CT, 360, 1, (rg_mix<360, 1>))                             \
    X(10, FAFD_KIND_CT, 1000, 0, (rg_mix<1000, 0>))                          \
    X(11, FAFD_KIND_CT, 1000, 1, (rg_mix<1000, 1>))                          \
    X(12, FAFD_KIND_CT, 3000, 0, (rg_mix<3000, 0>))                          \
    X(13, FAFD_KIND_CT, 3000, 1, (rg_mix<3000, 1>))                          \
    X(14, FAFD_KIND_CT, 3000, 2, (rg_mix<3000, 2>))                          \
    X(15, FAFD_KIND_CT, 4800, 0, (rg_mix<4800, 0>))                          \
    X(16, FAFD_KIND_CT, 4800, 1, (rg_mix<4800, 1>))                          \
    X(17, FAFD_KIND_RADER_CT, 11, 0, (rg_rader<11, 0>))                      \
    X(18, FAFD_KIND_RADER_CT, 11, 1, (rg_rader<11, 1>))                      \
    X(19, FAFD_KIND_RADER_CT, 17, 0, (rg_rader<17, 0>))                      \
    X(20, FAFD_KIND_RADER_CT, 31, 0, (rg_rader<31, 0>))                      \
    X(21, FAFD_KIND_RADER_CT, 31, 1, (rg_rader<31, 1>))                      \
    X(22, FAFD_KIND_RADER_CT, 97, 0, (rg_rader<97, 0>))                      \
    X(23, FAFD_KIND_RADER_CT, 97, 1, (rg_rader<97, 1>))                      \
    X(24, FAFD_KIND_RADER_CT, 1009, 0, (rg_rader<1009, 0>))                  \
    X(25, FAFD_KIND_RADER_CT, 1009, 1, (rg_rader<1009, 1>))

/* One unit: up to spb signals of one plan.  The switch is uniform across
 * the workgroup, so the barriers inside every path stay uniform.  GEN: also
 * the runtime Stockham, Rader and Bluestein paths.  Without them the kernel
 * needs far fewer registers (more workgroups per CU); the host picks it
 * when every length of the batch has compile-time passes. */
template <bool GEN>
FAF_DEV void rg_unit(int path, const fafd_rplan *pl, const RgIn &src, const RgOut &dst,
                     int sid, int lt, int active) {
    switch (path) {
#define RG_CASE(id, kind, n, v, body) \
    case id: body(pl, src, dst, sid, lt, active); break;
        RG_PATHS(RG_CASE)
#undef RG_CASE
    default:
        if constexpr (GEN) {
            if (pl->kind == FAFD_KIND_STOCKHAM) {
                Lds lds{faf_smem + sid * pl->fwd.plane};
                fft_run<false>(pl->fwd, src, lds, dst, lt, active);
            } else if (pl->kind == FAFD_KIND_RADER) {
                const int per = pl->fwd.plane + lpad(pl->rader.p) + 1;
                Lds xa{faf_smem + sid * per};
                Lds aa{faf_smem + sid * per + lpad(pl->rader.p) + 1};
                rader_body(pl->fwd, pl->inv, pl->rader, src, dst, xa, aa, lt, pl->threads,
                           active);
            } else if (pl->kind == FAFD_KIND_BLUESTEIN) {
                Lds aa{faf_smem + sid * pl->fwd.plane};
                blue_body(pl->fwd, pl->inv, pl->blue, src, dst, aa, lt, active);
            }
        }
        break;
    }
}

/* the code path of (kind, length, variant), or -1 if the kernel has none */
extern "C" int fafk_rg_path(int kind, int n, int v) {
#define RG_FIND(id, k, len, var, body) \
    if (kind == k && n == len && v == var) return id;
    RG_PATHS(RG_FIND)
#undef RG_FIND
    return -1;
}

/* Two launch shapes over the same unit body.
 * PERSIST false: one workgroup per unit (grid = units, unit = blockIdx);
 * the hardware dispatcher hands out the cost-ordered units in order.
 * PERSIST true: the grid is the workgroups resident at once; workgroup b
 * starts on unit b and then takes the next from counter[0] (set to the grid
 * size) until the queue is empty; the last workgroup to leave resets
 * counter[0] to the grid size and counter[1] to 0 for the next launch.
 * uoff[i] is the offset of the signal in slot i of the unit list. */
template <bool GEN>
FAF_DEV void rg_run_unit(const fafd_rplan *plans, const fafd_unit *units, int u, const long *uoff,
                         const float *in_re, const float *in_im, float *out_re, float *out_im,
                         int tid) {
    /* everything the data loads need comes from the unit record: a unit of
     * one signal reads no offset list and no plan record before them */
    const fafd_unit un = units[u];
    const fafd_rplan *pl = plans + un.plan;
    const int T = un.threads;
    const int sid = tid / T, lt = tid - sid * T;
    const int active = sid < un.count;
    uint32_t d = 0;
    if (un.count > 1) d = (uint32_t)(uoff[un.first + (active ? sid : 0)] - un.off0);
    RgIn src{in_re + un.off0, in_im + un.off0, d};
    RgOut dst{out_re + un.off0, out_im + un.off0, d};
    rg_unit<GEN>(un.path, pl, src, dst, sid, lt, active);
}

template <int MT, bool GEN, bool PERSIST>
__global__ __launch_bounds__(MT, GEN ? 1 : 4) void kg_ragged(const fafd_rplan *plans, const fafd_unit *units,
                                                int nunits, const long *uoff,
                                                const float *in_re, const float *in_im,
                                                float *out_re, float *out_im,
                                                unsigned *counter) {
    const int tid = (int)threadIdx.x;
    if constexpr (!PERSIST) {
        rg_run_unit<GEN>(plans, units, (int)blockIdx.x, uoff, in_re, in_im, out_re, out_im, tid);
    } else {
        __shared__ int s_unit;
        int u = (int)blockIdx.x;
        while (u < nunits) {
            rg_run_unit<GEN>(plans, units, u, uoff, in_re, in_im, out_re, out_im, tid);
            __syncthreads();
            if (tid == 0) s_unit = (int)atomicAdd(counter, 1u);
            __syncthreads();
            u = s_unit;
        }
        if (tid == 0) {
            __threadfence();
            if (atomicAdd(counter + 1, 1u) == gridDim.x - 1) {
                atomicExch(counter, gridDim.x);
                atomicExch(counter + 1, 0u);
            }
        }
    }
}

#define RG_KERNELS(X) X(true, false) X(false, false) X(true, true) X(false, true)

extern "C" int fafk_rg_init(void) {
#define RG_ATTR(G, P)                                                                  \
    (void)hipFuncSetAttribute((const void *)kg_ragged<512, G, P>,                      \
                              hipFuncAttributeMaxDynamicSharedMemorySize,              \
                              FAFD_MAX_LDS_BYTES - 64);
    RG_KERNELS(RG_ATTR)
#undef RG_ATTR
    (void)hipGetLastError();
    return 0;
}

/* counter[0] = grid, counter[1] = 0: the state kg_ragged expects (and
 * leaves behind) */
extern "C" int fafk_ragged_counter_init(unsigned *counter, int grid, void *stream) {
    hipLaunchKernelGGL(kg_fill_u32, dim3(1), dim3(64), 0, S(stream), counter, 1, (unsigned)grid);
    hipLaunchKernelGGL(kg_fill_u32, dim3(1), dim3(64), 0, S(stream), counter + 1, 1, 0u);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rg_fail("kg_fill_u32", e);
}

/* grid = nunits when not persistent */
extern "C" int fafk_ragged(const void *plans, const void *units, int nunits, const long *uoff,
                           const float *in_re, const float *in_im, float *out_re,
                           float *out_im, unsigned *counter, int grid, int block, size_t lds,
                           int generic, int persist, void *stream) {
    if (block > 512) {               /* the host plans every unit for 512 threads */
        snprintf(g_rg_err, sizeof g_rg_err, "ragged launch of %d threads per workgroup", block);
        return -1;
    }
    if (!persist) grid = nunits;
#define RG_LAUNCH(G, P)                                                                \
    if (generic == G && (persist != 0) == P)                                           \
        hipLaunchKernelGGL((kg_ragged<512, G, P>), dim3(grid), dim3(block), lds, S(stream), \
                           (const fafd_rplan *)plans, (const fafd_unit *)units, nunits,  \
                           uoff, in_re, in_im, out_re, out_im, counter);
    RG_KERNELS(RG_LAUNCH)
#undef RG_LAUNCH
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : rg_fail("kg_ragged", e);
}

extern "C" const char *fafk_rg_last_error(void) { return g_rg_err; }

/* LDS per signal (float2 slots) of the compile-time path of length n when it
 * is not the plan's plane, else 0 */
extern "C" int fafk_rg_ct_slots(int n) { return n == 8192 ? rs8::LDS_BYTES / 8 : 0; }

/* Compile-time pass sequences of the ragged launch.  For length n
 * (Rader: prime n, the sequence of its n - 1 inner points) and variant v:
 * the threads per signal, or 0 if there is no such sequence; with order
 * non-NULL also its radices (returned in *nr). */
template <class Seq, int T> static int seq_out(int *order, int *nr) {
    if (order && nr) *nr = radices_out(typename Seq::type{}, order);
    return T;
}
#define RG_SEQ(S, N, V)                                                      \
    if (n == N && v == V) return seq_out<S<N, V>, S<N, V>::T>(order, nr);
extern "C" int fafk_rg_rader_seq(int n, int v, int *order, int *nr) {
    RG_SEQ(RaderSeq, 11, 0) RG_SEQ(RaderSeq, 17, 0) RG_SEQ(RaderSeq, 31, 0)
    RG_SEQ(RaderSeq, 97, 0) RG_SEQ(RaderSeq, 1009, 0)
    RG_SEQ(RaderSeq, 11, 1) RG_SEQ(RaderSeq, 31, 1) RG_SEQ(RaderSeq, 97, 1)
    RG_SEQ(RaderSeq, 1009, 1)
    return 0;
}

extern "C" int fafk_rg_ct_seq(int n, int v, int *order, int *nr) {
    if (v == 0) {
        switch (n) {
        case 256: case 512: case 1024: case 2048: case 4096: return n / 16;
        case 8192:                       /* the single pass: 8 x 32 x 32 in registers */
            if (order && nr) { order[0] = 8; order[1] = 32; order[2] = 32; *nr = 3; }
            return rs8::T;
        default: break;
        }
    }
    RG_SEQ(MixSeq, 60, 0) RG_SEQ(MixSeq, 360, 0) RG_SEQ(MixSeq, 1000, 0)
    RG_SEQ(MixSeq, 3000, 0) RG_SEQ(MixSeq, 4800, 0)
    RG_SEQ(MixSeq, 60, 1) RG_SEQ(MixSeq, 360, 1) RG_SEQ(MixSeq, 1000, 1)
    RG_SEQ(MixSeq, 3000, 1) RG_SEQ(MixSeq, 4800, 1) RG_SEQ(MixSeq, 3000, 2)
    return 0;
}
#undef RG_SEQ
