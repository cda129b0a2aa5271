/*
 * faf_engine_ct.h -- compile-time specialised Stockham passes for the
 * power-of-two lengths the batched paths live on (256 ... 16384).
 *
 * Same index scheme, accessors and FP32 arithmetic as faf_engine.h, but the
 * radix sequence, span, butterflies per thread and thread count are template
 * constants: no radix dispatch, span division folds to masks, and the kernel
 * carries __launch_bounds__ of its real block size, so the compiler keeps
 * every pass in registers (about 62 VGPRs instead of 128 plus spills).
 * Radix sequence for n = 2^a: 16, 16, ... then the leftover 2/4/8 last;
 * threads per signal T = n/16 (one radix-16 butterfly per thread).  The
 * twiddle loads of all passes are issued at the start, next to the
 * signal's HBM reads, so no pass waits on a table round trip.
 */
#ifndef FAF_ENGINE_CT_H
#define FAF_ENGINE_CT_H
/* the plain (non-template) kernels are defined once, in faf_kernels.hip */
#define FAF_ENGINE_NO_PLAIN_KERNELS
#include "faf_engine.h"

/* The twiddles of each butterfly come from the loaded powers W^(s 2^i)
 * (TwBits / tw_expand in faf_engine.h), so every pass's loads can be issued
 * up front, beside the signal's own HBM reads. */
/* The passes of one length as a chain: pass (NS, R) then the rest.
 * load_tw() issues the twiddle loads of every pass with span > 1; run()
 * performs the passes src -> LDS ... LDS -> dst (callers barrier after an
 * LDS sink). */
template <bool INV, int N, int T, int NS, int... Rs> struct CtChain {
    FAF_DEV void load_tw(int, int, const float2 *) {}
};

/* run_lazy() is run() with each pass's twiddle loads issued beside that
 * pass's own data reads instead of all at the start: fewer registers held
 * across the passes, for kernels whose occupancy is register-bound. */

template <bool INV, int N, int T, int NS, int R, int... Rest>
struct CtChain<INV, N, T, NS, R, Rest...> {
    static constexpr int NB = N / R, B = (NB + T - 1) / T, K = TwBits<R>::K;
    float2 tb[NS > 1 ? B : 1][NS > 1 ? K : 1];
    CtChain<INV, N, T, NS * R, Rest...> next;

    FAF_DEV void load_tw(int lt, int active, const float2 *__restrict__ tw) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lt + b * T;
                if (active && j < NB) {
                    const int step = (j % NS) * (N / (NS * R));
#pragma unroll
                    for (int i = 0; i < K; ++i) tb[b][i] = tw[step << i];
                }
            }
        }
        next.load_tw(lt, active, tw);
    }

    template <class S, class D>
    FAF_DEV void stage(const S &src, const D &dst, int lt, int active, bool sync) {
        float2 v[B][R];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = lt + b * T;
            if (active && j < NB) {
#pragma unroll
                for (int t = 0; t < R; ++t) v[b][t] = src.ld(j + t * NB);
            }
        }
        if (sync) __syncthreads();
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = lt + b * T;
            if (active && j < NB) {
                const int k = j % NS;
                if constexpr (NS > 1) {
                    float2 w[R];
                    tw_expand<R>(tb[b], w);
#pragma unroll
                    for (int t = 1; t < R; ++t) v[b][t] = cmul(v[b][t], w[t]);
                }
                dft<R, INV>(v[b]);
                const int base = (j - k) * R + k;
#pragma unroll
                for (int t = 0; t < R; ++t) dst.st(base + t * NS, v[b][t]);
            }
        }
    }

    template <class S, class D>
    FAF_DEV void run(const S &src, const Lds &lds, const D &dst, int lt, int active) {
        if constexpr (sizeof...(Rest) == 0) {
            stage(src, dst, lt, active, S::is_lds || (!S::is_lds && !D::is_lds));
        } else {
            stage(src, lds, lt, active, S::is_lds);
            __syncthreads();
            next.run(lds, lds, dst, lt, active);
        }
    }

    template <class S, class D, class TP>
    FAF_DEV void stage_lazy(const S &src, const D &dst, int lt, int active, bool sync,
                            TP __restrict__ tw) {
        float2 v[B][R];
        float2 tl[NS > 1 ? B : 1][NS > 1 ? K : 1];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = lt + b * T;
            if (active && j < NB) {
#pragma unroll
                for (int t = 0; t < R; ++t) v[b][t] = src.ld(j + t * NB);
                if constexpr (NS > 1) {
                    const int step = (j % NS) * (N / (NS * R));
#pragma unroll
                    for (int i = 0; i < K; ++i) tl[b][i] = twld(tw, step << i);
                }
            }
        }
        if (sync) __syncthreads();
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = lt + b * T;
            if (active && j < NB) {
                const int k = j % NS;
                if constexpr (NS > 1) {
                    float2 w[R];
                    tw_expand<R>(tl[b], w);
#pragma unroll
                    for (int t = 1; t < R; ++t) v[b][t] = cmul(v[b][t], w[t]);
                }
                dft<R, INV>(v[b]);
                const int base = (j - k) * R + k;
#pragma unroll
                for (int t = 0; t < R; ++t) dst.st(base + t * NS, v[b][t]);
            }
        }
    }

    template <class S, class D, class TP>
    FAF_DEV void run_lazy(const S &src, const Lds &lds, const D &dst, int lt, int active,
                          TP tw) {
        if constexpr (sizeof...(Rest) == 0) {
            stage_lazy(src, dst, lt, active, S::is_lds || (!S::is_lds && !D::is_lds), tw);
        } else {
            stage_lazy(src, lds, lt, active, S::is_lds, tw);
            __syncthreads();
            next.run_lazy(lds, lds, dst, lt, active, tw);
        }
    }
};

/* radix sequence of a power-of-two length, as a type */
template <int... Rs> struct Radices {};

template <int N> struct Pow2Seq;
template <> struct Pow2Seq<128> { using type = Radices<16, 8>; };
template <> struct Pow2Seq<256> { using type = Radices<16, 16>; };
template <> struct Pow2Seq<512> { using type = Radices<16, 16, 2>; };
template <> struct Pow2Seq<1024> { using type = Radices<16, 16, 4>; };
template <> struct Pow2Seq<2048> { using type = Radices<16, 16, 8>; };
template <> struct Pow2Seq<4096> { using type = Radices<16, 16, 16>; };
template <> struct Pow2Seq<8192> { using type = Radices<16, 16, 16, 2>; };
template <> struct Pow2Seq<16384> { using type = Radices<16, 16, 16, 4>; };

/* 5-smooth lengths with compile-time passes in the ragged launch.
 * Variant 0: the runtime planner's radix order (powers of two first, then
 * 3s, then 5s) and its thread count.  Variant 1 (and 2): composite
 * coprime radices (prime-factor codelets, no inner twiddles) that cut the
 * passes, and with them the LDS round trips and barriers, from 4-5 to 2-3;
 * the thread count is the largest butterfly count of a pass, or less where
 * a pass takes two butterflies per thread so that more signals fill a
 * 512-thread unit. */
template <int N, int V = 0> struct MixSeq { static constexpr int T = 0; };
template <> struct MixSeq<4800, 0> { using type = Radices<16, 4, 3, 5, 5>; static constexpr int T = 320; };
template <> struct MixSeq<3000, 0> { using type = Radices<8, 3, 5, 5, 5>; static constexpr int T = 200; };
template <> struct MixSeq<1000, 0> { using type = Radices<8, 5, 5, 5>; static constexpr int T = 67; };
template <> struct MixSeq<360, 0> { using type = Radices<8, 3, 3, 5>; static constexpr int T = 24; };
template <> struct MixSeq<60, 0> { using type = Radices<4, 3, 5>; static constexpr int T = 4; };
template <> struct MixSeq<4800, 1> { using type = Radices<12, 20, 20>; static constexpr int T = 240; };
template <> struct MixSeq<3000, 1> { using type = Radices<10, 20, 15>; static constexpr int T = 150; };
template <> struct MixSeq<3000, 2> { using type = Radices<30, 10, 10>; static constexpr int T = 100; };
template <> struct MixSeq<1000, 1> { using type = Radices<10, 10, 10>; static constexpr int T = 100; };
template <> struct MixSeq<360, 1> { using type = Radices<24, 15>; static constexpr int T = 24; };
template <> struct MixSeq<60, 1> { using type = Radices<15, 4>; static constexpr int T = 4; };

/* Rader primes of the ragged launch with compile-time inner passes over
 * p - 1 points.  Variant 0: the runtime planner's radix order for p - 1
 * (7s allowed) and threads per signal; variant 1: composite radices. */
template <int P, int V = 0> struct RaderSeq { static constexpr int T = 0; };
template <> struct RaderSeq<11, 0> { using type = Radices<2, 5>; static constexpr int T = 1; };
template <> struct RaderSeq<17, 0> { using type = Radices<16>; static constexpr int T = 1; };
template <> struct RaderSeq<31, 0> { using type = Radices<2, 3, 5>; static constexpr int T = 2; };
template <> struct RaderSeq<97, 0> { using type = Radices<16, 2, 3>; static constexpr int T = 7; };
template <> struct RaderSeq<1009, 0> { using type = Radices<16, 3, 3, 7>; static constexpr int T = 72; };
template <> struct RaderSeq<11, 1> { using type = Radices<10>; static constexpr int T = 1; };
template <> struct RaderSeq<31, 1> { using type = Radices<6, 5>; static constexpr int T = 6; };
template <> struct RaderSeq<97, 1> { using type = Radices<8, 12>; static constexpr int T = 12; };
template <> struct RaderSeq<1009, 1> { using type = Radices<12, 12, 7>; static constexpr int T = 84; };

/* the radices of a sequence into o[], returning their count */
template <int... Rs> constexpr int radices_out(Radices<Rs...>, int *o) {
    int k = 0;
    ((o[k++] = Rs), ...);
    return k;
}

template <bool INV, int N, int T, class Seq> struct MixFFT;
template <bool INV, int N, int T, int... Rs> struct MixFFT<INV, N, T, Radices<Rs...>> {
    using Chain = CtChain<INV, N, T, 1, Rs...>;
};

template <bool INV, int N, class Seq> struct CtFFT;
template <bool INV, int N, int... Rs> struct CtFFT<INV, N, Radices<Rs...>> {
    static constexpr int T = N / 16;
    using Chain = CtChain<INV, N, T, 1, Rs...>;
};

/* Ct<INV, N>::T threads per signal; Ct<INV, N>::Chain holds the passes */
template <bool INV, int N> using Ct = CtFFT<INV, N, typename Pow2Seq<N>::type>;

template <int T> constexpr int ct_spb() { return T >= 256 ? 1 : 256 / T; }

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */

template <bool INV, bool INTER, int N>
__global__ __launch_bounds__((Ct<INV, N>::T * ct_spb<Ct<INV, N>::T>()))
void kc_c2c(fafd_fft P, const float *in_re, const float *in_im, long in_dist, float *out_re,
            float *out_im, long out_dist, long count, float scale) {
    constexpr int T = Ct<INV, N>::T, SPB = ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * P.plane};
    const float2 *tw = (const float2 *)P.tw;
    typename Ct<INV, N>::Chain ch;
    ch.load_tw(lt, active, tw);
    if (INTER) {
        GInterIn src{(const float2 *)in_re + sig * in_dist};
        GInterOutNT dst{(float2 *)out_re + sig * out_dist, scale};
        ch.run(src, lds, dst, lt, active);
    } else {
        GSplitIn src{in_re + sig * in_dist, in_im + sig * in_dist};
        GSplitOutNT dst{out_re + sig * out_dist, out_im + sig * out_dist, scale};
        ch.run(src, lds, dst, lt, active);
    }
}

/* SPBX > 0 overrides the signals per workgroup (rfft 4096: 4, so that a
 * workgroup's packed rows make a run of whole aligned float4s) */
template <bool INTER_OUT, int NC, int SPBX = 0>
__global__ __launch_bounds__((Ct<false, NC>::T * (SPBX > 0 ? SPBX : ct_spb<Ct<false, NC>::T>())))
void kc_r2c(fafd_fft P, const float2 *post, const float *x, long in_dist, float *out_re,
            float *out_im, long out_dist, long count, float scale) {
    constexpr int T = Ct<false, NC>::T, SPB = SPBX > 0 ? SPBX : ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * P.plane};
    GPairInNT src{x + sig * in_dist};
    typename Ct<false, NC>::Chain ch;
    ch.load_tw(lt, active, (const float2 *)P.tw);
    /* the post-pass super-twiddles too are loaded up front, except at
     * NC = 8192, where the extra registers would cost a wave per SIMD
     * (measured: 4.1 -> 2.95 TB/s); elsewhere LDS, not registers, bounds
     * the occupancy */
    constexpr int NP = (NC / 2 + T) / T;
    constexpr bool HOIST = NC != 8192;
    float2 pw[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = lt + i * T;
        if (HOIST && active && k > 0 && k <= NC / 2) pw[i] = post[k];
    }
    ch.run(src, lds, lds, lt, active);
    __syncthreads();
    /* packed split rows of a whole workgroup are one contiguous run of
     * SPB (NC + 1) bins per plane: staged through LDS and stored as aligned
     * float4 when the run allows it (as in kc_stft) */
    constexpr int NB = NC + 1, RUN = SPB * NB;
    if constexpr (!INTER_OUT && HOIST && RUN % 4 == 0 &&
                  RUN * 2 * 4 <= SPB * (NC + NC / 16 + 1) * 8) {
        /* the bins are formed once, by one copy of the code, and only then
         * stored either way: a signal's bits do not depend on whether its
         * workgroup staged */
        float2 xs[NP][2];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            if (k == 0) {
                float2 z = lds.ld(0);
                xs[i][0] = make_float2(z.x + z.y, 0.0f);
                xs[i][1] = make_float2(z.x - z.y, 0.0f);
            } else {
                kiss_post(lds.ld(k), lds.ld(NC - k), pw[i], xs[i][0], xs[i][1]);
            }
        }
        const long g0 = (long)blockIdx.x * SPB;
        float *re0 = out_re + g0 * NB, *im0 = out_im + g0 * NB;
        if (out_dist == NB && g0 + SPB <= count &&
            ((((uintptr_t)re0) | ((uintptr_t)im0)) & 15) == 0) {
            __syncthreads();                      /* every signal is out of LDS */
            float *sre = (float *)faf_smem, *sim = sre + RUN;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const int k = lt + i * T;
                if (k > NC / 2) break;
                sre[sid * NB + k] = xs[i][0].x * scale; sim[sid * NB + k] = xs[i][0].y * scale;
                sre[sid * NB + NC - k] = xs[i][1].x * scale;
                sim[sid * NB + NC - k] = xs[i][1].y * scale;
            }
            __syncthreads();
            const float4 *s4r = (const float4 *)sre, *s4i = (const float4 *)sim;
            float4 *o4r = (float4 *)re0, *o4i = (float4 *)im0;
            /* streaming (non-temporal) stores: the rows are written once
             * (headline rfft 4096: 5.33-5.35 -> 5.41-5.43 TB/s, same bits) */
            for (int q = tid; q < RUN / 4; q += SPB * T) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                __builtin_nontemporal_store(*(const f4 *)&s4r[q], (f4 *)&o4r[q]);
                __builtin_nontemporal_store(*(const f4 *)&s4i[q], (f4 *)&o4i[q]);
            }
            return;
        }
        if (!active) return;
        float *ore = out_re + sig * out_dist, *oim = out_im + sig * out_dist;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            ore[k] = xs[i][0].x * scale; oim[k] = xs[i][0].y * scale;
            ore[NC - k] = xs[i][1].x * scale; oim[NC - k] = xs[i][1].y * scale;
        }
        return;
    }
    if (!active) return;
    float *ore = out_re + sig * out_dist;
    float *oim = out_im + sig * out_dist;
    float2 *oi = (float2 *)out_re + sig * out_dist;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = lt + i * T;
        if (k > NC / 2) break;
        const int m = NC - k;
        float2 xk, xm;
        if (k == 0) {
            float2 z = lds.ld(0);
            xk = make_float2(z.x + z.y, 0.0f);
            xm = make_float2(z.x - z.y, 0.0f);
        } else {
            kiss_post(lds.ld(k), lds.ld(m), HOIST ? pw[i] : post[k], xk, xm);
        }
        if (INTER_OUT) {
            oi[k] = cscale(xk, scale);
            oi[m] = cscale(xm, scale);
        } else {
            ore[k] = xk.x * scale; oim[k] = xk.y * scale;
            ore[m] = xm.x * scale; oim[m] = xm.y * scale;
        }
    }
}

template <bool INTER_IN, int NC>
__global__ __launch_bounds__((Ct<true, NC>::T * ct_spb<Ct<true, NC>::T>()))
void kc_c2r(fafd_fft P, const float2 *pre, const float *in_re, const float *in_im,
            long in_dist, float *y, long out_dist, long count, float scale) {
    constexpr int T = Ct<true, NC>::T, SPB = ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * P.plane};
    typename Ct<true, NC>::Chain ch;
    ch.load_tw(lt, active, (const float2 *)P.tw);
    if (active) {
        const float *ire = in_re + sig * in_dist;
        const float *iim = in_im + sig * in_dist;
        const float2 *ii = (const float2 *)in_re + sig * in_dist;
#pragma unroll
        for (int i = 0; i < (NC / 2 + T) / T; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            const int m = NC - k;
            float2 a = INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
            float2 b = INTER_IN ? ii[m] : make_float2(ire[m], iim[m]);
            if (k == 0) {
                lds.st(0, make_float2(a.x + b.x, a.x - b.x));
            } else {
                float2 fk, fm;
                kiss_pre(a, b, pre[k], fk, fm);
                lds.st(k, fk);
                if (m != k) lds.st(m, fm);
            }
        }
    }
    __syncthreads();
    GPairOut dst{y + sig * out_dist, scale};
    ch.run(lds, lds, dst, lt, active);
}

/* "Wide" rfft / irfft: NC/8 threads per signal and a radix-8 first pass, so
 * the whole workgroup (one signal) reads it with 8 loads per thread, and the
 * Kiss pass spreads its stores over twice the threads.  The later radix-16
 * passes leave half the threads idle, but one signal per workgroup halves
 * the LDS per workgroup, and with lazily loaded twiddles (63 VGPRs) eight
 * waves fit per SIMD, twice the staged kernel's.  Measured, rfft 4096:
 * 5.14 -> 5.21 TB/s; irfft 4096: 5.13 -> 5.27 TB/s; irfft 1024 +3%, rfft
 * 16384 +1.3%.  The launchers keep the staged kernels where these lose
 * (rfft 1024 -1%, irfft 16384 -20%). */
template <int NC> struct WideSeq;
template <> struct WideSeq<512> { using type = Radices<8, 8, 8>; };
template <> struct WideSeq<1024> { using type = Radices<8, 16, 8>; };
template <> struct WideSeq<2048> { using type = Radices<8, 16, 16>; };
template <> struct WideSeq<8192> { using type = Radices<8, 16, 16, 4>; };
template <bool INV, int NC> using Wide =
    MixFFT<INV, NC, NC / 8, typename WideSeq<NC>::type>;

template <bool INTER_OUT, int NC>
__global__ __launch_bounds__(NC / 8)
void kc_r2cw(fafd_fft P, const float2 *post, const float *x, long in_dist, float *out_re,
             float *out_im, long out_dist, long count, float scale) {
    constexpr int T = NC / 8;
    const int lt = (int)threadIdx.x;
    const long sig = blockIdx.x;
    Lds lds{faf_smem};
    GPairIn src{x + sig * in_dist};
    typename Wide<false, NC>::Chain ch;
    constexpr int NP = (NC / 2 + T) / T;
    ch.run_lazy(src, lds, lds, lt, 1, (const float2 *)P.tw);
    __syncthreads();
    float *ore = out_re + sig * out_dist;
    float *oim = out_im + sig * out_dist;
    float2 *oi = (float2 *)out_re + sig * out_dist;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = lt + i * T;
        if (k > NC / 2) break;
        const int m = NC - k;
        float2 xk, xm;
        if (k == 0) {
            float2 z = lds.ld(0);
            xk = make_float2(z.x + z.y, 0.0f);
            xm = make_float2(z.x - z.y, 0.0f);
        } else {
            kiss_post(lds.ld(k), lds.ld(m), post[k], xk, xm);
        }
        if (INTER_OUT) {
            oi[k] = cscale(xk, scale);
            oi[m] = cscale(xm, scale);
        } else {
            ore[k] = xk.x * scale; oim[k] = xk.y * scale;
            ore[m] = xm.x * scale; oim[m] = xm.y * scale;
        }
    }
}

template <bool INTER_IN, int NC>
__global__ __launch_bounds__(NC / 8)
void kc_c2rw(fafd_fft P, const float2 *pre, const float *in_re, const float *in_im,
             long in_dist, float *y, long out_dist, long count, float scale) {
    constexpr int T = NC / 8;
    const int lt = (int)threadIdx.x;
    const long sig = blockIdx.x;
    Lds lds{faf_smem};
    typename Wide<true, NC>::Chain ch;
    const float *ire = in_re + sig * in_dist;
    const float *iim = in_im + sig * in_dist;
    const float2 *ii = (const float2 *)in_re + sig * in_dist;
#pragma unroll
    for (int i = 0; i < (NC / 2 + T) / T; ++i) {
        const int k = lt + i * T;
        if (k > NC / 2) break;
        const int m = NC - k;
        float2 a = INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
        float2 b = INTER_IN ? ii[m] : make_float2(ire[m], iim[m]);
        if (k == 0) {
            lds.st(0, make_float2(a.x + b.x, a.x - b.x));
        } else {
            float2 fk, fm;
            kiss_pre(a, b, pre[k], fk, fm);
            lds.st(k, fk);
            if (m != k) lds.st(m, fm);
        }
    }
    __syncthreads();
    GPairOutNT dst{y + sig * out_dist, scale};
    ch.run_lazy(lds, lds, dst, lt, 1, (const float2 *)P.tw);
}


/* MUL1: ops is one MUL (op_mul); otherwise the general apply_ops */
template <int NC, bool MUL1>
__global__ __launch_bounds__((Ct<false, NC>::T * ct_spb<Ct<false, NC>::T>()))
void kc_pipe(fafd_fft Pf, fafd_fft Pi, const float2 *post, const float2 *pre, fafd_ops ops,
             const float *x, long in_dist, float *y, long out_dist, long count, float scale) {
    constexpr int T = Ct<false, NC>::T, SPB = ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * Pf.plane};
    GPairIn src{x + sig * in_dist};
    typename Ct<false, NC>::Chain fw;
    fw.load_tw(lt, active, (const float2 *)Pf.tw);
    fw.run(src, lds, lds, lt, active);
    typename Ct<true, NC>::Chain bw;
    bw.load_tw(lt, active, (const float2 *)Pi.tw);
    __syncthreads();
    if (active) {
#pragma unroll
        for (int i = 0; i < (NC / 2 + T) / T; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            const int m = NC - k;
            float2 xk, xm;
            auto op = [&](int b, float2 v) { return MUL1 ? op_mul(ops, b, v) : apply_ops(ops, b, v); };
            if (k == 0) {
                float2 z = lds.ld(0);
                xk = op(0, make_float2(z.x + z.y, 0.0f));
                xm = op(NC, make_float2(z.x - z.y, 0.0f));
                lds.st(0, make_float2(xk.x + xm.x, xk.x - xm.x));
            } else {
                kiss_post(lds.ld(k), lds.ld(m), post[k], xk, xm);
                xk = op(k, xk);
                xm = (m != k) ? op(m, xm) : xk;
                float2 fk, fm;
                kiss_pre(xk, xm, pre[k], fk, fm);
                lds.st(k, fk);
                if (m != k) lds.st(m, fm);
            }
        }
    }
    __syncthreads();
    GPairOut dst{y + sig * out_dist, scale};
    bw.run(lds, lds, dst, lt, active);
}

/* The single-MUL pipeline with one signal per workgroup of NC/8 threads
 * (the kc_r2cw layout for both FFTs, twiddles per pass) */
template <int NC>
__global__ __launch_bounds__(NC / 8)
void kc_pipew(fafd_fft Pf, fafd_fft Pi, const float2 *post, const float2 *pre, fafd_ops ops,
              const float *x, long in_dist, float *y, long out_dist, long count, float scale) {
    constexpr int T = NC / 8;
    const int lt = (int)threadIdx.x;
    const long sig = blockIdx.x;
    Lds lds{faf_smem};
    GPairIn src{x + sig * in_dist};
    typename Wide<false, NC>::Chain fw;
    fw.run_lazy(src, lds, lds, lt, 1, (const float2 *)Pf.tw);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < (NC / 2 + T) / T; ++i) {
        const int k = lt + i * T;
        if (k > NC / 2) break;
        const int m = NC - k;
        float2 xk, xm;
        if (k == 0) {
            float2 z = lds.ld(0);
            xk = op_mul(ops, 0, make_float2(z.x + z.y, 0.0f));
            xm = op_mul(ops, NC, make_float2(z.x - z.y, 0.0f));
            lds.st(0, make_float2(xk.x + xm.x, xk.x - xm.x));
        } else {
            kiss_post(lds.ld(k), lds.ld(m), post[k], xk, xm);
            xk = op_mul(ops, k, xk);
            xm = (m != k) ? op_mul(ops, m, xm) : xk;
            float2 fk, fm;
            kiss_pre(xk, xm, pre[k], fk, fm);
            lds.st(k, fk);
            if (m != k) lds.st(m, fm);
        }
    }
    __syncthreads();
    GPairOut dst{y + sig * out_dist, scale};
    typename Wide<true, NC>::Chain bw;
    bw.run_lazy(lds, lds, dst, lt, 1, (const float2 *)Pi.tw);
}

/* FULL: every frame lies inside its record (record length >= window) */
template <bool FULL, int NC>
__global__ __launch_bounds__((Ct<false, NC>::T * ct_spb<Ct<false, NC>::T>()))
void kc_stft(fafd_fft P, const float2 *post, const float *win, const float *x, long rec_len,
             long rec_dist, long hop, long frames, float *out_re, float *out_im,
             long frame_dist, long rec_out_dist, long count) {
    constexpr int T = Ct<false, NC>::T, SPB = ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * P.plane};
    const long rec = sig / frames, f = sig - rec * frames;
    typename Ct<false, NC>::Chain ch;
    ch.load_tw(lt, active, (const float2 *)P.tw);
    constexpr int NP = (NC / 2 + T) / T;
    float2 pw[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = lt + i * T;
        if (active && k > 0 && k <= NC / 2) pw[i] = post[k];
    }
    if constexpr (FULL && NC <= 512) {
        /* the window (2 NC floats, at most 4 KB) is copied once into LDS
         * behind the frame planes; every frame of the workgroup reads it from
         * there instead of issuing its own window loads (STFT 64 x 2^22,
         * window 1024: 1.734 -> 1.442 ms).  Four workgroups still fit a CU. */
        float *wl = (float *)(faf_smem + SPB * P.plane);
        for (int i = tid; i < 2 * NC; i += SPB * T) wl[i] = win[i];
        __syncthreads();
        GFrameFull src{x + rec * rec_dist + f * hop, wl};
        ch.run(src, lds, lds, lt, active);
    } else if constexpr (FULL) {
        GFrameFull src{x + rec * rec_dist + f * hop, win};
        ch.run(src, lds, lds, lt, active);
    } else {
        GFrameIn src{x + rec * rec_dist + f * hop, win, rec_len - f * hop};
        ch.run(src, lds, lds, lt, active);
    }
    __syncthreads();
    /* A workgroup of whole frames writes one contiguous run of SPB (NC + 1)
     * bins per plane (frame rows are adjacent, record after record).  When
     * that run starts 16-byte aligned and fills float4s, the bins are staged
     * through LDS and leave as aligned float4 stores: four times fewer store
     * instructions, every one a full contiguous line set. */
    constexpr int NB = NC + 1, RUN = SPB * NB;
    if constexpr (FULL && RUN % 4 == 0 && RUN * 2 * 4 <= SPB * (NC + NC / 16 + 1) * 8) {
        /* bins formed once, then stored either way (as in kc_r2c) */
        float2 xs[NP][2];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            if (k == 0) {
                float2 z = lds.ld(0);
                xs[i][0] = make_float2(z.x + z.y, 0.0f);
                xs[i][1] = make_float2(z.x - z.y, 0.0f);
            } else {
                kiss_post(lds.ld(k), lds.ld(NC - k), pw[i], xs[i][0], xs[i][1]);
            }
        }
        const long g0 = (long)blockIdx.x * SPB;
        float *re0 = out_re + g0 * NB, *im0 = out_im + g0 * NB;
        if (frame_dist == NB && rec_out_dist == frames * NB && g0 + SPB <= count &&
            ((((uintptr_t)re0) | ((uintptr_t)im0)) & 15) == 0) {
            __syncthreads();                      /* every frame is out of LDS */
            float *sre = (float *)faf_smem, *sim = sre + RUN;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const int k = lt + i * T;
                if (k > NC / 2) break;
                sre[sid * NB + k] = xs[i][0].x; sim[sid * NB + k] = xs[i][0].y;
                sre[sid * NB + NC - k] = xs[i][1].x; sim[sid * NB + NC - k] = xs[i][1].y;
            }
            __syncthreads();
            const float4 *s4r = (const float4 *)sre, *s4i = (const float4 *)sim;
            float4 *o4r = (float4 *)re0, *o4i = (float4 *)im0;
            /* streaming (non-temporal) stores: the spectrum rows are 80% of
             * the traffic (STFT 64 x 2^22: 1.53-1.55 -> 1.47 ms, same bits) */
            for (int q = tid; q < RUN / 4; q += SPB * T) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                __builtin_nontemporal_store(*(const f4 *)&s4r[q], (f4 *)&o4r[q]);
                __builtin_nontemporal_store(*(const f4 *)&s4i[q], (f4 *)&o4i[q]);
            }
            return;
        }
        if (!active) return;
        float *ore = out_re + rec * rec_out_dist + f * frame_dist;
        float *oim = out_im + rec * rec_out_dist + f * frame_dist;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const int k = lt + i * T;
            if (k > NC / 2) break;
            ore[k] = xs[i][0].x; oim[k] = xs[i][0].y;
            ore[NC - k] = xs[i][1].x; oim[NC - k] = xs[i][1].y;
        }
        return;
    }
    if (!active) return;
    float *ore = out_re + rec * rec_out_dist + f * frame_dist;
    float *oim = out_im + rec * rec_out_dist + f * frame_dist;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int k = lt + i * T;
        if (k > NC / 2) break;
        const int m = NC - k;
        float2 xk, xm;
        if (k == 0) {
            float2 z = lds.ld(0);
            xk = make_float2(z.x + z.y, 0.0f);
            xm = make_float2(z.x - z.y, 0.0f);
        } else {
            kiss_post(lds.ld(k), lds.ld(m), pw[i], xk, xm);
        }
        ore[k] = xk.x; oim[k] = xk.y;
        ore[m] = xm.x; oim[m] = xm.y;
    }
}

/* CWT bank on one segment per workgroup (rows split over workgroups when
 * segments are few): rfft(N) into X, then for each row r the one-sided
 * X psi_r through an inverse N-point pass into W and out.  Both passes'
 * twiddles are loaded once and serve every row of the workgroup. */
template <bool REAL_OUT, int N>
__global__ __launch_bounds__(N / 16)
void kc_cwt(fafd_fft Ph, fafd_fft Pn, const float2 *post, const float *psi, long psi_stride,
            int rows, int rows_per_wg, const float *x, long in_dist, float *out,
            long out_dist, long count, float scale) {
    constexpr int NC = N / 2, TH = NC / 16;
    const int groups = (rows + rows_per_wg - 1) / rows_per_wg;
    const long seg = blockIdx.x / groups;
    const int g = (int)(blockIdx.x - seg * groups);
    const int tid = (int)threadIdx.x;
    const int ah = tid < TH && seg < count, an = seg < count;
    const int lh = tid < TH ? tid : 0;
    Lds X{faf_smem};
    Lds W{faf_smem + Ph.plane};
    typename Ct<false, NC>::Chain fh;
    typename Ct<true, N>::Chain fn;
    fh.load_tw(lh, ah, (const float2 *)Ph.tw);
    fn.load_tw(tid, an, (const float2 *)Pn.tw);
    fh.run(GPairIn{x + (seg < count ? seg : 0) * in_dist}, X, X, lh, ah);
    __syncthreads();
    if (ah) {
#pragma unroll
        for (int i = 0; i < (NC / 2 + TH) / TH; ++i) {
            const int k = tid + i * TH;
            if (k > NC / 2) break;
            const int m = NC - k;
            float2 xk, xm;
            if (k == 0) {
                float2 z = X.ld(0);
                xk = make_float2(z.x + z.y, 0.0f);
                xm = make_float2(z.x - z.y, 0.0f);
            } else {
                kiss_post(X.ld(k), X.ld(m), post[k], xk, xm);
            }
            X.st(k, xk);
            X.st(m, xm);
        }
    }
    __syncthreads();
    const int r0 = g * rows_per_wg;
    const int r1 = r0 + rows_per_wg < rows ? r0 + rows_per_wg : rows;
    for (int r = r0; r < r1; ++r) {
        OneSidedIn src{X, psi + (long)r * psi_stride, NC};
        const long o = (seg < count ? seg : 0) * out_dist + (long)r * N;
        if (REAL_OUT) fn.run(src, W, GRealOnlyOut{out + o, scale}, tid, an);
        else fn.run(src, W, GInterOutNT{(float2 *)out + o, scale}, tid, an);
        __syncthreads();
    }
}

/* ------------------------------------------------------------------ */
/* four-step for power-of-two lengths beyond one workgroup, n = N1 N2   */
/* ------------------------------------------------------------------ */

/* what the column pass reads: complex pairs of a real row (R2C), complex
 * (split or interleaved), or the Kiss pre-pass of a Hermitian spectrum
 * (C2R, split or interleaved) */
enum { F4_IN_PAIR = 0, F4_IN_SPLIT = 1, F4_IN_INTER = 2, F4_PRE_SPLIT = 3, F4_PRE_INTER = 4 };
/* what the row pass writes: complex (split / interleaved, natural order),
 * the real pairs of a C2R row, or the Kiss post-pass bins of an R2C row */
enum { F4_OUT_SPLIT = 0, F4_OUT_INTER = 1, F4_OUT_PAIR = 2, F4_POST_SPLIT = 3, F4_POST_INTER = 4 };

constexpr int F4_COLS = 32;   /* columns per column-pass workgroup */
constexpr int F4_ROWS = 32;   /* row-pass blocks hold 2 F4_ROWS rows (measured: 32 beats 16) */

/* element i of a Hermitian spectrum with nc+1 bins (split or interleaved) */
template <bool INTER> FAF_DEV float2 herm_ld(const float *re, const float *im, long i) {
    return INTER ? ((const float2 *)re)[i] : make_float2(re[i], im[i]);
}

/* Column pass: columns c0 .. c0+COLS of the [N1][N2] view, FFT N1, times
 * W_n^(k1 c), into ws[k1][N2] (interleaved).  in_dist in floats, ws_dist
 * in float2.  grid (N2/COLS, signals). */
template <int IN, bool INV, int N1, int N2>
__global__ __launch_bounds__(F4_COLS * (N1 / 16))
void kc4_col(fafd_fft P1, const float2 *twn, const void *aux, const float *in_re,
             const float *in_im, long in_dist, float2 *ws, long ws_dist) {
    constexpr int T = N1 / 16, NC = N1 * N2, PL = N1 + N1 / 16 + 1;
    const long b = blockIdx.y;
    const int c0 = blockIdx.x * F4_COLS, tid = (int)threadIdx.x;
    const int sid = tid / T, lt = tid - sid * T;
    typename Ct<INV, N1>::Chain ch;
    ch.load_tw(lt, 1, (const float2 *)P1.tw);
    const float *xr = in_re + b * in_dist;
    const float *xi = in_im ? in_im + b * in_dist : nullptr;
    for (int f = tid; f < F4_COLS * N1; f += F4_COLS * T) {
        const int c = f % F4_COLS, r = f / F4_COLS;
        const long i = (long)r * N2 + c0 + c;
        float2 v;
        if (IN == F4_IN_PAIR) {
            v = ((const float2 *)xr)[i];
        } else if (IN == F4_IN_SPLIT) {
            v = make_float2(xr[i], xi[i]);
        } else if (IN == F4_IN_INTER) {
            v = ((const float2 *)xr)[i];
        } else {                      /* Kiss pre-pass, F[i] from X[i], X[NC - i] */
            constexpr bool HI = IN == F4_PRE_INTER;
            const float2 *pre = (const float2 *)aux;
            if (i == 0) {
                const float2 a = herm_ld<HI>(xr, xi, 0), e = herm_ld<HI>(xr, xi, NC);
                v = make_float2(a.x + e.x, a.x - e.x);
            } else if (i <= NC / 2) {
                float2 fk, fm;
                kiss_pre(herm_ld<HI>(xr, xi, i), herm_ld<HI>(xr, xi, NC - i), pre[i], fk, fm);
                v = fk;
            } else {
                float2 fk, fm;
                kiss_pre(herm_ld<HI>(xr, xi, NC - i), herm_ld<HI>(xr, xi, i), pre[NC - i], fk, fm);
                v = fm;
            }
        }
        Lds{faf_smem + c * PL}.st(r, v);
    }
    __syncthreads();
    Lds lds{faf_smem + sid * PL};
    ch.run(lds, lds, lds, lt, 1);
    __syncthreads();
    float2 *w = ws + b * ws_dist;
    for (int f = tid; f < F4_COLS * N1; f += F4_COLS * T) {
        const int c = f % F4_COLS, k1 = f / F4_COLS;
        const float2 v = Lds{faf_smem + c * PL}.ld(k1);
        w[(long)k1 * N2 + c0 + c] = cmul(v, twn[(long)k1 * (c0 + c)]);
    }
}

/* Row pass (all row distances in floats): rows of ws [N1][N2], FFT N2,
 * results Z[k1 + N1 k2].  For the
 * R2C post-pass a workgroup holds a block A of rows k1 in [0, N1/2] and
 * their mirrors N1 - k1, so the partner Z[nc - k] of every bin it writes
 * is in its LDS: X[k] = post(Z[k], Z[nc - k]) is written straight to the
 * Hermitian output and the packed spectrum never goes back to HBM.
 * Other outputs use plain blocks of ROWS rows.  grid (blocks, signals). */
template <int OUT, bool INV, int N1, int N2, int RW>
__global__ __launch_bounds__(2 * RW * (N2 / 16))
void kc4_row(fafd_fft P2, const float2 *ws, long ws_dist, const float2 *post, float *out_re,
             float *out_im, long out_dist, float scale) {
    constexpr int T = N2 / 16, NC = N1 * N2, PL = N2 + N2 / 16 + 1;
    constexpr bool POST = OUT == F4_POST_SPLIT || OUT == F4_POST_INTER;
    __shared__ int s_row[2 * RW];
    __shared__ int s_slot_of_mirror[2 * RW];
    __shared__ int s_nrows;
    const long b = blockIdx.y;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x;
    if (tid == 0) {
        int n = 0;
        if (POST) {
            /* block x holds A = [x RW, x RW + RW) (block 0 also N1/2) and
             * the mirrors N1 - a of a != 0, N1/2: 2 RW rows every block */
            const int a0 = blockIdx.x * RW;
            for (int a = a0; a < a0 + RW; ++a) s_row[n++] = a;
            if (blockIdx.x == 0) s_row[n++] = N1 / 2;
            const int na = n;
            for (int i = 0; i < na; ++i) {
                const int a = s_row[i];
                if (a == 0 || 2 * a == N1) {
                    s_slot_of_mirror[i] = i;
                } else {
                    s_row[n] = N1 - a;
                    s_slot_of_mirror[i] = n;
                    s_slot_of_mirror[n] = i;
                    ++n;
                }
            }
        } else {
            const int a0 = blockIdx.x * 2 * RW;
            for (int a = a0; a < a0 + 2 * RW && a < N1; ++a) s_row[n++] = a;
        }
        s_nrows = n;
    }
    __syncthreads();
    const int nrows = s_nrows;
    const int sid = tid / T, lt = tid - sid * T;
    typename Ct<INV, N2>::Chain ch;
    ch.load_tw(lt, sid < nrows, (const float2 *)P2.tw);
    const float2 *w = ws + b * ws_dist;
    for (int f = tid; f < nrows * N2; f += bd) {
        const int s = f / N2, c = f - s * N2;
        Lds{faf_smem + s * PL}.st(c, w[(long)s_row[s] * N2 + c]);
    }
    __syncthreads();
    Lds lds{faf_smem + sid * PL};
    ch.run(lds, lds, lds, lt, sid < nrows);
    __syncthreads();
    /* k2-major order: consecutive threads take consecutive rows, so each
     * store instruction covers runs of adjacent output bins */
    for (int f = tid; f < nrows * N2; f += bd) {
        const int s = f % nrows, k2 = f / nrows;
        const int r = s_row[s];
        const long k = (long)r + (long)N1 * k2;
        const float2 z = Lds{faf_smem + s * PL}.ld(k2);
        if (POST) {
            /* partner Z[nc - k]: row (N1 - r) mod N1, column N2-1-k2, or
             * N2-k2 (mod N2) on row 0 */
            const int ms = s_slot_of_mirror[s];
            const int mk2 = r == 0 ? (N2 - k2) & (N2 - 1) : N2 - 1 - k2;
            const float2 zm = Lds{faf_smem + ms * PL}.ld(mk2);
            float2 xk, xm;
            if (k == 0) {
                xk = make_float2(z.x + z.y, 0.0f);
                xm = make_float2(z.x - z.y, 0.0f);   /* Nyquist bin NC */
            } else if (k <= NC / 2) {
                kiss_post(z, zm, post[k], xk, xm);
            } else {
                float2 t;
                kiss_post(zm, z, post[NC - k], t, xk);
            }
            xk = cscale(xk, scale);
            if (OUT == F4_POST_INTER) {
                float2 *o = (float2 *)(out_re + b * out_dist);
                o[k] = xk;
                if (k == 0) o[NC] = cscale(xm, scale);
            } else {
                out_re[b * out_dist + k] = xk.x;
                out_im[b * out_dist + k] = xk.y;
                if (k == 0) {
                    out_re[b * out_dist + NC] = xm.x * scale;
                    out_im[b * out_dist + NC] = 0.0f;
                }
            }
        } else {
            const float2 v = cscale(z, scale);
            if (OUT == F4_OUT_INTER || OUT == F4_OUT_PAIR)
                ((float2 *)(out_re + b * out_dist))[k] = v;
            else {
                out_re[b * out_dist + k] = v.x;
                out_im[b * out_dist + k] = v.y;
            }
        }
    }
}

#endif
