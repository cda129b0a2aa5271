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
};

/* radix sequence of a power-of-two length, as a type */
template <int... Rs> struct Radices {};

template <int N> struct Pow2Seq;
template <> struct Pow2Seq<256> { using type = Radices<16, 16>; };
template <> struct Pow2Seq<512> { using type = Radices<16, 16, 2>; };
template <> struct Pow2Seq<1024> { using type = Radices<16, 16, 4>; };
template <> struct Pow2Seq<2048> { using type = Radices<16, 16, 8>; };
template <> struct Pow2Seq<4096> { using type = Radices<16, 16, 16>; };
template <> struct Pow2Seq<8192> { using type = Radices<16, 16, 16, 2>; };
template <> struct Pow2Seq<16384> { using type = Radices<16, 16, 16, 4>; };

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
        GInterOut dst{(float2 *)out_re + sig * out_dist, scale};
        ch.run(src, lds, dst, lt, active);
    } else {
        GSplitIn src{in_re + sig * in_dist, in_im + sig * in_dist};
        GSplitOut dst{out_re + sig * out_dist, out_im + sig * out_dist, scale};
        ch.run(src, lds, dst, lt, active);
    }
}

template <bool INTER_OUT, int NC>
__global__ __launch_bounds__((Ct<false, NC>::T * ct_spb<Ct<false, NC>::T>()))
void kc_r2c(fafd_fft P, const float2 *post, const float *x, long in_dist, float *out_re,
            float *out_im, long out_dist, long count, float scale) {
    constexpr int T = Ct<false, NC>::T, SPB = ct_spb<T>();
    const int tid = (int)threadIdx.x, sid = tid / T, lt = tid % T;
    const long g = (long)blockIdx.x * SPB + sid;
    const int active = g < count;
    const long sig = active ? g : 0;
    Lds lds{faf_smem + sid * P.plane};
    GPairIn src{x + sig * in_dist};
    typename Ct<false, NC>::Chain ch;
    ch.load_tw(lt, active, (const float2 *)P.tw);
    ch.run(src, lds, lds, lt, active);
    __syncthreads();
    if (!active) return;
    float *ore = out_re + sig * out_dist;
    float *oim = out_im + sig * out_dist;
    float2 *oi = (float2 *)out_re + sig * out_dist;
#pragma unroll
    for (int i = 0; i < (NC / 2 + T) / T; ++i) {
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

template <int NC>
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
            if (k == 0) {
                float2 z = lds.ld(0);
                xk = apply_ops(ops, 0, make_float2(z.x + z.y, 0.0f));
                xm = apply_ops(ops, NC, make_float2(z.x - z.y, 0.0f));
                lds.st(0, make_float2(xk.x + xm.x, xk.x - xm.x));
            } else {
                kiss_post(lds.ld(k), lds.ld(m), post[k], xk, xm);
                xk = apply_ops(ops, k, xk);
                xm = (m != k) ? apply_ops(ops, m, xm) : xk;
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

template <int NC>
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
    GFrameIn src{x + rec * rec_dist + f * hop, win, rec_len - f * hop};
    typename Ct<false, NC>::Chain ch;
    ch.load_tw(lt, active, (const float2 *)P.tw);
    ch.run(src, lds, lds, lt, active);
    __syncthreads();
    if (!active) return;
    float *ore = out_re + rec * rec_out_dist + f * frame_dist;
    float *oim = out_im + rec * rec_out_dist + f * frame_dist;
#pragma unroll
    for (int i = 0; i < (NC / 2 + T) / T; ++i) {
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
        else fn.run(src, W, GInterOut{(float2 *)out + o, scale}, tid, an);
        __syncthreads();
    }
}

#endif
