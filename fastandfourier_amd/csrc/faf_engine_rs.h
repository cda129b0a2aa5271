/*
 * faf_engine_rs.h -- single-pass 32768-point complex FFT held in registers:
 * the core of the N = 65536 rfft / irfft (Kiss packing halves it to 32768
 * complex points) and of C2C 32768.
 *
 * One workgroup of 1024 threads owns a whole signal, 32 complex values per
 * thread (64 VGPRs; 256 KB per signal, half the CU's register file), so the
 * signal makes one HBM read and one HBM write.  The four-step split of the
 * same length makes two of each.  32768 = 32 * 32 * 32, natural order in and
 * out:
 *   pass 1  thread t          DFT-32 over m of z[t + 1024 m], times W_32768^(t k1)
 *   pass 2  thread (t1, k1)   DFT-32 over t2 of B[t1 + 32 t2][k1], times W_1024^(t1 k2a)
 *   pass 3  thread (k1, k2a)  DFT-32 over t1, giving X[k1 + 32 k2a + 1024 k2b]
 * so after pass 3 thread tid holds X[tid + 1024 j], j = 0..31, and stores of
 * one register index are contiguous across the workgroup, as are the loads
 * of pass 1.  Between passes the values go through LDS one plane (re, then
 * im) at a time: 32 x 1056 floats = 132 KB.  The 1056-float row pitch puts
 * lanes 32..63 of a wave 32 banks away from lanes 0..31, and the pass-2 to
 * pass-3 layout uses a 33-float pitch between k1 rows, so every LDS access
 * of a wave touches 64 distinct banks.
 *
 * Twiddles: a thread loads the five powers W^(s 2^i) of its twiddle step s
 * from the FP64-derived W_32768 table and forms W^(s k) over the set bits of
 * k (at most four products), as the staged kernels do.
 */
#ifndef FAF_ENGINE_RS_H
#define FAF_ENGINE_RS_H
#define FAF_ENGINE_NO_PLAIN_KERNELS
#include "faf_engine.h"

namespace rs {

constexpr int NC = 32768;        /* complex points per signal */
constexpr int T = 1024;          /* threads per signal */
constexpr int P = 1056;          /* LDS row pitch in floats */
constexpr int LDS_BYTES = 32 * P * 4;

/* cos(pi j / 32), j = 0..16 */
FAF_DEV constexpr float cq(int j) {
    return j == 0 ? 1.0f
         : j == 1 ? 0.995184726672196886245f
         : j == 2 ? 0.980785280403230449126f
         : j == 3 ? 0.956940335732208864936f
         : j == 4 ? 0.923879532511286756128f
         : j == 5 ? 0.881921264348355029713f
         : j == 6 ? 0.831469612302545237079f
         : j == 7 ? 0.773010453362736960811f
         : j == 8 ? 0.707106781186547524401f
         : j == 9 ? 0.634393284163645498215f
         : j == 10 ? 0.555570233019602224743f
         : j == 11 ? 0.471396736825997648556f
         : j == 12 ? 0.382683432365089771728f
         : j == 13 ? 0.290284677254462367636f
         : j == 14 ? 0.195090322016128267848f
         : j == 15 ? 0.0980171403295606019942f
         : 0.0f;
}

/* exp(-+ 2 pi i k / 64): forward sign -, inverse + */
template <bool INV> FAF_DEV constexpr float2 w64(int k) {
    k &= 63;
    float c = 0.0f, s = 0.0f;                     /* cos, sin of pi k / 32 */
    if (k <= 16) { c = cq(k); s = cq(16 - k); }
    else if (k <= 32) { c = -cq(32 - k); s = cq(k - 16); }
    else if (k <= 48) { c = -cq(k - 32); s = -cq(48 - k); }
    else { c = cq(64 - k); s = -cq(k - 48); }
    return make_float2(c, INV ? s : -s);
}

template <bool INV> FAF_DEV void dft32(float2 *v) {
    float2 e[16], o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
    dft16<INV>(e);
    dft16<INV>(o);
#pragma unroll
    for (int k = 1; k < 16; ++k) o[k] = k == 8 ? rot<INV>(o[k]) : cmul(o[k], w64<INV>(2 * k));
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 16] = csub(e[k], o[k]);
    }
}

/* the powers W^(s 2^i), i < 5, of the table's root */
FAF_DEV void tw_bases(const float2 *tw, int s, float2 *b) {
#pragma unroll
    for (int i = 0; i < 5; ++i) b[i] = tw[(s << i) & (NC - 1)];
}

/* v[k] *= W^(s k) for k = 1..31, each power a product over the set bits of k */
FAF_DEV void tw_apply(float2 *v, const float2 *b) {
    float2 w[32];
#pragma unroll
    for (int k = 1; k < 32; ++k) {
        const int i = __builtin_ctz(k);
        w[k] = k == (1 << i) ? b[i] : cmul(w[k & (k - 1)], b[i]);
        v[k] = cmul(v[k], w[k]);
    }
}

/* Exchange through LDS: the thread writes v[r] at lds[wb + r WS] and then
 * holds lds[rb + r RS], one plane at a time.  The caller barriers between
 * the previous reads of lds and this call. */
template <int WS, int RS> FAF_DEV void xchg(float2 *v, float *lds, int wb, int rb) {
#pragma unroll
    for (int r = 0; r < 32; ++r) lds[wb + r * WS] = v[r].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r].x = lds[rb + r * RS];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) lds[wb + r * WS] = v[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r].y = lds[rb + r * RS];
}

/* In: v[m] = z[tid + 1024 m].  Out: v[j] = Z[tid + 1024 j].  b1 holds the
 * twiddle bases of step tid (loaded with the signal); those of pass 2 are
 * loaded here, while pass 1's values go through LDS.  Ends with LDS reads
 * outstanding in other waves: barrier before reusing lds. */
template <bool INV>
FAF_DEV void fft32768(float2 *v, float *lds, const float2 *tw, const float2 *b1, int tid) {
    dft32<INV>(v);
    tw_apply(v, b1);
    float2 b2[5];
    tw_bases(tw, 32 * (tid & 31), b2);
    xchg<P, 32>(v, lds, tid, (tid >> 5) * P + (tid & 31));
    dft32<INV>(v);
    tw_apply(v, b2);
    __syncthreads();
    xchg<P, 1>(v, lds, (tid >> 5) * 33 + (tid & 31), (tid >> 5) * P + (tid & 31) * 33);
    dft32<INV>(v);
}

}  // namespace rs

/* ------------------------------------------------------------------ */
/* 8192 points in one pass: 256 threads x 32 values, four workgroups    */
/* per CU (35 KB of LDS each).  8192 = 8 * 32 * 32 in natural order:    */
/*   pass 1  thread t: for n1 = t + 256 i (i < 4) a DFT-8 over n2 of     */
/*           z[n1 + 1024 n2], times W_8192^(n1 k2)                       */
/*   pass 2  thread (k2, n1a), tid = k2 + 8 n1a: DFT-32 over n1b of     */
/*           Y[n1a + 32 n1b][k2], times W_1024^(n1a ka)                 */
/*   pass 3  thread (k2, ka), tid = k2 + 8 ka: DFT-32 over n1a, giving  */
/*           X[k2 + 8 (ka + 32 kb)] = X[tid + 256 kb]                    */
/* The two LDS exchanges go one plane at a time with row pitches 1032  */
/* (k2 rows of Y) and 1096 / 33 (k2 blocks, ka rows): every wave access */
/* of either touches 64 distinct banks.                                 */
/* ------------------------------------------------------------------ */
namespace rs8 {

constexpr int NC = 8192;
constexpr int T = 256;
constexpr int P1 = 1032;
constexpr int P2 = 1096;
constexpr int Q = 33;
constexpr int LDS_BYTES = 8 * P2 * 4;

/* twiddle bases of pass 1: W^(n1 2^e), e < 3, for the four n1 of a thread */
template <class TP> FAF_DEV void tw1_bases(TP tw, int tid, float2 (*b)[3]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int n1 = tid + 256 * i;
#pragma unroll
        for (int e = 0; e < 3; ++e) b[i][e] = twld(tw, n1 << e);
    }
}

/* value r of pass 1 (r = i + 4 k2 after it) goes to Y row k2, column
 * n1 = tid + 256 i; value r of pass 2 (r = ka) to block k2, row ka */
struct WoY { static constexpr int at(int r) { return (r >> 2) * P1 + (r & 3) * 256; } };
struct RoY { static constexpr int at(int r) { return 32 * r; } };
struct WoE { static constexpr int at(int r) { return r * Q; } };
struct RoE { static constexpr int at(int r) { return r; } };

template <class WO, class RO> FAF_DEV void xchg(float2 *v, float *lds, int wb, int rb) {
#pragma unroll
    for (int r = 0; r < 32; ++r) lds[wb + WO::at(r)] = v[r].x;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r].x = lds[rb + RO::at(r)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) lds[wb + WO::at(r)] = v[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r].y = lds[rb + RO::at(r)];
}

/* In: v[m] = z[tid + 256 m], b1 from tw1_bases.  Out: v[j] = Z[tid + 256 j].
 * Ends with LDS reads outstanding in other waves: barrier before reusing lds. */
template <bool INV, class TP>
FAF_DEV void fft8192(float2 *v, float *lds, TP tw, const float2 (*b1)[3], int tid) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float2 g[8];
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) g[n2] = v[i + 4 * n2];
        dft<8, INV>(g);
        float2 w[8];
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            const int e = __builtin_ctz(k);
            w[k] = k == (1 << e) ? b1[i][e] : cmul(w[k & (k - 1)], b1[i][e]);
            g[k] = cmul(g[k], w[k]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) v[i + 4 * k] = g[k];
    }
    const int k2 = tid & 7, n1a = tid >> 3;
    float2 b2[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) b2[e] = twld(tw, (8 * n1a) << e);
    xchg<WoY, RoY>(v, lds, tid, k2 * P1 + n1a);
    rs::dft32<INV>(v);
    rs::tw_apply(v, b2);
    __syncthreads();
    xchg<WoE, RoE>(v, lds, k2 * P2 + n1a, k2 * P2 + n1a * Q);
    rs::dft32<INV>(v);
}

}  // namespace rs8

/* ------------------------------------------------------------------ */
/* kernels: persistent, one workgroup per CU, signals s = blockIdx.x,   */
/* blockIdx.x + gridDim.x, ...                                          */
/* ------------------------------------------------------------------ */

/* rfft N = 65536: x real (8-byte aligned rows), out N/2+1 bins, split or
 * interleaved.  post = forward Kiss table of 32768 (slots 1..16384). */
template <bool INTER_OUT>
__global__ __launch_bounds__(1024) void kr_r2c(const float2 *tw, const float2 *post,
                                               const float *x, long in_dist, float *out_re,
                                               float *out_im, long out_dist, long count,
                                               float scale) {
    using namespace rs;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;                         /* 16 x 1024 float2 for the post-pass */
    /* Kiss twiddle of bin k = tid + 1024 j is -i W_65536^k = pb * W_64^j */
    const float2 pb0 = threadIdx.x ? post[threadIdx.x] : make_float2(0.0f, -1.0f);
    for (long s = blockIdx.x; s < count; s += gridDim.x) {
        /* the thread index and pb are re-read every signal so that the
         * compiler does not keep every derived offset and twiddle alive (and
         * spilled) across the whole loop */
        int tid = (int)threadIdx.x;
        float2 pb = pb0;
        asm volatile("" : "+v"(tid), "+v"(pb.x), "+v"(pb.y));
        const int pt = (T - tid) & (T - 1);
        float2 v[32], b1[5];
        const float2 *z = (const float2 *)(x + s * in_dist);
#pragma unroll
        for (int m = 0; m < 32; ++m) v[m] = z[tid + T * m];
        tw_bases(tw, tid, b1);
        __syncthreads();                          /* the previous signal is out of LDS */
        fft32768<false>(v, lds, tw, b1, tid);
        float *ore = out_re + s * out_dist, *oim = out_im + s * out_dist;
        float2 *oi = (float2 *)out_re + s * out_dist;
        auto put = [&](int k, float2 xk) {
            if (INTER_OUT) oi[k] = cscale(xk, scale);
            else { ore[k] = xk.x * scale; oim[k] = xk.y * scale; }
        };
        /* Bin pairs (k, 32768 - k) with k = tid + 1024 j, j < 16: the mirror
         * value is thread pt's value 31 - j (thread 0: its own 32 - j), so the
         * upper halves go through LDS and each thread finishes the pairs of its
         * lower half, X[k] and X[32768 - k] from one Kiss butterfly. */
        __syncthreads();
#pragma unroll
        for (int j = 16; j < 32; ++j) q[(j - 16) * T + tid] = v[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            float2 xk, xm;
            if (tid == 0 && j == 0) {
                put(0, make_float2(v[0].x + v[0].y, 0.0f));
                put(NC, make_float2(v[0].x - v[0].y, 0.0f));
                const float2 h = q[0];            /* Z[16384], its own mirror */
                kiss_post(h, h, cmul(pb, w64<false>(16)), xk, xm);
                put(NC / 2, xk);
                continue;
            }
            kiss_post(v[j], q[tid ? (15 - j) * T + pt : (16 - j) * T],
                      cmul(pb, w64<false>(j)), xk, xm);
            put(tid + T * j, xk);
            put(tid ? pt + T * (31 - j) : T * (32 - j), xm);
            if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
    }
}

/* irfft N = 65536: in N/2+1 bins (split or interleaved), y real (8-byte
 * aligned rows).  pre = inverse Kiss table of 32768; tw = inverse W_32768. */
template <bool INTER_IN>
__global__ __launch_bounds__(1024) void kr_c2r(const float2 *tw, const float2 *pre,
                                               const float *in_re, const float *in_im,
                                               long in_dist, float *y, long out_dist,
                                               long count, float scale) {
    using namespace rs;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;                         /* 16 x 1024 float2 for the pre-pass */
    /* Kiss pre-twiddle of bin k = tid + 1024 j is i W_65536^-k = pb * conj(W_64^j) */
    const float2 pb0 = threadIdx.x ? pre[threadIdx.x] : make_float2(0.0f, 1.0f);
    for (long s = blockIdx.x; s < count; s += gridDim.x) {
        int tid = (int)threadIdx.x;               /* re-read per signal, as in kr_r2c */
        float2 pb = pb0;
        asm volatile("" : "+v"(tid), "+v"(pb.x), "+v"(pb.y));
        const int pt = (T - tid) & (T - 1);
        float2 v[32], b1[5];
        const float *ire = in_re + s * in_dist, *iim = in_im + s * in_dist;
        const float2 *ii = (const float2 *)in_re + s * in_dist;
        auto get = [&](int k) {
            return INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
        };
        tw_bases(tw, tid, b1);
        /* Bin pairs (k, 32768 - k) with k = tid + 1024 j, j < 16: one Kiss
         * butterfly gives F[k], this thread's value j, and F[32768 - k], thread
         * pt's value 31 - j (thread 0: its own 32 - j), which goes through LDS.
         * Every bin is read from HBM once. */
        __syncthreads();                          /* the previous signal is out of LDS */
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float2 a = get(tid + T * j);
            const float2 b = get(tid ? pt + T * (31 - j) : T * (32 - j));
            float2 fm;
            if (tid == 0 && j == 0) {
                v[0] = make_float2(a.x + b.x, a.x - b.x);
                const float2 h = get(NC / 2);     /* X[16384], its own mirror */
                kiss_pre(h, h, cmul(pb, w64<true>(16)), q[0], fm);
                continue;
            }
            kiss_pre(a, b, cmul(pb, w64<true>(j)), v[j], fm);
            q[tid ? (15 - j) * T + pt : (16 - j) * T] = fm;
            if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
#pragma unroll
        for (int j = 16; j < 32; ++j) v[j] = q[(j - 16) * T + tid];
        __syncthreads();
        fft32768<true>(v, lds, tw, b1, tid);
        float2 *yo = (float2 *)(y + s * out_dist);
#pragma unroll
        for (int j = 0; j < 32; ++j) nt_st(&yo[tid + T * j], cscale(v[j], scale));
    }
}

/* C2C 32768, split or interleaved in and out (interleaved dists in complex
 * elements, split dists in floats per plane) */
template <bool INV, bool INTER_IN, bool INTER_OUT>
__global__ __launch_bounds__(1024) void kr_c2c(const float2 *tw, const float *in_re,
                                               const float *in_im, long in_dist,
                                               float *out_re, float *out_im, long out_dist,
                                               long count, float scale) {
    using namespace rs;
    float *lds = (float *)faf_smem;
    for (long s = blockIdx.x; s < count; s += gridDim.x) {
        int tid = (int)threadIdx.x;               /* re-read per signal, as in kr_r2c */
        asm volatile("" : "+v"(tid));
        float2 v[32], b1[5];
        const float *ire = in_re + s * in_dist, *iim = in_im + s * in_dist;
        const float2 *ii = (const float2 *)in_re + s * in_dist;
#pragma unroll
        for (int m = 0; m < 32; ++m) {
            const int k = tid + T * m;
            v[m] = INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
        }
        tw_bases(tw, tid, b1);
        __syncthreads();
        fft32768<INV>(v, lds, tw, b1, tid);
        float *ore = out_re + s * out_dist, *oim = out_im + s * out_dist;
        float2 *oi = (float2 *)out_re + s * out_dist;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int k = tid + T * j;
            if (INTER_OUT) nt_st(&oi[k], cscale(v[j], scale));
            else { nt_st(&ore[k], v[j].x * scale); nt_st(&oim[k], v[j].y * scale); }
        }
    }
}

/* ------------------------------------------------------------------ */
/* 2048 points in one pass: one wave (64 threads x 32 values) per      */
/* signal.  2048 = 8 * 16 * 16 in natural order:                       */
/*   pass 1  thread t: for n1 = t + 64 i (i < 4) a DFT-8 over n2 of     */
/*           z[n1 + 256 n2], times W_2048^(n1 k2)                        */
/*   pass 2  thread (k2, a), tid = k2 + 8 a: for n1a = a, a + 8 a       */
/*           DFT-16 over n1b of Y[n1a + 16 n1b][k2], times W_256^(n1a ka) */
/*   pass 3  thread (k2, b), tid = k2 + 8 b: for ka = b, b + 8 a DFT-16  */
/*           over n1a, giving X[k2 + 8 (ka + 16 kb)] = X[tid + 64 j],    */
/*           j = (ka >> 3) + 2 kb                                       */
/* Exchanges one plane at a time: Y rows of pitch 280, then blocks of  */
/* 280 floats with ka rows of pitch 17; every wave access touches 64   */
/* distinct banks.  2240 floats of LDS per signal.                      */
/* ------------------------------------------------------------------ */
namespace rs2 {

constexpr int NC = 2048;
constexpr int T = 64;
constexpr int P1 = 280;
constexpr int P2 = 280;
constexpr int Q = 17;
constexpr int LDS_BYTES = 8 * P1 * 4;

struct WoY { static constexpr int at(int r) { return (r >> 2) * P1 + (r & 3) * 64; } };
struct RoY { static constexpr int at(int r) { return 8 * (r >> 4) + 16 * (r & 15); } };
struct WoE { static constexpr int at(int r) { return (r & 15) * Q + 8 * (r >> 4); } };
struct RoE { static constexpr int at(int r) { return 8 * Q * (r >> 4) + (r & 15); } };

/* twiddle bases of pass 1: W^(n1 2^e), e < 3, for the four n1 of a thread */
FAF_DEV void tw1_bases(const float2 *tw, int tid, float2 (*b)[3]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int e = 0; e < 3; ++e) b[i][e] = tw[(tid + 64 * i) << e];
    }
}

/* v[k] *= W^(s k), k = 1..15, from the bases W^(s 2^e), e < 4 */
FAF_DEV void tw_apply16(float2 *v, const float2 *b) {
    float2 w[16];
#pragma unroll
    for (int k = 1; k < 16; ++k) {
        const int e = __builtin_ctz(k);
        w[k] = k == (1 << e) ? b[e] : cmul(w[k & (k - 1)], b[e]);
        v[k] = cmul(v[k], w[k]);
    }
}

/* In: v[m] = z[tid + 64 m], b1 from tw1_bases.  Out: v[j] = Z[tid + 64 j].
 * Ends with LDS reads outstanding: barrier before reusing lds. */
template <bool INV>
FAF_DEV void fft2048(float2 *v, float *lds, const float2 *tw, const float2 (*b1)[3], int tid) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float2 g[8];
#pragma unroll
        for (int n2 = 0; n2 < 8; ++n2) g[n2] = v[i + 4 * n2];
        dft<8, INV>(g);
        float2 w[8];
#pragma unroll
        for (int k = 1; k < 8; ++k) {
            const int e = __builtin_ctz(k);
            w[k] = k == (1 << e) ? b1[i][e] : cmul(w[k & (k - 1)], b1[i][e]);
            g[k] = cmul(g[k], w[k]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) v[i + 4 * k] = g[k];
    }
    const int k2 = tid & 7, a = tid >> 3;
    float2 b2[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int e = 0; e < 4; ++e) b2[h][e] = tw[(8 * (a + 8 * h)) << e];
    }
    rs8::xchg<WoY, RoY>(v, lds, tid, k2 * P1 + a);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        dft16<INV>(v + 16 * h);
        tw_apply16(v + 16 * h, b2[h]);
    }
    __syncthreads();
    rs8::xchg<WoE, RoE>(v, lds, k2 * P2 + a, k2 * P2 + a * Q);
    float2 u[32];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        dft16<INV>(v + 16 * h);
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) u[h + 2 * kb] = v[16 * h + kb];
    }
#pragma unroll
    for (int j = 0; j < 32; ++j) v[j] = u[j];
}

}  // namespace rs2

/* rfft N = 16384 in one pass, one workgroup per signal: x real (8-byte
 * aligned rows), N/2+1 bins out.  The Kiss post-pass is kr_r2c's with 256
 * threads: bin k = tid + 256 j has twiddle -i W_16384^k = post[tid] W_64^j. */
template <bool INTER_OUT>
__global__ __launch_bounds__(256) void kr8_r2c(const float2 *tw, const float2 *post,
                                               const float *x, long in_dist, float *out_re,
                                               float *out_im, long out_dist, float scale) {
    using namespace rs8;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;                         /* 16 x 256 float2 for the post-pass */
    const int tid = (int)threadIdx.x;
    const long s = blockIdx.x;
    const int pt = (T - tid) & (T - 1);
    float2 v[32], b1[4][3];
    const float2 *z = (const float2 *)(x + s * in_dist);
#pragma unroll
    for (int m = 0; m < 32; ++m) v[m] = z[tid + T * m];
    tw1_bases(tw, tid, b1);
    const float2 pb = tid ? post[tid] : make_float2(0.0f, -1.0f);
    fft8192<false>(v, lds, tw, b1, tid);
    float *ore = out_re + s * out_dist, *oim = out_im + s * out_dist;
    float2 *oi = (float2 *)out_re + s * out_dist;
    auto put = [&](int k, float2 xk) {
        if (INTER_OUT) oi[k] = cscale(xk, scale);
        else { ore[k] = xk.x * scale; oim[k] = xk.y * scale; }
    };
    __syncthreads();
#pragma unroll
    for (int j = 16; j < 32; ++j) q[(j - 16) * T + tid] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        float2 xk, xm;
        if (tid == 0 && j == 0) {
            put(0, make_float2(v[0].x + v[0].y, 0.0f));
            put(NC, make_float2(v[0].x - v[0].y, 0.0f));
            const float2 h = q[0];                /* Z[4096], its own mirror */
            kiss_post(h, h, cmul(pb, rs::w64<false>(16)), xk, xm);
            put(NC / 2, xk);
            continue;
        }
        kiss_post(v[j], q[tid ? (15 - j) * T + pt : (16 - j) * T],
                  cmul(pb, rs::w64<false>(j)), xk, xm);
        put(tid + T * j, xk);
        put(tid ? pt + T * (31 - j) : T * (32 - j), xm);
    }
}

/* C2C 8192 in one pass, one workgroup per signal, split or interleaved in
 * and out (interleaved dists in complex elements, split in floats) */
template <bool INV, bool INTER_IN, bool INTER_OUT>
__global__ __launch_bounds__(256) void kr8_c2c(const float2 *tw, const float *in_re,
                                               const float *in_im, long in_dist,
                                               float *out_re, float *out_im, long out_dist,
                                               float scale) {
    using namespace rs8;
    float *lds = (float *)faf_smem;
    const int tid = (int)threadIdx.x;
    const long s = blockIdx.x;
    float2 v[32], b1[4][3];
    const float *ire = in_re + s * in_dist, *iim = in_im + s * in_dist;
    const float2 *ii = (const float2 *)in_re + s * in_dist;
#pragma unroll
    for (int m = 0; m < 32; ++m) {
        const int k = tid + T * m;
        v[m] = INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
    }
    tw1_bases(tw, tid, b1);
    fft8192<INV>(v, lds, tw, b1, tid);
    float *ore = out_re + s * out_dist, *oim = out_im + s * out_dist;
    float2 *oi = (float2 *)out_re + s * out_dist;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const int k = tid + T * j;
        if (INTER_OUT) oi[k] = cscale(v[j], scale);
        else { ore[k] = v[j].x * scale; oim[k] = v[j].y * scale; }
    }
}

/* irfft N = 16384 in one pass: N/2+1 bins in, y real (8-byte aligned
 * rows); the Kiss pre-pass is kr_c2r's with 256 threads */
template <bool INTER_IN>
__global__ __launch_bounds__(256) void kr8_c2r(const float2 *tw, const float2 *pre,
                                               const float *in_re, const float *in_im,
                                               long in_dist, float *y, long out_dist,
                                               float scale) {
    using namespace rs8;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;
    const int tid = (int)threadIdx.x;
    const long s = blockIdx.x;
    const int pt = (T - tid) & (T - 1);
    const float2 pb = tid ? pre[tid] : make_float2(0.0f, 1.0f);
    float2 v[32], b1[4][3];
    const float *ire = in_re + s * in_dist, *iim = in_im + s * in_dist;
    const float2 *ii = (const float2 *)in_re + s * in_dist;
    auto get = [&](int k) {
        return INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
    };
    tw1_bases(tw, tid, b1);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float2 a = get(tid + T * j);
        const float2 b = get(tid ? pt + T * (31 - j) : T * (32 - j));
        float2 fm;
        if (tid == 0 && j == 0) {
            v[0] = make_float2(a.x + b.x, a.x - b.x);
            const float2 h = get(NC / 2);         /* X[4096], its own mirror */
            kiss_pre(h, h, cmul(pb, rs::w64<true>(16)), q[0], fm);
            continue;
        }
        kiss_pre(a, b, cmul(pb, rs::w64<true>(j)), v[j], fm);
        q[tid ? (15 - j) * T + pt : (16 - j) * T] = fm;
    }
    __syncthreads();
#pragma unroll
    for (int j = 16; j < 32; ++j) v[j] = q[(j - 16) * T + tid];
    __syncthreads();
    fft8192<true>(v, lds, tw, b1, tid);
    float2 *yo = (float2 *)(y + s * out_dist);
#pragma unroll
    for (int j = 0; j < 32; ++j) nt_st(&yo[tid + T * j], cscale(v[j], scale));
}

/* irfft N = 4096 in one pass, one wave per signal */
template <bool INTER_IN>
__global__ __launch_bounds__(64) void kr2_c2r(const float2 *tw, const float2 *pre,
                                              const float *in_re, const float *in_im,
                                              long in_dist, float *y, long out_dist,
                                              float scale) {
    using namespace rs2;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;
    const int tid = (int)threadIdx.x;
    const long s = blockIdx.x;
    const int pt = (T - tid) & (T - 1);
    const float2 pb = tid ? pre[tid] : make_float2(0.0f, 1.0f);
    float2 v[32], b1[4][3];
    const float *ire = in_re + s * in_dist, *iim = in_im + s * in_dist;
    const float2 *ii = (const float2 *)in_re + s * in_dist;
    auto get = [&](int k) {
        return INTER_IN ? ii[k] : make_float2(ire[k], iim[k]);
    };
    tw1_bases(tw, tid, b1);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float2 a = get(tid + T * j);
        const float2 b = get(tid ? pt + T * (31 - j) : T * (32 - j));
        float2 fm;
        if (tid == 0 && j == 0) {
            v[0] = make_float2(a.x + b.x, a.x - b.x);
            const float2 h = get(NC / 2);         /* X[1024], its own mirror */
            kiss_pre(h, h, cmul(pb, rs::w64<true>(16)), q[0], fm);
            continue;
        }
        kiss_pre(a, b, cmul(pb, rs::w64<true>(j)), v[j], fm);
        q[tid ? (15 - j) * T + pt : (16 - j) * T] = fm;
    }
    __syncthreads();
#pragma unroll
    for (int j = 16; j < 32; ++j) v[j] = q[(j - 16) * T + tid];
    __syncthreads();
    fft2048<true>(v, lds, tw, b1, tid);
    float2 *yo = (float2 *)(y + s * out_dist);
#pragma unroll
    for (int j = 0; j < 32; ++j) nt_st(&yo[tid + T * j], cscale(v[j], scale));
}

#ifndef FAF_RS_CORE_ONLY   /* set by files that use only the cores above */
/* The single-MUL spectral pipeline at N = 4096 (the notch) in one pass, one
 * wave per signal: rfft, X[k] := op_mul(X[k]) on every bin, irfft.  After
 * the forward FFT each thread holds Z[k], k = tid + 64 j; for its pairs
 * (k, 2048 - k), j < 16, it forms X[k] and X[2048 - k] (Kiss post), applies
 * the MUL to both, and forms F[k] and F[2048 - k] (Kiss pre).  F[2048 - k]
 * goes back into the LDS slot its partner value came from, so the partner
 * finds it there; the inverse FFT then runs on F. */
__global__ __launch_bounds__(64) void kr2_pipe(const float2 *twf, const float2 *twi,
                                               const float2 *post, const float2 *pre,
                                               fafd_ops ops, const float *x, long in_dist,
                                               float *y, long out_dist, float scale) {
    using namespace rs2;
    float *lds = (float *)faf_smem;
    float2 *q = faf_smem;                         /* 16 x 64 float2 for the bin pairs */
    const int tid = (int)threadIdx.x;
    const long s = blockIdx.x;
    const int pt = (T - tid) & (T - 1);
    float2 v[32], b1[4][3];
    const float2 *z = (const float2 *)(x + s * in_dist);
#pragma unroll
    for (int m = 0; m < 32; ++m) v[m] = z[tid + T * m];
    tw1_bases(twf, tid, b1);
    const float2 pb = tid ? post[tid] : make_float2(0.0f, -1.0f);
    const float2 qb = tid ? pre[tid] : make_float2(0.0f, 1.0f);
    fft2048<false>(v, lds, twf, b1, tid);
    __syncthreads();
#pragma unroll
    for (int j = 16; j < 32; ++j) q[(j - 16) * T + tid] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        float2 xk, xm;
        if (tid == 0 && j == 0) {
            xk = op_mul(ops, 0, make_float2(v[0].x + v[0].y, 0.0f));
            xm = op_mul(ops, NC, make_float2(v[0].x - v[0].y, 0.0f));
            v[0] = make_float2(xk.x + xm.x, xk.x - xm.x);
            const float2 h = q[0];                /* Z[1024], its own mirror */
            float2 fm;
            kiss_post(h, h, cmul(pb, rs::w64<false>(16)), xk, xm);
            xk = op_mul(ops, NC / 2, xk);
            kiss_pre(xk, xk, cmul(qb, rs::w64<true>(16)), q[0], fm);
            continue;
        }
        const int slot = tid ? (15 - j) * T + pt : (16 - j) * T;
        const int k = tid + T * j, m = NC - k;
        kiss_post(v[j], q[slot], cmul(pb, rs::w64<false>(j)), xk, xm);
        xk = op_mul(ops, k, xk);
        xm = op_mul(ops, m, xm);
        kiss_pre(xk, xm, cmul(qb, rs::w64<true>(j)), v[j], q[slot]);
        /* keep the spectrum loads of later bins from being hoisted into
         * registers all at once */
        if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    tw1_bases(twi, tid, b1);
    __syncthreads();
#pragma unroll
    for (int j = 16; j < 32; ++j) v[j] = q[(j - 16) * T + tid];
    __syncthreads();
    fft2048<true>(v, lds, twi, b1, tid);
    float2 *yo = (float2 *)(y + s * out_dist);
#pragma unroll
    for (int j = 0; j < 32; ++j) nt_st(&yo[tid + T * j], cscale(v[j], scale));
}

#endif

#endif
