/*
 * faf_engine.h -- device code of the transform execution path (gfx950).
 *
 * Every kernel here is built from one piece: a self-sorting Stockham FFT
 * whose passes run over LDS (see faf_dev.h for the index scheme).  A pass
 * reads its inputs through a "source" accessor and writes through a
 * "sink" accessor, so the first pass can read HBM directly (coalesced,
 * natural order) and the last pass can write HBM directly (coalesced,
 * natural order); only the passes in between touch LDS.  R2C/C2R packing,
 * window, spectral ops, chirps and Rader gathers are accessors too, so
 * they fuse into the passes instead of costing extra HBM round trips.
 *
 * Arithmetic is FP32.  Twiddles and chirps are FP32 tables rounded from
 * FP64 values computed on the host with exact integer phase indices.
 */
#ifndef FAF_ENGINE_H
#define FAF_ENGINE_H

#include "faf_dev.h"

#define FAF_DEV __device__ __forceinline__

/* ------------------------------------------------------------------ */
/* complex helpers                                                     */
/* ------------------------------------------------------------------ */

FAF_DEV float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
FAF_DEV float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
/* Twiddle loads: from a generic pointer, or from a pointer typed global.
 * Kernels that read their table pointers out of device memory (the ragged
 * launch's plan records) pass the second kind: a generic pointer of unknown
 * origin becomes flat loads, which also hold up every LDS wait. */
typedef float faf_f2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const faf_f2v faf_gtw;
typedef __attribute__((address_space(1))) faf_f2v faf_gf2;
typedef __attribute__((address_space(1))) const float faf_gcf;
typedef __attribute__((address_space(1))) const int32_t faf_gi32;
FAF_DEV float2 twld(const float2 *p, int i) { return p[i]; }
FAF_DEV float2 twld(const faf_gtw *p, int i) {
    const faf_f2v v = p[i];
    return make_float2(v.x, v.y);
}

FAF_DEV float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
FAF_DEV float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
FAF_DEV float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

/* multiply by -i (forward) or +i (inverse) */
template <bool INV> FAF_DEV float2 rot(float2 v) {
    return INV ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
}

/* ------------------------------------------------------------------ */
/* small DFTs, natural order in and out, e^{-2 pi i/R} forward          */
/* ------------------------------------------------------------------ */

template <bool INV> FAF_DEV void dft2(float2 *v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV> FAF_DEV void dft3(float2 *v) {
    const float c = 0.86602540378443864676f;
    float2 s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    float2 m = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    float2 r = cscale(rot<INV>(d), c);
    v[0] = cadd(v[0], s);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <bool INV> FAF_DEV void dft4(float2 *v) {
    float2 s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    float2 s1 = cadd(v[1], v[3]), d1 = rot<INV>(csub(v[1], v[3]));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
}

template <bool INV> FAF_DEV void dft5(float2 *v) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 x0 = v[0];
    float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    float2 p2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    float2 q1 = rot<INV>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    float2 q2 = rot<INV>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
}

template <bool INV> FAF_DEV void dft7(float2 *v) {
    const float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f,
                c3 = -0.90096886790241912624f;
    const float s1 = 0.78183148246802980871f, s2 = 0.97492791218182360702f,
                s3 = 0.43388373911755812048f;
    float2 a1 = cadd(v[1], v[6]), b1 = csub(v[1], v[6]);
    float2 a2 = cadd(v[2], v[5]), b2 = csub(v[2], v[5]);
    float2 a3 = cadd(v[3], v[4]), b3 = csub(v[3], v[4]);
    float2 x0 = v[0];
    float2 p1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x,
                            x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y);
    float2 p2 = make_float2(x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x,
                            x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y);
    float2 p3 = make_float2(x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x,
                            x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y);
    /* sin(2 pi p k / 7) for (p,k): k=1: s1 s2 s3; k=2: s2 -s3 -s1; k=3: s3 -s1 s2 */
    float2 q1 = rot<INV>(make_float2(s1 * b1.x + s2 * b2.x + s3 * b3.x,
                                     s1 * b1.y + s2 * b2.y + s3 * b3.y));
    float2 q2 = rot<INV>(make_float2(s2 * b1.x - s3 * b2.x - s1 * b3.x,
                                     s2 * b1.y - s3 * b2.y - s1 * b3.y));
    float2 q3 = rot<INV>(make_float2(s3 * b1.x - s1 * b2.x + s2 * b3.x,
                                     s3 * b1.y - s1 * b2.y + s2 * b3.y));
    v[0] = make_float2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
    v[1] = cadd(p1, q1); v[6] = csub(p1, q1);
    v[2] = cadd(p2, q2); v[5] = csub(p2, q2);
    v[3] = cadd(p3, q3); v[4] = csub(p3, q3);
}

template <bool INV> FAF_DEV void dft8(float2 *v) {
    const float h = 0.70710678118654752440f;
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    float2 r1 = rot<INV>(o[1]), r3 = rot<INV>(o[3]);
    o[1] = make_float2(h * (o[1].x + r1.x), h * (o[1].y + r1.y));
    o[2] = rot<INV>(o[2]);
    o[3] = make_float2(h * (r3.x - o[3].x), h * (r3.y - o[3].y));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <bool INV> FAF_DEV void dft16(float2 *v) {
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
    const float h = 0.70710678118654752440f;
    const float sg = INV ? 1.0f : -1.0f;
    float2 e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
    dft8<INV>(e);
    dft8<INV>(o);
    /* o[k] *= W16^k */
    o[1] = cmul(o[1], make_float2(c1, sg * s1));
    {
        float2 r = rot<INV>(o[2]);
        o[2] = make_float2(h * (o[2].x + r.x), h * (o[2].y + r.y));
    }
    o[3] = cmul(o[3], make_float2(s1, sg * c1));
    o[4] = rot<INV>(o[4]);
    o[5] = cmul(o[5], make_float2(-s1, sg * c1));
    {
        float2 r = rot<INV>(o[6]);
        o[6] = make_float2(h * (r.x - o[6].x), h * (r.y - o[6].y));
    }
    o[7] = cmul(o[7], make_float2(-c1, sg * s1));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 8] = csub(e[k], o[k]);
    }
}

template <int R, bool INV> FAF_DEV void dft(float2 *v);

/* Composite R = A B with gcd(A, B) = 1 by the prime-factor (Good-Thomas)
 * map: input n = (B n1 + A n2) mod R, A-point DFTs over n1, B-point DFTs
 * over n2, output k = the CRT of (k1 mod A, k2 mod B).  W_R^{nk} splits
 * into W_A^{n1 k1} W_B^{n2 k2} exactly, so no twiddle is multiplied: every
 * index is a compile-time constant once unrolled, and the maps are register
 * renames. */
template <int A, int B> struct Pfa {
    static constexpr int R = A * B;
    static constexpr int inv_mod(int a, int m) {
        for (int x = 1; x < m; ++x)
            if ((a * x) % m == 1) return x;
        return 1;
    }
    /* k = k1 B (B^-1 mod A) + k2 A (A^-1 mod B)  (mod R) */
    static constexpr int EA = B * inv_mod(B % A, A), EB = A * inv_mod(A % B, B);
};

template <int A, int B, bool INV> FAF_DEV void dft_pfa(float2 *v) {
    using P = Pfa<A, B>;
    float2 m[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) m[n2][n1] = v[(B * n1 + A * n2) % P::R];
        dft<A, INV>(m[n2]);
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        float2 c[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) c[n2] = m[n2][k1];
        dft<B, INV>(c);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[(k1 * P::EA + k2 * P::EB) % P::R] = c[k2];
    }
}

template <int R, bool INV> FAF_DEV void dft(float2 *v) {
    if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else if constexpr (R == 6) dft_pfa<2, 3, INV>(v);
    else if constexpr (R == 7) dft7<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 10) dft_pfa<2, 5, INV>(v);
    else if constexpr (R == 12) dft_pfa<4, 3, INV>(v);
    else if constexpr (R == 14) dft_pfa<2, 7, INV>(v);
    else if constexpr (R == 15) dft_pfa<3, 5, INV>(v);
    else if constexpr (R == 16) dft16<INV>(v);
    else if constexpr (R == 20) dft_pfa<4, 5, INV>(v);
    else if constexpr (R == 21) dft_pfa<3, 7, INV>(v);
    else if constexpr (R == 24) dft_pfa<8, 3, INV>(v);
    else if constexpr (R == 28) dft_pfa<4, 7, INV>(v);
    else if constexpr (R == 30) dft_pfa<2, 15, INV>(v);
    else if constexpr (R == 40) dft_pfa<8, 5, INV>(v);
    else if constexpr (R == 48) dft_pfa<16, 3, INV>(v);
    else static_assert(R == 16, "no codelet of this radix");
}

/* ------------------------------------------------------------------ */
/* accessors: ld(i) for sources, st(i, v) for sinks                     */
/* ------------------------------------------------------------------ */

/* LDS signal buffer of float2 with one pad slot every 16 to spread banks */
FAF_DEV int lpad(int i) { return i + (i >> 4); }

struct Lds {
    float2 *p;
    static constexpr bool is_lds = true;
    FAF_DEV float2 ld(int i) const { return p[lpad(i)]; }
    FAF_DEV void st(int i, float2 v) const { p[lpad(i)] = v; }
};

struct GSplitIn {            /* split planes re[], im[] */
    const float *re, *im;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const { return make_float2(re[i], im[i]); }
};

struct GInterIn {            /* interleaved complex */
    const float2 *p;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const { return p[i]; }
};

struct GSplitOut {
    float *re, *im;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const { re[i] = v.x * s; im[i] = v.y * s; }
};

struct GInterOut {
    float2 *p;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const { p[i] = cscale(v, s); }
};

/* Sinks with non-temporal (streaming) stores, for outputs a kernel writes
 * once in whole lines: the batched C2C (C2C 4096 x 256: 6.4 -> 5.95 us),
 * irfft 1024 (5.33 -> 5.55 TB/s) and the CWT rows gain; the ragged launch,
 * whose rows end on partial lines, lost 10% with them and keeps plain
 * stores.  Same values either way. */
/* One streaming (non-temporal) store of a float / float2.  The
 * single-pass kernels whose output rows go out in whole lines use them
 * (A/B, same bits: irfft 4096 5.43 -> 5.72 TB/s, irfft 16384 5.34 -> 5.49,
 * irfft 65536 3.79 -> 4.02, C2C 32768 4.56 -> 4.70, notch +1%); the rfft
 * single passes store mirrored bin pairs on partial lines and lost 8-14%
 * with them, C2C 8192 1.5%: those keep plain stores. */
FAF_DEV void nt_st(float *p, float v) { __builtin_nontemporal_store(v, p); }
FAF_DEV void nt_st(float2 *p, float2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

struct GSplitOutNT {
    float *re, *im;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const {
        __builtin_nontemporal_store(v.x * s, &re[i]);
        __builtin_nontemporal_store(v.y * s, &im[i]);
    }
};

struct GInterOutNT {
    float2 *p;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const {
        __builtin_nontemporal_store(v.x * s, &p[i].x);
        __builtin_nontemporal_store(v.y * s, &p[i].y);
    }
};

struct GRealOnlyOut {        /* keep only the real part (real CWT rows) */
    float *p;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const { p[i] = v.x * s; }
};

/* R2C packing z[i] = x[2i] + i x[2i+1] of a real row */
struct GPairIn {
    const float *x;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const { return make_float2(x[2 * i], x[2 * i + 1]); }
};

/* GPairIn with streaming (non-temporal) loads: the headline rfft reads its
 * 1 GiB of input once (5.33-5.35 -> 5.40-5.45 TB/s with the streaming row
 * stores; same bits).  Elsewhere streaming loads measured 1-8% slower
 * (irfft, C2C) and are not used. */
struct GPairInNT {
    const float *x;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const f2 v = __builtin_nontemporal_load((const f2 *)(x + 2 * i));
        return make_float2(v.x, v.y);
    }
};

struct GPairOut {            /* C2R unpacking y[2i] = Re z, y[2i+1] = Im z */
    float *y;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const { y[2 * i] = v.x * s; y[2 * i + 1] = v.y * s; }
};

struct GPairOutNT {          /* GPairOut with streaming stores */
    float *y;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const {
        __builtin_nontemporal_store(v.x * s, &y[2 * i]);
        __builtin_nontemporal_store(v.y * s, &y[2 * i + 1]);
    }
};

/* ------------------------------------------------------------------ */
/* one radix-R Stockham pass                                           */
/* ------------------------------------------------------------------ */

/* Twiddles of a radix-R butterfly with step s are W^(t s), t = 1 .. R-1.
 * A butterfly loads only W^(s 2^i) (i < K, 2^(K-1) <= R-1) from the
 * FP64-built table and forms W^(t s) as the product of the loaded powers
 * for the set bits of t (at most three products for t < 16).  The few
 * loads go out with the pass's data reads, before its barrier. */
template <int R> struct TwBits {
    static constexpr int K = R <= 2 ? 1 : R <= 4 ? 2 : R <= 8 ? 3 : R <= 16 ? 4 : R <= 32 ? 5 : 6;
};

template <int R> FAF_DEV void tw_expand(const float2 *b, float2 *w) {
#pragma unroll
    for (int t = 1; t < R; ++t) {
        const int i = __builtin_ctz(t);
        w[t] = t == (1 << i) ? b[i] : cmul(w[t & (t - 1)], b[i]);
    }
}

template <int R, bool INV, class S, class D>
FAF_DEV void stage(const S &src, const D &dst, int n, int active, int lt, int T,
                   const fafd_stage &st, const faf_gtw *__restrict__ tw, bool sync) {
    constexpr int B = FAFD_MAXE / R, K = TwBits<R>::K;
    const int nb = n / R;
    const int nbe = active ? nb : 0;
    float2 v[B][R];
    float2 tb[B][K];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = lt + b * T;
        if (j < nbe) {
#pragma unroll
            for (int t = 0; t < R; ++t) v[b][t] = src.ld(j + t * nb);
            if (st.ns > 1) {
                const int k = j - (int)__umulhi((unsigned)j, st.ns_magic) * st.ns;
                const int step = k * st.tw_stride;
#pragma unroll
                for (int i = 0; i < K; ++i) tb[b][i] = twld(tw, step << i);
            }
        }
    }
    if (sync) __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = lt + b * T;
        if (j < nbe) {
            int k = 0;
            if (st.ns > 1) {
                k = j - (int)__umulhi((unsigned)j, st.ns_magic) * st.ns;
                float2 w[R];
                tw_expand<R>(tb[b], w);
#pragma unroll
                for (int t = 1; t < R; ++t) v[b][t] = cmul(v[b][t], w[t]);
            }
            dft<R, INV>(v[b]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) dst.st(base + t * st.ns, v[b][t]);
        }
    }
}

template <bool INV, class S, class D>
FAF_DEV void stage_any(const S &src, const D &dst, int n, int active, int lt, int T,
                       const fafd_stage &st, const faf_gtw *tw, bool sync) {
    switch (st.radix) {
    case 2: stage<2, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    case 3: stage<3, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    case 4: stage<4, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    case 5: stage<5, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    case 7: stage<7, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    case 8: stage<8, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    default: stage<16, INV>(src, dst, n, active, lt, T, st, tw, sync); break;
    }
}

/*
 * Full FFT: src -> (LDS passes) -> dst.  All threads of the workgroup must
 * call it (barriers inside); `active` is false for slots past the batch.
 * On return the sink has been written; if the sink is LDS the caller
 * barriers before reading it.
 */
template <bool INV, class S, class D>
FAF_DEV void fft_run(const fafd_fft &P, const S &src, const Lds &lds, const D &dst,
                     int lt, int active) {
    /* the plan may live in device memory (its pointers then have no known
     * address space): the table is device memory, read by global loads */
    const faf_gtw *tw = (const faf_gtw *)P.tw;
    const int ns = P.nstages, n = P.n, T = P.threads;
    if (ns == 1) {
        stage_any<INV>(src, dst, n, active, lt, T, P.st[0], tw, true);
        return;
    }
    stage_any<INV>(src, lds, n, active, lt, T, P.st[0], tw, S::is_lds);
    __syncthreads();
    for (int s = 1; s < ns - 1; ++s) {
        stage_any<INV>(lds, lds, n, active, lt, T, P.st[s], tw, true);
        __syncthreads();
    }
    stage_any<INV>(lds, dst, n, active, lt, T, P.st[ns - 1], tw, D::is_lds);
}

/* slot of this thread in a workgroup holding spb signals of T threads */
struct Slot {
    int sid, lt, sig, active;
};
FAF_DEV Slot slot_of(int T, int spb, long first, long count) {
    Slot s;
    const int tid = (int)threadIdx.x;
    s.sid = tid / T;
    s.lt = tid - s.sid * T;
    const long g = first + s.sid;
    s.sig = (int)g;
    s.active = (s.sid < spb) && (g < count);
    return s;
}

extern __shared__ float2 faf_smem[];

/* ------------------------------------------------------------------ */
/* batched C2C, LDS-resident                                           */
/* ------------------------------------------------------------------ */

template <bool INV, bool INTER, int MT>
__global__ __launch_bounds__(MT) void k_c2c(const fafd_fft *Pd, int spb, const float *in_re, const float *in_im,
                      long in_dist, float *out_re, float *out_im, long out_dist,
                      long count, float scale) {
    const fafd_fft &P = *Pd;
    Slot sl = slot_of(P.threads, spb, (long)blockIdx.x * spb, count);
    Lds lds{faf_smem + sl.sid * P.plane};
    const long sig = sl.active ? sl.sig : 0;
    if (INTER) {
        GInterIn src{(const float2 *)in_re + sig * in_dist};
        GInterOut dst{(float2 *)out_re + sig * out_dist, scale};
        fft_run<INV>(P, src, lds, dst, sl.lt, sl.active);
    } else {
        GSplitIn src{in_re + sig * in_dist, in_im + sig * in_dist};
        GSplitOut dst{out_re + sig * out_dist, out_im + sig * out_dist, scale};
        fft_run<INV>(P, src, lds, dst, sl.lt, sl.active);
    }
}

/* ------------------------------------------------------------------ */
/* R2C / C2R with the Kiss post-pass / pre-pass fused                   */
/* (reference src/faf_rfft.c:61-89 and :127-149)                        */
/* ------------------------------------------------------------------ */

/* X[k], X[nc-k] from Z[k], Z[nc-k]; tw = exp(-i pi (k/nc + 1/2)) */
FAF_DEV void kiss_post(float2 a, float2 b_raw, float2 tw, float2 &xk, float2 &xm) {
    float2 b = cconj(b_raw);
    float2 f1 = cadd(a, b), f2 = csub(a, b);
    float2 w = cmul(f2, tw);
    xk = make_float2(0.5f * (f1.x + w.x), 0.5f * (f1.y + w.y));
    xm = make_float2(0.5f * (f1.x - w.x), 0.5f * (w.y - f1.y));
}

/* F[k], F[nc-k] from X[k], X[nc-k]; tw = exp(+i pi (k/nc + 1/2)) */
FAF_DEV void kiss_pre(float2 a, float2 b_raw, float2 tw, float2 &fk, float2 &fm) {
    float2 b = cconj(b_raw);
    float2 e = cadd(a, b), d = csub(a, b);
    float2 o = cmul(d, tw);
    fk = cadd(e, o);
    fm = make_float2(e.x - o.x, o.y - e.y);
}

template <bool INTER_OUT, int MT>
__global__ __launch_bounds__(MT) void k_r2c(const fafd_fft *Pd, int spb, const float2 *post, const float *x,
                      long in_dist, float *out_re, float *out_im, long out_dist, long count,
                      float scale) {
    const fafd_fft &P = *Pd;
    Slot sl = slot_of(P.threads, spb, (long)blockIdx.x * spb, count);
    Lds lds{faf_smem + sl.sid * P.plane};
    const long sig = sl.active ? sl.sig : 0;
    GPairIn src{x + sig * in_dist};
    fft_run<false>(P, src, lds, lds, sl.lt, sl.active);
    __syncthreads();
    if (!sl.active) return;
    const int nc = P.n;
    float *ore = out_re + sig * out_dist;
    float *oim = out_im + sig * out_dist;
    float2 *oi = (float2 *)out_re + sig * out_dist;
    for (int k = sl.lt; k <= nc / 2; k += P.threads) {
        float2 xk, xm;
        int m = nc - k;
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

template <bool INTER_IN, int MT>
__global__ __launch_bounds__(MT) void k_c2r(const fafd_fft *Pd, int spb, const float2 *pre, const float *in_re,
                      const float *in_im, long in_dist, float *y, long out_dist, long count,
                      float scale) {
    const fafd_fft &P = *Pd;
    Slot sl = slot_of(P.threads, spb, (long)blockIdx.x * spb, count);
    Lds lds{faf_smem + sl.sid * P.plane};
    const long sig = sl.active ? sl.sig : 0;
    const int nc = P.n;
    if (sl.active) {
        const float *ire = in_re + sig * in_dist;
        const float *iim = in_im + sig * in_dist;
        const float2 *ii = (const float2 *)in_re + sig * in_dist;
        for (int k = sl.lt; k <= nc / 2; k += P.threads) {
            int m = nc - k;
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
    fft_run<true>(P, lds, lds, dst, sl.lt, sl.active);
}

/* ------------------------------------------------------------------ */
/* fused rfft -> spectral ops -> irfft  (reference src/faf_pipeline.c,  */
/* op semantics src/chirp.c:421-471, src/faf_core.c:150-166)            */
/* ------------------------------------------------------------------ */

/* the op list that is exactly one MUL by the bound spectrum (the notch and
 * filter pipelines): apply_ops without the per-bin interpretation */
FAF_DEV float2 op_mul(const fafd_ops &ops, int k, float2 v) {
    /* a bin index is an int: the bound compares in 32 bits (per bin, twice) */
    const int m = ops.h_bins < 0x7fffffffL ? (int)ops.h_bins : 0x7fffffff;
    if (k < m) {
        if (k == 0 || k == m - 1) v = make_float2(v.x * ops.h_re[k], 0.0f);
        else v = cmul(v, make_float2(ops.h_re[k], ops.h_im[k]));
    }
    return v;
}

FAF_DEV float2 apply_ops(const fafd_ops &ops, int k, float2 v) {
    for (int o = 0; o < ops.nops; ++o) {
        const int c = ops.code[o];
        if (c == FAFD_OP_CONJ) {
            v.y = -v.y;
        } else if (c == FAFD_OP_MUL) {
            const long m = ops.h_bins;
            if (k < m) {
                if (k == 0 || k == m - 1)
                    v = make_float2(v.x * ops.h_re[k], 0.0f);
                else
                    v = cmul(v, make_float2(ops.h_re[k], ops.h_im[k]));
            }
        } else if (c == FAFD_OP_BANDPASS) {
            if ((unsigned)k < ops.a1[o] || (unsigned)k > ops.a2[o]) v = make_float2(0.0f, 0.0f);
        }
    }
    return v;
}

#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_pipe(const fafd_fft *Pfd,
                                              const fafd_fft *Pid, int spb,
                                              const float2 *post, const float2 *pre,
                                              const fafd_ops *opsd, const float *x, long in_dist,
                       float *y, long out_dist, long count, float scale) {
    const fafd_fft &Pf = *Pfd;
    const fafd_fft &Pi = *Pid;
    const fafd_ops &ops = *opsd;
    Slot sl = slot_of(Pf.threads, spb, (long)blockIdx.x * spb, count);
    Lds lds{faf_smem + sl.sid * Pf.plane};
    const long sig = sl.active ? sl.sig : 0;
    GPairIn src{x + sig * in_dist};
    fft_run<false>(Pf, src, lds, lds, sl.lt, sl.active);
    __syncthreads();
    const int nc = Pf.n;
    if (sl.active) {
        /* each thread owns the bin pair (k, nc-k): post-pass, ops, pre-pass */
        for (int k = sl.lt; k <= nc / 2; k += Pf.threads) {
            const int m = nc - k;
            float2 xk, xm;
            if (k == 0) {
                float2 z = lds.ld(0);
                xk = apply_ops(ops, 0, make_float2(z.x + z.y, 0.0f));
                xm = apply_ops(ops, nc, make_float2(z.x - z.y, 0.0f));
                lds.st(0, make_float2(xk.x + xm.x, xk.x - xm.x));
            } else {
                kiss_post(lds.ld(k), lds.ld(m), post[k], xk, xm);
                xk = apply_ops(ops, k, xk);
                if (m != k) xm = apply_ops(ops, m, xm);
                else xm = xk;
                float2 fk, fm;
                kiss_pre(xk, xm, pre[k], fk, fm);
                lds.st(k, fk);
                if (m != k) lds.st(m, fm);
            }
        }
    }
    __syncthreads();
    GPairOut dst{y + sig * out_dist, scale};
    fft_run<true>(Pi, lds, lds, dst, sl.lt, sl.active);
}
#endif


/* ------------------------------------------------------------------ */
/* framed STFT: frame + symmetric Hann + rfft in one pass               */
/* ------------------------------------------------------------------ */

struct GFrameIn {            /* windowed frame packed as z[i] = (w x)[2i] + i (w x)[2i+1] */
    const float *x;          /* frame start */
    const float *w;
    long avail;              /* samples left in the record from the frame start */
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const {
        const long a = 2L * i, b = a + 1;
        if (b < avail)            /* both samples in the record: paired loads */
            return make_float2(x[a] * w[a], x[b] * w[b]);
        float u = a < avail ? x[a] * w[a] : 0.0f;
        return make_float2(u, 0.0f);
    }
};

/* the same when every frame of the launch lies inside its record (record
 * length >= window): no bounds test, and the pairs load as 8-byte words */
struct GFrameFull {
    const float *x;
    const float *w;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const {
        return make_float2(x[2 * i] * w[2 * i], x[2 * i + 1] * w[2 * i + 1]);
    }
};

#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_stft(const fafd_fft *Pd, int spb, const float2 *post, const float *win,
                       const float *x, long rec_len, long rec_dist, long hop, long frames,
                       float *out_re, float *out_im, long frame_dist, long rec_out_dist,
                       long count) {
    const fafd_fft &P = *Pd;
    Slot sl = slot_of(P.threads, spb, (long)blockIdx.x * spb, count);
    Lds lds{faf_smem + sl.sid * P.plane};
    const long sig = sl.active ? sl.sig : 0;
    const long rec = sig / frames, f = sig - rec * frames;
    GFrameIn src{x + rec * rec_dist + f * hop, win, rec_len - f * hop};
    fft_run<false>(P, src, lds, lds, sl.lt, sl.active);
    __syncthreads();
    if (!sl.active) return;
    const int nc = P.n;
    float *ore = out_re + rec * rec_out_dist + f * frame_dist;
    float *oim = out_im + rec * rec_out_dist + f * frame_dist;
    for (int k = sl.lt; k <= nc / 2; k += P.threads) {
        float2 xk, xm;
        const int m = nc - k;
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
#endif


/* ------------------------------------------------------------------ */
/* CWT bank: one rfft, then per row a one-sided multiply and a complex  */
/* inverse FFT of length n (analytic rows, Re z = reference real row)   */
/* ------------------------------------------------------------------ */

struct OneSidedIn {          /* Z[k] from the packed spectrum X and a real filter row */
    Lds X;
    const float *psi;
    int nc;                  /* n/2 */
    static constexpr bool is_lds = true;
    FAF_DEV float2 ld(int k) const {
        if (k > nc) return make_float2(0.0f, 0.0f);
        float2 v = X.ld(k);
        const float p = psi[k];
        if (k == 0 || k == nc) return make_float2(v.x * p, 0.0f);
        return make_float2(2.0f * v.x * p, 2.0f * v.y * p);
    }
};

template <bool REAL_OUT, int MT>
__global__ __launch_bounds__(MT) void k_cwt(const fafd_fft *Phd,
                                              const fafd_fft *Pnd, const float2 *post, const float *psi,
                      long psi_stride, int rows, int rows_per_wg, const float *x,
                      long in_dist, float *out, long out_dist, long count, float scale) {
    const fafd_fft &Ph = *Phd;
    const fafd_fft &Pn = *Pnd;
    /* one segment per workgroup; blockIdx.x = seg * groups + group */
    const int groups = (rows + rows_per_wg - 1) / rows_per_wg;
    const long seg = blockIdx.x / groups;
    const int g = (int)(blockIdx.x - seg * groups);
    const int tid = (int)threadIdx.x;
    const int nc = Ph.n, n = Pn.n;
    Lds X{faf_smem};
    Lds W{faf_smem + Ph.plane};
    const int active_h = (tid < Ph.threads) && seg < count;
    const int active_n = (tid < Pn.threads) && seg < count;
    GPairIn src{x + (seg < count ? seg : 0) * in_dist};
    fft_run<false>(Ph, src, X, X, tid < Ph.threads ? tid : 0, active_h);
    __syncthreads();
    if (active_h) {
        for (int k = tid; k <= nc / 2; k += Ph.threads) {
            const int m = nc - k;
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
        OneSidedIn osrc{X, psi + (long)r * psi_stride, nc};
        const long o = (seg < count ? seg : 0) * out_dist + (long)r * n;
        if (REAL_OUT) {
            GRealOnlyOut dst{out + o, scale};
            fft_run<true>(Pn, osrc, W, dst, tid < Pn.threads ? tid : 0, active_n);
        } else {
            GInterOut dst{(float2 *)out + o, scale};
            fft_run<true>(Pn, osrc, W, dst, tid < Pn.threads ? tid : 0, active_n);
        }
        __syncthreads();
    }
}

/* dual-frame inverse CWT (reference src/faf_cwt.c:934-981): per segment
 * acc = sum_r rfft(W_r) * dual_r, then y = irfft(acc), in one workgroup.
 * X holds the packed row spectrum, A the accumulator (nc+1 bins). */
#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_icwt(const fafd_fft *Pfd,
                                              const fafd_fft *Pid, const float2 *post, const float2 *pre,
                       const float *dual, long dual_stride, int rows, const float *w,
                       long in_dist, float *y, long out_dist, long count, float scale) {
    const fafd_fft &Pf = *Pfd;
    const fafd_fft &Pi = *Pid;
    const long seg = blockIdx.x;
    const int tid = (int)threadIdx.x, T = Pf.threads, nc = Pf.n;
    const int active = seg < count;
    Lds X{faf_smem};
    Lds A{faf_smem + Pf.plane};
    for (int k = tid; k <= nc; k += T) A.st(k, make_float2(0.0f, 0.0f));
    const float *wseg = w + (active ? seg : 0) * in_dist;
    for (int r = 0; r < rows; ++r) {
        GPairIn src{wseg + (long)r * 2 * nc};
        fft_run<false>(Pf, src, X, X, tid, active);
        __syncthreads();
        const float *d = dual + (long)r * dual_stride;
        if (active)
            for (int k = tid; k <= nc / 2; k += T) {
                const int m = nc - k;
                float2 xk, xm;
                if (k == 0) {
                    float2 z = X.ld(0);
                    xk = make_float2(z.x + z.y, 0.0f);
                    xm = make_float2(z.x - z.y, 0.0f);
                } else {
                    kiss_post(X.ld(k), X.ld(m), post[k], xk, xm);
                }
                A.st(k, cadd(A.ld(k), cscale(xk, d[k])));
                if (m != k) A.st(m, cadd(A.ld(m), cscale(xm, d[m])));
            }
        __syncthreads();
    }
    if (active)
        for (int k = tid; k <= nc / 2; k += T) {
            const int m = nc - k;
            float2 a = A.ld(k), b = A.ld(m);
            if (k == 0) {
                X.st(0, make_float2(a.x + b.x, a.x - b.x));   /* imag of DC/Nyquist dropped */
            } else {
                float2 fk, fm;
                kiss_pre(a, b, pre[k], fk, fm);
                X.st(k, fk);
                if (m != k) X.st(m, fm);
            }
        }
    __syncthreads();
    GPairOut dst{y + (active ? seg : 0) * out_dist, scale};
    fft_run<true>(Pi, X, X, dst, tid, active);
}
#endif


/* ------------------------------------------------------------------ */
/* four-step for lengths above one workgroup's LDS: n = n1 * n2         */
/* pass A: n2 column FFTs of length n1, times W_n^(k1*n2), into tmp     */
/* pass B: n1 row FFTs of length n2, written transposed to k1 + n1*k2   */
/* ------------------------------------------------------------------ */

/* all row distances of the four-step kernels are in floats */
/* input kinds of pass A */
enum { FS_IN_SPLIT = 0, FS_IN_PAIR = 1, FS_IN_INTER = 2 };
/* output kinds of pass B */
enum { FS_OUT_SPLIT = 0, FS_OUT_PAIR = 1, FS_OUT_INTER = 2 };

template <bool INV, int IN, int MT>
__global__ __launch_bounds__(MT) void k_fs_col(const fafd_fft *P1d, int n2, int cols, const float2 *twn,
                         const float *in_re, const float *in_im, long in_dist,
                         float *t_re, float *t_im, long t_dist) {
    const fafd_fft &P1 = *P1d;
    const long b = blockIdx.y;
    const int c0 = blockIdx.x * cols;
    const int n1 = P1.n, tid = (int)threadIdx.x, bd = (int)blockDim.x;
    const float *xr = in_re + b * in_dist;
    const float *xi = in_im ? in_im + b * in_dist : nullptr;
    for (int f = tid; f < cols * n1; f += bd) {
        const int c = f % cols, r = f / cols;
        const long i = (long)r * n2 + c0 + c;
        float2 v;
        if (c0 + c >= n2) v = make_float2(0.0f, 0.0f);
        else if (IN == FS_IN_SPLIT) v = make_float2(xr[i], xi[i]);
        else if (IN == FS_IN_PAIR) v = make_float2(xr[2 * i], xr[2 * i + 1]);
        else v = ((const float2 *)xr)[i];
        Lds{faf_smem + c * P1.plane}.st(r, v);
    }
    __syncthreads();
    const int sid = tid / P1.threads, lt = tid - sid * P1.threads;
    Lds lds{faf_smem + sid * P1.plane};
    fft_run<INV>(P1, lds, lds, lds, lt, sid < cols);
    __syncthreads();
    float *tr = t_re + b * t_dist, *ti = t_im + b * t_dist;
    for (int f = tid; f < cols * n1; f += bd) {
        const int c = f % cols, k1 = f / cols;
        if (c0 + c >= n2) continue;
        float2 v = Lds{faf_smem + c * P1.plane}.ld(k1);
        v = cmul(v, twn[(long)k1 * (c0 + c)]);
        const long o = (long)k1 * n2 + c0 + c;
        tr[o] = v.x;
        ti[o] = v.y;
    }
}

template <bool INV, int OUT, int MT>
__global__ __launch_bounds__(MT) void k_fs_row(const fafd_fft *P2d, int n1, int rows, const float *t_re,
                         const float *t_im, long t_dist, float *out_re, float *out_im,
                         long out_dist, float scale) {
    const fafd_fft &P2 = *P2d;
    const long b = blockIdx.y;
    const int k0 = blockIdx.x * rows;
    const int n2 = P2.n, tid = (int)threadIdx.x, bd = (int)blockDim.x;
    const float *tr = t_re + b * t_dist, *ti = t_im + b * t_dist;
    for (int r = 0; r < rows; ++r) {
        if (k0 + r >= n1) break;
        Lds row{faf_smem + r * P2.plane};
        const long base = (long)(k0 + r) * n2;
        for (int c = tid; c < n2; c += bd) row.st(c, make_float2(tr[base + c], ti[base + c]));
    }
    __syncthreads();
    const int sid = tid / P2.threads, lt = tid - sid * P2.threads;
    Lds lds{faf_smem + sid * P2.plane};
    fft_run<INV>(P2, lds, lds, lds, lt, sid < rows && k0 + sid < n1);
    __syncthreads();
    for (int f = tid; f < rows * n2; f += bd) {
        const int r = f % rows, k2 = f / rows;
        if (k0 + r >= n1) continue;
        float2 v = cscale(Lds{faf_smem + r * P2.plane}.ld(k2), scale);
        const long o = (long)(k0 + r) + (long)n1 * k2;
        if (OUT == FS_OUT_SPLIT) {
            out_re[b * out_dist + o] = v.x;
            out_im[b * out_dist + o] = v.y;
        } else if (OUT == FS_OUT_PAIR) {
            float *y = out_re + b * out_dist;
            y[2 * o] = v.x;
            y[2 * o + 1] = v.y;
        } else {
            ((float2 *)(out_re + b * out_dist))[o] = v;
        }
    }
}

/* Kiss post-pass over a natural-order packed spectrum in HBM (large R2C) */
template <bool INTER_OUT>
__global__ void k_r2c_post(int nc, const float2 *post, const float *z_re, const float *z_im,
                           long z_dist, float *out_re, float *out_im, long out_dist,
                           float scale) {
    const long b = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nc / 2) return;
    const float *zr = z_re + b * z_dist, *zi = z_im + b * z_dist;
    const int m = nc - k;
    float2 xk, xm;
    if (k == 0) {
        xk = make_float2(zr[0] + zi[0], 0.0f);
        xm = make_float2(zr[0] - zi[0], 0.0f);
    } else {
        kiss_post(make_float2(zr[k], zi[k]), make_float2(zr[m], zi[m]), post[k], xk, xm);
    }
    if (INTER_OUT) {
        float2 *o = (float2 *)out_re + b * out_dist;
        o[k] = cscale(xk, scale);
        o[m] = cscale(xm, scale);
    } else {
        float *ore = out_re + b * out_dist, *oim = out_im + b * out_dist;
        ore[k] = xk.x * scale; oim[k] = xk.y * scale;
        ore[m] = xm.x * scale; oim[m] = xm.y * scale;
    }
}

/* Kiss pre-pass into a natural-order packed spectrum in HBM (large C2R) */
template <bool INTER_IN>
__global__ void k_c2r_pre(int nc, const float2 *pre, const float *in_re, const float *in_im,
                          long in_dist, float *f_re, float *f_im, long f_dist) {
    const long b = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nc / 2) return;
    const int m = nc - k;
    float2 a, c;
    if (INTER_IN) {
        const float2 *ii = (const float2 *)in_re + b * in_dist;
        a = ii[k]; c = ii[m];
    } else {
        const float *ir = in_re + b * in_dist, *iq = in_im + b * in_dist;
        a = make_float2(ir[k], iq[k]); c = make_float2(ir[m], iq[m]);
    }
    float *fr = f_re + b * f_dist, *fi = f_im + b * f_dist;
    if (k == 0) {
        fr[0] = a.x + c.x;
        fi[0] = a.x - c.x;
        return;
    }
    float2 fk, fm;
    kiss_pre(a, c, pre[k], fk, fm);
    fr[k] = fk.x; fi[k] = fk.y;
    if (m != k) { fr[m] = fm.x; fi[m] = fm.y; }
}

/* ------------------------------------------------------------------ */
/* Rader (reference src/faf_fft_rader.c:136-232) and Bluestein          */
/* (reference src/faf_bluestein.c:165-223) inside one workgroup         */
/* ------------------------------------------------------------------ */

/* The tables of Rader and Bluestein plans are device memory; typed global,
 * their reads stay global loads when the plan itself was read from device
 * memory (serving and ragged launches), instead of flat loads. */
struct GatherLds {           /* a[j] = x[rin[j]] from the staged signal */
    Lds x;
    const faf_gi32 *rin;
    static constexpr bool is_lds = true;
    FAF_DEV float2 ld(int j) const { return x.ld(rin[j]); }
};

struct MulLds {              /* A[j] * B[j] */
    Lds a;
    const faf_gtw *b;
    static constexpr bool is_lds = true;
    FAF_DEV float2 ld(int j) const { return cmul(a.ld(j), twld(b, j)); }
};

struct ScatterLds {          /* y[rout[j]] = x0 + v  (B carries the 1/m) */
    Lds y;
    const faf_gi32 *rout;
    float2 x0;
    static constexpr bool is_lds = true;
    FAF_DEV void st(int j, float2 v) const { y.st(rout[j], cadd(x0, v)); }
};

/* Rader body for one signal slot.  xa: p slots, aa: p-1 slots (LDS).
 * Source of the signal and sink of the result are accessors. */
template <class S, class D>
FAF_DEV void rader_body(const fafd_fft &Pf, const fafd_fft &Pi, const fafd_rader &R,
                        const S &src, const D &dst, Lds xa, Lds aa, int lt, int T,
                        int active) {
    const int p = R.p;
    if (active)
        for (int i = lt; i < p; i += T) xa.st(i, src.ld(i));
    __syncthreads();
    float2 x0 = active ? xa.ld(0) : make_float2(0.0f, 0.0f);
    GatherLds g{xa, (const faf_gi32 *)R.rin};
    fft_run<false>(Pf, g, aa, aa, lt, active);
    __syncthreads();
    if (active && lt == 0) xa.st(0, cadd(x0, aa.ld(0)));   /* DC = x0 + sum x[1..] */
    MulLds mb{aa, (const faf_gtw *)R.b};
    ScatterLds sc{xa, (const faf_gi32 *)R.rout, x0};
    fft_run<true>(Pi, mb, aa, sc, lt, active);
    __syncthreads();
    if (active)
        for (int i = lt; i < p; i += T) dst.st(i, xa.ld(i));
}

template <class S> struct ChirpIn {   /* a[k] = x[k] chirp[k] for k < n, 0 up to m */
    S x;
    const faf_gtw *chirp;
    int n;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int k) const {
        return k < n ? cmul(x.ld(k), twld(chirp, k)) : make_float2(0.0f, 0.0f);
    }
};

template <class D> struct ChirpOut {
    D out;
    const faf_gtw *chirp;
    int n;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int k, float2 v) const {
        if (k < n) out.st(k, cmul(v, twld(chirp, k)));
    }
};

template <class S, class D>
FAF_DEV void blue_body(const fafd_fft &Pf, const fafd_fft &Pi, const fafd_blue &Bq,
                       const S &src, const D &dst, Lds aa, int lt, int active) {
    ChirpIn<S> ci{src, (const faf_gtw *)Bq.chirp, Bq.n};
    fft_run<false>(Pf, ci, aa, aa, lt, active);
    __syncthreads();
    MulLds mb{aa, (const faf_gtw *)Bq.b};
    ChirpOut<D> co{dst, (const faf_gtw *)Bq.chirp, Bq.n};
    fft_run<true>(Pi, mb, aa, co, lt, active);
}

#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_rader(const fafd_fft *Pfd,
                                               const fafd_fft *Pid, fafd_rader R, int spb, const float *in_re,
                        const float *in_im, long in_dist, float *out_re, float *out_im,
                        long out_dist, long count, float scale) {
    const fafd_fft &Pf = *Pfd;
    const fafd_fft &Pi = *Pid;
    Slot sl = slot_of(Pf.threads, spb, (long)blockIdx.x * spb, count);
    const int per = Pf.plane + lpad(R.p) + 1;
    Lds xa{faf_smem + sl.sid * per};
    Lds aa{faf_smem + sl.sid * per + lpad(R.p) + 1};
    const long sig = sl.active ? sl.sig : 0;
    GSplitIn src{in_re + sig * in_dist, in_im + sig * in_dist};
    GSplitOut dst{out_re + sig * out_dist, out_im + sig * out_dist, scale};
    rader_body(Pf, Pi, R, src, dst, xa, aa, sl.lt, Pf.threads, sl.active);
}
#endif


#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_blue(const fafd_fft *Pfd,
                                              const fafd_fft *Pid, fafd_blue Bq, int spb, const float *in_re,
                       const float *in_im, long in_dist, float *out_re, float *out_im,
                       long out_dist, long count, float scale) {
    const fafd_fft &Pf = *Pfd;
    const fafd_fft &Pi = *Pid;
    Slot sl = slot_of(Pf.threads, spb, (long)blockIdx.x * spb, count);
    Lds aa{faf_smem + sl.sid * Pf.plane};
    const long sig = sl.active ? sl.sig : 0;
    GSplitIn src{in_re + sig * in_dist, in_im + sig * in_dist};
    GSplitOut dst{out_re + sig * out_dist, out_im + sig * out_dist, scale};
    blue_body(Pf, Pi, Bq, src, dst, aa, sl.lt, sl.active);
}
#endif


/* ------------------------------------------------------------------ */
/* persistent ragged launch: units of same-plan signals pulled from an  */
/* atomic queue; signals live at offsets[] in concatenated split planes */
/* ------------------------------------------------------------------ */

/* the kernel itself, with compile-time passes for power-of-two lengths, is
 * kg_ragged in faf_kernels_rg.hip */


/* y = window * x on interleaved complex rows (single-frame STFT) */
#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
__global__ void k_window_inter(const float2 *x, long in_dist, const float *w, long n,
                               float2 *y, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * count) return;
    const long b = i / n, k = i - b * n;
    y[i] = cscale(x[b * in_dist + k], w[k]);
}
#endif


#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
__global__ void k_fill_u32(unsigned *p, int n, unsigned v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
#endif



/* ------------------------------------------------------------------ */
/* serving: one persistent launch per scheduling window; each unit is    */
/* up to `count` same-plan signals with their own in/out pointers        */
/* ------------------------------------------------------------------ */

/* The serving kernel reads its plans and request records from device
 * memory, so every pointer it holds is of unknown address space to the
 * compiler and its accesses would be flat (each also counts against the LDS
 * wait counter).  Its inputs, results, workspaces and tables are all device
 * memory: these accessors type them global. */
FAF_DEV void gst2(faf_gf2 *p, long i, float2 v) {
    faf_f2v r = {v.x, v.y};
    p[i] = r;
}

struct SvInterIn {           /* interleaved complex */
    const faf_gtw *p;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const { return twld(p, i); }
};

struct SvInterOut {
    faf_gf2 *p;
    float s;
    static constexpr bool is_lds = false;
    FAF_DEV void st(int i, float2 v) const { gst2(p, i, cscale(v, s)); }
};

struct SvPairIn {            /* real samples as pairs z[i] = x[2i] + i x[2i+1] */
    const faf_gcf *x;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const { return make_float2(x[2 * i], x[2 * i + 1]); }
};

struct SvFrameIn {           /* windowed frame, zero past the record (as GFrameIn) */
    const faf_gcf *x;
    const faf_gcf *w;
    long avail;
    static constexpr bool is_lds = false;
    FAF_DEV float2 ld(int i) const {
        const long a = 2L * i, b = a + 1;
        if (b < avail) return make_float2(x[a] * w[a], x[b] * w[b]);
        float u = a < avail ? x[a] * w[a] : 0.0f;
        return make_float2(u, 0.0f);
    }
};

struct SvOneSidedIn {        /* as OneSidedIn, with the filter row typed global */
    Lds X;
    const faf_gcf *psi;
    int nc;
    static constexpr bool is_lds = true;
    FAF_DEV float2 ld(int k) const {
        if (k > nc) return make_float2(0.0f, 0.0f);
        float2 v = X.ld(k);
        const float p = psi[k];
        if (k == 0 || k == nc) return make_float2(v.x * p, 0.0f);
        return make_float2(2.0f * v.x * p, 2.0f * v.y * p);
    }
};

/* Kiss post-pass from LDS Z into an interleaved Hermitian row */
FAF_DEV void post_pass_inter(const Lds &z, const faf_gtw *post, int nc, int lt, int T,
                             faf_gf2 *out) {
    for (int k = lt; k <= nc / 2; k += T) {
        const int m = nc - k;
        float2 xk, xm;
        if (k == 0) {
            float2 v = z.ld(0);
            xk = make_float2(v.x + v.y, 0.0f);
            xm = make_float2(v.x - v.y, 0.0f);
        } else {
            kiss_post(z.ld(k), z.ld(m), twld(post, k), xk, xm);
        }
        gst2(out, k, xk);
        gst2(out, m, xm);
    }
}

/* One unit of a four-step phase (lengths beyond one workgroup).  All
 * threads of the workgroup take part; FSCOL and FSROW hold fs_per FFTs of
 * pl->a side by side in LDS. */
FAF_DEV void serve_phase_unit(const fafd_splan *pl, const fafd_ssig &sg, int tid, int bd) {
    if (pl->kind == FAFD_SK_POST) {
        const int nc = pl->fs_other;
        const int k1 = sg.r0 + pl->fs_per < nc / 2 + 1 ? sg.r0 + pl->fs_per : nc / 2 + 1;
        const faf_gtw *z = (const faf_gtw *)sg.in;
        faf_gf2 *out = (faf_gf2 *)sg.out;
        const faf_gtw *post = (const faf_gtw *)pl->post;
        for (int k = sg.r0 + tid; k < k1; k += bd) {
            const int m = nc - k;
            float2 xk, xm;
            if (k == 0) {
                const float2 v = twld(z, 0);
                xk = make_float2(v.x + v.y, 0.0f);
                xm = make_float2(v.x - v.y, 0.0f);
            } else {
                kiss_post(twld(z, k), twld(z, m), twld(post, k), xk, xm);
            }
            gst2(out, k, xk);
            gst2(out, m, xm);
        }
        return;
    }
    const fafd_fft &P = pl->a;
    const int n = P.n, per = pl->fs_per, T = P.threads, other = pl->fs_other;
    const int sid = tid / T, lt = tid - sid * T;
    if (pl->kind == FAFD_SK_FSCOL) {
        /* columns c0 .. c0+per of x viewed as [n1][n2]: element (k1, c) */
        const int n2 = other, c0 = sg.r0;
        const faf_gtw *xc = (const faf_gtw *)sg.in;
        const faf_gcf *xin = (const faf_gcf *)sg.in;
        const faf_gcf *psi = (const faf_gcf *)pl->psi + sg.avail * pl->psi_stride;
        for (int f = tid; f < per * n; f += bd) {
            const int c = f % per, r = f / per;
            const long i = (long)r * n2 + c0 + c;
            float2 v = make_float2(0.0f, 0.0f);
            if (c0 + c < n2) {
                if (pl->fs_in == FAFD_FSIN_INTER) {
                    v = twld(xc, i);
                } else if (pl->fs_in == FAFD_FSIN_PAIR) {
                    v = make_float2(xin[2 * i], xin[2 * i + 1]);
                } else if (i <= pl->fs_nc) {   /* analytic row spectrum X psi_r */
                    float w = psi[i];
                    if (i != 0 && i != pl->fs_nc) w *= 2.0f;
                    v = cscale(twld(xc, i), w);
                }
            }
            Lds{faf_smem + c * P.plane}.st(r, v);
        }
        __syncthreads();
        Lds lds{faf_smem + sid * P.plane};
        if (pl->inv) fft_run<true>(P, lds, lds, lds, lt, sid < per);
        else fft_run<false>(P, lds, lds, lds, lt, sid < per);
        __syncthreads();
        const faf_gtw *tw = (const faf_gtw *)pl->twn;
        faf_gf2 *ws = (faf_gf2 *)sg.out;
        for (int f = tid; f < per * n; f += bd) {
            const int c = f % per, k1 = f / per;
            if (c0 + c >= n2) continue;
            const float2 v = Lds{faf_smem + c * P.plane}.ld(k1);
            gst2(ws, (long)k1 * n2 + c0 + c, cmul(v, twld(tw + (long)k1 * (c0 + c), 0)));
        }
        return;
    }
    /* FSROW: rows k0 .. k0+per of ws [n1][n2], FFT n2, out[k1 + n1 k2] */
    const int n1 = other, k0 = sg.r0;
    const faf_gtw *ws = (const faf_gtw *)sg.in;
    for (int r = 0; r < per && k0 + r < n1; ++r) {
        Lds row{faf_smem + r * P.plane};
        const faf_gtw *wr = ws + (long)(k0 + r) * n;
        for (int c = tid; c < n; c += bd) row.st(c, twld(wr, c));
    }
    __syncthreads();
    Lds lds{faf_smem + sid * P.plane};
    const int act = sid < per && k0 + sid < n1;
    if (pl->inv) fft_run<true>(P, lds, lds, lds, lt, act);
    else fft_run<false>(P, lds, lds, lds, lt, act);
    __syncthreads();
    faf_gf2 *out = (faf_gf2 *)sg.out;
    for (int f = tid; f < per * n; f += bd) {
        const int r = f % per, k2 = f / per;
        if (k0 + r >= n1) continue;
        gst2(out, (long)(k0 + r) + (long)n1 * k2,
             cscale(Lds{faf_smem + r * P.plane}.ld(k2), pl->scale));
    }
}

#ifndef FAF_ENGINE_NO_PLAIN_KERNELS
template <int MT>
__global__ __launch_bounds__(MT) void k_serve(const fafd_splan *plans, const fafd_sunit *units, int nunits,
                        const fafd_ssig *sigs, unsigned *counter) {
    /* with a counter, workgroups pull units from it until the table is
     * drained; without one, workgroup b runs unit b only */
    __shared__ int s_unit;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x;
    for (int it = 0;; ++it) {
        if (tid == 0)
            s_unit = counter ? (int)atomicAdd(counter, 1u) : (it == 0 ? (int)blockIdx.x : nunits);
        __syncthreads();
        const int u = s_unit;
        __syncthreads();
        if (u >= nunits) break;
        const fafd_sunit un = units[u];
        const fafd_splan *pl = plans + un.plan;
        if (pl->kind >= FAFD_SK_FSCOL) {
            serve_phase_unit(pl, sigs[un.first], tid, bd);
            __syncthreads();
            continue;
        }
        const int T = pl->threads;
        const int sid = tid / T, lt = tid - sid * T;
        const int active = sid < un.count && (sid + 1) * T <= bd;
        const fafd_ssig sg = sigs[un.first + (active ? sid : 0)];
        float2 *base = faf_smem + sid * pl->lds_per;
        switch (pl->kind) {
        case FAFD_SK_C2C: {
            Lds lds{base};
            fft_run<false>(pl->a, SvInterIn{(const faf_gtw *)sg.in}, lds,
                           SvInterOut{(faf_gf2 *)sg.out, 1.0f}, lt, active);
            break;
        }
        case FAFD_SK_RADER: {
            Lds xa{base};
            Lds aa{base + lpad(pl->rader.p) + 1};
            rader_body(pl->a, pl->b, pl->rader, SvInterIn{(const faf_gtw *)sg.in},
                       SvInterOut{(faf_gf2 *)sg.out, 1.0f}, xa, aa, lt, T, active);
            break;
        }
        case FAFD_SK_BLUE: {
            Lds aa{base};
            blue_body(pl->a, pl->b, pl->blue, SvInterIn{(const faf_gtw *)sg.in},
                      SvInterOut{(faf_gf2 *)sg.out, 1.0f}, aa, lt, active);
            break;
        }
        case FAFD_SK_R2C:
        case FAFD_SK_STFT: {
            Lds lds{base};
            if (pl->kind == FAFD_SK_R2C)
                fft_run<false>(pl->a, SvPairIn{(const faf_gcf *)sg.in}, lds, lds, lt, active);
            else
                fft_run<false>(pl->a,
                               SvFrameIn{(const faf_gcf *)sg.in, (const faf_gcf *)pl->win, sg.avail},
                               lds, lds, lt, active);
            __syncthreads();
            if (active) post_pass_inter(lds, (const faf_gtw *)pl->post, pl->a.n, lt, T,
                                        (faf_gf2 *)sg.out);
            break;
        }
        default: {              /* FAFD_SK_CWT: one segment per unit */
            const int nc = pl->a.n, n = pl->b.n;
            Lds X{faf_smem};
            Lds W{faf_smem + pl->a.plane};
            const int ah = tid < pl->a.threads, an = tid < pl->b.threads;
            fft_run<false>(pl->a, SvPairIn{(const faf_gcf *)sg.in}, X, X, ah ? tid : 0, ah);
            __syncthreads();
            if (ah)
                for (int k = tid; k <= nc / 2; k += pl->a.threads) {
                    const int m = nc - k;
                    float2 xk, xm;
                    if (k == 0) {
                        float2 z = X.ld(0);
                        xk = make_float2(z.x + z.y, 0.0f);
                        xm = make_float2(z.x - z.y, 0.0f);
                    } else {
                        kiss_post(X.ld(k), X.ld(m), twld((const faf_gtw *)pl->post, k), xk, xm);
                    }
                    X.st(k, xk);
                    X.st(m, xm);
                }
            __syncthreads();
            const float sc = 1.0f / (float)n;
            for (int r = sg.r0; r < sg.r1; ++r) {
                SvOneSidedIn os{X, (const faf_gcf *)pl->psi + (long)r * pl->psi_stride, nc};
                fft_run<true>(pl->b, os, W, SvInterOut{(faf_gf2 *)sg.out + (long)r * n, sc},
                              an ? tid : 0, an);
                __syncthreads();
            }
            break;
        }
        }
        __syncthreads();
    }
}

/* filtered one-sided CWT row spectra of long segments, for a separate
 * inverse FFT.  X: packed spectra [count][x_dist] (nc+1 bins used).
 * analytic: out[seg][r][k], k < 2 nc, holds X psi_r doubled on 0 < k < nc,
 * undoubled at k = 0 and nc, zero above nc (its inverse C2C gives the
 * analytic row).  Otherwise k <= nc holds X psi_r (for an inverse C2R).
 * grid: x = bins / 256, y = rows, z = segments. */
__global__ void k_cwt_spread(const float2 *X, long x_dist, const float *psi, long psi_stride,
                             int nc, int analytic, float2 *out, long row_len, long seg_dist,
                             int rows) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)blockIdx.y;
    const long seg = blockIdx.z;
    if (k >= row_len || r >= rows) return;
    float2 v = make_float2(0.0f, 0.0f);
    if (k <= nc) {
        const float2 z = X[seg * x_dist + k];
        float p = psi[(long)r * psi_stride + k];
        if (analytic && k != 0 && k != nc) p *= 2.0f;
        v = make_float2(z.x * p, z.y * p);
    }
    out[seg * seg_dist + (long)r * row_len + k] = v;
}
#endif

#endif
