/*
 * faf_kernels64.hip -- the FP64 execution variant (gfx950).
 *
 * One generic self-sorting Stockham kernel in double precision: the passes
 * of a runtime radix plan (2, 3, 4, 5, 7, 8, 16; the same fafd_fft records
 * as the FP32 runtime kernels) run over one LDS plane of double2 per
 * transform, reading HBM in the first pass and writing it in the last.  Each
 * transform is addressed through a strided accessor
 *     element i of transform s at (s / G) * gdist + (s % G) * tdist + i * estride
 * (complex elements, split or interleaved planes), so the same kernel runs
 * plain batches (G = 1), the column and row passes of a four-step split
 * (with the W_N^(k1 c) twiddle applied on output), and real signals read or
 * written as packed complex pairs.  Transforms either own consecutive lanes
 * (contiguous signals) or are interleaved lane by lane (strided columns, so
 * a wave reads whole rows).  Passes alternate between two LDS planes, so
 * a thread carries one butterfly at a time from load to store; a transform
 * has up to 4096 points in one workgroup (four-step beyond).
 *
 * Around it: the Kiss post-pass (R2C) and pre-pass (C2R) in FP64, and the
 * element-wise steps of Rader, Bluestein and the DCT/DST/MDCT family.
 * Twiddles are FP64 tables from exact integer phases.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include "faf_dev.h"
#include "faf_kern.h"

static thread_local char g_err64[256];
extern "C" const char *fafk64_last_error(void) { return g_err64; }

static int fail64(const char *what, hipError_t e) {
    snprintf(g_err64, sizeof g_err64, "%s: %s", what, hipGetErrorString(e));
    return -1;
}

#define D64 __device__ __forceinline__

D64 double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
D64 double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
D64 double2 dmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
D64 double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
D64 double2 dconj(double2 a) { return make_double2(a.x, -a.y); }
template <bool INV> D64 double2 drot(double2 v) {     /* * -i forward, * +i inverse */
    return INV ? make_double2(-v.y, v.x) : make_double2(v.y, -v.x);
}

/* ---- small DFTs in double, natural order, e^{-2 pi i / R} forward ---- */

template <bool INV> D64 void d2(double2 *v) {
    double2 a = v[0], b = v[1];
    v[0] = dadd(a, b);
    v[1] = dsub(a, b);
}

template <bool INV> D64 void d3(double2 *v) {
    const double c = 0.866025403784438646763723170752936183;
    double2 s = dadd(v[1], v[2]), d = dsub(v[1], v[2]);
    double2 m = make_double2(v[0].x - 0.5 * s.x, v[0].y - 0.5 * s.y);
    double2 r = dscale(drot<INV>(d), c);
    v[0] = dadd(v[0], s);
    v[1] = dadd(m, r);
    v[2] = dsub(m, r);
}

template <bool INV> D64 void d4(double2 *v) {
    double2 s0 = dadd(v[0], v[2]), e0 = dsub(v[0], v[2]);
    double2 s1 = dadd(v[1], v[3]), e1 = drot<INV>(dsub(v[1], v[3]));
    v[0] = dadd(s0, s1);
    v[2] = dsub(s0, s1);
    v[1] = dadd(e0, e1);
    v[3] = dsub(e0, e1);
}

template <bool INV> D64 void d5(double2 *v) {
    const double c1 = 0.309016994374947424102293417182819059;
    const double c2 = -0.809016994374947424102293417182819059;
    const double s1 = 0.951056516295153572116439333379382143;
    const double s2 = 0.587785252292473129168705954639072769;
    double2 a1 = dadd(v[1], v[4]), b1 = dsub(v[1], v[4]);
    double2 a2 = dadd(v[2], v[3]), b2 = dsub(v[2], v[3]);
    double2 x0 = v[0];
    double2 p1 = make_double2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    double2 p2 = make_double2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    double2 q1 = drot<INV>(make_double2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    double2 q2 = drot<INV>(make_double2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = make_double2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    v[1] = dadd(p1, q1);
    v[4] = dsub(p1, q1);
    v[2] = dadd(p2, q2);
    v[3] = dsub(p2, q2);
}

template <bool INV> D64 void d7(double2 *v) {
    const double c1 = 0.623489801858733530525004884004239810;
    const double c2 = -0.222520933956314404288902564496794759;
    const double c3 = -0.900968867902419126236102319507445051;
    const double s1 = 0.781831482468029808708444526674057750;
    const double s2 = 0.974927912181823607018131682993931217;
    const double s3 = 0.433883739117558120475768332848358754;
    double2 a1 = dadd(v[1], v[6]), b1 = dsub(v[1], v[6]);
    double2 a2 = dadd(v[2], v[5]), b2 = dsub(v[2], v[5]);
    double2 a3 = dadd(v[3], v[4]), b3 = dsub(v[3], v[4]);
    double2 x0 = v[0];
    double2 p1 = make_double2(x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x,
                              x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y);
    double2 p2 = make_double2(x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x,
                              x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y);
    double2 p3 = make_double2(x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x,
                              x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y);
    double2 q1 = drot<INV>(make_double2(s1 * b1.x + s2 * b2.x + s3 * b3.x,
                                        s1 * b1.y + s2 * b2.y + s3 * b3.y));
    double2 q2 = drot<INV>(make_double2(s2 * b1.x - s3 * b2.x - s1 * b3.x,
                                        s2 * b1.y - s3 * b2.y - s1 * b3.y));
    double2 q3 = drot<INV>(make_double2(s3 * b1.x - s1 * b2.x + s2 * b3.x,
                                        s3 * b1.y - s1 * b2.y + s2 * b3.y));
    v[0] = make_double2(x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y);
    v[1] = dadd(p1, q1); v[6] = dsub(p1, q1);
    v[2] = dadd(p2, q2); v[5] = dsub(p2, q2);
    v[3] = dadd(p3, q3); v[4] = dsub(p3, q3);
}

/* W_16^m forward (cos, -sin); the inverse uses the conjugate */
__device__ __constant__ double c_w16[16][2] = {
    {1.0, 0.0},
    {0.923879532511286756128183189396788933, -0.382683432365089771728459984030398866},
    {0.707106781186547524400844362104849039, -0.707106781186547524400844362104849039},
    {0.382683432365089771728459984030398866, -0.923879532511286756128183189396788933},
    {0.0, -1.0},
    {-0.382683432365089771728459984030398866, -0.923879532511286756128183189396788933},
    {-0.707106781186547524400844362104849039, -0.707106781186547524400844362104849039},
    {-0.923879532511286756128183189396788933, -0.382683432365089771728459984030398866},
    {-1.0, 0.0},
    {-0.923879532511286756128183189396788933, 0.382683432365089771728459984030398866},
    {-0.707106781186547524400844362104849039, 0.707106781186547524400844362104849039},
    {-0.382683432365089771728459984030398866, 0.923879532511286756128183189396788933},
    {0.0, 1.0},
    {0.382683432365089771728459984030398866, 0.923879532511286756128183189396788933},
    {0.707106781186547524400844362104849039, 0.707106781186547524400844362104849039},
    {0.923879532511286756128183189396788933, 0.382683432365089771728459984030398866}};

template <bool INV> D64 double2 w16(int m) {
    m &= 15;
    return make_double2(c_w16[m][0], INV ? -c_w16[m][1] : c_w16[m][1]);
}

/* 8 = 2 x 4: n = 2 n2 + n1, k = k1 + 4 k2 */
template <bool INV> D64 void d8(double2 *v) {
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    d4<INV>(e);
    d4<INV>(o);
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) o[k1] = dmul(o[k1], w16<INV>(2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        v[k1] = dadd(e[k1], o[k1]);
        v[k1 + 4] = dsub(e[k1], o[k1]);
    }
}

/* 16 = 4 x 4: n = 4 n2 + n1, k = k1 + 4 k2 */
template <bool INV> D64 void d16(double2 *v) {
    double2 y[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) y[n1][n2] = v[4 * n2 + n1];
        d4<INV>(y[n1]);
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1)
            if (n1) y[n1][k1] = dmul(y[n1][k1], w16<INV>(n1 * k1));
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        double2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        d4<INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

template <int R, bool INV> D64 void dft_r(double2 *v) {
    if (R == 2) d2<INV>(v);
    else if (R == 3) d3<INV>(v);
    else if (R == 4) d4<INV>(v);
    else if (R == 5) d5<INV>(v);
    else if (R == 7) d7<INV>(v);
    else if (R == 8) d8<INV>(v);
    else d16<INV>(v);
}

/* ---- the strided accessor ---- */

struct Acc64 {
    double *re, *im;      /* im == NULL: interleaved double2 at re */
    long G, gdist, tdist, estride;
};

D64 long acc_base(const Acc64 &a, long s) { return (s / a.G) * a.gdist + (s % a.G) * a.tdist; }

D64 double2 acc_ld(const Acc64 &a, long b, long i) {
    const long o = b + i * a.estride;
    if (!a.im) return ((const double2 *)a.re)[o];
    return make_double2(a.re[o], a.im[o]);
}

D64 void acc_st(const Acc64 &a, long b, long i, double2 v) {
    const long o = b + i * a.estride;
    if (!a.im) ((double2 *)a.re)[o] = v;
    else { a.re[o] = v.x; a.im[o] = v.y; }
}

D64 int lp64(int i) { return i + (i >> 4); }        /* one pad slot every 16 */

/* One pass of radix R over the transform: butterflies j = lt, lt + T, ...
 * read from HBM (first pass) or LDS plane src, written to LDS plane dst or
 * to HBM (last pass).  The planes alternate, so a butterfly's values go
 * from load to store without waiting on the others. */
template <int R, bool INV>
D64 void pass64(const fafd_stage &st, int n, int lt, int T, bool active, bool first, bool last,
                const double2 *__restrict__ tw, const double2 *src, double2 *dst,
                const Acc64 &in, long ib, const Acc64 &out, long ob, long sidx,
                const double2 *__restrict__ twn, long twN, double scale) {
    const int nb = n / R, NS = st.ns, tws = st.tw_stride;
    if (active) {
        for (int j = lt; j < nb; j += T) {
            double2 v[R];
#pragma unroll
            for (int t = 0; t < R; ++t)
                v[t] = first ? acc_ld(in, ib, j + t * nb) : src[lp64(j + t * nb)];
            const int k = j % NS;
#pragma unroll
            for (int t = 1; t < R; ++t) v[t] = dmul(v[t], tw[t * k * tws]);
            dft_r<R, INV>(v);
            const int o0 = (j - k) * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const int o = o0 + t * NS;
                if (last) {
                    double2 y = dscale(v[t], scale);
                    if (twn) {                /* four-step column pass: W_N^(k1 c) */
                        const long c = sidx % out.G;
                        y = dmul(y, twn[((long)o * c) % twN]);
                    }
                    acc_st(out, ob, o, y);
                } else {
                    dst[lp64(o)] = v[t];
                }
            }
        }
    }
    if (!last) __syncthreads();
}

template <bool INV>
__global__ __launch_bounds__(512) void k64_fft(const fafd_fft *__restrict__ Pd,
                                                const double2 *__restrict__ tw, long count,
                                                int C, int T, int lane_major, Acc64 in,
                                                Acc64 out, const double2 *__restrict__ twn,
                                                long twN, double scale) {
    extern __shared__ double2 lds64[];
    const fafd_fft &P = *Pd;
    const int n = P.n, nst = P.nstages;
    int c, lt;
    if (lane_major) { c = threadIdx.x % C; lt = threadIdx.x / C; }
    else { c = threadIdx.x / T; lt = threadIdx.x % T; }
    const long s = (long)blockIdx.x * C + c;
    const bool active = s < count;
    const int plane = lp64(n) + 1;
    double2 *pa = lds64 + (long)c * 2 * plane, *pb = pa + plane;
    const long ib = active ? acc_base(in, s) : 0, ob = active ? acc_base(out, s) : 0;
    for (int q = 0; q < nst; ++q) {
        const fafd_stage &st = P.st[q];
        const bool first = q == 0, last = q == nst - 1;
        const double2 *src = (q & 1) ? pa : pb;     /* written by pass q - 1 */
        double2 *dst = (q & 1) ? pb : pa;
        switch (st.radix) {
        case 2: pass64<2, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        case 3: pass64<3, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        case 4: pass64<4, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        case 5: pass64<5, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        case 7: pass64<7, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        case 8: pass64<8, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        default: pass64<16, INV>(st, n, lt, T, active, first, last, tw, src, dst, in, ib, out, ob, s, twn, twN, scale); break;
        }
    }
}

/* threads per transform: a few butterflies each, 64..512 */
extern "C" int fafk64_threads(int n) {
    int t = ((n / 8 + 63) / 64) * 64;
    return t < 64 ? 64 : (t > 512 ? 512 : t);
}

/* LDS bytes of one transform: two planes */
static size_t plane_bytes(int n) { return 2 * (size_t)((n + (n >> 4)) + 1) * sizeof(double2); }

extern "C" int fafk64_init(void) {
    (void)hipFuncSetAttribute((const void *)k64_fft<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, FAFD_MAX_LDS_BYTES);
    (void)hipFuncSetAttribute((const void *)k64_fft<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, FAFD_MAX_LDS_BYTES);
    (void)hipGetLastError();
    return 0;
}

extern "C" int fafk64_fft(const void *Pdev, int n, const void *tw, int inv, long count,
                          int lane_major, const fafk_acc64 *in, const fafk_acc64 *out,
                          const void *twn, long twN, double scale, void *stream) {
    if (n < 2 || n > 4096 || count <= 0) {
        snprintf(g_err64, sizeof g_err64, "k64_fft: unsupported length %d (count %ld)", n, count);
        return -1;
    }
    /* columns: 16 transforms interleaved lane by lane with few threads
     * each; contiguous signals: whole waves per transform */
    int T = lane_major ? (n / 8 < 16 ? 16 : (n / 8 > 32 ? 32 : n / 8)) : fafk64_threads(n);
    const size_t pb = plane_bytes(n);
    int C = lane_major ? 16 : (256 / T > 1 ? 256 / T : 1);
    while (C > 1 && ((size_t)C * pb > FAFD_MAX_LDS_BYTES || C * T > 512)) --C;
    if ((size_t)C * pb > FAFD_MAX_LDS_BYTES) {
        snprintf(g_err64, sizeof g_err64, "k64_fft: length %d does not fit LDS", n);
        return -1;
    }
    const long grid = (count + C - 1) / C;
    if (grid > 0x7fffffffL) {
        snprintf(g_err64, sizeof g_err64, "k64_fft: %ld transforms in one launch", count);
        return -1;
    }
    Acc64 a{(double *)in->re, (double *)in->im, in->G, in->gdist, in->tdist, in->estride};
    Acc64 b{(double *)out->re, (double *)out->im, out->G, out->gdist, out->tdist, out->estride};
    if (inv)
        hipLaunchKernelGGL(k64_fft<true>, dim3((unsigned)grid), dim3(C * T), C * pb,
                           (hipStream_t)stream, (const fafd_fft *)Pdev, (const double2 *)tw, count,
                           C, T, lane_major, a, b, (const double2 *)twn, twN, scale);
    else
        hipLaunchKernelGGL(k64_fft<false>, dim3((unsigned)grid), dim3(C * T), C * pb,
                           (hipStream_t)stream, (const fafd_fft *)Pdev, (const double2 *)tw, count,
                           C, T, lane_major, a, b, (const double2 *)twn, twN, scale);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_fft", e);
}

/* ---- element-wise kernels: one thread per element, bps blocks per signal ---- */

static int grid64(long len, long count, long *bps, const char *what) {
    *bps = (len + 255) / 256;
    if (len <= 0 || count <= 0 || *bps * count > 0x7fffffffL) {
        snprintf(g_err64, sizeof g_err64, "%s: unsupported shape len=%ld count=%ld", what, len,
                 count);
        return -1;
    }
    return 0;
}

#define ELEM_IDX(LEN)                                                   \
    const long s = blockIdx.x / bps;                                    \
    const long k = (blockIdx.x - s * bps) * 256 + threadIdx.x;          \
    if (k >= (LEN)) return;

/* Kiss post-pass: Z (interleaved, nc per signal, zdist) -> X[0..nc] (split
 * re/im or interleaved at re), scale */
__global__ __launch_bounds__(256) void k64_post(const double2 *__restrict__ Z, long zdist,
                                                const double2 *__restrict__ tw, long nc,
                                                double *__restrict__ xr, double *__restrict__ xi,
                                                long xdist, double scale, long bps) {
    ELEM_IDX(nc + 1)
    const double2 *z = Z + s * zdist;
    double2 X;
    if (k == 0 || k == nc) {
        const double2 z0 = z[0];
        X = make_double2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0);
    } else {
        const double2 a = z[k], b = dconj(z[nc - k]);
        const double2 A = dadd(a, b), Bv = dsub(a, b);
        X = dscale(dadd(A, dmul(tw[k], Bv)), 0.5);      /* tw[k] = -i W_n^k */
    }
    X = dscale(X, scale);
    if (xi) { xr[s * xdist + k] = X.x; xi[s * xdist + k] = X.y; }
    else ((double2 *)xr)[s * xdist + k] = X;
}

/* Kiss pre-pass: X[0..nc] -> Z[k] = (X[k] + conj X[nc-k]) + i W^-k (X[k] - conj X[nc-k]),
 * DC and Nyquist imaginary parts dropped; tw[k] = i W_n^-k */
__global__ __launch_bounds__(256) void k64_pre(const double *__restrict__ xr,
                                               const double *__restrict__ xi, long xdist,
                                               const double2 *__restrict__ tw, long nc,
                                               double2 *__restrict__ Z, long zdist, long bps) {
    ELEM_IDX(nc)
    auto ld = [&](long i) {
        double2 v = xi ? make_double2(xr[s * xdist + i], xi[s * xdist + i])
                       : ((const double2 *)xr)[s * xdist + i];
        if (i == 0 || i == nc) v.y = 0.0;
        return v;
    };
    const double2 a = ld(k), b = dconj(ld(nc - k));
    Z[s * zdist + k] = dadd(dadd(a, b), dmul(tw[k], dsub(a, b)));
}

extern "C" int fafk64_post(const void *Z, long zdist, const void *tw, long nc, double *xr,
                           double *xi, long xdist, double scale, long count, void *stream) {
    long bps;
    if (grid64(nc + 1, count, &bps, "k64_post")) return -1;
    hipLaunchKernelGGL(k64_post, dim3((unsigned)(bps * count)), dim3(256), 0, (hipStream_t)stream,
                       (const double2 *)Z, zdist, (const double2 *)tw, nc, xr, xi, xdist, scale,
                       bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_post", e);
}

extern "C" int fafk64_pre(const double *xr, const double *xi, long xdist, const void *tw, long nc,
                          void *Z, long zdist, long count, void *stream) {
    long bps;
    if (grid64(nc, count, &bps, "k64_pre")) return -1;
    hipLaunchKernelGGL(k64_pre, dim3((unsigned)(bps * count)), dim3(256), 0, (hipStream_t)stream,
                       xr, xi, xdist, (const double2 *)tw, nc, (double2 *)Z, zdist, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_pre", e);
}

/* Rader gather / Bluestein chirp-in: a[j] = x[idx[j]] (idx) or
 * x[j] chirp[j] (j < n, zero to L); split planes */
__global__ __launch_bounds__(256) void k64_in(const double *__restrict__ xr,
                                              const double *__restrict__ xi, long xdist,
                                              const int32_t *__restrict__ idx,
                                              const double2 *__restrict__ chirp, long n, long L,
                                              double *__restrict__ ar, double *__restrict__ ai,
                                              long adist, long bps) {
    ELEM_IDX(L)
    double2 v = make_double2(0.0, 0.0);
    if (idx) {
        const long o = s * xdist + idx[k];
        v = make_double2(xr[o], xi[o]);
    } else if (k < n) {
        v = dmul(make_double2(xr[s * xdist + k], xi[s * xdist + k]), chirp[k]);
    }
    ar[s * adist + k] = v.x;
    ai[s * adist + k] = v.y;
}

/* A *= B; save0 (optional) gets A[0] first */
__global__ __launch_bounds__(256) void k64_mul(double *__restrict__ ar, double *__restrict__ ai,
                                               long adist, const double2 *__restrict__ B, long m,
                                               double2 *__restrict__ save0, long bps) {
    ELEM_IDX(m)
    const double2 a = make_double2(ar[s * adist + k], ai[s * adist + k]);
    if (save0 && k == 0) save0[s] = a;
    const double2 v = dmul(a, B[k]);
    ar[s * adist + k] = v.x;
    ai[s * adist + k] = v.y;
}

/* Rader scatter (idx) or Bluestein chirp-out; out length n (Rader: the
 * m = n - 1 outputs at idx, plus y[0] = x0 + save0) */
__global__ __launch_bounds__(256) void k64_out(const double *__restrict__ ar,
                                               const double *__restrict__ ai, long adist,
                                               const int32_t *__restrict__ idx,
                                               const double2 *__restrict__ chirp, long len,
                                               const double *__restrict__ xr,
                                               const double *__restrict__ xi, long xdist,
                                               const double2 *__restrict__ save0,
                                               double *__restrict__ yr, double *__restrict__ yi,
                                               long ydist, double scale, long bps) {
    ELEM_IDX(len)
    const double2 a = make_double2(ar[s * adist + k], ai[s * adist + k]);
    if (idx) {
        const double2 x0 = make_double2(xr[s * xdist], xi[s * xdist]);
        const double2 v = dadd(x0, a);
        const long d = s * ydist + idx[k];
        yr[d] = v.x * scale;
        yi[d] = v.y * scale;
        if (k == 0) {
            const double2 dc = dadd(x0, save0[s]);
            yr[s * ydist] = dc.x * scale;
            yi[s * ydist] = dc.y * scale;
        }
    } else {
        const double2 v = dmul(a, chirp[k]);
        yr[s * ydist + k] = v.x * scale;
        yi[s * ydist + k] = v.y * scale;
    }
}

extern "C" int fafk64_in(const double *xr, const double *xi, long xdist, const int32_t *idx,
                         const void *chirp, long n, long L, double *ar, double *ai, long adist,
                         long count, void *stream) {
    long bps;
    if (grid64(L, count, &bps, "k64_in")) return -1;
    hipLaunchKernelGGL(k64_in, dim3((unsigned)(bps * count)), dim3(256), 0, (hipStream_t)stream,
                       xr, xi, xdist, idx, (const double2 *)chirp, n, L, ar, ai, adist, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_in", e);
}

extern "C" int fafk64_mul(double *ar, double *ai, long adist, const void *B, long m, void *save0,
                          long count, void *stream) {
    long bps;
    if (grid64(m, count, &bps, "k64_mul")) return -1;
    hipLaunchKernelGGL(k64_mul, dim3((unsigned)(bps * count)), dim3(256), 0, (hipStream_t)stream,
                       ar, ai, adist, (const double2 *)B, m, (double2 *)save0, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_mul", e);
}

extern "C" int fafk64_out(const double *ar, const double *ai, long adist, const int32_t *idx,
                          const void *chirp, long len, const double *xr, const double *xi,
                          long xdist, const void *save0, double *yr, double *yi, long ydist,
                          double scale, long count, void *stream) {
    long bps;
    if (grid64(len, count, &bps, "k64_out")) return -1;
    hipLaunchKernelGGL(k64_out, dim3((unsigned)(bps * count)), dim3(256), 0, (hipStream_t)stream,
                       ar, ai, adist, idx, (const double2 *)chirp, len, xr, xi, xdist,
                       (const double2 *)save0, yr, yi, ydist, scale, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_out", e);
}

/* DCT / DST / MDCT around the complex FFT, FP64 (as k_r2r_pre / post) */
__global__ __launch_bounds__(256) void k64_r2r_pre(const double *__restrict__ x, long xdist,
                                                   long nin, const double2 *__restrict__ pre,
                                                   int shift, long L, double *__restrict__ ur,
                                                   double *__restrict__ ui, long udist,
                                                   long bps) {
    ELEM_IDX(L)
    const long i = k - shift;
    double2 v = make_double2(0.0, 0.0);
    if (i >= 0 && i < nin) v = dscale(pre[k], x[s * xdist + i]);
    ur[s * udist + k] = v.x;
    ui[s * udist + k] = v.y;
}

__global__ __launch_bounds__(256) void k64_r2r_post(const double *__restrict__ ur,
                                                    const double *__restrict__ ui, long udist,
                                                    const double2 *__restrict__ post, int off,
                                                    int use_im, long nout,
                                                    double *__restrict__ y, long ydist,
                                                    long bps) {
    ELEM_IDX(nout)
    const double2 u = make_double2(ur[s * udist + k + off], ui[s * udist + k + off]);
    const double2 v = dmul(post[k], u);
    y[s * ydist + k] = use_im ? v.y : v.x;
}

extern "C" int fafk64_r2r_pre(const double *x, long xdist, long nin, const void *pre, int shift,
                              long L, double *ur, double *ui, long udist, long count,
                              void *stream) {
    long bps;
    if (grid64(L, count, &bps, "k64_r2r_pre")) return -1;
    hipLaunchKernelGGL(k64_r2r_pre, dim3((unsigned)(bps * count)), dim3(256), 0,
                       (hipStream_t)stream, x, xdist, nin, (const double2 *)pre, shift, L, ur,
                       ui, udist, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_r2r_pre", e);
}

extern "C" int fafk64_r2r_post(const double *ur, const double *ui, long udist, const void *post,
                               int off, int use_im, long nout, double *y, long ydist, long count,
                               void *stream) {
    long bps;
    if (grid64(nout, count, &bps, "k64_r2r_post")) return -1;
    hipLaunchKernelGGL(k64_r2r_post, dim3((unsigned)(bps * count)), dim3(256), 0,
                       (hipStream_t)stream, ur, ui, udist, (const double2 *)post, off, use_im,
                       nout, y, ydist, bps);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : fail64("k64_r2r_post", e);
}
