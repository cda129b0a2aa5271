/*
 * faf_dev.h -- device-plan records shared by the C11 host code and the HIP
 * kernels.  Plain C, fixed-width fields, passed to kernels by value.
 *
 * A Stockham plan describes one complex FFT of length n computed as a
 * sequence of radix-r passes over LDS.  For pass s with radix r and span
 * ns (product of the radices before it), butterfly j in [0, n/r) reads
 * x[j + t*n/r] (t < r), multiplies leg t by W_n^(t*k*tw_stride) with
 * k = j mod ns, runs an r-point DFT and writes y[(j-k)*r + k + t*ns].
 * The first pass reads in natural order and the last pass writes in
 * natural order (self-sorting), so no digit-reversal permutation exists.
 */
#ifndef FAF_DEV_H
#define FAF_DEV_H
#include <stdint.h>

#define FAFD_MAX_STAGES 16
#define FAFD_MAXE 16                 /* complex values per thread per pass */
#define FAFD_MAX_LDS_BYTES 163840    /* gfx950 LDS per workgroup */
#define FAFD_MAX_OPS 8               /* spectral ops in a fused pipeline */

typedef struct fafd_stage {
    int32_t radix;
    int32_t ns;          /* span before this pass */
    uint32_t ns_magic;   /* ceil(2^32/ns): j/ns == umulhi(j, ns_magic) for j,ns <= 2^16 */
    int32_t tw_stride;   /* n / (ns*radix) */
} fafd_stage;

typedef struct fafd_fft {
    int32_t n;           /* complex length */
    int32_t nstages;
    int32_t threads;     /* threads cooperating on one signal */
    int32_t plane;       /* LDS floats per plane (n plus bank padding) */
    fafd_stage st[FAFD_MAX_STAGES];
    const void *tw;      /* device float2[n]: W_n^m for this direction */
} fafd_fft;

/* spectral op codes of the fused pipeline (reference include/chirp.h:36-38) */
enum { FAFD_OP_NONE = 0, FAFD_OP_CONJ = 1, FAFD_OP_MUL = 2, FAFD_OP_BANDPASS = 3 };

typedef struct fafd_ops {
    int32_t nops;
    int32_t code[FAFD_MAX_OPS];
    uint32_t a1[FAFD_MAX_OPS];
    uint32_t a2[FAFD_MAX_OPS];
    const float *h_re;   /* bound spectrum for MUL (device) */
    const float *h_im;
    int64_t h_bins;
} fafd_ops;

/* plan kinds of the persistent ragged launch */
enum { FAFD_KIND_STOCKHAM = 0, FAFD_KIND_RADER = 1, FAFD_KIND_BLUESTEIN = 2,
       FAFD_KIND_CT = 3,        /* Stockham through compile-time passes of its length */
       FAFD_KIND_RADER_CT = 4   /* Rader with compile-time inner passes */ };

typedef struct fafd_rader {
    int32_t p;           /* prime length */
    const int32_t *rin;  /* device: g^j mod p, j < p-1 */
    const int32_t *rout; /* device: g^-j mod p */
    const void *b;       /* device float2[p-1]: FFT_{p-1}(w^rout) / (p-1) */
} fafd_rader;

typedef struct fafd_blue {
    int32_t n;           /* signal length */
    int32_t m;           /* padded 5-smooth length >= 2n-1 */
    const void *chirp;   /* device float2[n]: exp(-+ i pi k^2/n), k^2 reduced mod 2n */
    const void *b;       /* device float2[m]: FFT_m(conj chirp pattern) / m */
} fafd_blue;

/* one entry of the ragged plan table (device resident) */
typedef struct fafd_rplan {
    int32_t kind;
    int32_t len;         /* signal length */
    int32_t threads;     /* threads per signal */
    int32_t spb;         /* signals per workgroup */
    int32_t variant;     /* compile-time kinds: which pass sequence of the length */
    int32_t path;        /* compile-time kinds: the kernel's code path, -1 otherwise */
    fafd_fft fwd;        /* main or inner forward FFT */
    fafd_fft inv;        /* inner inverse FFT (Rader / Bluestein) */
    fafd_rader rader;
    fafd_blue blue;
} fafd_rplan;

/* serving: kinds of work a persistent workgroup can pick up */
enum {
    FAFD_SK_C2C = 0,      /* Stockham C2C, interleaved complex in/out */
    FAFD_SK_RADER = 1,
    FAFD_SK_BLUE = 2,
    FAFD_SK_R2C = 3,      /* real in, interleaved Hermitian out */
    FAFD_SK_STFT = 4,     /* one windowed frame -> interleaved Hermitian */
    FAFD_SK_CWT = 5,      /* one segment: rfft then analytic rows [r0, r1) */
    /* phases of lengths beyond one workgroup (four-step n = n1 n2), one
     * signal-table entry per unit: in, out, avail = CWT filter row, r0 =
     * first column / row / bin of the unit */
    FAFD_SK_FSCOL = 6,    /* fs_per columns: gather stride n2, FFT n1, twiddle, ws[k1][n2] */
    FAFD_SK_FSROW = 7,    /* fs_per rows of ws: FFT n2, write out[k1 + n1 k2] */
    FAFD_SK_POST = 8      /* Kiss post-pass over fs_per bins of a packed spectrum */
};

/* what an FSCOL unit gathers */
enum { FAFD_FSIN_INTER = 0, FAFD_FSIN_PAIR = 1, FAFD_FSIN_SPREAD = 2 };

typedef struct fafd_splan {
    int32_t kind;
    int32_t len;         /* C2C length, real length (R2C/CWT) or window (STFT) */
    int32_t threads;     /* threads per signal */
    int32_t hop;         /* STFT */
    fafd_fft a;          /* main FFT, R2C inner, Rader/Bluestein inner forward */
    fafd_fft b;          /* Rader/Bluestein inner inverse; CWT length-n inverse */
    fafd_rader rader;
    fafd_blue blue;
    const void *post;    /* Kiss super-twiddles (R2C, STFT, CWT) */
    const float *win;    /* STFT window */
    const float *psi;    /* CWT filter rows */
    int64_t psi_stride;
    int32_t rows;        /* CWT rows */
    int32_t lds_per;     /* LDS float2 slots one signal needs */
    /* four-step phases */
    int32_t inv;         /* inverse transforms in this phase */
    int32_t fs_other;    /* FSCOL: n2; FSROW: n1; POST: packed length nc */
    int32_t fs_per;      /* columns / rows / bins per unit */
    int32_t fs_in;       /* FAFD_FSIN_* */
    int32_t fs_nc;       /* SPREAD: one-sided spectrum holds bins 0..fs_nc */
    float scale;         /* FSROW output scale */
    const void *twn;     /* FSCOL: W_n^m, n = n1 n2, in this phase's direction */
} fafd_splan;

typedef struct fafd_sunit {
    int32_t plan;
    int32_t count;       /* signals in this unit (all of one plan) */
    int32_t first;       /* index of the first signal in the signal table */
    int32_t pad;
} fafd_sunit;

typedef struct fafd_ssig {
    const float *in;     /* signal (or frame) start */
    float *out;          /* output start */
    int64_t avail;       /* STFT: samples left in the record from the frame start */
    int32_t r0, r1;      /* CWT: row range of this unit */
} fafd_ssig;

/* one unit of ragged work: `count` signals of one plan */
/* one unit of the ragged launch; the path, thread count and first offset
 * are copies of what the plan and the offset list hold, so a unit's data
 * loads depend on this one record only */
typedef struct fafd_unit {
    int32_t plan;
    int32_t count;
    int32_t first;       /* index into the sorted signal list */
    int32_t path;        /* the plan's compiled path (-1: runtime passes) */
    int32_t threads;     /* threads per signal */
    int32_t pad;
    int64_t off0;        /* offset of the unit's first signal */
} fafd_unit;

#endif
