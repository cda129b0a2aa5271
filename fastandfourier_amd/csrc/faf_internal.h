/*
 * faf_internal.h -- host-side internals of the MI355X transform library
 * (C11).  The public ABI is include/fastandfourier.h + include/faf_amd.h.
 */
#ifndef FAF_INTERNAL_H
#define FAF_INTERNAL_H
#include "fastandfourier.h"
#include "faf_amd.h"
#include "faf_dev.h"
#include "faf_kern.h"

/* ---- errors ---- */
void fafi_err(const char *fmt, ...);

/* ---- integer size policy (reference src/faf_fft_stage.c:25-135) ---- */
int fafi_is_5_smooth(size_t n);
int fafi_is_7_smooth(size_t n);
int fafi_is_prime(size_t n);
int fafi_rader_eligible(size_t n);
size_t fafi_primitive_root(size_t p);
size_t fafi_next_5_smooth(size_t n);
int fafi_is_codelet_size(size_t n);

/* ---- device runtime ---- */
int fafi_runtime_ready(void);         /* 0 ok, -1 with error set */
int fafi_create_ready(void);          /* as above, and no graph capture open */
int fafi_capturing(void);             /* 1 while this thread captures a graph */
int fafi_noalloc(void);               /* 1 inside faf_execute_batch_noalloc */
void *fafi_stream(void);              /* current thread's stream (NULL = default) */
int fafi_sync_mode(void);
int fafi_ct_enabled(void);            /* plans may pick compile-time kernels */
size_t fafi_cus(void);

/* ---- plan building (faf_plan.c) ---- */
int fafi_radix_fits(size_t n, int allow7);             /* one-workgroup LDS plan exists */
int fafi_make_fft(size_t n, int inv, int allow7, const int *order, int norder,
                  fafd_fft *out);                       /* 0 ok */
const void *fafi_twiddles(size_t n, int inv);           /* device float2[n], cached */
const void *fafi_kiss_table(size_t nc, int inv);        /* device float2[nc/2+1], cached */
int fafi_default_spb(const fafd_fft *P, size_t extra_per_signal);
int fafi_fft64(size_t n, int inv, double *re, double *im);   /* host FP64 FFT, in place */
/* exp(-+ 2 pi i m / n) in FP64 from the exact integer phase m mod n
 * (inv selects +) */
void fafi_unit_root(size_t m, size_t n, int inv, double *c, double *s);
int fafi_split_4step(size_t n, int *n1, int *n2);


/* ---- the device plan behind faf_transform.amd ---- */
typedef enum {
    PK_NONE = 0, PK_C2C, PK_C2C_4S, PK_RADER, PK_BLUE, PK_R2C, PK_C2R, PK_R2C_4S, PK_C2R_4S,
    PK_PIPE, PK_STFT, PK_CWT, PK_ICWT,
    PK_RADER_L, PK_BLUE_L,     /* Rader / Bluestein with inner FFTs beyond one workgroup */
    PK_R2R,                    /* DCT / DST / MDCT through a complex FFT */
    /* the FP64 variant (faf_fp64.c) */
    PK_C2C64, PK_C2C64_4S, PK_RADER64, PK_BLUE64, PK_R2C64, PK_C2R64
} fafh_kind;

typedef struct fafh_plan {
    fafh_kind kind;
    int inverse;
    size_t n;                 /* logical length */
    fafd_fft a, b;            /* main / inner forward, inner inverse, or 4-step n1,n2 */
    int n1, n2, cols, rows;
    const void *twn;          /* four-step W_n table */
    const void *post, *pre;   /* Kiss tables */
    fafd_rader rader;
    fafd_blue blue;
    void *rader_mem[3];
    void *blue_mem[2];
    int spb;
    int ct;                   /* use the compile-time specialised kernel */
    int ct4;                  /* four-step with specialised column/row passes */
    int rs;                   /* single-pass register-resident kernels: 32768 points
                                 (four-step plans) or 8192 (rfft / irfft 16384) */
    float scale;
    int interleaved;          /* buffers are interleaved complex */
    void *ws;                 /* device workspace */
    size_t ws_bytes;
    int ws_captured;          /* a graph holds ws: growth retires it, not frees it */
    void **ws_retired;        /* earlier workspaces kept for captured graphs */
    int n_retired;
    void *stg;                /* device staging for host buffers */
    size_t stg_bytes;
    /* pipeline */
    fafd_ops ops;
    void *h_dev;
    size_t h_dev_cap;
    const void *h_host_re, *h_host_im;
    size_t h_host_bins;
    /* framed STFT */
    void *win_dev;
    size_t win, hop;
    /* CWT */
    void *bank;               /* fafi_cwt_bank* */
    void *psi_dev;            /* device float[rows][stride] (forward) */
    void *dual_dev;           /* device float[rows][stride] (dual inverse) */
    long psi_stride;
    int rows_total, rows_per_wg;
    fafd_fft full_inv;        /* length-n inverse for analytic CWT rows */
    /* CWT segments longer than one workgroup holds: rfft, filter spread,
     * then a batched inverse over the rows (C2C for analytic rows, C2R for
     * real rows), each a plan of its own */
    struct fafh_plan *sub_r2c, *sub_c2c, *sub_c2r;
    /* Rader / Bluestein beyond one workgroup: the inner forward and
     * (unnormalised) inverse C2C of length p-1 or M */
    struct fafh_plan *inner_f, *inner_i;
    /* DCT / DST / MDCT (faf_r2r.c): y[k] = Re or Im of post[k] * U[k + off],
     * U = C2C_L(u), u[j] = pre[j] * x[j - shift] (zero outside the input) */
    void *r2r_pre, *r2r_post;
    /* FP64 plans: computed in double; tw64 / tw64b the double2 twiddle
     * tables of a / b (R2C / C2R: the Kiss table), twn64 the four-step
     * W_n table, scale64 the norm */
    int fp64;
    const void *tw64, *tw64b, *twn64;
    double scale64;
    long r2r_L, r2r_nin, r2r_nout;
    int r2r_shift, r2r_off, r2r_im;
    /* caller-lent device workspace, used instead of ws while set (the
     * serving loop lends each request its own so streams never share) */
    void *ws_ext;
    size_t ws_ext_bytes;
} fafh_plan;

/* ---- autotuned launch shapes (faf_tune.c), kept for the process and
 * keyed by (plan kind, length, direction, interleaved layout) ---- */
typedef struct fafi_shape {
    int kind;
    size_t n;
    int inverse, interleaved;
    int specialised;          /* the compile-time / single-pass kernel */
    int norder;               /* otherwise: runtime radix order ... */
    int order[FAFD_MAX_STAGES];
    int spb;                  /* ... and signals per workgroup */
    double ms;                /* measured median time of the tuning batch */
} fafi_shape;
int fafi_tune_find(int kind, size_t n, int inverse, int interleaved, fafi_shape *out);
/* applies the recorded shape of p's key, if any (end of plan building) */
int fafi_apply_tuned(fafh_plan *p);
/* times p's candidate shapes on `count` signals, applies and records the
 * fastest; returns its spb or -1 */
int fafi_autotune_plan(fafh_plan *p, size_t count);

fafh_plan *fafi_plan_new(void);
void fafi_plan_free(fafh_plan *p);
void *fafi_workspace(fafh_plan *p, size_t bytes);
void *fafi_staging(fafh_plan *p, size_t bytes);

/* build device plans for the transform kinds (faf_plan.c) */
int fafi_plan_c2c(fafh_plan *p, size_t n, int inv, int bluestein);
int fafi_plan_real(fafh_plan *p, size_t n, int inv);
int fafi_plan_pipe(fafh_plan *p, size_t n);
int fafi_plan_stft(fafh_plan *p, size_t win, size_t hop);

/* launches on device buffers (faf_launch.c); dists in elements */
int fafi_run_c2c(fafh_plan *p, const float *ir, const float *ii, long idist, float *orr,
                 float *oi, long odist, long count, void *stream);
int fafi_run_r2c(fafh_plan *p, const float *x, long idist, float *orr, float *oi, long odist,
                 long count, void *stream);
int fafi_run_c2r(fafh_plan *p, const float *ir, const float *ii, long idist, float *y,
                 long odist, long count, void *stream);
int fafi_run_pipe(fafh_plan *p, const float *x, long idist, float *y, long odist, long count,
                  void *stream);
int fafi_pipe_refresh(fafh_plan *p, void *stream);
size_t fafi_run_ws_bytes(const fafh_plan *p, long count);   /* workspace one run needs */

/* DCT / DST / MDCT (faf_r2r.c): kind 0 DCT, 1 DST (type 1..4), 2 MDCT */
int fafi_plan_r2r(fafh_plan *p, int kind, int type, size_t n, int inverse, int norm, int fp64);
int fafi_run_r2r(fafh_plan *p, const float *x, long idist, float *y, long odist, long count,
                 void *stream);

/* FP64 variant (faf_fp64.c); buffers hold doubles, dists in elements */
int fafi_plan_c2c64(fafh_plan *p, size_t n, int inv, int bluestein, int allow7);
int fafi_plan_real64(fafh_plan *p, size_t n, int inv);
size_t fafi_ws64_bytes(const fafh_plan *p, long count);
int fafi_run_c2c64(fafh_plan *p, const double *ir, const double *ii, long idist, double *orr,
                   double *oi, long odist, long count, void *stream);
int fafi_run_r2c64(fafh_plan *p, const double *x, long idist, double *orr, double *oi,
                   long odist, long count, void *stream);
int fafi_run_c2r64(fafh_plan *p, const double *ir, const double *ii, long idist, double *y,
                   long odist, long count, void *stream);
int fafi_run_r2r64(fafh_plan *p, const double *x, long idist, double *y, long odist, long count,
                   void *stream);

/* CWT bank (faf_cwt.c) */
int fafi_cwt_build(fafh_plan *p, const faf_cwt_config *cfg, int inverse_kind);
void fafi_cwt_free(fafh_plan *p);
int fafi_run_cwt(fafh_plan *p, const float *x, long idist, float *out, long odist,
                 long count, int complex_out, void *stream);
int fafi_run_icwt(fafh_plan *p, const float *w, long idist, float *y, long odist,
                  long count, void *stream);
size_t fafi_cwt_ws_bytes(const fafh_plan *p, long count, int complex_out);

#endif
