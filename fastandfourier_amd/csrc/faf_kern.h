/*
 * faf_kern.h -- C declarations of the kernel launch layer (faf_kernels.hip).
 * Included by the C11 host code; all arguments are plain pointers, sizes
 * and the plan records of faf_dev.h.
 */
#ifndef FAF_KERN_H
#define FAF_KERN_H
#include <stddef.h>
#include <stdint.h>
#include "faf_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *fafk_last_error(void);
int fafk_init(void);
/* the static scratch of the allocation-free entry point: its device address
 * on the current device (resolved once per device) and its size */
#define FAFK_SCRATCH_BYTES ((size_t)32 << 20)   /* = FAF_NOALLOC_SCRATCH_BYTES */
void *fafk_scratch(size_t *bytes);
/* thread-local: while on, fafk_malloc and the plan-record cache refuse to
 * allocate (they fail with a message instead) */
void fafk_set_noalloc(int on);
void *fafk_malloc(size_t bytes);
void fafk_free(void *p);
int fafk_h2d(void *dst, const void *src, size_t bytes, void *stream);
int fafk_d2h(void *dst, const void *src, size_t bytes, void *stream);
int fafk_d2d(void *dst, const void *src, size_t bytes, void *stream);
int fafk_sync(void *stream);
int fafk_is_device_ptr(const void *p);
void *fafk_stream_create(void);
void fafk_stream_destroy(void *s);
void *fafk_event_create(void);
void fafk_event_destroy(void *ev);
int fafk_event_record(void *ev, void *stream);
int fafk_event_query(void *ev);
int fafk_event_sync(void *ev);
float fafk_event_ms(void *a, void *b);
int fafk_device_info(int *cus, size_t *lds_per_block, size_t *total_mem);
/* single-pass 32768-point kernels (faf_kernels_rs.hip); grid 0 = one
 * workgroup per CU */
const char *fafk_rs_last_error(void);
int fafk_rs_init(void);
int fafk_rs_pair_ok(const void *p, long dist);
/* rfft / irfft N = 16384 in one pass: tw = W_8192 table of the direction,
 * post / pre = Kiss tables of 8192; rows must be 8-byte aligned */
/* irfft N = 4096 in one pass: tw = inverse W_2048 table, pre = inverse
 * Kiss table of 2048; rows must be 8-byte aligned */
/* the single-MUL spectral pipeline N = 4096 in one pass: twf / twi = forward
 * / inverse W_2048 tables; x and y rows must be 8-byte aligned */
int fafk_rs2_pipe(const void *twf, const void *twi, const void *post, const void *pre,
                  const fafd_ops *ops, const float *x, long in_dist, float *y, long out_dist,
                  long count, float scale, void *stream);
int fafk_rs2_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                 const float *in_im, long in_dist, float *y, long out_dist, long count,
                 float scale, void *stream);
/* C2C 8192 in one pass: tw = W_8192 table of the direction */
int fafk_rs8_c2c(int inv, int inter, const void *tw, const float *in_re, const float *in_im,
                 long in_dist, float *out_re, float *out_im, long out_dist, long count,
                 float scale, void *stream);
int fafk_rs8_r2c(int inter_out, const void *tw, const void *post, const float *x, long in_dist,
                 float *out_re, float *out_im, long out_dist, long count, float scale,
                 void *stream);
int fafk_rs8_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                 const float *in_im, long in_dist, float *y, long out_dist, long count,
                 float scale, void *stream);
int fafk_rs_r2c(int inter_out, const void *tw, const void *post, const float *x, long in_dist,
                float *out_re, float *out_im, long out_dist, long count, float scale, int grid,
                void *stream);
int fafk_rs_c2r(int inter_in, const void *tw, const void *pre, const float *in_re,
                const float *in_im, long in_dist, float *y, long out_dist, long count,
                float scale, int grid, void *stream);
int fafk_rs_c2c(int inv, int inter, const void *tw, const float *in_re, const float *in_im,
                long in_dist, float *out_re, float *out_im, long out_dist, long count,
                float scale, int grid, void *stream);
int fafk_capture_begin(void *stream);
void *fafk_capture_end(void *stream, int *nodes);
int fafk_graph_launch(void *exec, void *stream);
void fafk_graph_destroy(void *exec);
int fafk_fill_u32(unsigned *p, int n, unsigned v, void *stream);
int fafk_copy16(void *dst, const void *src, size_t bytes, int grid, int streaming,
                void *stream);
int fafk_window_inter(const float *x, long in_dist, const float *w, long n, float *y,
                      long count, void *stream);

int fafk_c2c(const fafd_fft *P, int inv, int inter, int spb, const float *in_re,
             const float *in_im, long in_dist, float *out_re, float *out_im, long out_dist,
             long count, float scale, void *stream);
int fafk_r2c(const fafd_fft *P, int inter_out, int spb, const void *post, const float *x,
             long in_dist, float *out_re, float *out_im, long out_dist, long count,
             float scale, void *stream);
int fafk_c2r(const fafd_fft *P, int inter_in, int spb, const void *pre, const float *in_re,
             const float *in_im, long in_dist, float *y, long out_dist, long count,
             float scale, void *stream);
int fafk_pipe(const fafd_fft *Pf, const fafd_fft *Pi, int spb, const void *post,
              const void *pre, const fafd_ops *ops, const float *x, long in_dist, float *y,
              long out_dist, long count, float scale, void *stream);
int fafk_stft(const fafd_fft *P, int spb, const void *post, const float *win, const float *x,
              long rec_len, long rec_dist, long hop, long frames, float *out_re,
              float *out_im, long frame_dist, long rec_out_dist, long count, void *stream);
int fafk_cwt(const fafd_fft *Ph, const fafd_fft *Pn, int real_out, const void *post,
             const float *psi, long psi_stride, int rows, int rows_per_wg, const float *x,
             long in_dist, float *out, long out_dist, long count, float scale, void *stream);
int fafk_icwt(const fafd_fft *Pf, const fafd_fft *Pi, const void *post, const void *pre,
              const float *dual, long dual_stride, int rows, const float *w, long in_dist,
              float *y, long out_dist, long count, float scale, void *stream);
/* Rader / Bluestein beyond one workgroup: element-wise steps around the
 * inner FFTs (split planes, dists in elements) */
int fafk_big_gather(const float *xr, const float *xi, long xdist, const int32_t *rin, long m,
                    float *ar, float *ai, long adist, long count, void *stream);
int fafk_big_chirp_in(const float *xr, const float *xi, long xdist, const void *chirp, long n,
                      long M, float *ar, float *ai, long adist, long count, void *stream);
int fafk_big_mul(float *ar, float *ai, long adist, const void *B, long m, void *save0,
                 long count, void *stream);
int fafk_big_rader_out(const float *ar, const float *ai, long adist, const int32_t *rout,
                       long m, const float *xr, const float *xi, long xdist, const void *save0,
                       float *yr, float *yi, long ydist, float scale, long count, void *stream);
int fafk_big_chirp_out(const float *ar, const float *ai, long adist, const void *chirp, long n,
                       float *yr, float *yi, long ydist, float scale, long count, void *stream);
/* DCT / DST / MDCT: the steps around the complex FFT (faf_r2r.c) */
int fafk_r2r_pre(const float *x, long xdist, long nin, const void *pre, int shift, long L,
                 float *ur, float *ui, long udist, long count, void *stream);
int fafk_r2r_post(const float *ur, const float *ui, long udist, const void *post, int off,
                  int use_im, long nout, float *y, long ydist, long count, void *stream);
/* long-segment dual inverse CWT: acc = sum_r Y_r * dual_r over interleaved
 * row spectra Y [count][rows][nb]; acc interleaved [count][nb] */
int fafk_icwt_mulacc(const float *Y, long nb, int rows, const float *dual, long stride,
                     float *acc, long count, void *stream);
/* L1 one-integral inverse CWT over real rows [lowpass?][n_scales wavelet rows] */
int fafk_icwt_l1(const float *w, long in_dist, float *y, long out_dist, long n, int n_scales,
                 int has_lp, float dlog, float c, long count, void *stream);
int fafk_fs_col(const fafd_fft *P1, int inv, int in_kind, int n2, int cols, const void *twn,
                const float *in_re, const float *in_im, long in_dist, float *t_re,
                float *t_im, long t_dist, long count, void *stream);
int fafk_fs_row(const fafd_fft *P2, int inv, int out_kind, int n1, int rows,
                const float *t_re, const float *t_im, long t_dist, float *out_re,
                float *out_im, long out_dist, long count, float scale, void *stream);
int fafk_r2c_post(int nc, const void *post, int inter_out, const float *z_re,
                  const float *z_im, long z_dist, float *out_re, float *out_im,
                  long out_dist, long count, float scale, void *stream);
int fafk_c2r_pre(int nc, const void *pre, int inter_in, const float *in_re,
                 const float *in_im, long in_dist, float *f_re, float *f_im, long f_dist,
                 long count, void *stream);
int fafk_rader(const fafd_fft *Pf, const fafd_fft *Pi, const fafd_rader *R, int spb,
               const float *in_re, const float *in_im, long in_dist, float *out_re,
               float *out_im, long out_dist, long count, float scale, void *stream);
int fafk_blue(const fafd_fft *Pf, const fafd_fft *Pi, const fafd_blue *Bq, int spb,
              const float *in_re, const float *in_im, long in_dist, float *out_re,
              float *out_im, long out_dist, long count, float scale, void *stream);
int fafk_rg_init(void);   /* faf_kernels_rg.hip */
const char *fafk_rg_last_error(void);
int fafk_rg_ct_slots(int n);
/* compile-time sequences of the ragged launch: threads per signal of
 * length n (Rader: prime n) in variant v, 0 if none; radices into order */
int fafk_rg_ct_seq(int n, int v, int *order, int *nr);
int fafk_rg_rader_seq(int p, int v, int *order, int *nr);
int fafk_rg_path(int kind, int n, int v);
int fafk_ragged(const void *plans, const void *units, int nunits, const long *uoff,
                const float *in_re, const float *in_im, float *out_re, float *out_im,
                unsigned *counter, int grid, int block, size_t lds, int generic, int persist,
                void *stream);
int fafk_ragged_counter_init(unsigned *counter, int grid, void *stream);

int fafk_serve(const void *plans, const void *units, int nunits, const void *sigs,
               unsigned *counter, int grid, int block, size_t lds, void *stream);
int fafk_cwt_spread(const float *X, long x_dist, const float *psi, long psi_stride, int nc,
                    int analytic, float *out, long row_len, long seg_dist, int rows,
                    long count, void *stream);   /* dists in complex elements */
void *fafk_host_alloc(size_t bytes);     /* pinned */
const void *fafk_plan_record(const fafd_fft *P);   /* device copy, uploaded once */
void fafk_host_free(void *p);

/* compile-time specialised kernels (faf_kernels_ct.hip): power-of-two
 * inner lengths 256 ... 16384 */
int fafk_ct_init(void);
int fafk_ct_supports(int kind, int n);
const char *fafk_ct_last_error(void);
int fafk_ct_c2c(const fafd_fft *P, int inv, int inter, const float *in_re, const float *in_im,
                long in_dist, float *out_re, float *out_im, long out_dist, long count,
                float scale, void *stream);
int fafk_ct_r2c(const fafd_fft *P, int inter_out, const void *post, const float *x,
                long in_dist, float *out_re, float *out_im, long out_dist, long count,
                float scale, void *stream);
int fafk_ct_c2r(const fafd_fft *P, int inter_in, const void *pre, const float *in_re,
                const float *in_im, long in_dist, float *y, long out_dist, long count,
                float scale, void *stream);
int fafk_ct_pipe(const fafd_fft *Pf, const fafd_fft *Pi, const void *post, const void *pre,
                 const fafd_ops *ops, const float *x, long in_dist, float *y, long out_dist,
                 long count, float scale, void *stream);
int fafk_ct_cwt(const fafd_fft *Ph, const fafd_fft *Pn, int real_out, const void *post,
                const float *psi, long psi_stride, int rows, int rows_per_wg, const float *x,
                long in_dist, float *out, long out_dist, long count, float scale, void *stream);
int fafk_ct_stft(const fafd_fft *P, const void *post, const float *win, const float *x,
                 long rec_len, long rec_dist, long hop, long frames, float *out_re,
                 float *out_im, long frame_dist, long rec_out_dist, long count, void *stream);

/* specialised four-step (n1 x n2 = 128 x 256 or 256 x 256): kinds of the
 * engine's F4_IN_* / F4_OUT_* enums */
enum { FAFK_F4_IN_PAIR = 0, FAFK_F4_IN_SPLIT = 1, FAFK_F4_IN_INTER = 2, FAFK_F4_PRE_SPLIT = 3,
       FAFK_F4_PRE_INTER = 4 };
enum { FAFK_F4_OUT_SPLIT = 0, FAFK_F4_OUT_INTER = 1, FAFK_F4_OUT_PAIR = 2,
       FAFK_F4_POST_SPLIT = 3, FAFK_F4_POST_INTER = 4 };
int fafk_ct4_supported(int n1, int n2);
int fafk_ct4(int in_kind, int out_kind, int inv, const fafd_fft *P1, const fafd_fft *P2,
             const void *twn, const void *pre, const void *post, const float *in_re,
             const float *in_im, long in_dist, void *ws, long ws_dist, float *out_re,
             float *out_im, long out_dist, long count, float scale, void *stream);

/* four-step I/O kinds (match the engine's enums) */
enum { FAFK_FS_IN_SPLIT = 0, FAFK_FS_IN_PAIR = 1, FAFK_FS_IN_INTER = 2 };
enum { FAFK_FS_OUT_SPLIT = 0, FAFK_FS_OUT_PAIR = 1, FAFK_FS_OUT_INTER = 2 };

/* ---- the FP64 variant (faf_kernels64.hip) ----
 * strided accessor: element i of transform s at (s / G) * gdist +
 * (s % G) * tdist + i * estride complex elements; im == NULL: interleaved
 * double2 at re */
typedef struct fafk_acc64 {
    const void *re, *im;
    long G, gdist, tdist, estride;
} fafk_acc64;
const char *fafk64_last_error(void);
int fafk64_init(void);
int fafk64_threads(int n);
/* count transforms of length n (<= 8192) of the plan record Pdev (device
 * copy) with double2 twiddles tw; lane_major interleaves transforms lane by
 * lane (strided columns); twn (optional) multiplies output k of transform s
 * by twn[(k * (s % out->G)) % twN] (four-step column pass) */
int fafk64_fft(const void *Pdev, int n, const void *tw, int inv, long count, int lane_major,
               const fafk_acc64 *in, const fafk_acc64 *out, const void *twn, long twN,
               double scale, void *stream);
int fafk64_post(const void *Z, long zdist, const void *tw, long nc, double *xr, double *xi,
                long xdist, double scale, long count, void *stream);
int fafk64_pre(const double *xr, const double *xi, long xdist, const void *tw, long nc, void *Z,
               long zdist, long count, void *stream);
int fafk64_in(const double *xr, const double *xi, long xdist, const int32_t *idx,
              const void *chirp, long n, long L, double *ar, double *ai, long adist, long count,
              void *stream);
int fafk64_mul(double *ar, double *ai, long adist, const void *B, long m, void *save0,
               long count, void *stream);
int fafk64_out(const double *ar, const double *ai, long adist, const int32_t *idx,
               const void *chirp, long len, const double *xr, const double *xi, long xdist,
               const void *save0, double *yr, double *yi, long ydist, double scale, long count,
               void *stream);
int fafk64_r2r_pre(const double *x, long xdist, long nin, const void *pre, int shift, long L,
                   double *ur, double *ui, long udist, long count, void *stream);
int fafk64_r2r_post(const double *ur, const double *ui, long udist, const void *post, int off,
                    int use_im, long nout, double *y, long ydist, long count, void *stream);

#ifdef __cplusplus
}
#endif
#endif
