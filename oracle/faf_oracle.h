/*
 * faf_oracle.h -- CPU oracle for the transform execution path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker or as the timed CPU baseline.
 *
 * Two families of functions:
 *   orc_*_f64  -- FP64 ground truth: direct DFT with exact (j*k mod n)
 *                 phase indexing, and an FP64 mixed-radix FFT for large n.
 *   orc_*_f32  -- a C11 restatement of the reference's FP32 CPU algorithms
 *                 (codelets, split-radix DIF, mixed-radix DIT bytecode,
 *                 Rader, Bluestein, Kiss R2C/C2R, norms, Hermitian
 *                 multiply, pipeline ops, Hann window, CWT execution).
 * Every function cites the reference file:line it follows.
 */
#ifndef FAF_ORACLE_H
#define FAF_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* norm codes follow faf_norm: 1 NONE, 2 ORTHO, 3 FORWARD */

/* integer size policy (reference src/faf_fft_stage.c:25-135) */
int orc_is_5_smooth(size_t n);
int orc_is_7_smooth(size_t n);
int orc_is_prime(size_t n);
int orc_rader_eligible(size_t n);
size_t orc_primitive_root(size_t p);
int orc_factor_7smooth(size_t n, int *factors, int *n_factors);
size_t orc_digit_reverse(size_t i, const int *factors, int n_factors);
size_t orc_next_5_smooth(size_t n);
int orc_is_codelet_size(size_t n);
int orc_fft_size_legal(size_t n, int bluestein_flag);

/* FP64 ground truth */
void orc_dft_f64(size_t n, int inverse, const double *xr, const double *xi,
                 double *yr, double *yi);
void orc_dft_bin_f64(size_t n, int inverse, const double *xr, const double *xi,
                     size_t k, double *yr, double *yi);
int orc_fft_f64(size_t n, int inverse, const double *xr, const double *xi,
                double *yr, double *yi);

/* FP32 restatement of the reference C2C planner path, unnormalised core
 * plus the norm pass.  Returns 0 on success, -1 on an illegal size. */
int orc_fft_f32(size_t n, int inverse, int norm, int bluestein_flag,
                const float *xr, const float *xi, float *yr, float *yi);

/* Kiss R2C / C2R (reference src/faf_rfft.c:57-172, 337-395) */
int orc_rfft_f32(size_t n, int norm, const float *x, float *re, float *im);
int orc_irfft_f32(size_t n, int norm, const float *re, const float *im, float *y);

/* Hermitian multiply (reference src/faf_core.c:150-220) and pipeline ops
 * (reference src/chirp.c:421-471).  op codes: 1 conj, 2 mul, 3 bandpass. */
void orc_herm_mul_f32(float *yr, float *yi, const float *xr, const float *xi,
                      const float *hr, const float *hi, size_t nb, int conj_h);
int orc_pipeline_f32(size_t n, int nops, const int *ops, const uint32_t *a1,
                     const uint32_t *a2, const float *hr, const float *hi,
                     size_t h_bins, const float *x, float *y);

/* windows and framed STFT composed from rfft (reference
 * src/faf_twiddles.c:159-163 for the symmetric Hann) */
void orc_hann_f32(float *w, size_t n);
size_t orc_stft_frames(size_t len, size_t win, size_t hop);
int orc_stft_f32(size_t len, size_t win, size_t hop, const float *x,
                 float *re, float *im);
int orc_stft_f64(size_t len, size_t win, size_t hop, const float *x,
                 double *re, double *im);

/* CWT forward given the live filter table psi[rows][stride]
 * (reference src/faf_cwt.c:769-843 and src/faf_cwt_kernels.c:24-35):
 * real rows  W_r = irfft(rfft(x) * psi_r)            (FP32 restatement)
 * analytic rows z_r with Re z_r = W_r                  (FP64 truth)      */
int orc_cwt_real_f32(size_t n, size_t rows, const float *psi, size_t stride,
                     const float *x, float *w);
int orc_cwt_complex_f64(size_t n, size_t rows, const float *psi, size_t stride,
                        const float *x, double *zr, double *zi);

/* L1 one-integral inverse CWT (reference src/faf_cwt.c:878-915), FP32 as
 * there: y[i] = (sum over wavelet rows j in order of w_j[i] * (float)(ln2/V))
 * / (float)C, plus the lowpass row 0 when has_lp.  w holds has_lp + n_scales
 * rows of n samples.  C <= 0 is taken as 1. */
int orc_icwt_l1_f32(size_t n, size_t n_scales, int has_lp, unsigned voices, double c,
                    const float *w, float *y);

/* CPU baseline: batched rfft over `count` signals spread on `threads`
 * POSIX threads; returns wall seconds of the timed repeat. */
double orc_time_rfft_batch(size_t n, size_t count, int threads, const float *x,
                           float *re, float *im);
/* the same through the compiled reference library at path `so` (-1 on failure) */
double orc_time_ref_rfft_batch(const char *so, size_t n, size_t count, int threads,
                               int backend, const float *x, float *re, float *im);

#ifdef __cplusplus
}
#endif
#endif
