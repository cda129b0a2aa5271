/*
 * faf_amd.h -- additive entry points of the MI355X build.
 *
 * None of these exist in the reference (FastAndFourier 1.2.0 is
 * single-signal and single-threaded); they are new symbols, so programs
 * written against the reference are unaffected.  All buffers may be device
 * pointers (used in place) or host pointers (staged through device memory).
 */
#ifndef FAF_AMD_H
#define FAF_AMD_H
#include "fastandfourier.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stream and synchronisation control (thread-local) ----
 * faf_execute* launch on the calling thread's stream (a hipStream_t,
 * NULL = the default stream).  With sync on (the default) every execute
 * returns after its results are complete, as in the reference. */
int faf_amd_set_stream(void *hip_stream);
void *faf_amd_get_stream(void);
int faf_amd_set_sync(int on);
int faf_amd_synchronize(void);
int faf_amd_device_ready(void);              /* 0 if a gfx950 device is usable */

/* ---- batched execute ----
 * Runs `count` independent signals of the plan in one launch.  `in` and
 * `out` describe signal 0 exactly as for faf_execute; signal i starts
 * in_dist / out_dist elements (complex elements for INTERLEAVED) later in
 * every plane.  dist 0 means packed (the buffer length). */
int faf_execute_batch(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                      size_t count, size_t in_dist, size_t out_dist);

/* ---- ragged C2C batch in one persistent launch ----
 * Signals of different lengths are concatenated in split planes at
 * offsets = exclusive prefix sum of lengths.  Every length must be legal
 * for faf_create_fft (5-smooth, codelet, or Rader prime; others need
 * FAF_FLAG_BLUESTEIN in flags).  Forward, NumPy norm (unscaled). */
typedef struct faf_ragged faf_ragged;
faf_ragged *faf_ragged_create(const size_t *lengths, size_t count, uint32_t flags);
size_t faf_ragged_total(const faf_ragged *r);
int faf_ragged_execute(faf_ragged *r, float *out_re, float *out_im, const float *in_re,
                       const float *in_im);
void faf_ragged_destroy(faf_ragged *r);

/* ---- framed real STFT ----
 * faf_create_stft with cfg.layout == FAF_LAYOUT_REAL (an error in the
 * reference) builds a framed STFT: cfg.win_length (default n) samples per
 * frame, cfg.hop_length (default win/4), symmetric Hann window
 * (reference src/faf_twiddles.c:159-163), one rfft per frame, no centre
 * padding; a record shorter than the window is one zero-padded frame.
 * Output: re/im planes [records][frames][win/2+1]. */
size_t faf_stft_frames(const faf_transform *t, size_t len);
int faf_execute_stft(const faf_transform *t, float *out_re, float *out_im, const float *x,
                     size_t len, size_t records, size_t rec_dist);

/* ---- CWT with analytic complex64 rows ----
 * out: interleaved complex64 [count][n_rows][n], with Re(out row) equal to
 * the reference's real row W = irfft(rfft(x) * psi). */
int faf_execute_cwt_complex(const faf_transform *t, float *out, const float *x, size_t count,
                            size_t in_dist);

/* ---- fused spectral pipeline without the DSL ----
 * rfft(n) -> ops -> irfft(n) in one kernel.  op codes: 1 conj, 2 mul by the
 * spectrum bound with chirp_bind(t, "H", ...), 3 bandpass keeping bins
 * [a1[i], a2[i]].  Same semantics as reference src/chirp.c:421-471. */
faf_transform *faf_create_pipeline(const faf_config *cfg, int nops, const int *ops,
                                   const uint32_t *a1, const uint32_t *a2);

/* ---- autotuner ----
 * Times the candidate launch shapes of the plan on `count` synthetic
 * signals and keeps the fastest (cached per kind and length for the
 * process).  Returns the chosen signals-per-workgroup, or -1. */
int faf_amd_autotune(faf_transform *t, size_t count);

/* ---- kernel selection (thread-local, read at create time) ----
 * Power-of-two lengths 256..16384 (C2C) and 512..32768 (R2C/C2R, fused
 * pipeline, STFT window) use compile-time specialised kernels by default;
 * 0 makes later creates use the runtime-generic radix kernels instead. */
int faf_amd_set_specialised(int on);

/* ---- dual-frame filter row of a CWT/ICWT bank (packed Hermitian, n/2+1
 * bins; row 0 is the lowpass when present).  Valid until destroy. ---- */
const float *faf_amd_cwt_dual_f32(const faf_transform *t, size_t row);

/* ---- device-side timing of the last execute on this thread (ms) ---- */
float faf_amd_last_kernel_ms(void);
int faf_amd_set_timing(int on);

#ifdef __cplusplus
}
#endif
#endif
