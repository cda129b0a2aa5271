/*
 * fastandfourier.h -- public C ABI of the MI355X transform library.
 *
 * Source-compatible with FastAndFourier 1.2.0 for the transform execution
 * path: the enum values, struct field order and the static-inline helpers
 * below match include/fastandfourier.h of the reference so that a program
 * written against that header compiles and links against this library
 * unchanged.  The transform-execution subset is implemented: FFT, RFFT,
 * STFT, CWT/ICWT, the spectral pipeline, and the DCT/DST/MDCT family built
 * on the same device FFT.  The DWT creator is declared for ABI completeness
 * and returns NULL with an error message.
 *
 * Execution runs on the GPU (gfx950).  faf_execute accepts host or device
 * pointers: device pointers are used in place, host buffers are staged
 * through device memory.  faf_execute is synchronous unless
 * faf_amd_set_sync(0) is in effect (see faf_amd.h for the additive batch,
 * ragged, framed-STFT, complex-CWT and serving entry points).
 */
#ifndef FASTANDFOURIER_H
#define FASTANDFOURIER_H

#ifdef __cplusplus
extern "C" {
#endif

#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>

#ifndef FAF_RESTRICT
#  if defined(__cplusplus)
#    define FAF_RESTRICT __restrict__
#  else
#    define FAF_RESTRICT restrict
#  endif
#endif

#define FASTANDFOURIER_VERSION_MAJOR 1
#define FASTANDFOURIER_VERSION_MINOR 2
#define FASTANDFOURIER_VERSION_PATCH 0
#define FASTANDFOURIER_VERSION_STRING "1.2.0"

#define FAF_ALIGN 64

/* Precision (include/fastandfourier.h:116-125 of the reference). */
typedef enum {
    FAF_PREC_FP8 = 0, FAF_PREC_FP16 = 1, FAF_PREC_BF16 = 2, FAF_PREC_FP32 = 3,
    FAF_PREC_FP64 = 4, FAF_PREC_INT8 = 5, FAF_PREC_INT16 = 6, FAF_PREC_INT32 = 7
} faf_precision;

/* Transform type (reference include/fastandfourier.h:130-154). */
typedef enum {
    FAF_TRANSFORM_FFT, FAF_TRANSFORM_IFFT, FAF_TRANSFORM_RFFT, FAF_TRANSFORM_IRFFT,
    FAF_TRANSFORM_DCT_I, FAF_TRANSFORM_DCT_II, FAF_TRANSFORM_DCT_III, FAF_TRANSFORM_DCT_IV,
    FAF_TRANSFORM_DST_I, FAF_TRANSFORM_DST_II, FAF_TRANSFORM_DST_III, FAF_TRANSFORM_DST_IV,
    FAF_TRANSFORM_STFT, FAF_TRANSFORM_MDCT, FAF_TRANSFORM_IMDCT, FAF_TRANSFORM_HAAR,
    FAF_TRANSFORM_DAUBECHIES4, FAF_TRANSFORM_CDF53, FAF_TRANSFORM_CDF97, FAF_TRANSFORM_SYM4,
    FAF_TRANSFORM_PIPELINE, FAF_TRANSFORM_CWT, FAF_TRANSFORM_ICWT
} faf_transform_type;

/* DWT families and conventions: values only (reference :159-249). */
typedef enum {
    FAF_WAVELET_HAAR = 0, FAF_WAVELET_D4, FAF_WAVELET_CDF53, FAF_WAVELET_CDF97,
    FAF_WAVELET_SYM4, FAF_WAVELET_COUNT
} faf_wavelet_family;

typedef enum {
    FAF_CONV_UNSPEC = 0, FAF_CONV_HAAR_ORTHO, FAF_CONV_HAAR_LAZY, FAF_CONV_HAAR_MEAN,
    FAF_CONV_D4_ORTHO, FAF_CONV_SYM4_ORTHO, FAF_CONV_CDF53_INT, FAF_CONV_CDF53_ENERGY,
    FAF_CONV_CDF97_JPEG, FAF_CONV_CDF97_ORTHO, FAF_CONV_CUSTOM_PR, FAF_CONV_ANALYSIS_ONLY
} faf_wavelet_convention;

typedef enum {
    FAF_DWT_BACKEND_AUTO = 0, FAF_DWT_BACKEND_LIFT = 1, FAF_DWT_BACKEND_FIR = 2
} faf_dwt_backend;

/* CWT enums (reference :182-232). */
typedef enum {
    FAF_CWT_WAVELET_MORLET = 0, FAF_CWT_WAVELET_MORSE = 1, FAF_CWT_WAVELET_BUMP = 2,
    FAF_CWT_WAVELET_SHANNON = 3, FAF_CWT_WAVELET_MEYER = 4
} faf_cwt_wavelet;

typedef enum {
    FAF_CWT_NORM_L1 = 0, FAF_CWT_NORM_L2 = 1, FAF_CWT_NORM_BANDPASS = 2
} faf_cwt_norm_kind;

typedef enum { FAF_CWT_SCALE_GEOMETRIC = 0, FAF_CWT_SCALE_LINEAR = 1 } faf_cwt_scale_kind;
typedef enum { FAF_CWT_CENTER_PEAK = 0, FAF_CWT_CENTER_ENERGY = 1 } faf_cwt_center_kind;
typedef enum { FAF_CWT_INV_DUAL = 0, FAF_CWT_INV_L1 = 1 } faf_cwt_inverse_kind;

#define FAF_CWT_FLAG_INCLUDE_LOWPASS  0x0001u
#define FAF_CWT_FLAG_ALLOW_UNTILED    0x0002u
#define FAF_CWT_FLAG_FIXED_LOWPASS    0x0004u
#define FAF_CWT_FLAG_KEEP_SPECTRUM    0x0008u
#define FAF_CWT_FLAG_REAL_OUTPUT      0x0010u
#define FAF_CWT_FLAG_VALIDATE_STRICT  0x0020u

/* Layout, norm, direction, backend (reference :275-304). */
typedef enum {
    FAF_LAYOUT_DEFAULT = 0, FAF_LAYOUT_SPLIT, FAF_LAYOUT_HERMITIAN, FAF_LAYOUT_REAL,
    FAF_LAYOUT_INTERLEAVED
} faf_layout;

typedef enum {
    FAF_NORM_DEFAULT = 0, FAF_NORM_NONE, FAF_NORM_ORTHO, FAF_NORM_FORWARD, FAF_NORM_LAZY,
    FAF_NORM_JPEG2000
} faf_norm;

typedef enum { FAF_DIR_FORWARD = 0, FAF_DIR_INVERSE } faf_direction;
typedef enum { FAF_BACKEND_AUTO = 0, FAF_BACKEND_VM, FAF_BACKEND_JIT } faf_backend;

/* CWT create-time configuration (reference :314-336, same field order). */
typedef struct faf_cwt_config {
    size_t n; double fs; faf_precision precision; faf_backend backend;
    faf_cwt_wavelet wavelet; faf_cwt_norm_kind norm; faf_cwt_scale_kind scale_kind;
    faf_cwt_center_kind center_kind; unsigned voices; double fmin; double fmax;
    double morlet_mu; double morse_gamma; double morse_beta; double bump_center;
    double bump_width; double bandpass_peak; double lp_alpha; double lp_beta;
    double lp_floor; uint32_t flags;
} faf_cwt_config;

/* Littlewood-Paley report (reference :344-364, same field order). */
typedef struct faf_cwt_lp_report {
    size_t n; size_t n_bins; size_t n_scales; int has_lowpass; int passed;
    double alpha, beta; double max_abs_dev; double mean_abs_dev; double min_A, max_A;
    double frame_cond; size_t k_lo, k_hi; size_t hole_bins_dc; size_t hole_bins_nyq;
    double max_dc_wavelet; double max_wrap_energy; double admissibility_C;
    const double *A;
} faf_cwt_lp_report;

/* Shared create-time configuration (reference :372-388, same field order). */
typedef struct faf_config {
    size_t n; faf_precision precision; faf_layout layout; faf_norm norm;
    faf_direction dir; faf_backend backend; uint32_t flags;
    faf_wavelet_family family; size_t levels; faf_wavelet_convention conv;
    faf_dwt_backend dwt_backend; int dct_type; size_t hop_length; size_t win_length;
} faf_config;

/* Execute-time buffer descriptor (reference :399-404). */
typedef struct faf_buffer {
    void *re; void *im; size_t n; faf_layout layout;
} faf_buffer;

typedef struct { uint32_t packed; uint32_t a0; uint32_t a1; uint32_t a2; } faf_inst;

typedef struct {
    void *handle; void *kernel_fn; void *native_code; char *so_path; char *c_path;
} faf_jit_cache;

/*
 * The plan.  Field order follows the reference's public struct
 * (include/fastandfourier.h:475-506) so that callers reading t->n, t->type,
 * t->cfg, t->flags, t->inner or t->precision keep working.  The device plan
 * hangs off the trailing `amd` pointer; execute_func is always NULL here.
 */
typedef struct faf_transform {
    size_t n; size_t n_inst; faf_inst *code; void *twiddles[8]; size_t twiddle_sizes[8];
    faf_precision precision; faf_transform_type type; uint32_t flags;
    faf_jit_cache *jit_cache; size_t hop_length; size_t win_length; size_t levels;
    faf_wavelet_family family; faf_config cfg; void *scratch; size_t scratch_size;
    int (*execute_func)(const struct faf_transform *t, void *out_re, void *out_im,
                        const void *in_re, const void *in_im);
    struct faf_transform *inner; struct faf_transform *inner_inv;
    void *user_aux; void *user_aux_im; size_t user_aux_n;
    const float *custom_h; const float *custom_g; const float *custom_ht;
    const float *custom_gt; int custom_len_hg; int custom_len_syn;
    void *amd;   /* device plan (this library only) */
} faf_transform;

/* Transform flags (reference include/faf.h:59-67 and :1022-1038). */
#define FAF_FLAG_INVERSE     0x0001u
#define FAF_FLAG_NORMALIZE   0x0002u
#define FAF_FLAG_INPLACE     0x0004u
#define FAF_FLAG_REAL        0x0008u
#define FAF_FLAG_PARALLEL    0x0010u
#define FAF_FLAG_BLUESTEIN   (1u << 12)
#define FAF_FLAG_MEASURE     (1u << 13)

/* --- lifecycle, errors, names --- */
int faf_init(void);
void faf_cleanup(void);
const char *faf_version(void);
const char *faf_arch_name(void);
const char *faf_get_error(void);
void faf_clear_error(void);
void faf_set_error(const char *fmt, ...);
const char *faf_layout_name(faf_layout layout);
const char *faf_norm_name(faf_norm norm);
const char *faf_precision_name(faf_precision prec);
const char *faf_transform_name(faf_transform_type type);
size_t faf_precision_size(faf_precision prec);
size_t faf_get_alignment(void);

/* --- static-inline helpers (reference :564-657, same semantics) --- */
static inline void *faf_aligned_alloc(size_t size) {
    if (size == 0) return NULL;
    return aligned_alloc(64, (size + (size_t)63) & ~(size_t)63);
}
static inline void faf_aligned_free(void *p) { free(p); }

static inline faf_config faf_config_init(size_t n) {
    faf_config c;
    memset(&c, 0, sizeof(c));
    c.n = n;
    c.precision = FAF_PREC_FP32;   /* FP8 is 0, so the zero default is not FP32 */
    return c;
}
static inline faf_config faf_config_from(const faf_transform *t) {
    faf_config c;
    memset(&c, 0, sizeof(c));
    return t ? t->cfg : c;
}
static inline faf_config faf_config_inverse(faf_config cfg) {
    cfg.dir = (cfg.dir == FAF_DIR_INVERSE) ? FAF_DIR_FORWARD : FAF_DIR_INVERSE;
    return cfg;
}
static inline faf_buffer faf_buffer_real(void *data, size_t n) {
    faf_buffer b; memset(&b, 0, sizeof(b));
    b.re = data; b.n = n; b.layout = FAF_LAYOUT_REAL; return b;
}
static inline faf_buffer faf_buffer_split(void *re, void *im, size_t n) {
    faf_buffer b; memset(&b, 0, sizeof(b));
    b.re = re; b.im = im; b.n = n; b.layout = FAF_LAYOUT_SPLIT; return b;
}
static inline faf_buffer faf_buffer_interleaved(void *data, size_t n) {
    faf_buffer b; memset(&b, 0, sizeof(b));
    b.re = data; b.n = n; b.layout = FAF_LAYOUT_INTERLEAVED; return b;
}
static inline faf_buffer faf_buffer_hermitian(void *re, void *im, size_t n_bins) {
    faf_buffer b; memset(&b, 0, sizeof(b));
    b.re = re; b.im = im; b.n = n_bins; b.layout = FAF_LAYOUT_HERMITIAN; return b;
}

/* --- size policy (reference src/faf_core.c:374-429) --- */
size_t faf_spectrum_len(const faf_transform *t);
bool faf_is_size_supported(faf_transform_type type, size_t n);
size_t faf_get_recommended_size(faf_transform_type type, size_t min_size);

/* --- creation (reference src/faf_core.c:568-1158, src/faf_rfft.c:174) --- */
faf_transform *faf_create(faf_transform_type type, const faf_config *cfg);
faf_transform *faf_create_fft(const faf_config *cfg);
faf_transform *faf_create_rfft(const faf_config *cfg);
faf_transform *faf_create_stft(const faf_config *cfg);
faf_transform *faf_create_inverse(const faf_transform *fwd);
/* DCT / DST types I-IV (cfg->dct_type 1-4, 0 = II) and MDCT (n even, n -> n/2),
   REAL layout, FP32 or FP64; reference include/fastandfourier.h:703-718 */
faf_transform *faf_create_dct(const faf_config *cfg);
faf_transform *faf_create_dst(const faf_config *cfg);
faf_transform *faf_create_mdct(const faf_config *cfg);
faf_transform *faf_create_dwt(const faf_config *cfg);   /* not on this path: NULL */
void faf_destroy_transform(faf_transform *t);

/* --- CWT bank (reference src/faf_cwt.c:27-1079) --- */
faf_cwt_config faf_cwt_config_init(size_t n);
faf_transform *faf_create_cwt(const faf_cwt_config *cfg);
faf_transform *faf_create_icwt(const faf_cwt_config *cfg, faf_cwt_inverse_kind kind);
const faf_cwt_lp_report *faf_cwt_bank_report(const faf_transform *t);
size_t faf_cwt_n_scales(const faf_transform *t);
size_t faf_cwt_n_rows(const faf_transform *t);
size_t faf_cwt_n_bins(const faf_transform *t);
int faf_cwt_has_lowpass(const faf_transform *t);
int faf_cwt_freqs(const faf_transform *t, double *hz, size_t cap);
int faf_cwt_scales(const faf_transform *t, double *s, size_t cap);
void faf_cwt_report_fprint(FILE *fp, const faf_cwt_lp_report *r);
const float *faf_cwt_psi_f32(const faf_transform *t, size_t scale_index);
const double *faf_cwt_psi_f64(const faf_transform *t, size_t scale_index);
const float *faf_cwt_phi_f32(const faf_transform *t);
const double *faf_cwt_phi_f64(const faf_transform *t);

/* --- execution (reference src/faf_core.c:1298-1367) --- */
int faf_execute(const faf_transform *t, faf_buffer *out, const faf_buffer *in);
int faf_execute_f32(const faf_transform *t, float *FAF_RESTRICT out,
                    const float *FAF_RESTRICT in);

/* --- edge converters and Hermitian multiply (reference src/faf_core.c:96-220) --- */
int faf_deinterleave_f32(float *FAF_RESTRICT re, float *FAF_RESTRICT im,
                         const float *FAF_RESTRICT interleaved, size_t n);
int faf_interleave_f32(float *FAF_RESTRICT interleaved, const float *FAF_RESTRICT re,
                       const float *FAF_RESTRICT im, size_t n);
int faf_deinterleave_f64(double *FAF_RESTRICT re, double *FAF_RESTRICT im,
                         const double *FAF_RESTRICT interleaved, size_t n);
int faf_interleave_f64(double *FAF_RESTRICT interleaved, const double *FAF_RESTRICT re,
                       const double *FAF_RESTRICT im, size_t n);
void faf_herm_mul_f32(float *yr, float *yi, const float *xr, const float *xi,
                      const float *hr, const float *hi, size_t n_bins);
void faf_herm_mul_f64(double *yr, double *yi, const double *xr, const double *xi,
                      const double *hr, const double *hi, size_t n_bins);
void faf_herm_mul_conj_f32(float *yr, float *yi, const float *xr, const float *xi,
                           const float *hr, const float *hi, size_t n_bins);
void faf_herm_mul_conj_f64(double *yr, double *yi, const double *xr, const double *xi,
                           const double *hr, const double *hi, size_t n_bins);

/* --- window generator (reference src/faf_twiddles.c:159-163) --- */
void faf_gen_hann_window_f32(float *win, size_t n);

#ifdef __cplusplus
}
#endif
#endif /* FASTANDFOURIER_H */
