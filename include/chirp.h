/*
 * chirp.h -- spectral-pipeline subset of the Chirp API.
 *
 * Source-compatible with the reference's include/chirp.h for the calls the
 * fused rfft -> spectral ops -> irfft pipeline needs (chirp_compile,
 * chirp_bind, chirp_register_spectral(_ex), chirp_cleanup) with the same
 * kind enum, callback types and reserved op codes (reference
 * include/chirp.h:18-38).  The general Chirp language is not part of this
 * library.
 */
#ifndef CHIRP_H
#define CHIRP_H
#include "fastandfourier.h"
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CHIRP_KIND_UNARY = 0,
    CHIRP_KIND_OTHER = 1,
    CHIRP_KIND_COMPLEX,
    CHIRP_KIND_VECTOR,
    CHIRP_KIND_SPECTRAL,
    CHIRP_KIND_HERMITIAN
} chirp_kind;

typedef void (*chirp_spectral_fn)(float *re, float *im, size_t n_bins, void *ctx);
typedef void (*chirp_spectral_fn_f64)(double *re, double *im, size_t n_bins, void *ctx);

#define CHIRP_OP_CONJ      0x7ffffffdu
#define CHIRP_OP_MUL       0x7ffffffeu
#define CHIRP_OP_BANDPASS  0x7fffffffu

int chirp_register_spectral(const char *name, chirp_spectral_fn fn, void *ctx);
int chirp_register_spectral_ex(const char *name, chirp_spectral_fn f32,
                               chirp_spectral_fn_f64 f64, void *ctx);
int chirp_bind(faf_transform *t, const char *name, void *re, void *im, size_t n_bins);
void chirp_cleanup(void);
int chirp_count(void);
faf_transform *chirp_compile(const char *source);

#ifdef __cplusplus
}
#endif
#endif
