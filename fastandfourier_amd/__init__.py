"""Python view of the MI355X transform library.

The product is the C ABI in ``include/fastandfourier.h`` / ``include/faf_amd.h``
implemented by ``libfastandfourier_amd.so`` (C11 host code plus gfx950 HIP
kernels).  This module only binds that ABI with ctypes so Python callers and
the test-suite can drive it; it computes nothing itself.  Buffers are passed
as raw pointers: device pointers (e.g. ``torch.Tensor.data_ptr()`` of a CUDA
tensor) are used in place, host pointers are staged by the library.

Call sites mirror the reference's C usage (reference include/fastandfourier.h):
``config_init`` -> ``create_*`` -> ``execute`` -> ``destroy``.
"""
from __future__ import annotations

import ctypes as C
import os
from ctypes import POINTER, Structure, byref, c_char_p, c_double, c_float, c_int
from ctypes import c_size_t, c_uint32, c_void_p

__all__ = [
    "lib", "load", "FafConfig", "FafBuffer", "FafCwtConfig", "FafLpReport", "FafTransform",
    "config_init", "buffer_real", "buffer_split", "buffer_hermitian", "buffer_interleaved",
    "FafError", "check", "LAYOUT", "NORM", "TYPE", "PREC", "FLAG_BLUESTEIN", "EXPORTED",
    "FafServerConfig", "FafServerStats", "REQ", "REQ_DISCARD",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfastandfourier_amd.so")

# enum values of the reference ABI (include/fastandfourier.h:116-304)
PREC = {"fp8": 0, "fp16": 1, "bf16": 2, "fp32": 3, "fp64": 4}
TYPE = {"fft": 0, "ifft": 1, "rfft": 2, "irfft": 3, "stft": 12, "pipeline": 20, "cwt": 21,
        "icwt": 22}
LAYOUT = {"default": 0, "split": 1, "hermitian": 2, "real": 3, "interleaved": 4}
NORM = {"default": 0, "none": 1, "ortho": 2, "forward": 3}
DIR = {"forward": 0, "inverse": 1}
FLAG_BLUESTEIN = 1 << 12
FLAG_MEASURE = 1 << 13
CWT_FLAG_INCLUDE_LOWPASS = 0x1
CWT_FLAG_ALLOW_UNTILED = 0x2
CWT_FLAG_VALIDATE_STRICT = 0x20


class FafConfig(Structure):
    _fields_ = [("n", c_size_t), ("precision", c_int), ("layout", c_int), ("norm", c_int),
                ("dir", c_int), ("backend", c_int), ("flags", c_uint32), ("family", c_int),
                ("levels", c_size_t), ("conv", c_int), ("dwt_backend", c_int),
                ("dct_type", c_int), ("hop_length", c_size_t), ("win_length", c_size_t)]


class FafBuffer(Structure):
    _fields_ = [("re", c_void_p), ("im", c_void_p), ("n", c_size_t), ("layout", c_int)]


class FafCwtConfig(Structure):
    _fields_ = [("n", c_size_t), ("fs", c_double), ("precision", c_int), ("backend", c_int),
                ("wavelet", c_int), ("norm", c_int), ("scale_kind", c_int),
                ("center_kind", c_int), ("voices", C.c_uint), ("fmin", c_double),
                ("fmax", c_double), ("morlet_mu", c_double), ("morse_gamma", c_double),
                ("morse_beta", c_double), ("bump_center", c_double), ("bump_width", c_double),
                ("bandpass_peak", c_double), ("lp_alpha", c_double), ("lp_beta", c_double),
                ("lp_floor", c_double), ("flags", c_uint32)]


class FafLpReport(Structure):
    _fields_ = [("n", c_size_t), ("n_bins", c_size_t), ("n_scales", c_size_t),
                ("has_lowpass", c_int), ("passed", c_int), ("alpha", c_double),
                ("beta", c_double), ("max_abs_dev", c_double), ("mean_abs_dev", c_double),
                ("min_A", c_double), ("max_A", c_double), ("frame_cond", c_double),
                ("k_lo", c_size_t), ("k_hi", c_size_t), ("hole_bins_dc", c_size_t),
                ("hole_bins_nyq", c_size_t), ("max_dc_wavelet", c_double),
                ("max_wrap_energy", c_double), ("admissibility_C", c_double),
                ("A", POINTER(c_double))]


class FafTransform(Structure):
    pass


FafTransform._fields_ = [
    ("n", c_size_t), ("n_inst", c_size_t), ("code", c_void_p), ("twiddles", c_void_p * 8),
    ("twiddle_sizes", c_size_t * 8), ("precision", c_int), ("type", c_int),
    ("flags", c_uint32), ("jit_cache", c_void_p), ("hop_length", c_size_t),
    ("win_length", c_size_t), ("levels", c_size_t), ("family", c_int), ("cfg", FafConfig),
    ("scratch", c_void_p), ("scratch_size", c_size_t), ("execute_func", c_void_p),
    ("inner", POINTER(FafTransform)), ("inner_inv", POINTER(FafTransform)),
    ("user_aux", c_void_p), ("user_aux_im", c_void_p), ("user_aux_n", c_size_t),
    ("custom_h", c_void_p), ("custom_g", c_void_p), ("custom_ht", c_void_p),
    ("custom_gt", c_void_p), ("custom_len_hg", c_int), ("custom_len_syn", c_int),
    ("amd", c_void_p)]

PT = POINTER(FafTransform)
PB = POINTER(FafBuffer)

# serving loop (include/faf_amd.h)
REQ = {"rfft": 0, "c2c": 1, "stft": 2, "cwt": 3}
REQ_DISCARD = 0x1


class FafServerConfig(Structure):
    _fields_ = [("hbm_budget", c_size_t), ("window_us", c_double), ("streams", c_int),
                ("max_batch", c_int), ("plain", c_int), ("stft_win", c_size_t),
                ("stft_hop", c_size_t), ("cwt_fs", c_double), ("cwt_voices", C.c_uint),
                ("cwt_octaves", c_double)]


class FafServerStats(Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("launches", C.c_uint64), ("mean_batch", c_double), ("p50_us", c_double),
                ("p99_us", c_double), ("max_us", c_double), ("mean_us", c_double),
                ("span_s", c_double), ("req_per_s", c_double), ("gb_per_s", c_double),
                ("peak_hbm_bytes", c_size_t), ("admission_waits", C.c_uint64),
                ("long_requests", C.c_uint64), ("host_us_per_batch", c_double),
                ("device_us_per_batch", c_double)]


PS = c_void_p

# every symbol the public headers declare, with its ctypes signature
_SIGS = {
    "faf_init": (c_int, []), "faf_cleanup": (None, []), "faf_version": (c_char_p, []),
    "faf_arch_name": (c_char_p, []), "faf_get_error": (c_char_p, []),
    "faf_clear_error": (None, []), "faf_set_error": (None, [c_char_p]),
    "faf_layout_name": (c_char_p, [c_int]), "faf_norm_name": (c_char_p, [c_int]),
    "faf_precision_name": (c_char_p, [c_int]), "faf_transform_name": (c_char_p, [c_int]),
    "faf_precision_size": (c_size_t, [c_int]), "faf_get_alignment": (c_size_t, []),
    "faf_spectrum_len": (c_size_t, [PT]), "faf_is_size_supported": (C.c_bool, [c_int, c_size_t]),
    "faf_get_recommended_size": (c_size_t, [c_int, c_size_t]),
    "faf_create": (PT, [c_int, POINTER(FafConfig)]),
    "faf_create_fft": (PT, [POINTER(FafConfig)]), "faf_create_rfft": (PT, [POINTER(FafConfig)]),
    "faf_create_stft": (PT, [POINTER(FafConfig)]), "faf_create_inverse": (PT, [PT]),
    "faf_create_dct": (PT, [POINTER(FafConfig)]), "faf_create_dst": (PT, [POINTER(FafConfig)]),
    "faf_create_mdct": (PT, [POINTER(FafConfig)]), "faf_create_dwt": (PT, [POINTER(FafConfig)]),
    "faf_destroy_transform": (None, [PT]),
    "faf_cwt_config_init": (FafCwtConfig, [c_size_t]),
    "faf_create_cwt": (PT, [POINTER(FafCwtConfig)]),
    "faf_create_icwt": (PT, [POINTER(FafCwtConfig), c_int]),
    "faf_cwt_bank_report": (POINTER(FafLpReport), [PT]),
    "faf_cwt_n_scales": (c_size_t, [PT]), "faf_cwt_n_rows": (c_size_t, [PT]),
    "faf_cwt_n_bins": (c_size_t, [PT]), "faf_cwt_has_lowpass": (c_int, [PT]),
    "faf_cwt_freqs": (c_int, [PT, POINTER(c_double), c_size_t]),
    "faf_cwt_scales": (c_int, [PT, POINTER(c_double), c_size_t]),
    "faf_cwt_report_fprint": (None, [c_void_p, POINTER(FafLpReport)]),
    "faf_cwt_psi_f32": (POINTER(c_float), [PT, c_size_t]),
    "faf_cwt_psi_f64": (POINTER(c_double), [PT, c_size_t]),
    "faf_cwt_phi_f32": (POINTER(c_float), [PT]), "faf_cwt_phi_f64": (POINTER(c_double), [PT]),
    "faf_execute": (c_int, [PT, PB, PB]),
    "faf_execute_f32": (c_int, [PT, c_void_p, c_void_p]),
    "faf_deinterleave_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "faf_interleave_f32": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "faf_deinterleave_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "faf_interleave_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "faf_herm_mul_f32": (None, [c_void_p] * 6 + [c_size_t]),
    "faf_herm_mul_f64": (None, [c_void_p] * 6 + [c_size_t]),
    "faf_herm_mul_conj_f32": (None, [c_void_p] * 6 + [c_size_t]),
    "faf_herm_mul_conj_f64": (None, [c_void_p] * 6 + [c_size_t]),
    "faf_gen_hann_window_f32": (None, [c_void_p, c_size_t]),
    # include/chirp.h
    "chirp_register_spectral": (c_int, [c_char_p, c_void_p, c_void_p]),
    "chirp_register_spectral_ex": (c_int, [c_char_p, c_void_p, c_void_p, c_void_p]),
    "chirp_bind": (c_int, [PT, c_char_p, c_void_p, c_void_p, c_size_t]),
    "chirp_cleanup": (None, []), "chirp_count": (c_int, []),
    "chirp_compile": (PT, [c_char_p]),
    # include/faf_amd.h
    "faf_amd_set_stream": (c_int, [c_void_p]), "faf_amd_get_stream": (c_void_p, []),
    "faf_amd_set_sync": (c_int, [c_int]), "faf_amd_synchronize": (c_int, []),
    "faf_amd_device_ready": (c_int, []),
    "faf_graph_begin": (c_int, []), "faf_graph_end": (c_void_p, []),
    "faf_graph_launch": (c_int, [c_void_p]), "faf_graph_nodes": (c_int, [c_void_p]),
    "faf_graph_destroy": (None, [c_void_p]),
    "faf_execute_batch": (c_int, [PT, PB, PB, c_size_t, c_size_t, c_size_t]),
    "faf_execute_batch_noalloc": (c_int, [PT, PB, PB, c_size_t, c_size_t, c_size_t]),
    "faf_ragged_create": (c_void_p, [POINTER(c_size_t), c_size_t, c_uint32]),
    "faf_ragged_total": (c_size_t, [c_void_p]),
    "faf_ragged_execute": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "faf_ragged_destroy": (None, [c_void_p]),
    "faf_ragged_autotune": (c_int, [c_void_p]),
    "faf_amd_copy_ceiling": (c_int, [c_void_p, c_void_p, c_size_t, c_int]),
    "faf_amd_copy_ceiling_ex": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_int]),
    "faf_ragged_choice": (c_int, [c_void_p, c_size_t, POINTER(c_int), POINTER(c_int),
                                  POINTER(c_int), POINTER(c_int)]),
    "faf_ragged_set_choice": (c_int, [c_size_t, c_int, c_int, POINTER(c_int), c_int]),
    "faf_ragged_persistent": (c_int, [c_void_p]),
    "faf_ragged_set_persistent": (c_int, [c_void_p, c_int]),
    "faf_ragged_block": (c_int, [c_void_p]),
    "faf_ragged_set_block": (c_int, [c_void_p, c_int]),
    "faf_amd_plan_shape": (c_int, [PT, POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                   POINTER(c_int)]),
    "faf_stft_frames": (c_size_t, [PT, c_size_t]),
    "faf_execute_stft": (c_int, [PT, c_void_p, c_void_p, c_void_p, c_size_t, c_size_t,
                                 c_size_t]),
    "faf_execute_cwt_complex": (c_int, [PT, c_void_p, c_void_p, c_size_t, c_size_t]),
    "faf_create_pipeline": (PT, [POINTER(FafConfig), c_int, POINTER(c_int),
                                 POINTER(c_uint32), POINTER(c_uint32)]),
    "faf_amd_autotune": (c_int, [PT, c_size_t]),
    "faf_amd_last_kernel_ms": (c_float, []), "faf_amd_set_timing": (c_int, [c_int]),
    "faf_amd_cwt_dual_f32": (POINTER(c_float), [PT, c_size_t]),
    "faf_amd_set_specialised": (c_int, [c_int]),
    "faf_server_config_init": (FafServerConfig, []),
    "faf_server_create": (PS, [POINTER(FafServerConfig)]),
    "faf_server_destroy": (None, [PS]),
    "faf_server_prepare": (c_int, [PS, c_int, c_size_t]),
    "faf_server_prepare_many": (c_int, [PS, c_void_p, c_void_p, c_size_t]),
    "faf_server_set_plain": (c_int, [PS, c_int]),
    "faf_server_set_window": (c_int, [PS, c_double]),
    "faf_server_result_bytes": (c_size_t, [PS, c_int, c_size_t]),
    "faf_server_submit": (C.c_int64, [PS, c_int, c_size_t, c_void_p, c_uint32]),
    "faf_server_poll": (c_int, [PS, C.c_int64]),
    "faf_server_wait": (c_int, [PS, C.c_int64]),
    "faf_server_result": (c_void_p, [PS, C.c_int64, POINTER(c_size_t)]),
    "faf_server_release": (c_int, [PS, C.c_int64]),
    "faf_server_drain": (c_int, [PS]),
    "faf_server_stats_get": (c_int, [PS, POINTER(FafServerStats)]),
    "faf_server_stats_reset": (None, [PS]),
    "faf_server_replay": (c_int, [PS, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                  c_uint32, c_void_p]),
}

EXPORTED = sorted(_SIGS)

_lib = None


class FafError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load the library (once) and attach prototypes.  Raises when absent:
    there is no fallback implementation."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("FAF_AMD_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise FafError(f"{p} is missing: build it with __graft_entry__.build()")
    h = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(h, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = h
    return h


def lib():
    return load()


def check(rc: int, what: str = "call"):
    if rc != 0:
        msg = lib().faf_get_error().decode(errors="replace")
        raise FafError(f"{what} failed ({rc}): {msg}")


def config_init(n: int, **kw) -> FafConfig:
    """Same semantics as the header's static inline faf_config_init: zero,
    n, precision FP32; keyword fields are then set by name (layout/norm/dir
    accept their names)."""
    c = FafConfig()
    C.memset(byref(c), 0, C.sizeof(c))
    c.n = n
    c.precision = PREC["fp32"]
    for k, v in kw.items():
        if k == "layout" and isinstance(v, str):
            v = LAYOUT[v]
        elif k == "norm" and isinstance(v, str):
            v = NORM[v]
        elif k == "dir" and isinstance(v, str):
            v = DIR[v]
        setattr(c, k, v)
    return c


def _buf(re, im, n, layout):
    b = FafBuffer()
    b.re = re or None
    b.im = im or None
    b.n = n
    b.layout = layout
    return b


def buffer_real(p, n):
    return _buf(p, None, n, LAYOUT["real"])


def buffer_split(re, im, n):
    return _buf(re, im, n, LAYOUT["split"])


def buffer_hermitian(re, im, nb):
    return _buf(re, im, nb, LAYOUT["hermitian"])


def buffer_interleaved(p, n):
    return _buf(p, None, n, LAYOUT["interleaved"])
