"""Write the golden fixtures of the transform path from the compiled reference.

Run in the build container, where ``oracle/_ref/libfaf_ref.so`` has been
compiled from the reference's own sources (``make -C oracle ref``):

    python tests/golden/gen_golden.py

Every fixture is data: seeded inputs and the reference library's FP32
outputs for them (SPLIT / HERMITIAN / REAL layouts, VM backend so the
reference's run-time C JIT is not involved).  The GPU tests compare the
device results with these outputs; the CPU tests compare the oracle's
restatement with them.

Inputs use numpy's PCG64 generator (``np.random.default_rng(seed)``),
uniform[-1, 1) for complex data and N(0, 1) for real data, as float32.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import fastandfourier_amd as fa  # noqa: E402  (struct layouts only)
import oracle as O  # noqa: E402

C2C_LENGTHS = [60, 360, 1000, 3000, 4800, 8192, 11, 17, 31, 97, 1009, 4096, 7, 12]
BLUESTEIN_LENGTHS = [23, 47, 307]
RFFT_LENGTHS = [1024, 4096, 16384]


def uniform_c(rng, count, n):
    re = rng.uniform(-1.0, 1.0, (count, n)).astype(np.float32)
    im = rng.uniform(-1.0, 1.0, (count, n)).astype(np.float32)
    return (re + 1j * im).astype(np.complex64)


def notch_mask(n):
    """Real mask zeroing bins [ceil(0.2 n/2), 0.25 n/2) = [ceil(n/10), n/8)."""
    nb = n // 2 + 1
    h = np.ones(nb, np.float32)
    h[-(-n // 10): n // 8] = 0.0
    return h


def ref_pipeline_notch(x, n):
    h = O.ref()
    t = h.chirp_compile(f"(pipeline (rfft :size {n}) (mul-spectrum) (irfft))".encode())
    assert t, h.faf_get_error()
    tp = C.cast(t, C.POINTER(fa.FafTransform))
    tp.contents.cfg.backend = 1          # VM: no run-time C compiler
    hr = notch_mask(n)
    hi = np.zeros_like(hr)
    assert h.chirp_bind(t, b"H", hr.ctypes.data, hi.ctypes.data, hr.size) == 0
    out = np.empty_like(x)
    for r in range(x.shape[0]):
        xr = np.ascontiguousarray(x[r])
        y = np.zeros(n, np.float32)
        bi = fa.buffer_real(xr.ctypes.data, n)
        bo = fa.buffer_real(y.ctypes.data, n)
        assert h.faf_execute(t, C.byref(bo), C.byref(bi)) == 0, h.faf_get_error()
        out[r] = y
    h.faf_destroy_transform(t)
    h.chirp_cleanup()
    return out


def ref_cwt(n, fs, voices=None, fmin=None, fmax=None):
    h = O.ref()
    cc = h.faf_cwt_config_init(n)
    cc.fs = fs
    if voices:
        cc.voices = voices
    if fmin:
        cc.fmin = fmin
    if fmax:
        cc.fmax = fmax
    cc.backend = 1
    t = h.faf_create_cwt(C.byref(cc))
    assert t, h.faf_get_error()
    rows = h.faf_cwt_n_rows(t)
    nb = h.faf_cwt_n_bins(t)
    ns = h.faf_cwt_n_scales(t)
    psi = np.zeros((rows, nb), np.float32)
    psi[0] = np.ctypeslib.as_array(h.faf_cwt_phi_f32(t), (nb,))
    for j in range(ns):
        psi[j + 1] = np.ctypeslib.as_array(h.faf_cwt_psi_f32(t, j), (nb,))
    freqs = np.zeros(ns)
    scales = np.zeros(ns)
    h.faf_cwt_freqs(t, freqs.ctypes.data_as(C.POINTER(C.c_double)), ns)
    h.faf_cwt_scales(t, scales.ctypes.data_as(C.POINTER(C.c_double)), ns)
    rep = h.faf_cwt_bank_report(t).contents
    report = np.array([rep.n_scales, rep.passed, rep.max_abs_dev, rep.mean_abs_dev,
                       rep.min_A, rep.max_A, rep.frame_cond, rep.k_lo, rep.k_hi,
                       rep.max_dc_wavelet, rep.max_wrap_energy, rep.admissibility_C])
    return t, psi, freqs, scales, report


def ref_icwt_dual(n, fs, rows_in, voices=None, fmin=None, fmax=None):
    """The reference's dual-frame inverse (FAF_CWT_INV_DUAL) of real rows."""
    h = O.ref()
    cc = h.faf_cwt_config_init(n)
    cc.fs = fs
    if voices:
        cc.voices = voices
    if fmin:
        cc.fmin = fmin
    if fmax:
        cc.fmax = fmax
    cc.backend = 1
    ti = h.faf_create_icwt(C.byref(cc), 0)
    assert ti, h.faf_get_error()
    w = np.ascontiguousarray(rows_in, dtype=np.float32)
    y = np.zeros(n, np.float32)
    bi = fa.buffer_real(w.ctypes.data, n)
    bo = fa.buffer_real(y.ctypes.data, n)
    assert h.faf_execute(ti, C.byref(bo), C.byref(bi)) == 0, h.faf_get_error()
    h.faf_destroy_transform(ti)
    return y


def ref_cwt_rows(t, x):
    h = O.ref()
    n = x.size
    rows = h.faf_cwt_n_rows(t)
    w = np.zeros(rows * n, np.float32)
    bi = fa.buffer_real(x.ctypes.data, n)
    bo = fa.buffer_real(w.ctypes.data, rows * n)
    assert h.faf_execute(t, C.byref(bo), C.byref(bi)) == 0, h.faf_get_error()
    return w.reshape(rows, n)


def add_dual_and_serving_fixtures(fx):
    """Fixtures added after the first set, from their own generator so the
    earlier ones keep their bytes:
      * the reference's dual-frame inverse of the n=512 default-bank rows,
        and of the C4 bank's rows of one 4096-sample segment;
      * the serving bank (n=1024, fs=16 kHz, 8 voices, 32 scales from
        fmin = 16 fs/n) tables, frequencies and report;
      * the reference's single-frame windowed STFT (interleaved, VM path)
        for n=1000, win 800."""
    rng = np.random.default_rng(100)
    h = O.ref()
    fx["cwt512_dual_out"] = ref_icwt_dual(512, 1.0, fx["cwt512_rows"])
    f0 = 62.5
    t, psi, _, _, _ = ref_cwt(4096, 16000.0, 11, f0, f0 * 2 ** (63.5 / 11))
    x = rng.standard_normal(4096).astype(np.float32)
    w = ref_cwt_rows(t, x)
    h.faf_destroy_transform(t)
    fx["cwt4096_seg_in"] = x
    fx["cwt4096_seg_rows"] = w
    fx["cwt4096_dual_out"] = ref_icwt_dual(4096, 16000.0, w, 11, f0, f0 * 2 ** (63.5 / 11))
    n, fs = 1024, 16000.0
    fmin = 16 * fs / n
    t, psi, freqs, scales, rep = ref_cwt(n, fs, 8, fmin, fmin * 2 ** (31.5 / 8))
    h.faf_destroy_transform(t)
    fx["cwt1024v8_psi"] = psi
    fx["cwt1024v8_freqs"] = freqs
    fx["cwt1024v8_report"] = rep
    n, win = 1000, 800
    cfg = fa.config_init(n, layout="interleaved", win_length=win, backend=1)
    ts = h.faf_create_stft(C.byref(cfg))
    assert ts, h.faf_get_error()
    z = uniform_c(rng, 1, n)[0]
    zin = np.ascontiguousarray(z.view(np.float32))
    zout = np.zeros(2 * n, np.float32)
    assert h.faf_execute_vm(ts, zout.ctypes.data, zin.ctypes.data) == 0, h.faf_get_error()
    h.faf_destroy_transform(ts)
    fx["stft1_1000_in"] = z
    fx["stft1_1000_out"] = zout.view(np.complex64)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libfaf_ref.so missing: run make -C oracle ref")
    rng = np.random.default_rng(0)
    fx = {}
    for n in C2C_LENGTHS:
        cnt = 2 if n <= 1009 else 1
        x = uniform_c(rng, cnt, n)
        fx[f"c2c_{n}_in"] = x
        fx[f"c2c_{n}_fwd"] = O.ref_fft(x)
        fx[f"c2c_{n}_inv"] = O.ref_fft(x, inverse=True)
    for n in BLUESTEIN_LENGTHS:
        x = uniform_c(rng, 2, n)
        fx[f"blue_{n}_in"] = x
        fx[f"blue_{n}_fwd"] = O.ref_fft(x, bluestein=True)
    for n in RFFT_LENGTHS:
        x = rng.standard_normal((2, n)).astype(np.float32)
        X = O.ref_rfft(x)
        fx[f"rfft_{n}_in"] = x
        fx[f"rfft_{n}_fwd"] = X
        fx[f"rfft_{n}_inv"] = O.ref_irfft(X, n)
        fx[f"rfft_{n}_ortho"] = O.ref_rfft(x, norm="ortho")
    n = 4096
    x = rng.standard_normal((2, n)).astype(np.float32)
    fx["notch_4096_in"] = x
    fx["notch_4096_out"] = ref_pipeline_notch(x, n)
    fx["notch_4096_mask"] = notch_mask(n)
    # CWT banks: the C4 bank (n=4096, fs=16 kHz, 11 voices, 64 scales) and the
    # library default bank at n=512, plus the default bank's forward rows
    f0 = 62.5
    t, psi, freqs, scales, rep = ref_cwt(4096, 16000.0, 11, f0, f0 * 2 ** (63.5 / 11))
    O.ref().faf_destroy_transform(t)
    fx["cwt4096_psi"] = psi
    fx["cwt4096_freqs"] = freqs
    fx["cwt4096_scales"] = scales
    fx["cwt4096_report"] = rep
    t, psi, freqs, scales, rep = ref_cwt(512, 1.0)
    x = rng.standard_normal(512).astype(np.float32)
    h = O.ref()
    rows = psi.shape[0]
    w = np.zeros(rows * 512, np.float32)
    bi = fa.buffer_real(x.ctypes.data, 512)
    bo = fa.buffer_real(w.ctypes.data, rows * 512)
    assert h.faf_execute(t, C.byref(bo), C.byref(bi)) == 0, h.faf_get_error()
    h.faf_destroy_transform(t)
    fx["cwt512_psi"] = psi
    fx["cwt512_freqs"] = freqs
    fx["cwt512_scales"] = scales
    fx["cwt512_report"] = rep
    fx["cwt512_in"] = x
    fx["cwt512_rows"] = w.reshape(rows, 512)
    # the reference's L1 one-integral inverse of those rows (the default bank
    # is L1-normalised), with the bank's voice count
    cc = h.faf_cwt_config_init(512)
    cc.fs = 1.0
    cc.backend = 1
    ti = h.faf_create_icwt(C.byref(cc), 1)
    assert ti, h.faf_get_error()
    y = np.zeros(512, np.float32)
    bi = fa.buffer_real(w.ctypes.data, 512)
    bo = fa.buffer_real(y.ctypes.data, 512)
    assert h.faf_execute(ti, C.byref(bo), C.byref(bi)) == 0, h.faf_get_error()
    h.faf_destroy_transform(ti)
    fx["cwt512_l1_out"] = y
    fx["cwt512_voices"] = np.array([cc.voices])
    add_dual_and_serving_fixtures(fx)
    out = os.path.join(HERE, "golden.npz")
    np.savez_compressed(out, **fx)
    print(f"wrote {out}: {len(fx)} arrays, {os.path.getsize(out) / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
