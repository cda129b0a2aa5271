"""CPU tests: pin the oracle before it is trusted as the checker.

Pins used (each chosen so a plausible slip -- wrong sign, dropped term,
transposed index, wrong norm -- fails at least one):
  * the reference's own known-answer tests (size policy, factor lists,
    Hermitian multiply KATs, impulse/constant/tone answers);
  * numpy's FFT as an independent library routine for the FP64 ground truth;
  * the reference library compiled from its sources: golden outputs in
    tests/golden/golden.npz, and live calls when oracle/_ref is present;
  * closed forms and invariants (Parseval, linearity, shift, Hermitian
    symmetry, conj = time reversal, Re(analytic CWT row) = real row).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, tol

rng = np.random.default_rng(1234)


def cplx(n, k=1):
    return (rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n)))


# ---------------------------------------------------------------- size policy

def test_size_policy_known_answers():
    L = O.lib()
    # reference tests/test_mixed_radix.cpp:49-64, tests/test_sizes.cpp:56-115,
    # tests/test_rader.cpp:74-85, tests/test_core.cpp:106-116
    assert L.orc_is_5_smooth(1) and L.orc_is_5_smooth(12) and L.orc_is_5_smooth(4000)
    assert not L.orc_is_5_smooth(7) and not L.orc_is_5_smooth(14)
    assert L.orc_next_5_smooth(7) == 8 and L.orc_next_5_smooth(11) == 12
    assert L.orc_next_5_smooth(23) == 24
    fac = np.zeros(16, np.int32)
    nf = C.c_int()
    assert L.orc_factor_7smooth(12, fac, C.byref(nf)) == 0 and list(fac[:nf.value]) == [4, 3]
    assert L.orc_factor_7smooth(28, fac, C.byref(nf)) == 0
    assert fac[0] == 4 and fac[nf.value - 1] == 7
    assert L.orc_factor_7smooth(4800, fac, C.byref(nf)) == 0
    assert list(fac[:nf.value]) == [4, 4, 4, 3, 5, 5]
    for p in (11, 13, 17, 19, 29, 31, 37, 1009):
        assert L.orc_rader_eligible(p), p
    for p in (7, 23, 12):
        assert not L.orc_rader_eligible(p), p
    assert L.orc_is_codelet_size(7) and not L.orc_is_codelet_size(11)
    assert L.orc_fft_size_legal(14, 0) == 0 and L.orc_fft_size_legal(23, 0) == 0
    assert L.orc_fft_size_legal(4097, 0) == 0 and L.orc_fft_size_legal(4097, 1) == 3
    assert L.orc_fft_size_legal(1009, 0) == 2          # 1009 goes to Rader, not Bluestein
    # primitive roots checked by brute force on the group order
    for p in (11, 17, 31, 97, 1009):
        g = L.orc_primitive_root(p)
        seen = {pow(int(g), e, p) for e in range(p - 1)}
        assert len(seen) == p - 1


# ---------------------------------------------------------------- FP64 truth

@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 11, 12, 60, 97, 360, 1000, 1009, 4096])
def test_dft64_matches_numpy_and_closed_forms(n):
    x = cplx(n)[0]
    assert rel_l2(O.dft64(x), np.fft.fft(x)) < 1e-13
    assert rel_l2(O.dft64(x, inverse=True), np.fft.ifft(x) * n) < 1e-13
    imp = np.zeros(n, complex)
    imp[0] = 1
    assert np.allclose(O.dft64(imp), 1.0, atol=1e-12)               # impulse -> ones
    assert np.allclose(O.dft64(np.ones(n))[0], n)                  # constant -> n at DC
    k = n // 3
    tone = np.exp(2j * np.pi * k * np.arange(n) / n)
    y = O.dft64(tone)
    assert abs(y[k] - n) < 1e-9 * n and np.linalg.norm(np.delete(y, k)) < 1e-8 * n


@pytest.mark.parametrize("n", [16, 1009, 4096, 4800, 16384, 65536])
def test_fft64_matches_numpy(n):
    x = cplx(n)[0]
    assert rel_l2(O.fft64(x)[0], np.fft.fft(x)) < 1e-13
    # Parseval and the shift theorem pin the sign convention independently
    X = O.fft64(x)[0]
    assert abs(np.sum(abs(X) ** 2) / n - np.sum(abs(x) ** 2)) < 1e-9 * np.sum(abs(x) ** 2)
    Xs = O.fft64(np.roll(x, 1))[0]
    assert rel_l2(Xs, X * np.exp(-2j * np.pi * np.arange(n) / n)) < 1e-12


# ------------------------------------------- FP32 restatement of the reference

C2C = [2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 32, 60, 64, 360, 1000, 1024, 3000, 4096, 4800,
       8192, 11, 17, 31, 97, 1009]


@pytest.mark.parametrize("n", C2C)
def test_fft32_restatement_accuracy(n):
    x = cplx(n, 2).astype(np.complex64)
    truth = np.fft.fft(x.astype(np.complex128), axis=1)
    y = O.fft32(x)
    assert rel_l2(y, truth) < tol(n)
    yi = O.fft32(x, inverse=True)
    assert rel_l2(yi, np.fft.ifft(x.astype(np.complex128), axis=1)) < tol(n)
    yo = O.fft32(x, norm="ortho")
    assert rel_l2(yo, truth / np.sqrt(n)) < tol(n)
    yf = O.fft32(x, norm="forward")
    assert rel_l2(yf, truth / n) < tol(n)


@pytest.mark.parametrize("n", [23, 47, 307, 4097])
def test_bluestein_restatement(n):
    x = cplx(n).astype(np.complex64)
    with pytest.raises(ValueError):
        O.fft32(x)
    y = O.fft32(x, bluestein=True)
    assert rel_l2(y, np.fft.fft(x.astype(np.complex128))) < 2 * tol(n)


def test_oracle_matches_reference_golden(golden):
    """FP32 restatement vs the reference library's own FP32 output."""
    for key in golden.files:
        if key.startswith("c2c_") and key.endswith("_in"):
            n = int(key.split("_")[1])
            x = golden[key]
            assert rel_l2(O.fft32(x), golden[f"c2c_{n}_fwd"]) < tol(n), n
            assert rel_l2(O.fft32(x, inverse=True), golden[f"c2c_{n}_inv"]) < tol(n), n
        if key.startswith("blue_") and key.endswith("_in"):
            n = int(key.split("_")[1])
            assert rel_l2(O.fft32(golden[key], bluestein=True), golden[f"blue_{n}_fwd"]) < 2 * tol(n)


@pytest.mark.parametrize("n", [4, 8, 30, 64, 1000, 1024, 4096, 16384, 65536])
def test_rfft_restatement(n):
    x = rng.standard_normal((2, n)).astype(np.float32)
    X = np.fft.rfft(x.astype(np.float64), axis=1)
    assert rel_l2(O.rfft32(x), X) < tol(n)
    assert rel_l2(O.rfft32(x, norm="ortho"), X / np.sqrt(n)) < tol(n)
    assert rel_l2(O.irfft32(X.astype(np.complex64), n), x) < tol(n)
    # Hermitian DC / Nyquist bins are real
    Y = O.rfft32(x)
    assert np.all(Y[:, 0].imag == 0) and np.all(Y[:, -1].imag == 0)


def test_rfft_matches_reference_golden(golden):
    for n in (1024, 4096, 16384):
        x = golden[f"rfft_{n}_in"]
        assert rel_l2(O.rfft32(x), golden[f"rfft_{n}_fwd"]) < tol(n)
        assert rel_l2(O.rfft32(x, norm="ortho"), golden[f"rfft_{n}_ortho"]) < tol(n)
        assert rel_l2(O.irfft32(golden[f"rfft_{n}_fwd"], n), golden[f"rfft_{n}_inv"]) < tol(n)


def test_herm_mul_known_answers():
    # reference tests/test_layout.cpp:126-153
    f = lambda v: np.asarray(v, np.float32)
    yr = np.zeros(5, np.float32)
    yi = np.zeros(5, np.float32)
    O.lib().orc_herm_mul_f32(yr, yi, f([2, 1, 0, 3, 4]), f([7, 1, 0, 2, 9]), f([3, 1, 0, 0, 5]),
                             f([8, 0, 0, 1, 6]), 5, 0)
    assert (yr[0], yi[0], yr[4], yi[4], yr[1], yi[1]) == (6, 0, 20, 0, 1, 1)
    yr = np.zeros(3, np.float32)
    yi = np.zeros(3, np.float32)
    O.lib().orc_herm_mul_f32(yr, yi, f([1, 2, 3]), f([0, 4, 0]), f([5, 1, 7]), f([0, 1, 0]), 3, 1)
    assert list(yr) == [5, 6, 21] and list(yi) == [0, 2, 0]


def test_pipeline_semantics():
    n = 64
    x = rng.standard_normal((1, n)).astype(np.float32)
    X = np.fft.rfft(x[0].astype(np.float64))
    # conj in frequency = time reversal of a real signal (tests/test_spectral.cpp:144-161)
    y = O.pipeline32(x, n, [1])[0]
    assert np.allclose(y, np.r_[x[0, 0], x[0, :0:-1]], atol=1e-5)
    # bandpass keeps bins [lo, hi] inclusive
    y = O.pipeline32(x, n, [3], [4], [9])[0]
    keep = np.zeros_like(X)
    keep[4:10] = X[4:10]
    assert rel_l2(y, np.fft.irfft(keep, n)) < tol(n)
    # mul-spectrum by an identity kernel is the identity (tests/test_spectral.cpp:117-142)
    h = np.ones(n // 2 + 1, np.complex64)
    assert rel_l2(O.pipeline32(x, n, [2], h=h)[0], x[0]) < tol(n)


def test_pipeline_matches_reference_golden(golden):
    n = 4096
    x = golden["notch_4096_in"]
    y = O.pipeline32(x, n, [2], h=golden["notch_4096_mask"].astype(np.complex64))
    assert rel_l2(y, golden["notch_4096_out"]) < tol(n)
    # and the notch really removed the band
    X = np.fft.rfft(golden["notch_4096_out"][0].astype(np.float64))
    lo, hi = -(-n // 10), n // 8
    assert np.abs(X[lo:hi]).max() < 1e-3 * np.abs(X).max()


def test_hann_is_symmetric_hann():
    for n in (8, 512, 1024):
        assert np.allclose(O.hann32(n), np.hanning(n), atol=1e-6)


def test_stft_composition():
    x = rng.standard_normal(5000).astype(np.float32)
    win, hop = 512, 128
    S = O.stft64(x, win, hop)
    nf = (5000 - win) // hop + 1
    assert S.shape == (nf, win // 2 + 1)
    w = np.hanning(win)
    for f in (0, 7, nf - 1):
        ref = np.fft.rfft(x[f * hop:f * hop + win].astype(np.float64) * w)
        assert rel_l2(S[f], ref) < 1e-6
    assert rel_l2(O.stft32(x, win, hop), S) < tol(win)
    # a record shorter than the window is one zero-padded frame
    s = O.stft64(x[:100], win, hop)
    assert s.shape[0] == 1
    assert rel_l2(s[0], np.fft.rfft(np.r_[x[:100], np.zeros(win - 100)] * w)) < 1e-6


def test_cwt_oracle_against_reference_rows(golden):
    psi = golden["cwt512_psi"]
    rows, nb = psi.shape
    stride = (nb + 15) & ~15
    tab = np.zeros((rows, stride), np.float32)
    tab[:, :nb] = psi
    x = golden["cwt512_in"]
    w = O.cwt_real32(x, tab, stride)
    assert rel_l2(w, golden["cwt512_rows"]) < tol(512)
    z = O.cwt_complex64(x, tab, stride)
    assert rel_l2(z.real, golden["cwt512_rows"]) < tol(512)     # Re(analytic row) = real row
    # analytic: wavelet rows have no negative-frequency content
    Z = np.fft.fft(z[1:], axis=1)
    assert np.abs(Z[:, 512 // 2 + 1:]).max() < 1e-9 * np.abs(Z).max()


def test_icwt_l1_oracle_against_reference(golden):
    """The L1 one-integral inverse restatement reproduces the compiled
    reference's FP32 output on the default n=512 bank's forward rows."""
    w = golden["cwt512_rows"]
    C_ = golden["cwt512_report"][11]
    v = int(golden["cwt512_voices"][0])
    y = O.icwt_l1_f32(w, w.shape[0] - 1, True, v, C_)
    ref = golden["cwt512_l1_out"]
    assert rel_l2(y, ref) < 1e-6
    # each term matters: the lowpass row, the voice weight and C
    assert rel_l2(O.icwt_l1_f32(w[1:], w.shape[0] - 1, False, v, C_), ref) > 1e-2
    assert rel_l2(O.icwt_l1_f32(w, w.shape[0] - 1, True, v + 1, C_), ref) > 1e-2
    assert rel_l2(O.icwt_l1_f32(w, w.shape[0] - 1, True, v, 2 * C_), ref) > 1e-2


@pytest.mark.parametrize("n", [512, 4096])
def test_icwt_dual_oracle_against_reference(golden, n):
    """The oracle's FP64 dual-frame inverse against the compiled reference's
    FP32 dual inverse (FAF_CWT_INV_DUAL) of the reference's own rows: the
    default n=512 bank and the 64-scale bank at n=4096."""
    if n == 512:
        w, psi, ref = golden["cwt512_rows"], golden["cwt512_psi"], golden["cwt512_dual_out"]
    else:
        w, psi, ref = (golden["cwt4096_seg_rows"], golden["cwt4096_psi"],
                       golden["cwt4096_dual_out"])
    y = O.icwt_dual64(w, psi)
    assert rel_l2(y, ref) < tol(n)
    # the pieces matter: without the lowpass row, with an unregularised
    # dual, or with a row order mistake the result moves off the reference
    assert rel_l2(O.icwt_dual64(w[1:], psi[1:]), ref) > 1e-3
    assert rel_l2(O.icwt_dual64(w[::-1].copy(), psi), ref) > 1e-2
    # and it reconstructs the segment the rows came from (n=4096 input kept)
    if n == 4096:
        assert rel_l2(y, golden["cwt4096_seg_in"]) < 1e-4


def test_single_frame_stft_reference_is_windowed_fft(golden):
    """The reference's single-frame STFT (VM path, interleaved, n=1000,
    win 800) is the FFT of the input times a symmetric Hann window of 800
    samples, zero past it (src/faf_core.c:994-1032, 1369-1383)."""
    z = golden["stft1_1000_in"]
    n, win = z.size, 800
    w = np.zeros(n)
    w[:win] = O.hann32(win)
    truth = O.fft64(z.astype(np.complex128) * w)[0]
    assert rel_l2(golden["stft1_1000_out"], truth) < tol(n)
    assert rel_l2(golden["stft1_1000_out"], O.fft64(z.astype(np.complex128))[0]) > 0.1


def test_serving_bank_golden_invariants(golden):
    """The serving bank (n=1024, 8 voices, 32 scales from 16 fs/n) written by
    the reference: 32 scales at ratio 2^(1/8), frame condition < 1.2 and
    passed (reference tests/test_cwt_lp.cpp:67-90, 193-215)."""
    f = golden["cwt1024v8_freqs"]
    rep = golden["cwt1024v8_report"]
    assert f.size == 32 and golden["cwt1024v8_psi"].shape == (33, 513)
    assert np.allclose(f[:-1] / f[1:], 2 ** (1 / 8), rtol=1e-9) or \
        np.allclose(f[1:] / f[:-1], 2 ** (1 / 8), rtol=1e-9)
    assert rep[1] == 1 and rep[6] < 1.2


def test_icwt_l1_oracle_sine_round_trip(golden):
    """Reference tests/test_cwt.cpp:172-204: a sine at the bank's middle
    frequency through the forward CWT and the L1 inverse comes back within
    0.25 relative L2."""
    psi = golden["cwt512_psi"]
    rows, nb = psi.shape
    stride = (nb + 15) & ~15
    tab = np.zeros((rows, stride), np.float32)
    tab[:, :nb] = psi
    freqs = golden["cwt512_freqs"]
    f = freqs[freqs.size // 2]
    x = np.sin(2 * np.pi * f * np.arange(512)).astype(np.float32)
    w = O.cwt_real32(x, tab, stride)
    y = O.icwt_l1_f32(w, rows - 1, True, int(golden["cwt512_voices"][0]),
                      golden["cwt512_report"][11])
    assert rel_l2(y, x) <= 0.25


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
@pytest.mark.parametrize("n", [64, 360, 4800, 1009, 16384, 65536])
def test_oracle_matches_live_reference(n):
    x = cplx(n).astype(np.complex64)
    assert rel_l2(O.fft32(x), O.ref_fft(x)) < tol(n)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_cpu_baseline_timer_runs_the_reference():
    """The CPU-baseline timer drives the compiled reference (AUTO and VM
    backends) through its C API on several threads: its spectra are the
    reference's rfft, which matches FP64 and the restatement."""
    rng = np.random.default_rng(42)
    x = rng.standard_normal((24, 4096)).astype(np.float32)
    truth = np.fft.rfft(x.astype(np.float64), axis=1)
    for backend in (0, 1):
        dt, y = O.time_ref_rfft_batch(x, 3, backend)
        assert dt > 0
        assert rel_l2(y, truth) < tol(4096)
        assert rel_l2(y, O.ref_rfft(x)) < 1e-6
    dtp, yp = O.time_rfft_batch(x, 3)
    assert rel_l2(yp, truth) < tol(4096)


# ------------------------------------------------- DCT / DST / MDCT oracle


@pytest.mark.parametrize("N", [2, 5, 8, 33, 100])
@pytest.mark.parametrize("kind", ["dct", "dst"])
@pytest.mark.parametrize("typ", [1, 2, 3, 4])
@pytest.mark.parametrize("norm", ["none", "ortho", "forward"])
def test_r2r_oracle_matches_scipy(N, kind, typ, norm):
    """The oracle's direct sums against scipy.fft's dct / dst (a library
    routine of the same definitions): forward, and the inverse side, which
    for type T is scipy's idct / idst of the partner type (II <-> III)."""
    sf = pytest.importorskip("scipy.fft")
    if kind == "dct" and typ == 1 and N < 2:
        return
    x = np.random.default_rng(N * 10 + typ).standard_normal((2, N))
    snorm = "backward" if norm == "none" else norm
    fwd = (sf.dct if kind == "dct" else sf.dst)(x, type=typ, norm=snorm, axis=-1)
    assert rel_l2(O.r2r64(x, kind, typ, norm), fwd) < 1e-12
    partner = {1: 1, 2: 3, 3: 2, 4: 4}[typ]
    inv = (sf.idct if kind == "dct" else sf.idst)(x, type=partner, norm=snorm, axis=-1)
    assert rel_l2(O.r2r64(x, kind, typ, norm, inverse=True), inv) < 1e-12


def test_r2r_oracle_closed_forms():
    """Closed forms that do not go through scipy: DCT-II of a constant is
    2N at k = 0 only; DCT-II of cos(pi m (2n+1)/(2N)) is N at k = m; the
    ortho DCT-IV matrix is orthogonal and its own inverse; a swapped sign
    or index (DST-II read as the k-th rather than (k+1)-th sine) fails."""
    N = 16
    y = O.r2r64(np.ones(N), "dct", 2)[0]
    assert abs(y[0] - 2 * N) < 1e-12 and np.max(np.abs(y[1:])) < 1e-12
    m = 5
    x = np.cos(np.pi * m * (2 * np.arange(N) + 1) / (2 * N))
    y = O.r2r64(x, "dct", 2)[0]
    assert abs(y[m] - N) < 1e-12 and np.max(np.abs(np.delete(y, m))) < 1e-12
    A = O.r2r64(np.eye(N), "dct", 4, "ortho")
    assert np.allclose(A @ A.T, np.eye(N), atol=1e-12)
    x = np.random.default_rng(0).standard_normal(N)
    wrong = 2 * np.sin(np.pi * np.outer(np.arange(N), 2 * np.arange(N) + 1) / (2 * N)) @ x
    assert rel_l2(O.r2r64(x, "dst", 2)[0], wrong) > 0.1


def test_mdct_oracle_tdac_reconstruction():
    """MDCT -> IMDCT with a sine (Princen-Bradley) window and 50% overlap-add
    gives the signal back (time-domain aliasing cancellation); a wrong phase
    offset (n + 1/2 instead of n + 1/2 + M/2) or IMDCT scale breaks it."""
    M = 32
    N = 2 * M
    w = np.sin(np.pi * (np.arange(N) + 0.5) / N)
    sig = np.random.default_rng(1).standard_normal(6 * M)
    for norm in ("none", "ortho", "forward"):
        out = np.zeros_like(sig)
        for b in range(5):
            seg = sig[b * M: b * M + N]
            out[b * M: b * M + N] += w * O.imdct64(O.mdct64(w * seg, norm), norm)[0]
        assert np.max(np.abs(out[M:5 * M] - sig[M:5 * M])) < 1e-12, norm
    # the planted mistakes
    n = np.arange(N)
    bad = np.cos(np.pi / M * np.outer(np.arange(M) + 0.5, n + 0.5))
    out = np.zeros_like(sig)
    for b in range(5):
        seg = sig[b * M: b * M + N]
        out[b * M: b * M + N] += w * O.imdct64((bad @ (w * seg))[None])[0]
    assert np.max(np.abs(out[M:5 * M] - sig[M:5 * M])) > 0.1
    out = np.zeros_like(sig)
    for b in range(5):
        seg = sig[b * M: b * M + N]
        out[b * M: b * M + N] += w * O.imdct64(O.mdct64(w * seg))[0] / 2
    assert np.max(np.abs(out[M:5 * M] - sig[M:5 * M])) > 0.1


@pytest.mark.parametrize("M", [2, 6, 16, 250])
def test_mdct_oracle_is_folded_dct4(M):
    """The textbook reduction of the MDCT to a DCT-IV of half the length:
    with x = (a, b, c, d) in quarters, MDCT(x) = DCT-IV(-c' - d, a - b') / 2
    (' = reversed; scipy's unnormalised DCT-IV carries a factor 2).  This
    pins the oracle's direct sum on a library routine, independently of the
    TDAC check; the IMDCT is pinned as its transpose times 2/M."""
    sf = pytest.importorskip("scipy.fft")
    N = 2 * M
    x = np.random.default_rng(M).standard_normal(N)
    h = M // 2
    a, b, c, d = x[:h], x[h:M], x[M:M + h], x[M + h:]
    u = np.concatenate([-c[::-1] - d, a - b[::-1]])
    assert rel_l2(O.mdct64(x[None])[0], sf.dct(u, type=4) / 2) < 1e-12
    A = O.mdct64(np.eye(N))                   # rows: MDCT of unit vectors
    X = np.random.default_rng(M + 1).standard_normal(M)
    assert rel_l2(O.imdct64(X[None])[0], (2.0 / M) * (A @ X)) < 1e-12
