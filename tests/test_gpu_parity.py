"""GPU parity tests: the gfx950 path through the C ABI against the oracle.

Bars (stated per the transform path's contract):
  * against FP64 ground truth: relative L2 <= 1e-6 * log2(N);
  * against the reference library's own FP32 output (tests/golden): the same
    bar (both sides are FP32 computations of the same transform);
  * integer/layout behaviour (error codes, Hermitian DC/Nyquist imag = 0,
    batch membership) exactly.
Inputs: numpy PCG64, uniform[-1, 1) complex / N(0, 1) real, as float32.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import fastandfourier_amd as fa
import oracle as O
from conftest import rel_l2, tol

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import gpu_util as G  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cplx(rng, k, n):
    return (rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))).astype(np.complex64)


def truth_fft(x, inverse=False):
    y = O.fft64(x.astype(np.complex128), inverse=inverse)
    return y


def test_device_ready():
    say("case: device ready")
    assert torch.cuda.is_available()
    fa.check(G.L().faf_amd_device_ready(), "device")


# ------------------------------------------------------------------ C2C

C2C_SIZES = [2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16, 32, 60, 64, 128, 256, 360, 512, 1000, 1024,
             2048, 3000, 3840, 4000, 4050, 4096, 4800, 8192, 11, 13, 17, 19, 29, 31, 37, 97, 1009,
             16384, 20000, 32768, 65536]


@pytest.mark.parametrize("n", C2C_SIZES)
def test_c2c_matches_fp64(n):
    say(f"case: c2c n={n}")
    rng = np.random.default_rng(n)
    x = cplx(rng, 3, n)
    t = G.create("fft", n)
    ti = G.create("fft", n, dir="inverse")
    try:
        y = G.c2c(t, x)
        assert rel_l2(y, truth_fft(x)) < tol(n)
        z = G.c2c(ti, y)
        assert rel_l2(z, x) < 2 * tol(n)                     # NumPy norm: round trip
    finally:
        G.destroy(t)
        G.destroy(ti)


def test_c1_batch256_n4096_forward_inverse():
    """Config C1: N=4096, batch 256, uniform[-1,1] from seed 0, NumPy norm."""
    say("case: C1 c2c 4096 x 256")
    n, b = 4096, 256
    rng = np.random.default_rng(0)
    x = cplx(rng, b, n)
    t = G.create("fft", n)
    ti = G.create("fft", n, dir="inverse")
    try:
        y = G.c2c(t, x)
        assert rel_l2(y, truth_fft(x)) < tol(n)
        assert rel_l2(y[:4], O.fft32(x[:4])) < tol(n)        # beside the reference's CPU path
        z = G.c2c(ti, y)
        assert rel_l2(z, truth_fft(y.astype(np.complex128), inverse=True) / n) < tol(n)
        assert rel_l2(z, x) < 2 * tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)


@pytest.mark.parametrize("norm", ["ortho", "forward"])
def test_c2c_norms(norm):
    say(f"case: c2c norm {norm}")
    n = 1024
    rng = np.random.default_rng(3)
    x = cplx(rng, 2, n)
    t = G.create("fft", n, norm=norm)
    ti = G.create("fft", n, norm=norm, dir="inverse")
    try:
        s = 1 / np.sqrt(n) if norm == "ortho" else 1 / n
        assert rel_l2(G.c2c(t, x), truth_fft(x) * s) < tol(n)
        si = 1 / np.sqrt(n) if norm == "ortho" else 1.0
        assert rel_l2(G.c2c(ti, x), truth_fft(x, inverse=True) * si) < tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)


@pytest.mark.parametrize("n", [2048, 8192])
@pytest.mark.parametrize("inter", [False, True])
def test_c2c_batches_2048_8192(n, inter):
    """Batched C2C at 2048 (staged passes) and 8192 (the single-pass
    register-resident core), split and interleaved, both directions, against
    FP64; a signal alone gives the same bits as inside the batch."""
    say(f"case: single-pass c2c n={n} inter={inter}")
    rng = np.random.default_rng(n + (7 if inter else 0))
    x = (rng.uniform(-1, 1, (9, n)) + 1j * rng.uniform(-1, 1, (9, n))).astype(np.complex64)
    kw = {"layout": "interleaved"} if inter else {}
    t = G.create("fft", n, **kw)
    ti = G.create("fft", n, dir="inverse", **kw)
    try:
        X = G.c2c(t, x, inter=inter)
        assert rel_l2(X, np.fft.fft(x.astype(np.complex128), axis=1)) < tol(n)
        y = G.c2c(ti, X, inter=inter)
        assert rel_l2(y, x) < 2 * tol(n)
        for i in (0, 8):
            assert np.array_equal(G.c2c(t, x[i:i + 1], inter=inter)[0], X[i])
    finally:
        G.destroy(t)
        G.destroy(ti)


@pytest.mark.parametrize("n", [64, 4096, 32768, 65536])
def test_c2c_interleaved(n):
    say(f"case: c2c interleaved n={n}")
    rng = np.random.default_rng(5)
    x = cplx(rng, 2, n)
    t = G.create("fft", n, layout="interleaved")
    try:
        assert rel_l2(G.c2c(t, x, inter=True), truth_fft(x)) < tol(n)
    finally:
        G.destroy(t)


def test_c2c_golden_reference_outputs(golden):
    """Against the reference library's FP32 outputs (tests/golden)."""
    for key in golden.files:
        if not (key.startswith("c2c_") and key.endswith("_in")):
            continue
        n = int(key.split("_")[1])
        say(f"case: golden c2c n={n}")
        x = golden[key]
        t = G.create("fft", n)
        ti = G.create("fft", n, dir="inverse")
        try:
            assert rel_l2(G.c2c(t, x), golden[f"c2c_{n}_fwd"]) < tol(n), n
            assert rel_l2(G.c2c(ti, x), golden[f"c2c_{n}_inv"]) < tol(n), n
        finally:
            G.destroy(t)
            G.destroy(ti)


@pytest.mark.parametrize("n", [23, 47, 307, 4097])
def test_bluestein(n, golden):
    say(f"case: bluestein n={n}")
    rng = np.random.default_rng(n)
    x = cplx(rng, 2, n)
    with pytest.raises(fa.FafError):
        G.create("fft", n)                                       # opt-in only
    t = G.create("fft", n, flags=fa.FLAG_BLUESTEIN)
    ti = G.create("fft", n, flags=fa.FLAG_BLUESTEIN, dir="inverse")
    try:
        assert rel_l2(G.c2c(t, x), truth_fft(x)) < tol(n)
        assert rel_l2(G.c2c(ti, x), truth_fft(x, inverse=True) / n) < tol(n)
        if f"blue_{n}_in" in golden.files:
            assert rel_l2(G.c2c(t, golden[f"blue_{n}_in"]), golden[f"blue_{n}_fwd"]) < 2 * tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)


def test_host_pointers_are_staged():
    """The drop-in path: host numpy buffers through faf_execute."""
    say("case: host pointers")
    lib = G.L()
    n = 360
    rng = np.random.default_rng(9)
    x = cplx(rng, 1, n)[0]
    xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    yr, yi = np.zeros(n, np.float32), np.zeros(n, np.float32)
    t = G.create("fft", n)
    try:
        bi = fa.buffer_split(xr.ctypes.data, xi.ctypes.data, n)
        bo = fa.buffer_split(yr.ctypes.data, yi.ctypes.data, n)
        fa.check(lib.faf_execute(t, C.byref(bo), C.byref(bi)), "execute")
        assert rel_l2(yr + 1j * yi, truth_fft(x)[0]) < tol(n)
        # in place on host memory
        fa.check(lib.faf_execute(t, C.byref(bi), C.byref(bi)), "execute in place")
        assert rel_l2(xr + 1j * xi, truth_fft(x)[0]) < tol(n)
    finally:
        G.destroy(t)


def test_error_paths():
    say("case: error paths")
    lib = G.L()
    t = G.create("fft", 64)
    try:
        a = G.empty((2, 64))
        bi = fa.buffer_interleaved(a.data_ptr(), 64)
        bo = fa.buffer_split(a[0].data_ptr(), a[1].data_ptr(), 64)
        assert lib.faf_execute(t, C.byref(bo), C.byref(bi)) == -1
        assert b"layout" in lib.faf_get_error()
    finally:
        G.destroy(t)
    r = G.create("rfft", 64)
    try:
        a = G.empty((3, 64))
        bi = fa.buffer_real(a[0].data_ptr(), 63)
        bo = fa.buffer_hermitian(a[1].data_ptr(), a[2].data_ptr(), 33)
        assert lib.faf_execute(r, C.byref(bo), C.byref(bi)) == -1
        assert b"expected 64" in lib.faf_get_error()
    finally:
        G.destroy(r)
    for bad in (14, 23, 4097):
        with pytest.raises(fa.FafError):
            G.create("fft", bad)
    with pytest.raises(fa.FafError):
        G.create("rfft", 15)


# ------------------------------------------------------------------ R2C / C2R

@pytest.mark.parametrize("n", [4, 8, 30, 64, 1000, 1024, 2048, 4096, 8192, 16384, 65536])
def test_rfft_irfft(n):
    say(f"case: rfft n={n}")
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal((3, n)).astype(np.float32)
    t = G.create("rfft", n)
    ti = G.create("rfft", n, dir="inverse")
    try:
        X = G.r2c(t, x)
        truth = np.stack([O.fft64(r.astype(np.float64))[0][: n // 2 + 1] for r in x])
        assert rel_l2(X, truth) < tol(n)
        assert np.all(X[:, 0].imag == 0) and np.all(X[:, -1].imag == 0)
        y = G.c2r(ti, X, n)
        assert rel_l2(y, x) < 2 * tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)


@pytest.mark.parametrize("n", [1024, 2048, 4096, 16384, 65536])
@pytest.mark.parametrize("norm", ["none", "ortho", "forward"])
def test_rfft_interleaved_layout_and_norms(n, norm):
    """INTERLEAVED spectra (nb complex64) and every norm, including the
    specialised four-step at 65536 whose Kiss passes are fused."""
    say(f"case: rfft interleaved n={n} norm={norm}")
    lib = G.L()
    rng = np.random.default_rng(n + 7)
    cnt, nb = 3, n // 2 + 1
    x = rng.standard_normal((cnt, n)).astype(np.float32)
    t = G.create("rfft", n, layout="interleaved", norm=norm)
    ti = G.create("rfft", n, layout="interleaved", norm=norm, dir="inverse")
    try:
        dx = G.dev(x)
        X = G.empty((cnt, 2 * nb))
        y = G.empty((cnt, n))
        torch.cuda.synchronize()
        bx = fa.buffer_real(dx.data_ptr(), n)
        bX = fa.buffer_interleaved(X.data_ptr(), nb)
        by = fa.buffer_real(y.data_ptr(), n)
        fa.check(lib.faf_execute_batch(t, C.byref(bX), C.byref(bx), cnt, 0, 0), "rfft inter")
        fa.check(lib.faf_execute_batch(ti, C.byref(by), C.byref(bX), cnt, 0, 0), "irfft inter")
        Xh = G.host(X).view(np.complex64)
        fwd_s = {"none": 1.0, "ortho": 1 / np.sqrt(n), "forward": 1 / n}[norm]
        truth = np.stack([O.fft64(r.astype(np.float64))[0][:nb] for r in x]) * fwd_s
        assert rel_l2(Xh, truth) < tol(n)
        assert rel_l2(G.host(y), x) < 2 * tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)


def test_rfft_golden_reference_outputs(golden):
    for n in (1024, 4096, 16384):
        say(f"case: golden rfft n={n}")
        x = golden[f"rfft_{n}_in"]
        t = G.create("rfft", n)
        to = G.create("rfft", n, norm="ortho")
        ti = G.create("rfft", n, dir="inverse")
        try:
            assert rel_l2(G.r2c(t, x), golden[f"rfft_{n}_fwd"]) < tol(n)
            assert rel_l2(G.r2c(to, x), golden[f"rfft_{n}_ortho"]) < tol(n)
            assert rel_l2(G.c2r(ti, golden[f"rfft_{n}_fwd"], n), golden[f"rfft_{n}_inv"]) < tol(n)
        finally:
            G.destroy(t)
            G.destroy(to)
            G.destroy(ti)


@pytest.mark.parametrize("n", [1024, 4096, 16384, 65536])
def test_c2_full_batch_round_trip(n):
    """Config C2 at its full size: batch = 2^28 / N signals (1 GiB in).

    Size-independent properties: irfft(rfft(x)) == x, Hermitian DC/Nyquist
    imag exactly 0, and bins of a sample of rows against FP64."""
    say(f"case: C2 full batch n={n}")
    b = (1 << 28) // n
    nb = n // 2 + 1
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn((b, n), generator=gen, device="cuda", dtype=torch.float32)
    re = torch.empty((b, nb), device="cuda")
    im = torch.empty((b, nb), device="cuda")
    y = torch.empty((b, n), device="cuda")
    lib = G.L()
    t = G.create("rfft", n)
    ti = G.create("rfft", n, dir="inverse")
    try:
        torch.cuda.synchronize()
        bi = fa.buffer_real(x.data_ptr(), n)
        bo = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), b, 0, 0), "rfft")
        by = fa.buffer_real(y.data_ptr(), n)
        fa.check(lib.faf_execute_batch(ti, C.byref(by), C.byref(bo), b, 0, 0), "irfft")
        torch.cuda.synchronize()
        err = (torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x)).item()
        assert err < 2 * tol(n), err
        assert float(im[:, 0].abs().max()) == 0.0 and float(im[:, -1].abs().max()) == 0.0
        rows = [0, b // 2, b - 1]
        xs = x[rows].cpu().numpy()
        Xs = re[rows].cpu().numpy() + 1j * im[rows].cpu().numpy()
        truth = np.stack([O.fft64(r.astype(np.float64))[0][:nb] for r in xs])
        assert rel_l2(Xs, truth) < tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)
        del x, re, im, y
        torch.cuda.empty_cache()


def test_single_pass_32768_many_signals():
    """C2C 32768 runs in one pass, one workgroup per CU looping over the
    signals: 300 signals make the first 44 workgroups take two."""
    say("case: c2c 32768 x 300")
    n, b = 32768, 300
    rng = np.random.default_rng(n)
    x = cplx(rng, b, n)
    t = G.create("fft", n)
    ti = G.create("fft", n, dir="inverse")
    try:
        y = G.c2c(t, x)
        sel = [0, 1, 43, 44, 255, 256, 299]
        assert rel_l2(y[sel], truth_fft(x[sel])) < tol(n)
        z = G.c2c(ti, y)
        assert rel_l2(z, x) < 2 * tol(n)
    finally:
        G.destroy(t)
        G.destroy(ti)
    t = G.create("fft", n, layout="interleaved")
    try:
        yi = G.c2c(t, x[:2], inter=True)
        assert rel_l2(yi, truth_fft(x[:2])) < tol(n)
    finally:
        G.destroy(t)


@pytest.mark.parametrize("n", [4096, 16384, 65536])
def test_rfft_unaligned_rows_fall_back(n):
    """Rows at an odd float distance are not 8-byte aligned, so the rfft and
    irfft leave the single pass for the staged (4096, 16384) or four-step
    (65536) path; both paths agree with FP64."""
    say(f"case: rfft {n} odd row distance")
    lib = G.L()
    cnt = 3
    nb = n // 2 + 1
    rng = np.random.default_rng(n + 1)
    xs = rng.standard_normal((cnt, n + 1)).astype(np.float32)
    t = G.create("rfft", n)
    try:
        dx = G.dev(xs)
        re, im = G.empty((cnt, nb)), G.empty((cnt, nb))
        torch.cuda.synchronize()
        bi = fa.buffer_real(dx.data_ptr(), n)
        bo = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, n + 1, 0), "rfft")
        X = G.host(re) + 1j * G.host(im)
        truth = np.stack([O.fft64(r[:n].astype(np.float64))[0][:nb] for r in xs])
        assert rel_l2(X, truth) < tol(n)
        assert rel_l2(G.r2c(t, np.ascontiguousarray(xs[:, :n])), truth) < tol(n)
        # the inverse into rows at an odd distance
        ti = G.create("rfft", n, dir="inverse")
        try:
            y = G.empty((cnt, n + 1))
            torch.cuda.synchronize()
            by = fa.buffer_real(y.data_ptr(), n)
            fa.check(lib.faf_execute_batch(ti, C.byref(by), C.byref(bo), cnt, 0, n + 1), "irfft")
            assert rel_l2(G.host(y)[:, :n], xs[:, :n]) < tol(n)
        finally:
            G.destroy(ti)
    finally:
        G.destroy(t)


def test_batch_member_equals_alone():
    """A signal gives the same bits alone and inside a batch."""
    say("case: batch membership")
    n = 4096
    rng = np.random.default_rng(11)
    x = rng.standard_normal((37, n)).astype(np.float32)
    t = G.create("rfft", n)
    try:
        allb = G.r2c(t, x)
        for i in (0, 17, 36):
            alone = G.r2c(t, x[i:i + 1])
            assert np.array_equal(alone[0], allb[i])
    finally:
        G.destroy(t)


# ------------------------------------------------------------------ pipeline

def test_notch_pipeline_golden(golden):
    """C2 notch: rfft -> real mask zeroing bins [ceil(N/10), N/8) -> irfft, one kernel."""
    say("case: notch pipeline")
    lib = G.L()
    n = 4096
    t = lib.chirp_compile(f"(pipeline (rfft :size {n}) (mul-spectrum) (irfft))".encode())
    assert t, lib.faf_get_error()
    try:
        m = golden["notch_4096_mask"]
        hr = G.dev(m)
        hi = G.dev(np.zeros_like(m))
        assert lib.chirp_bind(t, b"H", hr.data_ptr(), hi.data_ptr(), m.size) == 0
        x = golden["notch_4096_in"]
        y = G.real_to_real(t, x)
        assert rel_l2(y, golden["notch_4096_out"]) < tol(n)
        Y = np.fft.rfft(y.astype(np.float64), axis=1)
        assert np.abs(Y[:, -(-n // 10): n // 8]).max() < 1e-3 * np.abs(Y).max()
    finally:
        lib.faf_destroy_transform(t)


@pytest.mark.parametrize("n", [16, 64, 1000, 4096, 16384])
def test_pipeline_ops_match_oracle(n):
    say(f"case: pipeline ops n={n}")
    lib = G.L()
    nb = n // 2 + 1
    rng = np.random.default_rng(n)
    h = (rng.uniform(-1, 1, nb) + 1j * rng.uniform(-1, 1, nb)).astype(np.complex64)
    ops = (C.c_int * 3)(2, 1, 3)
    a1 = (C.c_uint32 * 3)(0, 0, 2)
    a2 = (C.c_uint32 * 3)(0, 0, nb - 3)
    t = lib.faf_create_pipeline(C.byref(fa.config_init(n)), 3, ops, a1, a2)
    assert t, lib.faf_get_error()
    try:
        hr = np.ascontiguousarray(h.real)
        hi = np.ascontiguousarray(h.imag)
        assert lib.chirp_bind(t, b"H", hr.ctypes.data, hi.ctypes.data, nb) == 0  # host spectrum
        x = rng.standard_normal((4, n)).astype(np.float32)
        y = G.real_to_real(t, x)
        ref = O.pipeline32(x, n, [2, 1, 3], [0, 0, 2], [0, 0, nb - 3], h)
        assert rel_l2(y, ref) < 2 * tol(n)
    finally:
        lib.faf_destroy_transform(t)


@pytest.mark.parametrize("bins", [2049, 1949, 100])
def test_single_mul_pipeline_4096(bins):
    """The single-MUL pipeline at N=4096 (one pass) against the oracle, for a
    complex spectrum bound over all bins and over fewer (the last bound bin
    is then multiplied as a real Nyquist bin, reference src/faf_core.c:150-220),
    on packed rows and on rows at an odd distance (the staged fallback)."""
    say(f"case: single-MUL pipeline 4096, {bins} bound bins")
    lib = G.L()
    n = 4096
    rng = np.random.default_rng(bins)
    h = (rng.uniform(-1, 1, bins) + 1j * rng.uniform(-1, 1, bins)).astype(np.complex64)
    ops = (C.c_int * 1)(2)
    a0 = (C.c_uint32 * 1)(0)
    t = lib.faf_create_pipeline(C.byref(fa.config_init(n)), 1, ops, a0, a0)
    assert t, lib.faf_get_error()
    try:
        hr = G.dev(np.ascontiguousarray(h.real))
        hi = G.dev(np.ascontiguousarray(h.imag))
        assert lib.chirp_bind(t, b"H", hr.data_ptr(), hi.data_ptr(), bins) == 0
        x = rng.standard_normal((5, n + 1)).astype(np.float32)
        ref = O.pipeline32(np.ascontiguousarray(x[:, :n]), n, [2], [0], [0], h)
        y = G.real_to_real(t, np.ascontiguousarray(x[:, :n]))
        assert rel_l2(y, ref) < 2 * tol(n)
        dx = G.dev(x)
        dy = G.empty((5, n + 1))
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_batch(t, C.byref(fa.buffer_real(dy.data_ptr(), n)),
                                       C.byref(fa.buffer_real(dx.data_ptr(), n)), 5, n + 1,
                                       n + 1), "pipeline odd rows")
        assert rel_l2(G.host(dy)[:, :n], ref) < 2 * tol(n)
    finally:
        lib.faf_destroy_transform(t)


def test_chirp_forms():
    say("case: chirp forms")
    lib = G.L()
    n = 16
    t = lib.chirp_compile(b"(pipeline (rfft :size 16) conj (irfft))")
    assert t, lib.faf_get_error()
    try:
        x = (np.arange(n, dtype=np.float32) - 8)[None]
        y = G.real_to_real(t, x)[0]
        assert np.allclose(y, np.r_[x[0, 0], x[0, :0:-1]], atol=1e-4)   # conj = time reversal
    finally:
        lib.faf_destroy_transform(t)
    t = lib.chirp_compile(b"(pipeline (rfft :size 32) (bandpass :lo 4 :hi 4) (irfft))")
    assert t
    try:
        i = np.arange(32)
        x = (np.cos(2 * np.pi * 4 * i / 32) + np.cos(2 * np.pi * 7 * i / 32)).astype(np.float32)
        y = G.real_to_real(t, x[None])[0]
        assert np.allclose(y, np.cos(2 * np.pi * 4 * i / 32), atol=1e-5)
    finally:
        lib.faf_destroy_transform(t)
    assert not lib.chirp_compile(b"(pipeline (rfft :size 64) (spectral notch) (irfft))")


# ------------------------------------------------------------------ ragged (C3)

C3_LENGTHS = [60, 360, 1000, 3000, 4800, 11, 17, 31, 97, 8192, 1009]


def test_c3_ragged_persistent_launch():
    """Config C3 shape: mixed lengths (incl. Rader primes) in one persistent launch."""
    say("case: C3 ragged 2000 signals")
    lib = G.L()
    rng = np.random.default_rng(0)
    lens = rng.choice(C3_LENGTHS, size=2000)
    arr = (C.c_size_t * len(lens))(*[int(v) for v in lens])
    r = lib.faf_ragged_create(arr, len(lens), 0)
    assert r, lib.faf_get_error()
    try:
        tot = lib.faf_ragged_total(r)
        assert tot == int(lens.sum())
        xr = rng.uniform(-1, 1, tot).astype(np.float32)
        xi = rng.uniform(-1, 1, tot).astype(np.float32)
        dxr, dxi = G.dev(xr), G.dev(xi)
        yr, yi = G.empty(tot), G.empty(tot)
        torch.cuda.synchronize()
        fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(), dxr.data_ptr(),
                                        dxi.data_ptr()), "ragged")
        y = G.host(yr) + 1j * G.host(yi)
        off = np.concatenate([[0], np.cumsum(lens)])
        x = xr + 1j * xi
        worst = 0.0
        for i in range(0, len(lens), 7):
            s = slice(off[i], off[i + 1])
            e = rel_l2(y[s], O.fft64(x[s])[0]) / tol(int(lens[i]))
            worst = max(worst, e)
        assert worst < 1.0, worst
        # each signal equals the same length run alone (batched plan)
        for n in (11, 1009, 4800):
            i = int(np.nonzero(lens == n)[0][0])
            s = slice(off[i], off[i + 1])
            t = G.create("fft", n)
            try:
                alone = G.c2c(t, x[s][None])[0]
            finally:
                G.destroy(t)
            assert rel_l2(y[s], alone) < tol(n)
    finally:
        lib.faf_ragged_destroy(r)


@pytest.mark.parametrize("lengths,flags", [
    (C3_LENGTHS, 0),
    ([4050, 128, 13, 19, 2000, 23, 97, 8192, 24, 1009, 60, 4096], 1)])
def test_ragged_launch_shapes_agree(lengths, flags):
    """One workgroup per unit and persistent workgroups, of 256 or 512 threads,
    run the same unit code: every signal has the same bits in every shape
    (lean kernel, and the one with the runtime paths), over several launches
    of each.  A batch with a length that needs more than 256 threads refuses
    the smaller workgroup and keeps running as before."""
    say(f"case: ragged launch shapes {lengths[:3]}...")
    lib = G.L()
    rng = np.random.default_rng(5)
    lens = [int(v) for v in rng.choice(lengths, 1500)]
    arr = (C.c_size_t * len(lens))(*lens)
    r = lib.faf_ragged_create(arr, len(lens), fa.FLAG_BLUESTEIN if flags else 0)
    assert r, lib.faf_get_error()
    try:
        tot = lib.faf_ragged_total(r)
        x = (rng.uniform(-1, 1, tot) + 1j * rng.uniform(-1, 1, tot)).astype(np.complex64)
        dxr, dxi = G.dev(x.real.copy()), G.dev(x.imag.copy())
        auto = lib.faf_ragged_block(r)
        say(f"  automatic workgroup size {auto}")
        assert auto in (256, 512)
        if flags == 0:
            assert auto == 256, "every C3 length runs on 256 threads"
        shapes = [(0, 0), (1, 0), (0, 512), (1, 512), (0, 0)]
        if auto == 256:
            shapes.append((1, 256))
        else:
            assert lib.faf_ragged_set_block(r, 256) == -1
            assert lib.faf_ragged_block(r) == 512
        outs = []
        for persist, block in shapes:
            fa.check(lib.faf_ragged_set_block(r, block), "set block")
            assert lib.faf_ragged_block(r) == (block or auto)
            fa.check(lib.faf_ragged_set_persistent(r, persist), "set persistent")
            yr, yi = G.empty(tot), G.empty(tot)
            for _ in range(3):
                fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(),
                                                dxr.data_ptr(), dxi.data_ptr()), "ragged")
            outs.append(G.host(yr) + 1j * G.host(yi))
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
        off = np.concatenate([[0], np.cumsum(lens)])
        for i in list(range(0, len(lens), 13)) + [len(lens) - 1]:
            a, e = off[i], off[i + 1]
            assert rel_l2(outs[1][a:e], O.fft64(x[a:e].astype(np.complex128))[0]) < tol(lens[i])
    finally:
        lib.faf_ragged_destroy(r)


def test_ragged_runtime_paths():
    """Lengths without compile-time passes (and a Bluestein length) put the
    launch on the kernel that carries the runtime Stockham, Rader and
    Bluestein paths; every signal still matches FP64."""
    say("case: ragged runtime paths")
    lib = G.L()
    rng = np.random.default_rng(12)
    pool = [4050, 128, 13, 19, 2000, 23, 97, 8192, 24, 1009, 60, 4096]
    lens = rng.choice(pool, size=400)
    arr = (C.c_size_t * len(lens))(*[int(v) for v in lens])
    r = lib.faf_ragged_create(arr, len(lens), fa.FLAG_BLUESTEIN)
    assert r, lib.faf_get_error()
    try:
        tot = lib.faf_ragged_total(r)
        x = (rng.uniform(-1, 1, tot) + 1j * rng.uniform(-1, 1, tot)).astype(np.complex64)
        dxr, dxi = G.dev(x.real.copy()), G.dev(x.imag.copy())
        yr, yi = G.empty(tot), G.empty(tot)
        torch.cuda.synchronize()
        fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(), dxr.data_ptr(),
                                        dxi.data_ptr()), "ragged")
        y = G.host(yr) + 1j * G.host(yi)
        off = np.concatenate([[0], np.cumsum(lens)])
        seen = set()
        for i in range(len(lens)):
            n = int(lens[i])
            if n in seen and i % 5:
                continue
            seen.add(n)
            s = slice(off[i], off[i + 1])
            assert rel_l2(y[s], O.fft64(x[s].astype(np.complex128))[0]) < tol(n), (i, n)
        assert seen == set(pool)
    finally:
        lib.faf_ragged_destroy(r)


def test_framed_stft_short_and_exact_records():
    """A record shorter than the window is one zero-padded frame (the loader
    with bounds tests); a record exactly one window long is one whole frame."""
    say("case: stft short and exact records")
    lib = G.L()
    win, hop = 1024, 256
    t = G.create("stft", win, layout="real", win_length=win, hop_length=hop)
    try:
        rng = np.random.default_rng(21)
        nb = win // 2 + 1
        for length in (700, 1024):
            x = rng.standard_normal((2, length)).astype(np.float32)
            assert lib.faf_stft_frames(t, length) == 1
            dx = G.dev(x)
            re, im = G.empty((2, 1, nb)), G.empty((2, 1, nb))
            torch.cuda.synchronize()
            fa.check(lib.faf_execute_stft(t, re.data_ptr(), im.data_ptr(), dx.data_ptr(),
                                          length, 2, 0), "stft")
            S = G.host(re) + 1j * G.host(im)
            for r in range(2):
                assert rel_l2(S[r], O.stft64(x[r], win, hop)) < tol(win), length
    finally:
        G.destroy(t)


# ------------------------------------------------------------------ STFT / CWT (C4)

def test_framed_stft():
    say("case: framed stft")
    lib = G.L()
    win, hop, length, recs = 1024, 256, 20000, 3
    t = G.create("stft", win, layout="real", win_length=win, hop_length=hop)
    try:
        rng = np.random.default_rng(4)
        x = rng.standard_normal((recs, length)).astype(np.float32)
        nf = lib.faf_stft_frames(t, length)
        assert nf == (length - win) // hop + 1
        nb = win // 2 + 1
        dx = G.dev(x)
        re, im = G.empty((recs, nf, nb)), G.empty((recs, nf, nb))
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_stft(t, re.data_ptr(), im.data_ptr(), dx.data_ptr(), length,
                                      recs, 0), "stft")
        S = G.host(re) + 1j * G.host(im)
        for r in range(recs):
            assert rel_l2(S[r], O.stft64(x[r], win, hop)) < tol(win)
    finally:
        G.destroy(t)


@pytest.mark.parametrize("win, hop", [(1024, 256), (512, 128), (2048, 512)])
def test_stft_staged_and_per_bin_stores_agree(win, hop):
    """Whole workgroups of frames store their bins through LDS as aligned
    float4 when the output planes are 16-byte aligned, and per bin otherwise;
    both give the same bits, and both match FP64 (the frame count leaves a
    partial last workgroup)."""
    say(f"case: stft staged vs per-bin stores win={win}")
    lib = G.L()
    length, recs = 9 * win + 77, 3
    t = G.create("stft", win, layout="real", win_length=win, hop_length=hop)
    try:
        rng = np.random.default_rng(win)
        x = rng.standard_normal((recs, length)).astype(np.float32)
        nf = lib.faf_stft_frames(t, length)
        nb = win // 2 + 1
        tot = recs * nf * nb
        dx = G.dev(x)
        out = {}
        for shift in (0, 1):
            re, im = G.empty(tot + 4), G.empty(tot + 4)
            torch.cuda.synchronize()
            fa.check(lib.faf_execute_stft(t, re.data_ptr() + 4 * shift, im.data_ptr() + 4 * shift,
                                          dx.data_ptr(), length, recs, 0), "stft")
            out[shift] = (G.host(re)[shift:shift + tot], G.host(im)[shift:shift + tot])
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        S = (out[0][0] + 1j * out[0][1]).reshape(recs, nf, nb)
        for r in range(recs):
            assert rel_l2(S[r], O.stft64(x[r], win, hop)) < tol(win)
    finally:
        G.destroy(t)


@pytest.mark.parametrize("n", [512, 1024, 4096])
def test_rfft_staged_and_per_bin_stores_agree(n):
    """The rfft kernels that stage whole workgroups of packed rows through
    LDS (aligned float4 stores) give the same bits into unaligned planes,
    where they store per bin, and match FP64."""
    say(f"case: rfft staged vs per-bin stores n={n}")
    lib = G.L()
    cnt = 37
    nb = n // 2 + 1
    rng = np.random.default_rng(n + 3)
    x = rng.standard_normal((cnt, n)).astype(np.float32)
    t = G.create("rfft", n)
    try:
        dx = G.dev(x)
        out = {}
        for shift in (0, 1):
            re, im = G.empty(cnt * nb + 4), G.empty(cnt * nb + 4)
            torch.cuda.synchronize()
            bo = fa.buffer_hermitian(re.data_ptr() + 4 * shift, im.data_ptr() + 4 * shift, nb)
            fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(fa.buffer_real(dx.data_ptr(), n)),
                                           cnt, 0, 0), "rfft")
            out[shift] = (G.host(re)[shift:shift + cnt * nb], G.host(im)[shift:shift + cnt * nb])
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        X = (out[0][0] + 1j * out[0][1]).reshape(cnt, nb)
        assert rel_l2(X, np.fft.rfft(x.astype(np.float64), axis=1)) < tol(n)
    finally:
        G.destroy(t)


def _cwt_create(n, fs, voices=None, fmin=None, fmax=None):
    lib = G.L()
    cc = lib.faf_cwt_config_init(n)
    cc.fs = fs
    if voices:
        cc.voices = voices
    if fmin:
        cc.fmin = fmin
    if fmax:
        cc.fmax = fmax
    t = lib.faf_create_cwt(C.byref(cc))
    assert t, lib.faf_get_error()
    return t


def _table(t, lib):
    rows = lib.faf_cwt_n_rows(t)
    nb = lib.faf_cwt_n_bins(t)
    stride = (nb + 15) & ~15
    tab = np.zeros((rows, stride), np.float32)
    tab[0, :nb] = np.ctypeslib.as_array(lib.faf_cwt_phi_f32(t), (nb,))
    for j in range(rows - 1):
        tab[j + 1, :nb] = np.ctypeslib.as_array(lib.faf_cwt_psi_f32(t, j), (nb,))
    return tab, stride


def test_cwt_bank_matches_reference(golden):
    """C4 bank (n=4096, fs=16 kHz, 11 voices, fmin 62.5 Hz): 64 scales + lowpass."""
    say("case: cwt bank tables")
    lib = G.L()
    f0 = 62.5
    t = _cwt_create(4096, 16000.0, 11, f0, f0 * 2 ** (63.5 / 11))
    try:
        assert lib.faf_cwt_n_scales(t) == 64 and lib.faf_cwt_n_rows(t) == 65
        tab, _ = _table(t, lib)
        ref = golden["cwt4096_psi"]
        assert rel_l2(tab[:, :ref.shape[1]], ref) < 1e-6
        fr = np.zeros(64)
        lib.faf_cwt_freqs(t, fr.ctypes.data_as(C.POINTER(C.c_double)), 64)
        assert np.allclose(fr, golden["cwt4096_freqs"], rtol=1e-12)
        rep = lib.faf_cwt_bank_report(t).contents
        g = golden["cwt4096_report"]
        assert rep.n_scales == g[0] and rep.passed == g[1] and rep.k_lo == g[7] and rep.k_hi == g[8]
        assert abs(rep.max_abs_dev - g[2]) < 1e-9 and abs(rep.frame_cond - g[6]) < 1e-9
        assert abs(rep.admissibility_C - g[11]) < 1e-9 * abs(g[11])
        assert rep.frame_cond < 1.2 and rep.max_wrap_energy <= 1e-3
    finally:
        lib.faf_destroy_transform(t)


def test_cwt_complex_rows_and_reference_real_rows(golden):
    say("case: cwt rows")
    lib = G.L()
    t = _cwt_create(512, 1.0)
    try:
        tab, stride = _table(t, lib)
        rows = tab.shape[0]
        assert rel_l2(tab[:, :257], golden["cwt512_psi"]) < 1e-6
        x = golden["cwt512_in"]
        # reference's REAL output through faf_execute
        dx = G.dev(x)
        w = G.empty((rows, 512))
        torch.cuda.synchronize()
        bi = fa.buffer_real(dx.data_ptr(), 512)
        bo = fa.buffer_real(w.data_ptr(), rows * 512)
        fa.check(lib.faf_execute(t, C.byref(bo), C.byref(bi)), "cwt")
        assert rel_l2(G.host(w), golden["cwt512_rows"]) < tol(512)
        # analytic complex64 rows for a batch
        rng = np.random.default_rng(2)
        xs = rng.standard_normal((5, 512)).astype(np.float32)
        dxs = G.dev(xs)
        z = G.empty((5, rows, 512, 2))
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_cwt_complex(t, z.data_ptr(), dxs.data_ptr(), 5, 0), "cwt complex")
        zc = G.host(z)
        zc = zc[..., 0] + 1j * zc[..., 1]
        for s in range(5):
            assert rel_l2(zc[s], O.cwt_complex64(xs[s], tab, stride)) < tol(512)
    finally:
        lib.faf_destroy_transform(t)


def test_cwt_c4_segments_and_dual_inverse():
    """C4 CWT bank on 4096-sample segments; dual-frame inverse reconstructs."""
    say("case: cwt C4 segments + icwt")
    lib = G.L()
    f0 = 62.5
    t = _cwt_create(4096, 16000.0, 11, f0, f0 * 2 ** (63.5 / 11))
    ti = lib.faf_create_inverse(t)
    assert ti, lib.faf_get_error()
    try:
        tab, stride = _table(t, lib)
        rows = 65
        rng = np.random.default_rng(8)
        xs = rng.standard_normal((3, 4096)).astype(np.float32)
        dxs = G.dev(xs)
        z = G.empty((3, rows, 4096, 2))
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_cwt_complex(t, z.data_ptr(), dxs.data_ptr(), 3, 0), "cwt")
        zc = G.host(z)
        zc = zc[..., 0] + 1j * zc[..., 1]
        assert rel_l2(zc[1], O.cwt_complex64(xs[1], tab, stride)) < tol(4096)
        # real rows -> dual inverse -> input (reference tests/test_cwt.cpp:348-399: <= 1e-4)
        w = G.dev(np.ascontiguousarray(zc[0].real.astype(np.float32)))
        y = G.empty(4096)
        torch.cuda.synchronize()
        bi = fa.buffer_real(w.data_ptr(), rows * 4096)
        bo = fa.buffer_real(y.data_ptr(), 4096)
        fa.check(lib.faf_execute(ti, C.byref(bo), C.byref(bi)), "icwt")
        assert rel_l2(G.host(y), xs[0]) < 1e-4
    finally:
        lib.faf_destroy_transform(t)
        lib.faf_destroy_transform(ti)


def test_icwt_l1_matches_reference_and_round_trips(golden):
    """L1 one-integral inverse on the device: the compiled reference's
    output for the default n=512 bank, a batch of three segments against the
    oracle, and the reference's sine round trip (tests/test_cwt.cpp:172-204,
    <= 0.25) at n=4096 through the device forward CWT."""
    say("case: icwt L1")
    lib = G.L()
    cc = lib.faf_cwt_config_init(512)
    cc.fs = 1.0
    ti = lib.faf_create_icwt(C.byref(cc), 1)
    assert ti, lib.faf_get_error()
    try:
        rows = golden["cwt512_rows"]
        nr = rows.shape[0]
        assert lib.faf_cwt_n_rows(ti) == nr
        rng = np.random.default_rng(21)
        ws = np.stack([rows, 0.5 * rows, rng.standard_normal(rows.shape).astype(np.float32)])
        dw = G.dev(ws)
        y = G.empty((3, 512))
        torch.cuda.synchronize()
        bi = fa.buffer_real(dw.data_ptr(), nr * 512)
        bo = fa.buffer_real(y.data_ptr(), 512)
        fa.check(lib.faf_execute_batch(ti, C.byref(bo), C.byref(bi), 3, 0, 0), "icwt l1 batch")
        yh = G.host(y)
        assert rel_l2(yh[0], golden["cwt512_l1_out"]) < 1e-6
        v, c = int(golden["cwt512_voices"][0]), golden["cwt512_report"][11]
        for s in range(3):
            assert rel_l2(yh[s], O.icwt_l1_f32(ws[s], nr - 1, True, v, c)) < 1e-6
    finally:
        lib.faf_destroy_transform(ti)
    cc = lib.faf_cwt_config_init(4096)
    t = lib.faf_create_cwt(C.byref(cc))
    ti = lib.faf_create_icwt(C.byref(cc), 1)
    assert t and ti, lib.faf_get_error()
    try:
        ns, nr = lib.faf_cwt_n_scales(t), lib.faf_cwt_n_rows(t)
        freqs = np.zeros(ns)
        lib.faf_cwt_freqs(t, freqs.ctypes.data_as(C.POINTER(C.c_double)), ns)
        x = np.sin(2 * np.pi * freqs[ns // 2] * np.arange(4096)).astype(np.float32)
        dx = G.dev(x)
        w = G.empty((nr, 4096))
        y = G.empty(4096)
        torch.cuda.synchronize()
        fa.check(lib.faf_execute(t, C.byref(fa.buffer_real(w.data_ptr(), nr * 4096)),
                                 C.byref(fa.buffer_real(dx.data_ptr(), 4096))), "cwt")
        fa.check(lib.faf_execute(ti, C.byref(fa.buffer_real(y.data_ptr(), 4096)),
                                 C.byref(fa.buffer_real(w.data_ptr(), nr * 4096))), "icwt l1")
        assert rel_l2(G.host(y), x) <= 0.25
    finally:
        lib.faf_destroy_transform(t)
        lib.faf_destroy_transform(ti)


@pytest.mark.parametrize("n", [256, 1024, 4096, 16384])
def test_generic_and_specialised_kernels_agree(n):
    """Power-of-two lengths run on compile-time specialised kernels by
    default; the runtime-generic radix kernels must give the same answers."""
    say(f"case: generic vs specialised n={n}")
    lib = G.L()
    rng = np.random.default_rng(n + 7)
    x = cplx(rng, 3, n)
    xr = rng.standard_normal((3, 2 * n)).astype(np.float32)
    out = {}
    for spec in (1, 0):
        lib.faf_amd_set_specialised(spec)
        t = G.create("fft", n)
        ti = G.create("fft", n, dir="inverse")
        r = G.create("rfft", 2 * n)
        ri = G.create("rfft", 2 * n, dir="inverse")
        try:
            y = G.c2c(t, x)
            X = G.r2c(r, xr)
            out[spec] = (y, G.c2c(ti, y), X, G.c2r(ri, X, 2 * n))
        finally:
            for h in (t, ti, r, ri):
                G.destroy(h)
    lib.faf_amd_set_specialised(1)
    truth = truth_fft(x)
    rtruth = np.stack([O.fft64(v.astype(np.float64))[0][: n + 1] for v in xr])
    for spec in (1, 0):
        y, z, X, w = out[spec]
        assert rel_l2(y, truth) < tol(n), spec
        assert rel_l2(z, x) < 2 * tol(n), spec
        assert rel_l2(X, rtruth) < tol(2 * n), spec
        assert rel_l2(w, xr) < 2 * tol(2 * n), spec


def test_autotune_keeps_results():
    say("case: autotune")
    lib = G.L()
    n = 4096
    t = G.create("rfft", n)
    try:
        rng = np.random.default_rng(1)
        x = rng.standard_normal((64, n)).astype(np.float32)
        before = G.r2c(t, x)
        spb = lib.faf_amd_autotune(t, 4096)
        assert spb >= 1, lib.faf_get_error()
        after = G.r2c(t, x)
        assert rel_l2(after, before) < tol(n)
    finally:
        G.destroy(t)


# ------------------------------------------- faf_execute_f32 and single-frame STFT

@pytest.mark.parametrize("n", [1000, 4096])
def test_execute_f32_interleaved_wrapper(n):
    """faf_execute_f32 (reference src/faf_vm.c:976-1083): interleaved complex
    in/out for FFT/IFFT, REAL -> interleaved n/2+1 bins for RFFT and back for
    IRFFT; one signal, on device pointers and on host pointers."""
    say(f"case: faf_execute_f32 n={n}")
    lib = G.L()
    rng = np.random.default_rng(n + 5)
    x = cplx(rng, 1, n)[0]
    xr = rng.standard_normal(n).astype(np.float32)
    nb = n // 2 + 1
    t = G.create("fft", n, layout="interleaved")
    ti = G.create("fft", n, layout="interleaved", dir="inverse")
    r = G.create("rfft", n, layout="interleaved")
    ri = lib.faf_create_inverse(r)
    assert ri, lib.faf_get_error()
    try:
        for on_dev in (True, False):
            xin = x.view(np.float32).copy()
            if on_dev:
                a, y, z = G.dev(xin), G.empty(2 * n), G.empty(2 * n)
                pa, py, pz = a.data_ptr(), y.data_ptr(), z.data_ptr()
            else:
                a, y, z = xin, np.empty(2 * n, np.float32), np.empty(2 * n, np.float32)
                pa, py, pz = a.ctypes.data, y.ctypes.data, z.ctypes.data
            fa.check(lib.faf_execute_f32(t, py, pa), "f32 fft")
            fa.check(lib.faf_execute_f32(ti, pz, py), "f32 ifft")
            yh = (G.host(y) if on_dev else y).view(np.complex64)
            zh = (G.host(z) if on_dev else z).view(np.complex64)
            assert rel_l2(yh, truth_fft(x[None])[0]) < tol(n), on_dev
            assert rel_l2(zh, x) < 2 * tol(n), on_dev
            # real: REAL n in, interleaved bins out, and back
            if on_dev:
                a2, X, w = G.dev(xr), G.empty(2 * nb), G.empty(n)
                p2, pX, pw = a2.data_ptr(), X.data_ptr(), w.data_ptr()
            else:
                a2, X, w = xr.copy(), np.empty(2 * nb, np.float32), np.empty(n, np.float32)
                p2, pX, pw = a2.ctypes.data, X.ctypes.data, w.ctypes.data
            fa.check(lib.faf_execute_f32(r, pX, p2), "f32 rfft")
            fa.check(lib.faf_execute_f32(ri, pw, pX), "f32 irfft")
            Xh = (G.host(X) if on_dev else X).view(np.complex64)
            wh = G.host(w) if on_dev else w
            assert rel_l2(Xh, O.fft64(xr.astype(np.float64))[0][:nb]) < tol(n), on_dev
            assert rel_l2(wh, xr) < 2 * tol(n), on_dev
        # a SPLIT transform is refused by the interleaved wrapper
        s_ = G.create("fft", n)
        try:
            assert lib.faf_execute_f32(s_, py, pa) != 0
            assert b"INTERLEAVED" in lib.faf_get_error()
        finally:
            G.destroy(s_)
    finally:
        for h in (t, ti, r, ri):
            G.destroy(h)


@pytest.mark.parametrize("n,win", [(1000, 800), (1024, 1024), (4096, 3000)])
def test_single_frame_stft_reference_semantics(n, win):
    """faf_create_stft with INTERLEAVED layout keeps the reference's meaning
    (src/faf_core.c:994-1032, 1369-1383): one forward C2C of length n of the
    input times a symmetric Hann window of win_length samples, zero past it;
    dir is forced forward; batched rows each get the same window."""
    say(f"case: single-frame stft n={n} win={win}")
    lib = G.L()
    rng = np.random.default_rng(win)
    x = cplx(rng, 3, n)
    w = np.zeros(n, np.float64)
    w[:win] = 0.5 * (1.0 - np.cos(2.0 * np.pi * np.arange(win) / (win - 1)))
    t = G.create("stft", n, layout="interleaved", win_length=win, dir="inverse")
    try:
        assert lib.faf_transform_name(C.cast(t, C.POINTER(fa.FafTransform)).contents.type) \
            == b"stft"
        y = G.c2c(t, x, inter=True)
        truth = O.fft64((x.astype(np.complex128) * w[None, :]))
        assert rel_l2(y, truth) < tol(n)
        # the symmetric Hann the window comes from matches the oracle's
        assert np.allclose(O.hann32(win), w[:win].astype(np.float32), atol=1e-7)
        y1 = G.c2c(t, x[1:2], inter=True)
        assert np.array_equal(y1[0], y[1])                      # alone == in the batch
    finally:
        G.destroy(t)


# ------------------------------------------------------------ dual-frame ICWT

def _icwt_create(n, fs, voices=None, fmin=None, fmax=None):
    lib = G.L()
    cc = lib.faf_cwt_config_init(n)
    cc.fs = fs
    if voices:
        cc.voices = voices
    if fmin:
        cc.fmin = fmin
    if fmax:
        cc.fmax = fmax
    t = lib.faf_create_icwt(C.byref(cc), 0)
    assert t, lib.faf_get_error()
    return t


def _run_icwt(t, w, count=1, in_dist=0):
    """w: host float32 rows [count][rows][n] (or with in_dist padding)."""
    lib = G.L()
    n = t_n = C.cast(t, C.POINTER(fa.FafTransform)).contents.n
    rows = lib.faf_cwt_n_rows(t)
    dw = G.dev(np.ascontiguousarray(w, dtype=np.float32).ravel())
    y = G.empty((count, n))
    torch.cuda.synchronize()
    bi = fa.buffer_real(dw.data_ptr(), rows * t_n)
    bo = fa.buffer_real(y.data_ptr(), n)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), count, in_dist, 0), "icwt")
    return G.host(y)


@pytest.mark.parametrize("n", [512, 4096])
def test_icwt_dual_matches_reference(golden, n):
    """The device dual-frame inverse of the reference's own rows against the
    compiled reference's FP32 dual inverse of the same rows and the oracle's
    FP64 dual inverse (bar 1e-6 * log2 n)."""
    say(f"case: icwt dual vs reference n={n}")
    if n == 512:
        t = _icwt_create(512, 1.0)
        w, psi, ref = golden["cwt512_rows"], golden["cwt512_psi"], golden["cwt512_dual_out"]
    else:
        f0 = 62.5
        t = _icwt_create(4096, 16000.0, 11, f0, f0 * 2 ** (63.5 / 11))
        w, psi, ref = (golden["cwt4096_seg_rows"], golden["cwt4096_psi"],
                       golden["cwt4096_dual_out"])
    try:
        y = _run_icwt(t, w)[0]
        assert rel_l2(y, ref) < tol(n)
        assert rel_l2(y, O.icwt_dual64(w, psi)) < tol(n)
    finally:
        G.L().faf_destroy_transform(t)


@pytest.mark.parametrize("n", [32768, 65536])
def test_icwt_dual_long_segments(n):
    """Segments longer than one workgroup holds: batched row rffts, the dual
    accumulation kernel and a batched irfft.  Device forward rows of two
    segments go back through the device dual inverse; checked against the
    oracle's FP64 dual inverse of the same rows and as a reconstruction
    (reference tests/test_cwt.cpp:348-399, <= 1e-4 in FP32); also with a
    padded segment distance (rows not one contiguous batch)."""
    say(f"case: icwt dual long n={n}")
    lib = G.L()
    fs = 16000.0
    tf = _cwt_create(n, fs)
    ti = lib.faf_create_inverse(tf)
    assert ti, lib.faf_get_error()
    try:
        rows = lib.faf_cwt_n_rows(tf)
        tab, stride = _table(tf, lib)
        rng = np.random.default_rng(n)
        x = rng.standard_normal((2, n)).astype(np.float32)
        dx = G.dev(x)
        wd = G.empty((2, rows, n))
        torch.cuda.synchronize()
        bi = fa.buffer_real(dx.data_ptr(), n)
        bo = fa.buffer_real(wd.data_ptr(), rows * n)
        fa.check(lib.faf_execute_batch(tf, C.byref(bo), C.byref(bi), 2, 0, 0), "cwt rows")
        w = G.host(wd)
        y = _run_icwt(ti, w, count=2)
        for s in range(2):
            assert rel_l2(y[s], O.icwt_dual64(w[s], tab)) < tol(n), s
            assert rel_l2(y[s], x[s]) < 1e-4, s
        # padded distance between segments: per-segment row batches
        pad = rows * n + 64
        wp = np.zeros((2, pad), np.float32)
        wp[:, :rows * n] = w.reshape(2, -1)
        yp = _run_icwt(ti, wp, count=2, in_dist=pad)
        assert np.array_equal(yp, y)
    finally:
        lib.faf_destroy_transform(tf)
        lib.faf_destroy_transform(ti)


def test_serving_bank_matches_reference(golden):
    """The bank the serving loop builds for CWT requests (n=1024, fs=16 kHz,
    8 voices, 32 scales from 16 fs/n) against the compiled reference's
    tables, frequencies and report."""
    say("case: serving cwt bank vs reference")
    lib = G.L()
    n, fs = 1024, 16000.0
    fmin = 16 * fs / n
    t = _cwt_create(n, fs, 8, fmin, fmin * 2 ** (31.5 / 8))
    try:
        assert lib.faf_cwt_n_scales(t) == 32
        tab, _ = _table(t, lib)
        ref = golden["cwt1024v8_psi"]
        assert rel_l2(tab[:, :ref.shape[1]], ref) < 1e-6
        fr = np.zeros(32)
        lib.faf_cwt_freqs(t, fr.ctypes.data_as(C.POINTER(C.c_double)), 32)
        assert np.allclose(fr, golden["cwt1024v8_freqs"], rtol=1e-12)
        rep = lib.faf_cwt_bank_report(t).contents
        g = golden["cwt1024v8_report"]
        assert rep.n_scales == g[0] and rep.passed == g[1] and rep.k_lo == g[7]
        assert abs(rep.frame_cond - g[6]) < 1e-9
    finally:
        lib.faf_destroy_transform(t)


def test_single_frame_stft_matches_reference_golden(golden):
    say("case: single-frame stft vs reference golden")
    z = golden["stft1_1000_in"]
    t = G.create("stft", 1000, layout="interleaved", win_length=800)
    try:
        y = G.c2c(t, z[None], inter=True)[0]
        assert rel_l2(y, golden["stft1_1000_out"]) < tol(1000)
    finally:
        G.destroy(t)


# ------------------------------------- Rader and Bluestein beyond one workgroup

@pytest.mark.parametrize("n,blue", [(12289, False), (30241, False), (40961, False),
                                    (65537, False), (10007, True), (30011, True)])
def test_rader_bluestein_beyond_one_workgroup(n, blue):
    """Primes whose Rader inner length p-1 (12288 in one workgroup but not
    beside the signal; 30240 = 2^5 3^3 5 7 four-step with radix 7; 40960 and
    65536 four-step) or Bluestein padded length M (20250, 60750) is too long
    for one workgroup: gather / chirp, inner FFTs through HBM, scatter /
    chirp.  Forward and inverse (ortho), a batch of 3 with padded rows,
    against FP64; batch members equal the signal alone."""
    say(f"case: {'bluestein' if blue else 'rader'} n={n} beyond one workgroup")
    lib = G.L()
    rng = np.random.default_rng(n)
    flags = fa.FLAG_BLUESTEIN if blue else 0
    x = cplx(rng, 3, n)
    t = G.create("fft", n, flags=flags)
    ti = G.create("fft", n, flags=flags, dir="inverse", norm="ortho")
    try:
        # padded row distance: signal i starts i * (n + 5) elements in
        d = n + 5
        xp = np.zeros((3, d), np.complex64)
        xp[:, :n] = x
        xr, xi = G.dev(xp.real.copy()), G.dev(xp.imag.copy())
        yr, yi = G.empty((3, d)), G.empty((3, d))
        bi = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
        bo = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), 3, d, d), "fwd")
        y = (G.host(yr) + 1j * G.host(yi))[:, :n]
        truth = truth_fft(x)
        assert rel_l2(y, truth) < tol(n)
        y1 = G.c2c(t, x[2:3])
        assert np.array_equal(y1[0], y[2])
        z = G.c2c(ti, y)
        assert rel_l2(z, x * np.sqrt(n)) < 2 * tol(n)        # ortho inverse of unscaled fwd
    finally:
        G.destroy(t)
        G.destroy(ti)


# compile-time pass sequences the ragged kernel carries per length: the
# planner's order first, then the composite-radix ones (prime-factor codelets)
RAGGED_CT_SEQS = {
    4800: [[16, 4, 3, 5, 5], [12, 20, 20]],
    3000: [[8, 3, 5, 5, 5], [10, 20, 15], [30, 10, 10]],
    1000: [[8, 5, 5, 5], [10, 10, 10]],
    360: [[8, 3, 3, 5], [24, 15]],
    60: [[4, 3, 5], [15, 4]],
}
RAGGED_RADER_SEQS = {
    11: [[2, 5], [10]],
    17: [[16]],
    31: [[2, 3, 5], [6, 5]],
    97: [[16, 2, 3], [8, 12]],
    1009: [[16, 3, 3, 7], [12, 12, 7]],
}


def test_ragged_every_compiled_sequence():
    """Every compile-time pass sequence of every C3 length, forced through
    faf_ragged_set_choice, in a batch of 300 signals of that length plus a
    mixed batch: every checked signal matches FP64, and faf_ragged_choice
    reports the sequence that ran.  The records are forgotten afterwards."""
    lib = G.L()
    rng = np.random.default_rng(31)
    cases = [(n, 3, o) for n, seqs in RAGGED_CT_SEQS.items() for o in seqs]
    cases += [(p, 4, o) for p, seqs in RAGGED_RADER_SEQS.items() for o in seqs]
    try:
        for n, kind, order in cases:
            say(f"case: ragged sequence n={n} kind={kind} order={order}")
            arr_o = (C.c_int * len(order))(*order)
            fa.check(lib.faf_ragged_set_choice(n, kind, 0, arr_o, len(order)), "set choice")
            lens = [n] * 300 + [int(v) for v in rng.choice([8192, 17, 4096], 40)]
            rng.shuffle(lens)
            arr = (C.c_size_t * len(lens))(*lens)
            r = lib.faf_ragged_create(arr, len(lens), 0)
            assert r, lib.faf_get_error()
            try:
                k_, spb_, no_ = C.c_int(), C.c_int(), C.c_int()
                ord_ = (C.c_int * 16)()
                fa.check(lib.faf_ragged_choice(r, n, C.byref(k_), C.byref(spb_), ord_,
                                               C.byref(no_)), "choice")
                assert k_.value == kind and list(ord_[: no_.value]) == order
                off = np.concatenate([[0], np.cumsum(lens)])
                x = (rng.uniform(-1, 1, int(off[-1])) +
                     1j * rng.uniform(-1, 1, int(off[-1]))).astype(np.complex64)
                dxr, dxi = G.dev(x.real.copy()), G.dev(x.imag.copy())
                yr, yi = G.empty(int(off[-1])), G.empty(int(off[-1]))
                fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(),
                                                dxr.data_ptr(), dxi.data_ptr()), "ragged")
                y = G.host(yr) + 1j * G.host(yi)
                for i in list(range(0, len(lens), 11)) + [len(lens) - 1]:
                    a, e = off[i], off[i + 1]
                    truth = O.fft64(x[a:e].astype(np.complex128))[0]
                    assert rel_l2(y[a:e], truth) < tol(lens[i]), (n, order, i, lens[i])
            finally:
                lib.faf_ragged_destroy(r)
        # an order the kernel does not carry is refused with a message
        bad = (C.c_int * 3)(20, 12, 20)
        assert lib.faf_ragged_set_choice(4800, 3, 0, bad, 3) == -1
        assert b"order" in lib.faf_get_error()
    finally:
        for n, kind, _ in cases:
            lib.faf_ragged_set_choice(n, -1, 0, None, 0)
