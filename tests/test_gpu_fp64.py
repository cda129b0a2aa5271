"""GPU tests of the FP64 execution variant (precision FAF_PREC_FP64).

The reference plans FFT, RFFT and the DCT family in FP64 and holds its FP64
rfft round trip to 1e-12 (tests/test_rfft.cpp:239-261).  Here every FP64
transform is checked against the oracle's FP64 truth (fft64 / r2r64 /
mdct64) at relative L2 <= 1e-13 * log2(N) -- five orders tighter than the
FP32 bar -- and round trips to 1e-12, on double buffers in device memory.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import fastandfourier_amd as fa
import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import gpu_util as G  # noqa: E402

FP64 = 4


def tol64(n):
    return 1e-13 * max(np.log2(max(n, 2)), 1.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def empty64(shape):
    return torch.empty(shape, dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def create(kind, n, **kw):
    lib = G.L()
    cfg = fa.config_init(n, **kw)
    cfg.precision = FP64
    fn = {"fft": lib.faf_create_fft, "rfft": lib.faf_create_rfft, "dct": lib.faf_create_dct,
          "mdct": lib.faf_create_mdct, "stft": lib.faf_create_stft}[kind]
    return fn(C.byref(cfg))


def c2c64(t, x, inter=False):
    lib = G.L()
    x = np.atleast_2d(x)
    cnt, n = x.shape
    if inter:
        xi = dev64(np.stack([x.real, x.imag], -1).reshape(cnt, 2 * n))
        yo = empty64((cnt, 2 * n))
        bi = fa.buffer_interleaved(xi.data_ptr(), n)
        bo = fa.buffer_interleaved(yo.data_ptr(), n)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "c2c64")
        y = host(yo).reshape(cnt, n, 2)
        return y[..., 0] + 1j * y[..., 1]
    xr, xi = dev64(x.real), dev64(x.imag)
    yr, yi = empty64((cnt, n)), empty64((cnt, n))
    bi = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
    bo = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "c2c64")
    return host(yr) + 1j * host(yi)


SIZES = [2, 7, 8, 12, 64, 360, 1000, 4096, 4800, 8192, 20000, 65536, 11, 97, 1009, 12289]


@pytest.mark.parametrize("n", SIZES)
def test_fp64_c2c_matches_fp64_truth(n):
    say(f"case: fp64 c2c n={n}")
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (3, n)) + 1j * rng.uniform(-1, 1, (3, n))
    t = create("fft", n)
    ti = create("fft", n, dir="inverse")
    assert t and ti, G.L().faf_get_error()
    try:
        y = c2c64(t, x)
        assert rel_l2(y, O.fft64(x)) < tol64(n)
        z = c2c64(ti, y)
        assert rel_l2(z, x) < 2 * tol64(n)
        if n not in (11, 97, 1009, 12289):                  # Rader needs SPLIT
            ti2 = create("fft", n, layout="interleaved", norm="ortho")
            try:
                assert rel_l2(c2c64(ti2, x, inter=True), O.fft64(x) / np.sqrt(n)) < tol64(n)
            finally:
                G.destroy(ti2)
    finally:
        G.destroy(t)
        G.destroy(ti)


def test_fp64_bluestein():
    say("case: fp64 bluestein 23 and 10007")
    for n in (23, 10007):
        rng = np.random.default_rng(n)
        x = rng.uniform(-1, 1, (2, n)) + 1j * rng.uniform(-1, 1, (2, n))
        t = create("fft", n, flags=fa.FLAG_BLUESTEIN)
        assert t, G.L().faf_get_error()
        try:
            assert rel_l2(c2c64(t, x), O.fft64(x)) < tol64(n)
        finally:
            G.destroy(t)


@pytest.mark.parametrize("n", [256, 1024, 4096, 16384, 65536])
def test_fp64_rfft_irfft(n):
    """Forward against FP64 truth; the round trip at 1e-12 (the reference's
    own FP64 bar, tests/test_rfft.cpp:239-261, restated on its 4-cycle sine
    at n=256 and on N(0,1) rows elsewhere); split and interleaved bins."""
    say(f"case: fp64 rfft n={n}")
    lib = G.L()
    nb = n // 2 + 1
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n))
    if n == 256:
        x[0] = np.sin(2 * np.pi * 4 * np.arange(n) / n)
    t = create("rfft", n)
    assert t, lib.faf_get_error()
    ti = lib.faf_create_inverse(t)
    assert ti, lib.faf_get_error()
    try:
        dx = dev64(x)
        re, im = empty64((3, nb)), empty64((3, nb))
        y = empty64((3, n))
        bi = fa.buffer_real(dx.data_ptr(), n)
        bh = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
        by = fa.buffer_real(y.data_ptr(), n)
        fa.check(lib.faf_execute_batch(t, C.byref(bh), C.byref(bi), 3, 0, 0), "rfft64")
        fa.check(lib.faf_execute_batch(ti, C.byref(by), C.byref(bh), 3, 0, 0), "irfft64")
        X = host(re) + 1j * host(im)
        assert rel_l2(X, O.fft64(x)[:, :nb]) < tol64(n)
        assert np.max(np.abs(host(y) - x)) < 1e-12
        # interleaved bins
        ti_ = create("rfft", n, layout="interleaved")
        try:
            Xi = empty64((3, 2 * nb))
            bo = fa.buffer_interleaved(Xi.data_ptr(), nb)
            fa.check(lib.faf_execute_batch(ti_, C.byref(bo), C.byref(bi), 3, 0, 0), "rfft64 i")
            v = host(Xi).reshape(3, nb, 2)
            assert rel_l2(v[..., 0] + 1j * v[..., 1], X) < tol64(n)
        finally:
            G.destroy(ti_)
    finally:
        G.destroy(ti)
        G.destroy(t)


def test_fp64_dct_and_mdct():
    say("case: fp64 dct / mdct")
    lib = G.L()
    rng = np.random.default_rng(4)
    for typ in (1, 2, 3, 4):
        n = 100
        x = rng.standard_normal((2, n))
        cfg = fa.config_init(n, norm="ortho")
        cfg.precision = FP64
        cfg.dct_type = typ
        t = lib.faf_create_dct(C.byref(cfg))
        assert t, lib.faf_get_error()
        try:
            dx, y = dev64(x), empty64((2, n))
            bi = fa.buffer_real(dx.data_ptr(), n)
            bo = fa.buffer_real(y.data_ptr(), n)
            fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), 2, 0, 0), "dct64")
            assert rel_l2(host(y), O.r2r64(x, "dct", typ, "ortho")) < tol64(2 * n)
        finally:
            G.destroy(t)
    n = 512
    x = rng.standard_normal((2, n))
    t = create("mdct", n)
    try:
        dx, X = dev64(x), empty64((2, n // 2))
        bi = fa.buffer_real(dx.data_ptr(), n)
        bo = fa.buffer_real(X.data_ptr(), n // 2)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), 2, 0, 0), "mdct64")
        assert rel_l2(host(X), O.mdct64(x)) < tol64(n)
    finally:
        G.destroy(t)


def test_fp64_refusals():
    say("case: fp64 refusals")
    lib = G.L()
    assert not create("stft", 1024, layout="real")
    assert b"FP64" in lib.faf_get_error()
    t = create("fft", 64, layout="interleaved")
    try:
        a = empty64(128)
        assert lib.faf_execute_f32(t, a.data_ptr(), a.data_ptr()) != 0
        assert b"FP64" in lib.faf_get_error()
    finally:
        G.destroy(t)
    # odd row distance: signals are read as complex pairs
    t = create("rfft", 64)
    try:
        x, re, im = empty64(200), empty64(100), empty64(100)
        bi = fa.buffer_real(x.data_ptr(), 64)
        bo = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), 33)
        assert lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), 2, 65, 33) != 0
        assert b"even row distance" in lib.faf_get_error()
    finally:
        G.destroy(t)
