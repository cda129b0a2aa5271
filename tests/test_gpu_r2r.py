"""GPU tests of DCT I-IV, DST I-IV and MDCT / IMDCT.

The reference names these transforms but computes an FFT in their place
(src/faf_transforms.c:195-222), so they are checked against the oracle's
FP64 direct sums of the definitions (oracle.r2r64 / mdct64 / imdct64, which
the CPU tests pin to scipy.fft and to closed forms), at relative L2 <=
1e-6 * log2(L) with L the inner FFT length (2N, 2(N-1) or 2(N+1)).
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

T_DCT = {1: 4, 2: 5, 3: 6, 4: 7}       # FAF_TRANSFORM_DCT_I .. IV
T_DST = {1: 8, 2: 9, 3: 10, 4: 11}
T_MDCT, T_IMDCT = 13, 14


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def create(kind, n, typ=2, norm="none", inverse=False):
    lib = G.L()
    cfg = fa.config_init(n, norm=norm, dir="inverse" if inverse else "forward")
    cfg.dct_type = typ
    fn = {"dct": lib.faf_create_dct, "dst": lib.faf_create_dst, "mdct": lib.faf_create_mdct}[kind]
    t = fn(C.byref(cfg))
    if not t:
        raise fa.FafError(lib.faf_get_error().decode())
    return t


def run(t, x, nout, in_dist=0, out_dist=0):
    """Batched REAL -> REAL execute over the rows of x (device buffers)."""
    lib = G.L()
    x = np.atleast_2d(np.asarray(x, np.float32))
    cnt, nin = x.shape
    if in_dist:
        xp = np.zeros((cnt, in_dist), np.float32)
        xp[:, :nin] = x
    else:
        xp = x
    dx = G.dev(xp)
    od = out_dist or nout
    y = G.empty((cnt, od))
    torch.cuda.synchronize()
    bi = fa.buffer_real(dx.data_ptr(), nin)
    bo = fa.buffer_real(y.data_ptr(), nout)
    fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), cnt, in_dist, out_dist), "r2r")
    return G.host(y)[:, :nout]


def inner_len(kind, typ, n):
    if typ == 1:
        return 2 * (n - 1) if kind == "dct" else 2 * (n + 1)
    return 2 * n


@pytest.mark.parametrize("n", [7, 64, 100, 1000])
@pytest.mark.parametrize("kind", ["dct", "dst"])
@pytest.mark.parametrize("typ", [1, 2, 3, 4])
def test_r2r_all_norms_and_directions(n, kind, typ):
    say(f"case: {kind}-{typ} n={n}")
    rng = np.random.default_rng(n * 7 + typ)
    x = rng.standard_normal((3, n)).astype(np.float32)
    bar = tol(inner_len(kind, typ, n))
    for norm in ("none", "ortho", "forward"):
        for inv in (False, True):
            t = create(kind, n, typ, norm, inv)
            try:
                y = run(t, x, n)
                truth = O.r2r64(x, kind, typ, norm, inverse=inv)
                assert rel_l2(y, truth) < bar, (norm, inv, rel_l2(y, truth))
            finally:
                G.destroy(t)


@pytest.mark.parametrize("kind", ["dct", "dst"])
@pytest.mark.parametrize("typ", [1, 2, 3, 4])
def test_r2r_create_inverse_round_trip_and_batches(kind, typ):
    """faf_create_inverse swaps II <-> III and flips dir (as the reference):
    forward then inverse gives x back; padded row distances and a batch
    member alone give the same bits."""
    say(f"case: {kind}-{typ} inverse round trip")
    lib = G.L()
    n = 4096
    rng = np.random.default_rng(typ)
    x = rng.standard_normal((4, n)).astype(np.float32)
    t = create(kind, n, typ)
    ti = lib.faf_create_inverse(t)
    assert ti, lib.faf_get_error()
    try:
        y = run(t, x, n, in_dist=n + 3, out_dist=n + 5)
        y1 = run(t, x[2:3], n)
        assert np.array_equal(y1[0], y[2])
        assert rel_l2(y, O.r2r64(x, kind, typ)) < tol(inner_len(kind, typ, n))
        z = run(ti, y, n)
        assert rel_l2(z, x) < 4 * tol(inner_len(kind, typ, n))
    finally:
        G.destroy(ti)
        G.destroy(t)


@pytest.mark.parametrize("n", [1009, 65536])
def test_dct2_long_and_prime_lengths(n):
    """Sizes are unrestricted: 2N = 2018 runs through Bluestein, 2N = 131072
    four-step.  Checked on sampled output bins by direct FP64 sums and by
    the round trip through faf_create_inverse."""
    say(f"case: dct-2 n={n}")
    lib = G.L()
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)).astype(np.float32)
    t = create("dct", n, 2, "ortho")
    ti = lib.faf_create_inverse(t)
    try:
        y = run(t, x, n)
        ks = np.array([0, 1, 2, n // 3, n // 2, n - 2, n - 1])
        c, _ = O._cs(ks[:, None] * (2 * np.arange(n)[None, :] + 1), 2 * n)
        w = np.full(len(ks), np.sqrt(2.0 / n))
        w[ks == 0] = np.sqrt(1.0 / n)
        for r in range(2):
            truth = (c @ x[r].astype(np.float64)) * w
            assert rel_l2(y[r, ks], truth) < 10 * tol(2 * n)
        z = run(ti, y, n)
        assert rel_l2(z, x) < 4 * tol(2 * n)
    finally:
        G.destroy(ti)
        G.destroy(t)


@pytest.mark.parametrize("n", [64, 1024, 480])
def test_mdct_imdct_and_tdac(n):
    """MDCT (n -> n/2) and IMDCT (n/2 -> n) against the oracle for every
    norm; frames of a signal through the device MDCT, IMDCT, sine window
    and 50% overlap-add reconstruct it (time-domain aliasing cancellation)."""
    say(f"case: mdct n={n}")
    lib = G.L()
    M = n // 2
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)).astype(np.float32)
    for norm in ("none", "ortho", "forward"):
        t = create("mdct", n, norm=norm)
        ti = lib.faf_create_inverse(t)
        assert ti, lib.faf_get_error()
        try:
            X = run(t, x, M)
            assert rel_l2(X, O.mdct64(x, norm)) < tol(n), norm
            y = run(ti, X, n)
            assert rel_l2(y, O.imdct64(X, norm)) < tol(n), norm
            if norm == "none":
                frames = 6
                sig = rng.standard_normal((frames + 1) * M).astype(np.float32)
                w = np.sin(np.pi * (np.arange(n) + 0.5) / n).astype(np.float32)
                seg = np.stack([sig[b * M: b * M + n] * w for b in range(frames)])
                yy = run(ti, run(t, seg, M), n)
                out = np.zeros_like(sig, dtype=np.float64)
                for b in range(frames):
                    out[b * M: b * M + n] += w * yy[b]
                assert rel_l2(out[M:frames * M], sig[M:frames * M]) < 4 * tol(n)
        finally:
            G.destroy(ti)
            G.destroy(t)


def test_r2r_create_paths_and_errors():
    say("case: r2r create paths and errors")
    lib = G.L()
    n = 60
    x = np.random.default_rng(3).standard_normal((1, n)).astype(np.float32)
    # faf_create with a DCT/DST type sets the type from it
    for typ in (1, 2, 3, 4):
        for kind, tab in (("dct", T_DCT), ("dst", T_DST)):
            t = lib.faf_create(tab[typ], C.byref(fa.config_init(n)))
            assert t, lib.faf_get_error()
            try:
                assert rel_l2(run(t, x, n), O.r2r64(x, kind, typ)) < tol(2 * n)
            finally:
                G.destroy(t)
    t = lib.faf_create(T_IMDCT, C.byref(fa.config_init(n)))
    assert t
    assert C.cast(t, C.POINTER(fa.FafTransform)).contents.type == T_IMDCT
    G.destroy(t)
    # faf_execute_f32 takes REAL arrays
    t = create("dct", n, 2)
    try:
        dx, y = G.dev(x[0]), G.empty(n)
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_f32(t, y.data_ptr(), dx.data_ptr()), "f32")
        assert rel_l2(G.host(y), O.r2r64(x, "dct", 2)[0]) < tol(2 * n)
    finally:
        G.destroy(t)
    # errors: type, odd MDCT, layout, buffer length
    cfg = fa.config_init(n)
    cfg.dct_type = 5
    assert not lib.faf_create_dct(C.byref(cfg))
    assert b"type must be 1-4" in lib.faf_get_error()
    assert not lib.faf_create_mdct(C.byref(fa.config_init(63)))
    assert b"MDCT size must be even" in lib.faf_get_error()
    assert not lib.faf_create_dst(C.byref(fa.config_init(n, layout="split")))
    t = create("mdct", n)
    try:
        d = G.empty(n)
        bi = fa.buffer_real(d.data_ptr(), n)
        bo = fa.buffer_real(d.data_ptr(), n)              # MDCT output is n/2
        assert lib.faf_execute(t, C.byref(bo), C.byref(bi)) != 0
        assert b"expected" in lib.faf_get_error()
    finally:
        G.destroy(t)
