"""GPU tests of faf_execute_batch_noalloc, the batched execute that allocates
nothing: its workspace is the library's static per-GPU scratch.

Each case runs the same plan through faf_execute_batch and through the
allocation-free call on the same device buffers and requires the same bits
(the same kernels run; only where the workspace lives and how a batch is cut
into chunks differ, and chunks hold whole signals).  Every case makes two
calls in a row, the second with a smaller batch, so the second call runs on
scratch that still holds the first call's contents.  Plans here are the ones
whose runs need workspace: four-step C2C 65536 (specialised column/row
passes) and 20000 (runtime passes), and the four-step fallback of rfft/irfft
65536 on rows that are not 8-byte aligned.
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

SCRATCH = 32 << 20


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def both(fn):
    """fn(call) runs one execute with `call` and returns the output planes;
    returns (faf_execute_batch result, faf_execute_batch_noalloc result)."""
    lib = G.L()
    a = fn(lib.faf_execute_batch)
    b = fn(lib.faf_execute_batch_noalloc)
    return a, b


@pytest.mark.parametrize("n,counts", [(65536, (70, 5)), (20000, (230, 17))])
def test_noalloc_c2c_four_step_matches_batch(n, counts):
    t = G.create("fft", n)
    try:
        for cnt in counts:
            say(f"case: noalloc c2c n={n} count={cnt}")
            rng = np.random.default_rng(cnt)
            x = (rng.uniform(-1, 1, (cnt, n)) + 1j * rng.uniform(-1, 1, (cnt, n))).astype(np.complex64)
            xr, xi = G.dev(x.real.copy()), G.dev(x.imag.copy())
            yr, yi = G.empty((cnt, n)), G.empty((cnt, n))
            bi = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
            bo = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)

            def run(call):
                yr.fill_(float("nan"))
                yi.fill_(float("nan"))
                torch.cuda.synchronize()
                fa.check(call(t, C.byref(bo), C.byref(bi), cnt, 0, 0), "c2c")
                return G.host(yr).copy(), G.host(yi).copy()

            (ar, ai), (br, bi_) = both(run)
            assert np.array_equal(ar.view(np.uint32), br.view(np.uint32))
            assert np.array_equal(ai.view(np.uint32), bi_.view(np.uint32))
            # and the result is the transform (first and last signal)
            for k in (0, cnt - 1):
                truth = O.fft64(x[k].astype(np.complex128))
                assert rel_l2(br[k] + 1j * bi_[k], truth) < tol(n)
    finally:
        G.destroy(t)


def test_noalloc_rfft_irfft_unaligned_rows_match_batch():
    """rfft / irfft 65536 on odd row distances take the four-step kernels,
    which need workspace."""
    n, nb = 65536, 65536 // 2 + 1
    t = G.create("rfft", n)
    ti = G.L().faf_create_inverse(t)
    assert ti
    try:
        for cnt in (131, 3):
            say(f"case: noalloc rfft/irfft n={n} count={cnt} odd dist")
            rng = np.random.default_rng(cnt)
            idist = n + 1
            x = rng.standard_normal((cnt, idist)).astype(np.float32)
            dx = G.dev(x)
            re, im = G.empty((cnt, nb)), G.empty((cnt, nb))
            y = G.empty((cnt, idist))
            bx = fa.buffer_real(dx.data_ptr(), n)
            bh = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
            by = fa.buffer_real(y.data_ptr(), n)

            def fwd(call):
                re.fill_(float("nan"))
                im.fill_(float("nan"))
                torch.cuda.synchronize()
                fa.check(call(t, C.byref(bh), C.byref(bx), cnt, idist, 0), "rfft")
                return G.host(re).copy(), G.host(im).copy()

            (ar, ai), (br, bi) = both(fwd)
            assert np.array_equal(ar.view(np.uint32), br.view(np.uint32))
            assert np.array_equal(ai.view(np.uint32), bi.view(np.uint32))
            truth = O.fft64(x[cnt - 1, :n].astype(np.complex128))[0][:nb]
            assert rel_l2(br[cnt - 1] + 1j * bi[cnt - 1], truth) < tol(n)

            def inv(call):
                y.fill_(float("nan"))
                torch.cuda.synchronize()
                fa.check(call(ti, C.byref(by), C.byref(bh), cnt, 0, idist), "irfft")
                return G.host(y).copy()

            ya, yb = both(inv)
            assert np.array_equal(ya[:, :n].view(np.uint32), yb[:, :n].view(np.uint32))
            assert rel_l2(yb[0, :n], x[0, :n]) < 2 * tol(n)
    finally:
        G.destroy(ti)
        G.destroy(t)


def test_noalloc_refuses_what_would_allocate():
    lib = G.L()
    n = 65536
    t = G.create("fft", n)
    try:
        say("case: noalloc host buffers refused")
        xh = np.zeros((2, n), np.float32)
        yh = np.zeros((2, n), np.float32)
        bi = fa.buffer_split(xh.ctypes.data, xh[1:].ctypes.data, n)
        bo = fa.buffer_split(yh.ctypes.data, yh[1:].ctypes.data, n)
        assert lib.faf_execute_batch_noalloc(t, C.byref(bo), C.byref(bi), 1, 0, 0) != 0
        assert b"device memory" in lib.faf_get_error()
    finally:
        G.destroy(t)
    # a single-frame windowed STFT copies its windowed input into workspace:
    # count * n * 8 bytes, which past 32 MiB the scratch cannot hold
    say("case: noalloc run larger than the scratch refused")
    w = 4096
    ts = G.create("stft", w, layout="interleaved")
    try:
        cnt = SCRATCH // (w * 8) + 1
        x = G.empty((cnt, 2 * w))
        y = G.empty((cnt, 2 * w))
        bi = fa.buffer_interleaved(x.data_ptr(), w)
        bo = fa.buffer_interleaved(y.data_ptr(), w)
        torch.cuda.synchronize()
        assert lib.faf_execute_batch_noalloc(ts, C.byref(bo), C.byref(bi), cnt, 0, 0) != 0
        msg = lib.faf_get_error()
        assert b"static scratch" in msg, msg
        # one signal fewer than the limit fits, and matches the allocating call
        ok = SCRATCH // (w * 8)
        x.uniform_(-1, 1)
        fa.check(lib.faf_execute_batch_noalloc(ts, C.byref(bo), C.byref(bi), ok, 0, 0), "stft")
        a = G.host(y[:ok]).copy()
        fa.check(lib.faf_execute_batch(ts, C.byref(bo), C.byref(bi), ok, 0, 0), "stft")
        b = G.host(y[:ok]).copy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        G.destroy(ts)
