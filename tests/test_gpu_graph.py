"""Graph capture of executes (faf_graph_*) on the GPU.

Every replay is checked against an eager execute of the same plans on the
same inputs (bit-exact: the same kernels on the same data) and against the
FP64 oracle at relative L2 <= 1e-6 * log2(N).  Replays are interleaved with
the program's own eager work and synchronisations, and every replay gets
new inputs written into the captured buffers.
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


def put(dst, a):
    dst.copy_(torch.from_numpy(np.ascontiguousarray(a)))


def err():
    return G.L().faf_get_error().decode()


def capture(calls):
    """Capture the executes that calls() makes; returns the graph handle."""
    lib = G.L()
    fa.check(lib.faf_graph_begin(), "graph begin")
    rcs = calls()
    g = lib.faf_graph_end()
    assert all(rc == 0 for rc in rcs), err()
    assert g, err()
    return g


def test_graph_c1_forward_inverse_replays():
    """C1 shape (C2C 4096 x 256, split planes): forward then inverse in one graph."""
    say("case: graph C1 fwd+inv 4096x256")
    lib = G.L()
    n, b = 4096, 256
    rng = np.random.default_rng(0)
    t = G.create("fft", n)
    ti = G.create("fft", n, dir="inverse")
    xr, xi, yr, yi, zr, zi = (G.empty((b, n)) for _ in range(6))
    er, ei, fr, fi = (G.empty((b, n)) for _ in range(4))
    bx = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
    by = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
    bz = fa.buffer_split(zr.data_ptr(), zi.data_ptr(), n)
    be = fa.buffer_split(er.data_ptr(), ei.data_ptr(), n)
    bf = fa.buffer_split(fr.data_ptr(), fi.data_ptr(), n)
    g = None
    try:
        x = cplx(rng, b, n)
        put(xr, x.real)
        put(xi, x.imag)
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0), "eager fwd")
        fa.check(lib.faf_execute_batch(ti, C.byref(bz), C.byref(by), b, 0, 0), "eager inv")
        g = capture(lambda: [lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0),
                             lib.faf_execute_batch(ti, C.byref(bz), C.byref(by), b, 0, 0)])
        assert lib.faf_graph_nodes(g) >= 2
        for rep in range(4):
            x = cplx(rng, b, n)
            put(xr, x.real)
            put(xi, x.imag)
            # eager reference into other buffers, then the replay
            fa.check(lib.faf_execute_batch(t, C.byref(be), C.byref(bx), b, 0, 0), "eager")
            fa.check(lib.faf_execute_batch(ti, C.byref(bf), C.byref(be), b, 0, 0), "eager")
            fa.check(lib.faf_graph_launch(g), "graph launch")
            y = G.host(yr) + 1j * G.host(yi)
            z = G.host(zr) + 1j * G.host(zi)
            assert np.array_equal(y, G.host(er) + 1j * G.host(ei)), f"replay {rep}"
            assert np.array_equal(z, G.host(fr) + 1j * G.host(fi)), f"replay {rep}"
            sel = rng.choice(b, 4, replace=False)
            assert rel_l2(y[sel], O.fft64(x[sel].astype(np.complex128))) < tol(n)
            assert rel_l2(z, x) < 2 * tol(n)
    finally:
        lib.faf_graph_destroy(g)
        G.destroy(ti)
        G.destroy(t)


@pytest.mark.parametrize("n", [65536, 16384])
def test_graph_rfft_irfft_replays(n):
    """rfft then irfft in one graph (65536: the single-pass kernels)."""
    say(f"case: graph rfft+irfft n={n}")
    lib = G.L()
    b = 6
    nb = n // 2 + 1
    rng = np.random.default_rng(n + 5)
    t = G.create("rfft", n)
    ti = lib.faf_create_inverse(t)
    assert ti, err()
    x = G.empty((b, n))
    re, im, y = G.empty((b, nb)), G.empty((b, nb)), G.empty((b, n))
    ere, eim, ey = G.empty((b, nb)), G.empty((b, nb)), G.empty((b, n))
    bx = fa.buffer_real(x.data_ptr(), n)
    bs = fa.buffer_hermitian(re.data_ptr(), im.data_ptr(), nb)
    by = fa.buffer_real(y.data_ptr(), n)
    bes = fa.buffer_hermitian(ere.data_ptr(), eim.data_ptr(), nb)
    bey = fa.buffer_real(ey.data_ptr(), n)
    g = None
    try:
        put(x, rng.standard_normal((b, n)).astype(np.float32))
        torch.cuda.synchronize()
        # the eager run sizes the workspaces the capture will use
        fa.check(lib.faf_execute_batch(t, C.byref(bs), C.byref(bx), b, 0, 0), "eager")
        fa.check(lib.faf_execute_batch(ti, C.byref(by), C.byref(bs), b, 0, 0), "eager")
        g = capture(lambda: [lib.faf_execute_batch(t, C.byref(bs), C.byref(bx), b, 0, 0),
                             lib.faf_execute_batch(ti, C.byref(by), C.byref(bs), b, 0, 0)])
        for rep in range(3):
            xs = rng.standard_normal((b, n)).astype(np.float32)
            put(x, xs)
            fa.check(lib.faf_execute_batch(t, C.byref(bes), C.byref(bx), b, 0, 0), "eager")
            fa.check(lib.faf_execute_batch(ti, C.byref(bey), C.byref(bes), b, 0, 0), "eager")
            torch.cuda.synchronize()
            fa.check(lib.faf_graph_launch(g), "graph launch")
            X = G.host(re) + 1j * G.host(im)
            assert np.array_equal(X, G.host(ere) + 1j * G.host(eim)), f"replay {rep}"
            assert np.array_equal(G.host(y), G.host(ey)), f"replay {rep}"
            ref = O.fft64(xs.astype(np.complex128))[:, :nb]
            assert rel_l2(X, ref) < tol(n)
            assert rel_l2(G.host(y), xs) < 2 * tol(n)
    finally:
        lib.faf_graph_destroy(g)
        G.destroy(ti)
        G.destroy(t)


def test_graph_c2c_four_step_replays():
    """C2C 65536 runs as a column pass and a row pass per workspace chunk:
    a graph of several kernels sharing the plan's workspace."""
    say("case: graph c2c 65536 four-step")
    lib = G.L()
    n, b = 65536, 5
    rng = np.random.default_rng(n + 9)
    t = G.create("fft", n)
    xr, xi, yr, yi, er, ei = (G.empty((b, n)) for _ in range(6))
    bx = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
    by = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
    be = fa.buffer_split(er.data_ptr(), ei.data_ptr(), n)
    g = None
    try:
        fa.check(lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0), "eager")
        g = capture(lambda: [lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0)])
        assert lib.faf_graph_nodes(g) >= 2
        for rep in range(3):
            x = cplx(rng, b, n)
            put(xr, x.real)
            put(xi, x.imag)
            fa.check(lib.faf_execute_batch(t, C.byref(be), C.byref(bx), b, 0, 0), "eager")
            fa.check(lib.faf_graph_launch(g), "graph launch")
            y = G.host(yr) + 1j * G.host(yi)
            assert np.array_equal(y, G.host(er) + 1j * G.host(ei)), f"replay {rep}"
            assert rel_l2(y, O.fft64(x.astype(np.complex128))) < tol(n)
    finally:
        lib.faf_graph_destroy(g)
        G.destroy(t)


def test_graph_survives_workspace_growth():
    """A graph keeps the workspace address it captured.  An eager execute at
    a larger batch afterwards grows the plan's workspace: the captured buffer
    is retired (kept until the plan is destroyed), not freed, so replays
    still run on valid memory and give the eager result."""
    say("case: graph c2c 65536 after workspace growth")
    lib = G.L()
    n, b = 65536, 2
    rng = np.random.default_rng(n + 13)
    t = G.create("fft", n)
    xr, xi, yr, yi, er, ei = (G.empty((b, n)) for _ in range(6))
    bx = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), n)
    by = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), n)
    be = fa.buffer_split(er.data_ptr(), ei.data_ptr(), n)
    g = None
    try:
        fa.check(lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0), "eager")
        g = capture(lambda: [lib.faf_execute_batch(t, C.byref(by), C.byref(bx), b, 0, 0)])
        for big in (8, 24):                   # each grows the workspace past the captured one
            ur, ui, vr, vi = (G.empty((big, n)) for _ in range(4))
            u = cplx(rng, big, n)
            put(ur, u.real)
            put(ui, u.imag)
            bu = fa.buffer_split(ur.data_ptr(), ui.data_ptr(), n)
            bv = fa.buffer_split(vr.data_ptr(), vi.data_ptr(), n)
            fa.check(lib.faf_execute_batch(t, C.byref(bv), C.byref(bu), big, 0, 0), "grow")
            v = G.host(vr) + 1j * G.host(vi)
            assert rel_l2(v[big - 1], O.fft64(u[big - 1].astype(np.complex128))) < tol(n)
            x = cplx(rng, b, n)
            put(xr, x.real)
            put(xi, x.imag)
            fa.check(lib.faf_execute_batch(t, C.byref(be), C.byref(bx), b, 0, 0), "eager")
            fa.check(lib.faf_graph_launch(g), "graph launch")
            y = G.host(yr) + 1j * G.host(yi)
            assert np.array_equal(y, G.host(er) + 1j * G.host(ei)), f"after growth to {big}"
            assert rel_l2(y, O.fft64(x.astype(np.complex128))) < tol(n)
    finally:
        lib.faf_graph_destroy(g)
        G.destroy(t)


@pytest.mark.parametrize("persist", [0, 1])
def test_graph_ragged_replays(persist):
    """Replays of a captured ragged launch match eager runs interleaved with
    them, in both launch shapes: the persistent kernel's last workgroup resets
    the work counter for the next launch, so every replay starts from a full
    queue."""
    say(f"case: graph ragged persist={persist}")
    lib = G.L()
    rng = np.random.default_rng(77)
    lens = [int(v) for v in rng.choice([60, 360, 1000, 11, 17, 97, 1009, 4096], 300)]
    arr = (C.c_size_t * len(lens))(*lens)
    r = lib.faf_ragged_create(arr, len(lens), 0)
    assert r, err()
    fa.check(lib.faf_ragged_set_persistent(r, persist), "set persistent")
    assert lib.faf_ragged_persistent(r) == persist
    tot = lib.faf_ragged_total(r)
    assert tot == sum(lens)
    off = np.concatenate([[0], np.cumsum(lens)])
    xr, xi, yr, yi, er, ei = (G.empty(tot) for _ in range(6))
    g = None
    try:
        fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(), xr.data_ptr(),
                                        xi.data_ptr()), "eager")
        g = capture(lambda: [lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(),
                                                    xr.data_ptr(), xi.data_ptr())])
        for rep in range(3):
            x = cplx(rng, 1, tot)[0]
            put(xr, x.real)
            put(xi, x.imag)
            fa.check(lib.faf_ragged_execute(r, er.data_ptr(), ei.data_ptr(), xr.data_ptr(),
                                            xi.data_ptr()), "eager")
            fa.check(lib.faf_graph_launch(g), "graph launch")
            y = G.host(yr) + 1j * G.host(yi)
            assert np.array_equal(y, G.host(er) + 1j * G.host(ei)), f"replay {rep}"
            for i in rng.choice(len(lens), 12, replace=False):
                a, e = off[i], off[i + 1]
                ref = O.fft64(x[a:e].astype(np.complex128)).ravel()
                assert rel_l2(y[a:e], ref) < tol(lens[i]), f"signal {i} n={lens[i]}"
    finally:
        lib.faf_graph_destroy(g)
        lib.faf_ragged_destroy(r)


def test_graph_refusals_leave_state_clean():
    """What a capture cannot hold fails with a message, and the thread's stream
    and sync mode are back to normal afterwards."""
    say("case: graph refusals")
    lib = G.L()
    n = 1024
    nb = n // 2 + 1
    t = G.create("rfft", n)
    big = G.create("fft", 65536)           # four-step, never executed: no workspace yet
    try:
        # host buffers cannot be captured
        xh = np.random.default_rng(3).standard_normal(n).astype(np.float32)
        reh, imh = np.zeros(nb, np.float32), np.zeros(nb, np.float32)
        bi = fa.buffer_real(xh.ctypes.data, n)
        bo = fa.buffer_hermitian(reh.ctypes.data, imh.ctypes.data, nb)
        fa.check(lib.faf_graph_begin(), "begin")
        assert lib.faf_execute(t, C.byref(bo), C.byref(bi)) == -1
        assert "host buffers" in err()
        assert not lib.faf_graph_end()
        # plans cannot be created inside a capture
        fa.check(lib.faf_graph_begin(), "begin")
        assert not lib.faf_create_rfft(C.byref(fa.config_init(2048)))
        assert "capture" in err()
        assert lib.faf_amd_synchronize() == -1
        assert lib.faf_graph_begin() == -1                 # no nesting
        assert not lib.faf_graph_end()                     # nothing captured
        # a workspace cannot grow inside a capture
        xr, xi, yr, yi = (G.empty((2, 65536)) for _ in range(4))
        bi2 = fa.buffer_split(xr.data_ptr(), xi.data_ptr(), 65536)
        bo2 = fa.buffer_split(yr.data_ptr(), yi.data_ptr(), 65536)
        fa.check(lib.faf_graph_begin(), "begin")
        assert lib.faf_execute_batch(big, C.byref(bo2), C.byref(bi2), 2, 0, 0) == -1
        assert "workspace" in err()
        assert not lib.faf_graph_end()
        assert lib.faf_graph_end() is None                  # no capture open any more
        # the thread is back to synchronous eager executes on host buffers
        fa.check(lib.faf_execute(t, C.byref(bo), C.byref(bi)), "eager after refusals")
        ref = O.fft64(xh.astype(np.complex128)).ravel()[:nb]
        assert rel_l2(reh + 1j * imh, ref) < tol(n)
    finally:
        G.destroy(big)
        G.destroy(t)
