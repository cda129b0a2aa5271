"""GPU tests of the autotuners.

Plan tuner (faf_amd_autotune, FAF_FLAG_MEASURE at create): the chosen
launch shape is recorded per (kind, length, direction, layout) and a plan of
the same key created afterwards starts from it, gives the same bits as the
tuned plan and matches FP64; a plan of another layout keeps its own shape.

Ragged tuner (faf_ragged_autotune, FAF_FLAG_MEASURE on faf_ragged_create):
per-length choices are applied in place, recorded for later ragged plans,
and every signal still matches FP64 at relative L2 <= 1e-6 * log2 N.

The records are process-wide, so each test uses lengths no later test in
this process plans.
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


def shape(t):
    sp, no, spb = C.c_int(), C.c_int(), C.c_int()
    order = (C.c_int * 16)()
    fa.check(G.L().faf_amd_plan_shape(t, C.byref(sp), order, C.byref(no), C.byref(spb)), "shape")
    return sp.value, tuple(order[: no.value]), spb.value


def cplx(rng, k, n):
    return (rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))).astype(np.complex64)


def test_autotune_shape_is_recorded_per_key():
    say("case: autotune c2c 2400 recorded per key")
    lib = G.L()
    n = 2400                                  # 5-smooth, runtime radix kernels only
    rng = np.random.default_rng(5)
    x = cplx(rng, 64, n)
    inter0 = G.create("fft", n, layout="interleaved")
    t1 = G.create("fft", n)
    try:
        d_inter = shape(inter0)
        spb = lib.faf_amd_autotune(t1, 2048)
        assert spb >= 1, lib.faf_get_error()
        s1 = shape(t1)
        y1 = G.c2c(t1, x)
        assert rel_l2(y1, O.fft64(x.astype(np.complex128))) < tol(n)
        t2 = G.create("fft", n)                # same key: starts from the record
        try:
            assert shape(t2) == s1
            assert np.array_equal(G.c2c(t2, x), y1)
        finally:
            G.destroy(t2)
        t3 = G.create("fft", n, layout="interleaved")   # other key: untouched
        try:
            assert shape(t3) == d_inter
            assert rel_l2(G.c2c(t3, x, inter=True), O.fft64(x.astype(np.complex128))) < tol(n)
        finally:
            G.destroy(t3)
        say(f"  tuned shape {s1}")
    finally:
        G.destroy(t1)
        G.destroy(inter0)


def test_measure_flag_tunes_at_create():
    say("case: FAF_FLAG_MEASURE rfft 6000")
    lib = G.L()
    n = 6000
    rng = np.random.default_rng(6)
    x = rng.standard_normal((16, n)).astype(np.float32)
    t = G.create("rfft", n, flags=fa.FLAG_MEASURE)
    try:
        s = shape(t)
        X = G.r2c(t, x)
        truth = np.stack([O.fft64(v.astype(np.float64))[0][: n // 2 + 1] for v in x])
        assert rel_l2(X, truth) < tol(n)
        t2 = G.create("rfft", n)               # no flag: the recorded shape
        try:
            assert shape(t2) == s
            assert np.array_equal(G.r2c(t2, x), X)
        finally:
            G.destroy(t2)
        # the inverse is another key, measured on its own when asked
        ti = G.create("rfft", n, dir="inverse", flags=fa.FLAG_MEASURE)
        try:
            y = G.c2r(ti, X, n)
            assert rel_l2(y, x) < 2 * tol(n)
        finally:
            G.destroy(ti)
    finally:
        G.destroy(t)


def _ragged_choices(r, lens):
    lib = G.L()
    out = {}
    for L_ in sorted(set(lens)):
        k, spb, no = C.c_int(), C.c_int(), C.c_int()
        order = (C.c_int * 16)()
        fa.check(lib.faf_ragged_choice(r, L_, C.byref(k), C.byref(spb), order, C.byref(no)),
                 "choice")
        out[L_] = (k.value, spb.value, tuple(order[: no.value]))
    return out


def _ragged_run(r, x):
    lib = G.L()
    tot = x.size
    xr, xi = G.dev(x.real.copy()), G.dev(x.imag.copy())
    yr, yi = G.empty(tot), G.empty(tot)
    fa.check(lib.faf_ragged_execute(r, yr.data_ptr(), yi.data_ptr(), xr.data_ptr(),
                                    xi.data_ptr()), "ragged")
    return G.host(yr) + 1j * G.host(yi)


def test_ragged_autotune_per_length():
    say("case: ragged autotune")
    lib = G.L()
    rng = np.random.default_rng(9)
    lset = [120, 720, 1500, 13, 61, 2048]     # Stockham (CT and runtime), Rader
    lens = [int(v) for v in rng.choice(lset, 900)]
    arr = (C.c_size_t * len(lens))(*lens)
    r = lib.faf_ragged_create(arr, len(lens), 0)
    assert r, lib.faf_get_error()
    off = np.concatenate([[0], np.cumsum(lens)])
    x = cplx(rng, 1, int(off[-1]))[0]
    try:
        before = _ragged_run(r, x)
        tuned = lib.faf_ragged_autotune(r)
        assert tuned == len(lset), lib.faf_get_error()
        ch = _ragged_choices(r, lens)
        say(f"  choices {ch}")
        after = _ragged_run(r, x)
        for i in list(range(0, len(lens), 37)) + [len(lens) - 1]:
            a, e = off[i], off[i + 1]
            truth = O.fft64(x[a:e].astype(np.complex128))[0]
            assert rel_l2(after[a:e], truth) < tol(lens[i]), (i, lens[i])
            assert rel_l2(before[a:e], truth) < tol(lens[i]), (i, lens[i])
        # a second plan over these lengths starts from the recorded choices,
        # and a second autotune has nothing left to time
        r2 = lib.faf_ragged_create(arr, len(lens), 0)
        assert r2, lib.faf_get_error()
        try:
            assert _ragged_choices(r2, lens) == ch
            assert np.array_equal(_ragged_run(r2, x), after)
            assert lib.faf_ragged_autotune(r2) == 0
        finally:
            lib.faf_ragged_destroy(r2)
    finally:
        lib.faf_ragged_destroy(r)


def test_ragged_measure_flag_at_create():
    say("case: ragged FAF_FLAG_MEASURE")
    lib = G.L()
    rng = np.random.default_rng(10)
    lens = [int(v) for v in rng.choice([240, 37, 1024], 300)]
    arr = (C.c_size_t * len(lens))(*lens)
    r = lib.faf_ragged_create(arr, len(lens), fa.FLAG_MEASURE)
    assert r, lib.faf_get_error()
    try:
        ch = _ragged_choices(r, lens)
        assert set(ch) == {240, 37, 1024}
        assert lib.faf_ragged_autotune(r) == 0           # measured at create
        off = np.concatenate([[0], np.cumsum(lens)])
        x = cplx(rng, 1, int(off[-1]))[0]
        y = _ragged_run(r, x)
        for i in range(0, len(lens), 29):
            a, e = off[i], off[i + 1]
            assert rel_l2(y[a:e], O.fft64(x[a:e].astype(np.complex128))[0]) < tol(lens[i])
    finally:
        lib.faf_ragged_destroy(r)
