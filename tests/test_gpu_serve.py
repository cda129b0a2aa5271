"""GPU tests of the serving loop and of long-segment CWT.

Every served result is checked against the oracle's FP64 transform at the
stated bar (relative L2 <= 1e-6 * log2 N; CWT rows as for faf_execute_cwt_complex),
and the same request is shown to give the same bits alone and inside a mix
of other requests that overlap it on the device.
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

FS, VOICES, OCT = 16000.0, 8, 31.5 / 8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def server(**kw):
    lib = G.L()
    cfg = lib.faf_server_config_init()
    for k, v in kw.items():
        setattr(cfg, k, v)
    s = lib.faf_server_create(C.byref(cfg))
    assert s, lib.faf_get_error()
    return s


def make_input(rng, kind, n):
    if kind == "c2c":
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
        return x, G.dev(x.view(np.float32))
    x = rng.uniform(-1, 1, n).astype(np.float32)
    return x, G.dev(x)


def fetch(s, t):
    """wait for ticket t and copy its result to the host"""
    lib = G.L()
    fa.check(lib.faf_server_wait(s, t), "serve wait")
    nb = C.c_size_t()
    p = lib.faf_server_result(s, t, C.byref(nb))
    assert p
    y = np.empty(nb.value // 4, np.float32)
    # the HIP runtime the library itself links (dlsym through its handle)
    memcpy = lib["hipMemcpy"]
    memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    memcpy.restype = C.c_int
    assert memcpy(y.ctypes.data, p, nb.value, 2) == 0     # device to host
    return y[0::2] + 1j * y[1::2]


def cwt_table(n):
    lib = G.L()
    cc = lib.faf_cwt_config_init(n)
    cc.fs, cc.voices = FS, VOICES
    cc.fmin = 16 * FS / n
    cc.fmax = cc.fmin * 2 ** OCT
    t = lib.faf_create_cwt(C.byref(cc))
    assert t, lib.faf_get_error()
    try:
        rows, nb = lib.faf_cwt_n_rows(t), lib.faf_cwt_n_bins(t)
        stride = (nb + 15) & ~15
        tab = np.zeros((rows, stride), np.float32)
        tab[0, :nb] = np.ctypeslib.as_array(lib.faf_cwt_phi_f32(t), (nb,))
        for j in range(rows - 1):
            tab[j + 1, :nb] = np.ctypeslib.as_array(lib.faf_cwt_psi_f32(t, j), (nb,))
        return tab, stride
    finally:
        lib.faf_destroy_transform(t)


def truth(kind, n, x):
    if kind == "rfft":
        return O.fft64(x.astype(np.complex128)).ravel()[: n // 2 + 1]
    if kind == "c2c":
        return O.fft64(x.astype(np.complex128)).ravel()
    if kind == "stft":
        return O.stft64(x, 512, 128).ravel()
    tab, stride = cwt_table(n)
    return O.cwt_complex64(x, tab, stride).ravel()


def bar(kind, n):
    return tol(512 if kind == "stft" else n)


# short lengths run inside the serving kernel; the last four are long
# (four-step C2C and R2C, CWT beyond one workgroup) and take the plan path
MIX = [("rfft", 64), ("rfft", 4096), ("rfft", 32768), ("c2c", 60), ("c2c", 1000),
       ("c2c", 16384), ("c2c", 11), ("c2c", 1009), ("stft", 300), ("stft", 4000),
       ("cwt", 512), ("cwt", 8192), ("c2c", 20000), ("rfft", 65536), ("cwt", 16384),
       ("c2c", 65536)]


def test_serve_each_kind_matches_fp64():
    say("case: serve mix vs fp64")
    lib = G.L()
    rng = np.random.default_rng(11)
    s = server(hbm_budget=1 << 30, streams=3)
    try:
        data = [make_input(rng, k, n) for k, n in MIX]
        torch.cuda.synchronize()
        tks = [lib.faf_server_submit(s, fa.REQ[k], n, d.data_ptr(), 0)
               for (k, n), (_, d) in zip(MIX, data)]
        assert min(tks) >= 0, lib.faf_get_error()
        for (k, n), (x, _), t in zip(MIX, data, tks):
            say(f"  {k} n={n}")
            y = fetch(s, t)
            ref = truth(k, n, x)
            assert y.shape == ref.shape
            assert rel_l2(y, ref) <= bar(k, n), (k, n, rel_l2(y, ref))
            fa.check(lib.faf_server_release(s, t), "release")
        st = fa.FafServerStats()
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.completed == len(MIX) and st.failed == 0
    finally:
        lib.faf_server_destroy(s)


def test_serve_result_same_alone_and_in_mix():
    """Each probe request alone, then inside a burst of other requests."""
    say("case: serve alone vs mix")
    lib = G.L()
    rng = np.random.default_rng(12)
    probes = [("rfft", 4096), ("c2c", 360), ("stft", 2000), ("cwt", 1024), ("c2c", 20000),
              ("rfft", 65536)]
    pdata = [make_input(rng, k, n) for k, n in probes]
    s = server(hbm_budget=2 << 30, streams=3, window_us=200.0)
    try:
        torch.cuda.synchronize()
        alone = []
        for (k, n), (_, d) in zip(probes, pdata):
            t = lib.faf_server_submit(s, fa.REQ[k], n, d.data_ptr(), 0)
            assert t >= 0
            alone.append(fetch(s, t))
            lib.faf_server_release(s, t)
        # the mix: 300 other requests around the probes, in one burst
        ks = ["rfft", "c2c", "stft", "cwt"]
        others = []
        for i in range(300):
            k = ks[i % 4]
            if k == "cwt":
                n = int(rng.choice([512, 640, 1024, 2048, 4800, 8192, 10000]))
            else:
                n = int(rng.choice([64, 120, 512, 1000, 2048, 4800, 8192,
                                    18000 if k == "c2c" else 640]))
            others.append((k, n, make_input(rng, k, n)[1]))
        torch.cuda.synchronize()
        tks, probe_t = [], []
        for i, (k, n, d) in enumerate(others):
            tks.append(lib.faf_server_submit(s, fa.REQ[k], n, d.data_ptr(), 0))
            if i % 50 == 25 and len(probe_t) < len(probes):
                j = len(probe_t)
                kk, nn = probes[j]
                probe_t.append(lib.faf_server_submit(s, fa.REQ[kk], nn, pdata[j][1].data_ptr(), 0))
        assert min(tks + probe_t) >= 0, lib.faf_get_error()
        st = fa.FafServerStats()
        for j, t in enumerate(probe_t):
            y = fetch(s, t)
            k, n = probes[j]
            say(f"  probe {k} n={n}")
            assert np.array_equal(y, alone[j]), (k, n, rel_l2(y, alone[j]))
            assert rel_l2(y, truth(k, n, pdata[j][0])) <= bar(k, n)
        lib.faf_server_drain(s)
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.failed == 0
        assert st.mean_batch > 1.0      # the mix really was batched
        for t in tks + probe_t:
            lib.faf_server_release(s, t)
    finally:
        lib.faf_server_destroy(s)


def test_serve_plain_mode_same_bits():
    say("case: serve plain vs batched")
    lib = G.L()
    rng = np.random.default_rng(13)
    reqs = [("rfft", 2048), ("c2c", 4000), ("stft", 1500), ("cwt", 2048), ("c2c", 32768)]
    data = [make_input(rng, k, n) for k, n in reqs]
    outs = []
    for plain in (0, 1):
        s = server(hbm_budget=1 << 30, plain=plain)
        try:
            torch.cuda.synchronize()
            tks = [lib.faf_server_submit(s, fa.REQ[k], n, d.data_ptr(), 0)
                   for (k, n), (_, d) in zip(reqs, data)]
            outs.append([fetch(s, t) for t in tks])
            st = fa.FafServerStats()
            lib.faf_server_stats_get(s, C.byref(st))
            if plain:
                assert st.mean_batch == 1.0
        finally:
            lib.faf_server_destroy(s)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_serve_admission_waits_for_budget():
    """Budget for about three results: the rest wait until results are released."""
    say("case: serve admission budget")
    lib = G.L()
    rng = np.random.default_rng(14)
    n = 4096            # short C2C: the result is all it holds (no workspace)
    one = lib.faf_server_result_bytes(None, 0, n)
    assert one == 0                                    # no server -> not servable
    s = server(hbm_budget=3 * (n * 8) + 4096, streams=2)
    try:
        per = lib.faf_server_result_bytes(s, fa.REQ["c2c"], n)
        assert per == n * 8
        xs = [make_input(rng, "c2c", n) for _ in range(10)]
        torch.cuda.synchronize()
        tks = [lib.faf_server_submit(s, fa.REQ["c2c"], n, d.data_ptr(), 0) for _, d in xs]
        done = []
        import time
        for _ in range(200):
            done = [t for t in tks if lib.faf_server_poll(s, t) == 1]
            if len(done) >= 3:
                break
            time.sleep(0.005)
        time.sleep(0.05)
        done = [t for t in tks if lib.faf_server_poll(s, t) == 1]
        assert len(done) == 3                          # the budget holds three results
        for i, t in enumerate(tks):
            y = fetch(s, t)
            assert rel_l2(y, O.fft64(xs[i][0].astype(np.complex128)).ravel()) <= tol(n)
            lib.faf_server_release(s, t)
            # one release admits exactly one more request; waiting for it before the
            # next release makes every later head-of-line request meet a full budget
            if i + 3 < len(tks):
                assert lib.faf_server_wait(s, tks[i + 3]) == 0
        st = fa.FafServerStats()
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.admission_waits >= 7 and st.peak_hbm_bytes <= 3 * n * 8 + 4096
    finally:
        lib.faf_server_destroy(s)


def test_serve_replay_poisson_stream():
    """A short Poisson stream at 200k req/s through the replay entry point."""
    say("case: serve replay")
    lib = G.L()
    rng = np.random.default_rng(15)
    count = 4000
    kinds = rng.choice(4, size=count, p=[0.5, 0.25, 0.15, 0.10]).astype(np.int32)
    ns = np.array([[256, 1000, 4096, 640][k] for k in kinds], dtype=np.uint64)
    pool = G.dev(rng.uniform(-1, 1, 1 << 20).astype(np.float32))
    ins = np.array([pool.data_ptr() + 4096 * (i % 128) for i in range(count)], dtype=np.uint64)
    arr = np.cumsum(rng.exponential(5e3, count)).astype(np.uint64)
    lat = np.zeros(count, np.uint64)
    s = server(hbm_budget=1 << 30)
    try:
        for k, n in {(int(k), int(n)) for k, n in zip(kinds, ns)}:
            fa.check(lib.faf_server_prepare(s, k, n), "prepare")
        torch.cuda.synchronize()
        rc = lib.faf_server_replay(s, kinds.ctypes.data, ns.ctypes.data, ins.ctypes.data,
                                   arr.ctypes.data, count, fa.REQ_DISCARD, lat.ctypes.data)
        fa.check(rc, "replay")
        st = fa.FafServerStats()
        lib.faf_server_stats_get(s, C.byref(st))
        say(f"  p50 {st.p50_us:.1f} us p99 {st.p99_us:.1f} us batch {st.mean_batch:.1f}")
        assert st.completed == count and st.failed == 0
        assert np.all(lat > 0)
        assert abs(np.percentile(lat, 50) / 1e3 - st.p50_us) < 0.25 * st.p50_us + 5
    finally:
        lib.faf_server_destroy(s)


def test_serve_failed_submit_issues_no_ticket():
    """A submission that fails returns -1 and leaves no ticket behind, so
    drain, stats and later submissions go on as if it never happened."""
    say("case: serve failed submissions")
    lib = G.L()
    rng = np.random.default_rng(21)
    s = server(hbm_budget=1 << 28)
    try:
        x, dx = make_input(rng, "rfft", 1024)
        assert lib.faf_server_submit(s, 0, 1023, dx.data_ptr(), 0) == -1      # odd rfft
        assert lib.faf_server_submit(s, 0, 1024, None, 0) == -1               # no input
        assert lib.faf_server_submit(s, 9, 1024, dx.data_ptr(), 0) == -1      # no such kind
        fa.check(lib.faf_server_drain(s), "drain after failures")
        t0 = lib.faf_server_submit(s, 0, 1024, dx.data_ptr(), 0)
        assert t0 == 0, "failed submissions must not consume tickets"
        y = fetch(s, t0)
        assert rel_l2(y, O.fft64(x.astype(np.complex128)).ravel()[:513]) < tol(1024)
        fa.check(lib.faf_server_release(s, t0), "release")
        fa.check(lib.faf_server_drain(s), "drain")
        st = fa.FafServerStats()
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.submitted == 1 and st.completed == 1 and st.failed == 0
        lib.faf_server_stats_reset(s)
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.submitted == 0
        assert lib.faf_server_poll(s, 5) == -1                                # never issued
    finally:
        lib.faf_server_destroy(s)


def test_serve_plans_build_beside_ready_requests():
    """Requests whose plans are new are submitted without waiting for the
    build; they complete once their plan is ready, with the same results as
    prepared ones, while a request with a ready plan goes ahead.  A plan whose
    build fails (C2C 65521, a prime beyond the serving kernel's Rader and
    Bluestein shapes) fails its ticket, and later submissions of that length
    return -1 at once."""
    say("case: serve plans built on builder threads")
    lib = G.L()
    rng = np.random.default_rng(33)
    s = server(hbm_budget=1 << 30)
    try:
        fa.check(lib.faf_server_prepare(s, 0, 4096), "prepare")
        cases = [("cwt", 8192), ("stft", 4000), ("c2c", 1000), ("rfft", 4096), ("c2c", 1009),
                 ("rfft", 32768), ("cwt", 512)]
        kid = {"rfft": 0, "c2c": 1, "stft": 2, "cwt": 3}
        subs = []
        for kind, n in cases:
            x, dx = make_input(rng, kind, n)
            t = lib.faf_server_submit(s, kid[kind], n, dx.data_ptr(), 0)
            assert t >= 0, lib.faf_get_error()
            subs.append((kind, n, x, dx, t))
        xb, dxb = make_input(rng, "c2c", 65521)
        tbad = lib.faf_server_submit(s, 1, 65521, dxb.data_ptr(), 0)
        assert tbad >= 0, "the build of a new plan is not waited for at submission"
        first = []
        for kind, n, x, _, t in subs:
            y = fetch(s, t)
            err = rel_l2(y, truth(kind, n, x))
            say(f"  {kind} {n}: rel l2 {err:.2e}")
            assert err < bar(kind, n), (kind, n, err)
            first.append(y)
            fa.check(lib.faf_server_release(s, t), "release")
        assert lib.faf_server_wait(s, tbad) == -1
        assert lib.faf_server_submit(s, 1, 65521, dxb.data_ptr(), 0) == -1
        assert lib.faf_get_error()
        assert lib.faf_server_prepare(s, 1, 65521) == -1
        fa.check(lib.faf_server_drain(s), "drain")
        st = fa.FafServerStats()
        lib.faf_server_stats_get(s, C.byref(st))
        assert st.submitted == len(cases) + 1
        assert st.completed == len(cases) and st.failed == 1
        # the same requests with every plan now ready give the same bits
        for (kind, n, _, dx, _), y0 in zip(subs, first):
            t = lib.faf_server_submit(s, kid[kind], n, dx.data_ptr(), 0)
            assert t >= 0, lib.faf_get_error()
            assert np.array_equal(fetch(s, t), y0), (kind, n)
            fa.check(lib.faf_server_release(s, t), "release")
    finally:
        lib.faf_server_destroy(s)


@pytest.mark.parametrize("n", [16384, 32768, 65536])
def test_cwt_long_segments_match_fp64(n):
    """CWT segments longer than one workgroup: rfft, spread, batched inverse."""
    say(f"case: cwt long n={n}")
    lib = G.L()
    cc = lib.faf_cwt_config_init(n)
    cc.fs, cc.voices = FS, VOICES
    cc.fmin = 16 * FS / n
    cc.fmax = cc.fmin * 2 ** OCT
    t = lib.faf_create_cwt(C.byref(cc))
    assert t, lib.faf_get_error()
    try:
        tab, stride = cwt_table(n)
        rows = tab.shape[0]
        rng = np.random.default_rng(n)
        xs = rng.standard_normal((2, n)).astype(np.float32)
        dxs = G.dev(xs)
        z = G.empty((2, rows, n, 2))
        w = G.empty((2, rows, n))
        torch.cuda.synchronize()
        fa.check(lib.faf_execute_cwt_complex(t, z.data_ptr(), dxs.data_ptr(), 2, 0), "cwt")
        bi = fa.buffer_real(dxs.data_ptr(), n)
        bo = fa.buffer_real(w.data_ptr(), rows * n)
        fa.check(lib.faf_execute_batch(t, C.byref(bo), C.byref(bi), 2, 0, 0), "cwt real")
        zc = G.host(z)
        zc = zc[..., 0] + 1j * zc[..., 1]
        wr = G.host(w)
        for sgm in range(2):
            ref = O.cwt_complex64(xs[sgm], tab, stride)
            assert rel_l2(zc[sgm], ref) <= tol(n)
            assert rel_l2(wr[sgm], ref.real) <= tol(n)
    finally:
        lib.faf_destroy_transform(t)


def test_serve_prepare_many():
    """faf_server_prepare_many builds a list of plans side by side and returns
    when all are ready; a plan that cannot be built makes it return -1 with
    that plan's message while the others are still built and serve at once."""
    say("case: serve prepare many")
    lib = G.L()
    rng = np.random.default_rng(41)
    s = server(hbm_budget=1 << 30)
    try:
        cases = [("rfft", 4096), ("c2c", 1000), ("stft", 4000), ("cwt", 8192), ("c2c", 1009)]
        kid = {"rfft": 0, "c2c": 1, "stft": 2, "cwt": 3}
        ks = np.array([kid[k] for k, _ in cases] + [1], dtype=np.int32)
        ns = np.array([n for _, n in cases] + [65521], dtype=np.uint64)
        assert lib.faf_server_prepare_many(s, ks.ctypes.data, ns.ctypes.data, len(ks)) == -1
        assert lib.faf_get_error()
        fa.check(lib.faf_server_prepare_many(s, ks.ctypes.data, ns.ctypes.data, len(ks) - 1),
                 "prepare many")
        for kind, n in cases:
            x, dx = make_input(rng, kind, n)
            t = lib.faf_server_submit(s, kid[kind], n, dx.data_ptr(), 0)
            assert t >= 0, lib.faf_get_error()
            err = rel_l2(fetch(s, t), truth(kind, n, x))
            say(f"  {kind} {n}: rel l2 {err:.2e}")
            assert err < bar(kind, n), (kind, n, err)
            fa.check(lib.faf_server_release(s, t), "release")
        assert lib.faf_server_prepare_many(s, None, None, 0) == 0
    finally:
        lib.faf_server_destroy(s)


def _smooth5(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def test_serve_destroy_while_plans_build():
    """Destroying a server while new plans are still queued for its builder
    threads fails the requests waiting on them, drops the queued builds and
    returns promptly; a new server then serves normally."""
    import time
    say("case: serve destroy during builds")
    lib = G.L()
    rng = np.random.default_rng(43)
    s = server(hbm_budget=1 << 30)
    x, dx = make_input(rng, "rfft", 65536)
    smooth = [n for n in range(20000, 65537, 2) if _smooth5(n)]
    tickets = []
    for n in smooth[-40:]:                           # 40 new CWT banks (slow builds)
        t = lib.faf_server_submit(s, 3, n, dx.data_ptr(), 0)
        assert t >= 0, lib.faf_get_error()
        tickets.append(t)
    t0 = time.perf_counter()
    lib.faf_server_destroy(s)
    took = time.perf_counter() - t0
    say(f"  destroy took {took:.2f} s with {len(tickets)} requests submitted")
    assert took < 20.0
    s = server(hbm_budget=1 << 28)
    try:
        t = lib.faf_server_submit(s, 0, 4096, dx.data_ptr(), 0)
        assert t >= 0, lib.faf_get_error()
        err = rel_l2(fetch(s, t), truth("rfft", 4096, x[:4096]))
        assert err < bar("rfft", 4096), err
    finally:
        lib.faf_server_destroy(s)
