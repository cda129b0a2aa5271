"""CPU tests of the C ABI: the library loads, exports every function the
public headers declare, keeps the reference's struct layouts and enum
values, and its host-only entry points behave as the reference's do.
No kernel is launched here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import fastandfourier_amd as fa
from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("fastandfourier.h", "faf_amd.h", "chirp.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"static inline[^{]*\{.*?\n\}", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:faf|chirp)_\w+)\s*\(", src, flags=re.M):
            line = src[m.start():src.find("\n", m.start())]
            if line.lstrip().startswith(("typedef", "#", "static")):
                continue
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    so = fa.LIB_PATH
    assert os.path.exists(so), "build first: __graft_entry__.build()"
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    want = declared_functions()
    assert len(want) > 60
    missing = sorted(want - exported)
    assert not missing, missing
    # and the Python binding knows every one of them
    assert not sorted(want - set(fa.EXPORTED)), sorted(want - set(fa.EXPORTED))


def test_library_loads_and_host_entry_points():
    L = fa.load()
    assert L.faf_version() == b"1.2.0"
    assert L.faf_init() == 0
    # size policy: reference tests/test_sizes.cpp:56-115, tests/test_core.cpp:106-116
    T = fa.TYPE
    for n in (3840, 4000, 4050, 4096, 7, 11):
        assert L.faf_is_size_supported(T["fft"], n)
    for n in (14, 23, 4097, 21, 49, 0):
        assert not L.faf_is_size_supported(T["fft"], n)
    for n in (15, 9, 25, 27):
        assert not L.faf_is_size_supported(T["rfft"], n)
    for a, b in [(1, 1), (2, 2), (3, 3), (7, 7), (100, 100), (1000, 1000), (11, 11), (23, 24),
                 (3000, 3000)]:
        assert L.faf_get_recommended_size(T["fft"], a) == b
    nr = L.faf_get_recommended_size(T["rfft"], 3000)
    assert nr >= 3000 and nr % 2 == 0
    assert L.faf_layout_name(1) == b"split" and L.faf_norm_name(2) == b"ortho"
    assert L.faf_transform_name(21) == b"cwt"


def test_herm_mul_and_converters_known_answers():
    # reference tests/test_layout.cpp:120-153
    L = fa.load()
    f = lambda v: np.asarray(v, np.float32)
    xr, xi, hr, hi = f([2, 1, 0, 3, 4]), f([7, 1, 0, 2, 9]), f([3, 1, 0, 0, 5]), f([8, 0, 0, 1, 6])
    yr, yi = np.zeros(5, np.float32), np.zeros(5, np.float32)
    L.faf_herm_mul_f32(yr.ctypes.data, yi.ctypes.data, xr.ctypes.data, xi.ctypes.data,
                       hr.ctypes.data, hi.ctypes.data, 5)
    assert (yr[0], yi[0], yr[4], yi[4], yr[1], yi[1]) == (6, 0, 20, 0, 1, 1)
    xr, xi, hr, hi = f([1, 2, 3]), f([0, 4, 0]), f([5, 1, 7]), f([0, 1, 0])
    yr, yi = np.zeros(3, np.float32), np.zeros(3, np.float32)
    L.faf_herm_mul_conj_f32(yr.ctypes.data, yi.ctypes.data, xr.ctypes.data, xi.ctypes.data,
                            hr.ctypes.data, hi.ctypes.data, 3)
    assert list(yr) == [5, 6, 21] and list(yi) == [0, 2, 0]
    iv = np.arange(8, dtype=np.float32)
    a, b = np.zeros(4, np.float32), np.zeros(4, np.float32)
    assert L.faf_deinterleave_f32(a.ctypes.data, b.ctypes.data, iv.ctypes.data, 4) == 0
    assert list(a) == [0, 2, 4, 6] and list(b) == [1, 3, 5, 7]
    assert L.faf_interleave_f32(None, a.ctypes.data, b.ctypes.data, 4) != 0
    w = np.zeros(1024, np.float32)
    L.faf_gen_hann_window_f32(w.ctypes.data, 1024)
    assert np.allclose(w, np.hanning(1024), atol=1e-7)


def test_cwt_config_defaults():
    # reference src/faf_cwt.c:27-51
    c = fa.load().faf_cwt_config_init(4096)
    assert (c.n, c.fs, c.precision, c.wavelet, c.norm, c.voices) == (4096, 1.0, 3, 1, 0, 10)
    assert (c.morse_gamma, c.morse_beta, c.lp_alpha, c.flags) == (3.0, 20.0, 0.05, 0x21)


@pytest.mark.skipif(__import__("importlib").util.find_spec("torch") is None, reason="no torch")
def test_creation_fails_loudly_without_a_device():
    """No silent CPU path: without a GPU every create returns NULL with a message."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = fa.load()
    t = L.faf_create_fft(C.byref(fa.config_init(64)))
    assert not t
    assert b"GPU runtime unavailable" in L.faf_get_error()
    # graph capture needs the device too; the null-argument paths need nothing
    assert L.faf_graph_begin() == -1
    assert b"GPU runtime unavailable" in L.faf_get_error()
    assert not L.faf_graph_end()
    assert L.faf_graph_launch(None) == -1 and L.faf_graph_nodes(None) == 0
    L.faf_graph_destroy(None)


def test_server_config_defaults_and_no_silent_cpu_path():
    """Serving defaults follow the stated request mix; without a GPU the
    server refuses to start instead of computing on the CPU."""
    import torch
    L = fa.load()
    c = L.faf_server_config_init()
    assert c.hbm_budget == 64 << 30 and c.stft_win == 512 and c.stft_hop == 128
    assert c.cwt_voices == 8 and abs(c.cwt_octaves - 31.5 / 8) < 1e-15 and c.cwt_fs == 16000.0
    assert c.plain == 0 and c.streams >= 1 and c.max_batch >= 1
    assert L.faf_server_submit(None, 0, 64, None, 0) == -1
    assert L.faf_server_result_bytes(None, 0, 64) == 0
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert not L.faf_server_create(C.byref(c))
    assert b"GPU runtime unavailable" in L.faf_get_error()


def test_struct_layout_matches_reference_header(tmp_path):
    """offsetof/sizeof of the public structs and enum values are identical to
    the reference header's (compiled side by side when the reference tree
    is present)."""
    ref_inc = "/root/reference/include"
    if not os.path.isdir(ref_inc):
        pytest.skip("reference headers not present on this host")
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "fastandfourier.h"
#define O(T, f) printf(#T "." #f "=%zu\n", offsetof(T, f))
int main(void) {
  printf("cfg=%zu buf=%zu cwt=%zu rep=%zu\n", sizeof(faf_config), sizeof(faf_buffer),
         sizeof(faf_cwt_config), sizeof(faf_cwt_lp_report));
  O(faf_config, layout); O(faf_config, norm); O(faf_config, dir); O(faf_config, flags);
  O(faf_config, hop_length); O(faf_config, win_length);
  O(faf_cwt_config, voices); O(faf_cwt_config, fmin); O(faf_cwt_config, flags);
  O(faf_cwt_lp_report, frame_cond); O(faf_cwt_lp_report, admissibility_C);
  O(faf_transform, n); O(faf_transform, type); O(faf_transform, cfg); O(faf_transform, flags);
  O(faf_transform, inner); O(faf_transform, inner_inv); O(faf_transform, user_aux_n);
  O(faf_transform, precision); O(faf_transform, custom_len_syn);
  printf("%d %d %d %d %d %d\n", FAF_PREC_FP32, FAF_TRANSFORM_CWT, FAF_LAYOUT_INTERLEAVED,
         FAF_NORM_FORWARD, FAF_CWT_WAVELET_MEYER, (int)FAF_FLAG_BLUESTEIN);
  return 0;
}
"""
    src = tmp_path / "abi.c"
    src.write_text(prog)
    outs = []
    for inc in (os.path.join(ROOT, "include"), ref_inc):
        exe = tmp_path / ("abi_" + str(len(outs)))
        subprocess.run(["gcc", "-std=c11", "-I", inc, str(src), "-o", str(exe)], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1], (outs[0], outs[1])
