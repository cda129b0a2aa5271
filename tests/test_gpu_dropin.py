"""A C program written against the reference API, compiled against
include/ and linked to this library unchanged, runs on the GPU."""
import os
import shutil
import subprocess
import sys

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "c", "dropin_rfft.c")
LIBDIR = os.path.join(ROOT, "fastandfourier_amd")


def build(out):
    cc = shutil.which("cc") or shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    subprocess.run([cc, "-std=c11", "-O2", SRC, "-I", os.path.join(ROOT, "include"),
                    "-L", LIBDIR, "-lfastandfourier_amd", f"-Wl,-rpath,{LIBDIR}", "-lm",
                    "-o", out], check=True)


REF_EXAMPLES = "/root/reference/examples"
# the reference's own example programs that stay on the transform path
EXAMPLES = ["simple_rfft", "simple_fft", "spectral_notch", "fft_convolve", "cwt_filterbank",
            "cwt_reconstruct", "split_corr", "analytic_envelope", "matched_filter"]


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference tree not present")
@pytest.mark.parametrize("name", EXAMPLES)
def test_reference_example_links_unchanged(tmp_path, name):
    """CPU: the reference's own example program, compiled from where it lies
    in the reference tree against include/ (nothing is copied), links
    against this library with no change."""
    cc = shutil.which("cc") or shutil.which("gcc")
    if not cc:
        pytest.skip("no C compiler")
    src = os.path.join(REF_EXAMPLES, name + ".c")
    r = subprocess.run([cc, "-std=c11", "-O1", src, "-I", os.path.join(ROOT, "include"),
                        "-L", LIBDIR, "-lfastandfourier_amd", f"-Wl,-rpath,{LIBDIR}", "-lm",
                        "-o", str(tmp_path / name)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_program_links(tmp_path):
    """CPU: the reference-API program compiles and links against this library."""
    build(str(tmp_path / "dropin"))


@pytest.mark.gpu
def test_dropin_program_runs_on_gpu(tmp_path):
    print("case: C drop-in program", flush=True)
    exe = str(tmp_path / "dropin")
    build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0 and "dropin OK" in r.stdout
