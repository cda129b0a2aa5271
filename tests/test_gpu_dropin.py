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
