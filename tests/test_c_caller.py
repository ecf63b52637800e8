"""A C program written against include/taucs.h, compiled with gcc and linked
with -ltaucs_amd like existing TAUCS callers (tests/c_caller/drop_in.c): the
protocol of the reference's progs/test_linsolve.c, the direct driver's flow
(progs/direct.c:540-550) and the symbolic/numeric split, through the C-ABI
with no Python in between.  The build is checked on the CPU; the run needs
the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c_caller", "drop_in.c")
LIBDIR = os.path.join(ROOT, "taucs_amd")


def build(out):
    cmd = ["gcc", "-O2", "-std=gnu99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-o", str(out), "-L", LIBDIR, "-ltaucs_amd", f"-Wl,-rpath,{LIBDIR}", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_c_caller_compiles_and_links_against_the_header(tmp_path):
    exe = build(tmp_path / "drop_in")
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_c_caller_runs_the_reference_protocol(tmp_path):
    print("case: C caller (gcc, -ltaucs_amd): test_linsolve protocol, direct flow", flush=True)
    exe = build(tmp_path / "drop_in")
    r = subprocess.run([str(exe), "20"], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "test succeeded" in r.stdout
