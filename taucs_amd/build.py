"""Build libtaucs_amd.so in-tree: host C with gcc, device code with hipcc for
gfx950, linked by hipcc into one shared library exporting the TAUCS C-ABI.

Usage: python taucs_amd/build.py  (the package itself needs the built library)
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libtaucs_amd.so")

HOST_SOURCES = ["util.c", "ccs.c", "mmd.c", "symbolic.c", "order.c", "factor.c", "linsolve.c", "io.c",
                "batch.c"]
DEVICE_SOURCES = ["factor_kernels.hip", "solve_kernels.hip", "driver.hip", "refine.hip",
                  "ccs_solve.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled in all device code.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]
# Per-file additions.  The factor kernels keep MFMA accumulators in
# ordinary VGPRs: a 64x64 FP64 tile wave then needs 104 registers instead of
# 128, so two of them leave room on a SIMD for the 288-register panel POTRF
# wave, which must run beside the trailing update.
FILE_FLAGS = {"factor_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}
C_FLAGS = ["-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wno-unused-function",
           "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        s = os.path.join(CSRC, "host", src)
        o = os.path.join(OBJ, src + ".o")
        jobs.append(["gcc"] + C_FLAGS + ["-c", s, "-o", o])
    for src in DEVICE_SOURCES:
        s = os.path.join(CSRC, "device", src)
        o = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lm", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
