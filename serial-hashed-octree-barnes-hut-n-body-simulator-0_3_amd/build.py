"""Builds libbhx.so (HIP kernels for gfx950 + the C ABI) in this directory.

Usage: python build.py            (or build() from Python)
Every HIP translation unit is compiled with hipcc for gfx950 only; the
parity-critical per-body kernels additionally with -ffp-contract=off.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbhx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-result", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
UNITS = {
    "bhx_body.hip": ["-ffp-contract=off"],
    "bhx_sort.hip": [],
    "bhx_tree.hip": [],
    "bhx_walk.hip": [],
    "bhx_api.hip": [],
    "bhx_grad.hip": [],
    "bhx_noalloc.hip": [],
    "bhx_dropin.hip": ["-ffp-contract=off"],
}
HOST_UNITS = {"bhx_ic.cpp": ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread"],
              "bhx_io.cpp": ["-O2", "-fPIC", "-std=c++17", "-Wall"]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = []
    for src, extra in UNITS.items():
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        jobs.append([HIPCC, "-x", "hip", *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    for src, extra in HOST_UNITS.items():
        obj = os.path.join(OUT_DIR, src.replace(".cpp", ".o"))
        jobs.append(["g++", *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stderr or r.stdout):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(OUT_DIR, s.rsplit(".", 1)[0] + ".o") for s in list(UNITS) + list(HOST_UNITS)]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
