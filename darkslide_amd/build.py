"""Builds the gfx950 shared library libdarkslide_gpu.so in-tree.

Every .hip file under csrc/ is compiled for gfx950 with FMA contraction off
(parity kernels keep the reference's rounding; files that allow contraction
say so with a pragma) and without packed-FP32 instructions, then linked into
one shared library next to this file.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
LIB = os.path.join(HERE, "libdarkslide_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is expected
    noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    return src, obj, r.returncode, "\n".join(noise)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    failed = [r for r in results if r[2] != 0]
    for src, _, rc, err in results:
        if err and (verbose or rc):
            print(f"--- {os.path.basename(src)}\n{err}", file=sys.stderr)
    if failed:
        raise RuntimeError("HIP compile failed: " + ", ".join(os.path.basename(r[0]) for r in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + [r[1] for r in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_mirror_smoke():
    """Links tests/cpp/mirror_smoke against the library (plain g++, C++17)."""
    src = os.path.join(ROOT, "tests", "cpp", "mirror_smoke.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "mirror_smoke")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-o", out,
           "-L" + HERE, "-ldarkslide_gpu", "-Wl,-rpath," + "$ORIGIN/../../darkslide_amd", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("mirror_smoke build failed:\n" + r.stderr)
    return out


def build_libm_exhaustive():
    """Builds tests/cpp/libm_exhaustive: the host glibc against the kernels'
    libm restatements (csrc/libm_ref.h compiled for the CPU), no GPU code."""
    src = os.path.join(ROOT, "tests", "cpp", "libm_exhaustive.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "libm_exhaustive")
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-I" + CSRC, src, "-o", out, "-pthread", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libm_exhaustive build failed:\n" + r.stderr)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
