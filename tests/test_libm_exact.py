"""The kernels' restatements of glibc's cbrtf and expf (darkslide_amd/csrc/
libm_ref.h), which Rust's f32::cbrt (Oklab, color.rs:408-412) and f32::exp
(NLM weights, enhance.rs:57) call, compared with the host glibc on every one
of the 2^32 float inputs.  The same header is compiled for the CPU by g++;
the GPU build is checked against the same glibc by hash in
test_gpu_exact.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "libm_exhaustive")


def libm_exe():
    if not os.path.exists(EXE):
        from darkslide_amd import build as hipbuild
        hipbuild.build_libm_exhaustive()
    return EXE


@pytest.mark.slow
def test_cbrtf_expf_bit_equal_to_glibc_on_all_inputs():
    r = subprocess.run([libm_exe(), "check"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    counts = dict(line.split(" mismatches ") for line in r.stdout.strip().splitlines())
    assert counts == {"cbrtf": "0", "expf": "0"}, r.stdout + r.stderr
    assert r.returncode == 0
