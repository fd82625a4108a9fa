"""The band-pruned register FFT (boostfractor.jl_amd/csrc/bf_fftreg.h) on the CPU.

The header's device code is plain C++ over a complex struct; a copy with the
device qualifiers removed is compiled with g++ into tests/fft_host/harness.cpp,
which runs the threads of a line as pthreads and checks, bit for bit, the
masked transforms against the unmasked ones (inputs zero outside the band,
outputs compared on the band), the whole-field completion (band outputs, then
the complement from the same last-pass input) on every element, and the
unmasked transform against a direct DFT.  g++ contracts no products here
(-ffp-contract=off), so "bit for bit" tests the arithmetic itself; the device
build may fuse products differently in the two variants (see the GPU test
test_dancer_pruned_transforms_match_unpruned).
"""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "boostfractor.jl_amd", "csrc", "bf_fftreg.h")


def _host_copy(text):
    text = text.replace('#include "bf_device.h"', "")
    text = text.replace("__host__ __device__ constexpr", "constexpr")
    text = text.replace("__device__ __forceinline__", "inline")
    text = text.replace("__device__ constexpr", "constexpr")
    assert "__device__" not in text
    return text


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pruned_register_fft_matches_unpruned(tmp_path):
    hdr = tmp_path / "fftreg_host.h"
    hdr.write_text(_host_copy(open(HEADER).read()))
    exe = tmp_path / "harness"
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-ffp-contract=off", f'-DFFTREG_HOST="{hdr}"',
                    os.path.join(HERE, "fft_host", "harness.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    m = re.search(r"total mismatches (\d+) max_dft_err ([\d.e+-]+)", out.stdout)
    assert m, out.stdout + out.stderr
    assert int(m.group(1)) == 0, out.stdout
    assert float(m.group(2)) < 1e-13, out.stdout
    assert out.returncode == 0
