"""Accuracy of the per-voxel logarithm (cuda_sto_pf_amd/csrc/pf_log.hpp).

The chemical potentials (thermo_utils.h:364-377) take log(x) and log(1-x);
the reference calls the CUDA library log.  pf_log is a short atanh-series
logarithm; this compiles its header for the host (the device build is the
same source with every multiply-add written as an explicit fma) and measures
it against an 80-bit long-double log over the whole normal range, the
chemical-potential domain [1e-12, 1 - 1e-12] and the neighbourhoods of 1 and
of the reduction boundary 1/sqrt2.  Bar: below one ulp everywhere, and
correctly rounded for at least 95% of the arguments; special values (zero,
negatives, infinities, NaN, subnormals) follow the library log.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "cuda_sto_pf_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pf_log_within_one_ulp(tmp_path):
    exe = str(tmp_path / "log_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "log_check.cpp"),
                    "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.split()
    max_ulp, frac_exact, n, specials = float(out[0]), float(out[1]), int(out[2]), int(out[3])
    assert n > 6_000_000
    assert max_ulp < 1.0, max_ulp
    assert frac_exact > 0.95, frac_exact
    assert specials == 1
