"""The index arithmetic of the wave-level FFTs (snf_wfft.h: 64 lanes, N =
64 R0; snf_wfft32.h: 32-lane halves, N = 32 R0) on the host: every stage is a
__host__ __device__ function of one lane, so tests/wfft_host_check.hip runs
them lane by lane in the order of the device drivers and compares both
directions (lane layout -> task layout and back) with a direct DFT in long
double.  Bar: relative L2 <= 1e-13 (FP64 rounding of a few passes)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not found")
def test_wave_fft_stages_on_host(tmp_path):
    exe = tmp_path / "wfchk"
    subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17",
                    os.path.join(HERE, "wfft_host_check.hip"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "N=96" in r.stdout and "N=384" in r.stdout
