"""Host emulation of the register-resident FFT (chapsim1.0_amd/csrc/chs_fftr.h): the same
header compiled with g++, the lanes that own a line run as threads meeting at the points where
the device code orders its LDS accesses.  Every planned length (64 ... 1024), both directions,
FP64 and FP32, with the complex LDS exchange and with the split one (real parts, then
imaginary parts), against a long-double DFT (FP64 error < 1e-14, FP32 < 2e-6 relative); the
outputs left in the lanes' registers must equal those handed to the output callback."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_register_fft_plans_match_dft(tmp_path):
    exe = tmp_path / "fftr_emu"
    subprocess.run(["g++", "-std=c++20", "-O1", "-pthread", os.path.join(HERE, "native", "fftr_emu.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 20
