"""CPU check of the FP64 phase sin/cos used by the assembly kernels: the same
header compiled for the host, compared with the C library over the phase range
the kernels see and across quadrant boundaries (absolute error < 4e-16); the
form that takes the phase in pi/64 steps is compared with the exact phase of its
argument, and the wavenumber scale k 64/pi with its exact value (one rounding)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_sincos_phase_matches_libm(tmp_path):
    exe = tmp_path / "sincos_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-mfma", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "native", "sincos_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
