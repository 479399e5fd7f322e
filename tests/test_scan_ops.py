"""The scan conversion's integer steps on the host (csrc/scan_ops.hpp, the
code the band kernels run): the stepped crossing column of every edge and
row against a direct 128-bit division per row, and the 8-pixel coverage
groups against bit-by-bit extraction (tests/cpu/scan_check.cpp, built with
g++)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_edge_crossings_and_bits8(tmp_path):
    exe = str(tmp_path / "scan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "raster-join_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpu", "scan_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
