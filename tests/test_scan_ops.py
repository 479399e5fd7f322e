"""The scan conversion's integer steps on the host (csrc/scan_ops.hpp, the
code the band kernels run): the stepped crossing column of every edge and
row against a direct 128-bit division per row, the stepped touched x-extent
and area deposits of every row against the same values divided afresh per
row, and the 8-pixel coverage groups against bit-by-bit extraction
(tests/cpu/scan_check.cpp, built with g++)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_edge_crossings_and_bits8(tmp_path):
    exe = str(tmp_path / "scan_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "raster-join_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpu", "scan_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")


def test_edge_area_host_build_equals_oracle(tmp_path):
    """The per-pixel area sums of the error-bounds kernel (csrc/scan_ops.hpp
    edge_area, host build tests/cpu/area_check.cpp) equal the oracle's
    orc_area bit for bit on random rings, including shallow and steep edges
    and vertices on pixel corners."""
    import numpy as np
    from oracle import pyoracle as O
    from common import star_polygons

    exe = str(tmp_path / "area_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "raster-join_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpu", "area_check.cpp"), "-o", exe], check=True)
    p = O.plan(np.array([0.0, 0.0, 1000.0, 1000.0]), 1000.0 * np.sqrt(2) / 80)
    xy, off = star_polygons(6, seed=11, box=1000.0, rmin=20, rmax=450, vmin=3, vmax=80)
    corner = np.array([0, 0, 250, 12.5, 987.5, 0, 999.9, 999.9, 12.5, 500], np.float32)
    rings = [xy[2 * off[k]:2 * off[k + 1]] for k in range(len(off) - 1)] + [corner]
    for v in rings:
        X, Y = O.snap(p, v)
        src = f"{p.W} {p.H} {len(X)}\n" + "\n".join(f"{int(a)} {int(b)}" for a, b in zip(X, Y))
        r = subprocess.run([exe], input=src.encode(), capture_output=True, timeout=60)
        assert r.returncode == 0
        got = np.frombuffer(r.stdout, np.int32).reshape(p.H, p.W)
        assert np.array_equal(got, O.area(p, v))
