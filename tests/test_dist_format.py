"""The .dist writer's number formatting (Output_Matrix, cms_cuda.cu:1305-1342,
prints through std::ostream with precision 6, i.e. printf %.6g). The writer
formats values in [1e-15, 1e5) with an integer routine (the binary value's
six significant digits, rounded half to even, then %g's notation rules);
tests/g6_check.cpp compares it with snprintf("%.6g") on hundreds of millions
of values, dyadic ties included."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "coordinated-meta-storms_amd", "csrc")


def test_g6_formatter_matches_printf(tmp_path):
    exe = str(tmp_path / "g6_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", CSRC, os.path.join(ROOT, "tests", "g6_check.cpp"),
                    os.path.join(CSRC, "host_tree.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
