"""Tile order of the persistent all-pairs launches (csrc/tile_order.h),
checked on the host by tests/tile_order_check.cpp: grouped bands of block
rows visit every tile of a whole-row launch exactly once, and partial-row
launches keep the row-major order. No GPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "coordinated-meta-storms_amd", "csrc")


def test_tile_order(tmp_path):
    exe = str(tmp_path / "tile_order_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", CSRC,
                    os.path.join(ROOT, "tests", "tile_order_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "ok"
