"""The spring-once assembly's tile strips (CPU only): every owner outside a
tile whose spring reaches into it has one LDS column, and every partner in
the tile finds its owner's column, for each owned direction and several tile
shapes.  A small C++ program over the library's header checks it."""
import os
import subprocess

from opencloth_amd.build import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def test_strip_tables(tmp_path):
    exe = str(tmp_path / "strip_map")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "opencloth_amd", "csrc"),
           os.path.join(HERE, "cpp", "strip_map.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
