"""The assembly launcher's guard on the fused update (CPU only).

A fused update (the previous step's v += dv, x += h v applied inside the
next assembly) is only correct with a warm CG start: a cold assembly zeroes
dv while neighbouring blocks still read it for the update.  It also needs at
least one CG iteration, since the previous solve's status is taken from
partials that a solve without iterations leaves where the assembly writes
its own.  launch_assemble refuses both cases with hipErrorInvalidValue
before any device call; a small C++ program linked against the library
checks that without a GPU."""
import os
import subprocess

from opencloth_amd.build import LIB, LIBDIR, ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fused_update_refused_without_warm_start(tmp_path):
    exe = str(tmp_path / "assemble_guard")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "opencloth_amd", "csrc"),
           os.path.join(HERE, "cpp", "assemble_guard.cpp"), "-o", exe, "-L", LIBDIR, "-lopencloth",
           f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"]
    assert os.path.exists(LIB)
    subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    invalid = int(lines[3].split()[1])
    assert invalid != 0
    for line in lines[:3]:
        assert int(line.split("->")[1]) == invalid, line
