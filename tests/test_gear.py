"""C++ unit tests of the gear layer (memory manager, compile-time scheduler),
built with g++ and run on the host."""
import subprocess
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_gear_unit_tests():
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "test_gear"
        build = subprocess.run(
            ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
             f"-I{ROOT / 'cuda_hlt_amd' / 'csrc' / 'host'}", "-I/opt/rocm/include",
             str(ROOT / "tests" / "cpp" / "test_gear.cpp"), "-o", str(exe), "-L/opt/rocm/lib", "-lamdhip64",
             "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
        assert build.returncode == 0, build.stderr
        run = subprocess.run([str(exe)], capture_output=True, text=True)
        assert run.returncode == 0, run.stdout + run.stderr
        assert "ok" in run.stdout
