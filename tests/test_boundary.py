"""The C++ drop-in for the reference class (include/FluidSolver3D_hip.h),
compiled here with g++ into a program that follows main.cpp:9-60, linked
against the CPU oracle (it exports the same fs_* ABI as the HIP library) and
run in a scene directory: the files it writes agree with main.cpp's sequence
made through the ABI directly (within 1e-5; the oracle's OpenMP reductions
do not fix their summation order)."""
import os
import subprocess

import pytest

from scene_files import compare_outputs, run_like_main, scene_dir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_style_program_against_wrapper(oracle, tmp_path, monkeypatch):
    exe = tmp_path / "main_like"
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "boundary", "main_like.cpp"),
           "-o", str(exe), oracle.path, "-Wl,-rpath," + os.path.dirname(oracle.path)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    N = 20
    d = tmp_path / "program"
    scene_dir(d, N)
    r = subprocess.run([str(exe)], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "alpha = 0.05" in r.stdout and "Number of particles is" in r.stdout
    o = tmp_path / "abi"
    scene_dir(o, N)
    monkeypatch.chdir(o)
    run_like_main(oracle, "InputData/config.json")
    compare_outputs(d, o, 3)


def test_wrapper_rethrows_abi_errors(oracle, tmp_path):
    """A missing label file: init() throws std::runtime_error with the ABI's
    message, as the reference's init throws (FluidSolver3D.cu:84)."""
    src = tmp_path / "bad.cpp"
    src.write_text('#include "FluidSolver3D_hip.h"\n#include <iostream>\n'
                   'int main() { FluidSolver3D s(5, 5, 5, 0.1f, 0.02f);\n'
                   '  try { s.init("no/such/labels.txt"); } catch (const std::runtime_error& e) {\n'
                   '    std::cout << e.what() << std::endl; return 3; }\n  return 0; }\n')
    exe = tmp_path / "bad"
    r = subprocess.run(["g++", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe), oracle.path,
                        "-Wl,-rpath," + os.path.dirname(oracle.path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "Cannot open file: no/such/labels.txt" in r.stdout
