"""The C++ facade (include/skgemm.hpp) used the way a caller of the
reference classes uses them: tests/cpp/facade_check.cpp is built with g++
against libskgemm.so and run; it checks gemm<double>, sme_syrk<double>,
sme_gemm_trp_3d<double> and linalg_pack/unpack_c16_f64 against closed forms
and naive loops and exits non-zero on a mismatch."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stream-k-gemm_amd")


def build_facade_check(out_dir):
    exe = os.path.join(out_dir, "facade_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"),
                    "-L", PKG, "-lskgemm", "-Wl,-rpath," + PKG, "-o", exe],
                   check=True, capture_output=True, text=True)
    return exe


def test_facade_builds_against_the_c_abi(skg, tmp_path):
    assert os.path.exists(build_facade_check(str(tmp_path)))


@pytest.mark.gpu
def test_facade_program_runs(skg, tmp_path):
    exe = build_facade_check(str(tmp_path))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all facade checks passed" in r.stdout
