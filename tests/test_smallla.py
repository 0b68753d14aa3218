"""Host small dense linear algebra of the deflated solver (cudalatticegauge_amd/csrc/clg_smallla.h),
compiled with g++ and checked against the defining identity of every routine."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_small_dense_linear_algebra(tmp_path):
    exe = str(tmp_path / "smallla_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cudalatticegauge_amd", "csrc"),
                           os.path.join(ROOT, "tests", "cpp", "smallla_check.cpp"), "-o", exe])
    out = subprocess.check_output([exe]).decode().split()
    vals = {out[i]: float(out[i + 1]) for i in range(0, len(out), 2)}
    print(vals)
    assert vals["lsq"] < 1e-10
    assert vals["qr"] < 1e-12
    assert vals["herm"] < 1e-12
    assert vals["gen"] < 1e-10
    assert vals["geneig"] < 1e-8
    assert vals["select"] < 1e-12
