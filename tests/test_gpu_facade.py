"""CLGTest sections of the reference's test suite replayed through the C++ facade (CLGLib class
names over the HIP library), checked by the reference tests' own criteria and expected values."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cudalatticegauge_amd", "clg_test_runner")
YAML = os.path.join(ROOT, "tests", "golden", "clgtest_sections.yaml")


def run(section, test, cwd=None):
    print(f"running {section} {test}", flush=True)
    out = subprocess.run([RUNNER, YAML, section, test], capture_output=True, timeout=240, cwd=cwd)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    res = json.loads(out.stdout.decode().strip().splitlines()[-1])
    print(res)
    return res


@pytest.mark.parametrize("section", ["TestSolverBiCGStab", "TestSolverGCRODR",
                                     "TestSolverGMRESLowMode", "TestSolverGCRODRLowMode", "TestSolverBiCGStabLowMode",
                                     "TestEOSolverBiCGStab", "TestEOSolverGMRES"])
def test_solver_sections(section):
    # (TestSolverGMRES runs the rotating-frame fermion CFieldFermionWilsonSquareSU3DR with Dirichlet
    # boundaries, a special-frame variant this library does not build.)
    # TestSolver (TestSolver.cpp:12-121): every |.|^2 error below ExpectedErr (default 1e-4)
    r = run(section, "solver")
    assert all(e < r["expected"] for e in r["errors"]), r


@pytest.mark.parametrize("section", ["TestFermionUpdatorKS", "TestFermionUpdatorKSNestedForceGradient",
                                     "TestFermionUpdatorNestedForceGradient"])
def test_updator_sections(section):
    # TestUpdateCommon (TestUpdator.cpp:12-140): plaquette within 0.15 of ExpectedRes, at most
    # ExpectMiss rejections, H_diff < 0.12 and |last H_diff| < 0.1
    r = run(section, "update")
    assert abs(r["plaquette_avg"] - r["expected"]) < 0.15, r
    assert r["accepted"] >= r["expect_accept"], r
    assert r["hdiff"] < 0.12 and r["last_hdiff"] < 0.1, r


@pytest.mark.parametrize("section", ["TestFileIOBridgePPText", "TestFileIOBridgePPBin"])
def test_fileio_sections(section, tmp_path):
    # TestFileIO (TestConfigurationFileIO.cpp:12-37): the loaded configuration's average plaquette
    # within 1e-6 of ExpectedRes.  The sections name ../Debug/<file>, as the reference's Bin/Release
    # working directory sees them; the fixtures stand in for that folder.
    (tmp_path / "Debug").symlink_to(os.path.join(ROOT, "tests", "golden", "configs"))
    (tmp_path / "Release").mkdir()
    r = run(section, "fileio", cwd=str(tmp_path / "Release"))
    assert abs(r["plaquette"] - r["expected"]) < 1e-6, r


@pytest.mark.parametrize("section", ["TestMSKSSolverFOM", "TestMSKSSolverBiCGStab", "TestMSKSSolverNested"])
def test_multishift_ks_sections(section):
    # TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): every |R^-1/n R^1/n phi - phi|^2
    # below ExpectedErr (default 1e-5)
    r = run(section, "mssolver_ks")
    assert all(e < r["expected"] for e in r["errors"]), r
