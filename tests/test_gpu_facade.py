"""CLGTest sections of the reference's test suite replayed through the C++ facade (CLGLib class
names over the HIP library), checked by the reference tests' own criteria and expected values."""
import hashlib
import json
import os
import subprocess

import numpy as np

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


@pytest.mark.parametrize("section,test", [("TestSaveConfigurationDouble", "fileio_double"),
                                          ("TestSaveConfigurationFloat", "fileio_float")])
def test_fileio_double_float_sections(section, test, tmp_path):
    # TestFileDOUBLE / TestFileFloat (TestConfigurationFileIO.cpp:101-162): Dot(U, U).x of the reference's
    # 8^4 configuration within 1e-6 of 49152; the files sit in ../Debug of the working directory
    (tmp_path / "Debug").symlink_to(os.path.join(ROOT, "tests", "golden", "configs"))
    (tmp_path / "Release").mkdir()
    r = run(section, test, cwd=str(tmp_path / "Release"))
    assert abs(r["dot"][0] - r["expected"]) <= 1e-6 and abs(r["dot"][1]) <= 1e-6, r


def test_fileio_save_section(tmp_path):
    # TestFileIOCLG (TestConfigurationFileIO.cpp:55-99): the identity gauge field and a Gaussian Wilson
    # field saved as EFFT_CLGBin and read back: <U', U> / (V Dir) = 3 and |psi' - psi|^2 = 0 within 1e-6
    r = run("TestSaveConfiguration", "fileio_save", cwd=str(tmp_path))
    assert abs(r["gauge"][0] - 3.0) <= 1e-6 and abs(r["gauge"][1]) <= 1e-6, r
    assert abs(r["fermion"][0]) <= 1e-6 and abs(r["fermion"][1]) <= 1e-6, r
    for key, name in (("md5_gauge", "testGauge.con"), ("md5_fermion", "testFermion.con")):
        assert r[key] == hashlib.md5((tmp_path / name).read_bytes()).hexdigest().upper(), r


@pytest.mark.parametrize("section", ["TestFileIOCLGCompressedDebug", "TestFileIOCLGCompressed"])
def test_fileio_compressed_sections(section, tmp_path):
    # TestFileIOCLGCompressed (TestConfigurationFileIO.cpp:133-160): a random gauge field (4^4, and 32^4
    # in the release section) saved as per-link logarithms and read back differs by |U' - U|^2 < 1e-8
    r = run(section, "fileio_compressed", cwd=str(tmp_path))
    assert abs(complex(*r["res"])) < 1e-8, r
    raw = (tmp_path / "testGaugeCompressed.con").read_bytes()
    assert len(raw) % 72 == 0 and r["md5"] == hashlib.md5(raw).hexdigest().upper()


@pytest.mark.parametrize("section", ["TestMSKSSolverFOM", "TestMSKSSolverBiCGStab", "TestMSKSSolverNested"])
def test_multishift_ks_sections(section):
    # TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): every |R^-1/n R^1/n phi - phi|^2
    # below ExpectedErr (default 1e-5)
    r = run(section, "mssolver_ks")
    assert all(e < r["expected"] for e in r["errors"]), r


def test_operators_section():
    # TestOperators (TestCommon.cpp:13-88): the Gaussian-source norm estimators f.f / 12V near 1 (cross
    # terms near 0), and <f2, D f1> = <D^dagger f2, f1> within 0.02 (real) / 0.01 (imaginary).
    # The reference bounds the estimators by 0.01, but on this 4x4x4x8 lattice (n = 6144 complex
    # components of unit variance) their own standard deviation is 1/sqrt(n) = 0.0128 for f.f / n and
    # sqrt(1/2n) = 0.009 for each part of f1.f2 / n, so 0.01 passes or fails with the random stream.
    # Here they are held to 5 standard deviations; the generator's variance itself is checked on 16^4 to
    # 5 standard deviations there, 0.0056 (tests/test_oracle_core.py; the device generator equals the
    # oracle's bit for bit, tests/test_gpu_wilson.py).
    r = run("TestOperators", "operators")
    n = r["n"]
    s_nn, s_x = 1.0 / np.sqrt(n), np.sqrt(0.5 / n)
    for key in ("f1f1", "f2f2"):
        assert abs(1 - r[key][0]) <= 5 * s_nn and r[key][1] == 0.0, r
    assert max(abs(v) for v in r["f1f2"] + r["f2f1"]) <= 5 * s_x, r
    a, b = r["f2_Df1"], r["Ddag_f2_f1"]
    assert abs(a[0] - b[0]) <= 0.02 and abs(a[1] - b[1]) <= 0.01, r
    # the identity itself holds to FP64 rounding
    assert abs(complex(*a) - complex(*b)) <= 1e-12 * abs(complex(*a)), r


@pytest.mark.parametrize("gamma5_test", [0, 1])
def test_gamma5_hermiticity_section(gamma5_test, tmp_path):
    # TestGamma5Hermiticity (TestCommon.cpp:40-50, TestGamm5Hermiticity section): no element of the
    # dense operator breaks gamma5 D = (gamma5 D)^dagger (GAMM5Test 0) or D^dagger = (D)^dagger
    # (GAMM5Test 1); the runner reports both forms
    r = run("TestGamm5Hermiticity", "gamma5")
    assert r["gamma5_test"] == 1
    errs = {1: (r["errors"], r["max_rel_dev"]), 0: (r["other_errors"], r["other_max_rel_dev"])}[gamma5_test]
    assert errs[0] == 0 and errs[1] < 1e-12, r


def test_anti_hermiticity_section():
    # TestAnitiHermiticity (TestCommon.cpp:153-163): the dense massless staggered D is anti-Hermitian
    r = run("TestAnitiHermiticity", "antihermitian")
    assert r["errors"] == 0 and r["max_rel_dev"] < 1e-12, r


def test_gauge_invariance_section():
    # TestGaugeInvarience (TestCommon.cpp:206-239): every action's energy changes by at most 1e-7 under a
    # random gauge transformation of the gauge field and the pseudofermion, which really moves the links
    r = run("TestGaugeInvarience", "gauge_invariance")
    assert r["gauge_moved"] > 1.0, r
    assert len(r["before"]) == len(r["after"]) >= 2
    for b, a in zip(r["before"], r["after"]):
        assert abs(b - a) <= 1e-7, r
        assert abs(b - a) <= 1e-12 * abs(b), r


def test_fermion_updator_large_scale_section():
    # TestFermionUpdatorL on TestFermionUpdatorLargeScale (TestFermionUpdator.cpp:160-219): 32^3x16 Wilson
    # Nf=2, kappa 0.1355, beta 2.5, nested force gradient 8 x 6, GMRES(12) force solves; one trajectory,
    # then Updates-1 in the H_diff test mode.  The FP64 criterion on H_diff (< 1e-8 H) is checked; the
    # plaquette is reported only: ExpectedRes (0.166393, to 1e-5) is the value the reference's own random
    # stream reaches from its EFIT_Random start, which a different generator does not reproduce.
    r = run("TestFermionUpdatorLargeScale", "update_l")
    print(f"seconds per trajectory {r['seconds_per_trajectory']} (the reference README quotes 404 s FP64 on a "
          f"GTX 1070 laptop GPU); plaquette {r['plaquette_avg']:.6f} (reference stream: {r['expected']})")
    assert r["trajectories"] == 5 and len(r["seconds_per_trajectory"]) == 5, r
    assert r["h"] > 1e6, r
    assert r["hdiff"] < 1e-8 * r["h"], r
    assert 0.0 < r["plaquette_avg"] < 1.0, r


def test_fermion_updator_large_scale_float_section():
    # TestFermionUpdatorL on TestFermionUpdatorLargeScaleFloat (TestFermionUpdator.cpp:160-219, the
    # single-precision build's branch): 32^3x32 Wilson Nf=2, nested force gradient 12 x 8 with an inner
    # leapfrog, exact link exponential (ExponentialPrecision 0).  This library runs the section in FP64,
    # which the single-precision criteria bound from above: plaquette within 0.01 of ExpectedRes (0.17)
    # and H_diff < 3e-7 H.
    r = run("TestFermionUpdatorLargeScaleFloat", "update_l")
    print(f"seconds per trajectory {r['seconds_per_trajectory']} (the reference README quotes 214 s for one "
          f"32^3x16 trajectory of its single-precision build on a GTX 1070 laptop GPU; this section is 32^3x32)")
    assert r["trajectories"] == 5 and len(r["seconds_per_trajectory"]) == 5, r
    assert r["h"] > 1e6, r
    assert r["hdiff"] < 3e-7 * r["h"], r
    assert abs(r["plaquette_avg"] - r["expected"]) <= 0.01, r
