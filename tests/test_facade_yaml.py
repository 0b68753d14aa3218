"""The CLGLib facade's YAML reader on the reference's own CLGTest sections (no device needed)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "cudalatticegauge_amd", "clg_test_runner")
YAML = os.path.join(ROOT, "tests", "golden", "clgtest_sections.yaml")


def dump(section):
    if not os.path.exists(RUNNER):
        pytest.skip("facade runner not built")
    return json.loads(subprocess.check_output([RUNNER, YAML, section, "dump"]).decode())


def test_updator_section_parses():
    s = dump("TestFermionUpdatorKS")
    assert s["LatticeLength"] == "[4, 4, 4, 8]"
    assert s["Updator"]["IntegratorType"] == "CIntegratorLeapFrog"
    assert s["Updator"]["IntegratorStep"] == "60"
    assert s["FermionField1"]["FieldName"] == "CFieldFermionKSSU3"
    md = [float(v) for v in s["FermionField1"]["MD"].strip("[]").split(",")]
    assert len(md) == 7 and abs(md[0] - 0.39046039002765764) < 1e-15
    assert s["MSSolver"]["SolverName"] == "CMultiShiftBiCGStab"
    assert s["Action1"]["Beta"] == "5.0" and s["Action2"]["ActionName"] == "CActionFermionKS"
    assert s["ExpectedRes"] == "[0.420549]"


def test_solver_and_nested_sections_parse():
    s = dump("TestSolverGCRODR")
    assert s["Solver"]["SolverName"] == "CSLASolverGCRODR"
    assert s["Solver"]["MDim"] == "6" and s["Solver"]["KDim"] == "3"
    assert s["Solver"]["DeflateSpace"] == "EEDT_SVD"
    w = dump("TestFermionUpdatorNestedForceGradient")
    assert w["Updator"]["NestedStep"] == "3" and w["FermionField1"]["Hopping"] == "0.1575"
    # comment-only and blank lines inside a section do not end it
    assert w["Measure1"]["MeasureName"] == "CMeasurePlaqutteEnergy"
