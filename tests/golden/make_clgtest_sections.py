"""Extracts the CLGTest YAML sections that the facade runner replays (inputs and the reference's
expected results) from the reference's test-suite files into clgtest_sections.yaml.

    python tests/golden/make_clgtest_sections.py /root/reference/Bin/Debug
"""
import os
import sys

SECTIONS = {
    "TestSuit_Solver.yaml": ["TestSolverBiCGStab", "TestSolverGCRODR", "TestSolverGMRESLowMode",
                             "TestSolverGCRODRLowMode", "TestSolverBiCGStabLowMode", "TestMSKSSolverFOM",
                             "TestMSKSSolverBiCGStab", "TestMSKSSolverNested"],
    "TestSuit_EvenOdd.yaml": ["TestEOSolverBiCGStab", "TestEOSolverGMRES"],
    "TestSuit_FermionUpdatorKS.yaml": ["TestFermionUpdatorKS", "TestFermionUpdatorKSNestedForceGradient"],
    "TestSuit_FermionUpdator.yaml": ["TestFermionUpdatorNestedForceGradient", "TestFermionUpdatorLargeScale",
                                     "TestFermionUpdatorLargeScaleFloat"],
    "TestSuit_Common.yaml": ["TestOperators", "TestGamm5Hermiticity", "TestAnitiHermiticity", "TestGaugeInvarience"],
    "TestSuit_FileIO.yaml": ["TestFileIOBridgePPText", "TestFileIOBridgePPBin", "TestSaveConfiguration",
                             "TestFileIOCLGCompressed", "TestFileIOCLGCompressedDebug", "TestSaveConfigurationDouble",
                             "TestSaveConfigurationFloat"],
}


def section(lines, name):
    out, inside = [], False
    for ln in lines:
        top = ln and not ln[0].isspace() and ln.rstrip().endswith(":")
        if top:
            if inside:
                break
            inside = ln.strip() == name + ":"
        if inside:
            out.append(ln.rstrip("\n"))
    return out


def main(src):
    text = ["# CLGTest sections copied verbatim from the reference test suite (data: inputs and expected",
            "# results), see make_clgtest_sections.py", ""]
    for fn, names in SECTIONS.items():
        lines = open(os.path.join(src, fn)).read().splitlines(True)
        for n in names:
            sec = section(lines, n)
            assert sec, (fn, n)
            text += [f"# from {fn}"] + sec + [""]
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "clgtest_sections.yaml"), "w") as f:
        f.write("\n".join(text) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Bin/Debug")
