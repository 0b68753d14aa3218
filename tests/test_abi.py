"""CPU-side checks of the C-ABI library: it loads, exports every entry point that
include/clg_hip.h declares, and rejects bad arguments without touching a GPU."""
import ctypes as C
import os
import re

import pytest

import cudalatticegauge_amd as clg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "clg_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(clg_\w+)\s*\(", txt, re.M)))


def test_library_is_built_and_loads():
    assert os.path.exists(clg.LIB_PATH), "run __graft_entry__.build() first"
    clg.lib()


def test_exports_every_declared_symbol():
    L = C.CDLL(clg.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/clg_hip.h but not exported"
    assert set(syms) <= set(clg.EXPORTED) | set(syms)
    for s in clg.EXPORTED:
        assert s in syms, f"binding {s} is not declared in the header"


def test_status_strings_and_null_checks():
    L = clg.lib()
    assert L.clg_status_string(0) == b"ok"
    assert L.clg_status_string(1).startswith(b"solver did not converge")
    # null output handle is rejected before any device call
    assert L.clg_init(0, None) == 2
    assert b"null" in L.clg_last_error()
    assert L.clg_field_create(None, 0, 0, 0, None, None) == 2
    assert L.clg_field_destroy(None) == 0


def test_solver_defaults_follow_reference_yaml_defaults():
    p = clg.solver_params()
    assert p.max_dim == 20 and p.restart == 3 and p.accuracy == pytest.approx(1e-8)
    assert p.stop_rule == clg.STOP_REFERENCE and p.inner_prec == clg.F64


def test_tuner_and_graph_entry_points_check_arguments_without_a_device():
    L = clg.lib()
    assert L.clg_autotune_enable(2) == 2 and L.clg_autotune_enable(-2) == 2
    assert L.clg_autotune_force(9, 256, 8, 1) == 2 and L.clg_autotune_force(1, 96, 8, 1) == 2
    assert L.clg_autotune_force(1, 128, 3, 1) == 2 and L.clg_autotune_force(1, 128, 8, 3) == 2
    assert L.clg_autotune_force(0, 0, 0, 0) == 0   # clearing needs no device
    assert L.clg_autotune_enable(-1) == 0
    assert L.clg_autotune_count() >= 0
    buf = C.create_string_buffer(64)
    assert L.clg_autotune_entry(-1, buf, 64) == 2
    assert L.clg_graph_begin(None) == 2
    assert L.clg_graph_launch(None) == 2
    assert L.clg_graph_nodes(None) == -1
    assert L.clg_graph_destroy(None) == 0
    assert L.clg_multidot(None, 1, None, None, None) == 2
    assert L.clg_vector_multiply_matrix(None, 1, None, 0, None, None) == 2
    assert L.clg_ks_hop(None, None, None, None, 1.0, None, 0.0) == 2
    assert L.clg_ks_schur(None, None, None, None, 0.1) == 2


def test_integration_snippet_compiles(tmp_path):
    """The CLGTest-style body in INTEGRATION.md section 1 compiles and links against the facade."""
    import re
    import shutil
    import subprocess
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"## 1\. CLGLib-side adapter.*?```cpp\n(.*?)```", txt, re.S)
    assert m, "no C++ block in INTEGRATION.md section 1"
    src = tmp_path / "snippet.cpp"
    src.write_text(m.group(1))
    pkg = os.path.join(ROOT, "cudalatticegauge_amd")
    if not os.path.exists(os.path.join(pkg, "libclg_facade.so")):
        pytest.skip("libclg_facade.so is not built")
    gxx = shutil.which("g++")
    assert gxx
    out = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(pkg, "facade"), "-I",
                          os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "snippet"), "-L", pkg,
                          "-lclg_facade", "-lclg_hip", "-Wl,-rpath," + pkg], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-3000:]
