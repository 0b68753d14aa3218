"""Device fields loaded from and saved to configuration files (CField::InitialFieldWithFile /
SaveToFile), against the reference's expected values (TestConfigurationFileIO.cpp:12-110):
the Bridge++ 4^3x8 configuration has average plaquette 0.625129946974942 on the device after the
load's re-unitarisation, and a saved gauge or fermion field loads back unchanged with the MD5
string of the file's bytes."""
import hashlib
import os

import numpy as np
import pytest

import cudalatticegauge_amd as clg

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden", "configs")


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name,ftype", [("bridgepp_conf_text.dat", clg.FILE_BRIDGEPP_TXT),
                                        ("bridgepp_conf_binary.dat", clg.FILE_BRIDGEPP_BIN)])
@pytest.mark.parametrize("kind,prec", [(clg.GAUGE, clg.F64), (clg.GAUGE, clg.F32), (clg.GAUGE12, clg.F32)])
def test_bridgepp_load_gives_the_reference_plaquette(ctx, name, ftype, kind, prec):
    dims = (4, 4, 4, 8)
    g = clg.Field(ctx, kind, prec, dims)
    g.load_file(os.path.join(HERE, name), ftype)
    host = clg.file_read(os.path.join(HERE, name), ftype, clg.GAUGE, dims)
    assert np.abs(g.download() - host).max() < (1e-13 if prec == clg.F64 else 1e-6)
    g64 = g
    if kind != clg.GAUGE or prec != clg.F64:   # the plaquette energy reads FP64 18-real links
        g64 = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g64.upload(g.download())
    e = clg.plaquette_energy(ctx, g64, 1.0)        # (1/3) sum (3 - ReTr U_p)
    plaq = 1.0 - e / (6 * np.prod(dims))
    assert abs(plaq - 0.625129946974942) < (1e-12 if prec == clg.F64 else 1e-6), plaq


def test_clg_gauge_files_load_and_save(ctx, tmp_path):
    dims = (8, 8, 8, 8)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.load_file(os.path.join(HERE, "testGaugeDouble.con_"), clg.FILE_CLG_BIN_DOUBLE)
    d = g.download()
    assert abs(np.dot(d, d) - 49152.0) < 1e-6            # TestFileDOUBLE
    for ftype in (clg.FILE_CLG_BIN, clg.FILE_CLG_BIN_FLOAT):
        p = str(tmp_path / f"g{ftype}.con")
        md5 = g.save_file(p, ftype)
        assert md5 == hashlib.md5(open(p, "rb").read()).hexdigest().upper()
        g2 = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g2.load_file(p, ftype)
        tol = 0.0 if ftype == clg.FILE_CLG_BIN else 1e-7
        assert np.abs(g2.download() - d).max() <= tol


@pytest.mark.parametrize("kind", [clg.WILSON, clg.KS])
def test_fermion_save_load_round_trip(ctx, tmp_path, kind):
    """TestFileIOCLG: a saved fermion field loads back with zero difference."""
    dims = (4, 4, 6, 8)
    f = clg.Field(ctx, kind, clg.F64, dims)
    f.gaussian(9)
    p = str(tmp_path / "f.con")
    md5 = f.save_file(p)
    assert md5 == hashlib.md5(open(p, "rb").read()).hexdigest().upper()
    f2 = clg.Field(ctx, kind, clg.F64, dims)
    f2.load_file(p)
    clg.axpy(ctx, f2, -1.0, f)
    assert clg.norm2(ctx, f2) == 0.0
    with pytest.raises(clg.CLGError):                     # fermions load CLG binary files only
        f2.load_file(p, clg.FILE_CLG_BIN_FLOAT)
    with pytest.raises(clg.CLGError):                     # files hold full-lattice fields
        f.parity(clg.EVEN).save_file(p)


@pytest.mark.parametrize("fname,kind,expected", [("__f2.con_", clg.GAUGE, 2508152.599227),
                                                  ("__f18.con_", clg.KS, 1189.146596)])
def test_reference_save_load_fixtures_on_the_device(ctx, fname, kind, expected):
    # TestFileIOConsistency (TestConfigurationFileIO.cpp:174-204): the single-precision build's 8^4
    # files loaded into device fields, Dot(f, f).x / V against the section's Expected entries (1e-5)
    dims = (8, 8, 8, 8)
    f = clg.Field(ctx, kind, clg.F64, dims) if kind == clg.GAUGE else clg.Field(ctx, kind, clg.F64, dims,
                                                                               bc=(1, 1, 1, -1))
    f.load_file(os.path.join(HERE, fname), clg.FILE_CLG_BIN_FLOAT)
    if kind == clg.GAUGE:
        d = f.download()
        amp = float(np.dot(d, d)) / np.prod(dims)
    else:
        amp = clg.norm2(ctx, f) / np.prod(dims)
    assert abs(amp - expected) <= 1e-5 * expected, amp
