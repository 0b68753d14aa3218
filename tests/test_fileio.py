"""Configuration files and the MD5 string, host side (no device): the reference's own configuration
fixtures (Bin/Debug/bridgepp_conf_text.dat, bridgepp_conf_binary.dat, testGaugeDouble.con_,
testGaugeSingle.con_, copied unchanged under tests/golden/configs) and its expected values
(TestConfigurationFileIO.cpp:12-37, 119-157, TestSuit_FileIO.yaml).

- Bridge++ text and binary hold one 4^3x8 configuration: both readers must give the same links,
  and its average plaquette is 0.625129946974942 (TestFileIO).  A wrong site order (x fastest in
  the file, t fastest in the host order) or a wrong byte order changes that value.
- The CLG double and float files hold one 8^4 SU(3) configuration: sum |U|^2 = 49152 =
  3 x 4 links x 8^4 (TestFileDOUBLE, TestFileFloat), every link unitary, the two precisions agree.
- MD5: RFC 1321's test suite and Python's hashlib on the bytes written.
- TestFileIOConsistency (TestConfigurationFileIO.cpp:174-204, section TestSaveLoad): the reference's
  __f2.con_ (an SU(3) gauge-type field) and __f18.con_ (a CFieldFermionKSSU3 field), 8^4, written by
  its single-precision build, load with Dot(f, f) / V equal to its Expected entries 2508152.599227
  and 1189.146596 within the test's 1e-5 relative bound.
- The compressed format (CFieldGaugeSU3.cu:1225-1265, 1651-1690) stores 9 reals of the principal
  logarithm A = log U per link.  The reference ships no compressed file (its test writes one and
  reads it back), so the logarithm is pinned by closed forms (diagonal links, links conjugated by a
  known unitary, a trace of 2 pi i) and by scipy's logm / expm; a file written here from scipy's
  logm must read back as U, and a round trip must meet TestFileIOCLGCompressed's |U' - U|^2 < 1e-8.
"""
import hashlib
import os

import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

HERE = os.path.join(os.path.dirname(__file__), "golden", "configs")
BRIDGE_DIMS = (4, 4, 4, 8)
CLG_DIMS = (8, 8, 8, 8)


def links(U):
    z = U[0::2] + 1j * U[1::2]
    return z.reshape(-1, 3, 3)


def test_md5_rfc1321_suite_and_hashlib():
    rfc = {
        b"": "D41D8CD98F00B204E9800998ECF8427E",
        b"a": "0CC175B9C0F1B6A831C399E269772661",
        b"abc": "900150983CD24FB0D6963F7D28E17F72",
        b"message digest": "F96B697D7CB7938D525A2F31AAF161D0",
        b"abcdefghijklmnopqrstuvwxyz": "C3FCD3D76192E4007DFB496CCA67E13B",
        b"12345678901234567890123456789012345678901234567890123456789012345678901234567890":
            "57EDF4A22BE3C955AC49DA2E2107B67A",
    }
    for msg, want in rfc.items():
        assert clg.md5(msg) == want, msg
    rng = np.random.default_rng(5)
    # every tail length around the one- and two-block padding boundaries, and a longer buffer
    for n in list(range(0, 140)) + [1 << 20]:
        b = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
        assert clg.md5(b) == hashlib.md5(b).hexdigest().upper(), n


def test_bridgepp_text_and_binary_give_the_reference_plaquette():
    txt = clg.file_read(os.path.join(HERE, "bridgepp_conf_text.dat"), clg.FILE_BRIDGEPP_TXT, clg.GAUGE, BRIDGE_DIMS)
    binr = clg.file_read(os.path.join(HERE, "bridgepp_conf_binary.dat"), clg.FILE_BRIDGEPP_BIN, clg.GAUGE, BRIDGE_DIMS)
    assert np.abs(txt - binr).max() < 1e-14
    V = int(np.prod(BRIDGE_DIMS))
    for U in (txt, binr):
        u = links(U)
        assert np.abs(u @ u.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
        plaq = 1.0 - orc.plaquette_sum(BRIDGE_DIMS, U) / 3.0 / (6 * V)
        assert abs(plaq - 0.625129946974942) < 1e-12, plaq
    # the same values taken in the host's own order (no reordering) do not give that plaquette
    raw = np.fromfile(os.path.join(HERE, "bridgepp_conf_binary.dat"), dtype=">f8").astype(np.float64)
    assert abs(1.0 - orc.plaquette_sum(BRIDGE_DIMS, raw) / 3.0 / (6 * V) - 0.625129946974942) > 1e-3


def test_clg_double_and_float_configurations():
    d = clg.file_read(os.path.join(HERE, "testGaugeDouble.con_"), clg.FILE_CLG_BIN_DOUBLE, clg.GAUGE, CLG_DIMS)
    s = clg.file_read(os.path.join(HERE, "testGaugeSingle.con_"), clg.FILE_CLG_BIN_FLOAT, clg.GAUGE, CLG_DIMS)
    assert abs(np.dot(d, d) - 49152.0) < 1e-6 and abs(np.dot(s, s) - 49152.0) < 1e-2
    ud = links(d)
    assert np.abs(ud @ ud.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    assert np.abs(d - s).max() < 1e-6
    # EFFT_CLGBin of a double build is the double layout
    assert np.array_equal(d, clg.file_read(os.path.join(HERE, "testGaugeDouble.con_"), clg.FILE_CLG_BIN, clg.GAUGE,
                                           CLG_DIMS))


@pytest.mark.parametrize("kind,per_site", [(clg.GAUGE, 72), (clg.WILSON, 24), (clg.KS, 6)])
def test_clg_binary_round_trip_and_md5(tmp_path, kind, per_site):
    dims = (4, 2, 6, 4)
    host = np.random.default_rng(3).normal(size=int(np.prod(dims)) * per_site)
    for ftype, dt in ((clg.FILE_CLG_BIN, "<f8"), (clg.FILE_CLG_BIN_DOUBLE, "<f8"), (clg.FILE_CLG_BIN_FLOAT, "<f4")):
        p = str(tmp_path / f"f{ftype}.con")
        md5 = clg.file_write(p, ftype, kind, dims, host)
        raw = open(p, "rb").read()
        assert md5 == hashlib.md5(raw).hexdigest().upper()
        assert np.array_equal(np.frombuffer(raw, dtype=dt), host.astype(dt))
        back = clg.file_read(p, ftype, kind, dims)
        assert np.array_equal(back, host.astype(dt).astype(np.float64))


@pytest.mark.parametrize("fname,kind,expected", [("__f2.con_", clg.GAUGE, 2508152.599227),
                                                  ("__f18.con_", clg.KS, 1189.146596)])
def test_reference_save_load_fixtures(fname, kind, expected):
    # the single-precision build's EFFT_CLGBin files are float; TestFileIOConsistency's amplitude
    # Dot(f, f).x / V against its Expected list
    dims = (8, 8, 8, 8)
    V = int(np.prod(dims))
    f = clg.file_read(os.path.join(HERE, fname), clg.FILE_CLG_BIN_FLOAT, kind, dims)
    amp = np.dot(f, f) / V
    assert abs(amp - expected) <= 1e-5 * expected, amp
    # read as doubles, the same bytes are the wrong size for the lattice
    with pytest.raises(clg.CLGError, match="bytes"):
        clg.file_read(os.path.join(HERE, fname), clg.FILE_CLG_BIN, kind, dims)


def test_file_errors(tmp_path):
    p = str(tmp_path / "g.con")
    clg.file_write(p, clg.FILE_CLG_BIN, clg.GAUGE, (2, 2, 2, 2), np.zeros(16 * 72))
    with pytest.raises(clg.CLGError, match="bytes"):       # size does not match the lattice
        clg.file_read(p, clg.FILE_CLG_BIN, clg.GAUGE, (2, 2, 2, 4))
    with pytest.raises(clg.CLGError, match="values"):      # text file with too few values
        clg.file_read(os.path.join(HERE, "bridgepp_conf_text.dat"), clg.FILE_BRIDGEPP_TXT, clg.GAUGE, (4, 4, 4, 16))
    with pytest.raises(clg.CLGError, match="values"):      # and with more values than the lattice holds
        clg.file_read(os.path.join(HERE, "bridgepp_conf_text.dat"), clg.FILE_BRIDGEPP_TXT, clg.GAUGE, (4, 4, 4, 4))
    with pytest.raises(clg.CLGError, match="fermion"):
        clg.file_read(p, clg.FILE_BRIDGEPP_BIN, clg.WILSON, (2, 2, 2, 2))
    with pytest.raises(clg.CLGError):
        clg.file_read(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, (2, 2, 2, 2))
    with pytest.raises(clg.CLGError):
        clg.file_write(p, clg.FILE_BRIDGEPP_BIN, clg.GAUGE, (2, 2, 2, 2), np.zeros(16 * 72))
    with pytest.raises(clg.CLGError):
        clg.file_read(str(tmp_path / "missing.con"), clg.FILE_CLG_BIN, clg.GAUGE, (2, 2, 2, 2))


def compressed_to_a(c):
    """9 reals per link (Re/Im A01, A02, A12, Im A00, A11, A22) -> anti-Hermitian 3x3 matrices."""
    c = c.reshape(-1, 9)
    A = np.zeros((c.shape[0], 3, 3), dtype=complex)
    A[:, 0, 1] = c[:, 0] + 1j * c[:, 1]
    A[:, 0, 2] = c[:, 2] + 1j * c[:, 3]
    A[:, 1, 2] = c[:, 4] + 1j * c[:, 5]
    A[:, 1, 0] = -np.conj(A[:, 0, 1])
    A[:, 2, 0] = -np.conj(A[:, 0, 2])
    A[:, 2, 1] = -np.conj(A[:, 1, 2])
    A[:, 0, 0], A[:, 1, 1], A[:, 2, 2] = 1j * c[:, 6], 1j * c[:, 7], 1j * c[:, 8]
    return A


def a_to_compressed(A):
    return np.stack([A[:, 0, 1].real, A[:, 0, 1].imag, A[:, 0, 2].real, A[:, 0, 2].imag, A[:, 1, 2].real,
                     A[:, 1, 2].imag, A[:, 0, 0].imag, A[:, 1, 1].imag, A[:, 2, 2].imag], axis=1).ravel()


def host_from_links(u):
    out = np.empty(u.shape[0] * 18)
    z = u.reshape(-1)
    out[0::2], out[1::2] = z.real, z.imag
    return out


def random_unitary(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3)))
    return q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]


def test_compressed_closed_forms(tmp_path):
    dims = (1, 1, 1, 3)  # 12 links
    rng = np.random.default_rng(21)
    # the last four put an eigenvalue at e^{-2 i f} for the library's first, second and third angle f
    # (0.7236, 2.1817, -1.3109), where that angle's Hermitian combination has a double eigenvalue, and
    # one at two of them at once
    th = np.array([[0.3, -0.5, 0.2], [2.5, 2.5, 2 * np.pi - 5.0], [-3.0, 1.0, 2.0 - 2 * np.pi + 2 * np.pi],
                   [1e-9, -2e-9, 1e-9], [0.0, 0.0, 0.0], [1.2, 1.2, -2.4], [3.1, -1.5, -1.6], [-2.0, -2.0, 4.0 - 2 * np.pi],
                   [-1.4472, 0.3, 1.1472], [-4.3634, -0.5, 4.8634], [2.6218, -1.0, -1.6218], [-1.4472, -4.3634, 5.8106]])
    th = np.angle(np.exp(1j * th))  # principal values in (-pi, pi]
    assert abs(th[1].sum() - 2 * np.pi) < 1e-12  # a link whose logarithm has trace 2 pi i
    W = random_unitary(rng, 12)
    W[:4] = np.eye(3)  # the first four links diagonal
    A = np.einsum("nij,nj,nkj->nik", W, 1j * th, W.conj())
    U = np.einsum("nij,nj,nkj->nik", W, np.exp(1j * th), W.conj())
    p = str(tmp_path / "c.con")
    md5 = clg.file_write(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims, host_from_links(U))
    raw = np.fromfile(p, dtype="<f8")
    assert raw.size == 12 * 9 and md5 == hashlib.md5(raw.tobytes()).hexdigest().upper()
    # diagonal links store exactly their phases, zero off the diagonal
    assert np.abs(raw.reshape(12, 9)[:4, :6]).max() < 1e-15
    assert np.abs(raw.reshape(12, 9)[:4, 6:] - th[:4]).max() < 1e-14
    assert np.abs(compressed_to_a(raw) - A).max() < 1e-13
    back = links(clg.file_read(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims))
    assert np.abs(back - U).max() < 1e-14


@pytest.mark.parametrize("eps", [0.5, 2.0])
def test_compressed_against_scipy_logm_expm(tmp_path, eps):
    from scipy.linalg import expm, logm
    dims = (2, 2, 4, 4)
    U = orc.gauge_random(dims, eps, 4)
    u = links(U)
    p = str(tmp_path / "c.con")
    clg.file_write(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims, U)
    A = compressed_to_a(np.fromfile(p, dtype="<f8"))
    assert A.shape[0] == u.shape[0]
    ev = np.linalg.eigvals(u)
    far = np.abs(np.angle(ev)).max(axis=1) < np.pi - 1e-6  # principal logarithm unambiguous
    assert far.sum() > 0.9 * len(far)
    for i in range(len(u)):
        assert np.abs(expm(A[i]) - u[i]).max() < 1e-13, i
        if far[i]:
            assert np.abs(A[i] - logm(u[i])).max() < 1e-11, i
    back = clg.file_read(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims)
    assert np.abs(back - U).max() < 1e-13
    # a file written independently from scipy's logm reads back as U
    q = str(tmp_path / "s.con")
    a_to_compressed(np.stack([logm(x) for x in u])).astype("<f8").tofile(q)
    assert np.abs(clg.file_read(q, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims) - U).max() < 1e-13


def test_compressed_round_trip_criterion(tmp_path):
    # TestFileIOCLGCompressed (TestConfigurationFileIO.cpp:133-160): save compressed, load, |U' - U|^2 < 1e-8
    dims = (8, 8, 8, 8)
    U = orc.gauge_random(dims, 1.0, 9)
    p = str(tmp_path / "c.con")
    clg.file_write(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims, U)
    d = clg.file_read(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, dims) - U
    assert np.dot(d, d) < 1e-8 and np.abs(d).max() < 1e-14
    with pytest.raises(clg.CLGError, match="gauge"):
        clg.file_write(p, clg.FILE_CLG_BIN_COMPRESSED, clg.WILSON, dims, np.zeros(int(np.prod(dims)) * 24))
    with pytest.raises(clg.CLGError, match="bytes"):
        clg.file_read(p, clg.FILE_CLG_BIN_COMPRESSED, clg.GAUGE, (8, 8, 8, 4))
