"""Krylov BLAS entry points (multi-dot, multi-axpy, VectorMultiplyMatrix), the autotuner controls
and graph recording/replay, checked through the C ABI against numpy on the downloaded fields.

Reference behaviour: CField::Dot / Axpy (CField.h:127-170), CFieldMatrixOperation::VectorMultiplyMatrix
(_kernelMatrixMultiply, CFieldFermionWilsonSquareSU3.cu:1533-1560: res_i <- sum_j W_j M[j*Y+i]).
"""
import numpy as np
import pytest

import cudalatticegauge_amd as clg

pytestmark = pytest.mark.gpu

DIMS = (4, 4, 4, 8)


def cvec(f):
    a = f.download()
    return a[0::2] + 1j * a[1::2]


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


def fields(ctx, n, kind=clg.WILSON, prec=clg.F64, subset=clg.FULL, seed=1):
    out = []
    for i in range(n):
        f = clg.Field(ctx, kind, prec, DIMS, subset=subset)
        f.gaussian(seed + 17 * i)
        out.append(f)
    return out


@pytest.mark.parametrize("kind,prec,subset,n", [(clg.WILSON, clg.F64, clg.FULL, 5), (clg.WILSON, clg.F64, clg.EVEN, 20),
                                                (clg.KS, clg.F64, clg.FULL, 3), (clg.WILSON, clg.F32, clg.ODD, 7),
                                                (clg.KS, clg.F64, clg.FULL, 33), (clg.WILSON, clg.F32, clg.EVEN, 40),
                                                (clg.KS, clg.F32, clg.FULL, 64), (clg.KS, clg.F64, clg.FULL, 70)])
def test_multidot_matches_numpy(ctx, kind, prec, subset, n):
    V = fields(ctx, n, kind, prec, subset, seed=3)
    w = fields(ctx, 1, kind, prec, subset, seed=99)[0]
    got = clg.multidot(ctx, V, w)
    wv = cvec(w)
    want = np.array([np.vdot(cvec(v), wv) for v in V])
    tol = 1e-12 if prec == clg.F64 else 1e-5
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want)) * 10 + tol
    # single-vector case equals clg_dot
    assert abs(got[0] - clg.dot(ctx, V[0], w)) <= 1e-12 * abs(want[0]) + 1e-12


@pytest.mark.parametrize("n", [1, 4, 17, 33, 61, 70])
def test_multiaxpy_matches_numpy(ctx, n):
    V = fields(ctx, n, seed=5)
    w = fields(ctx, 1, seed=77)[0]
    rng = np.random.default_rng(n)
    c = rng.normal(size=n) + 1j * rng.normal(size=n)
    want = cvec(w) + sum(ck * cvec(v) for ck, v in zip(c, V))
    clg.multiaxpy(ctx, w, V, c)
    assert np.max(np.abs(cvec(w) - want)) <= 1e-13 * np.max(np.abs(want))
    with pytest.raises(clg.CLGError):
        clg.multiaxpy(ctx, w, [w], [1.0])


@pytest.mark.parametrize("nres,nleft", [(1, 0), (5, 4), (8, 8), (12, 20), (31, 1)])
def test_vector_multiply_matrix(ctx, nres, nleft):
    res = fields(ctx, nres, seed=11)
    left = fields(ctx, nleft, seed=500)
    W = np.array([cvec(f) for f in res + left])
    rng = np.random.default_rng(nres * 100 + nleft)
    M = rng.normal(size=(nres + nleft, nres)) + 1j * rng.normal(size=(nres + nleft, nres))
    want = M.T @ W  # res_i = sum_j W_j M[j, i]
    clg.vector_multiply_matrix(ctx, res, left, M)
    for i in range(nres):
        assert np.max(np.abs(cvec(res[i]) - want[i])) <= 1e-13 * np.max(np.abs(want[i])) * (nres + nleft)
    # left fields are untouched
    for j, f in enumerate(left):
        assert np.array_equal(cvec(f), W[nres + j])


def test_vector_multiply_matrix_rejects_bad_shapes(ctx):
    res = fields(ctx, 2, seed=1)
    f32 = fields(ctx, 1, prec=clg.F32, seed=1)
    with pytest.raises(clg.CLGError):
        clg.vector_multiply_matrix(ctx, res, f32, np.ones((3, 2)))
    with pytest.raises(clg.CLGError):
        clg.vector_multiply_matrix(ctx, res, [res[0]], np.ones((3, 2)))


def test_autotune_controls(ctx):
    dims = (8, 8, 8, 16)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.5, 3)
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    x.gaussian(4)
    y0 = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    clg.autotune_reset()
    assert clg.lib().clg_autotune_count() == 0
    clg.autotune_enable(1)
    clg.wilson_apply(ctx, y0, x, g, 0.12, clg.EFO_F_D)
    ref = cvec(y0)
    ent = clg.autotune_entries()
    assert len(ent) >= 1 and any(e.startswith("wilson_hop") and "dims=8x8x8x16" in e for e in ent)
    for pm, blk, xcd, spt in [(1, 64, 0, 1), (2, 128, 8, 1), (1, 256, 8, 1), (3, 128, 8, 1), (4, 256, 0, 1),
                              (3, 128, 1, 1), (4, 64, 1, 1), (1, 256, 1, 1), (5, 64, 8, 1), (6, 64, 1, 1),
                              (2, 128, 8, 2), (4, 256, 8, 4), (1, 64, 0, 4), (6, 64, 8, 2), (8, 64, 8, 4)]:
        clg.autotune_force(pm, blk, xcd, spt)
        y = clg.Field(ctx, clg.WILSON, clg.F64, dims)
        clg.wilson_apply(ctx, y, x, g, 0.12, clg.EFO_F_D)
        assert np.max(np.abs(cvec(y) - ref)) <= 1e-13 * np.max(np.abs(ref))
    clg.autotune_force(0)
    with pytest.raises(clg.CLGError):
        clg.autotune_force(9, 256, 8)
    with pytest.raises(clg.CLGError):
        clg.autotune_enable(2)
    clg.autotune_enable(-1)
    clg.autotune_reset()
    assert clg.lib().clg_autotune_count() == 0


def test_graph_replays_match_eager(ctx):
    """A recorded sequence (Wilson D, axpy, scale, zero, hop) replayed several times, with eager
    work and synchronisations between replays, equals the eager result every time."""
    dims = (8, 8, 8, 16)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.4, 8)
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    y = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    t = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    e_in = clg.Field(ctx, clg.WILSON, clg.F64, dims, subset=clg.EVEN)
    o_out = clg.Field(ctx, clg.WILSON, clg.F64, dims, subset=clg.ODD)

    def step():
        clg.wilson_apply(ctx, t, x, g, 0.13, clg.EFO_F_D)  # t = D x
        clg.axpy(ctx, y, 0.5 - 0.25j, t)                   # y += c t
        clg.scale(ctx, y, 0.9)                             # y *= 0.9
        t.zero()                                           # zero fill kernel
        clg.wilson_hop(ctx, o_out, e_in, g, 0, 1.0, None, 0.0)

    # eager reference over three steps
    x.gaussian(21)
    y.gaussian(22)
    e_in.gaussian(23)
    refs = []
    for _ in range(3):
        step()
        ctx.sync()
        refs.append((cvec(y), cvec(t), cvec(o_out)))
    # record one step
    x.gaussian(21)
    y.gaussian(22)
    e_in.gaussian(23)
    ctx.sync()
    clg.Graph.begin(ctx)
    step()
    with pytest.raises(clg.CLGError):
        clg.dot(ctx, x, y)  # blocking calls are refused while recording
    graph = clg.Graph.end(ctx)
    assert graph.nodes >= 5
    for k in range(3):
        graph.launch()
        ctx.sync()
        yv, tv, ov = cvec(y), cvec(t), cvec(o_out)
        assert np.max(np.abs(yv - refs[k][0])) <= 1e-13 * np.max(np.abs(refs[k][0])), k
        assert np.max(np.abs(tv)) == 0.0
        assert np.max(np.abs(ov - refs[k][2])) <= 1e-13 * np.max(np.abs(refs[k][2])), k
        # eager work between replays
        _ = clg.norm2(ctx, x)
        clg.wilson_apply(ctx, t, x, g, 0.2, clg.EFO_F_Ddagger)
    graph.close()


def test_autotune_cache_file(ctx, tmp_path, monkeypatch):
    """CLG_TUNE_CACHE: every tuned shape is appended to the file (keyed by the device architecture),
    and a fresh tuner table takes its configurations from the file instead of timing again."""
    path = tmp_path / "tune.txt"
    monkeypatch.setenv("CLG_TUNE_CACHE", str(path))
    dims = (8, 8, 8, 16)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.5, 3)
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    x.gaussian(4)
    k = clg.Field(ctx, clg.KS, clg.F64, dims)
    k.gaussian(5)
    try:
        clg.autotune_reset()
        clg.autotune_enable(1)
        y0 = clg.Field(ctx, clg.WILSON, clg.F64, dims)
        clg.wilson_apply(ctx, y0, x, g, 0.12, clg.EFO_F_D)
        k0 = clg.Field(ctx, clg.KS, clg.F64, dims)
        clg.ks_apply(ctx, k0, k, g, 0.05, clg.EFO_F_D)
        lines = path.read_text().splitlines()
        arch = ctx.info().split()[0].split(":")[0]
        assert any(" wilson_hop " in ln and " 8 8 8 16 :" in ln for ln in lines), lines
        assert any(" staggered_d " in ln for ln in lines), lines
        assert all(ln.startswith(arch) for ln in lines), (arch, lines)
        # replace the tuned configurations; a fresh table loads these instead of tuning
        new = []
        for ln in lines:
            head = ln.split(" :")[0]
            if " wilson_hop " in ln:
                ln = head + " : 1 64 0 2"
            elif " staggered_d " in ln:
                ln = head + " : 64 0 2"
            new.append(ln)
        path.write_text("\n".join(new) + "\nnot a cache line\n")
        clg.autotune_reset()
        y1 = clg.Field(ctx, clg.WILSON, clg.F64, dims)
        clg.wilson_apply(ctx, y1, x, g, 0.12, clg.EFO_F_D)
        k1 = clg.Field(ctx, clg.KS, clg.F64, dims)
        clg.ks_apply(ctx, k1, k, g, 0.05, clg.EFO_F_D)
        ent = clg.autotune_entries()
        assert any(e.startswith("wilson_hop") and "pin=1 block=64 xcd=0 spt=2" in e for e in ent), ent
        assert any(e.startswith("staggered_d") and "block=64 xcd=0 pin=2" in e for e in ent), ent
        assert len(path.read_text().splitlines()) == len(new) + 1   # nothing was tuned again
        assert np.max(np.abs(cvec(y1) - cvec(y0))) <= 1e-13 * np.max(np.abs(cvec(y0)))
        assert np.max(np.abs(k1.download() - k0.download())) <= 1e-13 * np.max(np.abs(k0.download()))
    finally:
        clg.autotune_enable(-1)
        clg.autotune_reset()
