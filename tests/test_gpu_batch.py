"""Batched Wilson hops (several right-hand sides and lattices per launch) and the batched solver,
against the single-system operators and solver (themselves checked against the oracle) and
against the oracle's D."""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

BC = orc.BC_FERMION
KAPPA = 0.1355
# three systems on one 8^4 lattice, one on 4^4 and one on a 6x4x4x6 box
SHAPES = [(8, 8, 8, 8)] * 3 + [(4, 4, 4, 4), (6, 4, 4, 6)]


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def gauges(ctx):
    out = {}
    for dims in sorted(set(SHAPES)):
        g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g.gauge_random(0.4, 1)
        out[dims] = g
    return out


def field(ctx, dims, prec, subset=clg.FULL, seed=None):
    f = clg.Field(ctx, clg.WILSON, prec, dims, subset=subset, bc=BC)
    if seed is not None:
        if prec == clg.F64 and subset == clg.FULL:
            f.gaussian(seed)
        else:
            t = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
            t.gaussian(seed)
            if subset != clg.FULL:
                t = t.parity(subset)
            f.copy_from(t)
    return f


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def gauge_in(ctx, gauges, dims, prec, recon12=False):
    g64 = gauges[dims]
    if prec == clg.F64:
        return g64
    kind = clg.GAUGE12 if recon12 else clg.GAUGE
    g = clg.Field(ctx, kind, clg.F16 if recon12 else clg.F32, dims)
    g.copy_from(g64)
    return g


@pytest.mark.parametrize("prec,links,tile", [
    (clg.F64, "f64", 1), (clg.F64, "f64", 2),
    (clg.F32, "f32", 1), (clg.F32, "f32", 2), (clg.F32, "f32", 4), (clg.F32, "f16r12", 4), (clg.F32, "f32", 0),
    (clg.F16, "f16r12", 1), (clg.F16, "f16r12", 2), (clg.F16, "f16r12", 4), (clg.F16, "f16r12", 0),
])
@pytest.mark.parametrize("dagger", [False, True])
def test_batched_half_hop_matches_single(ctx, gauges, prec, links, tile, dagger):
    gs = {d: gauge_in(ctx, gauges, d, prec, links == "f16r12") for d in set(SHAPES)}
    ins = [field(ctx, d, prec, clg.ODD, seed=10 + i) for i, d in enumerate(SHAPES)]
    xs = [field(ctx, d, prec, clg.EVEN, seed=30 + i) for i, d in enumerate(SHAPES)]
    outs = [field(ctx, d, prec, clg.EVEN) for d in SHAPES]
    a = [0.3 + 0.1 * i for i in range(len(SHAPES))]
    clg.wilson_hop_batch(ctx, outs, ins, [gs[d] for d in SHAPES], a, dagger=dagger, xs=xs, b=-0.7, tile=tile)
    # FP16 storage: both paths compute in FP32 and round the outputs the same way
    tol = 1e-14 if prec == clg.F64 else 2e-6 if prec == clg.F32 else 2e-3
    for i, d in enumerate(SHAPES):
        ref = field(ctx, d, prec, clg.EVEN)
        clg.wilson_hop(ctx, ref, ins[i], gs[d], dagger=dagger, a=a[i], x=xs[i], b=-0.7)
        assert rel(outs[i].download(), ref.download()) < tol, (i, d)


@pytest.mark.parametrize("tile", [1, 2])
def test_batched_full_hop_is_wilson_d(ctx, gauges, tile):
    # out = -kappa H psi + psi = D psi on FULL fields
    ins = [field(ctx, d, clg.F64, seed=50 + i) for i, d in enumerate(SHAPES)]
    outs = [field(ctx, d, clg.F64) for d in SHAPES]
    clg.wilson_hop_batch(ctx, outs, ins, [gauges[d] for d in SHAPES], -KAPPA, xs=ins, b=1.0, tile=tile)
    for i, d in enumerate(SHAPES):
        U = np.zeros(orc.volume(d) * 72)
        U = gauges[d].download()
        want = orc.wilson_op(d, BC, KAPPA, U, ins[i].download(), orc.D)
        assert rel(outs[i].download(), want) < 1e-12, (i, d)


@pytest.mark.parametrize("inner", [clg.F32, clg.F64, clg.F16, clg.INNER_F16_VECTORS])
def test_solve_batch_matches_single_solves(ctx, gauges, inner):
    bs, xs = [], []
    for i, d in enumerate(SHAPES):
        b = clg.Field(ctx, clg.WILSON, clg.F64, d, bc=BC)
        if i < 3:
            b.point_source((0, 0, 0, 0), i % 4, i % 3)
        else:
            b.gaussian(70 + i)
        bs.append(b)
        xs.append(clg.Field(ctx, clg.WILSON, clg.F64, d, bc=BC))
    p = clg.solver_params(max_dim=12, restart=100, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                          inner_prec=inner, inner_accuracy={clg.F64: 1e-11, clg.INNER_F16_VECTORS: 1e-2}.get(inner, 1e-4),
                          max_outer=40)
    st, reps = clg.solve_batch(ctx, p, xs, bs, [gauges[d] for d in SHAPES], KAPPA)
    print([(r.cycles, r.outer, r.true_residual) for r in reps])
    assert st == clg.OK
    for i, d in enumerate(SHAPES):
        assert reps[i].converged and reps[i].true_residual < 1e-10
        ref = clg.Field(ctx, clg.WILSON, clg.F64, d, bc=BC)
        st1, rep1 = clg.solve(ctx, p, ref, bs[i], gauges[d], KAPPA, clg.EFO_F_InverseD)
        assert st1 == clg.OK
        assert rel(xs[i].download(), ref.download()) < 1e-8, i
        # and the oracle's operator maps the solution back to b
        x = xs[i].download()
        Dx = orc.wilson_op(d, BC, KAPPA, gauges[d].download(), x, orc.D)
        bb = bs[i].download()
        assert np.linalg.norm(Dx - bb) < 2e-10 * np.linalg.norm(bb)


def test_solve_batch_dagger_and_errors(ctx, gauges):
    d = (4, 4, 4, 4)
    b = clg.Field(ctx, clg.WILSON, clg.F64, d, bc=BC)
    b.gaussian(5)
    x = clg.Field(ctx, clg.WILSON, clg.F64, d, bc=BC)
    p = clg.solver_params(max_dim=10, restart=100, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                          inner_prec=clg.F32, inner_accuracy=1e-4)
    st, reps = clg.solve_batch(ctx, p, [x], [b], [gauges[d]], KAPPA, op=clg.EFO_F_InverseDdagger)
    assert st == clg.OK
    Dx = orc.wilson_op(d, BC, KAPPA, gauges[d].download(), x.download(), orc.DDAG)
    assert np.linalg.norm(Dx - b.download()) < 2e-10 * np.linalg.norm(b.download())
    with pytest.raises(clg.CLGError):
        clg.solve_batch(ctx, clg.solver_params(max_dim=40, stop_rule=clg.STOP_TRUE_RESIDUAL), [x], [b],
                        [gauges[d]], KAPPA)
    with pytest.raises(clg.CLGError):
        clg.solve_batch(ctx, p, [x], [x], [gauges[d]], KAPPA)


def test_batch_rejects_cross_slot_aliasing(ctx, gauges):
    """Every workgroup of a batched hop reads the inputs and x terms of all slots while others write the
    outputs: storage shared across slots (an output that is another slot's input, output or x, or a
    parity view of one) is refused before any launch, and so are overlapping x or x/b fields of a
    batched solve.  An output sharing only its own slot's x stays allowed (same site, same thread)."""
    dims = (4, 4, 4, 4)
    g = gauges[dims]
    a, b, c = (field(ctx, dims, clg.F64, seed=s) for s in (1, 2, 3))
    o1, o2 = field(ctx, dims, clg.F64), field(ctx, dims, clg.F64)
    bad = [
        dict(outs=[o1, a], ins=[a, b]),              # out of slot 1 is the input of slot 0
        dict(outs=[o1, o1], ins=[a, b]),             # two slots write one field
        dict(outs=[o1, o2], ins=[a, b], xs=[o2, b]),  # x of slot 0 is the output of slot 1
    ]
    for kw in bad:
        with pytest.raises(clg.CLGError):
            clg.wilson_hop_batch(ctx, kw["outs"], kw["ins"], [g, g], -KAPPA, xs=kw.get("xs"), b=1.0 if "xs" in kw
                                 else 0.0)
    # half-lattice hops: the output of slot 1 is a parity view of the field whose same parity view is the
    # x term of slot 0 (two view objects, one storage)
    oe = clg.Field(ctx, clg.WILSON, clg.F64, dims, subset=clg.EVEN, bc=BC)
    d = field(ctx, dims, clg.F64, seed=4)
    with pytest.raises(clg.CLGError):
        clg.wilson_hop_batch(ctx, [oe, c.parity(clg.EVEN)], [a.parity(clg.ODD), b.parity(clg.ODD)], [g, g], -KAPPA,
                             xs=[c.parity(clg.EVEN), d.parity(clg.EVEN)], b=1.0)
    # while the even view of one field and the odd view of the same field share no storage
    clg.wilson_hop_batch(ctx, [oe, a.parity(clg.EVEN)], [c.parity(clg.ODD), b.parity(clg.ODD)], [g, g], -KAPPA)
    # allowed: out of each slot is its own x
    want = []
    for o, i in ((o1, a), (o2, b)):
        o.copy_from(c)
        t = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
        clg.wilson_hop_batch(ctx, [t], [i], [g], -KAPPA, xs=[c], b=1.0)
        want.append(t.download())
    clg.wilson_hop_batch(ctx, [o1, o2], [a, b], [g, g], -KAPPA, xs=[o1, o2], b=1.0)
    assert rel(o1.download(), want[0]) < 1e-15 and rel(o2.download(), want[1]) < 1e-15
    # batched solve: x shared between systems, or x of one system is b of another
    p = clg.solver_params(max_dim=8, restart=20, accuracy=1e-8, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    for xs, bs in (([o1, o1], [a, b]), ([o1, o2], [a, o1])):
        with pytest.raises(clg.CLGError):
            clg.solve_batch(ctx, p, xs, bs, [g, g], KAPPA)
