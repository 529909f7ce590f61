"""Generic transfers, Picasso::G2P / P2G value, gradient and divergence
(Picasso_ParticleInterpolation.hpp:110-279), against the oracle: every
operation and component count, spline orders 1-3, collocated and staggered
lattices, in input order and in sorted order (bins).  Bars as elsewhere:
FP64 within 1e-12 and FP32 within 1e-5 of the oracle on the FP32-rounded
inputs (||.||_inf / ||ref||_inf)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

VALUE, GRADIENT, DIVERGENCE = 0, 1, 2
G2P_CASES = [(VALUE, 1), (VALUE, 2), (VALUE, 3), (GRADIENT, 1), (GRADIENT, 3), (DIVERGENCE, 3)]
P2G_CASES = [(VALUE, 1), (VALUE, 3), (GRADIENT, 1), (DIVERGENCE, 3), (DIVERGENCE, 9)]
H = 0.25
LOW = np.array([-1.0, -0.5, -0.75])
LO = np.array([1, 2, 0], np.int32)
N = np.array([16, 15, 17], np.int32)


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _g2p_nout(op, nc):
    return nc if op == VALUE else (3 * nc if op == GRADIENT else 1)


def _p2g_ngrid(op, nc):
    return nc if op == VALUE else (3 if op == GRADIENT else nc // 3)


def _round(a, dtype):
    return a.astype(np.float32).astype(np.float64) if dtype == "f32" else a


def _spline(pa, gu, entity, order, x, tdt, sorted_order):
    mesh = pa.Mesh(LOW, H)
    xd = gu.soa(x, tdt)
    bins = None
    if sorted_order:
        bins = pa.bin_sort(gu.ctx(), mesh, xd, pa.Binning((4, 4, 4), (8, 8, 8), (2, 2, 2), 0))
    return pa.create_spline(entity, order, mesh, xd, bins)


@pytest.mark.parametrize("sorted_order", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("entity", [0, 1, 2, 7])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_g2p_interp_vs_oracle(pa, order, entity, dtype, sorted_order):
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    rng = np.random.default_rng(1000 * order + 10 * entity + (dtype == "f32"))
    x = _round(LOW + H * rng.uniform(4.05, 11.95, size=(400, 3)), dtype)
    sd = _spline(pa, gu, entity, order, x, tdt, sorted_order)
    for op, nc in G2P_CASES:
        f = _round(rng.normal(size=tuple(int(k) for k in N) + (nc,)), dtype)
        ref = orpic.g2p_interp(op, order, entity, LOW, H, LO, N, x, f, nc)
        field = pa.GridField(gu.dev(f, tdt), tuple(int(k) for k in LO))
        out = torch.zeros((len(x), _g2p_nout(op, nc)), dtype=tdt, device=gu.DEV)
        if op == VALUE:
            pa.G2P.value(gu.ctx(), sd, field, out)
        elif op == GRADIENT:
            pa.G2P.gradient(gu.ctx(), sd, field, out, ncomp=nc)
        else:
            pa.G2P.divergence(gu.ctx(), sd, field, out)
        gu.ctx().synchronize()
        assert gu.rel_inf(gu.host(out), ref) < tol, (op, nc)


@pytest.mark.parametrize("sorted_order", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("entity", [0, 1, 2, 7])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_p2g_interp_vs_oracle(pa, order, entity, dtype, sorted_order):
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    rng = np.random.default_rng(2000 * order + 10 * entity + (dtype == "f32"))
    x = _round(LOW + H * rng.uniform(4.05, 11.95, size=(400, 3)), dtype)
    sd = _spline(pa, gu, entity, order, x, tdt, sorted_order)
    shape = tuple(int(k) for k in N)
    for op, nc in P2G_CASES:
        v = _round(rng.normal(size=(len(x), nc)), dtype)
        ref = orpic.p2g_interp(op, order, entity, LOW, H, LO, N, x, v, nc)
        field = pa.GridField(torch.zeros(shape + (_p2g_ngrid(op, nc),), dtype=tdt,
                                         device=gu.DEV), tuple(int(k) for k in LO))
        vd = gu.soa(v, tdt)
        if op == VALUE:
            pa.P2G.value(gu.ctx(), sd, vd, field)
        elif op == GRADIENT:
            pa.P2G.gradient(gu.ctx(), sd, vd, field)
        else:
            pa.P2G.divergence(gu.ctx(), sd, vd, field)
        gu.ctx().synchronize()
        assert gu.rel_inf(gu.host(field.tensor), ref) < tol, (op, nc)


def test_p2g_interp_zero_first_and_accumulate(pa):
    """P2G adds into the field (ScatterView semantics); PIC_P2G_ZERO_FIRST
    clears it first."""
    import gpu_util as gu
    rng = np.random.default_rng(5)
    x = LOW + H * rng.uniform(4, 12, size=(300, 3))
    v = rng.normal(size=(300, 3))
    ref = orpic.p2g_interp(VALUE, 2, 0, LOW, H, LO, N, x, v, 3)
    sd = _spline(pa, gu, 0, 2, x, torch.float64, False)
    field = pa.GridField(gu.dev(np.ones(tuple(int(k) for k in N) + (3,))),
                         tuple(int(k) for k in LO))
    pa.P2G.value(gu.ctx(), sd, gu.soa(v), field)
    gu.ctx().synchronize()
    assert gu.rel_inf(gu.host(field.tensor), ref + 1.0) < 1e-12
    pa.P2G.value(gu.ctx(), sd, gu.soa(v), field, zero_first=True)
    gu.ctx().synchronize()
    assert gu.rel_inf(gu.host(field.tensor), ref) < 1e-12


@pytest.mark.parametrize("binned", [False, True])
def test_grid_operator_reset_then_accumulate(pa, binned):
    """GridOperator's scatter semantics as tstGridOperator3d.hpp:390-429 pins
    them: scatter fields holding wrong data (-1.1, -2.2) are reset to zero
    and then hold ppc * value per cell; gather fields (2.0, 3.0) reach every
    particle unchanged.  Particles sit at cell centres (ppc copies each) on
    the Cell lattice with linear splines, so every weight is exactly 1."""
    import gpu_util as gu
    ppc = 3
    cells = np.stack(np.meshgrid(*[np.arange(4, 12)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = np.repeat(LOW + H * (cells + 0.5), ppc, axis=0)
    P = len(x)
    mesh = pa.Mesh(LOW, H)
    xd = gu.soa(x)
    bins = None
    if binned:
        bins = pa.bin_sort(gu.ctx(), mesh, xd,
                           pa.Binning((3, 3, 3), (12, 12, 12), (4, 4, 4),
                                      pa.aligned_dual(1, pa.FieldLocation.Cell)))
    sd = pa.create_spline(pa.FieldLocation.Cell, 1, mesh, xd, bins)
    shp = tuple(int(k) for k in N)
    lo = tuple(int(k) for k in LO)
    foo_out = pa.GridField(gu.dev(np.full(shp + (3,), -1.1)), lo)
    bar_out = pa.GridField(gu.dev(np.full(shp + (1,), -2.2)), lo)
    pa.P2G.value(gu.ctx(), sd, gu.soa(np.full((P, 3), 2.0)), foo_out, zero_first=True)
    pa.P2G.value(gu.ctx(), sd, gu.soa(np.full((P, 1), 3.0)), bar_out, zero_first=True)
    foo_p = gu.dev(np.zeros((P, 3)))
    bar_p = gu.dev(np.zeros((P, 1)))
    pa.G2P.value(gu.ctx(), sd, pa.GridField(gu.dev(np.full(shp + (3,), 2.0)), lo), foo_p)
    pa.G2P.value(gu.ctx(), sd, pa.GridField(gu.dev(np.full(shp + (1,), 3.0)), lo), bar_p)
    gu.ctx().synchronize()
    np.testing.assert_array_equal(gu.host(foo_p), 2.0)
    np.testing.assert_array_equal(gu.host(bar_p), 3.0)
    own = tuple(slice(4 - LO[d], 12 - LO[d]) for d in range(3))
    foo = gu.host(foo_out.tensor)
    bar = gu.host(bar_out.tensor)
    np.testing.assert_array_equal(foo[own], ppc * 2.0)
    np.testing.assert_array_equal(bar[own], ppc * 3.0)
    mask = np.ones(shp, bool)
    mask[own] = False
    assert np.all(foo[mask] == 0.0) and np.all(bar[mask] == 0.0)


@pytest.mark.parametrize("binned", [False, True])
def test_reference_interpolation_test_periodic_box(pa, binned):
    """tstParticleInterpolation.hpp:351-600, on this library's pieces: a
    periodic 18 x 22 x 39 box of 0.05 cells with one ghost layer
    (particle_interpolation_test.json), one particle at every cell centre
    (InitUniform, ppc 1), linear node splines.  Constant particle values
    3.5 give 3.5 at every owned node after P2G and the scatter-sum halo, and
    zero gradients and divergences; a constant grid 3.5 (halo gathered)
    gives 3.5 at every particle and zero gradients and divergences."""
    import gpu_util as gu
    h = 0.05
    nc = np.array([18, 22, 39])
    low = np.array([-1.2, 0.1, 1.1]) - h  # one ghost cell
    mesh = pa.Mesh(low, h)
    ctx = gu.ctx()
    x, _ = pa.initialize_particles(ctx, pa.InitUniform, 1, mesh, (1, 1, 1), nc,
                                   dtype=torch.float64)
    P = x.shape[0]
    assert P == int(nc.prod())
    bins = None
    if binned:
        bins = pa.bin_sort(ctx, mesh, x, pa.Binning((1, 1, 1), tuple(int(v) for v in nc),
                                                    (4, 4, 4), 0))
    sd = pa.create_spline(pa.FieldLocation.Node, 1, mesh, x, bins)
    shp = tuple(int(v) + 2 for v in nc)  # nodes 0 .. nc + 1: owned 1 .. nc, one ghost each side
    own_lo, own_n = (1, 1, 1), tuple(int(v) for v in nc)
    own = (slice(1, 1 + nc[0]), slice(1, 1 + nc[1]), slice(1, 1 + nc[2]))

    def node_field(ncomp, value=0.0):
        return pa.GridField(gu.dev(np.full(shp + (ncomp,), value)))

    def p2g(op, pv, ncomp_field):
        f = node_field(ncomp_field)
        getattr(pa.P2G, op)(ctx, sd, gu.dev(pv), f, zero_first=True)
        pa.halo_scatter(ctx, f, own_lo, own_n)
        ctx.synchronize()
        return gu.host(f.tensor)[own]

    np.testing.assert_allclose(p2g("value", np.full((P, 1), 3.5), 1), 3.5, rtol=1e-12)
    np.testing.assert_allclose(p2g("value", np.full((P, 3), 3.5), 3), 3.5, rtol=1e-12)
    np.testing.assert_allclose(p2g("gradient", np.full((P, 1), 3.5), 3), 0.0, atol=1e-12)
    np.testing.assert_allclose(p2g("divergence", np.full((P, 3), 3.5), 1), 0.0, atol=1e-12)
    np.testing.assert_allclose(p2g("divergence", np.full((P, 9), 3.5), 3), 0.0, atol=1e-12)

    def g2p(op, ncomp_field, nout, ncomp):
        f = node_field(ncomp_field, 3.5)
        pa.halo_gather(ctx, f, own_lo, own_n)
        out = gu.dev(np.zeros((P, nout)))
        if op == "divergence":
            pa.G2P.divergence(ctx, sd, f, out)
        else:
            getattr(pa.G2P, op)(ctx, sd, f, out, ncomp=ncomp)
        ctx.synchronize()
        return gu.host(out)

    np.testing.assert_allclose(g2p("value", 1, 1, 1), 3.5, rtol=1e-12)
    np.testing.assert_allclose(g2p("value", 3, 3, 3), 3.5, rtol=1e-12)
    np.testing.assert_allclose(g2p("gradient", 1, 3, 1), 0.0, atol=1e-12)
    np.testing.assert_allclose(g2p("gradient", 3, 9, 3), 0.0, atol=1e-12)
    np.testing.assert_allclose(g2p("divergence", 3, 1, 3), 0.0, atol=1e-12)


def test_interp_rejects_unsupported_shapes(pa):
    import gpu_util as gu
    x = LOW + H * np.full((4, 3), 6.5)
    sd = _spline(pa, gu, 0, 2, x, torch.float64, False)
    f2 = pa.GridField(gu.dev(np.zeros(tuple(int(k) for k in N) + (2,))), tuple(int(k) for k in LO))
    with pytest.raises(pa.PicError):
        pa.G2P.gradient(gu.ctx(), sd, f2, gu.dev(np.zeros((4, 6))), ncomp=2)
    f3 = pa.GridField(gu.dev(np.zeros(tuple(int(k) for k in N) + (3,))), tuple(int(k) for k in LO))
    with pytest.raises(pa.PicError):
        pa.P2G.divergence(gu.ctx(), sd, gu.dev(np.zeros((4, 2))), f3)


def test_interp_stencil_outside_window_reported(pa):
    """A particle whose stencil leaves the window is skipped and reported."""
    import gpu_util as gu
    x = np.array([LOW + H * np.array([0.5, 6.5, 6.5])])  # i stencil below lo = 1
    sd = _spline(pa, gu, 0, 2, x, torch.float64, False)
    field = pa.GridField(gu.dev(np.zeros(tuple(int(k) for k in N) + (1,))),
                         tuple(int(k) for k in LO))
    pa.G2P.value(gu.ctx(), sd, field, gu.dev(np.zeros((1, 1))))
    with pytest.raises(pa.PicError):
        gu.ctx().synchronize()


def _aligned_spline(pa, gu, entity, order, x, tdt, x_bins=None):
    """Spline data with 4^3-brick bins on the knot lattice of (order,
    entity), so P2G runs on the brick kernel; bins are sorted from x_bins
    (default x), so moved particles exercise the stale list."""
    mesh = pa.Mesh(LOW, H)
    geom = pa.Binning((3, 3, 3), (12, 12, 12), (4, 4, 4), pa.aligned_dual(order, entity))
    xb = gu.soa(x if x_bins is None else x_bins, tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, xb, geom)
    return pa.create_spline(entity, order, mesh, gu.soa(x, tdt), bins)


@pytest.mark.parametrize("moved", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("entity", [0, 1, 3, 7])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_p2g_interp_brick_vs_oracle(pa, order, entity, dtype, moved):
    """Binned P2G on the brick kernel (and, with moved particles, its stale
    list) against the oracle, on node, cell and staggered (Face<J>,
    Edge<K>: bins with a dual flag per dimension) lattices."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-12 if dtype == "f64" else 1e-5
    rng = np.random.default_rng(3000 * order + 10 * entity + (dtype == "f32") + 5 * moved)
    x0 = _round(LOW + H * rng.uniform(4.05, 11.95, size=(3000, 3)), dtype)
    x = x0
    if moved:  # a third of the particles move up to half a cell after the sort
        x = x0.copy()
        sel = rng.random(len(x)) < 0.33
        x[sel] += H * rng.uniform(-0.5, 0.5, size=(int(sel.sum()), 3))
        x = _round(x, dtype)
    sd = _aligned_spline(pa, gu, entity, order, x, tdt, x0)
    shape = tuple(int(k) for k in N)
    for op, nc in P2G_CASES:
        v = _round(rng.normal(size=(len(x), nc)), dtype)
        ref = orpic.p2g_interp(op, order, entity, LOW, H, LO, N, x, v, nc)
        field = pa.GridField(torch.zeros(shape + (_p2g_ngrid(op, nc),), dtype=tdt,
                                         device=gu.DEV), tuple(int(k) for k in LO))
        vd = gu.soa(v, tdt)
        ctx = gu.ctx()
        ctx.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        try:
            if op == VALUE:
                pa.P2G.value(ctx, sd, vd, field)
            elif op == GRADIENT:
                pa.P2G.gradient(ctx, sd, vd, field)
            else:
                pa.P2G.divergence(ctx, sd, vd, field)
            ctx.synchronize()
            stale_runs = ctx.kernel_time("p2g_interp_stale")[1]
        finally:
            ctx.set_timing(False)
        assert stale_runs == 1  # the brick path ran (its stale-list pass is timed)
        assert gu.rel_inf(gu.host(field.tensor), ref) < tol, (op, nc)
