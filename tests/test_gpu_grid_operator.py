"""GridOperator::apply for particle operations (Picasso_GridOperator.hpp:
459-496) through pic_grid_operator_apply: gather halos, zeroed scatter
fields, the caller's transfers, scatter-sum halos.  The mesh and fields of
the reference's own test (tstGridOperator3d.hpp:318-429: 43 x 32 x 39 cells
of 0.23 from (1.2, 3.3, -2.8), periodic in x and z, two ghost cells, gather
fields FooIn = 2 (vector) and BarIn = 3, scatter fields pre-filled with
wrong data, 10 particles per cell)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

NCELL = np.array([43, 32, 39])
H = 0.23
HALO = 2
LOW = np.array([1.2, 3.3, -2.8]) - HALO * H   # local mesh low corner, ghosts included
PERIODIC = (True, False, True)
WIN = NCELL + 2 * HALO                          # Cell window [0, WIN)
CELL = 1                                        # FieldLocation::Cell
VALUE = 0


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _fields(pa, gu):
    lo = (0, 0, 0)
    shp = tuple(int(v) for v in WIN)
    dep = dict(own_lo=(HALO,) * 3, own_n=tuple(int(v) for v in NCELL), periodic=PERIODIC)
    foo_in = pa.FieldDep(pa.GridField(gu.dev(np.full(shp + (3,), 2.0)), lo), **dep)
    bar_in = pa.FieldDep(pa.GridField(gu.dev(np.full(shp + (1,), 3.0)), lo), **dep)
    foo_out = pa.FieldDep(pa.GridField(gu.dev(np.full(shp + (3,), -1.1)), lo), **dep)
    bar_out = pa.FieldDep(pa.GridField(gu.dev(np.full(shp + (1,), -2.2)), lo), **dep)
    return foo_in, bar_in, foo_out, bar_out


def _body(pa, sd, foo_in, bar_in, foo_out, bar_out, foo_p, bar_p):
    def body(ctx):
        pa.G2P.value(ctx, sd, foo_in.field, foo_p)
        pa.G2P.value(ctx, sd, bar_in.field, bar_p)
        pa.P2G.value(ctx, sd, foo_p, foo_out.field)
        pa.P2G.value(ctx, sd, bar_p, bar_out.field)
    return body


def _owned(a):
    s = tuple(slice(HALO, HALO + int(n)) for n in NCELL)
    return a[s]


@pytest.mark.parametrize("binned", [False, True])
def test_reference_grid_operator_particle_apply(pa, binned):
    """The reference's particle operation: G2P of the gather fields to the
    particles, P2G of them to the scatter fields.  Particles sit at cell
    centres (ppc per owned cell) with linear splines on cells, so each
    weighs 1 on its own cell: every particle gets (2, 2, 2) and 3, and every
    owned cell of the scatter fields ppc x those values -- the wrong data
    the scatter fields held is gone."""
    import gpu_util as gu
    ppc = 10
    ii = [np.arange(HALO, HALO + n) for n in NCELL]
    cells = np.stack(np.meshgrid(*ii, indexing="ij"), -1).reshape(-1, 3)
    x = np.repeat(LOW + H * (cells + 0.5), ppc, axis=0)
    P = len(x)
    mesh = pa.Mesh(LOW, H)
    xd = gu.soa(x)
    bins = None
    if binned:
        dual = pa.aligned_dual(1, CELL)  # node-centred bins: one more per dimension
        bins = pa.bin_sort(gu.ctx(), mesh, xd,
                           pa.Binning((HALO,) * 3,
                                      tuple(int(v) + f for v, f in zip(NCELL, pa.dual_dims(dual))),
                                      (4, 4, 4), dual))
    sd = pa.create_spline(CELL, 1, mesh, xd, bins)
    foo_in, bar_in, foo_out, bar_out = _fields(pa, gu)
    foo_p = gu.dev(np.zeros((P, 3)))
    bar_p = gu.dev(np.zeros((P, 1)))
    op = pa.GridOperator(gather=[foo_in, bar_in], scatter=[foo_out, bar_out])
    op.apply(gu.ctx(), _body(pa, sd, foo_in, bar_in, foo_out, bar_out, foo_p, bar_p))
    gu.ctx().synchronize()
    np.testing.assert_allclose(gu.host(foo_p), 2.0, rtol=1e-14)
    np.testing.assert_allclose(gu.host(bar_p), 3.0, rtol=1e-14)
    np.testing.assert_allclose(_owned(gu.host(foo_out.field.tensor)), ppc * 2.0, rtol=1e-13)
    np.testing.assert_allclose(_owned(gu.host(bar_out.field.tensor)), ppc * 3.0, rtol=1e-13)


def test_grid_operator_halos_against_the_oracle(pa):
    """Random positions, random gather data on the owned cells and garbage
    in the ghosts: the gather halo must fill the periodic ghosts the linear
    stencils reach (x and z), and the scatter-sum must fold the ghosts' P2G
    deposits into their owned images.  Checked against the oracle's gather
    halo -> G2P and P2G -> scatter-sum halo.  Particles avoid the outer cells
    of the non-periodic y dimension, whose ghosts no halo fills."""
    import gpu_util as gu
    rng = np.random.default_rng(318)
    P = 40000
    lo_c = np.array([HALO, HALO + 1, HALO])
    hi_c = HALO + NCELL - np.array([0, 1, 0])
    x = LOW + H * rng.uniform(lo_c, hi_c, size=(P, 3))
    shp = tuple(int(v) for v in WIN)
    f0 = rng.normal(size=shp + (3,))
    g0 = np.full(shp + (3,), -7.5)
    own = tuple(slice(HALO, HALO + int(n)) for n in NCELL)
    g0[own] = f0[own]
    lo = (0, 0, 0)
    dep = dict(own_lo=(HALO,) * 3, own_n=tuple(int(v) for v in NCELL), periodic=PERIODIC)
    fin = pa.FieldDep(pa.GridField(gu.dev(g0), lo), **dep)
    fout = pa.FieldDep(pa.GridField(gu.dev(np.full(shp + (3,), 9.0)), lo), **dep)
    mesh = pa.Mesh(LOW, H)
    xd = gu.soa(x)
    sd = pa.create_spline(CELL, 1, mesh, xd)
    vp = gu.dev(np.zeros((P, 3)))

    def body(ctx):
        pa.G2P.value(ctx, sd, fin.field, vp)
        pa.P2G.value(ctx, sd, vp, fout.field)

    pa.GridOperator(gather=[fin], scatter=[fout]).apply(gu.ctx(), body)
    gu.ctx().synchronize()
    win_lo, win_n = np.zeros(3, np.int32), WIN.astype(np.int32)
    per = tuple(int(v) for v in PERIODIC)
    gref = orpic.grid_halo(0, win_lo, g0.copy(), dep["own_lo"], dep["own_n"], per)
    vref = orpic.g2p_interp(VALUE, 1, CELL, LOW, H, win_lo, win_n, x, gref, 3)
    assert gu.rel_inf(gu.host(vp), vref) < 1e-12
    fref = orpic.p2g_interp(VALUE, 1, CELL, LOW, H, win_lo, win_n, x, vref, 3)
    fref = orpic.grid_halo(1, win_lo, np.ascontiguousarray(fref), dep["own_lo"], dep["own_n"], per)
    got = gu.host(fout.field.tensor)
    assert gu.rel_inf(got[own], fref[own]) < 1e-12
    # nothing deposited is lost: the owned cells hold every particle's value
    np.testing.assert_allclose(got[own].reshape(-1, 3).sum(0), gu.host(vp).sum(0), rtol=1e-12)


def test_grid_operator_body_error_stops_the_apply(pa):
    """An exception in the operation stops the apply before the scatter
    halo and reaches the caller."""
    import gpu_util as gu
    foo_in, bar_in, foo_out, bar_out = _fields(pa, gu)

    def body(ctx):
        raise RuntimeError("operation failed")

    with pytest.raises(RuntimeError, match="operation failed"):
        pa.GridOperator(gather=[foo_in], scatter=[foo_out]).apply(gu.ctx(), body)
    gu.ctx().synchronize()
    # the scatter field was zeroed before the operation ran
    assert gu.host(foo_out.field.tensor).sum() == 0.0
