"""pic_apic_step (sort, p2g with the grid zeroed first, normalised g2p in one
call) against the FP64 oracle and against the same three calls made one by
one.  The step clears the grid windows inside the sort's key pass and lists
the brick walk from the sort's scan, so every test starts from grids filled
with garbage: a node the clearing missed would show.

Bars as in the other parity tests: FP64 within 1e-12 relative, FP32 within
1e-5 of the FP64 oracle on the FP32-rounded inputs (B within 1e-4), bins
bit-exact against pic_bin_sort."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 1e-5


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.fixture(scope="module")
def wl():
    from picasso_amd import synth
    return synth.apic_sphere(32, 8)


def _geom(pa, wl, order, entity, brick=(4, 4, 4)):
    dual = pa.aligned_dual(order, entity)
    flags = pa.dual_dims(dual)
    cell_lo, ncell = wl.cell_region
    return pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), brick, dual)


def _inputs(wl, f32):
    rnd = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    return rnd(wl.x), rnd(wl.m), rnd(wl.c)


def _oracle(wl, order, entity, x, m, c):
    lo, n = wl.node_window
    shp = tuple(int(v) for v in n)
    stag = entity not in (orpic.NODE, orpic.CELL)
    rm = np.zeros(shp)
    rmu = np.zeros(shp if stag else shp + (3,))
    orpic.apic_p2g(order, entity, wl.low, wl.h, lo, n, x, m, c, rm, rmu)
    if stag:
        u = np.where(rm > 0, rmu / np.where(rm > 0, rm, 1.0), 0.0)
    else:
        u = orpic.grid_normalize(rm, rmu)
    rc = orpic.apic_g2p(order, entity, wl.low, wl.h, lo, n, x, u)
    return rm, rmu, rc


def _grids(pa, wl, tdt, ncomp, pad=0):
    """Grid windows filled with garbage; pad > 0 gives the momentum grid a
    padded row (a window that is not one dense range)."""
    import gpu_util as gu
    lo, n = wl.node_window
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.full(shp, 7.0, dtype=tdt, device=gu.DEV), tuple(int(v) for v in lo))
    if pad:
        big = torch.full(shp[:2] + (shp[2] + pad,) + ((ncomp,) if ncomp > 1 else ()), -3.0,
                         dtype=tdt, device=gu.DEV)
        t = big[:, :, :shp[2]]
    else:
        t = torch.full(shp + ((ncomp,) if ncomp > 1 else ()), -3.0, dtype=tdt, device=gu.DEV)
    return gm, pa.GridField(t, tuple(int(v) for v in lo))


def _run_step(pa, wl, order, entity, dtype, brick=(4, 4, 4), pad=0, presort=False):
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    xh, mh, ch = _inputs(wl, dtype == "f32")
    mesh = pa.Mesh(wl.low, wl.h)
    ctx = gu.ctx()
    x = gu.soa(xh, tdt)
    m = gu.soa(mh[:, None], tdt)
    c = gu.soa(ch.reshape(-1, 12), tdt)
    geom = _geom(pa, wl, order, entity, brick)
    order_h = np.arange(len(xh))
    if presort:
        # particles stored in bin order: the sort takes its identity path
        b0 = pa.bin_sort(ctx, mesh, x, geom)
        ctx.synchronize()
        perm = b0.perm.long()
        x, m, c = (t[perm].t().contiguous().t() for t in (x, m, c))
        order_h = gu.host(b0.perm).astype(np.int64)
    stag = entity not in (pa.NODE, pa.CELL)
    gm, gmu = _grids(pa, wl, tdt, 1 if stag else 3, pad)
    sd = pa.create_spline(entity, order, mesh, x)
    bins = pa.APIC.step(ctx, m, c, gm, gmu, sd, geom)
    ctx.synchronize()
    ref_bins = pa.bin_sort(ctx, mesh, x, geom)
    ctx.synchronize()
    for name in ("perm", "keys", "bin_start"):
        np.testing.assert_array_equal(gu.host(getattr(bins, name)),
                                      gu.host(getattr(ref_bins, name)))
    cc = np.empty((len(xh), 12))
    cc[order_h] = gu.host(c)
    return gu.host(gm.tensor), gu.host(gmu.tensor), cc.reshape(-1, 4, 3), (xh, mh, ch)


def _check(wl, order, entity, got, f32):
    gm, gmu, cc, (xh, mh, ch) = got
    rm, rmu, rc = _oracle(wl, order, entity, xh, mh, ch)
    tol = F32_TOL if f32 else F64_TOL
    import gpu_util as gu
    assert gu.rel_inf(gm, rm) < tol
    assert gu.rel_inf(gmu, rmu) < tol
    assert gu.rel_inf(cc[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(cc[:, 1:], rc[:, 1:]) < (1e-4 if f32 else tol)


@pytest.mark.parametrize("presort", [False, True])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_step_f64_node(pa, wl, order, presort):
    """Shuffled storage (the sort's full path) and bin-ordered storage (its
    identity path), every spline order, FP64."""
    _check(wl, order, pa.NODE, _run_step(pa, wl, order, pa.NODE, "f64", presort=presort), False)


@pytest.mark.parametrize("presort", [False, True])
def test_step_f32_node(pa, wl, presort):
    _check(wl, 2, pa.NODE, _run_step(pa, wl, 2, pa.NODE, "f32", presort=presort), True)


@pytest.mark.parametrize("entity", [1, 2, 5])
def test_step_cell_and_staggered(pa, wl, entity):
    """Cell and staggered Face/Edge lattices through the step (one momentum
    component; g2p rewrites that column and zeroes the others)."""
    _check(wl, 2, entity, _run_step(pa, wl, 2, entity, "f64"), False)


def test_step_strided_momentum_window(pa, wl):
    """A momentum window that is not one dense range is cleared by a fill of
    its own; the results do not change."""
    _check(wl, 2, pa.NODE, _run_step(pa, wl, 2, pa.NODE, "f64", pad=5), False)


def test_step_other_bricks_take_three_calls(pa, wl):
    """8^3 bricks (no p2g brick kernel): the step runs the three calls."""
    _check(wl, 2, pa.NODE, _run_step(pa, wl, 2, pa.NODE, "f64", brick=(8, 8, 8)), False)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_step_matches_three_calls(pa, wl, dtype):
    """The step against bin_sort + p2g(zero_first) + g2p_normalized on the
    same inputs (atomic summation order aside)."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-13 if dtype == "f64" else 1e-6
    xh, mh, ch = _inputs(wl, dtype == "f32")
    mesh = pa.Mesh(wl.low, wl.h)
    ctx = gu.ctx()
    x = gu.soa(xh, tdt)
    m = gu.soa(mh[:, None], tdt)
    geom = _geom(pa, wl, 2, pa.NODE)
    out = []
    for one_call in (True, False):
        c = gu.soa(ch.reshape(-1, 12), tdt)
        gm, gmu = _grids(pa, wl, tdt, 3)
        sd = pa.create_spline(pa.NODE, 2, mesh, x)
        if one_call:
            pa.APIC.step(ctx, m, c, gm, gmu, sd, geom)
        else:
            sd.bins = pa.bin_sort(ctx, mesh, x, geom)
            pa.APIC.p2g(ctx, m, c, gm, gmu, sd, zero_first=True)
            pa.APIC.g2p_normalized(ctx, gm, gmu, c, sd)
        ctx.synchronize()
        out.append((gu.host(gm.tensor), gu.host(gmu.tensor), gu.host(c)))
    for a, b in zip(*out):
        assert gu.rel_inf(a, b) < tol


def test_step_empty_particle_set_zeroes_grids(pa, wl):
    import gpu_util as gu
    ctx = gu.ctx()
    mesh = pa.Mesh(wl.low, wl.h)
    x = torch.zeros((0, 3), dtype=torch.float64, device=gu.DEV)
    m = torch.zeros((0, 1), dtype=torch.float64, device=gu.DEV)
    c = torch.zeros((0, 12), dtype=torch.float64, device=gu.DEV)
    gm, gmu = _grids(pa, wl, torch.float64, 3)
    sd = pa.create_spline(pa.NODE, 2, mesh, x)
    pa.APIC.step(ctx, m, c, gm, gmu, sd, _geom(pa, wl, 2, pa.NODE))
    ctx.synchronize()
    assert not gm.tensor.any() and not gmu.tensor.any()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_step_graph_replays(pa, wl, dtype):
    """The step captured as a step graph and replayed three times, with
    eager work and garbage-filled grids between the replays, against the
    eager step."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-13 if dtype == "f64" else 1e-6
    xh, mh, ch = _inputs(wl, dtype == "f32")
    mesh = pa.Mesh(wl.low, wl.h)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ctx = pa.Context(0)
        x = gu.soa(xh, tdt)
        m = gu.soa(mh[:, None], tdt)
        c0 = gu.soa(ch.reshape(-1, 12), tdt)
        c = c0.clone()
        geom = _geom(pa, wl, 2, pa.NODE)
        gm, gmu = _grids(pa, wl, tdt, 3)
        sd = pa.create_spline(pa.NODE, 2, mesh, x)
        bins = pa.APIC.step(ctx, m, c, gm, gmu, sd, geom)  # eager reference
        ctx.synchronize()
        ref = (gu.host(gm.tensor).copy(), gu.host(gmu.tensor).copy(), gu.host(c).copy())
        ctx.graph_begin()
        pa.APIC.step(ctx, m, c, gm, gmu, sd, geom, out=bins)
        graph = ctx.graph_end()
        for _ in range(3):
            c.copy_(c0)
            gm.tensor.fill_(5.0)
            gmu.tensor.fill_(-5.0)
            torch.cuda.synchronize()
            graph.launch()
            ctx.synchronize()
            assert gu.rel_inf(gu.host(gm.tensor), ref[0]) < tol
            assert gu.rel_inf(gu.host(gmu.tensor), ref[1]) < tol
            assert gu.rel_inf(gu.host(c), ref[2]) < tol * 10
        graph.close()
        ctx.close()


def test_step_without_active_list(pa, wl):
    """active may be NULL: the step keeps the list of non-empty bricks in its
    scratch, as pic_bin_sort does."""
    import gpu_util as gu
    ctx = gu.ctx()
    mesh = pa.Mesh(wl.low, wl.h)
    x = gu.soa(wl.x)
    m = gu.soa(wl.m[:, None])
    c = gu.soa(wl.c.reshape(-1, 12))
    geom = _geom(pa, wl, 2, pa.NODE)
    nb = pa.bin_count(geom)
    i32 = dict(dtype=torch.int32, device=gu.DEV)
    out = pa.BinSort(torch.empty(wl.np, **i32), torch.empty(wl.np, **i32),
                     torch.empty(nb + 1, **i32), geom, None)
    gm, gmu = _grids(pa, wl, torch.float64, 3)
    pa.APIC.step(ctx, m, c, gm, gmu, pa.create_spline(pa.NODE, 2, mesh, x), geom, out=out)
    ctx.synchronize()
    rm, rmu, rc = _oracle(wl, 2, pa.NODE, wl.x, wl.m, wl.c)
    assert gu.rel_inf(gu.host(gm.tensor), rm) < F64_TOL
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < F64_TOL
    assert gu.rel_inf(gu.host(c).reshape(-1, 4, 3), rc) < F64_TOL


def test_step_reports_particle_outside_bins(pa, wl):
    """A particle outside the binned cells is reported by synchronize, as
    after pic_bin_sort (the outputs are then undefined)."""
    import gpu_util as gu
    ctx = gu.ctx()
    mesh = pa.Mesh(wl.low, wl.h)
    geom = _geom(pa, wl, 2, pa.NODE)
    cell_lo, _ = wl.cell_region
    xh = wl.x.copy()
    # node coordinate 1.0: node-centred cell 1, one below the first binned cell (2)
    xh[5, 0] = wl.low[0] + (int(cell_lo[0]) - 1.0) * wl.h
    x = gu.soa(xh)
    m = gu.soa(wl.m[:, None])
    c = gu.soa(wl.c.reshape(-1, 12))
    gm, gmu = _grids(pa, wl, torch.float64, 3)
    pa.APIC.step(ctx, m, c, gm, gmu, pa.create_spline(pa.NODE, 2, mesh, x), geom)
    with pytest.raises(pa.PicError):
        ctx.synchronize()
    ctx.synchronize()  # the error was cleared
