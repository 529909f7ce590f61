"""HIP kernels against the CPU oracle on identical seeded inputs.

Bars: FP64 within 1e-12 relative (||.||_inf / ||ref||_inf; atomic summation
order differs), cell keys / permutations / bin offsets bit-exact, mass and
momentum conserved to round-off, FP32 paths within 1e-5 of the FP64 oracle
evaluated on the same (FP32-rounded) inputs.
"""
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
def wl0():
    from picasso_amd import synth
    return synth.apic_uniform(32, 8)


def _node_grids(pa, lo, n, dtype, ncomp=3):
    import gpu_util as gu
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.zeros(shp, dtype=dtype, device=gu.DEV), tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.zeros(shp + (ncomp,), dtype=dtype, device=gu.DEV),
                       tuple(int(v) for v in lo))
    return gm, gmu


def _oracle_step(wl, x, m, c, order=2):
    lo, n = wl.node_window
    gm = np.zeros(tuple(n))
    gmu = np.zeros(tuple(n) + (3,))
    orpic.apic_p2g(order, orpic.NODE, wl.low, wl.h, lo, n, x, m, c, gm, gmu)
    gu = orpic.grid_normalize(gm, gmu)
    cout = orpic.apic_g2p(order, orpic.NODE, wl.low, wl.h, lo, n, x, gu)
    return gm, gmu, gu, cout


@pytest.mark.parametrize("layout", ["soa", "aos"])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_apic_step_f64_uniform(pa, wl0, layout, order):
    """p2g -> normalise -> g2p on the 32^3 x 8 ppc workload, FP64."""
    import gpu_util as gu
    lo, n = wl0.node_window
    put = gu.soa if layout == "soa" else gu.dev
    x = put(wl0.x)
    m = put(wl0.m[:, None])
    c = put(wl0.c.reshape(-1, 12))
    mesh = pa.Mesh(wl0.low, wl0.h)
    sd = pa.create_spline(pa.FieldLocation.Node, order, mesh, x)
    gm, gmu = _node_grids(pa, lo, n, torch.float64)
    gu_ = pa.GridField(torch.empty_like(gmu.tensor), gmu.lo)
    pa.APIC.p2g(gu.ctx(), m, c, gm, gmu, sd, zero_first=True)
    pa.grid_normalize(gu.ctx(), gm, gmu, gu_)
    pa.APIC.g2p(gu.ctx(), gu_, c, sd)
    gu.ctx().synchronize()

    rm, rmu, ru, rc = _oracle_step(wl0, wl0.x, wl0.m, wl0.c, order)
    assert gu.rel_inf(gu.host(gm.tensor), rm) < F64_TOL
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < F64_TOL
    assert gu.rel_inf(gu.host(gu_.tensor), ru) < F64_TOL
    cc = gu.host(c).reshape(-1, 4, 3)
    assert gu.rel_inf(cc[:, 0], rc[:, 0]) < F64_TOL
    assert gu.rel_inf(cc[:, 1:], rc[:, 1:]) < F64_TOL
    # conservation
    assert abs(gu.host(gm.tensor).sum() - wl0.m.sum()) <= 1e-13 * wl0.m.sum()
    mom = (wl0.m[:, None] * wl0.c[:, 0, :]).sum(0)
    scale = (wl0.m[:, None] * np.abs(wl0.c[:, 0, :])).sum(0)
    np.testing.assert_array_less(np.abs(gu.host(gmu.tensor).reshape(-1, 3).sum(0) - mom),
                                 1e-12 * scale)


@pytest.mark.parametrize("order", [2])
def test_apic_step_f32_uniform(pa, wl0, order):
    """FP32 path against the FP64 oracle on the FP32-rounded inputs."""
    import gpu_util as gu
    lo, n = wl0.node_window
    x32 = wl0.x.astype(np.float32)
    m32 = wl0.m.astype(np.float32)
    c32 = wl0.c.astype(np.float32)
    x = gu.soa(x32, torch.float32)
    m = gu.soa(m32[:, None], torch.float32)
    c = gu.soa(c32.reshape(-1, 12), torch.float32)
    mesh = pa.Mesh(wl0.low, wl0.h)
    sd = pa.create_spline(pa.FieldLocation.Node, order, mesh, x)
    gm, gmu = _node_grids(pa, lo, n, torch.float32)
    gu_ = pa.GridField(torch.empty_like(gmu.tensor), gmu.lo)
    pa.APIC.p2g(gu.ctx(), m, c, gm, gmu, sd, zero_first=True)
    pa.grid_normalize(gu.ctx(), gm, gmu, gu_, rel_mass_eps=1e-6)
    pa.APIC.g2p(gu.ctx(), gu_, c, sd)
    gu.ctx().synchronize()
    rm, rmu, ru, rc = _oracle_step(wl0, x32.astype(np.float64), m32.astype(np.float64),
                                   c32.astype(np.float64), order)
    assert gu.rel_inf(gu.host(gm.tensor), rm) < F32_TOL
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < F32_TOL
    cc = gu.host(c).reshape(-1, 4, 3)
    assert gu.rel_inf(cc[:, 0], rc[:, 0]) < F32_TOL
    assert gu.rel_inf(cc[:, 1:], rc[:, 1:]) < 1e-4


@pytest.mark.parametrize("entity", range(8))
@pytest.mark.parametrize("order", [1, 2, 3])
def test_apic_entities_random(pa, entity, order):
    """All entity lattices, random particles, FP64, vs the oracle."""
    import gpu_util as gu
    rng = np.random.default_rng(100 * order + entity)
    h = 0.25
    low = np.array([-1.0, -0.5, -0.75])
    x = low + h * rng.uniform(4, 12, size=(500, 3))
    lo = np.array([1, 2, 0], np.int32)
    n = np.array([16, 15, 17], np.int32)
    m = rng.uniform(0.5, 1.5, 500)
    c = rng.normal(size=(500, 4, 3))
    stag = entity >= 2
    rm = np.zeros(tuple(n))
    rmu = np.zeros(tuple(n) + (() if stag else (3,)))
    orpic.apic_p2g(order, entity, low, h, lo, n, x, m, c, rm, rmu)
    gfield = rng.normal(size=tuple(n) + (() if stag else (3,)))
    rc = orpic.apic_g2p(order, entity, low, h, lo, n, x, gfield)

    mesh = pa.Mesh(low, h)
    sd = pa.create_spline(entity, order, mesh, gu.dev(x))
    gm = pa.GridField(gu.dev(np.zeros(tuple(n))), tuple(lo))
    gmu = pa.GridField(gu.dev(np.zeros(rmu.shape)), tuple(lo))
    pa.APIC.p2g(gu.ctx(), gu.dev(m[:, None]), gu.dev(c.reshape(-1, 12)), gm, gmu, sd)
    cg = gu.dev(np.zeros((500, 12)))
    pa.APIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield), tuple(lo)), cg, sd)
    gu.ctx().synchronize()
    assert gu.rel_inf(gu.host(gm.tensor), rm) < F64_TOL
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < F64_TOL
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3), rc) < F64_TOL


def _binning(pa, wl, order=2, entity=0, brick=(8, 8, 8)):
    """Bins aligned with the spline's knot base for (order, entity): one more
    bin in every dimension binned by node-centred cell."""
    dual = pa.aligned_dual(order, entity)
    flags = pa.dual_dims(dual)
    cell_lo, ncell = wl.cell_region
    return pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), brick, dual)


@pytest.mark.parametrize("brick", [(1, 1, 1), (4, 4, 4), (8, 8, 8), (2, 4, 8)])
@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("dual", [0, 1, 0x10 | 1, 0x10 | 6])
def test_bin_sort_bit_exact(pa, wl0, brick, dt, dual):
    import gpu_util as gu
    tdt = torch.float64 if dt == "f64" else torch.float32
    xn = wl0.x if dt == "f64" else wl0.x.astype(np.float32)
    x = gu.soa(xn, tdt)
    cell_lo, ncell = wl0.cell_region
    flags = pa.dual_dims(dual)
    geom = pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), brick, dual)
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl0.low, wl0.h), x, geom)
    gu.ctx().synchronize()
    keys, _ = orpic.cell_keys(xn.astype(np.float64), wl0.low, wl0.h, geom.cell_lo, geom.ncell,
                              brick, dual)
    nb = pa.bin_count(geom)
    bvol = int(np.prod(brick))
    perm, kstart = orpic.bin_sort(keys, nb * bvol)
    np.testing.assert_array_equal(gu.host(res.perm).view(np.uint32), perm)
    np.testing.assert_array_equal(gu.host(res.keys).view(np.uint32), keys[perm])
    np.testing.assert_array_equal(gu.host(res.bin_start).view(np.uint32), kstart[::bvol])
    _check_active(gu, res, kstart[::bvol])


def _check_active(gu, res, bin_start):
    """The active list is [count, ids of the non-empty bricks ascending]."""
    nonempty = np.flatnonzero(np.diff(bin_start.astype(np.int64)) > 0).astype(np.uint32)
    act = gu.host(res.active).view(np.uint32)
    assert act[0] == len(nonempty)
    np.testing.assert_array_equal(act[1:1 + len(nonempty)], nonempty)


def test_bin_sort_many_scan_tiles(pa, wl0):
    """Over 2048 x 2048 bricks: the tile sums of the counts get a scan pass
    of their own; perm, bin_start and the active list still match."""
    import gpu_util as gu
    x = gu.soa(wl0.x)
    cell_lo, ncell = wl0.cell_region
    big = (int(ncell[0]) + 300, int(ncell[1]) + 200, int(ncell[2]) + 60)
    geom = pa.Binning(tuple(int(v) for v in cell_lo), big, (1, 1, 1), 0)
    nb = pa.bin_count(geom)
    assert nb > 2048 * 2048
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl0.low, wl0.h), x, geom)
    gu.ctx().synchronize()
    keys, _ = orpic.cell_keys(wl0.x, wl0.low, wl0.h, geom.cell_lo, geom.ncell, (1, 1, 1), 0)
    perm, kstart = orpic.bin_sort(keys, nb)
    np.testing.assert_array_equal(gu.host(res.perm).view(np.uint32), perm)
    np.testing.assert_array_equal(gu.host(res.bin_start).view(np.uint32), kstart)
    _check_active(gu, res, kstart)


def test_bin_sort_out_of_range_reported(pa, wl0):
    """Particles outside the bins are reported; perm stays a permutation,
    with those particles after the last brick (key 0xffffffff) and the
    others sorted as the oracle sorts them."""
    import gpu_util as gu
    x = gu.soa(wl0.x)
    geom = pa.Binning((3, 2, 2), (31, 32, 32), (8, 8, 8), 0)  # cells at i = 2 left out
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl0.low, wl0.h), x, geom)
    with pytest.raises(pa.PicError):
        gu.ctx().synchronize()
    perm = gu.host(res.perm).view(np.uint32)
    keys = gu.host(res.keys).view(np.uint32)
    np.testing.assert_array_equal(np.sort(perm), np.arange(len(wl0.x), dtype=np.uint32))
    cells = np.floor((wl0.x - wl0.low) / wl0.h).astype(np.int64)
    outside = cells[:, 0] < 3
    nin = int((~outside).sum())
    assert outside.any()
    assert outside[perm[nin:]].all() and (keys[nin:] == 0xFFFFFFFF).all()
    ref_keys, _ = orpic.cell_keys(wl0.x[~outside], wl0.low, wl0.h, geom.cell_lo, geom.ncell,
                                  (8, 8, 8), 0)
    ref_perm, _ = orpic.bin_sort(ref_keys, pa.bin_count(geom) * 512)
    np.testing.assert_array_equal(perm[:nin], np.flatnonzero(~outside)[ref_perm])


def _check_brick(pa, wl, order, entity, dtype, edge):
    """Binned p2g/g2p (brick kernels) on workload wl against the oracle."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    xh = wl.x.astype(np.float32).astype(np.float64) if f32 else wl.x
    mh = wl.m.astype(np.float32).astype(np.float64) if f32 else wl.m
    ch = wl.c.astype(np.float32).astype(np.float64) if f32 else wl.c
    lo, n = wl.node_window
    mesh = pa.Mesh(wl.low, wl.h)
    x = gu.soa(xh, tdt)
    m = gu.soa(mh[:, None], tdt)
    c = gu.soa(ch.reshape(-1, 12), tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, x, _binning(pa, wl, order, entity, (edge,) * 3))
    sd = pa.create_spline(entity, order, mesh, x, bins)
    gm, gmu = _node_grids(pa, lo, n, tdt)
    pa.APIC.p2g(gu.ctx(), m, c, gm, gmu, sd, zero_first=True)
    rng = np.random.default_rng(order)
    gfield = rng.normal(size=tuple(n) + (3,))
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    cg = torch.zeros_like(c)
    pa.APIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cg, sd)
    gu.ctx().synchronize()
    rm = np.zeros(tuple(n))
    rmu = np.zeros(tuple(n) + (3,))
    orpic.apic_p2g(order, entity, wl.low, wl.h, lo, n, xh, mh, ch, rm, rmu)
    rc = orpic.apic_g2p(order, entity, wl.low, wl.h, lo, n, xh, gfield)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3)[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3)[:, 1:], rc[:, 1:]) < tol * 10


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", [0, 1])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("edge", [4, 8])
def test_brick_transfers_vs_oracle(pa, wl0, order, entity, dtype, edge):
    _check_brick(pa, wl0, order, entity, dtype, edge)


@pytest.fixture(scope="module")
def wl_dense():
    """16^3 cells: 2000 particles spread over the box plus 6000 packed into
    a 3-cell cube, so some bricks hold several staging chunks and some bins
    hundreds of particles."""
    from picasso_amd import synth
    rng = np.random.default_rng(21)
    n = 16
    h = 1.0 / n
    x = np.concatenate([rng.uniform(0, 1, size=(2000, 3)),
                        0.5 + rng.uniform(-1.5 * h, 1.5 * h, size=(6000, 3))])
    rng.shuffle(x)
    m = rng.uniform(0.5, 1.5, len(x)) * h ** 3
    c = rng.normal(0, 0.1, size=(len(x), 4, 3))
    c[:, 0] = rng.normal(size=(len(x), 3))
    return synth.Workload("dense_cluster_16^3", n, 0, h, np.full(3, -synth.HALO * h), x, m, c=c)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_brick_dense_cluster(pa, wl_dense, order, dtype):
    _check_brick(pa, wl_dense, order, 0, dtype, 4)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", range(2, 8))
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_brick_staggered_vs_oracle(pa, wl0, order, entity, dtype):
    """Face/Edge lattices binned by aligned_dual(order, entity), a dual flag
    per dimension, take the brick kernels: p2g fills the scalar mass and
    momentum grids of the staggered component, g2p (plain, and with the
    normalisation fused in) rebuilds that column of c_p and zeroes the
    others (Picasso_APIC.hpp:148-232, 477-544), all against the oracle."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    rnd = (lambda a: a.astype(np.float32).astype(np.float64)) if f32 else (lambda a: a)
    xh, mh, ch = rnd(wl0.x), rnd(wl0.m), rnd(wl0.c)
    lo, n = wl0.node_window
    shp = tuple(int(v) for v in n)
    mesh = pa.Mesh(wl0.low, wl0.h)
    x = gu.soa(xh, tdt)
    m = gu.soa(mh[:, None], tdt)
    c = gu.soa(ch.reshape(-1, 12), tdt)
    ctx = gu.ctx()
    bins = pa.bin_sort(ctx, mesh, x, _binning(pa, wl0, order, entity, (4, 4, 4)))
    sd = pa.create_spline(entity, order, mesh, x, bins)
    gm = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV), tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV), tuple(int(v) for v in lo))
    gfield = rnd(np.random.default_rng(entity).normal(size=shp))
    cg = torch.zeros_like(c)
    cn = torch.zeros_like(c)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        pa.APIC.p2g(ctx, m, c, gm, gmu, sd, zero_first=True)
        pa.APIC.g2p(ctx, pa.GridField(gu.dev(gfield, tdt), tuple(int(v) for v in lo)), cg, sd)
        pa.APIC.g2p_normalized(ctx, gm, gmu, cn, sd)
        ctx.synchronize()
        # The brick path ends with its stale-list pass, which the
        # per-particle kernels do not have.
        brick_runs = [ctx.kernel_time(k)[1] for k in ("apic_p2g_stale", "apic_g2p_stale")]
    finally:
        ctx.set_timing(False)
    assert brick_runs[0] == 1 and brick_runs[1] == 2
    rm = np.zeros(shp)
    rmu = np.zeros(shp)
    orpic.apic_p2g(order, entity, wl0.low, wl0.h, lo, n, xh, mh, ch, rm, rmu)
    rc = orpic.apic_g2p(order, entity, wl0.low, wl0.h, lo, n, xh, gfield)
    gmh, gmuh = gu.host(gm.tensor), gu.host(gmu.tensor)
    u = np.where(gmh > 0, gmuh / np.where(gmh > 0, gmh, 1.0), 0.0)
    rn = orpic.apic_g2p(order, entity, wl0.low, wl0.h, lo, n, xh, u)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gmh, rm) < tol
    assert gu.rel_inf(gmuh, rmu) < tol
    cgh = gu.host(cg).reshape(-1, 4, 3)
    cnh = gu.host(cn).reshape(-1, 4, 3)
    dim = (entity - 2) % 3
    others = [d for d in range(3) if d != dim]
    assert np.all(cgh[:, :, others] == 0) and np.all(cnh[:, :, others] == 0)
    assert gu.rel_inf(cgh[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(cgh[:, 1:], rc[:, 1:]) < tol * 10
    assert gu.rel_inf(cnh[:, 0], rn[:, 0]) < tol * 10
    assert gu.rel_inf(cnh[:, 1:], rn[:, 1:]) < tol * 100


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_brick_with_stale_bins(pa, wl0, dtype):
    """Bins computed before a third of the particles moved: the brick
    kernels route those particles through the slow path, results still
    match the oracle on the new positions."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    xh = wl0.x.astype(np.float32).astype(np.float64) if f32 else wl0.x.copy()
    x = gu.soa(xh, tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, x, _binning(pa, wl0))
    gu.ctx().synchronize()
    rng = np.random.default_rng(9)
    moved = xh.copy()
    sel = rng.random(len(xh)) < 0.33
    moved[sel] += rng.uniform(-1.5, 1.5, size=(sel.sum(), 3)) * wl0.h
    moved = np.clip(moved, 0.0, 1.0 - 1e-9)
    if f32:
        moved = moved.astype(np.float32).astype(np.float64)
    x.copy_(gu.soa(moved, tdt))
    mh = wl0.m.astype(np.float32).astype(np.float64) if f32 else wl0.m
    ch = wl0.c.astype(np.float32).astype(np.float64) if f32 else wl0.c
    m = gu.soa(mh[:, None], tdt)
    c = gu.soa(ch.reshape(-1, 12), tdt)
    sd = pa.create_spline(0, 2, mesh, x, bins)
    gm, gmu = _node_grids(pa, lo, n, tdt)
    pa.APIC.p2g(gu.ctx(), m, c, gm, gmu, sd, zero_first=True)
    gfield = np.random.default_rng(3).normal(size=tuple(n) + (3,))
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    cg = torch.zeros_like(c)
    pa.APIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cg, sd)
    gu.ctx().synchronize()
    rm = np.zeros(tuple(n))
    rmu = np.zeros(tuple(n) + (3,))
    orpic.apic_p2g(2, 0, wl0.low, wl0.h, lo, n, moved, mh, ch, rm, rmu)
    rc = orpic.apic_g2p(2, 0, wl0.low, wl0.h, lo, n, moved, gfield)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3), rc) < tol * 10


def test_unaligned_bins_fall_back(pa, wl0):
    """Mesh-cell bins with quadratic node splines are not knot-aligned: the
    transfers fall back to the scatter/gather kernels and stay exact."""
    import gpu_util as gu
    lo, n = wl0.node_window
    x = gu.soa(wl0.x)
    m = gu.soa(wl0.m[:, None])
    c = gu.soa(wl0.c.reshape(-1, 12))
    mesh = pa.Mesh(wl0.low, wl0.h)
    cell_lo, ncell = wl0.cell_region
    bins = pa.bin_sort(gu.ctx(), mesh, x, pa.Binning(tuple(cell_lo), tuple(ncell), (8, 8, 8), 0))
    gm, gmu = _node_grids(pa, lo, n, torch.float64)
    pa.APIC.p2g(gu.ctx(), m, c, gm, gmu, pa.create_spline(0, 2, mesh, x, bins))
    gu.ctx().synchronize()
    rm = np.zeros(tuple(n))
    rmu = np.zeros(tuple(n) + (3,))
    orpic.apic_p2g(2, 0, wl0.low, wl0.h, lo, n, wl0.x, wl0.m, wl0.c, rm, rmu)
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < F64_TOL


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_polypic_random(pa, order, ncomp, dtype):
    import gpu_util as gu
    rng = np.random.default_rng(7 * order + ncomp)
    h = 0.125
    low = np.zeros(3)
    P = 300
    x = low + h * rng.uniform(4, 12, size=(P, 3))
    lo = np.array([0, 0, 0], np.int32)
    n = np.array([18, 18, 18], np.int32)
    nm = orpic.polypic_nmode(order)
    m = rng.uniform(0.5, 1.5, P)
    u = rng.normal(size=(P, nm, 3))
    cm = rng.normal(size=(P, nm, ncomp))
    tdt = torch.float64 if dtype == "f64" else torch.float32
    if dtype == "f32":
        x, m, u, cm = (a.astype(np.float32).astype(np.float64) for a in (x, m, u, cm))
    dt = 1e-3
    rmu = np.zeros(tuple(n) + (ncomp,))
    rm = np.zeros(tuple(n))
    orpic.polypic_p2g(order, orpic.NODE, low, h, lo, n, x, m, u, cm, dt, rmu, rm)
    gfield = rng.normal(size=tuple(n) + (ncomp,))
    if dtype == "f32":
        gfield = gfield.astype(np.float32).astype(np.float64)
    rc = orpic.polypic_g2p(order, orpic.NODE, low, h, lo, n, x, gfield, ncomp)

    mesh = pa.Mesh(low, h)
    sd = pa.create_spline(0, order, mesh, gu.dev(x, tdt))
    gmu = pa.GridField(gu.dev(np.zeros(rmu.shape), tdt))
    gm = pa.GridField(gu.dev(np.zeros(rm.shape), tdt))
    pa.PolyPIC.p2g(gu.ctx(), gu.dev(m[:, None], tdt), gu.dev(u.reshape(P, -1), tdt),
                   gu.dev(cm.reshape(P, -1), tdt), gmu, gm, dt, sd, ncomp=ncomp)
    cg = gu.dev(np.zeros((P, nm * ncomp)), tdt)
    pa.PolyPIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cg, sd, ncomp=ncomp)
    gu.ctx().synchronize()
    tol = F64_TOL if dtype == "f64" else F32_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    assert gu.rel_inf(gu.host(cg).reshape(P, nm, ncomp), rc) < tol * 10


@pytest.mark.parametrize("order", [1, 2, 3])
def test_gradu_and_update_F(pa, order):
    import gpu_util as gu
    rng = np.random.default_rng(order)
    h = 0.125
    low = np.zeros(3)
    P = 400
    x = low + h * rng.uniform(4, 12, size=(P, 3))
    lo = np.array([0, 0, 0], np.int32)
    n = np.array([18, 18, 18], np.int32)
    gfield = rng.normal(size=tuple(n) + (3,))
    rL = orpic.g2p_gradu(order, orpic.NODE, low, h, lo, n, x, gfield)
    F0 = np.eye(3)[None] + 0.01 * rng.normal(size=(P, 3, 3))
    rF, rdet, rinv = orpic.update_F(rL, 1e-3, F0)

    sd = pa.create_spline(0, order, pa.Mesh(low, h), gu.dev(x))
    L = gu.dev(np.zeros((P, 9)))
    pa.g2p_gradu(gu.ctx(), pa.GridField(gu.dev(gfield)), L, sd)
    F = gu.dev(F0.reshape(P, 9))
    det = gu.dev(np.zeros((P, 1)))
    inv = gu.dev(np.zeros((P, 9)))
    pa.update_F(gu.ctx(), L, 1e-3, F, det, inv)
    gu.ctx().synchronize()
    assert gu.rel_inf(gu.host(L).reshape(P, 3, 3), rL) < F64_TOL
    assert gu.rel_inf(gu.host(F).reshape(P, 3, 3), rF) < F64_TOL
    assert gu.rel_inf(gu.host(det)[:, 0], rdet) < F64_TOL
    assert gu.rel_inf(gu.host(inv).reshape(P, 3, 3), rinv) < F64_TOL


def test_out_of_window_is_reported(pa):
    import gpu_util as gu
    x = gu.dev([[0.01, 0.5, 0.5]])
    mesh = pa.Mesh([0.0, 0.0, 0.0], 0.1)
    gm = pa.GridField(gu.dev(np.zeros((10, 10, 10))))
    gmu = pa.GridField(gu.dev(np.zeros((10, 10, 10, 3))))
    pa.APIC.p2g(gu.ctx(), gu.dev([[1.0]]), gu.dev(np.ones((1, 12))), gm, gmu,
                pa.create_spline(0, 2, mesh, x))
    with pytest.raises(pa.PicError) as ei:
        gu.ctx().synchronize()
    assert ei.value.code == -2
    # nothing was written and the error is cleared
    assert gu.host(gm.tensor).sum() == 0.0
    gu.ctx().synchronize()


def test_empty_inputs(pa):
    import gpu_util as gu
    x = gu.dev(np.zeros((0, 3)))
    mesh = pa.Mesh([0.0, 0.0, 0.0], 0.1)
    gm = pa.GridField(gu.dev(np.ones((4, 4, 4))))
    gmu = pa.GridField(gu.dev(np.ones((4, 4, 4, 3))))
    pa.APIC.p2g(gu.ctx(), gu.dev(np.zeros((0, 1))), gu.dev(np.zeros((0, 12))), gm, gmu,
                pa.create_spline(0, 2, mesh, x), zero_first=True)
    gu.ctx().synchronize()
    assert gu.host(gm.tensor).sum() == 0.0 and gu.host(gmu.tensor).sum() == 0.0
    res = pa.bin_sort(gu.ctx(), mesh, x, pa.Binning((0, 0, 0), (4, 4, 4), (2, 2, 2), 0))
    gu.ctx().synchronize()
    assert np.all(gu.host(res.bin_start) == 0)


@pytest.mark.parametrize("binned", [False, True])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_g2p_normalized_matches_normalize_then_g2p(pa, wl0, binned, dtype):
    """Fused normalisation + g2p against the oracle's normalise then g2p."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    xh = wl0.x.astype(np.float32).astype(np.float64) if f32 else wl0.x
    mh = wl0.m.astype(np.float32).astype(np.float64) if f32 else wl0.m
    ch = wl0.c.astype(np.float32).astype(np.float64) if f32 else wl0.c
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    x = gu.soa(xh, tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, x, _binning(pa, wl0)) if binned else None
    sd = pa.create_spline(0, 2, mesh, x, bins)
    gm, gmu = _node_grids(pa, lo, n, tdt)
    c = gu.soa(ch.reshape(-1, 12), tdt)
    pa.APIC.p2g(gu.ctx(), gu.soa(mh[:, None], tdt), c, gm, gmu, sd, zero_first=True)
    pa.APIC.g2p_normalized(gu.ctx(), gm, gmu, c, sd)
    gu.ctx().synchronize()
    rm, rmu, ru, rc = _oracle_step(wl0, xh, mh, ch, 2)
    tol = F32_TOL if f32 else F64_TOL
    cc = gu.host(c).reshape(-1, 4, 3)
    assert gu.rel_inf(cc[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(cc[:, 1:], rc[:, 1:]) < tol * 10


def test_active_brick_list(pa, wl0):
    """The sort's list of non-empty bricks equals the bricks with particles."""
    import gpu_util as gu
    from picasso_amd import synth
    wl = synth.apic_sphere(32, 4, radius=0.3)
    x = gu.soa(wl.x)
    geom = _binning(pa, wl)
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl.low, wl.h), x, geom)
    gu.ctx().synchronize()
    start = gu.host(res.bin_start).astype(np.int64)
    want = np.nonzero(np.diff(start) > 0)[0]
    act = gu.host(res.active)
    assert act[0] == len(want)
    np.testing.assert_array_equal(act[1:1 + len(want)], want)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("binned", [False, True])
@pytest.mark.parametrize("fused_norm", [False, True])
def test_g2p_update_F_matches_oracle(pa, wl0, order, dtype, binned, fused_norm):
    """pic_apic_g2p_update_F against the oracle's g2p, velocity gradient and
    FP64 F update run one after the other (FP64 F, det, F^-1 at 1e-12 for
    FP64 inputs; FP32 inputs within the FP32 bar)."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    xh = wl0.x.astype(np.float32).astype(np.float64) if f32 else wl0.x
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    x = gu.soa(xh, tdt)
    rng = np.random.default_rng(70 + order)
    P = len(xh)
    F0 = np.eye(3) + rng.normal(0, 0.01, size=(P, 3, 3))
    dt = 1e-3
    if fused_norm:
        gmh = rng.uniform(0.5, 1.5, size=tuple(n))
        gmh[rng.random(gmh.shape) < 0.05] = 0.0
        gmuh = rng.normal(size=tuple(n) + (3,))
        if f32:
            gmh = gmh.astype(np.float32).astype(np.float64)
            gmuh = gmuh.astype(np.float32).astype(np.float64)
        uh = orpic.grid_normalize(gmh, gmuh)
        m_i = pa.GridField(gu.dev(gmh, tdt))
        u_i = pa.GridField(gu.dev(gmuh, tdt))
    else:
        uh = rng.normal(size=tuple(n) + (3,))
        if f32:
            uh = uh.astype(np.float32).astype(np.float64)
        m_i = None
        u_i = pa.GridField(gu.dev(uh, tdt))
    bins = None
    if binned:
        bins = pa.bin_sort(gu.ctx(), mesh, x, _binning(pa, wl0, order, 0, (4, 4, 4)))
    sd = pa.create_spline(0, order, mesh, x, bins)
    cg = gu.soa(np.zeros((P, 12)), tdt)
    F = gu.soa(F0.reshape(P, 9))
    det = gu.soa(np.zeros((P, 1)))
    Finv = gu.soa(np.zeros((P, 9)))
    pa.APIC.g2p_update_F(gu.ctx(), u_i, cg, dt, F, det, Finv, sd, m_i=m_i)
    gu.ctx().synchronize()
    rc = orpic.apic_g2p(order, 0, wl0.low, wl0.h, lo, n, xh, uh)
    rg = orpic.g2p_gradu(order, 0, wl0.low, wl0.h, lo, n, xh, uh)
    rF, rdet, rFinv = orpic.update_F(rg, dt, F0)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3)[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(gu.host(cg).reshape(-1, 4, 3)[:, 1:], rc[:, 1:]) < tol * 10
    # F, det, F^-1 are FP64; with FP32 grid data the gradient carries FP32
    # rounding, scaled by dt in F.
    ftol = 1e-12 if not f32 else 1e-6
    assert gu.rel_inf(gu.host(F).reshape(-1, 3, 3), rF) < ftol
    assert gu.rel_inf(gu.host(det)[:, 0], rdet) < ftol
    assert gu.rel_inf(gu.host(Finv).reshape(-1, 3, 3), rFinv) < ftol


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", [2, 4, 6])
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_polypic_brick_staggered_vs_oracle(pa, wl0, order, entity, ncomp, dtype):
    """Staggered PolyPIC (Picasso_PolyPIC.hpp:138-229, 299-359) on bins with a
    dual flag per dimension takes the brick kernels: p2g projects column
    `dim` (entity dim for a 3-column field, else 0) onto the scalar grid,
    g2p rewrites only that column; against the oracle."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    rng = np.random.default_rng(300 + 10 * order + entity + ncomp)
    P = len(wl0.x)
    nm = orpic.polypic_nmode(order)
    x, m = wl0.x.copy(), wl0.m.copy()
    u = rng.normal(0, 0.5, size=(P, nm, 3))
    cm = rng.normal(size=(P, nm, ncomp))
    c0 = rng.normal(size=(P, nm, ncomp))
    if f32:
        x, m, u, cm, c0 = (a.astype(np.float32).astype(np.float64) for a in (x, m, u, cm, c0))
    lo, n = wl0.node_window
    shp = tuple(int(v) for v in n)
    mesh = pa.Mesh(wl0.low, wl0.h)
    xt = gu.soa(x, tdt)
    ctx = gu.ctx()
    bins = pa.bin_sort(ctx, mesh, xt, _binning(pa, wl0, order, entity, (4, 4, 4)))
    sd = pa.create_spline(entity, order, mesh, xt, bins)
    dt = 1e-3
    gmu = pa.GridField(gu.dev(np.zeros(shp), tdt))
    gm = pa.GridField(gu.dev(np.zeros(shp), tdt))
    gfield = rng.normal(size=shp)
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    cg = gu.soa(c0.reshape(P, -1), tdt)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        pa.PolyPIC.p2g(ctx, gu.soa(m[:, None], tdt), gu.soa(u.reshape(P, -1), tdt),
                       gu.soa(cm.reshape(P, -1), tdt), gmu, gm, dt, sd, ncomp=ncomp,
                       zero_first=True)
        pa.PolyPIC.g2p(ctx, pa.GridField(gu.dev(gfield, tdt)), cg, sd, ncomp=ncomp)
        ctx.synchronize()
        runs = [ctx.kernel_time(k)[1] for k in ("polypic_p2g_stale", "polypic_g2p_stale")]
    finally:
        ctx.set_timing(False)
    assert runs == [1, 1]  # the brick path ran (its stale-list pass is timed)
    rmu = np.zeros(shp)
    rm = np.zeros(shp)
    orpic.polypic_p2g(order, entity, wl0.low, wl0.h, lo, n, x, m, u, cm, dt, rmu, rm)
    rc = orpic.polypic_g2p(order, entity, wl0.low, wl0.h, lo, n, x, gfield, ncomp, c=c0.copy())
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    cgh = gu.host(cg).reshape(P, nm, ncomp)
    col = (entity - 2) % 3 if ncomp == 3 else 0
    others = [d for d in range(ncomp) if d != col]
    np.testing.assert_array_equal(cgh[:, :, others], c0[:, :, others].astype(cgh.dtype))
    for lo_r, hi_r in ((0, 1), (1, 4), (4, nm)):
        assert gu.rel_inf(cgh[:, lo_r:hi_r, col], rc[:, lo_r:hi_r, col]) < tol * 10


def _polypic_binned_check(pa, wl, order, dtype, stale=False):
    """PolyPIC p2g through the brick kernel (bins from pic_bin_sort) against
    the oracle, three grid components."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    rng = np.random.default_rng(90 + order)
    P = len(wl.x)
    nm = orpic.polypic_nmode(order)
    x = wl.x.copy()
    m = wl.m.copy()
    u = rng.normal(0, 0.5, size=(P, nm, 3))
    cm = rng.normal(size=(P, nm, 3))
    if f32:
        x, m, u, cm = (a.astype(np.float32).astype(np.float64) for a in (x, m, u, cm))
    lo, n = wl.node_window
    mesh = pa.Mesh(wl.low, wl.h)
    xt = gu.soa(x, tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, xt, _binning(pa, wl, order, 0, (4, 4, 4)))
    if stale:
        gu.ctx().synchronize()
        sel = rng.random(P) < 0.3
        x[sel] += rng.uniform(-1.5, 1.5, size=(sel.sum(), 3)) * wl.h
        x = np.clip(x, 0.0, 1.0 - 1e-9)
        if f32:
            x = x.astype(np.float32).astype(np.float64)
        xt.copy_(gu.soa(x, tdt))
    dt = 1e-3
    rmu = np.zeros(tuple(n) + (3,))
    rm = np.zeros(tuple(n))
    orpic.polypic_p2g(order, orpic.NODE, wl.low, wl.h, lo, n, x, m, u, cm, dt, rmu, rm)
    sd = pa.create_spline(0, order, mesh, xt, bins)
    gmu = pa.GridField(gu.dev(np.zeros(rmu.shape), tdt))
    gm = pa.GridField(gu.dev(np.zeros(rm.shape), tdt))
    pa.PolyPIC.p2g(gu.ctx(), gu.soa(m[:, None], tdt), gu.soa(u.reshape(P, -1), tdt),
                   gu.soa(cm.reshape(P, -1), tdt), gmu, gm, dt, sd, zero_first=True)
    gfield = rng.normal(size=tuple(n) + (3,))
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    rc = orpic.polypic_g2p(order, orpic.NODE, wl.low, wl.h, lo, n, x, gfield, 3)
    cg = gu.soa(np.zeros((P, nm * 3)), tdt)
    pa.PolyPIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cg, sd)
    gu.ctx().synchronize()
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    assert abs(gu.host(gm.tensor).sum() - m.sum()) <= (1e-6 if f32 else 1e-13) * m.sum()
    cgh = gu.host(cg).reshape(P, nm, 3)
    # per mode family (values, first and second derivatives differ in scale)
    for lo_r, hi_r in ((0, 1), (1, 4), (4, nm)):
        assert gu.rel_inf(cgh[:, lo_r:hi_r], rc[:, lo_r:hi_r]) < tol * 10


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_polypic_brick_p2g_vs_oracle(pa, wl0, order, dtype):
    _polypic_binned_check(pa, wl0, order, dtype)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_polypic_brick_p2g_stale_and_dense(pa, wl0, wl_dense, dtype):
    _polypic_binned_check(pa, wl0, 3, dtype, stale=True)
    _polypic_binned_check(pa, wl_dense, 3, dtype)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_step_graph_replays_match_eager(pa, wl0, dtype):
    """sort -> p2g -> fused normalise+g2p captured once (pic_graph_*) and
    replayed three times, with eager copies and synchronisations between the
    replays; every replay matches the eager step (atomic summation order
    aside)."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-13 if dtype == "f64" else 2e-6
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ctx = pa.Context(0)
        x = gu.soa(wl0.x, tdt)
        m = gu.soa(wl0.m[:, None], tdt)
        c0 = gu.soa(wl0.c.reshape(-1, 12), tdt)
        c = c0.clone()
        geom = _binning(pa, wl0, 2, 0, (4, 4, 4))
        bins = pa.bin_sort(ctx, mesh, x, geom)
        gm, gmu = _node_grids(pa, lo, n, tdt)

        def step():
            pa.bin_sort(ctx, mesh, x, geom, out=bins)
            sd = pa.create_spline(0, 2, mesh, x, bins)
            pa.APIC.p2g(ctx, m, c, gm, gmu, sd, zero_first=True)
            pa.APIC.g2p_normalized(ctx, gm, gmu, c, sd)

        step()  # eager reference (also sizes the scratch)
        ctx.synchronize()
        ref = (gu.host(gm.tensor).copy(), gu.host(gmu.tensor).copy(), gu.host(c).copy())
        ctx.graph_begin()
        step()
        graph = ctx.graph_end()
        for _ in range(3):
            c.copy_(c0)
            gm.tensor.fill_(-1.0)
            graph.launch()
            ctx.synchronize()
            assert gu.rel_inf(gu.host(gm.tensor), ref[0]) < tol
            assert gu.rel_inf(gu.host(gmu.tensor), ref[1]) < tol
            assert gu.rel_inf(gu.host(c), ref[2]) < tol * 10
        graph.close()
        ctx.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_permute_matches_gather(pa, dtype):
    """pic_permute against torch indexing, SoA and AoS fields."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(11)
    P = 5000
    perm = torch.as_tensor(rng.permutation(P).astype(np.int32), device=gu.DEV)
    a = gu.soa(rng.normal(size=(P, 3)), tdt)
    b = gu.dev(rng.normal(size=(P, 12)), tdt)
    a2 = torch.empty_like(a.t()).t()
    b2 = torch.empty_like(b)
    pa.permute(gu.ctx(), perm, [(a, a2), (b, b2)])
    gu.ctx().synchronize()
    idx = perm.long()
    assert torch.equal(a2, a[idx])
    assert torch.equal(b2, b[idx])
