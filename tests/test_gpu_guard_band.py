"""Guard band: p2g writes nothing outside the nodes its particles' stencils
touch, and g2p nothing past the particles it was given.

Every grid window is a strided view into a larger buffer whose every
element holds a sentinel; p2g accumulates into the window (no zeroing).
Afterwards every element outside the window, and every window node that
no particle's stencil contains (the stencils come from the oracle's
spline), must still hold the sentinel bit for bit, and the touched nodes
must hold the sentinel plus the oracle's deposit.  Covers the brick
kernels (with bins) and the per-particle kernels (without), APIC orders
1-3 on Node, Face and Edge lattices, PolyPIC (8 and 27 modes) and the
generic P2G, in FP64 and FP32."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

SENTINEL = 0.5 + 2.0 ** -20  # exact in FP32 and FP64
GUARD = 3                    # sentinel layers around each window
H = 0.25
LOW = np.array([-1.0, -0.5, -0.75])
LO = np.array([1, 2, 0], np.int32)
N = np.array([16, 15, 17], np.int32)
P = 600


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _particles(seed, f32):
    rng = np.random.default_rng(seed)
    # cells [4, 12) of the mesh: every stencil of every order and lattice
    # stays inside the window [LO, LO + N)
    x = LOW + H * rng.uniform(4, 12, size=(P, 3))
    return (x.astype(np.float32).astype(np.float64) if f32 else x), rng


def _guarded(ncomp, tdt):
    """A sentinel-filled buffer and its window view (a GridField)."""
    import gpu_util as gu
    shp = tuple(int(v) + 2 * GUARD for v in N) + ((ncomp,) if ncomp else ())
    buf = torch.full(shp, SENTINEL, dtype=tdt, device=gu.DEV)
    win = buf[GUARD:GUARD + int(N[0]), GUARD:GUARD + int(N[1]), GUARD:GUARD + int(N[2])]
    return buf, win


def _touched(order, entity, x):
    """Window nodes contained in some particle's stencil."""
    t = np.zeros(tuple(int(v) for v in N), bool)
    for p in range(len(x)):
        s = orpic.spline(order, entity, LOW, H, x[p])["s"] - LO[:, None]
        t[np.ix_(s[0], s[1], s[2])] = True
    return t


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _check(buf, ref, touched, f32):
    """Guard band and untouched nodes bit-identical to the sentinel; the
    touched nodes hold sentinel + ref."""
    import gpu_util as gu
    b = gu.host(buf)
    sent = _bits(np.full((), SENTINEL, b.dtype))
    outside = np.ones(b.shape[:3], bool)
    outside[GUARD:GUARD + int(N[0]), GUARD:GUARD + int(N[1]), GUARD:GUARD + int(N[2])] = False
    assert np.all(_bits(b[outside]) == sent), "write outside the grid window"
    w = b[GUARD:GUARD + int(N[0]), GUARD:GUARD + int(N[1]), GUARD:GUARD + int(N[2])]
    assert np.all(_bits(w[~touched]) == sent), "write to a node no stencil contains"
    scale = np.abs(ref).max() + SENTINEL
    tol = 1e-5 if f32 else 1e-13
    np.testing.assert_allclose(w[touched], SENTINEL + ref[touched], rtol=0, atol=tol * scale)


def _binning(pa, order, entity):
    dual = pa.aligned_dual(order, entity)
    flags = pa.dual_dims(dual)
    return pa.Binning((2, 2, 2), tuple(13 + f for f in flags), (4, 4, 4), dual)


def _ran_brick(ctx, family):
    return ctx.kernel_time(family + "_stale")[1] == 1


@pytest.mark.parametrize("binned", [True, False])
@pytest.mark.parametrize("entity", [0, 3, 7])  # Node, Face<J>, Edge<K>
@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_apic_p2g_guard_band(pa, dtype, order, entity, binned):
    import gpu_util as gu
    f32 = dtype == "f32"
    tdt = torch.float32 if f32 else torch.float64
    x, rng = _particles(10 * order + entity, f32)
    m = rng.uniform(0.5, 1.5, P)
    c = rng.normal(size=(P, 4, 3))
    if f32:
        m, c = (a.astype(np.float32).astype(np.float64) for a in (m, c))
    stag = entity >= 2
    ctx = gu.ctx()
    mesh = pa.Mesh(LOW, H)
    xt = gu.soa(x, tdt)
    bins = pa.bin_sort(ctx, mesh, xt, _binning(pa, order, entity)) if binned else None
    sd = pa.create_spline(entity, order, mesh, xt, bins)
    mbuf, mwin = _guarded(0, tdt)
    ubuf, uwin = _guarded(0 if stag else 3, tdt)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        pa.APIC.p2g(ctx, gu.soa(m[:, None], tdt), gu.soa(c.reshape(P, 12), tdt),
                    pa.GridField(mwin, tuple(LO)), pa.GridField(uwin, tuple(LO)), sd)
        ctx.synchronize()
        brick = _ran_brick(ctx, "apic_p2g")
    finally:
        ctx.set_timing(False)
    assert brick == binned  # bins aligned with the knot base take the brick kernel
    rm = np.zeros(tuple(int(v) for v in N))
    rmu = np.zeros(rm.shape + (() if stag else (3,)))
    orpic.apic_p2g(order, entity, LOW, H, LO, N, x, m, c, rm, rmu)
    touched = _touched(order, entity, x)
    _check(mbuf, rm, touched, f32)
    _check(ubuf, rmu, touched if stag else np.repeat(touched[..., None], 3, -1), f32)


@pytest.mark.parametrize("order", [1, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_polypic_p2g_guard_band(pa, dtype, order):
    import gpu_util as gu
    f32 = dtype == "f32"
    tdt = torch.float32 if f32 else torch.float64
    x, rng = _particles(50 + order, f32)
    nm = orpic.polypic_nmode(order)
    m = rng.uniform(0.5, 1.5, P)
    u = rng.normal(0, 0.5, size=(P, nm, 3))
    cm = rng.normal(size=(P, nm, 3))
    if f32:
        m, u, cm = (a.astype(np.float32).astype(np.float64) for a in (m, u, cm))
    ctx = gu.ctx()
    mesh = pa.Mesh(LOW, H)
    xt = gu.soa(x, tdt)
    bins = pa.bin_sort(ctx, mesh, xt, _binning(pa, order, 0))
    sd = pa.create_spline(0, order, mesh, xt, bins)
    mbuf, mwin = _guarded(0, tdt)
    ubuf, uwin = _guarded(3, tdt)
    dt = 1e-3
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        pa.PolyPIC.p2g(ctx, gu.soa(m[:, None], tdt), gu.soa(u.reshape(P, -1), tdt),
                       gu.soa(cm.reshape(P, -1), tdt), pa.GridField(uwin, tuple(LO)),
                       pa.GridField(mwin, tuple(LO)), dt, sd)
        ctx.synchronize()
        brick = _ran_brick(ctx, "polypic_p2g")
    finally:
        ctx.set_timing(False)
    assert brick
    rmu = np.zeros(tuple(int(v) for v in N) + (3,))
    rm = np.zeros(tuple(int(v) for v in N))
    orpic.polypic_p2g(order, orpic.NODE, LOW, H, LO, N, x, m, u, cm, dt, rmu, rm)
    touched = _touched(order, 0, x)
    _check(mbuf, rm, touched, f32)
    _check(ubuf, rmu, np.repeat(touched[..., None], 3, -1), f32)


@pytest.mark.parametrize("binned", [True, False])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_p2g_value_guard_band(pa, dtype, binned):
    """Generic P2G::value of a 3-vector on the Cell lattice, linear."""
    import gpu_util as gu
    f32 = dtype == "f32"
    tdt = torch.float32 if f32 else torch.float64
    x, rng = _particles(77, f32)
    v = rng.normal(size=(P, 3))
    if f32:
        v = v.astype(np.float32).astype(np.float64)
    ctx = gu.ctx()
    mesh = pa.Mesh(LOW, H)
    xt = gu.soa(x, tdt)
    bins = pa.bin_sort(ctx, mesh, xt, _binning(pa, 1, 1)) if binned else None
    sd = pa.create_spline(1, 1, mesh, xt, bins)
    fbuf, fwin = _guarded(3, tdt)
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    try:
        pa.P2G.value(ctx, sd, gu.soa(v, tdt), pa.GridField(fwin, tuple(LO)))
        ctx.synchronize()
        brick = _ran_brick(ctx, "p2g_interp")
    finally:
        ctx.set_timing(False)
    assert brick == binned
    rf = orpic.p2g_interp(orpic.INTERP_VALUE, 1, 1, LOW, H, LO, N, x, v, 3)
    touched = _touched(1, 1, x)
    _check(fbuf, rf.reshape(tuple(int(n) for n in N) + (3,)),
           np.repeat(touched[..., None], 3, -1), f32)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_g2p_writes_only_its_particles(pa, dtype):
    """APIC g2p (brick and per-particle) on the first P of P + 64 particle
    rows: the 64 rows after them keep their sentinel."""
    import gpu_util as gu
    f32 = dtype == "f32"
    tdt = torch.float32 if f32 else torch.float64
    x, rng = _particles(91, f32)
    ctx = gu.ctx()
    mesh = pa.Mesh(LOW, H)
    xt = gu.soa(x, tdt)
    gfield = rng.normal(size=tuple(int(v) for v in N) + (3,))
    for bins in (pa.bin_sort(ctx, mesh, xt, _binning(pa, 2, 0)), None):
        sd = pa.create_spline(0, 2, mesh, xt, bins)
        cbuf = torch.full((P + 64, 12), SENTINEL, dtype=tdt, device=gu.DEV)
        pa.APIC.g2p(ctx, pa.GridField(gu.dev(gfield, tdt), tuple(LO)), cbuf[:P], sd)
        ctx.synchronize()
        tail = gu.host(cbuf[P:])
        assert np.all(_bits(tail) == _bits(np.full((), SENTINEL, tail.dtype)))
        assert not np.any(_bits(gu.host(cbuf[:P])) == _bits(np.full((), SENTINEL, tail.dtype)))
