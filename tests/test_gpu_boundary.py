"""The library's boundary as a Picasso caller meets it: step-graph and
scratch lifetime, device-side bounds reporting on the brick paths, argument
validation in the host mirror, and Cabana AoSoA particle slices
(Picasso_ParticleList.hpp:90-138) through every transfer path against the
oracle."""
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


def _binning(pa, wl, order=2, entity=0, brick=(4, 4, 4)):
    dual = pa.aligned_dual(order, entity)
    flags = pa.dual_dims(dual)
    cell_lo, ncell = wl.cell_region
    return pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), brick, dual)


def test_graph_pins_scratch_until_destroyed(pa, wl0):
    """A step graph replays pointers into its context's scratch buffer, so
    while it is alive an eager call that needs a larger buffer is refused
    (PIC_ERR_ALLOC) instead of reallocating under it; the graph keeps
    replaying correctly, and once it is destroyed the larger call runs."""
    import gpu_util as gu
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ctx = pa.Context(0)
        x = gu.soa(wl0.x)
        m = gu.soa(wl0.m[:, None])
        c0 = gu.soa(wl0.c.reshape(-1, 12))
        c = c0.clone()
        geom = _binning(pa, wl0)
        bins = pa.bin_sort(ctx, mesh, x, geom)
        shp = tuple(int(v) for v in n)
        gm = pa.GridField(torch.zeros(shp, dtype=torch.float64, device=gu.DEV))
        gmu = pa.GridField(torch.zeros(shp + (3,), dtype=torch.float64, device=gu.DEV))

        def step():
            pa.bin_sort(ctx, mesh, x, geom, out=bins)
            sd = pa.create_spline(0, 2, mesh, x, bins)
            pa.APIC.p2g(ctx, m, c, gm, gmu, sd, zero_first=True)
            pa.APIC.g2p_normalized(ctx, gm, gmu, c, sd)

        step()
        ctx.synchronize()
        ref = (gu.host(gm.tensor).copy(), gu.host(c).copy())
        c.copy_(c0)
        ctx.graph_begin()
        step()
        graph = ctx.graph_end()
        # Eight copies of the particles need a larger sort scratch.
        x8 = gu.soa(np.tile(wl0.x, (8, 1)))
        with pytest.raises(pa.PicError) as ei:
            pa.bin_sort(ctx, mesh, x8, geom)
        assert ei.value.code == -4
        with pytest.raises(pa.PicError):
            ctx.close()  # refused while the graph is alive
        for _ in range(2):
            c.copy_(c0)
            graph.launch()
            ctx.synchronize()
            assert gu.rel_inf(gu.host(gm.tensor), ref[0]) < 1e-13
            assert gu.rel_inf(gu.host(c), ref[1]) < 1e-12
        graph.close()
        big = pa.bin_sort(ctx, mesh, x8, geom)
        ctx.synchronize()
        assert int(gu.host(big.bin_start)[-1]) == 8 * wl0.np
        ctx.close()


def _edge_case(pa, order):
    """One particle at x = 0.01 on a 0.1 mesh whose node window starts at
    node 0: the quadratic / cubic stencil reaches node -1, outside the
    window, while the particle itself lies inside the bins."""
    import gpu_util as gu
    x = gu.dev([[0.01, 0.45, 0.45]])
    mesh = pa.Mesh([0.0, 0.0, 0.0], 0.1)
    dual = pa.aligned_dual(order, pa.FieldLocation.Node)
    bins = pa.bin_sort(gu.ctx(), mesh, x, pa.Binning((0, 0, 0), (8, 8, 8), (4, 4, 4), dual))
    gu.ctx().synchronize()
    return mesh, x, bins


@pytest.mark.parametrize("kind", ["apic", "apic_normalized", "polypic"])
def test_binned_g2p_out_of_window_is_reported(pa, kind):
    """The brick g2p kernels stage out-of-window footprint nodes as zeros; a
    particle whose stencil reaches them is skipped and reported, as on the
    per-particle path (PIC_ERR_BOUNDS from pic_synchronize)."""
    import gpu_util as gu
    order = 2
    mesh, x, bins = _edge_case(pa, order)
    sd = pa.create_spline(0, order, mesh, x, bins)
    g3 = pa.GridField(gu.dev(np.ones((10, 10, 10, 3))))
    if kind == "apic":
        pa.APIC.g2p(gu.ctx(), g3, gu.dev(np.zeros((1, 12))), sd)
    elif kind == "apic_normalized":
        pa.APIC.g2p_normalized(gu.ctx(), pa.GridField(gu.dev(np.ones((10, 10, 10)))), g3,
                               gu.dev(np.zeros((1, 12))), sd)
    else:
        nm = pa.PolyPIC.nmode(order)
        pa.PolyPIC.g2p(gu.ctx(), g3, gu.dev(np.zeros((1, nm * 3))), sd)
    with pytest.raises(pa.PicError) as ei:
        gu.ctx().synchronize()
    assert ei.value.code == -2
    gu.ctx().synchronize()  # the latched error was cleared


def test_wrapper_rejects_mismatched_fields(pa, wl0):
    """Fields whose dtype, row count or component count does not fit the
    call are rejected in the host mirror before anything is launched."""
    import gpu_util as gu
    P = 64
    x = gu.soa(wl0.x[:P])
    mesh = pa.Mesh(wl0.low, wl0.h)
    lo, n = wl0.node_window
    shp = tuple(int(v) for v in n)
    sd = pa.create_spline(0, 2, mesh, x)
    gm = pa.GridField(gu.dev(np.zeros(shp)))
    gmu = pa.GridField(gu.dev(np.zeros(shp + (3,))))
    m = gu.soa(wl0.m[:P, None])
    c = gu.soa(wl0.c[:P].reshape(-1, 12))
    with pytest.raises(TypeError):
        pa.APIC.p2g(gu.ctx(), m, c.float(), gm, gmu, sd)
    with pytest.raises(ValueError):
        pa.APIC.p2g(gu.ctx(), m[:P - 1], c, gm, gmu, sd)
    with pytest.raises(ValueError):
        pa.APIC.p2g(gu.ctx(), m, c, gm, pa.GridField(gu.dev(np.zeros(shp + (1,)))), sd)
    with pytest.raises(TypeError):
        pa.APIC.g2p(gu.ctx(), pa.GridField(gu.dev(np.zeros(shp + (3,)), torch.float32)), c, sd)
    with pytest.raises(ValueError):
        pa.PolyPIC.g2p(gu.ctx(), gmu, gu.dev(np.zeros((P, 8 * 3))), sd)  # 27 modes at order 2
    geom = _binning(pa, wl0)
    nb = pa.bin_count(geom)
    i32 = dict(dtype=torch.int32, device=gu.DEV)
    short = pa.BinSort(torch.empty(P - 1, **i32), torch.empty(P, **i32),
                       torch.empty(nb + 1, **i32), geom, torch.empty(nb + 1, **i32))
    with pytest.raises(ValueError):
        pa.bin_sort(gu.ctx(), mesh, x, geom, out=short)
    few = pa.BinSort(torch.empty(P, **i32), torch.empty(P, **i32),
                     torch.empty(nb, **i32), geom, None)
    with pytest.raises(ValueError):
        pa.bin_sort(gu.ctx(), mesh, x, geom, out=few)
    assert gu.host(gm.tensor).sum() == 0.0 and gu.host(gmu.tensor).sum() == 0.0
    gu.ctx().synchronize()


# ---------------------------------------------------------------------------
# Cabana AoSoA slices: vector length 16, members (x, m, c_p) as Picasso's
# APIC particle list holds them; a ragged last tile (P not a multiple of V).
V = 16


def _aosoa_apic(pa, wl, tdt, P):
    import gpu_util as gu
    x = gu.dev(wl.x[:P], tdt)
    m = gu.dev(wl.m[:P, None], tdt)
    c = gu.dev(wl.c[:P].reshape(-1, 12), tdt)
    aos = pa.AoSoA.from_fields([x, m, c], V)
    return aos, aos.slice(0), aos.slice(1), aos.slice(2)


def test_aosoa_slices_match_their_fields(pa, wl0):
    """The slice formula reads back exactly the packed fields, and pic_permute
    gathers AoSoA members like torch indexing."""
    import gpu_util as gu
    P = 1000
    aos, xs, ms, cs = _aosoa_apic(pa, wl0, torch.float64, P)
    assert torch.equal(xs.to_tensor(), gu.dev(wl0.x[:P]))
    assert torch.equal(cs.to_tensor(), gu.dev(wl0.c[:P].reshape(-1, 12)))
    perm = torch.as_tensor(np.random.default_rng(3).permutation(P).astype(np.int32),
                           device=gu.DEV)
    dst = pa.AoSoA.from_fields([torch.zeros((P, 3), dtype=torch.float64, device=gu.DEV),
                                torch.zeros((P, 1), dtype=torch.float64, device=gu.DEV),
                                torch.zeros((P, 12), dtype=torch.float64, device=gu.DEV)], V)
    pa.permute(gu.ctx(), perm, [(xs, dst.slice(0)), (cs, dst.slice(2))])
    gu.ctx().synchronize()
    idx = perm.long()
    assert torch.equal(dst.slice(0).to_tensor(), xs.to_tensor()[idx])
    assert torch.equal(dst.slice(2).to_tensor(), cs.to_tensor()[idx])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("binned", [False, True])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_aosoa_apic_step_vs_oracle(pa, wl0, dtype, binned, order):
    """APIC p2g and g2p with x, m and c_p as members of one V = 16 AoSoA,
    through the brick kernels (binned) and the per-particle kernels."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    P = 40000 + 7
    xh, mh, ch = wl0.x[:P], wl0.m[:P], wl0.c[:P]
    if f32:
        xh, mh, ch = (a.astype(np.float32).astype(np.float64) for a in (xh, mh, ch))
    aos, xs, ms, cs = _aosoa_apic(pa, wl0, tdt, P)
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    bins = pa.bin_sort(gu.ctx(), mesh, xs, _binning(pa, wl0, order)) if binned else None
    sd = pa.create_spline(0, order, mesh, xs, bins)
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV))
    gmu = pa.GridField(torch.zeros(shp + (3,), dtype=tdt, device=gu.DEV))
    pa.APIC.p2g(gu.ctx(), ms, cs, gm, gmu, sd, zero_first=True)
    rng = np.random.default_rng(order)
    gfield = rng.normal(size=shp + (3,))
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    pa.APIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cs, sd)
    gu.ctx().synchronize()
    rm = np.zeros(shp)
    rmu = np.zeros(shp + (3,))
    orpic.apic_p2g(order, orpic.NODE, wl0.low, wl0.h, lo, n, xh, mh, ch, rm, rmu)
    rc = orpic.apic_g2p(order, orpic.NODE, wl0.low, wl0.h, lo, n, xh, gfield)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    got = gu.host(cs.to_tensor()).reshape(-1, 4, 3)
    assert gu.rel_inf(got[:, 0], rc[:, 0]) < tol
    assert gu.rel_inf(got[:, 1:], rc[:, 1:]) < tol * 10
    # the other members are untouched by g2p
    assert torch.equal(xs.to_tensor(), gu.dev(xh, tdt))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_aosoa_polypic_vs_oracle(pa, wl0, dtype):
    """PolyPIC (cubic, 27 modes) with x, m, u_p and c_p as members of one
    V = 16 AoSoA, through the brick kernels."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    f32 = dtype == "f32"
    order = 3
    P = 20000 + 3
    nm = orpic.polypic_nmode(order)
    rng = np.random.default_rng(77)
    x, m = wl0.x[:P].copy(), wl0.m[:P].copy()
    u = rng.normal(0, 0.5, size=(P, nm, 3))
    cm = rng.normal(size=(P, nm, 3))
    if f32:
        x, m, u, cm = (a.astype(np.float32).astype(np.float64) for a in (x, m, u, cm))
    aos = pa.AoSoA.from_fields([gu.dev(x, tdt), gu.dev(m[:, None], tdt),
                                gu.dev(u.reshape(P, -1), tdt), gu.dev(cm.reshape(P, -1), tdt)],
                               V)
    xs, ms, us, cs = (aos.slice(k) for k in range(4))
    lo, n = wl0.node_window
    mesh = pa.Mesh(wl0.low, wl0.h)
    bins = pa.bin_sort(gu.ctx(), mesh, xs, _binning(pa, wl0, order))
    sd = pa.create_spline(0, order, mesh, xs, bins)
    shp = tuple(int(v) for v in n)
    gmu = pa.GridField(torch.zeros(shp + (3,), dtype=tdt, device=gu.DEV))
    gm = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV))
    dt = 1e-3
    pa.PolyPIC.p2g(gu.ctx(), ms, us, cs, gmu, gm, dt, sd, zero_first=True)
    gfield = rng.normal(size=shp + (3,))
    if f32:
        gfield = gfield.astype(np.float32).astype(np.float64)
    pa.PolyPIC.g2p(gu.ctx(), pa.GridField(gu.dev(gfield, tdt)), cs, sd)
    gu.ctx().synchronize()
    rmu = np.zeros(shp + (3,))
    rm = np.zeros(shp)
    orpic.polypic_p2g(order, orpic.NODE, wl0.low, wl0.h, lo, n, x, m, u, cm, dt, rmu, rm)
    rc = orpic.polypic_g2p(order, orpic.NODE, wl0.low, wl0.h, lo, n, x, gfield, 3)
    tol = F32_TOL if f32 else F64_TOL
    assert gu.rel_inf(gu.host(gm.tensor), rm) < tol
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < tol
    got = gu.host(cs.to_tensor()).reshape(P, nm, 3)
    for lo_r, hi_r in ((0, 1), (1, 4), (4, nm)):
        assert gu.rel_inf(got[:, lo_r:hi_r], rc[:, lo_r:hi_r]) < tol * 10


def test_cpp_facade_drives_a_step_on_the_gpu():
    """The C++ facade program (tests/cpp/facade_step.cpp, built by build())
    runs sort -> p2g -> normalised g2p on the GPU with AoSoA particles and
    checks every field against the oracle; it runs as a child process."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_step")
    assert os.path.exists(exe), "tests/cpp/facade_step is built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rel err" in r.stdout
