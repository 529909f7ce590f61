"""PIC_P2G_DETERMINISTIC: APIC p2g sums that are bit-identical from run to
run (brick-colour passes without atomics, stale particles one after another
in slot order), checked against the default atomic mode and the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.fixture(scope="module")
def wl0():
    from picasso_amd import synth
    return synth.apic_uniform(32, 8)


def _binning(pa, wl, order, entity):
    dual = pa.aligned_dual(order, entity)
    flags = pa.dual_dims(dual)
    cell_lo, ncell = wl.cell_region
    return pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), (4, 4, 4), dual)


def _setup(pa, wl, order, entity, tdt, move=0.0, seed=5):
    """Device particles of wl (FP32-rounded for FP32), bins, and the host
    copies the oracle uses; with `move`, that fraction of the particles is
    displaced by up to 1.5 cells after binning (stale bins)."""
    import gpu_util as gu
    rnd = (lambda a: a.astype(np.float32).astype(np.float64)) if tdt is torch.float32 else (lambda a: a)
    x, m, c = rnd(wl.x.copy()), rnd(wl.m), rnd(wl.c)
    mesh = pa.Mesh(wl.low, wl.h)
    xt = gu.soa(x, tdt)
    bins = pa.bin_sort(gu.ctx(), mesh, xt, _binning(pa, wl, order, entity))
    if move:
        gu.ctx().synchronize()
        rng = np.random.default_rng(seed)
        sel = rng.random(len(x)) < move
        x[sel] += rng.uniform(-1.5, 1.5, size=(sel.sum(), 3)) * wl.h
        x = rnd(np.clip(x, 0.0, 1.0 - 1e-9))
        xt.copy_(gu.soa(x, tdt))
    sd = pa.create_spline(entity, order, mesh, xt, bins)
    return sd, x, m, c, gu.soa(m[:, None], tdt), gu.soa(c.reshape(-1, 12), tdt)


def _run(pa, wl, sd, md, cd, tdt, stag, deterministic):
    import gpu_util as gu
    lo, n = wl.node_window
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV), tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.zeros(shp + (() if stag else (3,)), dtype=tdt, device=gu.DEV),
                       tuple(int(v) for v in lo))
    pa.APIC.p2g(gu.ctx(), md, cd, gm, gmu, sd, zero_first=True, deterministic=deterministic)
    gu.ctx().synchronize()
    return gu.host(gm.tensor).copy(), gu.host(gmu.tensor).copy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64 if a.dtype == np.float64 else np.uint32)


@pytest.mark.parametrize("move", [0.0, 0.01])
@pytest.mark.parametrize("case", [(2, 0), (3, 0), (1, 0), (2, 3)])  # (order, entity)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_deterministic_p2g_repeats_bit_for_bit(pa, wl0, dtype, case, move):
    """Three deterministic runs, with default-mode runs between them, give
    the same bits; the result agrees with the default mode and the oracle
    (FP64 1e-12, FP32 1e-5).  move = 0.01: about 2,600 particles left their
    bins, so the stale list is sorted and applied one particle at a time."""
    import gpu_util as gu
    order, entity = case
    tdt = torch.float64 if dtype == "f64" else torch.float32
    stag = entity >= 2
    sd, x, m, c, md, cd = _setup(pa, wl0, order, entity, tdt, move)
    runs = []
    for _ in range(3):
        runs.append(_run(pa, wl0, sd, md, cd, tdt, stag, True))
        atomic = _run(pa, wl0, sd, md, cd, tdt, stag, False)
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(_bits(a), _bits(b)), "deterministic p2g differs between runs"
    tol = 1e-12 if dtype == "f64" else 1e-5
    for a, b in zip(runs[0], atomic):
        assert gu.rel_inf(a, b) < tol
    lo, n = wl0.node_window
    rm = np.zeros(tuple(int(v) for v in n))
    rmu = np.zeros(rm.shape + (() if stag else (3,)))
    orpic.apic_p2g(order, entity, wl0.low, wl0.h, lo, n, x, m, c, rm, rmu)
    assert gu.rel_inf(runs[0][0], rm) < tol
    assert gu.rel_inf(runs[0][1], rmu) < tol


def test_deterministic_p2g_without_bins_is_refused(pa, wl0):
    """The per-particle kernel's atomics land in any order: deterministic
    mode needs bins the brick kernels take, and leaves the grid untouched
    when it refuses."""
    import gpu_util as gu
    x = gu.soa(wl0.x)
    mesh = pa.Mesh(wl0.low, wl0.h)
    lo, n = wl0.node_window
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.full(shp, 3.0, dtype=torch.float64, device=gu.DEV), tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.full(shp + (3,), 3.0, dtype=torch.float64, device=gu.DEV),
                       tuple(int(v) for v in lo))
    sd = pa.create_spline(0, 2, mesh, x)
    with pytest.raises(pa.PicError) as e:
        pa.APIC.p2g(gu.ctx(), gu.soa(wl0.m[:, None]), gu.soa(wl0.c.reshape(-1, 12)), gm, gmu, sd,
                    zero_first=True, deterministic=True)
    assert e.value.code == pa._abi.PIC_ERR_UNSUPPORTED
    gu.ctx().synchronize()
    assert float(gm.tensor.min()) == 3.0 and float(gmu.tensor.max()) == 3.0


def test_deterministic_p2g_long_stale_list(pa, wl0):
    """More stale particles than the deterministic mode orders: every
    particle is still applied (the result matches the oracle) and
    synchronisation reports PIC_ERR_CAPACITY."""
    import gpu_util as gu
    sd, x, m, c, md, cd = _setup(pa, wl0, 2, 0, torch.float64, move=0.1, seed=9)
    lo, n = wl0.node_window
    shp = tuple(int(v) for v in n)
    gm = pa.GridField(torch.zeros(shp, dtype=torch.float64, device=gu.DEV), tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.zeros(shp + (3,), dtype=torch.float64, device=gu.DEV),
                       tuple(int(v) for v in lo))
    pa.APIC.p2g(gu.ctx(), md, cd, gm, gmu, sd, zero_first=True, deterministic=True)
    with pytest.raises(pa.PicError) as e:
        gu.ctx().synchronize()
    assert e.value.code == pa._abi.PIC_ERR_CAPACITY
    gu.ctx().synchronize()  # cleared
    rm = np.zeros(shp)
    rmu = np.zeros(shp + (3,))
    orpic.apic_p2g(2, 0, wl0.low, wl0.h, lo, n, x, m, c, rm, rmu)
    assert gu.rel_inf(gu.host(gm.tensor), rm) < 1e-12
    assert gu.rel_inf(gu.host(gmu.tensor), rmu) < 1e-12
