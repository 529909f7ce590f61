"""The moving-particle step and its tuner (picasso_amd/tune.py) on the GPU.

Every configuration the tuner chooses between (re-sort interval, re-permute
interval, g2p brick edge, g2p workgroup size) must give the same physics:
four FP64 steps of p2g (grid zeroed first), normalised g2p and x += dt v
on a 32^3 sphere are checked against the oracle running the same steps.
Storage order differs between configurations (the permutes), so particles
are compared after sorting both sides by position, and the last step's
grid (independent of particle order) is compared directly.  Bar: 1e-11
relative (four steps of FP64 with atomic summation order differences)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-11
STEPS = 4


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.fixture(scope="module")
def wl():
    from picasso_amd import synth
    w = synth.apic_sphere(32, 8)
    w.c[:, 1:, :] = 0.0  # B_p = 0 at the start, as in the moving bench leg
    return w


@pytest.fixture(scope="module")
def ref(wl):
    lo, n = wl.node_window
    x, c = wl.x.copy(), wl.c.copy()
    dt = 0.5 * wl.h
    for _ in range(STEPS):
        gm = np.zeros(tuple(n))
        gmu = np.zeros(tuple(n) + (3,))
        orpic.apic_p2g(2, orpic.NODE, wl.low, wl.h, lo, n, x, wl.m, c, gm, gmu)
        c = orpic.apic_g2p(2, orpic.NODE, wl.low, wl.h, lo, n, x, orpic.grid_normalize(gm, gmu))
        x = x + dt * c[:, 0, :]
    return x, c, gm, gmu


def _step(pa, wl, config):
    import gpu_util as gu
    from picasso_amd import tune
    lo, n = wl.node_window
    nn = tuple(int(v) for v in n)
    mesh = pa.Mesh(wl.low, wl.h)
    cell_lo, ncell = wl.cell_region
    lo_c, n_c = tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell)
    geom = pa.Binning(lo_c, n_c, (4, 4, 4), 1)
    geom8 = pa.Binning(lo_c, n_c, (8, 8, 8), 1)
    x = gu.soa(wl.x)
    m = gu.soa(wl.m[:, None])
    c = gu.soa(wl.c.reshape(-1, 12))
    gm = pa.GridField(torch.full(nn, 3.0, dtype=torch.float64, device=gu.DEV),
                      tuple(int(v) for v in lo))
    gmu = pa.GridField(torch.full((3,) + nn, -3.0, dtype=torch.float64,
                                  device=gu.DEV).permute(1, 2, 3, 0), tuple(int(v) for v in lo))
    return tune.MovingStep(pa, gu.ctx(), mesh, geom, geom8, x, m, c, gm, gmu, 0.5 * wl.h, config)


def _by_position(x, c):
    order = np.lexsort((x[:, 2], x[:, 1], x[:, 0]))
    return x[order], c.reshape(len(x), -1)[order]


@pytest.mark.parametrize("cfg", [(1, None, 4, 128), (1, 1, 4, 256), (2, 4, 8, 0),
                                 (1, 2, 8, 128), (2, None, 4, 256)])
def test_moving_step_configs_match_oracle(pa, wl, ref, cfg):
    import gpu_util as gu
    from picasso_amd import tune
    step = _step(pa, wl, tune.StepConfig(*cfg))
    try:
        for _ in range(STEPS):
            step.run()
        gu.ctx().synchronize()
    finally:
        gu.ctx().set_g2p_group(0)
    rx, rc, rm, rmu = ref
    assert gu.rel_inf(gu.host(step.gm.tensor), rm) < TOL
    assert gu.rel_inf(gu.host(step.gmu.tensor), rmu) < TOL
    gx, gc = _by_position(gu.host(step.x), gu.host(step.c))
    ox, oc = _by_position(rx, rc)
    assert gu.rel_inf(gx, ox) < TOL
    assert gu.rel_inf(gc, oc) < TOL


def test_tuner_picks_a_candidate(pa, wl):
    import gpu_util as gu
    from picasso_amd import tune
    step = _step(pa, wl, tune.StepConfig())
    tuner = tune.StepTuner(resort=(1,), permute=(4, None), g2p_brick=(4,), g2p_group=(128, 256))
    try:
        best, table = tuner.tune(step, window=2, warmup=1)
        gu.ctx().synchronize()
    finally:
        gu.ctx().set_g2p_group(0)
    assert len(table) == 4
    assert best in tuner.candidates and step.config == best
    assert all(t["ms_per_step"] > 0 for t in table)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_g2p_group_sizes_agree(pa, wl, dtype):
    """Both g2p workgroup shapes give bit-identical c_p (g2p has no atomics,
    and each particle's arithmetic does not depend on the shape)."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    lo, n = wl.node_window
    nn = tuple(int(v) for v in n)
    mesh = pa.Mesh(wl.low, wl.h)
    ctx = gu.ctx()
    x = gu.soa(wl.x, tdt)
    cell_lo, ncell = wl.cell_region
    geom = pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell),
                      (4, 4, 4), 1)
    bins = pa.bin_sort(ctx, mesh, x, geom)
    sd = pa.create_spline(pa.NODE, 2, mesh, x, bins)
    rng = np.random.default_rng(7)
    gm = pa.GridField(gu.dev(rng.uniform(0.5, 1.5, nn), tdt), tuple(int(v) for v in lo))
    gmu = pa.GridField(gu.dev(rng.normal(size=nn + (3,)), tdt), tuple(int(v) for v in lo))
    out = []
    try:
        for group in (128, 256, 0):
            ctx.set_g2p_group(group)
            c = torch.zeros((wl.np, 12), dtype=tdt, device=gu.DEV)
            pa.APIC.g2p_normalized(ctx, gm, gmu, c, sd)
            ctx.synchronize()
            out.append(gu.host(c))
    finally:
        ctx.set_g2p_group(0)
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_set_option_rejects_bad_values(pa):
    import gpu_util as gu
    with pytest.raises(pa.PicError):
        gu.ctx().set_g2p_group(64)
