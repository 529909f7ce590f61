"""The HIP kernels against the reference's Mathematica known answers.

Same set-up and tolerance (1e-11 absolute, FP64) as unit_test/tstAPIC.hpp
and unit_test/tstPolyPIC.hpp, run through the C ABI on the GPU.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import golden_cases as gc  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa_mod():
    import picasso_amd as pa
    return pa


def _grid(arr):
    from gpu_util import dev
    import picasso_amd as pa
    return pa.GridField(dev(arr), tuple(int(v) for v in gc.WIN_LO))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_apic_collocated_golden(pa_mod, order):
    from gpu_util import ctx, dev, host
    pa = pa_mod
    gold = gc.APIC["orders"][str(order)]
    ent = gc.CELL if order == 2 else gc.NODE
    mesh = pa.Mesh(gc.local_low(), gc.H)
    x = dev(gc.PX[None, :])
    sd = pa.create_spline(ent, order, mesh, x)
    c = dev(np.zeros((1, 12)))
    pa.APIC.g2p(ctx(), _grid(gc.grid_vector()), c, sd)
    ch = host(c).reshape(4, 3)
    np.testing.assert_allclose(ch[0], gold["particle_u"], rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(ch[1:].T, gold["particle_b"], rtol=0, atol=gc.EPS)

    gm = _grid(np.zeros(tuple(gc.WIN_N)))
    gmu = _grid(np.zeros(tuple(gc.WIN_N) + (3,)))
    pa.APIC.p2g(ctx(), dev([[gc.PM]]), c, gm, gmu, sd)
    ctx().synchronize()
    np.testing.assert_allclose(gc.window_values(host(gmu.tensor), order), gold["grid_momentum"],
                               rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(host(gm.tensor), order), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_apic_staggered_golden(pa_mod, order, dim):
    from gpu_util import ctx, dev, host
    pa = pa_mod
    gold = gc.APIC["orders"][str(order)]
    ent, low = gc.apic_staggered_setup(order, dim)
    sd = pa.create_spline(ent, order, pa.Mesh(low, gc.H), dev(gc.PX[None, :]))
    c = dev(np.full((1, 12), 7.0))  # g2p must overwrite everything
    pa.APIC.g2p(ctx(), _grid(gc.grid_function(dim)), c, sd)
    ch = host(c).reshape(4, 3)
    assert ch[0, dim] == pytest.approx(gold["particle_u"][dim], abs=gc.EPS)
    np.testing.assert_allclose(ch[1:, dim], gold["particle_b"][dim], rtol=0, atol=gc.EPS)
    others = [d for d in range(3) if d != dim]
    assert np.all(ch[:, others] == 0.0)

    gm = _grid(np.zeros(tuple(gc.WIN_N)))
    gs = _grid(np.zeros(tuple(gc.WIN_N)))
    pa.APIC.p2g(ctx(), dev([[gc.PM]]), c, gm, gs, sd)
    ctx().synchronize()
    np.testing.assert_allclose(gc.window_values(host(gs.tensor), order),
                               np.array(gold["grid_momentum"])[:, dim], rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(host(gm.tensor), order), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)


def test_polypic_collocated_golden(pa_mod):
    from gpu_util import ctx, dev, host
    pa = pa_mod
    gold = gc.POLY["orders"]["1"]
    sd = pa.create_spline(gc.NODE, 1, pa.Mesh(gc.local_low(), gc.H), dev(gc.PX[None, :]))
    modes = dev(np.zeros((1, 24)))
    pa.PolyPIC.g2p(ctx(), _grid(gc.grid_vector()), modes, sd)
    np.testing.assert_allclose(host(modes).reshape(8, 3), gold["particle_modes"], rtol=0,
                               atol=gc.EPS)
    gm = _grid(np.zeros(tuple(gc.WIN_N)))
    gmu = _grid(np.zeros(tuple(gc.WIN_N) + (3,)))
    pa.PolyPIC.p2g(ctx(), dev([[gc.PM]]), modes, modes, gmu, gm, gc.DT, sd)
    ctx().synchronize()
    np.testing.assert_allclose(gc.window_values(host(gmu.tensor), 1), gold["grid_momentum"],
                               rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(host(gm.tensor), 1), gold["grid_mass"], rtol=0,
                               atol=gc.EPS)


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_polypic_staggered_golden(pa_mod, dim):
    from gpu_util import ctx, dev, host
    pa = pa_mod
    gold = gc.POLY["orders"]["1"]
    ent, low = gc.polypic_staggered_setup(dim)
    sd = pa.create_spline(ent, 1, pa.Mesh(low, gc.H), dev(gc.PX[None, :]))
    modes = dev(np.full((1, 24), 5.0))
    pa.PolyPIC.g2p(ctx(), _grid(gc.grid_function(dim)), modes, sd)
    mh = host(modes).reshape(8, 3)
    np.testing.assert_allclose(mh[:, dim], np.array(gold["particle_modes"])[:, dim], rtol=0,
                               atol=gc.EPS)
    # staggered PolyPIC g2p leaves the other columns alone (Picasso_PolyPIC.hpp:328)
    assert np.all(mh[:, [d for d in range(3) if d != dim]] == 5.0)
    modes = dev(np.array(gold["set_particle_modes"]).reshape(1, 24))
    gm = _grid(np.zeros(tuple(gc.WIN_N)))
    gs = _grid(np.zeros(tuple(gc.WIN_N)))
    pa.PolyPIC.p2g(ctx(), dev([[gc.PM]]), modes, modes, gs, gm, gc.DT, sd)
    ctx().synchronize()
    np.testing.assert_allclose(gc.window_values(host(gs.tensor), 1),
                               np.array(gold["grid_momentum"])[:, dim], rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(host(gm.tensor), 1), gold["grid_mass"], rtol=0,
                               atol=gc.EPS)
