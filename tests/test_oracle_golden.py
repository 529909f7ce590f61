"""Pin the CPU oracle to the reference's Mathematica known answers.

Every check here mirrors an EXPECT_NEAR of unit_test/tstAPIC.hpp or
unit_test/tstPolyPIC.hpp at the reference's tolerance (1e-11 absolute).
"""
import numpy as np
import pytest

import golden_cases as gc
from oracle import orpic


@pytest.mark.parametrize("order", [1, 2, 3])
def test_grid_function_matches_fixture(order):
    """tstAPIC.hpp:594-599: the grid function itself against Mathematica."""
    gv = gc.grid_vector()
    want = np.array(gc.APIC["orders"][str(order)]["grid_velocity"])
    got = gc.window_values(gv, order)
    np.testing.assert_allclose(got, want, rtol=0, atol=gc.EPS)


def _collocated_entity(order):
    # tstAPIC.hpp:856,869,882: Node for orders 1 and 3, Cell for order 2.
    return gc.CELL if order == 2 else gc.NODE


@pytest.mark.parametrize("order", [1, 2, 3])
def test_apic_collocated(order):
    gold = gc.APIC["orders"][str(order)]
    ent = _collocated_entity(order)
    low = gc.local_low()
    x = gc.PX[None, :]
    c = orpic.apic_g2p(order, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, gc.grid_vector())
    np.testing.assert_allclose(c[0, 0, :], gold["particle_u"], rtol=0, atol=gc.EPS)
    # c(1+a, d) = B(d, a); mathematica_b[d][a]
    np.testing.assert_allclose(c[0, 1:, :].T, gold["particle_b"], rtol=0, atol=gc.EPS)

    gm = np.zeros(tuple(gc.WIN_N))
    gmu = np.zeros(tuple(gc.WIN_N) + (3,))
    orpic.apic_p2g(order, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, [gc.PM], c, gm, gmu)
    np.testing.assert_allclose(gc.window_values(gmu, order), gold["grid_momentum"],
                               rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(gm, order), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)
    # Nothing is written outside the stencil.
    a, b = gc.stencil_range(order)
    mask = np.ones(tuple(gc.WIN_N), bool)
    sl = tuple(slice(gc.CX[d] + a - gc.WIN_LO[d], gc.CX[d] + b - gc.WIN_LO[d]) for d in range(3))
    mask[sl] = False
    assert np.all(gm[mask] == 0.0)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("dim", [0, 1, 2])
def test_apic_staggered(order, dim):
    gold = gc.APIC["orders"][str(order)]
    ent, low = gc.apic_staggered_setup(order, dim)
    x = gc.PX[None, :]
    gs = gc.grid_function(dim)  # tstAPIC.hpp:753-760 uses (Dim+1)
    c = orpic.apic_g2p(order, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, gs)
    assert c[0, 0, dim] == pytest.approx(gold["particle_u"][dim], abs=gc.EPS)
    np.testing.assert_allclose(c[0, 1:, dim], gold["particle_b"][dim], rtol=0, atol=gc.EPS)
    # Staggered g2p zeroes the other components (Picasso_APIC.hpp:505-543).
    other = [d for d in range(3) if d != dim]
    assert np.all(c[0, :, other] == 0.0)

    gm = np.zeros(tuple(gc.WIN_N))
    gmu = np.zeros(tuple(gc.WIN_N))
    orpic.apic_p2g(order, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, [gc.PM], c, gm, gmu)
    want_mu = np.array(gold["grid_momentum"])[:, dim]
    np.testing.assert_allclose(gc.window_values(gmu, order), want_mu, rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(gm, order), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)


def test_polypic_linear_collocated():
    gold = gc.POLY["orders"]["1"]
    low = gc.local_low()
    x = gc.PX[None, :]
    modes = orpic.polypic_g2p(1, gc.NODE, low, gc.H, gc.WIN_LO, gc.WIN_N, x,
                              gc.grid_vector(), 3)
    np.testing.assert_allclose(modes[0], gold["particle_modes"], rtol=0, atol=gc.EPS)
    gm = np.zeros(tuple(gc.WIN_N))
    gmu = np.zeros(tuple(gc.WIN_N) + (3,))
    # tstPolyPIC.hpp:316: p2g(pm, modes, modes, ...) with the g2p output.
    orpic.polypic_p2g(1, gc.NODE, low, gc.H, gc.WIN_LO, gc.WIN_N, x, [gc.PM],
                      modes, modes, gc.DT, gmu, gm)
    np.testing.assert_allclose(gc.window_values(gmu, 1), gold["grid_momentum"],
                               rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(gm, 1), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)


@pytest.mark.parametrize("dim", [0, 1, 2])
def test_polypic_linear_staggered(dim):
    gold = gc.POLY["orders"]["1"]
    ent, low = gc.polypic_staggered_setup(dim)
    x = gc.PX[None, :]
    gs = gc.grid_function(dim)
    modes = orpic.polypic_g2p(1, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, gs, 3)
    np.testing.assert_allclose(modes[0, :, dim], np.array(gold["particle_modes"])[:, dim],
                               rtol=0, atol=gc.EPS)
    # tstPolyPIC.hpp:458-459: reset all modes to the Mathematica values.
    modes = np.array(gold["set_particle_modes"])[None]
    gm = np.zeros(tuple(gc.WIN_N))
    gmu = np.zeros(tuple(gc.WIN_N))
    orpic.polypic_p2g(1, ent, low, gc.H, gc.WIN_LO, gc.WIN_N, x, [gc.PM],
                      modes, modes, gc.DT, gmu, gm)
    np.testing.assert_allclose(gc.window_values(gmu, 1),
                               np.array(gold["grid_momentum"])[:, dim], rtol=0, atol=gc.EPS)
    np.testing.assert_allclose(gc.window_values(gm, 1), gold["grid_mass"],
                               rtol=0, atol=gc.EPS)
