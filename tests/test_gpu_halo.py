"""Grid halo on one GPU (GridOperator's gather / scatter-sum,
Picasso_GridOperator.hpp:203-215, 465-491, for periodic dimensions) against
the oracle, and a periodic p2g checked by translation invariance."""
import numpy as np
import pytest

from oracle import orpic

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

LO = np.array([-2, -1, -3])
N = np.array([14, 11, 15])
OLO = np.array([0, 1, 0])
ON = np.array([9, 8, 10])


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1), (0, 1, 0), (0, 0, 0)])
@pytest.mark.parametrize("ncomp", [1, 3])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_halo_matches_oracle(pa, periodic, ncomp, dtype):
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(7 * ncomp + sum(periodic))
    g0 = rng.normal(size=tuple(N) + ((ncomp,) if ncomp > 1 else ()))
    if dtype == "f32":
        g0 = g0.astype(np.float32).astype(np.float64)
    f = pa.GridField(gu.dev(g0, tdt), tuple(int(v) for v in LO))
    pa.halo_scatter(gu.ctx(), f, OLO, ON, periodic)
    gu.ctx().synchronize()
    want = orpic.grid_halo(1, LO, np.ascontiguousarray(g0).copy(), OLO, ON, periodic)
    got = gu.host(f.tensor)
    # Owned in every periodic dimension: the scatter's defined output.
    own = np.ones(tuple(N), bool)
    for d in range(3):
        if periodic[d]:
            ax = np.arange(LO[d], LO[d] + N[d])
            m = (ax >= OLO[d]) & (ax < OLO[d] + ON[d])
            own &= m.reshape([-1 if e == d else 1 for e in range(3)])
    if dtype == "f64":
        np.testing.assert_array_equal(got[own], want[own])
    else:
        np.testing.assert_allclose(got[own], want[own], rtol=1e-6, atol=1e-6)
    pa.halo_gather(gu.ctx(), f, OLO, ON, periodic)
    gu.ctx().synchronize()
    wg = orpic.grid_halo(0, LO, np.ascontiguousarray(got, dtype=np.float64), OLO, ON, periodic)
    np.testing.assert_array_equal(gu.host(f.tensor), wg.astype(got.dtype))


def test_halo_rejects_wide_ghost_layers(pa):
    import gpu_util as gu
    f = pa.GridField(gu.dev(np.zeros((14, 11, 15))), tuple(int(v) for v in LO))
    with pytest.raises(pa.PicError):
        pa.halo_scatter(gu.ctx(), f, OLO, (4, 8, 10))  # 10 ghost layers > 4 owned in i
    with pytest.raises(pa.PicError):
        pa.halo_gather(gu.ctx(), f, (-3, 1, 0), ON)  # owned range leaves the window


@pytest.mark.parametrize("shift", [(1, 0, 0), (3, 5, 0), (0, 0, 15)])
def test_periodic_p2g_is_translation_invariant(pa, shift):
    """Periodic box of 16^3 cells with two ghost layers: p2g into the window,
    scatter-sum halo, then the owned grid.  Moving every particle by whole
    cells (wrapped into the box) rolls the owned mass and momentum by the
    same number of nodes, and the owned mass adds up to the particles'."""
    import gpu_util as gu
    n, g = 16, 2
    h = 1.0 / n
    rng = np.random.default_rng(sum(shift))
    P = 5000
    x = rng.uniform(0, 1, size=(P, 3))
    m = rng.uniform(0.5, 1.5, P)
    c = rng.normal(size=(P, 4, 3))
    mesh = pa.Mesh(np.full(3, -g * h), h)  # local node g is the box's node 0
    lo = (0, 0, 0)
    nw = (n + 2 * g,) * 3

    def owned_grids(xp):
        sd = pa.create_spline(0, 2, mesh, gu.dev(xp))
        gm = pa.GridField(gu.dev(np.zeros(nw)), lo)
        gmu = pa.GridField(gu.dev(np.zeros(nw + (3,))), lo)
        pa.APIC.p2g(gu.ctx(), gu.dev(m[:, None]), gu.dev(c.reshape(-1, 12)), gm, gmu, sd,
                    zero_first=True)
        for f in (gm, gmu):
            pa.halo_scatter(gu.ctx(), f, (g, g, g), (n, n, n))
        gu.ctx().synchronize()
        sl = (slice(g, g + n),) * 3
        return gu.host(gm.tensor)[sl], gu.host(gmu.tensor)[sl]

    m0, mu0 = owned_grids(x)
    xs = np.mod(x + np.array(shift) * h, 1.0)
    m1, mu1 = owned_grids(xs)
    np.testing.assert_allclose(m1, np.roll(m0, shift, axis=(0, 1, 2)), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(mu1, np.roll(mu0, shift, axis=(0, 1, 2)), rtol=1e-12,
                               atol=1e-12)
    assert abs(m0.sum() - m.sum()) < 1e-11 * m.sum()
