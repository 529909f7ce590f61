"""End to end on the GPU, nothing generated on the host but the field
coefficients: particles initialised in every cell (initializeParticles),
filtered to a sphere (its creation functor, as a keep mask), binned, and
stepped through the brick kernels.  APIC carries an affine velocity field
exactly through p2g -> normalisation -> g2p (Picasso_APIC.hpp:55-140,
407-472), and mass and momentum are conserved."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.mark.parametrize("init, ppc", [("uniform", 2), ("random", 8)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_initialize_filter_sort_step(pa, init, ppc, dtype):
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    tol = 1e-11 if dtype == "f64" else 2e-5
    n, halo = 48, 2
    h = 1.0 / n
    mesh = pa.Mesh(np.full(3, -halo * h), h)  # local node `halo` sits at x = 0
    ctx = gu.ctx()
    mode = pa.InitUniform if init == "uniform" else pa.InitRandom
    x, vol = pa.initialize_particles(ctx, mode, ppc, mesh, (halo,) * 3, (n,) * 3, dtype=tdt,
                                     seed=2024)
    keep = ((x - 0.5) ** 2).sum(1) < 0.3 ** 2  # the creation functor: a sphere
    x, vol = pa.filter_particles(ctx, keep, x, vol)
    P = x.shape[0]
    assert P > 0.9 * (4 / 3 * np.pi * 0.3 ** 3) * n ** 3 * (ppc ** 3 if init == "uniform" else ppc)
    rng = np.random.default_rng(5)
    A = torch.as_tensor(rng.normal(size=(3, 3)), dtype=tdt, device=gu.DEV)
    b = torch.as_tensor(rng.normal(size=3), dtype=tdt, device=gu.DEV)
    D = h * h / 4.0
    m = (vol * 1000.0).contiguous().t().contiguous().t()  # density 1000
    c = torch.empty((12, P), dtype=tdt, device=gu.DEV).t()
    c[:, 0:3] = x @ A.T + b
    c[:, 3:] = (A * D).T.reshape(1, 9)  # c(1 + a, d) = B(d, a) = A[d, a] D
    c_in = c.clone()
    shp = (n + 2 * halo + 1,) * 3
    gm = pa.GridField(torch.zeros(shp, dtype=tdt, device=gu.DEV))
    gmu = pa.GridField(torch.zeros(shp + (3,), dtype=tdt, device=gu.DEV))
    dual = pa.aligned_dual(2, pa.FieldLocation.Node)
    geom = pa.Binning((halo,) * 3, (n + 1,) * 3, (4, 4, 4), dual)
    bins = pa.bin_sort(ctx, mesh, x, geom)
    sd = pa.create_spline(pa.FieldLocation.Node, 2, mesh, x, bins)
    pa.APIC.p2g(ctx, m, c, gm, gmu, sd, zero_first=True)
    pa.APIC.g2p_normalized(ctx, gm, gmu, c, sd)
    ctx.synchronize()
    mh = gu.host(m)[:, 0].astype(np.float64)
    cin = gu.host(c_in).astype(np.float64).reshape(P, 4, 3)
    gmh = gu.host(gm.tensor).astype(np.float64)
    gmuh = gu.host(gmu.tensor).astype(np.float64)
    assert abs(gmh.sum() - mh.sum()) <= tol * mh.sum()
    mom = (mh[:, None] * cin[:, 0, :]).sum(0)
    assert np.abs(gmuh.reshape(-1, 3).sum(0) - mom).max() <= tol * np.abs(mom).max() * 10
    cout = gu.host(c).astype(np.float64).reshape(P, 4, 3)
    scale = np.abs(cin[:, 0]).max()
    assert np.abs(cout[:, 0] - cin[:, 0]).max() <= 10 * tol * scale
    assert np.abs(cout[:, 1:] - cin[:, 1:]).max() <= 100 * tol * scale * D
