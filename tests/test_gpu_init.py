"""Particle initialisation and stable selection on the GPU against the
oracle (Picasso_ParticleInit.hpp:38-91 filterEmpties, :121-248 InitRandom,
:279-400 InitUniform)."""
import ctypes

import numpy as np
import pytest

from oracle import orpic

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

LOW = np.array([-0.25, -0.125, -0.375])
H = 0.0625


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.mark.parametrize("init, ppc", [(0, 1), (0, 2), (0, 3), (1, 1), (1, 8), (1, 27)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_init_particles_matches_oracle(pa, init, ppc, dtype):
    """Positions and volumes bit for bit: the kernel computes in FP64 with
    the oracle's operations and rounds once to the storage type."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    cell_lo, ncell = (2, 2, 2), (7, 5, 9)
    x, vol = pa.initialize_particles(gu.ctx(), init, ppc, pa.Mesh(LOW, H), cell_lo, ncell,
                                     dtype=tdt, seed=1234)
    gu.ctx().synchronize()
    rx, rvol = orpic.init_particles(init, LOW, H, cell_lo, ncell, ppc, seed=1234)
    cast = (lambda a: a) if dtype == "f64" else (lambda a: a.astype(np.float32))
    np.testing.assert_array_equal(gu.host(x), cast(rx))
    np.testing.assert_array_equal(gu.host(vol)[:, 0], cast(rvol))


def test_init_particles_capacity_and_arguments(pa):
    import gpu_util as gu
    from picasso_amd import _abi
    x = torch.empty((3, 10), dtype=torch.float64, device=gu.DEV).t()
    xf = pa._pfield(x, 3, "x")
    i3 = ctypes.c_int32 * 3
    mesh = pa.Mesh(LOW, H).c()
    call = lambda mode, ncell, ppc, cap: _abi.lib.pic_init_particles(  # noqa: E731
        gu.ctx()._bind_stream(), mode, _abi.F64, ctypes.byref(mesh), i3(0, 0, 0), i3(*ncell),
        ppc, 0, cap, ctypes.byref(xf), None)
    assert call(0, (2, 1, 1), 2, 10) == _abi.PIC_ERR_CAPACITY  # 16 particles > 10
    assert call(1, (2, 1, 1), 5, 10) == 10
    assert call(2, (2, 1, 1), 5, 10) == _abi.PIC_ERR_ARG
    assert call(1, (0, 1, 1), 5, 10) == _abi.PIC_ERR_ARG
    assert call(1, (1, 1, 1), 0, 10) == _abi.PIC_ERR_ARG
    gu.ctx().synchronize()


@pytest.mark.parametrize("n, p", [(0, 0.5), (1, 1.0), (1000, 0.0), (1000, 1.0), (77777, 0.3),
                                  (5_000_000, 0.6)])
def test_select_is_stable(pa, n, p):
    """The kept indices are np.flatnonzero of the mask, in order (the last
    case spans more than 2048 scan tiles, the two-level scan)."""
    import gpu_util as gu
    rng = np.random.default_rng(n)
    keep = rng.random(n) < p
    idx = pa.select(gu.ctx(), torch.as_tensor(keep, device=gu.DEV))
    np.testing.assert_array_equal(gu.host(idx).astype(np.int64), np.flatnonzero(keep))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_initialize_and_filter_like_a_creation_functor(pa, dtype):
    """initializeParticles with a functor that creates particles inside a
    sphere only: fill the box, mark the sphere, compact -- the kept
    particles are the oracle's in the same order, with their volumes."""
    import gpu_util as gu
    tdt = torch.float64 if dtype == "f64" else torch.float32
    cell_lo, ncell = (2, 2, 2), (16, 16, 16)
    x, vol = pa.initialize_particles(gu.ctx(), pa.InitRandom, 8, pa.Mesh(LOW, H), cell_lo,
                                     ncell, dtype=tdt, seed=99)
    centre = LOW + H * (np.array(cell_lo) + np.array(ncell) / 2)
    keep = ((x - torch.as_tensor(centre, dtype=tdt, device=gu.DEV)) ** 2).sum(1) < (5 * H) ** 2
    xk, vk = pa.filter_particles(gu.ctx(), keep, x, vol)
    gu.ctx().synchronize()
    rx, rvol = orpic.init_particles(1, LOW, H, cell_lo, ncell, 8, seed=99)
    cast = (lambda a: a) if dtype == "f64" else (lambda a: a.astype(np.float32))
    rkeep = gu.host(keep).astype(bool)
    assert 0 < rkeep.sum() < len(rkeep)
    np.testing.assert_array_equal(gu.host(xk), cast(rx)[rkeep])
    np.testing.assert_array_equal(gu.host(vk)[:, 0], cast(rvol)[rkeep])


@pytest.mark.parametrize("init, ppc, boundary, multiplier",
                         [(1, 17, 1, 1), (1, 9, 3, 1), (0, 3, 1, 1), (0, 2, 2, 1),
                          (0, 3, 1, 3), (0, 2, 3, 3), (1, 11, 1, 4), (1, 4, 4, 3)])
def test_reference_particle_init_test(pa, init, ppc, boundary, multiplier):
    """tstParticleInit.hpp:46-200 on this library: a 14 x 17 x 19 mesh of 0.23
    cells at (1.2, 3.3, -2.8) with two ghost layers
    (particle_init_test.json); the creation functor keeps particles strictly
    inside the box `boundary` cells in from every face; initialising
    `multiplier` times (a new seed each time) appends.  Expected: exactly
    multiplier * n * (cells left in the box) particles, all strictly inside
    the box, each with volume h^3 / n (n = ppc^3 uniform, ppc random)."""
    import gpu_util as gu
    h = 0.23
    nc = np.array([14, 17, 19])
    glow = np.array([1.2, 3.3, -2.8])
    ghigh = glow + h * nc
    mesh = pa.Mesh(glow - 2 * h, h)
    lo_box, hi_box = glow + h * boundary, ghigh - h * boundary
    xs, vs = [], []
    for it in range(multiplier):
        x, vol = pa.initialize_particles(gu.ctx(), init, ppc, mesh, (2, 2, 2), nc,
                                         dtype=torch.float64, seed=17 + it)
        lo_t = torch.as_tensor(lo_box, device=gu.DEV)
        hi_t = torch.as_tensor(hi_box, device=gu.DEV)
        keep = ((x > lo_t) & (x < hi_t)).all(1)
        xk, vk = pa.filter_particles(gu.ctx(), keep, x, vol)
        xs.append(gu.host(xk))
        vs.append(gu.host(vk)[:, 0])
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    n = ppc ** 3 if init == 0 else ppc
    assert len(x) > 0
    assert len(x) == multiplier * n * int(np.prod(nc - 2 * boundary))
    assert np.all(x > lo_box) and np.all(x < hi_box)
    np.testing.assert_allclose(v, h ** 3 / n, rtol=1e-15)
