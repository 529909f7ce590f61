"""The other full-size workloads, checked against the oracle on a box of
nodes and the particles around it:

- the cube [1/4, 3/4]^3 of a 256^3 mesh at 8 ppc (16,777,216 particles),
  FP32 particles and grid with the FP64 deformation-gradient update, ten
  consecutive steps;
- PolyPIC on 128^3 cells at 8 ppc (16,777,216 particles), cubic splines,
  27 velocity modes per component, FP32.

The whole grid cannot be formed by the oracle in the time of a test, but a
box of nodes can: every particle whose stencil reaches the box is given to
the oracle, so the oracle's sums at the box nodes are complete and must
match the GPU's (p2g).  The particles whose stencils lie inside the box
gather only box nodes, so they are checked against the oracle's g2p on the
GPU's own grid (g2p, velocity gradient, F update).  The boxes straddle the
edge of the particle region, where nodes go from full to empty.  Inputs are
drawn with numpy's generator (the shapes and distributions of the
workloads; the values do not matter for parity).  FP32 within 1e-5 of the
box's largest value (1e-4 for the affine and higher modes); the FP64 F
update and det F within 1e-6 (the gradient carries FP32 rounding, scaled by
dt), F^-1 within what its condition makes of that."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

HALO = 2


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _cells(lo, hi):
    ii = np.arange(lo, hi)
    return np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)


def _positions(rng, cells, ppc, h):
    """ppc uniform FP32 positions in each cell, in shuffled order."""
    x = (np.repeat(cells, ppc, axis=0) + rng.random((len(cells) * ppc, 3))) * h
    x = x[rng.permutation(len(x))]
    return x.astype(np.float32)


def _box_sets(x, low, h, order, L, U):
    """Particles whose stencil reaches nodes [L, U)^3 (touch) and whose
    stencil lies inside it (inside), by the knot base of the spline."""
    t = (x - low) / h
    base = np.floor(t + 0.5) if order == 2 else np.floor(t)
    first = base.astype(np.int64) - (0 if order == 1 else 1)
    last = first + order
    touch = np.all((last >= L) & (first < U), axis=1)
    inside = np.all((first >= L) & (last < U), axis=1)
    return np.nonzero(touch)[0], np.nonzero(inside)[0]


def _node_geom(pa, n):
    # Bins are the node-centred dual cells; n + 1 per dimension, since an
    # FP32 position in the last cell may round to exactly 1.0.
    return pa.Binning((HALO,) * 3, (n + 1,) * 3, (4, 4, 4), 1)


def test_cube_f_update_full_size(pa):
    import gpu_util as gu
    n, ppc, dt = 256, 8, 1e-3
    h = 1.0 / n
    low = np.full(3, -HALO * h)
    rng = np.random.default_rng(0x5EED0004)
    x = _positions(rng, _cells(n // 4, 3 * n // 4), ppc, h)
    P = len(x)
    assert P == 16_777_216
    c = rng.standard_normal((P, 12), dtype=np.float32)
    c[:, 3:] *= np.float32(0.1)
    F0 = np.eye(3)[None] + 0.01 * rng.standard_normal((P, 3, 3))
    m = np.full(P, h ** 3 / ppc, np.float32)
    nn = n + 1 + 2 * HALO
    f32 = torch.float32
    mesh = pa.Mesh(low, h)
    xd, md, cd = gu.soa(x, f32), gu.soa(m[:, None], f32), gu.soa(c, f32)
    F = gu.soa(F0.reshape(P, 9))
    det = gu.soa(np.zeros((P, 1)))
    Finv = gu.soa(np.zeros((P, 9)))
    gm = pa.GridField(torch.zeros((nn,) * 3, dtype=f32, device=gu.DEV))
    gmu = pa.GridField(torch.zeros((3,) + (nn,) * 3, dtype=f32, device=gu.DEV).permute(1, 2, 3, 0))
    geom = _node_geom(pa, n)

    def step():
        bins = pa.bin_sort(gu.ctx(), mesh, xd, geom)
        sd = pa.create_spline(pa.FieldLocation.Node, 2, mesh, xd, bins)
        pa.APIC.p2g(gu.ctx(), md, cd, gm, gmu, sd, zero_first=True)
        return sd

    sd = step()
    gu.ctx().synchronize()
    # The box crosses the cube's low faces (node HALO + n/4).
    L = HALO + n // 4 - 8
    U = L + 16
    lo_w, n_w = L - 2, U - L + 4
    win = slice(lo_w, lo_w + n_w)
    gmh = gu.host(gm.tensor[win, win, win]).astype(np.float64)
    gmuh = gu.host(gmu.tensor[win, win, win]).astype(np.float64)
    xh = x.astype(np.float64)
    touch, inside = _box_sets(xh, low, h, 2, L, U)
    # one octant of the box is inside the cube
    assert len(inside) > 1000 and (gmh == 0).any() and (gmh > 0).any()
    rm = np.zeros((n_w,) * 3)
    rmu = np.zeros((n_w,) * 3 + (3,))
    lo3, n3 = np.full(3, lo_w, np.int32), np.full(3, n_w, np.int32)
    orpic.apic_p2g(2, orpic.NODE, low, h, lo3, n3, xh[touch], m[touch].astype(np.float64),
                   c[touch].astype(np.float64).reshape(-1, 4, 3), rm, rmu)
    box = slice(2, 2 + U - L)
    assert gu.rel_inf(gmh[box, box, box], rm[box, box, box]) < 1e-5
    assert gu.rel_inf(gmuh[box, box, box], rmu[box, box, box]) < 1e-5

    pa.APIC.g2p_update_F(gu.ctx(), gmu, cd, dt, F, det, Finv, sd, m_i=gm)
    gu.ctx().synchronize()
    uh = orpic.grid_normalize(gmh, gmuh)  # the GPU's grid, as the fused g2p normalises it
    rc = orpic.apic_g2p(2, orpic.NODE, low, h, lo3, n3, xh[inside], uh)
    rg = orpic.g2p_gradu(2, orpic.NODE, low, h, lo3, n3, xh[inside], uh)
    rF, rdet, rFinv = orpic.update_F(rg, dt, F0[inside])
    idx = torch.as_tensor(inside, device=gu.DEV)
    cg = gu.host(cd[idx]).astype(np.float64).reshape(-1, 4, 3)
    assert gu.rel_inf(cg[:, 0], rc[:, 0]) < 1e-5
    assert gu.rel_inf(cg[:, 1:], rc[:, 1:]) < 1e-4
    Fg = gu.host(F[idx]).reshape(-1, 3, 3)
    Fig = gu.host(Finv[idx]).reshape(-1, 3, 3)
    assert gu.rel_inf(Fg, rF) < 1e-6
    assert gu.rel_inf(gu.host(det[idx])[:, 0], rdet) < 1e-6
    # F^-1 is the FP64 inverse of the F the GPU holds, and it differs from
    # the oracle's only as the inverse propagates the FP32 rounding of the
    # gradient: |d(F^-1)| <= |F^-1|^2 |dF| to first order (some of these
    # random-velocity steps leave F badly conditioned).
    assert np.abs(np.einsum("pij,pjk->pik", Fg, Fig) - np.eye(3)[None]).max() < 1e-10
    dF = np.abs(Fg - rF).max(axis=(1, 2))
    nFi = np.abs(rFinv).max(axis=(1, 2))
    assert (np.abs(Fig - rFinv).max(axis=(1, 2)) <= 10 * nFi ** 2 * dF + 1e-12 * nFi).all()

    # nine more steps; the FP64 update keeps det F and F^-1 consistent with F
    for _ in range(9):
        sd = step()
        pa.APIC.g2p_update_F(gu.ctx(), gmu, cd, dt, F, det, Finv, sd, m_i=gm)
    gu.ctx().synchronize()
    Fh = gu.host(F).reshape(P, 3, 3)
    deth = gu.host(det)[:, 0]
    Finvh = gu.host(Finv).reshape(P, 3, 3)
    assert np.isfinite(Fh).all() and np.isfinite(Finvh).all() and np.isfinite(deth).all()
    # Per particle, within the FP64 rounding of each product: the random
    # velocities stretch F by orders of magnitude over ten steps, so the
    # bounds scale with the entries.  det F sums products of three entries;
    # F^-1 = adj(F) / det F, so F F^-1 - I carries that rounding divided by
    # det F (a few eps nF^3 / |det F| for the adjugate formula).
    nF = np.abs(Fh).max(axis=(1, 2))
    assert (np.abs(deth - np.linalg.det(Fh)) <= 1e-13 * nF ** 3).all()
    eye_err = np.abs(np.einsum("pij,pjk->pik", Fh, Finvh) - np.eye(3)[None]).max(axis=(1, 2))
    assert (eye_err <= 1e-13 * nF ** 3 / np.abs(deth)).all()
    assert np.abs(Fh - F0).max() > 1e-6  # the ten steps moved F


def test_polypic_cubic_full_size(pa):
    import gpu_util as gu
    n, ppc, dt = 128, 8, 1e-3
    h = 1.0 / n
    low = np.full(3, -HALO * h)
    rng = np.random.default_rng(0x5EED0002)
    x = _positions(rng, _cells(0, n), ppc, h)
    P = len(x)
    assert P == 16_777_216
    m = np.full(P, h ** 3 / ppc, np.float32)
    nm = pa.PolyPIC.nmode(3)
    assert nm == 27
    g = torch.Generator(device=gu.DEV)
    g.manual_seed(0x5EED0002)
    # modes (r, d) at row 3 r + d of an SoA [81, P] array: N(0, 0.1), mode 0 N(0, 1)
    modes = torch.randn((nm * 3, P), generator=g, device=gu.DEV) * 0.1
    modes[:3] = torch.randn((3, P), generator=g, device=gu.DEV)
    modes = modes.t()
    nn = n + 1 + 2 * HALO
    f32 = torch.float32
    mesh = pa.Mesh(low, h)
    xd, md = gu.soa(x, f32), gu.soa(m[:, None], f32)
    gm = pa.GridField(torch.zeros((nn,) * 3, dtype=f32, device=gu.DEV))
    gmu = pa.GridField(torch.zeros((3,) + (nn,) * 3, dtype=f32, device=gu.DEV).permute(1, 2, 3, 0))
    gu_ = pa.GridField(torch.zeros((3,) + (nn,) * 3, dtype=f32, device=gu.DEV).permute(1, 2, 3, 0))
    cg = torch.empty_like(modes)
    # mesh-cell bins (the cubic knot base)
    geom = pa.Binning((HALO,) * 3, (n + 1,) * 3, (4, 4, 4), 0)
    bins = pa.bin_sort(gu.ctx(), mesh, xd, geom)
    sd = pa.create_spline(pa.FieldLocation.Node, 3, mesh, xd, bins)
    pa.PolyPIC.p2g(gu.ctx(), md, modes, modes, gmu, gm, dt, sd, zero_first=True)
    pa.grid_normalize(gu.ctx(), gm, gmu, gu_, rel_mass_eps=1e-6)
    pa.PolyPIC.g2p(gu.ctx(), gu_, cg, sd)
    gu.ctx().synchronize()
    # The box holds the ghost nodes of the low corner and the first cells.
    L, U = 0, 16
    lo_w, n_w = 0, U - L + 4
    win = slice(lo_w, lo_w + n_w)
    gmh = gu.host(gm.tensor[win, win, win]).astype(np.float64)
    gmuh = gu.host(gmu.tensor[win, win, win]).astype(np.float64)
    guh = gu.host(gu_.tensor[win, win, win]).astype(np.float64)
    xh = x.astype(np.float64)
    touch, inside = _box_sets(xh, low, h, 3, L, U)
    assert len(inside) > 10_000 and (gmh == 0).any() and (gmh > 0).any()

    def rows(sel):
        t = torch.as_tensor(sel, device=gu.DEV)
        return gu.host(modes[t]).astype(np.float64).reshape(-1, nm, 3)

    mt = rows(touch)
    rm = np.zeros((n_w,) * 3)
    rmu = np.zeros((n_w,) * 3 + (3,))
    lo3, n3 = np.full(3, lo_w, np.int32), np.full(3, n_w, np.int32)
    orpic.polypic_p2g(3, orpic.NODE, low, h, lo3, n3, xh[touch], m[touch].astype(np.float64),
                      mt, mt, dt, rmu, rm)
    box = slice(L - lo_w, U - lo_w)
    assert gu.rel_inf(gmh[box, box, box], rm[box, box, box]) < 1e-5
    assert gu.rel_inf(gmuh[box, box, box], rmu[box, box, box]) < 1e-5
    rc = orpic.polypic_g2p(3, orpic.NODE, low, h, lo3, n3, xh[inside], guh, 3)
    cgh = gu.host(cg[torch.as_tensor(inside, device=gu.DEV)]).astype(np.float64).reshape(-1, nm, 3)
    # per mode family (values, first and second derivatives differ in scale)
    assert gu.rel_inf(cgh[:, :1], rc[:, :1]) < 1e-5
    for lo_r, hi_r in ((1, 4), (4, nm)):
        assert gu.rel_inf(cgh[:, lo_r:hi_r], rc[:, lo_r:hi_r]) < 1e-4
