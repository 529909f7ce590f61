"""Pin the oracle functions the reference's known answers do not cover.

Each check uses mathematics that fixes the result independently of the
oracle's own formulas: closed forms, polynomial reproduction, conservation,
exact integer arithmetic and brute force.
"""
import itertools

import numpy as np
import pytest

from oracle import orpic

NODE, CELL = orpic.NODE, orpic.CELL
LOW = np.array([-0.25, -0.5, 0.125])
H = 0.125


def _rand_x(rng, n, lo_cell, hi_cell):
    return LOW + H * rng.uniform(lo_cell, hi_cell, size=(n, 3))


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", range(8))
def test_spline_partition_and_moments(order, entity):
    """Sum w = 1, sum w d = 0, sum g = 0, sum g x_n = 1, sum g2 = 0 and, for
    orders 2 and 3, D_p = sum w d^2 = h^2/4 resp. h^2/3 (the closed form behind
    APIC::inertialScaling, Picasso_APIC.hpp:33-47)."""
    rng = np.random.default_rng(order * 10 + entity)
    for x in _rand_x(rng, 50, 4, 12):
        s = orpic.spline(order, entity, LOW, H, x)
        for d in range(3):
            w, g, g2, dist = s["w"][d], s["g"][d], s["g2"][d], s["d"][d]
            assert abs(w.sum() - 1.0) < 1e-14
            assert abs(g.sum()) < 1e-12 / H
            assert abs(g2.sum()) < 1e-10 / H ** 2
            assert np.all(w >= 0.0)
            xn = x[d] + dist
            assert abs(np.dot(g, xn) - 1.0) < 1e-12
            if order >= 2:
                assert abs(np.dot(w, dist)) < 1e-15
                want = H * H / (4.0 if order == 2 else 3.0)
                assert abs(np.dot(w, dist * dist) - want) < 1e-15
            # nodes are consecutive, lattice-aligned, and bracket the particle
            off = orpic.lib().orc_entity_offset(entity, d, H)
            np.testing.assert_allclose(xn, LOW[d] + off + H * s["s"][d], atol=1e-13)
            assert xn[0] <= x[d] <= xn[-1]


def test_spline_values_brute_force():
    """Weights equal the B-spline kernels evaluated at the node distances."""
    def bspline(order, r):
        r = abs(r)
        if order == 1:
            return max(0.0, 1 - r)
        if order == 2:
            return 0.75 - r * r if r < 0.5 else (0.5 * (1.5 - r) ** 2 if r < 1.5 else 0.0)
        return (2 / 3 - r * r + 0.5 * r ** 3) if r < 1 else ((2 - r) ** 3 / 6 if r < 2 else 0.0)

    rng = np.random.default_rng(7)
    for order in (1, 2, 3):
        for x in _rand_x(rng, 30, 4, 12):
            s = orpic.spline(order, NODE, LOW, H, x)
            for d in range(3):
                for n in range(s["nk"]):
                    assert s["w"][d][n] == pytest.approx(bspline(order, s["d"][d][n] / H), abs=1e-14)


def _window_for(x, pad=3):
    cells = np.floor((x - LOW) / H).astype(int)
    lo = cells.min(0) - pad
    n = cells.max(0) - lo + pad + 2
    return lo.astype(np.int32), n.astype(np.int32)


def _node_coords(lo, n, entity=NODE):
    offs = [orpic.lib().orc_entity_offset(entity, d, H) for d in range(3)]
    ax = [LOW[d] + offs[d] + H * np.arange(lo[d], lo[d] + n[d]) for d in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_apic_affine_field_reproduction(order):
    """g2p of u(x) = u0 + A x gives v_p = u(x_p) and, for orders 2/3,
    B_p = A D_p; p2g of that state then gives mu_n = m_n u(x_n) so grid
    normalisation recovers u at every touched node.  Order 1 reproduces
    v_p and grad u via B_p = A D_p with the gradient-form p2g."""
    rng = np.random.default_rng(order)
    x = _rand_x(rng, 40, 5, 9)
    lo, n = _window_for(x)
    X = _node_coords(lo, n)
    u0 = rng.normal(size=3)
    A = rng.normal(size=(3, 3))
    gu = u0 + X @ A.T
    c = orpic.apic_g2p(order, NODE, LOW, H, lo, n, x, gu)
    np.testing.assert_allclose(c[:, 0, :], u0 + x @ A.T, atol=1e-12)
    if order >= 2:
        Dp = H * H / (4.0 if order == 2 else 3.0)
        for p in range(len(x)):
            np.testing.assert_allclose(c[p, 1:, :].T, A * Dp, atol=1e-13)
        m = rng.uniform(0.5, 2.0, size=len(x))
        gm = np.zeros(tuple(n))
        gmu = np.zeros(tuple(n) + (3,))
        orpic.apic_p2g(order, NODE, LOW, H, lo, n, x, m, c, gm, gmu)
        u = orpic.grid_normalize(gm, gmu)
        touched = gm > 0
        np.testing.assert_allclose(u[touched], gu[touched], atol=1e-11)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", [NODE, CELL, 2, 6])
def test_apic_p2g_conservation(order, entity):
    """sum_n m_n = sum_p m_p; sum_n mu_n = sum_p m_p v_p (the affine term
    cancels because sum w d = 0 for orders 2/3 and sum grad w = 0 for order 1)."""
    rng = np.random.default_rng(100 + order)
    x = _rand_x(rng, 64, 5, 9)
    lo, n = _window_for(x)
    m = rng.uniform(0.1, 1.0, size=len(x))
    c = rng.normal(size=(len(x), 4, 3))
    gm = np.zeros(tuple(n))
    stag = entity not in (NODE, CELL)
    gmu = np.zeros(tuple(n) + (() if stag else (3,)))
    orpic.apic_p2g(order, entity, LOW, H, lo, n, x, m, c, gm, gmu)
    assert gm.sum() == pytest.approx(m.sum(), rel=1e-14)
    mom = (m[:, None] * c[:, 0, :]).sum(0)
    if stag:
        dim = (entity - 2) % 3
        assert gmu.sum() == pytest.approx(mom[dim], rel=1e-12, abs=1e-12)
    else:
        np.testing.assert_allclose(gmu.reshape(-1, 3).sum(0), mom, rtol=1e-12, atol=1e-12)


def test_gradu_linear_field():
    """grad u of u(x) = u0 + A x is A for every order (gradu[i][j] = du_i/dx_j)."""
    rng = np.random.default_rng(3)
    x = _rand_x(rng, 30, 5, 9)
    lo, n = _window_for(x)
    X = _node_coords(lo, n)
    A = rng.normal(size=(3, 3))
    gu = 0.3 + X @ A.T
    for order in (1, 2, 3):
        L = orpic.g2p_gradu(order, NODE, LOW, H, lo, n, x, gu)
        for p in range(len(x)):
            np.testing.assert_allclose(L[p], A, atol=1e-11)
    # A dropped transpose fails: a non-symmetric A distinguishes A from A^T.
    assert np.abs(A - A.T).max() > 0.1


def test_det_inverse_exact_integers():
    """Integer matrices: det is exact in FP64 and A A^{-1} = I."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        a = rng.integers(-9, 10, size=(3, 3)).astype(float)
        want = round(np.linalg.det(a))
        assert orpic.det3(a) == want
        if want != 0:
            inv = orpic.inverse3(a)
            np.testing.assert_allclose(inv @ a, np.eye(3), atol=1e-12)
            np.testing.assert_allclose(a @ inv, np.eye(3), atol=1e-12)
            # adjugate entries are integers: inv*det is integral
            np.testing.assert_allclose(inv * want, np.round(inv * want), atol=1e-9)


def test_update_F_closed_form():
    """F_new = (I + dt L) F; det(F_new) = det(I + dt L) det(F);
    F_new^{-1} F_new = I; compared with numpy on random inputs."""
    rng = np.random.default_rng(5)
    n = 64
    L = rng.normal(size=(n, 3, 3))
    F = np.eye(3) + 0.01 * rng.normal(size=(n, 3, 3))
    dt = 1e-3
    Fn, det, Finv = orpic.update_F(L, dt, F)
    A = np.eye(3) + dt * L
    np.testing.assert_allclose(Fn, A @ F, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(det, np.linalg.det(A) * np.linalg.det(F), rtol=1e-12)
    np.testing.assert_allclose(Finv @ Fn, np.broadcast_to(np.eye(3), Fn.shape), atol=1e-13)


def test_normalize_closed_form():
    m = np.array([0.0, 2.0, 1e-30, 4.0])
    mu = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1e-30, 0.0, -2e-30], [-4.0, 0.0, 8.0]])
    u = orpic.grid_normalize(m, mu)
    np.testing.assert_array_equal(u, [[0, 0, 0], [1, 2, 3], [1, 0, -2], [-1, 0, 2]])


@pytest.mark.parametrize("brick", [(1, 1, 1), (4, 4, 4), (2, 3, 5)])
def test_cell_keys_and_stable_sort(brick):
    """Keys decode back to the cells of floor((x-low)/h); the permutation
    equals numpy's stable argsort; key_start is the exclusive scan of the
    histogram."""
    rng = np.random.default_rng(17)
    ncell = np.array([12, 10, 15], np.int32)
    cell_lo = np.array([2, 3, 1], np.int32)
    x = LOW + H * (cell_lo + rng.uniform(0, 1, size=(3000, 3)) * ncell)
    # many duplicates so stability matters
    x[::3] = x[1::3][: len(x[::3])]
    keys, cells = orpic.cell_keys(x, LOW, H, cell_lo, ncell, brick)
    np.testing.assert_array_equal(cells, np.floor((x - LOW) / H).astype(np.int32))
    b = np.array(brick)
    nb = -(-ncell // b)
    rel = cells - cell_lo
    bid = ((rel[:, 0] // b[0]) * nb[1] + rel[:, 1] // b[1]) * nb[2] + rel[:, 2] // b[2]
    lid = ((rel[:, 0] % b[0]) * b[1] + rel[:, 1] % b[1]) * b[2] + rel[:, 2] % b[2]
    np.testing.assert_array_equal(keys, bid * b.prod() + lid)
    nkeys = int(nb.prod() * b.prod())
    perm, start = orpic.bin_sort(keys, nkeys)
    np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable"))
    hist = np.bincount(keys, minlength=nkeys)
    np.testing.assert_array_equal(start, np.concatenate([[0], np.cumsum(hist)]))


@pytest.mark.parametrize("mask", range(8))
def test_cell_keys_dual_per_dimension(mask):
    """A per-dimension dual (0x10 | mask) bins each dimension exactly as the
    all-mesh (dual 0) or all-dual (dual 1) binning does in that dimension;
    values outside 0, 1 and 0x10 | mask are rejected."""
    rng = np.random.default_rng(40 + mask)
    ncell = np.array([12, 10, 15], np.int32)
    cell_lo = np.array([2, 3, 1], np.int32)
    x = LOW + H * (cell_lo + 0.5 + rng.uniform(0, 1, size=(2000, 3)) * (ncell - 1))
    _, c0 = orpic.cell_keys(x, LOW, H, cell_lo, ncell, (4, 4, 4), 0)
    _, c1 = orpic.cell_keys(x, LOW, H, cell_lo, ncell, (4, 4, 4), 1)
    _, cm = orpic.cell_keys(x, LOW, H, cell_lo, ncell, (4, 4, 4), 0x10 | mask)
    for d in range(3):
        np.testing.assert_array_equal(cm[:, d], (c1 if (mask >> d) & 1 else c0)[:, d])
    # the dual cell of a point half a cell past a node is the next one
    assert np.all(c1 - c0 >= 0) and np.all(c1 - c0 <= 1)
    for bad in (2, 0x20, 0x18, -1):
        with pytest.raises(orpic.OracleError):
            orpic.cell_keys(x, LOW, H, cell_lo, ncell, (4, 4, 4), bad)


def test_aligned_dual_matches_the_knot_base():
    """aligned_dual(order, entity) bins by the spline's knot base in every
    dimension: on its bins, the first stencil knot of every particle is the
    same offset from its bin (checked through the oracle's spline), for all
    orders and all eight entity lattices."""
    import picasso_amd as pa
    rng = np.random.default_rng(5)
    x = LOW + H * (6 + rng.uniform(0, 8, size=(400, 3)))
    for order in (1, 2, 3):
        for entity in range(8):
            dual = pa.aligned_dual(order, entity)
            flags = pa.dual_dims(dual)
            _, cells = orpic.cell_keys(x, LOW, H, [0, 0, 0], [40, 40, 40], (1, 1, 1), dual)
            base = np.array([orpic.spline(order, entity, LOW, H, xp)["s"][:, 0] for xp in x])
            off = base - cells
            assert np.all(off == off[0]), (order, entity, dual)
            if entity in (0, 1):  # the collocated rule the brick kernels always used
                assert dual == (1 if (order == 2) == (entity == 0) else 0)
            assert all(f in (0, 1) for f in flags)


@pytest.mark.parametrize("periodic", [(1, 1, 1), (1, 0, 1), (0, 1, 0), (0, 0, 0)])
def test_grid_halo_scatter_is_the_sum_over_preimages(periodic):
    """The sequential per-dimension fold equals the direct definition: each
    owned-in-the-periodic-dimensions node receives the sum of every window
    node whose image (all periodic dimensions mapped at once) it is; the
    total over those nodes is conserved; gather copies the images.  (Ghost
    values after the scatter are unspecified: an x-fold adds into nodes that
    are still y-ghosts.)"""
    rng = np.random.default_rng(sum(periodic))
    lo, n = np.array([-2, -1, -3]), np.array([14, 11, 15])
    olo, on = np.array([0, 1, 0]), np.array([9, 8, 10])
    g = rng.normal(size=tuple(n) + (2,))
    out = orpic.grid_halo(1, lo, g.copy(), olo, on, periodic)
    idx = np.stack(np.meshgrid(*[np.arange(lo[d], lo[d] + n[d]) for d in range(3)],
                               indexing="ij"), -1).reshape(-1, 3)
    img = idx.copy()
    for d in range(3):
        if periodic[d]:
            img[:, d] = olo[d] + np.mod(idx[:, d] - olo[d], on[d])
    ghost = np.any(img != idx, 1)
    want = g.reshape(-1, 2).copy()
    flat = lambda v: ((v[:, 0] - lo[0]) * n[1] + (v[:, 1] - lo[1])) * n[2] + (v[:, 2] - lo[2])  # noqa: E731
    np.add.at(want, flat(img[ghost]), g.reshape(-1, 2)[ghost])
    keep = ~ghost  # nodes that are their own image: owned in every periodic dimension
    np.testing.assert_allclose(out.reshape(-1, 2)[keep], want[keep], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(out.reshape(-1, 2)[keep].sum(0), g.reshape(-1, 2).sum(0),
                               rtol=1e-12)
    ga = orpic.grid_halo(0, lo, out.copy(), olo, on, periodic)
    np.testing.assert_array_equal(ga.reshape(-1, 2), out.reshape(-1, 2)[flat(img)])


def test_uniform01_is_splitmix64():
    """The random placement draws SplitMix64 outputs: for seed 0 the states
    1, 2, 3 x golden gamma give the generator's published first outputs
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    for d, v in enumerate((0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F)):
        assert orpic.uniform01(0, 0, d) == (v >> 11) / 2.0 ** 53


@pytest.mark.parametrize("ppd", [1, 2, 3])
def test_init_uniform_lattice(ppd):
    """InitUniform (Picasso_ParticleInit.hpp:279-400): each cell holds a
    ppd^3 lattice with spacing h/ppd, half a spacing in from the cell faces,
    centred on the cell; particle p belongs to cell p // ppd^3 with i
    fastest and, inside, ip fastest; volumes add up to the cells' volume."""
    cell_lo, ncell, h = (2, 1, 3), (3, 4, 2), 0.125
    x, vol = orpic.init_particles(0, LOW, H, cell_lo, ncell, ppd)
    npc = ppd ** 3
    assert x.shape == (np.prod(ncell) * npc, 3)
    c = np.arange(len(x)) // npc
    own = np.stack([c % ncell[0], (c // ncell[0]) % ncell[1], c // (ncell[0] * ncell[1])], 1)
    lc = LOW + (np.array(cell_lo) + own) * H
    frac = (x - lc) / H
    loc = np.arange(len(x)) % npc
    ip = np.stack([loc % ppd, (loc // ppd) % ppd, loc // ppd ** 2], 1)
    np.testing.assert_allclose(frac, (ip + 0.5) / ppd, rtol=0, atol=1e-12)
    np.testing.assert_allclose(x.reshape(-1, npc, 3).mean(1), lc[::npc] + 0.5 * H, atol=1e-12)
    assert abs(vol.sum() - np.prod(ncell) * H ** 3) < 1e-12


def test_init_random_in_cells():
    """InitRandom (Picasso_ParticleInit.hpp:121-248): every particle lies in
    its own cell (p // ppc, i fastest), the fractional coordinates are
    uniform (mean 1/2, variance 1/12, no dimension correlated), the same
    seed gives the same points and another seed different ones."""
    cell_lo, ncell, ppc = (1, 2, 3), (8, 6, 5), 40
    x, vol = orpic.init_particles(1, LOW, H, cell_lo, ncell, ppc, seed=7)
    c = np.arange(len(x)) // ppc
    own = np.stack([c % ncell[0], (c // ncell[0]) % ncell[1], c // (ncell[0] * ncell[1])], 1)
    frac = (x - (LOW + (np.array(cell_lo) + own) * H)) / H
    assert frac.min() >= 0.0 and frac.max() < 1.0
    assert np.all(np.abs(frac.mean(0) - 0.5) < 0.01)
    assert np.all(np.abs(frac.var(0) - 1 / 12) < 0.005)
    assert np.all(np.abs(np.corrcoef(frac.T) - np.eye(3)) < 0.03)
    x2, _ = orpic.init_particles(1, LOW, H, cell_lo, ncell, ppc, seed=7)
    x3, _ = orpic.init_particles(1, LOW, H, cell_lo, ncell, ppc, seed=8)
    np.testing.assert_array_equal(x, x2)
    assert np.mean(x == x3) < 1e-3
    np.testing.assert_allclose(vol, H ** 3 / ppc, rtol=1e-15)


def test_cell_keys_out_of_range_is_error():
    x = np.array([[LOW[0] - 0.01, LOW[1], LOW[2]]])
    with pytest.raises(orpic.OracleError):
        orpic.cell_keys(x, LOW, H, [0, 0, 0], [4, 4, 4])


def test_p2g_out_of_window_is_error():
    x = np.array([[LOW[0] + 0.3 * H, LOW[1] + 5 * H, LOW[2] + 5 * H]])
    gm = np.zeros((10, 10, 10))
    gmu = np.zeros((10, 10, 10, 3))
    with pytest.raises(orpic.OracleError):
        orpic.apic_p2g(2, NODE, LOW, H, [0, 0, 0], [10, 10, 10], x, [1.0],
                       np.zeros((1, 4, 3)), gm, gmu)


# ---------------------------------------------------------------------------
# 27-mode PolyPIC (this library's extension; no reference known answers).

def _poly_setup(order, seed, nmode_field=None):
    rng = np.random.default_rng(seed)
    x = _rand_x(rng, 32, 5, 9)
    lo, n = _window_for(x)
    return rng, x, lo, n


@pytest.mark.parametrize("order", [1, 2, 3])
def test_polypic_affine_modes(order):
    """g2p of u(x) = u0 + A x: mode 0 is u(x_p), the three linear modes are
    the columns of A, every other mode vanishes (cubic B-splines reproduce
    affine fields, so all second derivatives of the interpolant are 0)."""
    rng, x, lo, n = _poly_setup(order, 21 + order)
    X = _node_coords(lo, n)
    u0 = rng.normal(size=3)
    A = rng.normal(size=(3, 3))
    gu = u0 + X @ A.T
    modes = orpic.polypic_g2p(order, NODE, LOW, H, lo, n, x, gu, 3)
    np.testing.assert_allclose(modes[:, 0, :], u0 + x @ A.T, atol=1e-12)
    for a in range(3):
        np.testing.assert_allclose(modes[:, 1 + a, :], np.broadcast_to(A[:, a], (len(x), 3)),
                                   atol=1e-10)
    np.testing.assert_allclose(modes[:, 4:, :], 0.0, atol=1e-9)


def test_polypic27_quadratic_taylor_modes():
    """For u(x) = x_a x_b the cubic-spline interpolant is x_a x_b + const
    (B-spline quasi-interpolation error is a constant for quadratics), so
    the 27-mode g2p returns the Taylor coefficients of x_a x_b about x_p:
    the y_a y_b mode is 1 (a != b) and the y_a^2 mode is 1 (a == b)."""
    rng, x, lo, n = _poly_setup(3, 31)
    X = _node_coords(lo, n)
    exps = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    exps += [e for e in itertools.product(range(3), repeat=3) if max(e) == 2]
    for a, b in [(0, 1), (1, 2), (0, 0), (2, 2)]:
        gu = np.repeat((X[..., a] * X[..., b])[..., None], 1, axis=-1)
        modes = orpic.polypic_g2p(3, NODE, LOW, H, lo, n, x, gu, 1)
        e = [0, 0, 0]
        e[a] += 1
        e[b] += 1
        r = exps.index(tuple(e))
        np.testing.assert_allclose(modes[:, r, 0], 1.0, atol=1e-9)
        for q, ex in enumerate(exps):
            if sum(ex) > 2 or (sum(ex) == 2 and q != r):
                np.testing.assert_allclose(modes[:, q, 0], 0.0, atol=1e-8)


@pytest.mark.parametrize("order", [1, 3])
def test_polypic_p2g_conservation_and_constant_field(order):
    """Mass is conserved; with dt = 0 and only mode 0 set, the momentum is
    m_p c_0 and normalisation returns c_0 where all particles share it."""
    rng, x, lo, n = _poly_setup(order, 41)
    nmode = orpic.polypic_nmode(order)
    m = rng.uniform(0.1, 1.0, len(x))
    c = np.zeros((len(x), nmode, 3))
    c[:, 0, :] = [1.5, -2.0, 0.25]
    gm = np.zeros(tuple(n))
    gmu = np.zeros(tuple(n) + (3,))
    orpic.polypic_p2g(order, NODE, LOW, H, lo, n, x, m, c, c, 0.0, gmu, gm)
    assert gm.sum() == pytest.approx(m.sum(), rel=1e-14)
    u = orpic.grid_normalize(gm, gmu)
    np.testing.assert_allclose(u[gm > 0], np.broadcast_to([1.5, -2.0, 0.25], u[gm > 0].shape),
                               atol=1e-13)


@pytest.mark.parametrize("order", [1, 3])
def test_polypic_affine_roundtrip(order):
    """g2p of an affine field then p2g with dt = 0 reproduces the field at
    every touched node (the modes encode u(x_p + y) exactly)."""
    rng, x, lo, n = _poly_setup(order, 51)
    X = _node_coords(lo, n)
    A = rng.normal(size=(3, 3))
    gu = 0.7 + X @ A.T
    modes = orpic.polypic_g2p(order, NODE, LOW, H, lo, n, x, gu, 3)
    m = rng.uniform(0.5, 1.0, len(x))
    gm = np.zeros(tuple(n))
    gmu = np.zeros(tuple(n) + (3,))
    orpic.polypic_p2g(order, NODE, LOW, H, lo, n, x, m, modes, modes, 0.0, gmu, gm)
    u = orpic.grid_normalize(gm, gmu)
    t = gm > 0
    np.testing.assert_allclose(u[t], gu[t], atol=1e-9)


_EXPS27 = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1),
           (1, 1, 1)] + [e for e in itertools.product(range(3), repeat=3) if max(e) == 2]


@pytest.mark.parametrize("order", [2, 3])
def test_polypic27_quadratic_round_trip(order):
    """Pins the quadratic modes of p2g against those of g2p: the spline
    interpolant of u = x_a x_b is x_a x_b + delta_ab D (D = h^2/4 for
    quadratic, h^2/3 for cubic B-splines, the constant second moment), a
    quadratic, so its 27 Taylor modes at x_p represent it exactly; p2g at
    dt = 0 evaluates them at the nodes, and normalisation returns
    x_a x_b + delta_ab D at every node with mass.  A p2g whose mode
    convention differed from g2p's (exponent order, the 1/2 of the second
    derivative) would not."""
    rng, x, lo, n = _poly_setup(order, 61 + order)
    X = _node_coords(lo, n)
    field = np.stack([X[..., 0] * X[..., 1], X[..., 2] * X[..., 2], X[..., 0] * X[..., 2]], -1)
    D = H * H / (4.0 if order == 2 else 3.0)
    modes = orpic.polypic_g2p(order, NODE, LOW, H, lo, n, x, field, 3)
    m = rng.uniform(0.5, 1.0, len(x))
    gm = np.zeros(tuple(n))
    gmu = np.zeros(tuple(n) + (3,))
    orpic.polypic_p2g(order, NODE, LOW, H, lo, n, x, m, modes, modes, 0.0, gmu, gm)
    u = orpic.grid_normalize(gm, gmu)
    t = gm > 1e-6 * gm.max()
    want = field + np.array([0.0, D, 0.0])
    np.testing.assert_allclose(u[t], want[t], atol=1e-11)


@pytest.mark.parametrize("kind", ["diagonal", "nilpotent"])
def test_polypic27_backtrack_known_answer(kind):
    """p2g with dt != 0: y = (I + dt U)^-1 d with U(a, d) = u_p(1 + a, d).
    For a diagonal U, y_a = d_a / (1 + dt u_aa); for U with one entry s at
    (0, 1), (I + dt U)^-1 = I - dt U and y_0 = d_0 - dt s d_1 (fixing the
    orientation of U).  Field modes c_0, the y_0 and y_1 modes and the y_0^2
    mode; the node momentum is sum_p m_p w_n (c_0 + g_0 y_0 + g_1 y_1 +
    q y_0^2) with w_n and d_n from the spline evaluation."""
    order, dt = 3, 0.05
    rng, x, lo, n = _poly_setup(order, 71)
    P = len(x)
    nm = orpic.polypic_nmode(order)
    m = rng.uniform(0.5, 1.0, P)
    c = np.zeros((P, nm, 1))
    c0, g0, g1, q = rng.normal(size=(4, P))
    c[:, 0, 0], c[:, 1, 0], c[:, 2, 0] = c0, g0, g1
    c[:, _EXPS27.index((2, 0, 0)), 0] = q
    u = np.zeros((P, nm, 3))
    diag = rng.uniform(-2.0, 2.0, size=(P, 3))
    s = rng.uniform(-3.0, 3.0, P)
    if kind == "diagonal":
        for a in range(3):
            u[:, 1 + a, a] = diag[:, a]
    else:
        u[:, 1, 1] = s
    gm = np.zeros(tuple(n))
    gmu = np.zeros(tuple(n) + (1,))
    orpic.polypic_p2g(order, NODE, LOW, H, lo, n, x, m, u, c, dt, gmu, gm)
    want = np.zeros(tuple(n))
    for p in range(P):
        sp = orpic.spline(order, NODE, LOW, H, x[p])
        for i in range(4):
            for j in range(4):
                for k in range(4):
                    d = np.array([sp["d"][0][i], sp["d"][1][j], sp["d"][2][k]])
                    if kind == "diagonal":
                        y = d / (1.0 + dt * diag[p])
                    else:
                        y = np.array([d[0] - dt * s[p] * d[1], d[1], d[2]])
                    wn = sp["w"][0][i] * sp["w"][1][j] * sp["w"][2][k]
                    node = (sp["s"][0][i] - lo[0], sp["s"][1][j] - lo[1], sp["s"][2][k] - lo[2])
                    want[node] += m[p] * wn * (c0[p] + g0[p] * y[0] + g1[p] * y[1]
                                               + q[p] * y[0] * y[0])
    np.testing.assert_allclose(gmu[..., 0], want, atol=1e-12 * np.abs(want).max())
