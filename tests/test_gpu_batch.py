"""Batched independent instances (pic_batch_*) against the oracle, instance
by instance: uneven grids of 8^3 to 24^3 cells at 2 to 16 particles per
cell, every pass one launch over all instances.  Each instance's grid mass,
momentum and the c_p its normalised g2p writes match the oracle's
p2g -> normalise -> g2p on that instance alone (FP64 within 1e-12, FP32
within 1e-5 of the FP64 oracle on the rounded inputs, B within 1e-4 as for
the single-instance FP32 step)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

HALO = 2
F64_TOL = 1e-12
F32_TOL = 1e-5


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _workload(ninst, nmin, nmax, seed):
    """Instances with n ~ U{nmin..nmax} cells per dimension and ppc ~
    U{2..16}, positions uniform in their cells (shuffled), v ~ N(0, 1),
    B ~ N(0, 0.1), m = h^3 / ppc."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ninst):
        n = int(rng.integers(nmin, nmax + 1))
        ppc = int(rng.integers(2, 17))
        h = 1.0 / n
        ii = np.arange(n)
        cells = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
        x = (np.repeat(cells, ppc, axis=0) + rng.uniform(size=(len(cells) * ppc, 3))) * h
        x = x[rng.permutation(len(x))]
        P = len(x)
        c = np.empty((P, 4, 3))
        c[:, 0] = rng.normal(size=(P, 3))
        c[:, 1:] = rng.normal(0.0, 0.1, size=(P, 3, 3))
        out.append(dict(n=n, h=h, x=x, m=np.full(P, h ** 3 / ppc), c=c,
                        low=np.full(3, -HALO * h)))
    return out


def _build(pa, wl, order, tdt, f32):
    import gpu_util as gu
    xs, ms, cs, insts, first = [], [], [], [], 0
    dual = pa.aligned_dual(order, pa.FieldLocation.Node)
    for w in wl:
        x, m, c = w["x"], w["m"], w["c"]
        if f32:
            x, m, c = (a.astype(np.float32).astype(np.float64) for a in (x, m, c))
            w["x"], w["m"], w["c"] = x, m, c
        nn = w["n"] + 1 + 2 * HALO
        gm = pa.GridField(torch.full((nn, nn, nn), -3.0, dtype=tdt, device=gu.DEV))
        gmu = pa.GridField(torch.full((nn, nn, nn, 3), -3.0, dtype=tdt, device=gu.DEV))
        geom = pa.Binning((HALO,) * 3, (w["n"] + 1,) * 3, (4, 4, 4), dual)
        insts.append(pa.Instance(pa.Mesh(w["low"], w["h"]), first, len(x), geom, gm, gmu))
        first += len(x)
        xs.append(x)
        ms.append(m)
        cs.append(c.reshape(-1, 12))
    X = gu.soa(np.concatenate(xs), tdt)
    M = gu.soa(np.concatenate(ms)[:, None], tdt)
    C = gu.soa(np.concatenate(cs), tdt)
    return insts, X, M, C


def _check(gu, wl, insts, C, order, f32):
    cg = gu.host(C).reshape(-1, 4, 3)
    tol = F32_TOL if f32 else F64_TOL
    tol_b = 1e-4 if f32 else F64_TOL
    for w, inst in zip(wl, insts):
        nn = w["n"] + 1 + 2 * HALO
        lo, n = np.zeros(3, np.int32), np.full(3, nn, np.int32)
        rm = np.zeros((nn,) * 3)
        rmu = np.zeros((nn,) * 3 + (3,))
        orpic.apic_p2g(order, orpic.NODE, w["low"], w["h"], lo, n, w["x"], w["m"], w["c"],
                       rm, rmu)
        ru = orpic.grid_normalize(rm, rmu)
        rc = orpic.apic_g2p(order, orpic.NODE, w["low"], w["h"], lo, n, w["x"], ru)
        assert gu.rel_inf(gu.host(inst.m_i.tensor), rm) < tol, w["n"]
        assert gu.rel_inf(gu.host(inst.mu_i.tensor), rmu) < tol, w["n"]
        got = cg[inst.first:inst.first + inst.np]
        assert gu.rel_inf(got[:, 0], rc[:, 0]) < tol, w["n"]
        assert gu.rel_inf(got[:, 1:], rc[:, 1:]) < tol_b, w["n"]


@pytest.mark.parametrize("group", [0, 128])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_step_vs_oracle(pa, dtype, group):
    """32 instances, quadratic splines, one step from shuffled storage (the
    sort's full path), with the g2p workgroups of the default shape and of
    128 work-items (PIC_OPT_G2P_GROUP)."""
    import gpu_util as gu
    f32 = dtype == "f32"
    tdt = torch.float32 if f32 else torch.float64
    wl = _workload(32, 8, 24, 0x5EED0003)
    insts, X, M, C = _build(pa, wl, 2, tdt, f32)
    batch = pa.Batch(gu.ctx(), 2, tdt, insts)
    gu.ctx().set_g2p_group(group)
    try:
        batch.step(gu.ctx(), X, M, C)
        gu.ctx().synchronize()
    finally:
        gu.ctx().set_g2p_group(0)
    _check(gu, wl, insts, C, 2, f32)
    batch.close()


@pytest.mark.parametrize("order", [1, 3])
def test_batch_other_orders_vs_oracle(pa, order):
    """Linear and cubic splines (mesh-cell bins) through the same passes."""
    import gpu_util as gu
    wl = _workload(12, 8, 20, 100 + order)
    insts, X, M, C = _build(pa, wl, order, torch.float64, False)
    batch = pa.Batch(gu.ctx(), order, torch.float64, insts)
    batch.bin_sort(gu.ctx(), X)
    batch.p2g(gu.ctx(), X, M, C)
    batch.g2p_normalized(gu.ctx(), X, C)
    gu.ctx().synchronize()
    _check(gu, wl, insts, C, order, False)
    batch.close()


def test_batch_graph_replays_sorted_storage(pa):
    """Particles stored in bin order (the sort's identity path), the step
    captured once in a step graph and replayed three times with eager work
    between the replays: every replay matches the eager step and the
    oracle."""
    import gpu_util as gu
    wl = _workload(40, 8, 24, 77)
    for w in wl:  # store each instance's particles in bin order
        keys, _ = orpic.cell_keys(w["x"], w["low"], w["h"], (HALO,) * 3, (w["n"] + 1,) * 3,
                                  (4, 4, 4), 1)
        o = np.argsort(keys, kind="stable")
        w["x"], w["c"] = w["x"][o], w["c"][o]
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        ctx = pa.Context(0)
        insts, X, M, C = _build(pa, wl, 2, torch.float64, False)
        C0 = C.clone()
        batch = pa.Batch(ctx, 2, torch.float64, insts)
        batch.step(ctx, X, M, C)
        ctx.synchronize()
        ref_c = gu.host(C).copy()
        ref_m = [gu.host(i.m_i.tensor).copy() for i in insts]
        C.copy_(C0)
        ctx.graph_begin()
        batch.step(ctx, X, M, C)
        graph = ctx.graph_end()
        for _ in range(3):
            C.copy_(C0)
            for i in insts:
                i.m_i.tensor.fill_(7.0)
            graph.launch()
            ctx.synchronize()
            assert gu.rel_inf(gu.host(C), ref_c) < 1e-13
            for i, r in zip(insts, ref_m):
                assert gu.rel_inf(gu.host(i.m_i.tensor), r) < 1e-13
        _check(gu, wl, insts, C, 2, False)
        graph.close()
        batch.close()
        ctx.close()


def test_batch_rejects_bad_tables(pa):
    """Instances must tile the particles and use 4^3 bricks on the knot
    lattice."""
    import gpu_util as gu
    wl = _workload(2, 8, 8, 5)
    insts, X, M, C = _build(pa, wl, 2, torch.float64, False)
    gap = [insts[0], pa.Instance(insts[1].mesh, insts[1].first + 1, insts[1].np, insts[1].geom,
                                 insts[1].m_i, insts[1].mu_i)]
    with pytest.raises(pa.PicError):
        pa.Batch(gu.ctx(), 2, torch.float64, gap)
    wrong_dual = [pa.Instance(i.mesh, i.first, i.np, pa.Binning(i.geom.cell_lo, i.geom.ncell,
                                                                 (4, 4, 4), 0), i.m_i, i.mu_i)
                  for i in insts]
    with pytest.raises(pa.PicError):
        pa.Batch(gu.ctx(), 2, torch.float64, wrong_dual)
