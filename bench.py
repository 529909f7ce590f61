"""Throughput of the APIC particle<->grid transfer step on one MI355X.

Workload: 256^3 cells on the unit cube, quadratic B-splines on nodes,
particles at 8 per cell in the cells whose centre lies within 0.3 of the
cube centre (~15.2M particles), v = rigid rotation about z plus N(0, 0.05),
B ~ N(0, 0.1), shuffled.  One step = cell-bin sort -> p2g (grid zeroed
first) -> grid normalisation -> g2p, with inputs resident in HBM.

The headline value is the FP32 (compensated p2g) step in particles/s; the
plain FP64 step is timed beside it, and the FP32 grid momentum is compared
with the FP64 step on the same FP32-rounded inputs.  The roofline object
covers the kernel family with the largest share of the step, from HIP
events recorded around every launch on the stream the kernels run on.  The
CPU baseline is the oracle's OpenMP FP64 step on the box's host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--ppc 8]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("APIC p2g+g2p particles/s @256³×8ppc, FP32-comp vs FP64 at stated err; "
          "HBM GB/s")
HBM_PEAK_GBS = 8000.0
BRICK = (4, 4, 4)
CALIB_BYTES = 256 << 20  # device copy read once and written once
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Step:
    """Device buffers and one APIC step for one precision."""

    def __init__(self, pa, ctx, wl, dtype, x, m, c):
        import torch
        self.pa, self.ctx, self.torch = pa, ctx, torch
        self.dtype = dtype
        dev = "cuda:0"
        P = wl.np
        # SoA particle storage: [q, P] tensors viewed as [P, q].
        self.x = torch.as_tensor(np.ascontiguousarray(x.T), dtype=dtype, device=dev).t()
        self.m = torch.as_tensor(m[None, :], dtype=dtype, device=dev).t()
        self.c = torch.as_tensor(np.ascontiguousarray(c.reshape(P, 12).T), dtype=dtype,
                                 device=dev).t()
        self.c0 = self.c.clone()
        lo, n = wl.node_window
        nn = tuple(int(v) for v in n)
        self.gm = pa.GridField(torch.zeros(nn, dtype=dtype, device=dev))
        self.gmu = pa.GridField(torch.zeros((3,) + nn, dtype=dtype, device=dev).permute(1, 2, 3, 0))
        self.gu = pa.GridField(torch.zeros((3,) + nn, dtype=dtype, device=dev).permute(1, 2, 3, 0))
        self.mesh = pa.Mesh(wl.low, wl.h)
        # Bins are the node-centred dual cells: the knot base of quadratic
        # splines on nodes, so each bin's particles share one stencil.
        cell_lo, ncell = wl.cell_region
        self.geom = pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell),
                               BRICK, 1)
        nb = pa.bin_count(self.geom)
        i32 = dict(dtype=torch.int32, device=dev)
        self.bins = pa.BinSort(torch.empty(P, **i32), torch.empty(P, **i32),
                               torch.empty(nb + 1, **i32), self.geom, torch.empty(nb + 1, **i32))
        self.fused = True

    def restore(self):
        self.c.copy_(self.c0)

    def reorder(self, check=True):
        """Store the particles in bin order (what a PIC code does when it
        re-sorts its particle arrays: sort, then pic_permute into a second
        set of arrays); later sorts still run in full."""
        self.pa.bin_sort(self.ctx, self.mesh, self.x, self.geom, out=self.bins)
        if check:
            self.ctx.synchronize()  # raises if a particle fell outside the bins
        new = {name: self.torch.empty_like(getattr(self, name).t()).t()
               for name in ("x", "m", "c")}
        self.pa.permute(self.ctx, self.bins.perm, [(getattr(self, k), new[k]) for k in new])
        for k, v in new.items():
            setattr(self, k, v)
        self.c0 = self.c.clone()

    def run(self):
        pa, ctx = self.pa, self.ctx
        pa.bin_sort(ctx, self.mesh, self.x, self.geom, out=self.bins)
        sd = pa.create_spline(pa.FieldLocation.Node, 2, self.mesh, self.x, self.bins)
        pa.APIC.p2g(ctx, self.m, self.c, self.gm, self.gmu, sd, zero_first=True)
        if self.fused:
            # grid normalisation fused into g2p: u = mu/m formed in LDS
            pa.APIC.g2p_normalized(ctx, self.gm, self.gmu, self.c, sd)
        else:
            pa.grid_normalize(ctx, self.gm, self.gmu, self.gu,
                              rel_mass_eps=1e-6 if self.dtype is self.torch.float32 else 0.0)
            pa.APIC.g2p(ctx, self.gu, self.c, sd)


def time_steps(step, steps, warmup, ctx, families):
    import torch
    for _ in range(warmup):
        step.run()
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step.run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_timing(False)
    ctx.synchronize()
    kt = {}
    for f in families:
        ms, n = ctx.kernel_time(f)
        if n:
            kt[f] = {"ms_total": ms, "launches": n, "ms_avg": ms / n}
    ctx.reset_timing()
    return (t1 - t0) * 1e3 / steps, kt


def time_graph_steps(step, steps, warmup, ctx):
    """The same step captured once into a step graph (pic_graph_*) and
    replayed K times: identical kernels and arguments, without the host's
    per-launch gaps.  Runs on a stream of its own (capture needs one)."""
    import torch
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        step.run()  # eager, on this stream: the context's scratch is sized
        ctx.synchronize()
        ctx.graph_begin()
        step.run()
        g = ctx.graph_end()
        for _ in range(warmup):
            g.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.launch()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ctx.synchronize()
        g.close()
    return (t1 - t0) * 1e3 / steps


def cpu_baseline(wl, reps=3):
    """Oracle OpenMP FP64 step (atomic p2g) on the host cores."""
    from oracle import orpic
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    lo, n = wl.node_window
    x = np.ascontiguousarray(wl.x)
    c = np.ascontiguousarray(wl.c.copy())
    orpic.apic_step_omp(2, wl.low, wl.h, lo, n, x, wl.m, c, mode=1, nthreads=threads)  # warm-up
    ts = []
    for _ in range(reps):
        c[:] = wl.c
        t0 = time.perf_counter()
        orpic.apic_step_omp(2, wl.low, wl.h, lo, n, x, wl.m, c, mode=1, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"value": wl.np / t, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": f"full workload ({wl.np} particles), FP64 p2g(atomic)->normalise->g2p, "
                      f"median of {reps} after 1 warm-up, {t * 1e3:.1f} ms/step"}


def _time(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def _kernel_avgs(ctx, families):
    out = {}
    for f in families:
        ms, n = ctx.kernel_time(f)
        if n:
            out[f] = ms / n
    return out


def config_plumbing(pa, ctx, synth):
    """32^3 cells x 8 ppc, FP64: sort -> p2g -> g2p (normalisation fused)."""
    import torch
    wl = synth.apic_uniform(32, 8)
    st = Step(pa, ctx, wl, torch.float64, wl.x, wl.m, wl.c)
    st.reorder()
    ms = _time(st.run, 20, 3)
    return {"workload": "APIC uniform 32^3 cells x 8 ppc, quadratic node spline, FP64",
            "particles": wl.np, "ms_per_step": ms, "particles_per_s": wl.np / (ms * 1e-3)}


def config_f_update(pa, ctx, synth, steps=10):
    """Cube [1/4, 3/4]^3 of a 256^3 mesh at 8 ppc, FP32 particles and grid,
    FP64 F: sort -> p2g -> g2p fused with grad u and F <- (I + dt grad u) F,
    det F, F^-1 (FP64), 10 consecutive steps."""
    import torch
    wl = synth.apic_cube(256, 8)
    P = wl.np
    st = Step(pa, ctx, wl, torch.float32, wl.x.astype(np.float32), wl.m.astype(np.float32),
              wl.c.astype(np.float32))
    st.reorder()
    perm = st.bins.perm.long().cpu().numpy()
    F = torch.as_tensor(np.ascontiguousarray(wl.F.reshape(P, 9)[perm].T), device="cuda:0").t()
    F0 = F.clone()
    det = torch.empty((1, P), dtype=torch.float64, device="cuda:0").t()
    Finv = torch.empty((9, P), dtype=torch.float64, device="cuda:0").t()
    dt = 1e-3

    def step():
        pa.bin_sort(ctx, st.mesh, st.x, st.geom, out=st.bins)
        sd = pa.create_spline(pa.FieldLocation.Node, 2, st.mesh, st.x, st.bins)
        pa.APIC.p2g(ctx, st.m, st.c, st.gm, st.gmu, sd, zero_first=True)
        pa.APIC.g2p_update_F(ctx, st.gmu, st.c, dt, F, det, Finv, sd, m_i=st.gm)

    step()
    F.copy_(F0)
    st.restore()
    ctx.reset_timing()
    ctx.set_timing(True)
    ms = _time(step, steps, 0)
    ctx.set_timing(False)
    ctx.synchronize()
    kt = _kernel_avgs(ctx, ["bin_sort", "zero_grid", "apic_p2g", "apic_g2p"])
    ctx.reset_timing()
    finite = bool(torch.isfinite(F).all()) and bool(torch.isfinite(Finv).all())
    return {"workload": "APIC cube [0.25,0.75]^3 of 256^3 cells x 8 ppc, FP32 with FP64 F update, "
                        f"{steps} consecutive steps",
            "particles": P, "ms_per_step": ms, "particles_per_s": P / (ms * 1e-3),
            "kernels_ms": kt, "F_finite": finite}


def config_polypic(pa, ctx, synth, steps=5):
    """128^3 cells x 8 ppc, cubic spline, 27 velocity modes per component,
    FP32: PolyPIC p2g -> normalise -> g2p (per-particle kernels)."""
    import torch
    n, ppc = 128, 8
    st = synth.Stream(0x5EED0002)
    h = 1.0 / n
    ii = np.arange(n)
    cells = np.stack(np.meshgrid(ii, ii, ii, indexing="ij"), -1).reshape(-1, 3)
    x = synth._fill_cells(st, cells, ppc, h)
    P = x.shape[0]
    low = np.full(3, -synth.HALO * h)
    dev = "cuda:0"
    g = torch.Generator(device=dev)
    g.manual_seed(0x5EED0002)
    nm = pa.PolyPIC.nmode(3)
    # modes: N(0, 0.1), mode 0 ~ N(0, 1); SoA [27*3, P] viewed as [P, 81]
    modes = (torch.randn((nm * 3, P), generator=g, device=dev) * 0.1)
    modes[:3] = torch.randn((3, P), generator=g, device=dev)
    modes = modes.t()
    xt = torch.as_tensor(np.ascontiguousarray(x.T), dtype=torch.float32, device=dev).t()
    mt = torch.full((1, P), h ** 3 / ppc, dtype=torch.float32, device=dev).t()
    nn = n + 1 + 2 * synth.HALO
    gm = pa.GridField(torch.zeros((nn, nn, nn), dtype=torch.float32, device=dev))
    gmu = pa.GridField(torch.zeros((3, nn, nn, nn), dtype=torch.float32, device=dev).permute(1, 2, 3, 0))
    gu = pa.GridField(torch.zeros((3, nn, nn, nn), dtype=torch.float32, device=dev).permute(1, 2, 3, 0))
    cg = torch.empty_like(modes)
    mesh = pa.Mesh(low, h)
    # Mesh-cell bins (the cubic knot base), 4^3-bin bricks; particles stored
    # in bin order once, the sort redone every step.
    # n + 1 bins per dimension: a position in the last cell may round to
    # exactly 1.0 in FP32, which lies in the next cell.
    geom = pa.Binning((synth.HALO,) * 3, (n + 1,) * 3, (4, 4, 4), 0)
    bins = pa.bin_sort(ctx, mesh, xt, geom)
    ctx.synchronize()  # raises if a particle fell outside the bins (perm undefined then)
    perm = bins.perm.long()
    xt = xt.t()[:, perm].contiguous().t()
    modes = modes.t()[:, perm].contiguous().t()

    def step():
        pa.bin_sort(ctx, mesh, xt, geom, out=bins)
        sd = pa.create_spline(pa.FieldLocation.Node, 3, mesh, xt, bins)
        pa.PolyPIC.p2g(ctx, mt, modes, modes, gmu, gm, 1e-3, sd, zero_first=True)
        pa.grid_normalize(ctx, gm, gmu, gu, rel_mass_eps=1e-6)
        pa.PolyPIC.g2p(ctx, gu, cg, sd)

    step()
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    ms = _time(step, steps, 0)
    ctx.set_timing(False)
    ctx.synchronize()
    kt = _kernel_avgs(ctx, ["bin_sort", "zero_grid", "polypic_p2g", "polypic_p2g_stale", "normalize",
                            "max_mass", "polypic_g2p"])
    ctx.reset_timing()
    return {"workload": "PolyPIC uniform 128^3 cells x 8 ppc, cubic node spline, 27 modes x 3, "
                        "FP32, cell-bin sort every step, brick p2g and g2p",
            "particles": P, "ms_per_step": ms, "particles_per_s": P / (ms * 1e-3),
            "kernels_ms": kt}


def config_interp(pa, ctx, synth, steps=10):
    """Generic transfers (P2G::value of a vector, G2P::gradient of a vector
    field) on the 256^3 sphere, FP32, particles in bin order."""
    import torch
    wl = synth.apic_sphere(256, 8)
    P = wl.np
    dev = "cuda:0"
    mesh = pa.Mesh(wl.low, wl.h)
    x = torch.as_tensor(np.ascontiguousarray(wl.x.astype(np.float32).T), device=dev).t()
    cell_lo, ncell = wl.cell_region
    geom = pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell), BRICK, 1)
    bins = pa.bin_sort(ctx, mesh, x, geom)
    ctx.synchronize()
    xs = x.t()[:, bins.perm.long()].contiguous().t()  # particles stored in bin order
    bins = pa.bin_sort(ctx, mesh, xs, geom)
    v = torch.randn((3, P), device=dev).t()
    grad = torch.empty((9, P), device=dev).t()
    lo, n = wl.node_window
    f = pa.GridField(torch.zeros(tuple(int(k) for k in n) + (3,), device=dev),
                     tuple(int(k) for k in lo))
    sd = pa.create_spline(pa.FieldLocation.Node, 2, mesh, xs, bins)

    def step():
        pa.P2G.value(ctx, sd, v, f, zero_first=True)
        pa.G2P.gradient(ctx, sd, f, grad, ncomp=3)

    step()
    ctx.synchronize()
    ctx.reset_timing()
    ctx.set_timing(True)
    ms = _time(step, steps, 0)
    ctx.set_timing(False)
    ctx.synchronize()
    kt = _kernel_avgs(ctx, ["zero_grid", "p2g_interp", "g2p_interp"])
    ctx.reset_timing()
    return {"workload": "P2G::value (vector) + G2P::gradient (vector) on the APIC sphere "
                        "256^3 x 8 ppc, quadratic node spline, FP32, particles in bin order",
            "particles": P, "ms_per_step": ms, "particles_per_s": P / (ms * 1e-3),
            "kernels_ms": kt}


def config_mac(pa, ctx, synth, steps=10, steps_particles=3):
    """APIC on the three Face lattices of a MAC grid (Picasso_APIC.hpp:148-232,
    477-544) on the 256^3 sphere, FP32, quadratic splines, particles stored
    in node-bin order.  Per face: a sort on that face's knot-aligned bins (a
    dual flag per dimension), p2g into the face's mass and momentum grids,
    g2p with the normalisation fused into the face's own c_p (a staggered
    g2p zeroes the other columns).  The same transfers through the
    per-particle kernels (no bins) are timed beside it."""
    import torch
    wl = synth.apic_sphere(256, 8)
    P = wl.np
    dev = "cuda:0"
    mesh = pa.Mesh(wl.low, wl.h)
    x = torch.as_tensor(np.ascontiguousarray(wl.x.astype(np.float32).T), device=dev).t()
    m = torch.as_tensor(wl.m.astype(np.float32)[None, :], device=dev).t()
    c = torch.as_tensor(np.ascontiguousarray(wl.c.astype(np.float32).reshape(P, 12).T),
                        device=dev).t()
    cell_lo, ncell = wl.cell_region
    node_geom = pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell),
                           BRICK, 1)
    order = pa.bin_sort(ctx, mesh, x, node_geom).perm.long()
    ctx.synchronize()
    x = x.t()[:, order].contiguous().t()
    m = m.t()[:, order].contiguous().t()
    c = c.t()[:, order].contiguous().t()
    lo, n = wl.node_window
    shp, lot = tuple(int(k) for k in n), tuple(int(k) for k in lo)
    faces = []
    for d in range(3):
        ent = pa.FieldLocation.Face(d)
        dual = pa.aligned_dual(2, ent)
        geom = pa.Binning(tuple(int(v) for v in cell_lo),
                          tuple(int(v) + f for v, f in zip(ncell, pa.dual_dims(dual))), BRICK,
                          dual)
        bins = pa.bin_sort(ctx, mesh, x, geom)
        faces.append(dict(geom=geom, bins=bins, sd=pa.create_spline(ent, 2, mesh, x, bins),
                          sd_pp=pa.create_spline(ent, 2, mesh, x),
                          gm=pa.GridField(torch.zeros(shp, device=dev), lot),
                          gmu=pa.GridField(torch.zeros(shp, device=dev), lot),
                          cf=torch.zeros((12, P), device=dev).t()))

    def step_bricks():
        for f in faces:
            pa.bin_sort(ctx, mesh, x, f["geom"], out=f["bins"])
            pa.APIC.p2g(ctx, m, c, f["gm"], f["gmu"], f["sd"], zero_first=True)
            pa.APIC.g2p_normalized(ctx, f["gm"], f["gmu"], f["cf"], f["sd"])

    def step_particles():
        for f in faces:
            pa.APIC.p2g(ctx, m, c, f["gm"], f["gmu"], f["sd_pp"], zero_first=True)
            pa.APIC.g2p_normalized(ctx, f["gm"], f["gmu"], f["cf"], f["sd_pp"])

    fam = ["bin_sort", "zero_grid", "apic_p2g", "apic_g2p"]
    out = {"workload": "APIC MAC grid: the three Face lattices on the APIC sphere 256^3 x 8 ppc, "
                       "quadratic splines, FP32, per face sort -> p2g -> g2p (normalisation "
                       "fused), particles in node-bin order",
           "particles": P}
    for name, fn, k in (("bricks", step_bricks, steps), ("per_particle", step_particles,
                                                           steps_particles)):
        fn()
        ctx.synchronize()
        ctx.reset_timing()
        ctx.set_timing(True)
        ms = _time(fn, k, 0)
        ctx.set_timing(False)
        ctx.synchronize()
        out[name] = {"ms_per_step": ms, "particles_per_s": P / (ms * 1e-3),
                     "kernels_ms": _kernel_avgs(ctx, fam)}
        ctx.reset_timing()
    return out


def config_batched(pa, ctx, synth, n_inst=512, steps=3, lanes=4):
    """Independent instances: grid edge n_k = round(16 * 4^u_k), u_k ~ U(0, 1)
    (16^3 to 64^3 cells, log-uniform), ppc_k ~ U{2..16}, fields as the 32^3
    workload, FP32, quadratic node spline.  One step = for every instance:
    sort -> p2g -> g2p (normalisation fused).  The instances are dealt,
    largest first, to `lanes` streams with a context each (own scratch); each
    lane's part of the step is captured once into a step graph
    (pic_graph_*), and the lanes' graphs are replayed concurrently."""
    import torch
    dev = "cuda:0"
    rs = np.random.default_rng(0x5EED0003)
    ns = np.rint(16 * 4.0 ** rs.uniform(0, 1, n_inst)).astype(np.int64)
    ppcs = rs.integers(2, 17, n_inst)
    counts = ns ** 3 * ppcs
    P = int(counts.sum())
    offs = np.concatenate([[0], np.cumsum(counts)])
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x5EED0003)
    X = torch.empty((3, P), dtype=torch.float32, device=dev)
    M = torch.empty((1, P), dtype=torch.float32, device=dev)
    C = torch.empty((12, P), dtype=torch.float32, device=dev)
    side = torch.cuda.Stream()
    inst = []
    with torch.cuda.stream(side):
        for k in range(n_inst):
            n, ppc, a, b = int(ns[k]), int(ppcs[k]), int(offs[k]), int(offs[k + 1])
            h = 1.0 / n
            cell = torch.arange(n ** 3, device=dev).repeat_interleave(ppc)
            ijk = torch.stack([cell // (n * n), (cell // n) % n, cell % n]).to(torch.float64)
            x = ((ijk + torch.rand((3, b - a), generator=gen, device=dev, dtype=torch.float64)) * h)
            X[:, a:b] = x.to(torch.float32)
            M[:, a:b] = h ** 3 / ppc
            C[0:3, a:b] = torch.randn((3, b - a), generator=gen, device=dev)
            C[3:, a:b] = torch.randn((9, b - a), generator=gen, device=dev) * 0.1
            low = np.full(3, -synth.HALO * h)
            mesh = pa.Mesh(low, h)
            geom = pa.Binning((synth.HALO,) * 3, (n + 1,) * 3, (4, 4, 4), 1)
            xv = X[:, a:b].t()
            bins = pa.bin_sort(ctx, mesh, xv, geom)
            ctx.synchronize()  # raises if a particle fell outside the bins
            perm = bins.perm.long() + a
            for t in (X, M, C):
                t[:, a:b] = t[:, perm]
            nn = n + 1 + 2 * synth.HALO
            gm = pa.GridField(torch.zeros((nn, nn, nn), dtype=torch.float32, device=dev))
            gmu = pa.GridField(torch.zeros((3, nn, nn, nn), dtype=torch.float32,
                                           device=dev).permute(1, 2, 3, 0))
            inst.append(dict(mesh=mesh, geom=geom, bins=bins, x=X[:, a:b].t(), m=M[:, a:b].t(),
                             c=C[:, a:b].t(), gm=gm, gmu=gmu, np=b - a))
        C0 = C.clone()
    torch.cuda.synchronize()

    def run(cx, d):
        pa.bin_sort(cx, d["mesh"], d["x"], d["geom"], out=d["bins"])
        sd = pa.create_spline(pa.FieldLocation.Node, 2, d["mesh"], d["x"], d["bins"])
        pa.APIC.p2g(cx, d["m"], d["c"], d["gm"], d["gmu"], sd, zero_first=True)
        pa.APIC.g2p_normalized(cx, d["gm"], d["gmu"], d["c"], sd)

    # Eager pass on one stream: times every instance on its own.
    with torch.cuda.stream(side):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in inst]
        for d, (e0, e1) in zip(inst, ev):
            e0.record()
            run(ctx, d)
            e1.record()
        ctx.synchronize()
        per_inst = np.array([e0.elapsed_time(e1) for e0, e1 in ev])
    # Deal instances to lanes, largest first, to the least loaded lane.
    order = np.argsort(-counts)
    load = np.zeros(lanes)
    groups = [[] for _ in range(lanes)]
    for k in order:
        j = int(np.argmin(load))
        groups[j].append(int(k))
        load[j] += counts[k]
    streams = [torch.cuda.Stream() for _ in range(lanes)]
    ctxs = [pa.Context(0) for _ in range(lanes)]
    # Eager pass per lane: sizes each context's scratch before the capture.
    for cx, st, grp in zip(ctxs, streams, groups):
        with torch.cuda.stream(st):
            for k in grp:
                run(cx, inst[k])
    torch.cuda.synchronize()
    C.copy_(C0)
    torch.cuda.synchronize()
    # One step graph per lane, replayed on the lane's own stream so the
    # lanes run concurrently.
    t0 = time.perf_counter()
    graphs = []
    for cx, st, grp in zip(ctxs, streams, groups):
        with torch.cuda.stream(st):
            cx.graph_begin()
            for k in grp:
                run(cx, inst[k])
            graphs.append(cx.graph_end())
    t_capture = time.perf_counter() - t0

    def replay():
        for g_, st in zip(graphs, streams):
            with torch.cuda.stream(st):
                g_.launch()

    replay()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        replay()
        for st in streams:  # a step ends when every lane has finished
            side.wait_stream(st)
        for st in streams:
            st.wait_stream(side)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    for cx, g_ in zip(ctxs, graphs):
        cx.synchronize()
        g_.close()
        cx.close()
    return {"workload": f"{n_inst} APIC instances, 16^3-64^3 cells (log-uniform), 2-16 ppc, FP32, "
                        f"quadratic node spline, sort -> p2g -> g2p per instance, {lanes} concurrent "
                        "streams with one step graph each",
            "instances": n_inst, "particles": P, "ms_per_step": ms,
            "particles_per_s": P / (ms * 1e-3), "graph_capture_s": t_capture,
            "instance_ms_eager": {"min": float(per_inst.min()),
                                  "median": float(np.median(per_inst)),
                                  "max": float(per_inst.max())},
            "instance_particles": {"min": int(counts.min()), "median": int(np.median(counts)),
                                   "max": int(counts.max())}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-shuffled", action="store_true",
                    help="skip the reference step on shuffled storage (profiling runs)")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="time only the headline workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the transfer path is built and timed on a single "
              "MI355X; multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    import picasso_amd as pa
    from picasso_amd import synth

    t0 = time.perf_counter()
    wl = synth.apic_sphere(args.n, args.ppc)
    P = wl.np
    log(f"workload {wl.name}: {P} particles ({time.perf_counter() - t0:.1f} s)")
    n_touched = synth.touched_nodes(wl.x.astype(np.float32).astype(np.float64), wl.low, wl.h, 2)
    log(f"touched nodes {n_touched}")

    ctx = pa.Context(0)
    # One plain device copy of known size: under rocprofv3 --pmc its
    # FETCH_SIZE / WRITE_SIZE calibrate the counters against known bytes.
    calib = torch.empty(CALIB_BYTES // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    calib_dst = torch.empty_like(calib)
    calib_dst.copy_(calib)
    torch.cuda.synchronize()
    del calib, calib_dst
    x32, m32, c32 = wl.x.astype(np.float32), wl.m.astype(np.float32), wl.c.astype(np.float32)
    families = ["bin_sort", "zero_grid", "apic_p2g", "max_mass", "normalize", "apic_g2p"]

    s32 = Step(pa, ctx, wl, torch.float32, x32, m32, c32)
    shuffled_ms = shuffled_first_ms = None
    if not args.skip_shuffled:
        # The step on the shuffled storage order, for reference: reading the
        # particles through the permutation, and with the particle arrays
        # first stored in bin order (sort -> permute -> step).
        s32.run()
        s32.restore()
        torch.cuda.synchronize()
        t_sh = time.perf_counter()
        s32.run()
        torch.cuda.synchronize()
        shuffled_ms = (time.perf_counter() - t_sh) * 1e3
        s32.restore()
        log(f"fp32 step on shuffled storage, read through the permutation: {shuffled_ms:.3f} ms")
        keep = (s32.x, s32.m, s32.c, s32.c0)
        s32.reorder()
        s32.run()
        s32.x, s32.m, s32.c, s32.c0 = keep
        s32.restore()
        torch.cuda.synchronize()
        t_sh = time.perf_counter()
        s32.reorder(check=False)
        s32.run()
        torch.cuda.synchronize()
        shuffled_first_ms = (time.perf_counter() - t_sh) * 1e3
        s32.x, s32.m, s32.c, s32.c0 = keep
        log(f"fp32 first step from shuffled storage (sort, permute, step): "
            f"{shuffled_first_ms:.3f} ms")
    s32.reorder()
    ms32_eager, kt32 = time_steps(s32, args.steps, args.warmup, ctx, families)
    log(f"fp32: {ms32_eager:.3f} ms/step eager, kernels {json.dumps(kt32)}")
    ms32 = time_graph_steps(s32, args.steps, args.warmup, ctx)
    log(f"fp32: {ms32:.3f} ms/step as a step graph")

    s64 = Step(pa, ctx, wl, torch.float64, wl.x, wl.m, wl.c)
    s64.reorder()
    ms64_eager, kt64 = time_steps(s64, args.steps, args.warmup, ctx, families)
    log(f"fp64: {ms64_eager:.3f} ms/step eager, kernels {json.dumps(kt64)}")
    ms64 = time_graph_steps(s64, args.steps, args.warmup, ctx)
    log(f"fp64: {ms64:.3f} ms/step as a step graph")
    del s64

    # Error of the FP32 step against the FP64 step on the same rounded inputs.
    ref = Step(pa, ctx, wl, torch.float64, x32.astype(np.float64), m32.astype(np.float64),
               c32.astype(np.float64))
    ref.reorder()
    ref.run()
    s32.restore()
    s32.run()
    ctx.synchronize()

    def rel(a, b):
        a = a.double()
        b = b.double()
        return float((a - b).abs().max() / b.abs().max())

    err = {"grid_mass": rel(s32.gm.tensor, ref.gm.tensor),
           "grid_momentum": rel(s32.gmu.tensor, ref.gmu.tensor),
           "particle_velocity": rel(s32.c[:, :3], ref.c[:, :3]),
           "particle_B": rel(s32.c[:, 3:], ref.c[:, 3:])}
    log(f"fp32 error vs fp64: {err}")
    del ref

    s_alg32 = 4
    B_alg32 = P * 31 * s_alg32 + n_touched * 14 * s_alg32
    per_kernel_bytes = {
        "apic_p2g": P * 16 * s_alg32 + n_touched * 4 * s_alg32,
        "normalize": n_touched * 7 * s_alg32,
        "apic_g2p": P * 15 * s_alg32 + n_touched * 3 * s_alg32,
    }
    dom = max(kt32, key=lambda f: kt32[f]["ms_total"])
    roof = None
    # HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
    # on this workload, scaled by the calibration copy (see the .md it names).
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get(dom)
    if dom in per_kernel_bytes:
        gbs = per_kernel_bytes[dom] / (kt32[dom]["ms_avg"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": gbs, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "alg_bytes_per_launch": per_kernel_bytes[dom]}
        if traffic:
            roof["traffic_source"] = traffic["source"]
    else:
        roof = {"bound": "hbm", "kernel": dom, "achieved": None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None,
                "note": "dominant kernel has no algorithmic-byte definition (sort overhead)"}

    others = None
    if not args.no_other_configs:
        del s32
        torch.cuda.empty_cache()
        others = {}
        log("config: plumbing 32^3 FP64 ...")
        others["apic_uniform_32"] = config_plumbing(pa, ctx, synth)
        log(f"  {others['apic_uniform_32']}")
        log("config: cube 256^3 FP32 + FP64 F update ...")
        others["apic_cube_f_update"] = config_f_update(pa, ctx, synth)
        log(f"  {others['apic_cube_f_update']}")
        log("config: PolyPIC cubic 27 modes 128^3 FP32 ...")
        others["polypic_cubic27"] = config_polypic(pa, ctx, synth)
        log(f"  {others['polypic_cubic27']}")
        torch.cuda.empty_cache()
        log("config: generic transfers 256^3 FP32 ...")
        others["generic_interp"] = config_interp(pa, ctx, synth)
        log(f"  {others['generic_interp']}")
        torch.cuda.empty_cache()
        log("config: MAC grid, three Face lattices, 256^3 FP32 ...")
        others["apic_mac_faces"] = config_mac(pa, ctx, synth)
        log(f"  {others['apic_mac_faces']}")
        torch.cuda.empty_cache()
        log("config: 512 batched instances ...")
        others["batched_instances"] = config_batched(pa, ctx, synth)
        log(f"  {others['batched_instances']}")
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(wl)
        log(f"cpu baseline {cpu}")

    step_gbs = B_alg32 / (ms32 * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": P / (ms32 * 1e-3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64), inputs resident in HBM",
        "config": {"workload": f"APIC sphere r=0.3, {args.n}^3 cells x {args.ppc} ppc, "
                               "quadratic node spline, cell-bin sort every step, "
                               "step replayed as a captured HIP graph",
                   "particles": P, "cells": args.n, "ppc": args.ppc,
                   "touched_nodes": n_touched, "brick": list(BRICK)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "ms_per_step_eager": ms32_eager,
        "fp64": {"particles_per_s": P / (ms64 * 1e-3), "ms_per_step": ms64,
                 "ms_per_step_eager": ms64_eager, "kernels": kt64},
        "fp32_over_fp64": ms64 / ms32,
        "fp32_step_ms_shuffled_storage": shuffled_ms,
        "fp32_first_step_ms_from_shuffled": shuffled_first_ms,
        "fp32_error_vs_fp64": err,
        "step_alg_bytes": B_alg32,
        "step_alg_GBs": step_gbs,
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "kernels": kt32,
        "other_configs": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
