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
# The CPU baseline's OpenMP threads stay on their cores; the runtime reads
# these when it starts, before torch is imported.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
try:  # read before the OpenMP runtime binds this thread to one core
    CPUS_AVAILABLE = len(os.sched_getaffinity(0))
except AttributeError:
    CPUS_AVAILABLE = os.cpu_count()

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
        self.order = np.arange(P)  # input index of the particle in each storage slot
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
        # One pic_apic_step call per step (grid cleared inside the sort's key
        # pass, walk listed once); False: the three calls one after another.
        self.one_call = True

    def restore(self):
        self.c.copy_(self.c0)

    def reorder(self, check=True):
        """Store the particles in bin order (what a PIC code does when it
        re-sorts its particle arrays: sort, then pic_permute into a second
        set of arrays); later sorts still run in full.  With check=False
        only the device work is done (the timed form): no synchronisation,
        and the bench's own bookkeeping (the reset copy of c_p, the input
        order of each slot) is left stale."""
        self.pa.bin_sort(self.ctx, self.mesh, self.x, self.geom, out=self.bins)
        if check:
            self.ctx.synchronize()  # raises if a particle fell outside the bins
        new = {name: self.torch.empty_like(getattr(self, name).t()).t()
               for name in ("x", "m", "c")}
        self.pa.permute(self.ctx, self.bins.perm, [(getattr(self, k), new[k]) for k in new])
        for k, v in new.items():
            setattr(self, k, v)
        if check:
            self.c0 = self.c.clone()
            self.order = self.order[self.bins.perm.long().cpu().numpy()]

    def outputs(self):
        """Grid mass, momentum and c_p in input order, as float64 numpy."""
        c = np.empty((len(self.order), 12))
        c[self.order] = self.c.double().cpu().numpy()
        return (self.gm.tensor.double().cpu().numpy(),
                self.gmu.tensor.double().cpu().numpy(), c.reshape(-1, 4, 3))

    def run(self):
        pa, ctx = self.pa, self.ctx
        if self.one_call and self.fused:
            sd = pa.create_spline(pa.FieldLocation.Node, 2, self.mesh, self.x)
            pa.APIC.step(ctx, self.m, self.c, self.gm, self.gmu, sd, self.geom, out=self.bins)
            return
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


def _cpu_info():
    """CPU model and physical core count of this host, and the CPUs this
    process may run on."""
    model, cores = None, set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, len(cores) or None, CPUS_AVAILABLE


def oracle_step(wl, x, m, c, threads=0):
    """The oracle's FP64 p2g -> normalise -> g2p on a zeroed node window:
    grid mass, momentum (window shaped) and c_p."""
    from oracle import orpic
    lo, n = wl.node_window
    cc = np.ascontiguousarray(c.copy())
    gm, gmu, _ = orpic.apic_step_omp(2, wl.low, wl.h, lo, n, np.ascontiguousarray(x), m, cc,
                                     mode=1, nthreads=threads)
    shp = tuple(int(v) for v in n)
    return gm.reshape(shp), gmu.reshape(shp + (3,)), cc.reshape(-1, 4, 3)


def field_errors(got, ref):
    """||got - ref||_inf / ||ref||_inf of grid mass, grid momentum, particle
    velocity (c_p row 0) and B (rows 1-3)."""
    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())
    return {"grid_mass": rel(got[0], ref[0]), "grid_momentum": rel(got[1], ref[1]),
            "particle_velocity": rel(got[2][:, 0], ref[2][:, 0]),
            "particle_B": rel(got[2][:, 1:], ref[2][:, 1:])}


def cpu_baseline(wl, synth, reps=5):
    """The oracle's OpenMP FP64 step (atomic p2g) on the host cores of the
    process: the whole workload, median of `reps` runs after one warm-up,
    on every thread OpenMP may use here (OMP_NUM_THREADS, else the CPUs
    this process may run on), threads bound to cores; and a one-thread run
    on the 32^3 x 8 ppc workload.  Returns the baseline record and the
    oracle's outputs of the full workload."""
    from oracle import orpic
    model, phys, avail = _cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    lo, n = wl.node_window
    x = np.ascontiguousarray(wl.x)

    def run(wlx, xx, nt, keep=False):
        cc = np.ascontiguousarray(wlx.c.copy())
        lo_, n_ = wlx.node_window
        t0 = time.perf_counter()
        gm, gmu, _ = orpic.apic_step_omp(2, wlx.low, wlx.h, lo_, n_, xx, wlx.m, cc, mode=1,
                                         nthreads=nt)
        t = time.perf_counter() - t0
        return t, ((gm, gmu, cc) if keep else None)

    run(wl, x, threads)  # warm-up
    ts, out = [], None
    for r in range(reps):
        t, o = run(wl, x, threads, keep=(r == reps - 1))
        ts.append(t)
        out = o if o is not None else out
    t = float(np.median(ts))
    small = synth.apic_uniform(32, 8)
    xs = np.ascontiguousarray(small.x)
    run(small, xs, 1)
    t1 = float(np.median([run(small, xs, 1)[0] for _ in range(reps)]))
    shp = tuple(int(v) for v in n)
    ref = (out[0].reshape(shp), out[1].reshape(shp + (3,)), out[2].reshape(-1, 4, 3))
    rec = {"value": wl.np / t, "unit": "particles/s", "cores": threads, "kind": "port",
           "sample": f"full workload ({wl.np} particles), FP64 p2g(atomic)->normalise->g2p, "
                     f"{threads} OpenMP threads bound to cores, median of {reps} after 1 "
                     f"warm-up, {t * 1e3:.1f} ms/step",
           "cpu_model": model, "physical_cores_on_host": phys, "cpus_available": avail,
           "one_core": {"value": small.np / t1, "unit": "particles/s",
                        "sample": f"32^3 x 8 ppc ({small.np} particles), FP64, 1 thread, "
                                  f"median of {reps}, {t1 * 1e3:.1f} ms/step"}}
    return rec, ref


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


def config_batched(pa, ctx, synth, n_inst=512, steps=5):
    """Independent instances: grid edge n_k = round(16 * 4^u_k), u_k ~ U(0, 1)
    (16^3 to 64^3 cells, log-uniform), ppc_k ~ U{2..16}, fields as the 32^3
    workload, FP32, quadratic node spline.  One step = bin sort -> p2g (grids
    zeroed first) -> g2p with the normalisation fused, for every instance,
    through the batch entry points (pic_batch_*): each pass is one launch
    over all instances' bricks, and the whole step is captured once in a
    step graph and replayed.  The per-instance spread is the eager time of
    each instance run alone through the single-instance entry points."""
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
    inst, single = [], []
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
            mesh = pa.Mesh(np.full(3, -synth.HALO * h), h)
            geom = pa.Binning((synth.HALO,) * 3, (n + 1,) * 3, (4, 4, 4), 1)
            bins = pa.bin_sort(ctx, mesh, X[:, a:b].t(), geom)
            ctx.synchronize()  # raises if a particle fell outside the bins
            perm = bins.perm.long() + a  # store the instance's particles in bin order
            for t in (X, M, C):
                t[:, a:b] = t[:, perm]
            nn = n + 1 + 2 * synth.HALO
            gm = pa.GridField(torch.zeros((nn, nn, nn), dtype=torch.float32, device=dev))
            gmu = pa.GridField(torch.zeros((3, nn, nn, nn), dtype=torch.float32,
                                           device=dev).permute(1, 2, 3, 0))
            inst.append(pa.Instance(mesh, a, b - a, geom, gm, gmu))
            single.append(dict(mesh=mesh, geom=geom, bins=bins, x=X[:, a:b].t(),
                               m=M[:, a:b].t(), c=C[:, a:b].t(), gm=gm, gmu=gmu))
        C0 = C.clone()
        # Per-instance spread: each instance alone through the single-instance
        # entry points, eagerly.
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in single]
        for d, (e0, e1) in zip(single, ev):
            e0.record()
            pa.bin_sort(ctx, d["mesh"], d["x"], d["geom"], out=d["bins"])
            sd = pa.create_spline(pa.FieldLocation.Node, 2, d["mesh"], d["x"], d["bins"])
            pa.APIC.p2g(ctx, d["m"], d["c"], d["gm"], d["gmu"], sd, zero_first=True)
            pa.APIC.g2p_normalized(ctx, d["gm"], d["gmu"], d["c"], sd)
            e1.record()
        ctx.synchronize()
        per_inst = np.array([e0.elapsed_time(e1) for e0, e1 in ev])
        sequential_ms = float(per_inst.sum())
        C.copy_(C0)
        # The batch: one launch per pass over every instance, captured once.
        batch = pa.Batch(ctx, 2, torch.float32, inst)
        xs, ms_, cs = X.t(), M.t(), C.t()
        batch.step(ctx, xs, ms_, cs)  # eager warm-up
        ctx.synchronize()
        C.copy_(C0)
        t0 = time.perf_counter()
        ctx.graph_begin()
        batch.step(ctx, xs, ms_, cs)
        graph = ctx.graph_end()
        t_capture = time.perf_counter() - t0
        graph.launch()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            graph.launch()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        ctx.synchronize()
        # kernel families of one eager step
        ctx.reset_timing()
        ctx.set_timing(True)
        batch.step(ctx, xs, ms_, cs)
        ctx.set_timing(False)
        ctx.synchronize()
        kt = _kernel_avgs(ctx, ["batch_sort", "batch_zero", "batch_p2g", "batch_p2g_stale",
                                "batch_g2p", "batch_g2p_stale"])
        ctx.reset_timing()
        graph.close()
        batch.close()
    n_touched = int(((ns + 3) ** 3).sum())  # upper bound: every node of each window's cells
    alg = P * 31 * 4 + n_touched * 14 * 4
    return {"workload": f"{n_inst} APIC instances, 16^3-64^3 cells (log-uniform), 2-16 ppc, FP32, "
                        "quadratic node spline, particles stored in bin order per instance; "
                        "batched step (sort -> p2g -> normalised g2p, one launch per pass over "
                        "all instances) replayed as one step graph",
            "instances": n_inst, "particles": P, "ms_per_step": ms,
            "particles_per_s": P / (ms * 1e-3), "graph_capture_s": t_capture,
            "kernels_ms": kt,
            "alg_bytes_upper": alg, "alg_GBs_upper": alg / (ms * 1e-3) / 1e9,
            "sequential_single_instance_ms": sequential_ms,
            "instance_ms_eager": {"min": float(per_inst.min()),
                                  "median": float(np.median(per_inst)),
                                  "max": float(per_inst.max())},
            "instance_particles": {"min": int(counts.min()), "median": int(np.median(counts)),
                                   "max": int(counts.max())}}


def config_moving(pa, ctx, synth, steps=32):
    """The 256^3 sphere in FP32 with moving particles (B_p = 0 at the start):
    every step advects x += dt v (dt = h / 2, v from the g2p), so bins go
    stale and storage order decays.  The tuner (picasso_amd/tune.py) times every combination
    of re-sort interval, re-permute interval and g2p brick edge over short
    windows of live steps and keeps the fastest; `steps` further steps are
    then timed with it.  One step = [sort] -> p2g (zero first) -> normalised
    g2p -> advection [-> sort + permute into bin order]."""
    import torch
    from picasso_amd import tune
    wl = synth.apic_sphere(256, 8)
    P = wl.np
    dev = "cuda:0"
    x = torch.as_tensor(np.ascontiguousarray(wl.x.astype(np.float32).T), device=dev).t()
    m = torch.as_tensor(wl.m.astype(np.float32)[None, :], device=dev).t()
    # B_p starts at zero: the synthetic N(0, 0.1) entries are no velocity
    # gradient, and their affine term D^-1 B d ~ 0.4 / h would dominate the
    # velocity of nodes few particles reach; advected, it would carry those
    # particles out of the domain.  The first g2p makes B consistent.
    c0 = wl.c.astype(np.float32).copy()
    c0[:, 1:, :] = 0.0
    c = torch.as_tensor(np.ascontiguousarray(c0.reshape(P, 12).T), device=dev).t()
    mesh = pa.Mesh(wl.low, wl.h)
    cell_lo, ncell = wl.cell_region
    lo_c, n_c = tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell)
    geom = pa.Binning(lo_c, n_c, (4, 4, 4), 1)
    geom8 = pa.Binning(lo_c, n_c, (8, 8, 8), 1)
    lo, n = wl.node_window
    nn = tuple(int(v) for v in n)
    gm = pa.GridField(torch.zeros(nn, device=dev))
    gmu = pa.GridField(torch.zeros((3,) + nn, device=dev).permute(1, 2, 3, 0))
    step = tune.MovingStep(pa, ctx, mesh, geom, geom8, x, m, c, gm, gmu, 0.5 * wl.h)
    step.config = tune.StepConfig(1, 1, 4)
    step.run()  # stores the particles in bin order once
    ctx.synchronize()
    t0 = time.perf_counter()
    best, table = tune.StepTuner().tune(step)
    t_tune = time.perf_counter() - t0
    step.step_no = 0
    ctx.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step.run()
    e1.record()
    torch.cuda.synchronize()
    ctx.synchronize()  # raises if a particle left the bins or a window
    ms = e0.elapsed_time(e1) / steps
    return {"workload": "APIC sphere 256^3 x 8 ppc, FP32, moving particles (x += dt v, "
                        "dt = h/2): [sort] -> p2g -> normalised g2p -> advect [-> permute], "
                        f"{steps} steps with the tuned configuration",
            "particles": P, "ms_per_step": ms, "particles_per_s": P / (ms * 1e-3),
            "tuned": best.as_dict(), "tuning_s": t_tune, "candidates": table}


OTHER_CONFIGS = [
    ("plumbing", "apic_uniform_32", "plumbing 32^3 FP64", config_plumbing),
    ("f_update", "apic_cube_f_update", "cube 256^3 FP32 + FP64 F update", config_f_update),
    ("polypic", "polypic_cubic27", "PolyPIC cubic 27 modes 128^3 FP32", config_polypic),
    ("interp", "generic_interp", "generic transfers 256^3 FP32", config_interp),
    ("mac", "apic_mac_faces", "MAC grid, three Face lattices, 256^3 FP32", config_mac),
    ("batched", "batched_instances", "512 batched instances", config_batched),
    ("moving", "apic_moving_tuned", "moving particles 256^3 FP32, autotuned", config_moving),
]


def run_others(pa, ctx, synth, only=""):
    """The other workloads of the scope (all, or the comma-separated subset
    `only` by short name)."""
    import torch
    want = {w for w in only.split(",") if w}
    others = {}
    for short, key, title, fn in OTHER_CONFIGS:
        if want and short not in want:
            continue
        log(f"config: {title} ...")
        others[key] = fn(pa, ctx, synth)
        log(f"  {others[key]}")
        torch.cuda.empty_cache()
    return others


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the comparison of the headline outputs with the FP64 oracle")
    ap.add_argument("--skip-shuffled", action="store_true",
                    help="skip the reference step on shuffled storage (profiling runs)")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="time only the headline workload")
    ap.add_argument("--only", default="",
                    help="comma-separated subset of the other configs to time (plumbing, "
                         "f_update, polypic, interp, mac, batched, moving)")
    ap.add_argument("--skip-headline", action="store_true",
                    help="skip the headline workload (with --only, for a config's own run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the transfer path is built and timed on a single "
              "MI355X; multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch

    import picasso_amd as pa
    from picasso_amd import synth

    if args.skip_headline:
        ctx = pa.Context(0)
        print(json.dumps({"other_configs": run_others(pa, ctx, synth, args.only)}))
        return

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
    # The same copy timed: the HBM bandwidth a plain device copy reaches on
    # this box (bytes read + written), quoted beside the 8 TB/s peak.
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        calib_dst.copy_(calib)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * CALIB_BYTES * 20 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
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
        keep = (s32.x, s32.m, s32.c, s32.c0, s32.order)
        s32.reorder()
        s32.run()
        s32.x, s32.m, s32.c, s32.c0, s32.order = keep
        s32.restore()
        torch.cuda.synchronize()
        t_sh = time.perf_counter()
        s32.reorder(check=False)
        s32.run()
        torch.cuda.synchronize()
        shuffled_first_ms = (time.perf_counter() - t_sh) * 1e3
        s32.x, s32.m, s32.c, s32.c0, s32.order = keep
        log(f"fp32 first step from shuffled storage (sort, permute, step): "
            f"{shuffled_first_ms:.3f} ms")
    s32.reorder()
    ms32_eager, kt32 = time_steps(s32, args.steps, args.warmup, ctx, families)
    log(f"fp32: {ms32_eager:.3f} ms/step eager, kernels {json.dumps(kt32)}")
    ms32 = time_graph_steps(s32, args.steps, args.warmup, ctx)
    log(f"fp32: {ms32:.3f} ms/step as a step graph")
    s32.one_call = False
    ms32_three = time_graph_steps(s32, args.steps, args.warmup, ctx)
    s32.one_call = True
    log(f"fp32: {ms32_three:.3f} ms/step as a step graph of the three calls")

    s64 = Step(pa, ctx, wl, torch.float64, wl.x, wl.m, wl.c)
    s64.reorder()
    ms64_eager, kt64 = time_steps(s64, args.steps, args.warmup, ctx, families)
    log(f"fp64: {ms64_eager:.3f} ms/step eager, kernels {json.dumps(kt64)}")
    ms64 = time_graph_steps(s64, args.steps, args.warmup, ctx)
    log(f"fp64: {ms64:.3f} ms/step as a step graph")
    s64.one_call = False
    ms64_three = time_graph_steps(s64, args.steps, args.warmup, ctx)
    s64.one_call = True
    log(f"fp64: {ms64_three:.3f} ms/step as a step graph of the three calls")
    gpu_out = {}
    if not args.no_oracle_check:
        s64.restore()
        s64.run()
        ctx.synchronize()
        gpu_out["fp64"] = s64.outputs()
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
    if not args.no_oracle_check:
        gpu_out["fp32"] = s32.outputs()

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
                "alg_bytes_per_launch": per_kernel_bytes[dom],
                "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": gbs / copy_gbs}
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
        others = run_others(pa, ctx, synth, args.only)

    cpu = None
    oracle_err = {}
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu, ref64 = cpu_baseline(wl, synth)
        log(f"cpu baseline {cpu}")
        if "fp64" in gpu_out:
            oracle_err["fp64"] = field_errors(gpu_out["fp64"], ref64)
            log(f"fp64 error vs the FP64 oracle: {oracle_err['fp64']}")
        del ref64
    if "fp32" in gpu_out:
        # the FP64 oracle on the FP32-rounded inputs the FP32 step computed from
        ref32 = oracle_step(wl, x32.astype(np.float64), m32.astype(np.float64),
                            c32.astype(np.float64))
        oracle_err["fp32"] = field_errors(gpu_out["fp32"], ref32)
        log(f"fp32 error vs the FP64 oracle: {oracle_err['fp32']}")
        del ref32

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
        "ms_per_step_three_calls": ms32_three,
        "fp64": {"particles_per_s": P / (ms64 * 1e-3), "ms_per_step": ms64,
                 "ms_per_step_eager": ms64_eager, "ms_per_step_three_calls": ms64_three,
                 "kernels": kt64},
        "fp32_over_fp64": ms64 / ms32,
        "fp32_step_ms_shuffled_storage": shuffled_ms,
        "fp32_first_step_ms_from_shuffled": shuffled_first_ms,
        "fp32_error_vs_fp64": err,
        "fp32_error_vs_oracle": oracle_err.get("fp32"),
        "fp64_error_vs_oracle": oracle_err.get("fp64"),
        "step_alg_bytes": B_alg32,
        "step_alg_GBs": step_gbs,
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "step_frac_of_measured_copy": step_gbs / copy_gbs,
        "kernels": kt32,
        "other_configs": others,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
