"""Autotuning of a moving-particle APIC step.

Particles move every step (x += dt v), so a cell-bin sort goes stale and the
storage order drifts away from the bin order.  Three choices trade those
costs against each other:

- ``resort``: sort every ``resort`` steps.  In between, the brick kernels
  reuse the last bins; particles that left their bin take the per-particle
  path (correct, slower).
- ``permute``: every ``permute`` steps, store the particle arrays in bin
  order (pic_permute of x, m, c_p), after which the sort finds its keys
  already ordered and the brick kernels read particles with unit stride.
  ``None`` never does.
- ``g2p_brick``: the brick edge of the g2p binning: 4 reuses the p2g bins,
  8 sorts a second time into 8^3-bin bricks whose footprints share more
  nodes.
- ``g2p_group``: particles per g2p workgroup pass, as the workgroup size
  (128 work-items with four particles each, or 256 with three; the
  context option PIC_OPT_G2P_GROUP).

``StepTuner.tune`` times every candidate over a short window of real steps
on the live workload (HIP events around the window) and keeps the fastest.
The p2g brick shape and its staging capacity are compile-time constants of
the p2g kernels (4^3-bin bricks, 616 FP32 particles per chunk), not tuned
here.
"""
import itertools
from dataclasses import dataclass
from typing import Optional


@dataclass(frozen=True)
class StepConfig:
    resort: int = 1
    permute: Optional[int] = None
    g2p_brick: int = 4
    g2p_group: int = 0  # 0: the library's default (256)

    def as_dict(self):
        return {"resort": self.resort, "permute": self.permute, "g2p_brick": self.g2p_brick,
                "g2p_group": self.g2p_group}


class MovingStep:
    """One APIC step of a moving particle set (FP32 or FP64, quadratic Node
    splines): sort (every ``resort`` steps), p2g with the grid zeroed first,
    g2p of the normalised grid velocity (the three in one pic_apic_step call
    on the steps that sort with 4^3-bin g2p bricks), advection x += dt v,
    and the particle arrays stored in bin order every ``permute`` steps.

    x, m, c are [P, q] tensors (any strides); ``geom`` bins the quadratic
    knot base with 4^3-bin bricks; ``geom8`` the same cells in 8^3 bricks
    (for ``g2p_brick = 8``).  Every step sets the context's g2p workgroup
    option (``Context.set_g2p_group``) to the configuration's, and leaves it
    so."""

    def __init__(self, pa, ctx, mesh, geom, geom8, x, m, c, gm, gmu, dt, config=StepConfig()):
        import torch
        self.pa, self.ctx, self.torch = pa, ctx, torch
        self.mesh, self.geom, self.geom8 = mesh, geom, geom8
        self.x, self.m, self.c = x, m, c
        self.gm, self.gmu, self.dt = gm, gmu, dt
        self.config = config
        self.bins = pa.bin_sort(ctx, mesh, x, geom)
        self.bins8 = pa.bin_sort(ctx, mesh, x, geom8)
        self.step_no = 0

    def _permute(self):
        pa = self.pa
        new = {k: self.torch.empty_like(getattr(self, k).t()).t() for k in ("x", "m", "c")}
        pa.permute(self.ctx, self.bins.perm, [(getattr(self, k), new[k]) for k in new])
        for k, v in new.items():
            setattr(self, k, v)
        # Bins for the new storage order, unless the next step sorts anyway.
        if self.step_no % self.config.resort != 0:
            self.bins = pa.bin_sort(self.ctx, self.mesh, self.x, self.geom, out=self.bins)

    def run(self):
        pa, ctx, cf = self.pa, self.ctx, self.config
        ctx.set_g2p_group(cf.g2p_group)
        if self.step_no % cf.resort == 0 and cf.g2p_brick == 4:
            # sort, p2g and g2p in one call (pic_apic_step)
            sd = pa.create_spline(pa.FieldLocation.Node, 2, self.mesh, self.x)
            pa.APIC.step(ctx, self.m, self.c, self.gm, self.gmu, sd, self.geom, out=self.bins)
        else:
            if self.step_no % cf.resort == 0:
                pa.bin_sort(ctx, self.mesh, self.x, self.geom, out=self.bins)
            sd = pa.create_spline(pa.FieldLocation.Node, 2, self.mesh, self.x, self.bins)
            pa.APIC.p2g(ctx, self.m, self.c, self.gm, self.gmu, sd, zero_first=True)
            if cf.g2p_brick == 8:
                pa.bin_sort(ctx, self.mesh, self.x, self.geom8, out=self.bins8)
                sd = pa.create_spline(pa.FieldLocation.Node, 2, self.mesh, self.x, self.bins8)
            pa.APIC.g2p_normalized(ctx, self.gm, self.gmu, self.c, sd)
        self.x.add_(self.c[:, :3], alpha=self.dt)
        self.step_no += 1
        if cf.permute is not None and self.step_no % cf.permute == 0:
            pa.bin_sort(ctx, self.mesh, self.x, self.geom, out=self.bins)
            self._permute()


class StepTuner:
    """Picks the fastest StepConfig for a MovingStep by timing each
    candidate over `window` steps (after `warmup`) on the live particles,
    which keep moving throughout."""

    def __init__(self, resort=(1, 2), permute=(1, 4, 16, None), g2p_brick=(4, 8),
                 g2p_group=(128, 256)):
        self.candidates = [StepConfig(r, p, g, w) for r, p, g, w in
                           itertools.product(resort, permute, g2p_brick, g2p_group)]

    def tune(self, step, window=8, warmup=2):
        import torch
        timings = []
        for cf in self.candidates:
            step.config = cf
            step.step_no = 0
            for _ in range(warmup):
                step.run()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(window):
                step.run()
            e1.record()
            torch.cuda.synchronize()
            timings.append((e0.elapsed_time(e1) / window, cf))
        best_ms, best = min(timings, key=lambda t: t[0])
        step.config = best
        return best, [{"ms_per_step": ms, **cf.as_dict()} for ms, cf in timings]
