"""The dynamic scene: blob meshes rebuilt every frame under one instance accel, traced plainly and
with the learned proposal, the proposal fine-tuned online on the frame's own plain-traversal hits.

Every frame f:
  * every vertex = base + U(-amplitude, amplitude) (device generator, seed 7000 + f);
  * the triangle accels are rebuilt in one graph-captured batch (gprtAccelBuildBatch) and the
    instance accel on top of them (gprtAccelBuild);
  * side^2 pinhole primary rays are traced (gprtAccelTraceClosestHitDetail: hit, instance, top-level
    leaf), side^2 cosine-weighted bounce rays are generated from their hits
    (gprtSynthBounceRays) and traced the same way;
  * every `tune_every` frames the proposal is fine-tuned on that frame's rays, labelled by the
    plain traversal's top-level leaves (gprtProposalTrainLabels), and its confidence gate is
    calibrated on the same rays (gprtProposalCalibrateGate: the logit threshold that minimises
    their node fetches; +infinity when proposing to no ray is cheapest, and the learned traces then
    skip the inference);
  * both ray sets are traced again with the proposal first (instances and hits written to the
    `*_learned` buffers), which must give the plain traversal's hits.

Only C-ABI calls: the loop is a user of the library, shared by the benchmark and the tests.
"""
from __future__ import annotations

import numpy as np

from . import api, scenes


class DynamicScene:
    def __init__(self, ctx: api.Context, meshes: int = 64, per: int = 65536, side: int = 1024,
                 seed0: int = 5000, amplitude: float = 1e-3, max_batch: int = 4096, proposal_seed: int = 1):
        self.ctx, self.lib = ctx, ctx.lib
        self.meshes, self.per, self.amplitude = meshes, per, amplitude
        self.sizes = np.full(meshes, per, dtype=np.int64)
        self.base, self.ibuf = api.synth_blob_meshes(ctx, self.sizes, seed0, 0.002)
        self.cur = ctx.device_buffer(nbytes=self.base.nbytes)
        self.floats = self.base.nbytes // 4
        self.perturb(0)
        self.blas = api.mesh_accels(ctx, self.cur, self.ibuf, self.sizes)
        api.build_batch(ctx, self.blas)
        self.tlas = api.InstanceAccel(ctx, api.instance_records(self.blas))
        self.tlas.build()
        self.prim = scenes.pinhole_rays(side, side)
        n = self.n = len(self.prim)
        self.pb = ctx.device_buffer(self.prim)
        self.bb = ctx.device_buffer(nbytes=32 * n)
        # plain: hits, instances, top-level leaves; learned: hits, instances
        self.hp, self.hb, self.hp2, self.hb2 = (ctx.device_buffer(nbytes=16 * n) for _ in range(4))
        self.ip, self.ib, self.lp, self.lb, self.ip2, self.ib2 = (ctx.device_buffer(nbytes=4 * n) for _ in range(6))
        self.prop = api.Proposal(ctx, max_batch=max_batch, seed=proposal_seed)
        self.classes = self.prop.bind(self.tlas)
        self.cal = None  # the frame's primary and bounce rays together, for the gate calibration
        self.gate_keep = None  # share of the calibration rays the gate kept at the last fine-tune
        self.leaves_per_class = int(self.lib.gprtAccelGetPrimCount(self.tlas.h)) // self.classes

    @property
    def rays_per_frame(self) -> int:
        return 2 * self.n

    def perturb(self, f: int):
        self.lib.gprtSynthPerturb(self.ctx.h, self.base.h, self.cur.h, self.floats, 7000 + f, self.amplitude)

    def rebuild(self, f: int):
        self.perturb(f)
        api.build_batch(self.ctx, self.blas)
        self.tlas.build()

    def trace_plain(self, f: int):
        ctx, n = self.ctx, self.n
        api.trace_detail(ctx, self.tlas, self.pb, self.hp, n, self.ip, self.lp)
        api.bounce_rays(ctx, self.tlas, self.pb, self.hp, self.ip, n, self.bb, n, seed=7000 + f)
        api.trace_detail(ctx, self.tlas, self.bb, self.hb, n, self.ib, self.lb)

    def fine_tune(self, rays: int | None = None, epochs: int = 2, lr: float = 1e-3, calibrate: bool = True):
        """Minibatch Adam over the frame's primary and bounce rays (the first `rays` / 2 of each;
        all of them by default), labelled by the plain traversal's top-level leaves; then (with
        `calibrate`) the confidence gate from all of the frame's rays."""
        k = self.n if rays is None else min(self.n, rays // 2)
        l1 = self.prop.train_labels(self.pb, self.lp, k, epochs=epochs, lr=lr)
        l2 = self.prop.train_labels(self.bb, self.lb, k, epochs=epochs, lr=lr)
        if calibrate:
            n = self.n
            if self.cal is None:
                self.cal = self.ctx.device_buffer(nbytes=2 * 32 * n)
            self.lib.gprtBufferCopy(self.ctx.h, self.pb.h, self.cal.h, 0, 0, 32, n, 0, 0)
            self.lib.gprtBufferCopy(self.ctx.h, self.bb.h, self.cal.h, 0, n, 32, n, 0, 0)
            self.gate_keep = float(self.lib.gprtProposalCalibrateGate(self.ctx.h, self.prop.h, self.tlas.h,
                                                                      self.cal.h, 2 * n))
        return l1, l2

    def trace_learned(self):
        ctx, n = self.ctx, self.n
        api.trace_detail(ctx, self.tlas, self.pb, self.hp2, n, self.ip2, proposal=self.prop)
        api.trace_detail(ctx, self.tlas, self.bb, self.hb2, n, self.ib2, proposal=self.prop)

    def frame(self, f: int, tune_every: int = 16, tune_rays: int | None = None, epochs: int = 2,
              lr: float = 1e-3, on_tune=None, calibrate: bool = True):
        """One frame; returns the fine-tune's (loss, top-1) pairs in frames that have one.
        `on_tune(before: bool)` is called around the fine-tune (the benchmark pauses its kernel
        timer there)."""
        self.rebuild(f)
        self.trace_plain(f)
        tuned = None
        if tune_every and f % tune_every == 0:
            if on_tune:
                on_tune(True)
            tuned = self.fine_tune(tune_rays, epochs, lr, calibrate)
            if on_tune:
                on_tune(False)
        self.trace_learned()
        return tuned

    def mismatches(self) -> int:
        """Rays whose learned-trace hit or instance differs from the plain trace's (both sets)."""
        n, bad = self.n, 0
        for h1, h2, i1, i2 in ((self.hp, self.hp2, self.ip, self.ip2), (self.hb, self.hb2, self.ib, self.ib2)):
            a = h1.read(api.HIT_DTYPE, n).view(np.uint32).reshape(n, 4)
            b = h2.read(api.HIT_DTYPE, n).view(np.uint32).reshape(n, 4)
            bad += int(((a != b).any(1) | (i1.read(np.int32, n) != i2.read(np.int32, n))).sum())
        return bad

    def proposal_hit_rate(self) -> float:
        """Share of hit primary rays whose proposed class holds the closest hit's top-level leaf."""
        n = self.n
        cls = self.ctx.device_buffer(nbytes=4 * n)
        self.lib.gprtProposalInfer(self.ctx.h, self.prop.h, self.pb.h, n, cls.h)
        lab, c = self.lp.read(np.int32, n), cls.read(np.int32, n)
        cls.destroy()
        m = lab >= 0
        return float((c[m] == lab[m] // self.leaves_per_class).mean()) if m.any() else 0.0

    def destroy(self):
        for b in (self.pb, self.bb, self.hp, self.hb, self.hp2, self.hb2, self.ip, self.ib, self.lp, self.lb,
                  self.ip2, self.ib2, self.cur, self.cal):
            if b is not None:
                b.destroy()
        self.prop.destroy()
        self.tlas.destroy()
        for a in self.blas:
            a.destroy()
        self.base.destroy()
        self.ibuf.destroy()
