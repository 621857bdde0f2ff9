"""Ensembles of PlanktonSimulations on one GPU.

`Ensemble` steps many simulations together: at every kernel of the step one
launch covers all members (libpi_hip.so's pi_batch_*), so small members no
longer pay a full launch sequence each.  Every member is stepped exactly as
`update` (update!, Simulation/update.jl:11-48) steps it alone — same staging,
same kernels on its own buffers — so its results are bit-identical to a run on
its own.  The reference has no ensemble driver; it runs one simulation per
`update!` call.
"""
import numpy as np

from ._native import NativeBatch, CapacityError
from .simulation import time_offsets, stage_inputs, advance_clock


class Ensemble:
    """Simulations sharing one ΔT, stepped by batched launches.

    `step(n)` advances every member n steps; `update()` runs every member to
    its own `iterations` (members that finish leave the batch).  A member whose
    step fails on the device (capacity) is reported in `failed` and removed;
    the others continue."""

    def __init__(self, sims=(), time_offset=None, device=0, check_every=0):
        self.batch = NativeBatch(device)
        self.sims = []
        self.done = {}          # sim -> steps taken in this ensemble
        self.failed = {}        # sim -> exception
        self.ΔT = None
        self.off = time_offsets(time_offset)
        self.check_every = int(check_every)
        for s in sims:
            self.add(s)

    def add(self, sim):
        dt = np.float32(sim.ΔT)
        if self.ΔT is None:
            self.ΔT = dt
        elif dt != self.ΔT:
            raise ValueError("members of an Ensemble share one ΔT")
        self.batch.add(sim.model._native)
        self.sims.append(sim)
        self.done[sim] = 0

    def remove(self, sim):
        self.batch.remove(sim.model._native)
        self.sims.remove(sim)

    def __len__(self):
        return len(self.sims)

    def step(self, n=1):
        for _ in range(n):
            if not self.sims:
                return
            for s in self.sims:
                stage_inputs(s, self.done[s] + 1, self.off)
            self.batch.run(self.ΔT, 1)
            for s in self.sims:
                self.done[s] += 1
                advance_clock(s)
            if self.check_every and self.done[self.sims[0]] % self.check_every == 0:
                self.check()

    def autotune(self, steps=4, reps=3, shapes=None, log=None):
        """Searches the batch's block budget and the launch shapes of every
        member-shape class by timing real steps of the ensemble (the members
        advance; no choice changes any result): autotune.tune_batch.  The
        steps are taken with the inputs staged as `step` stages them."""
        from .autotune import tune_batch

        class _Runner:
            members = self.batch.members

            def __init__(r):
                r.budget = self.batch.budget

            def set_budget(r, b):
                self.batch.set_budget(b)

            def run(r, dt, k):
                self.step(k)

            def sync(r):
                self.batch.sync()

        res = tune_batch(_Runner(), self.ΔT, steps=steps, reps=reps, shapes=shapes, log=log)
        self.budget = res["budget"]
        return res

    def check(self):
        """Synchronises and moves members in a device error state to `failed`."""
        self.batch.sync()
        for s in list(self.sims):
            try:
                s.model._native.sync()
            except CapacityError as e:
                self.failed[s] = e
                self.remove(s)

    def update(self):
        """Every member to its own `iterations` (update! for all of them)."""
        while self.sims:
            left = [s.iterations - self.done[s] for s in self.sims]
            k = min(left)
            if k > 0:
                self.step(k)
            for s in list(self.sims):
                if self.done[s] >= s.iterations:
                    self.remove(s)
        self.check()

    def close(self):
        for s in list(self.sims):
            self.remove(s)
        self.batch.close()
