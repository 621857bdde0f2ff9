"""Serving an ensemble stream on one GPU: parameter-sweep requests arrive over
time, are admitted while their device footprint fits a memory budget, keep
their state resident on the device, and are stepped together with every other
admitted member by batched launches (continuous batching: members join and
leave the running batch between steps).  The reference has no counterpart; each
member is stepped exactly as `update` would step it alone.

Members are prepared (device allocation and the initial state drawn on the
device, pi_generate_state) by worker threads while the batch steps; a member
that finishes leaves the batch and its device memory is released by a worker.
"""
import collections
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native
from .ensemble import Ensemble
from .model import PlanktonModel, HIPArch
from .simulation import PlanktonSimulation, default_PARF, default_temperature


class Request:
    """One member run: `make()` returns (ModelSpec, velocities or None) and
    `steps` is its length; `arrival` is the time it entered the server."""

    def __init__(self, rid, make, steps, arrival=None, footprint=None, tag=None):
        self.rid, self.make, self.steps = rid, make, int(steps)
        self.arrival = arrival
        self.footprint = footprint
        self.tag = tag
        self.admitted = self.started = self.finished = None
        self.result = None
        self.error = None


def simulation_for(spec, vel, steps, dt=60.0, device=0):
    """PlanktonSimulation of a member: default diel PAR and temperature, steady
    velocity (if any)."""
    m = PlanktonModel(HIPArch(device), spec=spec)
    g = spec.grid
    kw = {}
    if vel is not None:
        kw = dict(vels=dict(u=vel[0][..., None], v=vel[1][..., None], w=vel[2][..., None]), ΔT_vel=dt * 1e6)
    return PlanktonSimulation(m, ΔT=dt, iterations=steps, PARF=default_PARF(g, dt, steps),
                              temp=default_temperature(g, dt, steps).astype(np.float32), **kw)


def member_footprint(spec):
    return _native.footprint(spec.grid, spec.n_species, spec.max_individuals, spec.species_params,
                             spec.bgc_vector, spec.seed, mode=spec.mode_id)


class EnsembleServer:
    """Admission by footprint (`budget` bytes), preparation on `workers`
    threads, one batched step per `step()` call.  `on_finish(req, sim)` (optional)
    runs before a finished member's memory is released."""

    def __init__(self, budget=200 * 2 ** 30, dt=60.0, device=0, workers=4, on_finish=None):
        self.budget, self.dt, self.device = int(budget), np.float32(dt), device
        self.ens = Ensemble(device=device)
        self.ens.batch.graphs(False)      # workers make HIP calls while the batch runs
        self.pool = ThreadPoolExecutor(workers)
        self.queue = collections.deque()
        self.preparing = []
        self.active = {}                  # sim -> request
        self.used = 0
        self.finished = []
        self.step_ms = []                 # wall time of each batched step
        self.member_steps = 0             # Σ members per step
        self.individual_steps = 0         # Σ over members and steps of the live individuals stepped
        self.on_finish = on_finish
        self._lock = threading.Lock()
        self._n = {}
        self.releasing = []

    def submit(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        spec, _ = req.make()
        req.footprint = member_footprint(spec)
        if req.footprint > self.budget:
            req.error = "footprint exceeds the budget"
            self.finished.append(req)
            return
        self.queue.append(req)

    def _prepare(self, req):
        spec, vel = req.make()
        sim = simulation_for(spec, vel, req.steps, float(self.dt), self.device)
        return req, sim, sum(spec.N)

    def _admit(self):
        while self.queue:
            with self._lock:
                if self.used + self.queue[0].footprint > self.budget:
                    return
                req = self.queue.popleft()
                self.used += req.footprint
            req.admitted = time.perf_counter()
            self.preparing.append((req, self.pool.submit(self._prepare, req)))

    def _collect(self):
        still = []
        for req, f in self.preparing:
            if not f.done():
                still.append((req, f))
                continue
            try:
                _, sim, n = f.result()
            except Exception as e:                       # creation failed: give the memory back
                req.error = repr(e)
                with self._lock:
                    self.used -= req.footprint
                self.finished.append(req)
                continue
            req.started = time.perf_counter()
            self.ens.add(sim)
            self.active[sim] = req
            self._n[sim] = n
        self.preparing = still

    def _release(self, sim, req):
        try:
            if self.on_finish is not None:
                self.on_finish(req, sim)
        finally:
            sim.model._native.close()
            with self._lock:
                self.used -= req.footprint

    def step(self):
        """Admit, add ready members, run one batched step, retire finished ones.
        Returns False when there is nothing to do."""
        self._admit()
        self._collect()
        if not self.ens.sims:
            return bool(self.queue or self.preparing)
        t0 = time.perf_counter()
        self.ens.step(1)
        self.ens.batch.sync()
        self.step_ms.append((time.perf_counter() - t0) * 1e3)
        self.member_steps += len(self.ens.sims)
        for sim in list(self.ens.sims):
            if self.ens.done[sim] >= sim.iterations:
                req = self.active.pop(sim)
                req.finished = time.perf_counter()
                self.ens.remove(sim)
                try:
                    sim.model._native.sync()
                except _native.CapacityError as e:
                    req.error = str(e)
                req.result = [sim.model._native.count_active(s) for s in range(sim.model.n_species)]
                # live individuals of every step it took (counted on the device, pi_individual_steps)
                self.individual_steps += sim.model._native.individual_steps()
                self.finished.append(req)
                del self._n[sim]
                self.releasing.append(self.pool.submit(self._release, sim, req))
        return True

    def drain(self, deadline=None):
        while self.step():
            if deadline is not None and time.perf_counter() > deadline:
                break
        for f in self.releasing:
            f.result()
        self.releasing = []

    def stats(self):
        """Step and end-to-end latency percentiles; individual_steps counts every
        step taken so far, members still running included."""
        st = np.array(self.step_ms) if self.step_ms else np.zeros(1)
        running = sum(sim.model._native.individual_steps() for sim in self.ens.sims)
        done = [r for r in self.finished if r.error is None and r.finished is not None]
        e2e = np.array([r.finished - r.arrival for r in done]) if done else np.zeros(1)
        return dict(steps=len(self.step_ms), p50_step_ms=float(np.percentile(st, 50)),
                    p99_step_ms=float(np.percentile(st, 99)), mean_step_ms=float(st.mean()),
                    finished=len(done), p50_e2e_s=float(np.percentile(e2e, 50)),
                    p99_e2e_s=float(np.percentile(e2e, 99)), member_steps=self.member_steps,
                    individual_steps=self.individual_steps + running)

    def close(self):
        for f in self.releasing:
            f.result()
        self.pool.shutdown(wait=True)
        for sim in list(self.ens.sims):
            self.ens.remove(sim)
            sim.model._native.close()
        self.ens.close()
