"""PlanktonSimulation, diagnostics and update! (reference: src/Simulation/simulations.jl,
src/Simulation/update.jl, src/Simulation/utils.jl, src/Parameters/param_default.jl:1-38,
src/Diagnostics/diagnostics_struct.jl)."""
import math

import numpy as np

from . import _abi
from ._abi import BOUNDED
from .output import write_output
from .model import validate_bcs


def default_PARF(grid, ΔT, iterations):
    """default_PARF(grid, ΔT, iterations) — param_default.jl:5-18 (Float64, hourly)."""
    PAR = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.3871666666666666, 87.10258333333333, 475.78150000000016,
           929.2737916666669, 1232.3633333333337, 1638.918916666667, 1823.7921666666664,
           1906.2769583333336, 1776.0280416666667, 1678.5026249999999, 1410.216666666667,
           815.4129583333336, 525.104, 135.993, 2.9493750000000003, 0.0, 0.0, 0.0]
    total_days = int((ΔT * iterations) // 86400 + 1)
    day = np.zeros((grid.Nx, grid.Ny, 24))
    for i in range(24):
        day[:, :, i] = PAR[i]
    return np.tile(day, (1, 1, total_days))


def default_temperature(grid, ΔT, iterations):
    """default_temperature(grid, ΔT, iterations) — param_default.jl:20-38.  Note the
    vertical gradient uses grid.zC[j] with the unpadded index j (the padded array's
    j-th element)."""
    temp = [26.646446609406727, 26.56698729810778, 26.517037086855467, 26.5, 26.517037086855467,
            26.56698729810778, 26.646446609406727, 26.75, 26.87059047744874, 27.0, 27.12940952255126,
            27.25, 27.353553390593273, 27.43301270189222, 27.482962913144533, 27.5, 27.482962913144533,
            27.43301270189222, 27.353553390593273, 27.25, 27.12940952255126, 27.0, 26.87059047744874, 26.75]
    total_days = int((ΔT * iterations) // 86400 + 1)
    day = np.zeros((grid.Nx, grid.Ny, grid.Nz, 24))
    for i in range(24):
        day[:, :, -1, i] = temp[i]
    for j in range(grid.Nz - 1, 0, -1):            # j = Nz-1 : -1 : 1 (1-based)
        day[:, :, j - 1, :] = day[:, :, j, :] + 4.0e-4 * np.float64(grid.zC[j - 1])
    return np.tile(day, (1, 1, 1, total_days))


class PlanktonDiagnostics:
    """PlanktonDiagnostics(model; tracer=(), phytoplankton=(:num,:graz,:mort,:dvid,:ptc),
    abiotic_particle=(:num,), iteration_interval=1) — diagnostics_struct.jl:55-137.

    Accumulators live on the device; attribute reads download them."""

    def __init__(self, model, tracer=(), phytoplankton=("num", "graz", "mort", "dvid", "ptc"),
                 abiotic_particle=("num",), iteration_interval=1):
        avail_tr = ("PAR",) + _abi.TRACER_NAMES
        for t in tracer:
            if t not in avail_tr:
                raise ValueError(f"{t} is not one of the diagnostics")
        for p in phytoplankton:
            if p not in ("num", "graz", "mort", "dvid", "ptc"):
                raise NotImplementedError(f"per-individual rate diagnostic {p} is not on the MI355X hot path")
        self.model = model
        self.iteration_interval = int(iteration_interval)
        self.tracer_names = tuple(t for t in tracer if t != "PAR")
        # key orders of merge(NamedTuple{tracer}, (PAR,)) and of
        # merge(NamedTuple{phytoplankton}, (num, graz, mort, dvid, ptc))
        self._tracer_keys = tuple(tracer) + (() if "PAR" in tracer else ("PAR",))
        self._phyto_keys = tuple(phytoplankton) + tuple(
            p for p in ("num", "graz", "mort", "dvid", "ptc") if p not in phytoplankton)
        model._native.set_diag_tracers([_abi.TRACER_ID[t] for t in self.tracer_names])

    @property
    def tracer(self):
        d = {k: self.model._native.diag_tracer(_abi.TRACER_ID[k]) for k in self.tracer_names}
        d["PAR"] = self.model._native.grid_field("diag_PAR")
        return d

    @property
    def phytos(self):
        out = {}
        for s in range(self.model.n_species):
            d = {k: self.model._native.diag(s, k) for k in ("num", "graz", "mort", "dvid")}
            d["ptc"] = self.model._native.diag(s, "ptc")     # zero unless IronEnergyMode with abiotics
            out[f"sp{s + 1}"] = d
        return out

    @property
    def abiotics(self):
        """diags.abiotics.saN.num: abiotic particles per cell, accumulated per step."""
        n = getattr(getattr(self.model, "spec", None), "n_abiotic", 0)
        return {f"sa{a + 1}": {"num": self.model._native.abiotic_diag(a)} for a in range(n)}

    def reset(self):
        self.model._native.reset_diags()

    def tracer_keys(self):
        """keys(diags.tracer), in the reference's order."""
        return self._tracer_keys

    def phyto_keys(self):
        """keys(diags.phytos.spN), in the reference's order."""
        return self._phyto_keys


class PlanktonSimulation:
    """PlanktonSimulation(model; ΔT, iterations, PARF=nothing, temp=nothing, diags=nothing,
    vels=(;), ΔT_vel=ΔT, ΔT_PAR=3600f0, ΔT_temp=3600f0, output_writer=nothing)
    — simulations.jl:244-288."""

    def __init__(self, model, *, ΔT, iterations, PARF=None, temp=None, diags=None, vels=None,
                 ΔT_vel=None, ΔT_PAR=3600.0, ΔT_temp=3600.0, output_writer=None):
        self.output_writer = output_writer
        self.model = model
        self.ΔT = np.float32(ΔT)
        self.iterations = int(iterations)
        g = model.grid
        if PARF is None:
            PARF = default_PARF(g, ΔT, iterations)
        if temp is None:
            temp = default_temperature(g, ΔT, iterations)
        self.PARF = np.asarray(PARF, dtype=np.float32)
        self.temp = np.asarray(temp, dtype=np.float32)
        self.vels = None
        if vels:
            self.vels = tuple(np.asarray(vels[k], dtype=np.float32) for k in ("u", "v", "w"))
        self.ΔT_vel = np.float32(ΔT if ΔT_vel is None else ΔT_vel)
        self.ΔT_PAR = np.float32(ΔT_PAR)
        self.ΔT_temp = np.float32(ΔT_temp)
        self.diags = diags if diags is not None else PlanktonDiagnostics(model)
        self._staged = {"vel": None, "par": None, "temp": None}
        self.validate_velocity()
        self.validate_PARF()
        self.validate_temp()
        validate_bcs(model.tracers, g, self.iterations)       # simulations.jl:81
        abiotics = getattr(getattr(model, "individuals", None), "abiotics", {})
        validate_bcs(abiotics, g, self.iterations)             # simulations.jl:82

    # ---- validators (Simulation/utils.jl:36-127)
    def _check_step(self, ΔT_in, what):
        if math.floor(float(ΔT_in) / float(self.ΔT)) * float(self.ΔT) != float(ΔT_in):
            raise ValueError(f"ΔT_{what} must be an integer times of {self.ΔT}.")

    def validate_temp(self):
        g = self.model.grid
        if self.temp.shape[:3] != (g.Nx, g.Ny, g.Nz):
            raise ValueError(f"Dimension mismatch: the size of temperature must be {(g.Nx, g.Ny, g.Nz)}.")
        self._check_step(self.ΔT_temp, "temp")
        if self.temp.shape[3] < math.floor(self.iterations * float(self.ΔT) / float(self.ΔT_temp)):
            raise ValueError("Temperature provided cannot cover the time period of (iterations+1)*ΔT")

    def validate_PARF(self):
        g = self.model.grid
        if self.PARF.shape[:2] != (g.Nx, g.Ny):
            raise ValueError(f"Dimension mismatch: the size of PARF must be {(g.Nx, g.Ny)}.")
        self._check_step(self.ΔT_PAR, "PAR")
        if self.PARF.shape[2] < math.floor(self.iterations * float(self.ΔT) / float(self.ΔT_PAR)):
            raise ValueError("Surface PAR provided cannot cover the time period of (iterations+1)*ΔT")

    def validate_velocity(self):
        if self.vels is None:
            return
        g = self.model.grid
        t = g.topology_ids
        us = (g.Nx + (t[0] == BOUNDED), g.Ny, g.Nz)
        vs = (g.Nx, g.Ny + (t[1] == BOUNDED), g.Nz)
        ws = (g.Nx, g.Ny, g.Nz + (t[2] == BOUNDED))
        for a, s, n in zip(self.vels, (us, vs, ws), "uvw"):
            if a.shape[:3] != s:
                raise ValueError(f"Dimension mismatch: the size of {n} must be {s}.")
        self._check_step(self.ΔT_vel, "vel")
        if self.vels[0].shape[3] < math.floor(self.iterations * float(self.ΔT) / float(self.ΔT_vel)):
            raise ValueError("Velocities provided cannot cover the time period of (iterations+1)*ΔT")

    def set_vels_fields(self, u, v, w):
        self.vels = tuple(np.asarray(a, dtype=np.float32) for a in (u, v, w))
        self._staged["vel"] = None
        self.validate_velocity()

    def set_PARF_fields(self, PARF):
        self.PARF = np.asarray(PARF, dtype=np.float32)
        self._staged["par"] = None
        self.validate_PARF()

    def set_temp_fields(self, temp):
        self.temp = np.asarray(temp, dtype=np.float32)
        self._staged["temp"] = None
        self.validate_temp()


def _index(model_t, ΔT_in):
    """floor(Int, model_t / ΔT_in) + 1 evaluated in Float32 (update.jl:18-26), 0-based."""
    return int(np.floor(np.float32(model_t) / np.float32(ΔT_in)))


def update(sim, time_offset=None, check_every=1):
    """update!(sim; time_offset=(vels=false, PARF=false, temp=false)) — update.jl:11-48.

    Staging copies the input slice to the device only when its index changes (the
    reference copies every step; the device result is the same).  `check_every`
    controls how often the device error state is checked (1 = after every step,
    like the reference's synchronous exception)."""
    off = time_offsets(time_offset)
    m = sim.model
    nat = m._native
    writer = getattr(sim, "output_writer", None)
    for i in range(1, sim.iterations + 1):
        stage_inputs(sim, i, off)
        nat.step(np.float32(sim.ΔT))
        advance_clock(sim)
        if (check_every and (i % check_every == 0)) or writer is not None:
            nat.sync()
        if writer is not None:
            write_output(writer, m, sim.diags)               # write_output! (update.jl:30, 45)
    nat.sync()


def time_offsets(time_offset=None):
    off = dict(vels=False, PARF=False, temp=False)
    if time_offset:
        off.update(time_offset)
    return off


def stage_inputs(sim, i, off):
    """The input staging of iteration i (1-based) of update! (update.jl:18-26)."""
    m = sim.model
    nat = m._native
    base = (i - 1) * float(sim.ΔT)
    if sim.vels is not None:
        tv = _index(np.float32(base) if off["vels"] else m.t, sim.ΔT_vel)
        if sim._staged["vel"] != tv:
            nat.set_velocity(sim.vels[0][..., tv], sim.vels[1][..., tv], sim.vels[2][..., tv])
            sim._staged["vel"] = tv
    tp = _index(np.float32(base) if off["PARF"] else m.t, sim.ΔT_PAR)
    if sim._staged["par"] != tp:
        nat.set_parf(sim.PARF[:, :, tp])
        sim._staged["par"] = tp
    tt = _index(np.float32(base) if off["temp"] else m.t, sim.ΔT_temp)
    if sim._staged["temp"] != tt:
        nat.set_temperature(sim.temp[:, :, :, tt])
        sim._staged["temp"] = tt


def advance_clock(sim):
    """model.t = model.iteration * ΔT (time_step.jl:15)."""
    m = sim.model
    m.t = np.float32(np.float32(m._native.iteration()) * sim.ΔT)
