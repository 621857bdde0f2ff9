"""Host-side tracer fields: halo fill and initial conditions
(reference: src/Biogeochemistry/HaloRegions/halo_regions.jl, tracer_fields.jl)."""
import numpy as np

from ._abi import PERIODIC, TRACER_NAMES, TRACER_ID
from . import stream


def _fill_dim(c, axis, N, H, topo):
    """fill_halo_{west,east}! etc. on one axis; every copy snapshots its source first,
    like Julia's broadcast unaliasing of `@views @. c[...] = c[...]` (halo_regions.jl:1-16)."""
    def sl(a, b):
        s = [slice(None)] * 3
        s[axis] = slice(a, b)
        return tuple(s)
    if topo == PERIODIC:
        c[sl(0, H)] = c[sl(N, N + H)].copy()
        c[sl(N + H, N + 2 * H)] = c[sl(H, 2 * H)].copy()
    else:
        c[sl(0, H)] = c[sl(H, H + 1)].copy()
        c[sl(N + H, N + 2 * H)] = c[sl(N + H - 1, N + H)].copy()


def fill_halo_tracer(c, grid):
    """fill_halo_tracer! for one padded field (halo_regions.jl:37-47): west, east,
    south, north, top, bottom."""
    t = grid.topology_ids
    _fill_dim(c, 0, grid.Nx, grid.Hx, t[0])
    _fill_dim(c, 1, grid.Ny, grid.Hy, t[1])
    _fill_dim(c, 2, grid.Nz, grid.Hz, t[2])
    return c


def default_tracer_init():
    """default_tracer_init() — tracer_fields.jl:41-45."""
    init = dict(DIC=20.0, NH4=0.5, NO3=0.8, PO4=0.10, DFe=1.0e-6, DOC=10.0, DON=0.1, DOP=0.05,
                PFe_inorg=0.0, POC=0.0, PON=0.0, POP=0.0, PFe_bio=0.0, Dust=0.0)
    noise = {k: 0.1 for k in init}
    return dict(initial_condition=init, rand_noise=noise)


def _range_values(lower, upper, step=np.float32(1.0e-4)):
    n = int(round((float(upper) - float(lower)) / float(step))) + 1
    return (np.float64(lower) + np.arange(n, dtype=np.float64) * np.float64(step)).astype(np.float32)


def tracer_ranges(source):
    """(c0, lower, n) per tracer of the values source: tracer q starts at
    c0[q] * (lower[q] + k * 1f-4) with k < n[q] drawn per padded cell (the ranges
    generate_tracers draws from; pi_generate_state draws the same on the device)."""
    if not (isinstance(source, dict) and "initial_condition" in source):
        raise ValueError("generating the state on the device needs the (initial_condition, rand_noise) source")
    c0 = np.zeros(len(TRACER_NAMES), np.float32)
    lower = np.zeros(len(TRACER_NAMES), np.float32)
    n = np.zeros(len(TRACER_NAMES), np.int32)
    for q, name in enumerate(TRACER_NAMES):
        ic = source["initial_condition"]
        if name not in ic:
            raise ValueError(f"TRAC_INIT: tracer not found {name}")
        if ic[name] < 0.0:
            raise ValueError("TRAC_INIT:  The initial condition should be none-negetive.")
        eps = source["rand_noise"][name]
        lo, up = np.float32(1.0 - eps), np.float32(1.0 + eps)
        c0[q] = np.float32(ic[name])
        lower[q] = lo
        n[q] = len(_range_values(lo, up))
    return c0, lower, n


def generate_tracers(grid, source, seed):
    """generate_tracers(arch, grid, source, FT) — tracer_fields.jl:93-134.

    `source` is the reference's NamedTuple form {initial_condition, rand_noise}
    (uniform value times `rand(lower:1f-4:upper)` per padded cell, drawn here from
    the shared stream), or a dict tracer -> (Nx,Ny,Nz) array (the reference's
    Dict form, with arrays instead of file paths).  Returns 14 padded float32
    arrays (Fortran order) with halos filled.
    """
    shape = grid.padded_shape
    ncell = grid.ncells_padded
    out = []
    is_values = isinstance(source, dict) and "initial_condition" in source
    for name in TRACER_NAMES:
        if is_values:
            ic = source["initial_condition"]
            if name not in ic:
                raise ValueError(f"TRAC_INIT: tracer not found {name}")
            if ic[name] < 0.0:
                raise ValueError("TRAC_INIT:  The initial condition should be none-negetive.")
            eps = source["rand_noise"][name]
            lower, upper = np.float32(1.0 - eps), np.float32(1.0 + eps)
            vals = _range_values(lower, upper)
            u, _, _ = stream.uniform3(seed, 0, 0, stream.KIND_TRACER + TRACER_ID[name], np.arange(ncell))
            pick = np.minimum((u.astype(np.float64) * len(vals)).astype(np.int64), len(vals) - 1)
            c = (np.float32(ic[name]) * vals[pick]).reshape(shape, order="F")
            c = np.asfortranarray(c)
        else:
            if name not in source:
                raise ValueError(f"TRAC_INIT: tracer not found {name}")
            tmp = np.asarray(source[name])
            if np.sum(tmp < 0.0) > 0:
                raise ValueError("TRAC_INIT: The initial condition should be none-negetive.")
            if tmp.shape != (grid.Nx, grid.Ny, grid.Nz):
                raise ValueError("TRAC_INIT:  grid mismatch")
            c = np.zeros(shape, np.float32, order="F")
            c[grid.Hx:grid.Hx + grid.Nx, grid.Hy:grid.Hy + grid.Ny, grid.Hz:grid.Hz + grid.Nz] = tmp.astype(np.float32)
        # `@views @. tracers[name].data *= g.landmask` (Bool mask keeps the sign of zero)
        c = np.where(grid.landmask, c, np.copysign(np.float32(0), c)).astype(np.float32)
        c = np.asfortranarray(c)
        fill_halo_tracer(c, grid)
        out.append(c)
    return out
