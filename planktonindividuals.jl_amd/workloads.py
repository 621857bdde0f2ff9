"""Synthetic workloads used by the benchmark and the parity tests.

  cellular_flow   2-D steady cellular flow from a streamfunction on cell
                  corners, u = -δψ/Δy, v = δψ/Δx (discretely divergence-free;
                  the pattern of the reference's examples/helper_functions.jl:38-59).
  fourier_flow    3-D field from a sum of random Fourier modes with a k^-5/3
                  energy spectrum: a rotational part (streamfunction on
                  corners) plus a divergent part whose vertical structure has
                  zero column mean, w from continuity (zero at the top and
                  bottom faces), rescaled to a target rms speed.
"""
import numpy as np

from .grids import RectilinearGrid, Periodic, Bounded
from .model import ModelSpec, phyto_setup


def cellular_flow(grid, U=0.1, L=16e3):
    """ψ = (U L/π) sin(πx/L) sin(πy/L) at corners; returns u (Nx,Ny,Nz), v, w (Nx,Ny,Nz+1)."""
    Nx, Ny, Nz = grid.Nx, grid.Ny, grid.Nz
    dx, dy = float(grid.Δx), float(grid.Δy)
    xc = np.arange(Nx + 1) * dx
    yc = np.arange(Ny + 1) * dy
    psi = (U * L / np.pi) * np.sin(np.pi * xc[:, None] / L) * np.sin(np.pi * yc[None, :] / L)
    u2 = -(psi[:Nx, 1:] - psi[:Nx, :-1]) / dy
    v2 = (psi[1:, :Ny] - psi[:-1, :Ny]) / dx
    u = np.repeat(u2[:, :, None], Nz, axis=2).astype(np.float32)
    v = np.repeat(v2[:, :, None], Nz, axis=2).astype(np.float32)
    w = np.zeros((Nx, Ny, Nz + 1), np.float32)
    return np.asfortranarray(u), np.asfortranarray(v), np.asfortranarray(w)


def fourier_flow(grid, nmodes=64, rms=0.05, seed=0, kmax=None):
    """Periodic-in-x/y, bounded-in-z synthetic velocity (see module doc)."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz = grid.Nx, grid.Ny, grid.Nz
    dx, dy = float(grid.Δx), float(grid.Δy)
    dz = grid.dzF[grid.Hz:grid.Hz + Nz].astype(np.float64)
    Lx, Ly = Nx * dx, Ny * dy
    kmax = kmax or max(2, min(Nx, Ny) // 4)
    xf = np.arange(Nx + 1) * dx             # x faces / corners
    yf = np.arange(Ny + 1) * dy
    xc = (np.arange(Nx) + 0.5) * dx
    yc = (np.arange(Ny) + 0.5) * dy
    kk = (np.arange(Nz) + 0.5) / Nz           # normalised depth of cell centres
    # every mode's vertical structure is one of three cosines (streamfunction)
    # or two (divergent part): the horizontal patterns are summed per vertical
    # structure and expanded in depth once
    psi2 = np.zeros((3, Nx + 1, Ny + 1))
    chi2 = np.zeros((2, Nx, Ny))
    for m in range(nmodes):
        while True:
            nx_, ny_ = rng.integers(-kmax, kmax + 1, size=2)
            if nx_ or ny_:
                break
        kx, ky = 2 * np.pi * nx_ / Lx, 2 * np.pi * ny_ / Ly
        kmag = np.hypot(kx, ky)
        amp = np.sqrt(np.hypot(nx_, ny_) ** (-5.0 / 3.0)) / kmag
        ph = rng.uniform(0, 2 * np.pi)
        nz_ = int(rng.integers(0, 3))
        psi2[nz_] += amp * np.cos(kx * xf[:, None] + ky * yf[None, :] + ph)
        ph2 = rng.uniform(0, 2 * np.pi)
        chi2[m % 2] += 0.3 * amp * np.cos(kx * xc[:, None] + ky * yc[None, :] + ph2)
    psi = np.zeros((Nx + 1, Ny + 1, Nz))
    for a in range(3):
        psi += psi2[a][:, :, None] * np.cos(np.pi * a * kk)[None, None, :]
    chi = np.zeros((Nx, Ny, Nz))
    for a in range(2):
        chi += chi2[a][:, :, None] * np.cos(2 * np.pi * (1 + a) * kk)[None, None, :]   # zero column mean
    # C-grid: u on x faces (i = 0..Nx-1, periodic), v on y faces
    u = -(psi[:Nx, 1:, :] - psi[:Nx, :-1, :]) / dy
    v = (psi[1:, :Ny, :] - psi[:-1, :Ny, :]) / dx
    u += (chi - np.roll(chi, 1, axis=0)) / dx
    v += (chi - np.roll(chi, 1, axis=1)) / dy
    s = rms / np.sqrt(np.mean(u ** 2 + v ** 2))
    u *= s
    v *= s
    div = (np.roll(u, -1, axis=0) - u) / dx + (np.roll(v, -1, axis=1) - v) / dy
    # w positive upward on z faces k = 0 (top) .. Nz (bottom); bottom = 0, integrate upward
    w = np.zeros((Nx, Ny, Nz + 1))
    for k in range(Nz - 1, -1, -1):
        w[:, :, k] = w[:, :, k + 1] - dz[k] * div[:, :, k]
    w[:, :, 0] = 0.0
    f = lambda a: np.asfortranarray(a.astype(np.float32))
    return f(u), f(v), f(w)


# ------------------------------------------------------------- configurations
def noisy_tracer_init(eps=0.05):
    """The default uniform tracer values (tracer_fields.jl:41-45) with noise of
    relative amplitude `eps` on every tracer, in the reference's own noise form:
    c0 * rand(1-eps : 1f-4 : 1+eps) per cell (generate_tracers,
    tracer_fields.jl:93-134).  The workloads use eps = 0.05."""
    from .fields import default_tracer_init
    init = default_tracer_init()
    init["rand_noise"] = {k: float(eps) for k in init["rand_noise"]}
    return init


def config_2d(N=32768, n=128, cap=None, seed=1, overrides=None, bgc=None, host_state=True):
    """Plain 2-D horizontal run: 128x128x1 periodic, Δ = 1 km, one species,
    tracers uniform plus 5 % noise."""
    grid = RectilinearGrid(size=(n, n, 1), x=(0, n * 1000), y=(0, n * 1000), z=(0, -10),
                           topology=(Periodic, Periodic, Bounded))
    params = dict(overrides or {})
    spec = ModelSpec(grid, max_individuals=cap or 2 * N, phyto=phyto_setup(params or None, [N], 1),
                     bgc_params=bgc, seed=seed, tracer_initial=noisy_tracer_init(), host_state=host_state)
    return spec, cellular_flow(grid, U=0.1, L=16e3 * n / 128)


# Overrides that make grazing, death and division happen within a short run
# (the defaults give no death above Sz = 0.5 and no grazing).
ACTIVE_PHYSIOLOGY = {"mort_reg": [1.0], "grz_P": [1.0e-6], "dvid_reg": [1.8]}


def config_3d(n=256, nz=64, N=4194304, nsp=2, cap=None, seed=2, nmodes=64, overrides=None,
              dz=10.0, kappa=(1e-2, 1e-5), host_state=True):
    """Plain 3-D run: n x n x nz, periodic x/y, bounded z, nsp species, Fourier
    flow, tracers uniform plus 5 % noise."""
    grid = RectilinearGrid(size=(n, n, nz), x=(0, n * 1000), y=(0, n * 1000), z=(0, -nz * dz),
                           topology=(Periodic, Periodic, Bounded))
    ov = dict(ACTIVE_PHYSIOLOGY)
    ov.update(overrides or {})
    params = {k: v * nsp if len(v) == 1 else v for k, v in ov.items()}
    bgc = {"κh": kappa[0], "κv": kappa[1], "κhP": kappa[0], "κvP": kappa[1]}
    spec = ModelSpec(grid, max_individuals=cap or 2 * N, phyto=phyto_setup(params, [N] * nsp, nsp),
                     bgc_params=bgc, seed=seed, tracer_initial=noisy_tracer_init(), host_state=host_state)
    return spec, fourier_flow(grid, nmodes=nmodes, rms=0.05, seed=seed)


# ------------------------------------------------------------- ensembles
# The ensemble workloads: members drawn from one seed, each a 0-D box (1x1x1,
# 1k-64k individuals), a 1-D column (1x1x64, 8k-256k) or a 2-D slab
# (64-256 squared, 16k-1M), one species, with VNH4max, VNO3max, VPO4max and
# dvid_P scaled by independent U[0.8, 1.2] factors.
PERTURBED = ("VNH4max", "VNO3max", "VPO4max", "dvid_P")


def ensemble_table(n=256, seed=4, lengths=None):
    """Member table: dicts with kind (0 = box, 1 = column, 2 = slab), size,
    N, parameter overrides, member seed and (when `lengths` = (lo, hi)) a
    number of steps drawn uniformly from [lo, hi]."""
    from .params import phyt_params_default, QuotaMode
    rng = np.random.default_rng(seed)
    base = phyt_params_default(1, QuotaMode())
    out = []
    for i in range(n):
        kind = int(rng.integers(0, 3))
        if kind == 0:
            size, N = (1, 1, 1), int(rng.integers(1000, 64001))
        elif kind == 1:
            size, N = (1, 1, 64), int(rng.integers(8000, 256001))
        else:
            m = int(rng.integers(64, 257))
            size, N = (m, m, 1), int(rng.integers(16000, 1000001))
        f = rng.uniform(0.8, 1.2, len(PERTURBED))
        over = {k: [float(np.float32(np.float32(base[k][0]) * np.float32(fk)))] for k, fk in zip(PERTURBED, f)}
        steps = int(rng.integers(lengths[0], lengths[1] + 1)) if lengths else None
        out.append(dict(id=i, kind=kind, size=size, N=N, params=over, seed=int(seed) * 100003 + i, steps=steps))
    return out


def ensemble_member(entry, host_state=False, cap_factor=2):
    """(ModelSpec, velocities or None) of one member: 1 km boxes and columns
    with 10 m levels and no flow; slabs of 1 km cells under the cellular flow
    (U = 0.1 m/s, L = 16 km per 128 cells); capacity cap_factor x N (a
    favourable population doubles in about 1.1 days of 60 s steps)."""
    size, N = entry["size"], entry["N"]
    if size[0] == 1:
        grid = RectilinearGrid(size=size, x=(0, 1e3), y=(0, 1e3), z=(0, -10.0 * size[2]))
        vel = None
    else:
        grid = RectilinearGrid(size=size, x=(0, size[0] * 1e3), y=(0, size[1] * 1e3), z=(0, -10.0))
        vel = cellular_flow(grid, U=0.1, L=16e3 * size[0] / 128)
    spec = ModelSpec(grid, max_individuals=cap_factor * N, phyto=phyto_setup(entry["params"], [N], 1),
                     seed=entry["seed"], host_state=host_state)
    return spec, vel
