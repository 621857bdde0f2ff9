"""GPU parity on LatLonGrid (Grids/lat_lon_grid.jl): the advection kernel with
the spacings of each corner column, the separate DST3FL sweeps and the z
sweep with per-cell areas and volumes, PAR with cell volumes, against the
oracle.  Bar: bit-identical individuals, tracers and diagnostics."""
import numpy as np
import pytest

from parity import compare_models
from test_oracle_pins_latlon import _grid_info, _streamfunction_flow

pytestmark = pytest.mark.gpu


def _run(pi, orc, spec, uvw, dt, n, bcs=()):
    g = spec.grid
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    PARF = np.full((g.Nx, g.Ny, n + 1), 700.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    for tracer, pos, v in bcs:
        pi.set_bc(gm, tracer=tracer, pos=pos, bc_value=v)
        om.set_tracer_bc(pi._abi.TRACER_ID[tracer], pos, v)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t, dt)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return gm, om


def _w_column(g, seed):
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal((g.Nx, g.Ny, g.Nz + 1)) * 2e-4).astype(np.float32)
    w[:, :, 0] = w[:, :, -1] = 0
    return w


def test_parity_latlon_global(pi, orc):
    """A global LatLonGrid (periodic in longitude, 36x20x4 over ±60°), two
    species, a divergence-free horizontal flow of ~1 m/s plus vertical motion,
    diffusion of individuals and tracers, 10 hourly steps."""
    g = pi.LatLonGrid(size=(36, 20, 4), lat=(-60, 60), lon=(-180, 180), z=[0.0, -5.0, -15.0, -30.0, -60.0])
    spec = pi.ModelSpec(g, max_individuals=6000, seed=71,
                        phyto=pi.phyto_setup({"dvid_P": [1e-4, 2e-4], "mort_P": [5e-5, 5e-5]}, [2500, 1500], 2),
                        bgc_params={"κhP": 5e3, "κvP": 1e-3, "κh": 1e4, "κv": 1e-4})
    u, v, _ = _streamfunction_flow(g, 72)
    scale = np.float32(1.0 / max(np.abs(u).max(), np.abs(v).max()))
    gm, om = _run(pi, orc, spec, (u * scale, v * scale, _w_column(g, 73)), 3600.0, 10)
    bad = compare_models(gm._native, om, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_loaded_latlon_land_and_bcs(pi, orc):
    """A LoadLatLonGrid (MITgcm-style spacings with dxF ≠ dxC, hFac-scaled
    areas and volumes), bounded in x, with land and flux boundary conditions
    on every face."""
    Nx, Ny, Nz = 14, 10, 3
    gi = _grid_info(Nx, Ny, Nz, 74)
    lm = np.ones((Nx, Ny, Nz), dtype=bool)
    lm[0:3, 4:6, 1:3] = False
    lm[9:12, 0:2, 0:2] = False
    g = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-30, 30), lon=(0, 70), landmask=lm)
    assert g.topology[0] is pi.Bounded
    spec = pi.ModelSpec(g, max_individuals=3000, seed=75,
                        phyto=pi.phyto_setup({"dvid_P": [2e-4], "mort_reg": [1.0]}, [1500], 1),
                        bgc_params={"κhP": 50.0, "κvP": 1e-3, "κh": 100.0, "κv": 1e-4})
    rng = np.random.default_rng(76)
    u = (rng.standard_normal((Nx + 1, Ny, Nz)) * 0.05).astype(np.float32)
    v = (rng.standard_normal((Nx, Ny + 1, Nz)) * 0.05).astype(np.float32)
    w = _w_column(g, 77)
    n = 8
    bcs = [("DIC", "west", np.float32(1e-3)), ("NO3", "east", (rng.random((Ny, Nz)) * 1e-3).astype(np.float32)),
           ("PO4", "south", (rng.random((Nx, Nz, n)) * 1e-4).astype(np.float32)),
           ("DOC", "north", np.float32(-5e-4)), ("NH4", "bottom", (rng.random((Nx, Ny)) * 1e-3).astype(np.float32)),
           ("Dust", "top", (rng.random((Nx, Ny, n)) * 1e-5).astype(np.float32))]
    gm, om = _run(pi, orc, spec, (u, v, w), 600.0, n, bcs=bcs)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_latlon_member_in_a_batch_matches_oracle(pi, orc):
    """A LatLonGrid member and a RectilinearGrid member stepped in one batch,
    each identical to its own oracle run."""
    g1 = pi.LatLonGrid(size=(24, 12, 3), lat=(-45, 45), lon=(-180, 180), z=(0, -30))
    s1 = pi.ModelSpec(g1, max_individuals=3000, seed=78, phyto=pi.phyto_setup({"dvid_P": [2e-4]}, [1200], 1),
                      bgc_params={"κhP": 2e3, "κh": 5e3})
    u, v, _ = _streamfunction_flow(g1, 79)
    k = np.float32(0.5 / max(np.abs(u).max(), np.abs(v).max()))
    uvw1 = (u * k, v * k, np.zeros((24, 12, 4), np.float32))
    g2 = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    s2 = pi.ModelSpec(g2, max_individuals=6000, seed=80,
                      phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                      bgc_params={"κhP": 1.0, "κh": 20.0})
    uvw2 = pi.workloads.cellular_flow(g2, U=2.0, L=5e3)
    batch = pi._native.NativeBatch()
    hs = []
    for s, uvw in ((s1, uvw1), (s2, uvw2)):
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        h.set_velocity(*uvw)
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        hs.append(h)
    batch.run(600.0, 8)
    batch.sync()
    for (s, uvw), h in zip(((s1, uvw1), (s2, uvw2)), hs):
        g = s.grid
        om = orc.OracleModel(s)
        om.update(600.0, 8, vel=uvw, parf=np.full((g.Nx, g.Ny), 550.0, np.float32),
                  temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(h, om, s.n_species)
        assert not bad, "\n".join(bad[:20])
    batch.close()
