"""GPU parity of the tracer flux boundary conditions (set_bc!,
HaloRegions/boundary_conditions.jl:12-40; apply_bcs!, HaloRegions/apply_bcs.jl)
against the oracle: every face, every kind of value (number, face array,
face array per iteration), with flow, individuals, diffusion, sinking and
forcing all on; alone, through update!, on land, graph-replayed and in a
batch.  Bar: bit-identical individuals, tracers and diagnostics."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu


def _bcs(g, n, seed):
    """(tracer, face, value) for a run of n steps: every face, several tracers
    (sinking and forced ones among them), all three kinds of value."""
    rng = np.random.default_rng(seed)
    fx, fy, fz = (g.Ny, g.Nz), (g.Nx, g.Nz), (g.Nx, g.Ny)
    r = lambda *s: (rng.random(s) * 2e-3).astype(np.float32)
    return [("DIC", "west", np.float32(3e-3)),
            ("NO3", "east", r(*fx)),
            ("PO4", "south", r(*fy, n)),
            ("DOC", "north", np.float32(-1e-3)),
            ("NH4", "bottom", r(*fz)),
            ("DFe", "top", r(*fz, n) * np.float32(1e-3)),
            ("Dust", "bottom", np.float32(5e-4)),
            ("POC", "top", r(*fz)),
            ("DIC", "east", r(*fx, n)),
            ("DIC", "top", np.float32(2e-3))]


def _apply(pi, gm, om, bcs):
    for tracer, pos, v in bcs:
        if gm is not None:
            pi.set_bc(gm, tracer=tracer, pos=pos, bc_value=v)
        if om is not None:
            om.set_tracer_bc(pi._abi.TRACER_ID[tracer], pos, v)


def _run(pi, orc, spec, uvw, dt, n, seed):
    g = spec.grid
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    PARF = np.full((g.Nx, g.Ny, n + 1), 600.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    bcs = _bcs(g, n, seed)
    _apply(pi, gm, om, bcs)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t, dt)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return gm, om


def test_parity_flux_bcs_3d(pi, orc):
    """Two species under a Fourier flow on a 16x16x8 grid with flux conditions
    on all six faces, 12 steps through update! (the steady steps replay the
    captured graph)."""
    spec, (u, v, w) = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=61, nmodes=12,
                                             overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    gm, om = _run(pi, orc, spec, (u * 20, v * 20, w * 20), 60.0, 12, seed=1)
    bad = compare_models(gm._native, om, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_flux_bcs_bounded_land(pi, orc):
    """A bounded box with land cells next to the faces (the landmask
    instantiation of the z sweep), flux conditions on all faces."""
    lm = np.ones((12, 10, 4), dtype=bool)
    lm[0:2, 3:6, 0:2] = False
    lm[5:8, 8:10, 2:4] = False
    g = pi.RectilinearGrid(size=(12, 10, 4), x=(0, 12e3), y=(0, 10e3), z=[0.0, -2.0, -5.0, -9.0, -14.0],
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded), landmask=lm)
    spec = pi.ModelSpec(g, max_individuals=1200, seed=62,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_reg": [1.0]}, [600], 1),
                        bgc_params={"κhP": 5.0, "κvP": 1e-3, "κh": 20.0, "κv": 1e-3})
    rng = np.random.default_rng(5)
    u = (rng.standard_normal((13, 10, 4)) * 3).astype(np.float32)
    v = (rng.standard_normal((12, 11, 4)) * 3).astype(np.float32)
    w = (rng.standard_normal((12, 10, 5)) * 1e-3).astype(np.float32)
    gm, om = _run(pi, orc, spec, (u, v, w), 60.0, 10, seed=2)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_flux_bc_series_past_its_end_is_refused(pi):
    """A step whose iteration has no slice of a time-dependent condition is
    refused before it runs (the reference would read past the array); the
    model keeps its state and steps again once the condition is replaced."""
    g = pi.RectilinearGrid(size=(8, 8, 2), x=(0, 8e3), y=(0, 8e3), z=(0, -20))
    spec = pi.ModelSpec(g, max_individuals=400, seed=63, phyto=pi.phyto_setup(None, [200], 1))
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    h = gm._native
    h.set_parf(np.full((8, 8), 500.0, np.float32))
    h.set_temperature(np.full((8, 8, 2), 20.0, np.float32))
    pi.set_bc(gm, tracer="NO3", pos="south", bc_value=np.ones((8, 2, 3), np.float32))
    h.run(60.0, 3)
    before = h.tracer(pi._abi.TRACER_ID["NO3"])
    with pytest.raises(ValueError):
        h.step(60.0)
    with pytest.raises(ValueError):
        h.run(60.0, 2)
    assert h.iteration() == 3
    assert np.array_equal(h.tracer(pi._abi.TRACER_ID["NO3"]), before)
    pi.set_bc(gm, tracer="NO3", pos="south", bc_value=None)
    h.step(60.0)
    assert h.iteration() == 4


def test_flux_bcs_in_a_batch_match_oracle(pi, orc):
    """A member with flux conditions beside one without, stepped in one batch,
    each against its own oracle run."""
    specs = []
    spec, uvw = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=64, nmodes=12,
                                       overrides={"dvid_P": [1e-3]}, kappa=(1.0, 1e-4))
    specs.append((spec, tuple(a * np.float32(20) for a in uvw), True))
    g = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    specs.append((pi.ModelSpec(g, max_individuals=6000, seed=65,
                               phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                               bgc_params={"κhP": 1.0, "κh": 20.0}),
                  pi.workloads.cellular_flow(g, U=2.0, L=5e3), False))
    n = 8
    models, oracles = [], []
    batch = pi._native.NativeBatch()
    for s, uvw, with_bc in specs:
        gm = pi.PlanktonModel(pi.HIPArch(), spec=s)
        om = orc.OracleModel(s)
        if with_bc:
            _apply(pi, gm, om, _bcs(s.grid, n, 3))
        h = gm._native
        h.set_velocity(*uvw)
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        models.append(gm)
        oracles.append(om)
    batch.run(60.0, n)
    batch.sync()
    for (s, uvw, _), gm, om in zip(specs, models, oracles):
        g = s.grid
        om.update(60.0, n, vel=uvw, parf=np.full((g.Nx, g.Ny), 550.0, np.float32),
                  temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(gm._native, om, s.n_species)
        assert not bad, "\n".join(bad[:20])
    batch.close()
