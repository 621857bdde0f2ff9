"""GPU parity of abiotic particles (Individuals/Abiotic; Model/time_step.jl:21-38,
152-159) against the oracle: creation at the faces by particle flux boundary
conditions, advection, the random walk, find_inds! and the num diagnostic,
beside phytoplankton; alone, on a LatLonGrid with land, with the capacity
error, and in a batch.  Bar: every column bit-identical."""
import numpy as np
import pytest

from parity import compare_models, first_diff, same

pytestmark = pytest.mark.gpu

ABIO_FIELDS = ("x", "y", "z", "xi", "yi", "zi", "ac")


def _compare_abiotic(gpu, ref, nsa):
    bad = []
    for a in range(nsa):
        for f in ABIO_FIELDS:
            x, y = gpu.abiotic_field(a, f), ref.abiotic_field(a, f)
            if not same(x, y):
                bad.append(f"sa{a + 1}.{f}: {first_diff(x, y)}")
        x, y = gpu.abiotic_diag(a), ref.abiotic_diag(a)
        if not same(x, y):
            bad.append(f"sa{a + 1}.num: {first_diff(x, y)}")
    return bad


def _pbcs(spec, dt, n, seed):
    """Fluxes through four faces of a RectilinearGrid that create a particle at
    about a third of the cells next to them per step (a number, a face array,
    a face array per iteration)."""
    g, rng = spec.grid, np.random.default_rng(seed)
    unit = lambda sa, A: spec.abiotic_matrix[sa, 3] * spec.abiotic_matrix[sa, 0] / (dt * A)
    Az, Ax, Ay = float(g.Δx * g.Δy), float(g.Δy * g.dzF[g.Hz]), float(g.Δx * g.dzF[g.Hz])
    return [(0, "top", np.float32(0.3 * unit(0, Az))),
            (0, "west", (rng.random((g.Ny, g.Nz)) * 0.6 * unit(0, Ax)).astype(np.float32)),
            (1, "bottom", (rng.random((g.Nx, g.Ny, n)) * 0.6 * unit(1, Az)).astype(np.float32)),
            (1, "north", np.float32(0.3 * unit(1, Ay)))]


def _run(pi, orc, spec, uvw, dt, n, pbcs):
    g = spec.grid
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    PARF = np.full((g.Nx, g.Ny, n + 1), 650.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    for sa, pos, v in pbcs:
        pi.set_bc_particle(gm, sa=f"sa{sa + 1}", pos=pos, bc_value=v)
        om.set_particle_bc(sa, pos, v)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t, dt)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return gm, om


def test_parity_abiotic_3d(pi, orc):
    """Two phytoplankton species and two abiotic species (1500 and 600 initial
    particles) under a Fourier flow; particle fluxes through four faces, one
    per iteration; 12 steps through update!."""
    spec0, uvw = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=5000, seed=81, nmodes=12,
                                        overrides={"dvid_P": [1e-3]}, kappa=(1.0, 1e-4))
    spec = pi.ModelSpec(spec0.grid, max_individuals=5000, seed=81,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3, 1e-3]}, [1500, 1500], 2),
                        abiotic=pi.abiotic_setup({"sz_min": [1e-3, 2e-3]}, [1500, 600], 2),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 1.0, "κv": 1e-4})
    n = 12
    u, v, w = uvw
    gm, om = _run(pi, orc, spec, (u * 20, v * 20, w * 20), 60.0, n, _pbcs(spec, 60.0, n, 1))
    assert om.abiotic_field(0, "ac").sum() > 1500 and om.abiotic_field(1, "ac").sum() > 600
    bad = compare_models(gm._native, om, 2) + _compare_abiotic(gm._native, om, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_abiotic_latlon_land(pi, orc):
    """Abiotic particles on a LoadLatLonGrid with land (the advection kernel with
    per-column spacings, face areas from the metric arrays)."""
    from test_oracle_pins_latlon import _grid_info
    Nx, Ny, Nz = 12, 9, 3
    lm = np.ones((Nx, Ny, Nz), dtype=bool)
    lm[4:7, 3:5, :] = False
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 82), size=(Nx, Ny, Nz), lat=(-30, 30),
                          lon=(-180, 180), landmask=lm)
    spec = pi.ModelSpec(g, max_individuals=3000, seed=83, phyto=pi.phyto_setup(None, [800], 1),
                        abiotic=pi.abiotic_setup({"Nsuper": [1e3]}, [700], 1),
                        bgc_params={"κhP": 50.0, "κvP": 1e-3, "κh": 100.0, "κv": 1e-4})
    rng = np.random.default_rng(84)
    u = (rng.standard_normal((Nx, Ny, Nz)) * 0.05).astype(np.float32)
    v = (rng.standard_normal((Nx, Ny + 1, Nz)) * 0.05).astype(np.float32)
    w = np.zeros((Nx, Ny, Nz + 1), np.float32)
    pbcs = [(0, "top", np.float32(1e-6)), (0, "south", np.float32(2e-6))]
    gm, om = _run(pi, orc, spec, (u, v, w), 600.0, 8, pbcs)
    bad = compare_models(gm._native, om, 1) + _compare_abiotic(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_abiotic_capacity_error(pi, orc):
    """More particles to create than free slots: both raise the capacity error
    at the same step, with the iteration already counted (TimeStep! increments
    it before particles_from_bcs! throws, time_step.jl:14); the GPU model keeps
    the particles of the step before."""
    g = pi.RectilinearGrid(size=(6, 5, 2), x=(0, 6e3), y=(0, 5e3), z=(0, -20))
    spec = pi.ModelSpec(g, max_individuals=60, seed=85, phyto=pi.phyto_setup(None, [20], 1),
                        abiotic=pi.abiotic_setup(None, [10], 1))
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    for h in (gm._native, om):
        h.set_particle_bc(0, "top", np.float32(1e6))          # 30 particles per step
        h.set_parf(np.full((6, 5), 500.0, np.float32))
        h.set_temperature(np.full((6, 5, 2), 20.0, np.float32))
    errs = []
    for h in (gm._native, om):
        k = 0
        try:
            for k in range(1, 5):
                h.step(60.0)
                if h is gm._native:
                    h.sync()
        except pi.CapacityError:
            errs.append((k, h.iteration()))
    assert errs[0] == errs[1] == (2, 2)
    assert same(gm._native.abiotic_field(0, "x"), om.abiotic_field(0, "x"))


def test_abiotic_member_in_a_batch(pi, orc):
    """A model with abiotic particles and one without, stepped in one batch,
    each identical to its own oracle run."""
    g1 = pi.RectilinearGrid(size=(12, 10, 3), x=(0, 12e3), y=(0, 10e3), z=(0, -30))
    s1 = pi.ModelSpec(g1, max_individuals=3000, seed=86, phyto=pi.phyto_setup({"dvid_P": [1e-3]}, [900], 1),
                      abiotic=pi.abiotic_setup(None, [400], 1), bgc_params={"κhP": 2.0, "κh": 10.0})
    uvw1 = pi.workloads.cellular_flow(g1, U=1.0, L=4e3)
    g2 = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    s2 = pi.ModelSpec(g2, max_individuals=6000, seed=87,
                      phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                      bgc_params={"κhP": 1.0, "κh": 20.0})
    uvw2 = pi.workloads.cellular_flow(g2, U=2.0, L=5e3)
    batch = pi._native.NativeBatch()
    hs, oms = [], []
    for s, uvw in ((s1, uvw1), (s2, uvw2)):
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        om = orc.OracleModel(s)
        if s is s1:
            for x in (h, om):
                x.set_particle_bc(0, "east", np.float32(3e-9))
        h.set_velocity(*uvw)
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        hs.append(h)
        oms.append(om)
    batch.run(60.0, 8)
    batch.sync()
    for (s, uvw), h, om in zip(((s1, uvw1), (s2, uvw2)), hs, oms):
        g = s.grid
        om.update(60.0, 8, vel=uvw, parf=np.full((g.Nx, g.Ny), 550.0, np.float32),
                  temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(h, om, s.n_species) + _compare_abiotic(h, om, s.n_abiotic)
        assert not bad, "\n".join(bad[:20])
    batch.close()
