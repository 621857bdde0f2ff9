"""GPU parity of IronEnergyMode (Individuals/Plankton/IronEnergyMode) against the
oracle: the energy budget (light reaction, respiration, carbon fixation,
nitrate reduction, N fixation), the four iron pools and their light- and
clock-driven exchanges, losses into ten tracers (NH4, NO3, DFe and PFe_bio
besides the organic pools), DFe uptake, the slots of dead individuals that keep
moving their iron, the device-drawn initial state, crowded cells (the
long-segment exchange kernel) and a member in a batch.  Bar: every column
(the eight extra ones included), tracer and diagnostic bit-identical."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu

EX = tuple(f"ex{k}" for k in range(8))
OVER = {"is_nr": [1.0, 0.0], "is_tric": [0.0, 1.0], "KSAFe": [2.77e-15, 5e-15], "dvid_P": [2e-3, 1e-3],
        "dvid_reg": [1.3, 1.3], "grz_P": [1e-4, 2e-4], "mort_P": [1e-3, 2e-3], "mort_reg": [1.4, 1.4],
        "PCmax": [7.8e-6, 7.8e-6]}


def _init(pi):
    init = pi.workloads.noisy_tracer_init()
    init["initial_condition"].update(DFe=1e-3, PFe_inorg=1e-4, Dust=1e-3)
    return init


def _run(pi, orc, spec, uvw, dt, parf_steps, temp_c=24.0):
    g = spec.grid
    n = len(parf_steps)
    PARF = np.stack([np.full((g.Nx, g.Ny), v, np.float32) for v in parf_steps] + [np.zeros((g.Nx, g.Ny), np.float32)],
                    axis=2)
    temp = np.full((g.Nx, g.Ny, g.Nz, 1), temp_c, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt, ΔT_temp=dt * 1e6,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    idx = lambda t: int(np.floor(np.float32(t) / np.float32(dt)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t)], temp=temp[..., 0])
    return gm, om


def test_parity_iron_3d_diel(pi, orc):
    """A non-diazotroph and a Trichodesmium-like species under a Fourier flow with
    the random walk, through night, sunrise, day and sunset (par = 0, rising
    and falling light), 16 steps: divisions, grazing and deaths happen."""
    spec0, uvw = pi.workloads.config_3d(n=16, nz=6, N=1200, nsp=1, cap=4000, seed=101, nmodes=12, kappa=(1.0, 1e-4))
    spec = pi.ModelSpec(spec0.grid, mode=pi.IronEnergyMode(), max_individuals=4000, seed=101,
                        phyto=pi.phyto_setup(OVER, [1500, 1200], 2), tracer_initial=_init(pi),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 1.0, "κv": 1e-4})
    u, v, w = uvw
    parf = [0, 0, 100, 300, 600, 800, 800, 600, 300, 100, 0, 0, 0, 200, 400, 400]
    gm, om = _run(pi, orc, spec, (u * 10, v * 10, w * 10), 60.0, parf)
    for s in range(2):
        assert om.diag(s, "dvid").sum() > 5 and om.diag(s, "mort").sum() > 5 and om.diag(s, "graz").sum() > 5
    bad = compare_models(gm._native, om, 2, extra_fields=EX)
    assert not bad, "\n".join(bad[:20])


def test_parity_iron_land_crocosphaera_long_night(pi, orc):
    """A bounded box with land, a Crocosphaera-like species (iron to N fixation
    early in the night, back after NF_clock) beside a non-diazotroph, hourly
    steps through a night longer than NF_clock, Timer division, shared_graz = 0."""
    lm = np.ones((8, 7, 3), dtype=bool)
    lm[2:4, 2:5, 1:3] = False
    g = pi.RectilinearGrid(size=(8, 7, 3), x=(0, 8e3), y=(0, 7e3), z=(0, -30),
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded), landmask=lm)
    over = dict(OVER, is_nr=[0.0, 1.0], is_tric=[0.0, 0.0], is_croc=[1.0, 0.0], dvid_type=[2, 1],
                dvid_reg=[3.0, 1.3], dvid_P=[2e-5, 5e-5], grz_P=[1e-6, 1e-6], mort_P=[2e-5, 2e-5])
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=3000, seed=102,
                        phyto=pi.phyto_setup(over, [800, 700], 2), tracer_initial=_init(pi),
                        bgc_params={"κhP": 5.0, "κh": 20.0, "shared_graz": 0.0})
    rng = np.random.default_rng(103)
    uvw = ((rng.standard_normal((9, 7, 3)) * 0.1).astype(np.float32),
           (rng.standard_normal((8, 8, 3)) * 0.1).astype(np.float32),
           (rng.standard_normal((8, 7, 4)) * 1e-5).astype(np.float32))
    parf = [300, 0, 0, 0, 0, 0, 0, 0, 0, 0, 200, 500]
    gm, om = _run(pi, orc, spec, uvw, 3600.0, parf)
    assert (om.field(0, "qFeNF") > 0).any() and om.field(0, "tdark").max() > 21600.0
    bad = compare_models(gm._native, om, 2, extra_fields=EX)
    assert not bad, "\n".join(bad[:20])


def test_parity_iron_crowded_cells(pi, orc):
    """Hundreds of individuals per cell: the long-segment exchange kernel with
    ten loss values and DFe consumption."""
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 4e3), y=(0, 4e3), z=(0, -20))
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=12000, seed=104,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "dvid_reg": [1.3], "grz_P": [1e-6],
                                              "mort_P": [2e-3], "mort_reg": [1.4], "KSAFe": [2.77e-15]}, [8000], 1),
                        tracer_initial=_init(pi), bgc_params={"κhP": 10.0, "κvP": 1e-3})
    uvw = pi.workloads.cellular_flow(g, U=1.0, L=2e3)
    gm, om = _run(pi, orc, spec, uvw, 60.0, [500, 500, 0, 0, 300, 300])
    assert om.diag(0, "mort").sum() > 50
    bad = compare_models(gm._native, om, 1, extra_fields=EX)
    assert not bad, "\n".join(bad[:20])


def test_iron_device_generated_state_matches_host(pi):
    """pi_generate_state for IronEnergyMode equals the host generator's arrays
    uploaded, before and after three steps."""
    from test_gpu_batch import _inputs
    g = pi.RectilinearGrid(size=(12, 10, 4), x=(0, 12e3), y=(0, 10e3), z=(0, -40))
    kw = dict(mode=pi.IronEnergyMode(), max_individuals=3000, seed=105, phyto=pi.phyto_setup(OVER, [1800, 900], 2))
    A = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, **kw))._native
    B = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, host_state=False, **kw))._native
    bad = compare_models(A, B, 2, diags=False, extra_fields=EX)
    assert not bad, "\n".join(bad[:20])
    for h in (A, B):
        _inputs(pi, h, pi.ModelSpec(g, **dict(kw, phyto=pi.phyto_setup(OVER, [1, 1], 2))), None)
        h.run(60.0, 3)
        h.sync()
    bad = compare_models(A, B, 2, extra_fields=EX)
    assert not bad, "\n".join(bad[:20])


def test_iron_member_in_a_batch(pi, orc):
    """An IronEnergyMode model, a MacroMolecularMode model and a QuotaMode model
    stepped in one batch, each identical to its own oracle run."""
    g1 = pi.RectilinearGrid(size=(12, 10, 3), x=(0, 12e3), y=(0, 10e3), z=(0, -30))
    s1 = pi.ModelSpec(g1, mode=pi.IronEnergyMode(), max_individuals=3000, seed=106,
                      phyto=pi.phyto_setup(OVER, [900, 700], 2), tracer_initial=_init(pi),
                      bgc_params={"κhP": 2.0, "κh": 10.0})
    s2 = pi.ModelSpec(g1, mode=pi.MacroMolecularMode(), max_individuals=3000, seed=107,
                      phyto=pi.phyto_setup({"k_dna": [2e-4], "dvid_P": [2e-3]}, [900], 1),
                      bgc_params={"κhP": 2.0, "κh": 10.0})
    g3 = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    s3 = pi.ModelSpec(g3, max_individuals=6000, seed=108,
                      phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                      bgc_params={"κhP": 1.0, "κh": 20.0})
    batch = pi._native.NativeBatch()
    hs = []
    for s in (s1, s2, s3):
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        h.set_velocity(*pi.workloads.cellular_flow(s.grid, U=1.0, L=4e3))
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        hs.append(h)
    batch.run(60.0, 10)
    batch.sync()
    for s, h, ex in ((s1, hs[0], EX), (s2, hs[1], ("ex0",)), (s3, hs[2], ())):
        g = s.grid
        om = orc.OracleModel(s)
        om.update(60.0, 10, vel=pi.workloads.cellular_flow(g, U=1.0, L=4e3),
                  parf=np.full((g.Nx, g.Ny), 550.0, np.float32), temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(h, om, s.n_species, extra_fields=ex)
        assert not bad, "\n".join(bad[:20])
    batch.close()
