"""GPU parity of the phytoplankton-particle interaction and release
(Individuals/Abiotic/particle_interaction.jl:1-138; Model/time_step.jl:119-159)
against the oracle: particles joining individuals of their cell closer than Rd
up to max_ptc (candidates by residue of max_candidates), release with the
capacity check, ptc carried by daughters and left in the slots of the dead,
the ptc and abiotic num diagnostics after both; the hand-built cells of the
oracle's pins, a 3D flow with particle creation at the faces, crowded cells
with shared residues, a LatLonGrid, the capacity error and a member in a
batch.  Bar: every column (ptc included), tracer and diagnostic bit-identical."""
import numpy as np
import pytest

from parity import compare_models, first_diff, same

pytestmark = pytest.mark.gpu

EX = tuple(f"ex{k}" for k in range(8))
ABIO_FIELDS = ("x", "y", "z", "xi", "yi", "zi", "ac")
OVER = {"is_nr": [1.0, 0.0], "is_tric": [0.0, 1.0], "KSAFe": [2.77e-15, 5e-15], "dvid_P": [2e-3, 1e-3],
        "dvid_reg": [1.3, 1.3], "grz_P": [1e-4, 2e-4], "mort_P": [1e-3, 2e-3], "mort_reg": [1.4, 1.4],
        "PCmax": [7.8e-6, 7.8e-6], "max_ptc": [3.0, 2.0]}


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


def _init(pi):
    init = pi.workloads.noisy_tracer_init()
    init["initial_condition"].update(DFe=1e-3, PFe_inorg=1e-4, Dust=1e-3)
    return init


def _run(pi, orc, spec, uvw, dt, n, parf=600.0, pbcs=()):
    g = spec.grid
    PARF = np.full((g.Nx, g.Ny, n + 1), parf, np.float32)
    temp = np.full((g.Nx, g.Ny, g.Nz, 1), 22.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    for sa, pos, v in pbcs:
        pi.set_bc_particle(gm, sa=f"sa{sa + 1}", pos=pos, bc_value=v)
        om.set_particle_bc(sa, pos, v)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt, ΔT_temp=dt * 1e6,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    om.update(dt, n, vel=(u, v, w), parf=PARF[:, :, 0], temp=temp[..., 0])
    return gm, om


def _check(gm, om, nsp, nsa):
    bad = compare_models(gm._native, om, nsp, extra_fields=EX) + _compare_abiotic(gm._native, om, nsa)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("case", range(4))
def test_hand_built_cells(pi, orc, case):
    """The oracle's hand-built cells (residues, replacement within a residue,
    max_ptc, cells, metres) on the device."""
    from test_oracle_pins_interact import IND, PTC, _forcing, _place, _spec, _still
    mc, maxptc, ptc0 = [(25, 5.0, [0, 0, 0, 0]), (25, 1.0, [0, 0, 0, 0]), (2, 5.0, [0, 0, 0, 0]),
                        (2, 1.0, [0, 0, 1, 0])][case]
    g = pi.RectilinearGrid(size=(2, 1, 1), x=(0, 8.0), y=(0, 2.0), z=(0, -1.0))
    spec = _spec(pi, g, [4], [5], 16, intac=[("sp1", "sa1")], mc=mc, over={"max_ptc": [maxptc]},
                 aover={"Rd": [0.9]})
    _place(spec.initial_individuals[0], range(4), IND, ptc0)
    spec.initial_individuals[0]["ex7"][9] = 3.0
    _place(spec.initial_abiotics[0], range(5), PTC)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    for h in (gm._native, om):
        h.set_velocity(*_still(g))
        h.set_parf(parf)
        h.set_temperature(temp)
        h.step(60.0)
    gm._native.sync()
    assert om.abiotic_field(0, "ac")[:5].sum() < 5
    _check(gm, om, 1, 1)


def test_parity_interaction_3d(pi, orc):
    """Two species and two particle species under a Fourier flow with the random
    walk: three interaction pairs, two release pairs, particles created through
    two faces, divisions (daughters take their slot's ptc), grazing and
    mortality; 14 steps through update!."""
    spec0, uvw = pi.workloads.config_3d(n=12, nz=4, N=900, nsp=1, cap=4000, seed=121, nmodes=10, kappa=(1.0, 1e-4))
    g = spec0.grid
    pal = pi.Palat(intac=[("sp1", "sa1"), ("sp2", "sa1"), ("sp2", "sa2")], release=[("sp1", "sa1"), ("sp2", "sa2")])
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=4000, seed=121,
                        phyto=pi.phyto_setup(OVER, [1100, 900], 2), tracer_initial=_init(pi),
                        abiotic=pi.abiotic_setup({"Rd": [350.0, 250.0], "release_P": [2e-3, 5e-3]}, [1500, 800], 2,
                                                 palat=pal),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 1.0, "κv": 1e-4})
    unit = float(spec.abiotic_matrix[0, 3] * spec.abiotic_matrix[0, 0]) / (60.0 * float(g.Δx * g.Δy))
    u, v, w = uvw
    gm, om = _run(pi, orc, spec, (u * 10, v * 10, w * 10), 60.0, 14,
                  pbcs=[(0, "top", np.float32(0.2 * unit)), (1, "bottom", np.float32(0.2 * unit))])
    ptc = om.field(0, "ptc")
    assert ptc.max() == 3 and om.field(1, "ptc").max() == 2 and om.diag(0, "dvid").sum() > 5
    assert om.diag(0, "ptc").sum() > 100
    _check(gm, om, 2, 2)


def test_parity_interaction_crowded(pi, orc):
    """A 3x3x1 grid with about 70 individuals and 100 particles per cell and
    max_candidates = 5: long candidate lists sharing residues, hundreds taken
    per step, releases into the same species."""
    g = pi.RectilinearGrid(size=(3, 3, 1), x=(0, 30.0), y=(0, 30.0), z=(0, -5.0))
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=1500, seed=122, max_candidates=5,
                        phyto=pi.phyto_setup({"dvid_P": [5e-4], "dvid_reg": [1.3], "grz_P": [1e-6], "mort_P": [5e-4],
                                              "mort_reg": [1.4], "max_ptc": [4.0]}, [640], 1),
                        tracer_initial=_init(pi),
                        abiotic=pi.abiotic_setup({"Rd": [3.0], "release_P": [1e-3]}, [900], 1,
                                                 palat=pi.Palat(intac=[("sp1", "sa1")], release=[("sp1", "sa1")])),
                        bgc_params={"κhP": 0.05, "κvP": 1e-4, "κh": 0.1, "κv": 1e-4})
    rng = np.random.default_rng(123)
    uvw = ((rng.standard_normal((3, 3, 1)) * 1e-3).astype(np.float32),
           (rng.standard_normal((3, 3, 1)) * 1e-3).astype(np.float32), np.zeros((3, 3, 2), np.float32))
    gm, om = _run(pi, orc, spec, uvw, 60.0, 8)
    assert om.diag(0, "ptc").sum() > 500
    _check(gm, om, 1, 1)


def test_parity_interaction_latlon(pi, orc):
    """On a LoadLatLonGrid with land: distances in metres from the per-column
    spacings of the metric arrays."""
    from test_oracle_pins_latlon import _grid_info
    Nx, Ny, Nz = 10, 8, 2
    lm = np.ones((Nx, Ny, Nz), dtype=bool)
    lm[3:5, 2:4, :] = False
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 124), size=(Nx, Ny, Nz), lat=(-30, 30),
                          lon=(-180, 180), landmask=lm)
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=3000, seed=125,
                        phyto=pi.phyto_setup(dict((k, v[:1]) for k, v in OVER.items()), [900], 1),
                        tracer_initial=_init(pi),
                        abiotic=pi.abiotic_setup({"Rd": [1.2e6], "release_P": [1e-3], "Nsuper": [1e3]}, [1000], 1,
                                                 palat=pi.Palat(intac=[("sp1", "sa1")], release=[("sp1", "sa1")])),
                        bgc_params={"κhP": 50.0, "κvP": 1e-3, "κh": 100.0, "κv": 1e-4})
    rng = np.random.default_rng(126)
    uvw = ((rng.standard_normal((Nx, Ny, Nz)) * 0.05).astype(np.float32),
           (rng.standard_normal((Nx, Ny + 1, Nz)) * 0.05).astype(np.float32), np.zeros((Nx, Ny, Nz + 1), np.float32))
    gm, om = _run(pi, orc, spec, uvw, 600.0, 6)
    assert om.diag(0, "ptc").sum() > 20
    _check(gm, om, 1, 1)


def test_release_capacity_error(pi, orc):
    """More releases than free particle slots: both raise the capacity error
    at the same step with the reference's message."""
    from test_oracle_pins_interact import _forcing, _release_case, _still
    g, dt, spec, _ = _release_case(pi, orc, dvid=False)
    spec.initial_individuals[0]["ex7"][:32] = 2.0
    ab = spec.initial_abiotics[0]
    ab["ac"][:] = True
    ab["ac"][[7, 40]] = False
    ab["x"][:], ab["y"][:], ab["z"][:] = 63.9, 0.9, 0.9
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    msgs = []
    for h in (gm._native, om):
        h.set_velocity(*_still(g))
        h.set_parf(parf)
        h.set_temperature(temp)
        with pytest.raises(pi.CapacityError) as e:
            h.step(dt)
            if h is gm._native:
                h.sync()
        msgs.append((str(e.value), h.iteration()))
    # the device's message names the species after the reference's text
    assert msgs[0][1] == msgs[1][1] == 1 and msgs[0][0].startswith(msgs[1][0])
    assert msgs[1][0] == "number of abiotic particles exceeds the capacity at timestep 0.000694444 days"


def test_release_case_parity(pi, orc):
    """The oracle's release case (two species, pre-set ptc, particles already
    in slots 0, 2, 5) for three steps."""
    from test_oracle_pins_interact import _forcing, _release_case, _still
    g, dt, spec, _ = _release_case(pi, orc)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    for h in (gm._native, om):
        h.set_velocity(*_still(g))
        h.set_temperature(temp)
        for _ in range(3):          # update! copies the PARF slice in before every step
            h.set_parf(parf)
            h.step(dt)
    gm._native.sync()
    assert om.abiotic_field(0, "ac").sum() > 10
    _check(gm, om, 2, 1)


def test_interaction_member_in_a_batch(pi, orc):
    """An IronEnergyMode model with interaction and release beside an
    IronEnergyMode model without particles, in one batch."""
    g1 = pi.RectilinearGrid(size=(10, 8, 2), x=(0, 10e3), y=(0, 8e3), z=(0, -20))
    s1 = pi.ModelSpec(g1, mode=pi.IronEnergyMode(), max_individuals=3000, seed=127,
                      phyto=pi.phyto_setup(OVER, [900, 700], 2), tracer_initial=_init(pi),
                      abiotic=pi.abiotic_setup({"Rd": [400.0], "release_P": [2e-3]}, [1200], 1,
                                               palat=pi.Palat(intac=[("sp1", "sa1"), ("sp2", "sa1")],
                                                              release=[("sp2", "sa1")])),
                      bgc_params={"κhP": 2.0, "κh": 10.0})
    s2 = pi.ModelSpec(g1, mode=pi.IronEnergyMode(), max_individuals=3000, seed=128,
                      phyto=pi.phyto_setup(OVER, [800, 600], 2), tracer_initial=_init(pi),
                      bgc_params={"κhP": 2.0, "κh": 10.0})
    batch = pi._native.NativeBatch()
    hs = []
    for s in (s1, s2):
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        h.set_velocity(*pi.workloads.cellular_flow(s.grid, U=1.0, L=4e3))
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        hs.append(h)
    batch.run(60.0, 8)
    batch.sync()
    for s, h in ((s1, hs[0]), (s2, hs[1])):
        g = s.grid
        om = orc.OracleModel(s)
        om.update(60.0, 8, vel=pi.workloads.cellular_flow(g, U=1.0, L=4e3),
                  parf=np.full((g.Nx, g.Ny), 550.0, np.float32), temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(h, om, s.n_species, extra_fields=EX) + _compare_abiotic(h, om, s.n_abiotic)
        assert not bad, "\n".join(bad[:20])
    batch.close()


def test_interaction_budget_at_size(pi):
    """1,048,576 individuals and 524,288 particles on a 128x128x32 grid with
    divisions and deaths, 20 steps: ptc summed over every slot (dead slots keep
    theirs, a daughter takes its slot's) plus the active particles stays the
    initial particle count; ptc integral in [0, max_ptc]; the ptc diagnostic is
    the living individuals' ptc at the last step."""
    g = pi.RectilinearGrid(size=(128, 128, 32), x=(0, 128e3), y=(0, 128e3), z=(0, -320))
    spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=2 ** 21, seed=131,
                        phyto=pi.phyto_setup({"max_ptc": [3.0], "dvid_P": [2e-3], "dvid_reg": [1.3], "mort_P": [1e-3],
                                              "mort_reg": [1.4], "grz_P": [1e-6]}, [2 ** 20], 1),
                        abiotic=pi.abiotic_setup({"Rd": [300.0], "release_P": [1e-4]}, [2 ** 19], 1,
                                                 palat=pi.Palat(intac=[("sp1", "sa1")], release=[("sp1", "sa1")])),
                        bgc_params={"κhP": 1.0, "κh": 1.0})
    m = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    m.set_velocity(*pi.workloads.cellular_flow(g, U=0.5, L=32e3))
    m.set_parf(np.full((128, 128), 500.0, np.float32))
    m.set_temperature(np.full((128, 128, 32), 20.0, np.float32))
    m.run(60.0, 19)
    m.reset_diags()
    m.run(60.0, 1)
    m.sync()
    ptc = m.field(0, "ptc").astype(np.float64)
    ac = m.field(0, "ac")
    npart = int(m.abiotic_field(0, "ac").sum())
    assert np.all(ptc == np.round(ptc)) and ptc.min() >= 0 and ptc.max() <= 3
    assert ptc.sum() + npart == 2 ** 19 and ptc.sum() > 1000
    assert m.diag(0, "ptc").sum() == ptc[ac].sum()
    assert m.diag(0, "dvid").sum() > 0 and m.diag(0, "mort").sum() > 0
