"""GPU parity of MacroMolecularMode (Individuals/Plankton/MacroMolecularMode) against
the oracle: protein/DNA/RNA/carbohydrate/N and P stores, division by DNA
content, mortality by age, the C/N/P of every pool in the losses, the slots of
dead individuals that keep synthesising, the device-drawn initial state, and a
MacroMolecularMode member in a batch.  Bar: every column (RNA included),
tracer and diagnostic bit-identical."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu

OVER = {"k_dna": [2e-4, 1.5e-4], "dvid_P": [2e-3, 3e-3], "grz_P": [1e-4, 2e-4], "mort_P": [2e-3, 1e-3],
        "mort_reg": [0.1, 0.15], "CHmax": [0.3, 0.4], "VDOCmax": [1e-6, 0.0], "KsatDOC": [0.2, 0.0]}


def _run(pi, orc, spec, uvw, dt, n):
    g = spec.grid
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    PARF = np.full((g.Nx, g.Ny, n + 1), 700.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    u, v, w = uvw
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t, dt)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return gm, om


def test_parity_mm_3d(pi, orc):
    """Two MacroMolecularMode species under a Fourier flow with the random walk,
    16 steps: divisions, grazing and age-driven deaths happen, daughters land in
    slots freed earlier and in this step, and the dead slots keep synthesising."""
    spec0, uvw = pi.workloads.config_3d(n=16, nz=6, N=1200, nsp=1, cap=4000, seed=91, nmodes=12, kappa=(1.0, 1e-4))
    spec = pi.ModelSpec(spec0.grid, mode=pi.MacroMolecularMode(), max_individuals=4000, seed=91,
                        phyto=pi.phyto_setup(OVER, [1500, 1200], 2),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 1.0, "κv": 1e-4})
    u, v, w = uvw
    gm, om = _run(pi, orc, spec, (u * 10, v * 10, w * 10), 60.0, 16)
    for s in range(2):
        assert om.diag(s, "dvid").sum() > 10 and om.diag(s, "mort").sum() > 10 and om.diag(s, "graz").sum() > 10
    bad = compare_models(gm._native, om, 2, extra_fields=("ex0",))
    assert not bad, "\n".join(bad[:20])


def test_parity_mm_land_species_grazing(pi, orc):
    """A bounded box with land and shared_graz = 0 (grazing by each species' own
    count), 12 steps."""
    lm = np.ones((10, 8, 3), dtype=bool)
    lm[3:6, 2:5, 1:3] = False
    g = pi.RectilinearGrid(size=(10, 8, 3), x=(0, 10e3), y=(0, 8e3), z=(0, -30),
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded), landmask=lm)
    spec = pi.ModelSpec(g, mode=pi.MacroMolecularMode(), max_individuals=3000, seed=92,
                        phyto=pi.phyto_setup(OVER, [1000, 800], 2),
                        bgc_params={"κhP": 5.0, "κh": 20.0, "shared_graz": 0.0})
    rng = np.random.default_rng(96)
    uvw = ((rng.standard_normal((11, 8, 3)) * 0.5).astype(np.float32),
           (rng.standard_normal((10, 9, 3)) * 0.5).astype(np.float32),
           (rng.standard_normal((10, 8, 4)) * 1e-4).astype(np.float32))
    gm, om = _run(pi, orc, spec, uvw, 60.0, 12)
    bad = compare_models(gm._native, om, 2, extra_fields=("ex0",))
    assert not bad, "\n".join(bad[:20])


def test_mm_device_generated_state_matches_host(pi):
    """pi_generate_state for MacroMolecularMode (DNA, RNA, protein, CH, NST, PST
    drawn on the device) equals the host generator's arrays uploaded, before
    and after three steps."""
    from test_gpu_batch import _inputs
    g = pi.RectilinearGrid(size=(12, 10, 4), x=(0, 12e3), y=(0, 10e3), z=(0, -40))
    kw = dict(mode=pi.MacroMolecularMode(), max_individuals=3000, seed=93,
              phyto=pi.phyto_setup(OVER, [1800, 900], 2))
    A = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, **kw))._native
    B = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, host_state=False, **kw))._native
    bad = compare_models(A, B, 2, diags=False, extra_fields=("ex0",))
    assert not bad, "\n".join(bad[:20])
    for h in (A, B):
        _inputs(pi, h, pi.ModelSpec(g, **dict(kw, phyto=pi.phyto_setup(OVER, [1, 1], 2))), None)
        h.run(60.0, 3)
        h.sync()
    bad = compare_models(A, B, 2, extra_fields=("ex0",))
    assert not bad, "\n".join(bad[:20])


def test_mm_member_in_a_batch(pi, orc):
    """A MacroMolecularMode model and a QuotaMode model stepped in one batch,
    each identical to its own oracle run."""
    g1 = pi.RectilinearGrid(size=(12, 10, 3), x=(0, 12e3), y=(0, 10e3), z=(0, -30))
    s1 = pi.ModelSpec(g1, mode=pi.MacroMolecularMode(), max_individuals=3000, seed=94,
                      phyto=pi.phyto_setup(OVER, [900, 700], 2), bgc_params={"κhP": 2.0, "κh": 10.0})
    g2 = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    s2 = pi.ModelSpec(g2, max_individuals=6000, seed=95,
                      phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                      bgc_params={"κhP": 1.0, "κh": 20.0})
    batch = pi._native.NativeBatch()
    hs = []
    for s in (s1, s2):
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        h.set_velocity(*pi.workloads.cellular_flow(s.grid, U=1.0, L=4e3))
        h.set_parf(np.full((s.grid.Nx, s.grid.Ny), 550.0, np.float32))
        h.set_temperature(np.full((s.grid.Nx, s.grid.Ny, s.grid.Nz), 21.0, np.float32))
        batch.add(h)
        hs.append(h)
    batch.run(60.0, 10)
    batch.sync()
    for s, h, ex in ((s1, hs[0], ("ex0",)), (s2, hs[1], ())):
        g = s.grid
        om = orc.OracleModel(s)
        om.update(60.0, 10, vel=pi.workloads.cellular_flow(g, U=1.0, L=4e3),
                  parf=np.full((g.Nx, g.Ny), 550.0, np.float32), temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        bad = compare_models(h, om, s.n_species, extra_fields=ex)
        assert not bad, "\n".join(bad[:20])
    batch.close()


def test_mm_log_lines_match_oracle(pi, orc, tmp_path):
    """write_species_dynamics for MacroMolecularMode (write_outputs.jl:77-96)
    through update! with write_log: every line equals the one the oracle's
    columns give at the same step."""
    from importlib import import_module
    out = import_module("pi_amd.output")
    g = pi.RectilinearGrid(size=(8, 8, 2), x=(0, 8e3), y=(0, 8e3), z=(0, -20))
    spec = pi.ModelSpec(g, mode=pi.MacroMolecularMode(), max_individuals=2000, seed=97,
                        phyto=pi.phyto_setup(OVER, [600, 500], 2))
    dt, n = 60.0, 5
    d = str(tmp_path / "log")
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    PARF = np.full((8, 8, n + 1), 600.0, np.float32)
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    uvw = pi.workloads.cellular_flow(g, U=0.5, L=4e3)
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt,
                                vels=dict(u=uvw[0][..., None], v=uvw[1][..., None], w=uvw[2][..., None]),
                                ΔT_vel=dt * 1e6, output_writer=pi.PlanktonOutputWriter(dir=d, write_log=True))
    pi.update(sim)
    om = orc.OracleModel(spec)
    om.set_velocity(*uvw)
    want = {0: [], 1: []}
    for it in range(1, n + 1):
        om.set_parf(PARF[:, :, it - 1])
        om.set_temperature(temp[:, :, :, 0])
        om.step(np.float32(dt))
        t = np.float32(np.float32(it) * np.float32(dt))
        for s in range(2):
            cols = {f: om.field(s, f) for f in out.MM_DYNAMICS_FIELDS + ("ac",)}
            want[s].append(out.dynamics_line(t, cols, 2))
    for s in range(2):
        lines = open(f"{d}/dynamic_species00{s + 1}.txt").read().splitlines()
        assert lines == want[s], s
        assert len(lines[0].split()) == 12
