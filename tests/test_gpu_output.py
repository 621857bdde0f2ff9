"""update! with an output writer on the GPU (write_outputs.jl): the log lines,
the diagnostics (interior ./ iteration_interval, then zeroed) and the saved
individual columns equal what the oracle's state gives at the same steps."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_update_with_output_writer_matches_oracle(pi, orc, tmp_path):
    from importlib import import_module
    out = import_module("pi_amd.output")
    try:
        out.hdf5()
    except RuntimeError as e:
        pytest.skip(str(e))
    g = pi.RectilinearGrid(size=(16, 16, 4), x=(0, 16e3), y=(0, 16e3), z=(0, -40))
    spec = pi.ModelSpec(g, max_individuals=4000, seed=29,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3, 1e-3], "mort_reg": [1.0, 1.0], "grz_P": [2e-4, 1e-4],
                                              "dvid_reg": [1.5, 1.6]}, [1500, 1200], 2),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    u, v, w = pi.workloads.cellular_flow(g, U=2.0, L=4e3)
    dt, n, parf_dt = 60.0, 8, 600.0
    rng = np.random.default_rng(2)
    PARF = (400.0 + 300.0 * rng.random((g.Nx, g.Ny, 3))).astype(np.float32)
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    d = str(tmp_path / "results")
    writer = pi.PlanktonOutputWriter(dir=d, write_log=True, save_diags=True, save_phytoplankton=True,
                                     phytoplankton_include=("x", "y", "Sz", "Chl", "ac", "xi"),
                                     phytoplankton_iteration_interval=2)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    diags = pi.PlanktonDiagnostics(gm, tracer=("NO3", "PAR"), iteration_interval=3)
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=parf_dt, diags=diags,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6,
                                output_writer=writer)
    pi.update(sim)

    om = orc.OracleModel(spec, diag_tracers=[pi._abi.TRACER_ID["NO3"]])
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    H = (g.Hx, g.Hy, g.Hz)
    inner = lambda a: a[H[0]:H[0] + g.Nx, H[1]:H[1] + g.Ny, H[2]:H[2] + g.Nz]
    logs = {s: [] for s in range(2)}
    t = np.float32(0.0)
    om.set_velocity(u, v, w)
    for it in range(1, n + 1):
        om.set_parf(PARF[:, :, idx(t, parf_dt)])
        om.set_temperature(temp[:, :, :, idx(t, 3600.0)])
        om.step(np.float32(dt))
        t = np.float32(np.float32(it) * np.float32(dt))
        for s in range(2):
            cols = {f: om.field(s, f) for f in out.DYNAMICS_FIELDS + ("ac",)}
            logs[s].append(out.dynamics_line(t, cols))
        if it % 3 == 0:
            f = os.path.join(d, "diags.jld2")
            assert out.read_dataset(f, f"timeseries/t/{it}") == t
            three = np.float32(3.0)
            assert np.array_equal(out.read_dataset(f, f"timeseries/NO3/{it}"),
                                  inner(om.diag_tracer(pi._abi.TRACER_ID["NO3"])) / three)
            assert np.array_equal(out.read_dataset(f, f"timeseries/PAR/{it}"), inner(om.grid_field("diag_PAR")) / three)
            for s in range(2):
                for proc in ("num", "graz", "mort", "dvid"):
                    got = out.read_dataset(f, f"timeseries/phyto/sp{s + 1}/{proc}/{it}")
                    assert np.array_equal(got, inner(om.diag(s, proc)) / three), (it, s, proc)
                assert not np.any(out.read_dataset(f, f"timeseries/phyto/sp{s + 1}/ptc/{it}"))
            om.reset_diags()
        if it % 2 == 0:
            f = os.path.join(d, "phytoplankton.jld2")
            for s in range(2):
                for att in ("x", "y", "Sz", "Chl", "ac", "xi"):
                    got = out.read_dataset(f, f"timeseries/sp{s + 1}/{att}/{it}")
                    assert np.array_equal(got, om.field(s, att)), (it, s, att)
    for s in range(2):
        lines = open(os.path.join(d, f"dynamic_species00{s + 1}.txt")).read().splitlines()
        assert lines == logs[s], s
    assert not out.has_dataset(os.path.join(d, "diags.jld2"), "timeseries/t/4")


def test_abiotic_particle_output_matches_oracle(pi, orc, tmp_path):
    """save_abiotic_particle and the abiotic num diagnostic through update! on the
    GPU: every saved column and timeseries/abiotic/saN/num equal the oracle's
    state at the same steps (write_outputs.jl:31-38, 159-169, 182-186)."""
    from importlib import import_module
    out = import_module("pi_amd.output")
    try:
        out.hdf5()
    except RuntimeError as e:
        pytest.skip(str(e))
    g = pi.RectilinearGrid(size=(12, 10, 3), x=(0, 12e3), y=(0, 10e3), z=(0, -30))
    spec = pi.ModelSpec(g, max_individuals=3000, seed=88, phyto=pi.phyto_setup({"dvid_P": [1e-3]}, [800], 1),
                        abiotic=pi.abiotic_setup({"sz_min": [1e-3, 2e-3]}, [300, 200], 2),
                        bgc_params={"κhP": 2.0, "κh": 10.0})
    u, v, w = pi.workloads.cellular_flow(g, U=1.0, L=4e3)
    dt, n = 60.0, 6
    PARF = np.full((g.Nx, g.Ny, n + 1), 600.0, np.float32)
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    d = str(tmp_path / "res")
    writer = pi.PlanktonOutputWriter(dir=d, save_diags=True, save_abiotic_particle=True,
                                     abiotic_particle_include=("x", "y", "z", "xi", "ac"),
                                     abiotic_particle_iteration_interval=2)
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    flux = np.float32(0.3 * spec.abiotic_matrix[0, 3] * spec.abiotic_matrix[0, 0] / (dt * g.Δy * 10.0))
    pi.set_bc_particle(gm, sa="sa1", pos="west", bc_value=flux)
    om.set_particle_bc(0, "west", flux)
    diags = pi.PlanktonDiagnostics(gm, iteration_interval=3)
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=dt, diags=diags,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6,
                                output_writer=writer)
    pi.update(sim)
    H = (g.Hx, g.Hy, g.Hz)
    inner = lambda a: a[H[0]:H[0] + g.Nx, H[1]:H[1] + g.Ny, H[2]:H[2] + g.Nz]
    om.set_velocity(u, v, w)
    t = np.float32(0.0)
    for it in range(1, n + 1):
        om.set_parf(PARF[:, :, it - 1])
        om.set_temperature(temp[:, :, :, int(np.floor(t / np.float32(3600.0)))])
        om.step(np.float32(dt))
        t = np.float32(np.float32(it) * np.float32(dt))
        if it % 2 == 0:
            f = os.path.join(d, "abiotic_particle.jld2")
            assert out.read_dataset(f, f"timeseries/t/{it}") == t
            for a in range(2):
                for att in ("x", "y", "z", "xi", "ac"):
                    got = out.read_dataset(f, f"timeseries/sa{a + 1}/{att}/{it}")
                    assert np.array_equal(got, om.abiotic_field(a, att)), (it, a, att)
        if it % 3 == 0:
            f = os.path.join(d, "diags.jld2")
            for a in range(2):
                got = out.read_dataset(f, f"timeseries/abiotic/sa{a + 1}/num/{it}")
                assert np.array_equal(got, inner(om.abiotic_diag(a)) / np.float32(3.0)), (it, a)
            om.reset_diags()
    assert om.abiotic_field(0, "ac").sum() > 300
