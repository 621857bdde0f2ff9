"""GPU parity: libpi_hip.so against the oracle (FP32 reference order) on the same
ModelSpec and inputs.  Bar: every individual field, every tracer, par/Chl/pop and
the per-species diagnostics identical (up to the sign of zero / NaN payload)."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu


def _drive(pi, orc, spec, uvw, dt, n, parf_dt=600.0, seed=3, diag_tracers=(), PARF=None):
    """The same run through update! on the GPU and through the oracle's literal
    update! (inputs re-staged every step).  `diag_tracers`: tracer names also
    accumulated into diags.tracer (PlanktonDiagnostics(model; tracer=...))."""
    g = spec.grid
    rng = np.random.default_rng(seed)
    nt = int(np.ceil(n * dt / parf_dt)) + 1
    if PARF is None:
        PARF = (400.0 + 300.0 * rng.random((g.Nx, g.Ny, nt))).astype(np.float32)
    temp = pi.default_temperature(g, dt, n).astype(np.float32)
    u, v, w = uvw
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    diags = pi.PlanktonDiagnostics(gm, tracer=("PAR",) + tuple(diag_tracers))
    sim = pi.PlanktonSimulation(gm, ΔT=dt, iterations=n, PARF=PARF, temp=temp, ΔT_PAR=parf_dt, diags=diags,
                                vels=dict(u=u[..., None], v=v[..., None], w=w[..., None]), ΔT_vel=dt * 1e6)
    pi.update(sim)
    om = orc.OracleModel(spec, diag_tracers=[pi._abi.TRACER_ID[t] for t in diag_tracers])
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om.update(dt, n, vel=(u, v, w), parf_fn=lambda t: PARF[:, :, idx(t, parf_dt)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return gm, om


def test_parity_2d_division_death(pi, orc):
    g = pi.RectilinearGrid(size=(32, 32, 1), x=(0, 32e3), y=(0, 32e3), z=(0, -10))
    spec = pi.ModelSpec(g, max_individuals=6000, seed=11,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_reg": [1.0], "mort_P": [1e-3],
                                              "grz_P": [2e-4], "dvid_reg": [1.5]}, [3000], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 50.0, "κv": 1e-4})
    uvw = pi.workloads.cellular_flow(g, U=2.0, L=8e3)
    gm, om = _drive(pi, orc, spec, uvw, 60.0, 24)
    n_gpu = gm._native.count_active(0)
    assert n_gpu == om.count_active(0)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_3d_two_species(pi, orc):
    spec, uvw = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=5, nmodes=12,
                                       overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    u, v, w = uvw
    gm, om = _drive(pi, orc, spec, (u * 20, v * 20, w * 20), 60.0, 16)
    bad = compare_models(gm._native, om, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_bounded_box(pi, orc):
    g = pi.RectilinearGrid(size=(12, 10, 4), x=(0, 12e3), y=(0, 10e3), z=[0.0, -2.0, -5.0, -9.0, -14.0],
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded))
    spec = pi.ModelSpec(g, max_individuals=1200, seed=21,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_reg": [1.0]}, [600], 1),
                        bgc_params={"κhP": 5.0, "κvP": 1e-3, "κh": 20.0, "κv": 1e-3})
    rng = np.random.default_rng(4)
    u = (rng.standard_normal((13, 10, 4)) * 3).astype(np.float32)
    v = (rng.standard_normal((12, 11, 4)) * 3).astype(np.float32)
    w = (rng.standard_normal((12, 10, 5)) * 1e-3).astype(np.float32)
    gm, om = _drive(pi, orc, spec, (u, v, w), 60.0, 12)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_crowded_cells(pi, orc):
    """Hundreds of individuals per cell: the long-segment sort path."""
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 4e3), y=(0, 4e3), z=(0, -20))
    spec = pi.ModelSpec(g, max_individuals=12000, seed=8,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "grz_P": [1e-6]}, [8000], 1),
                        bgc_params={"κhP": 10.0, "κvP": 1e-3})
    uvw = pi.workloads.cellular_flow(g, U=1.0, L=2e3)
    gm, om = _drive(pi, orc, spec, uvw, 60.0, 6)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_capacity_error_matches(pi, orc):
    g = pi.RectilinearGrid(size=(8, 8, 1), x=(0, 8e3), y=(0, 8e3), z=(0, -10))
    spec = pi.ModelSpec(g, max_individuals=120, seed=2,
                        phyto=pi.phyto_setup({"dvid_P": [0.05], "dvid_reg": [1.0e-3]}, [100], 1))
    gm = pi.PlanktonModel(pi.HIPArch(), spec=spec)
    om = orc.OracleModel(spec)
    parf = np.full((8, 8), 500.0, np.float32)
    temp = np.full((8, 8, 1), 25.0, np.float32)
    gm._native.set_parf(parf); gm._native.set_temperature(temp)
    steps_gpu = steps_ref = None
    for i in range(1, 30):
        om.set_parf(parf); om.set_temperature(temp)
        try:
            om.step(60.0)
        except pi.CapacityError:
            steps_ref = i
            break
    for i in range(1, 30):
        gm._native.step(60.0)
        try:
            gm._native.sync()
        except pi.CapacityError:
            steps_gpu = i
            break
    assert steps_ref is not None and steps_gpu == steps_ref


def test_parity_subnormal_quotients(pi, orc):
    """Velocities and tracer values so small that quotients of the interpolation,
    the random walk, the tracer diffusion and plk/volume fall in the subnormal
    range on half of the domain: the fast division flags them and the exact
    recomputation must give the oracle's values."""
    g = pi.RectilinearGrid(size=(16, 16, 4), x=(0, 16e3), y=(0, 16e3), z=(0, -40))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(NH4=1.0e-36, DOC=3.0e-35, PO4=2.0e-37, DON=1.0e-38)
    spec = pi.ModelSpec(g, max_individuals=3000, seed=17, tracer_initial=init,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "mort_reg": [1.0]}, [1500], 1),
                        bgc_params={"κhP": 1.0e-2, "κvP": 1.0e-5, "κh": 1.0, "κv": 1e-4})
    u, v, w = pi.workloads.cellular_flow(g, U=1.0, L=4e3)
    tiny = np.float32(1e-36)
    u = u * np.where(np.arange(u.shape[0])[:, None] < u.shape[0] // 2, tiny, np.float32(1))[..., None].astype(np.float32)
    v = (v * tiny).astype(np.float32)
    w = np.zeros((16, 16, 5), np.float32)
    w[:8, :, 1:4] = np.float32(3e-37)
    gm, om = _drive(pi, orc, spec, (u.astype(np.float32), v, w), 60.0, 8)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_0d_box(pi, orc):
    """The reference's 0-D box (test/example_0D_test.jl:3): one cell, Periodic x
    and y with N = 1 < H = 2, so every halo fill aliases."""
    g = pi.RectilinearGrid(size=(1, 1, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    spec = pi.ModelSpec(g, max_individuals=4096, seed=8,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_reg": [1.0], "mort_P": [1e-3]}, [1024], 1))
    z = np.zeros((1, 1, 1), np.float32)
    gm, om = _drive(pi, orc, spec, (z, z, np.zeros((1, 1, 2), np.float32)), 60.0, 10)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("size,n0,steps,grz", [((1, 1, 1), (3000, 2600), 6, 3e-8), ((1, 1, 24), (9000, 7000), 5, 2.5e-7)])
def test_parity_long_cells(pi, orc, size, n0, steps, grz):
    """Cells holding thousands of individuals (a 0-D box, a 1-D column), two
    species, frequent grazing, mortality and division: the one-block-per-cell
    paths (ordered sums over several LDS chunks, lists of grazed and dead
    entries, re-sorting by merging the arrivals into the stayers, and the full
    sort when the arrivals exceed the LDS sort) against the oracle."""
    g = pi.RectilinearGrid(size=size, x=(0, 32), y=(0, 32), z=(0, -32 * size[2]))
    spec = pi.ModelSpec(g, max_individuals=3 * max(n0), seed=31,
                        phyto=pi.phyto_setup({"dvid_P": [5e-3, 4e-3], "mort_reg": [1.0, 1.0], "mort_P": [2e-3, 3e-3],
                                              "grz_P": [grz, grz / 2], "dvid_reg": [1.5, 1.5]}, list(n0), 2),
                        bgc_params={"κhP": 1.0, "κvP": 1e-2, "κh": 10.0, "κv": 1e-3})
    nz = size[2]
    z = np.zeros((1, 1, nz), np.float32)
    gm, om = _drive(pi, orc, spec, (z, z, np.zeros((1, 1, nz + 1), np.float32)), 60.0, steps)
    for s_ in range(2):
        assert gm._native.count_active(s_) == om.count_active(s_)
    bad = compare_models(gm._native, om, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_aliased_halos_with_flow(pi, orc):
    """A column with Nx = 1 < Hx = 2 and Ny = 2 < Hy = 3, both Periodic, under
    non-zero u, v, w: the last face of each aliased direction takes the first
    face's flux and the halo fills read the halos they fill, as in the reference
    (halo_regions.jl:1-35, multi_dim_adv.jl:111-135)."""
    g = pi.RectilinearGrid(size=(1, 2, 12), x=(0, 1e3), y=(0, 2e3), z=(0, -60), halo=(2, 3, 2))
    spec = pi.ModelSpec(g, max_individuals=3000, seed=23,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_reg": [1.0]}, [1200], 1),
                        bgc_params={"κhP": 2.0, "κvP": 1e-3, "κh": 10.0, "κv": 1e-3})
    rng = np.random.default_rng(9)
    u = (rng.standard_normal((1, 2, 12)) * 2).astype(np.float32)
    v = (rng.standard_normal((1, 2, 12)) * 2).astype(np.float32)
    w = (rng.standard_normal((1, 2, 13)) * 1e-2).astype(np.float32)
    gm, om = _drive(pi, orc, spec, (u, v, w), 60.0, 12)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_graph_replay_matches_eager(pi):
    """pi_run replays one captured step graph for steady steps; its result must be
    the eager step-by-step result (pinned to the oracle by the other tests) bit
    for bit, also across a velocity change (graph dropped and captured again) and
    a change of Δt."""
    spec, (u, v, w) = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=5, nmodes=12,
                                             overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    g = spec.grid
    A = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    B = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    parf = np.full((g.Nx, g.Ny), 600.0, np.float32)
    temp = np.full((g.Nx, g.Ny, g.Nz), 22.0, np.float32)
    for h in (A, B):
        h.set_velocity(u * 20, v * 20, w * 20); h.set_parf(parf); h.set_temperature(temp)
    A.run(60.0, 7)
    for _ in range(7):
        B.step(60.0)
    for h in (A, B):
        h.set_velocity(u * 30, v * 30, w * 30)
    A.run(60.0, 4)
    A.run(30.0, 3)
    for _ in range(4):
        B.step(60.0)
    for _ in range(3):
        B.step(30.0)
    A.sync(); B.sync()
    assert A.iteration() == B.iteration() == 14
    bad = compare_models(A, B, 2)
    assert not bad, "\n".join(bad[:20])


def test_parity_velocity_cancellation(pi, orc):
    """A new velocity that nearly cancels the previous one: vel½ = (vel₀ + vel₁)/2
    holds tiny non-zero values (2^-54 m/s) in the step after the change, whose
    interpolation quotients reach the subnormal range.  Results identical to the
    oracle."""
    g = pi.RectilinearGrid(size=(16, 16, 4), x=(0, 16e3), y=(0, 16e3), z=(0, -40))
    spec = pi.ModelSpec(g, max_individuals=3000, seed=23,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "mort_reg": [1.0]}, [1500], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    u, v, w = pi.workloads.cellular_flow(g, U=2.0, L=4e3)
    u0, u1 = u.copy(), u.copy()
    u0[:9] = np.float32(2.0 ** -30)
    u1[:9] = np.float32(-(2.0 ** -30) + 2.0 ** -53)
    assert np.all((u0[:9] + u1[:9]) * np.float32(0.5) == np.float32(2.0 ** -54))
    parf = np.full((16, 16), 700.0, np.float32)
    temp = np.full((16, 16, 4), 22.0, np.float32)
    gm = pi.PlanktonModel(pi.HIPArch(0), spec=spec)
    om = orc.OracleModel(spec)
    for uu in (u0, u1):
        for h in (gm._native, om):
            h.set_velocity(uu, v, w)
        for _ in range(3):
            for h in (gm._native, om):
                h.set_parf(parf)
                h.set_temperature(temp)
                h.step(60.0)
    gm._native.sync()
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def _bits_same(a, b):
    """Bitwise equality, sign of zero included (NaN payloads aside)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | nan))


def test_parity_zero_tracers(pi, orc):
    """Tracers that are +0, -0 or subnormal on whole stencils next to ordinary
    values, in a flow of both signs in x, y and z with zero faces: face fluxes
    whose term of the sign of T is +0 or -0 (the zero rules of the fast flux)
    and the full formula elsewhere.  Tracers identical to the oracle bit for
    bit, sign of zero included."""
    g = pi.RectilinearGrid(size=(16, 16, 8), x=(0, 16e3), y=(0, 16e3), z=(0, -80))
    rng = np.random.default_rng(11)
    base = pi.fields.default_tracer_init()["initial_condition"]
    X, Y, Z = np.meshgrid(np.arange(16), np.arange(16), np.arange(8), indexing="ij")
    init = {}
    for q, name in enumerate(pi.fields.TRACER_NAMES):
        c = (max(base[name], 0.3) * (1.0 + 0.1 * rng.random((16, 16, 8)))).astype(np.float32)
        if q % 3 == 0:
            c[X < 6] = np.float32(0.0)
            c[(X >= 6) & (X < 10)] = np.float32(-0.0)
        elif q % 3 == 1:
            c[:] = np.float32(0.0)
            c[(Y + Z) % 5 == 0] = np.float32(-0.0)
            c[(X > 11) & (Y < 4)] = np.float32(1e-40)
        else:
            c[Z < 3] = np.float32(-0.0)
            c[(Z >= 3) & (Y > 9)] = np.float32(0.0)
        init[name] = c
    init["Dust"] = np.full((16, 16, 8), -0.0, np.float32)
    spec = pi.ModelSpec(g, max_individuals=3000, seed=29, tracer_initial=init,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "mort_reg": [1.0]}, [1500], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    u, v, _ = pi.workloads.cellular_flow(g, U=1.0, L=4e3)
    u[5] = np.float32(0.0)
    w = (1e-3 * (rng.random((16, 16, 9)) - 0.5)).astype(np.float32)
    w[:, :, 0] = 0.0
    w[:, :, 8] = 0.0
    w[:4, :4, 3] = np.float32(0.0)
    gm, om = _drive(pi, orc, spec, (u.astype(np.float32), v.astype(np.float32), w), 60.0, 6)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])
    strict = [q for q in range(14) if not _bits_same(gm._native.tracer(q), om.tracer(q))]
    assert not strict, f"sign of zero differs in tracers {strict}"


@pytest.mark.parametrize("nx,ny,nz,land", [(128, 6, 4, False), (256, 3, 2, False), (320, 5, 3, False),
                                            (128, 5, 3, True), (320, 4, 2, True), (512, 3, 2, False),
                                            (544, 3, 2, True)])
def test_parity_wide_rows(pi, orc, nx, ny, nz, land):
    """Rows of two and four waves (the fused x/y sweep with 256-thread blocks,
    rows of whole waves, and the z walk's lanes of one row per wave), rows of
    257-512 cells (the fused sweep with 512-thread blocks), and rows wider than
    that (the separate sweeps and line fills), periodic in x and y, a random
    flow of both signs, with and without land: identical to the oracle."""
    lm = None
    if land:
        lm = np.ones((nx, ny, nz), dtype=bool)
        lm[nx // 3: nx // 2, 1:3, :] = False
        lm[nx - 5:, 0, nz - 1] = False
    g = pi.RectilinearGrid(size=(nx, ny, nz), x=(0, nx * 1e3), y=(0, ny * 1e3), z=(0, -10.0 * nz), landmask=lm)
    ncell = nx * ny * nz
    spec = pi.ModelSpec(g, max_individuals=4 * ncell + 1000, seed=31,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "mort_reg": [1.0]}, [2 * ncell], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    rng = np.random.default_rng(nx)
    u = (rng.standard_normal((nx, ny, nz)) * 2).astype(np.float32)
    v = (rng.standard_normal((nx, ny, nz)) * 2).astype(np.float32)
    w = (rng.standard_normal((nx, ny, nz + 1)) * 1e-3).astype(np.float32)
    w[:, :, 0] = 0.0
    w[:, :, nz] = 0.0
    gm, om = _drive(pi, orc, spec, (u, v, w), 60.0, 6)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_deep_columns(pi, orc):
    """70 levels: the PAR column scan for columns deeper than its 64-level
    tiles, and z runs that do not divide the column: identical to the oracle."""
    nx, ny, nz = 4, 3, 70
    g = pi.RectilinearGrid(size=(nx, ny, nz), x=(0, nx * 1e3), y=(0, ny * 1e3), z=(0, -2.0 * nz))
    ncell = nx * ny * nz
    spec = pi.ModelSpec(g, max_individuals=4 * ncell + 1000, seed=37,
                        phyto=pi.phyto_setup({"dvid_P": [1e-3], "mort_reg": [1.0]}, [2 * ncell], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    rng = np.random.default_rng(70)
    u = (rng.standard_normal((nx, ny, nz)) * 2).astype(np.float32)
    v = (rng.standard_normal((nx, ny, nz)) * 2).astype(np.float32)
    w = (rng.standard_normal((nx, ny, nz + 1)) * 1e-3).astype(np.float32)
    w[:, :, 0] = 0.0
    w[:, :, nz] = 0.0
    gm, om = _drive(pi, orc, spec, (u, v, w), 60.0, 6)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_species_specific_grazing(pi, orc):
    """shared_graz != 1 (time_step.jl:80-118): pop is counted per species, so each
    species' grazing probability and uptake caps see only its own individuals
    per cell, and the pop field keeps the last species' counts.  Three species
    with different grazing rates on a 3-D grid, tracer diagnostics on."""
    g = pi.RectilinearGrid(size=(12, 10, 6), x=(0, 12e3), y=(0, 10e3), z=(0, -60))
    spec = pi.ModelSpec(g, max_individuals=5000, seed=41,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3, 1e-3, 3e-3], "mort_reg": [1.0, 1.0, 1.0],
                                              "grz_P": [4e-4, 1e-4, 8e-4], "dvid_reg": [1.5, 1.6, 1.4]},
                                             [2500, 1500, 2000], 3),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4, "shared_graz": 0.0})
    u, v, w = pi.workloads.cellular_flow(g, U=2.0, L=6e3)
    dtr = ("NH4", "DOC", "PFe_bio")
    gm, om = _drive(pi, orc, spec, (u, v, w), 60.0, 14, diag_tracers=dtr)
    for s in range(3):
        assert gm._native.count_active(s) == om.count_active(s)
    assert sum(float(om.diag(s, "graz").sum()) for s in range(3)) > 0
    bad = compare_models(gm._native, om, 3, diag_tracers=[pi._abi.TRACER_ID[t] for t in dtr])
    assert not bad, "\n".join(bad[:20])


@pytest.mark.parametrize("dvid_type", [2, 3])
def test_parity_division_timer(pi, orc, dvid_type):
    """Timer (type 2) and Sizer_Timer (type 3) division
    (division_death_probability.jl:9-12): the probability depends on the hour of
    the model day t % 86400 / 3600; 5-minute steps cross the 1-hour threshold
    of dvid_reg (type 2) or dvid_reg2 (type 3) halfway through the run."""
    g = pi.RectilinearGrid(size=(16, 16, 2), x=(0, 16e3), y=(0, 16e3), z=(0, -20))
    over = {"dvid_type": [dvid_type], "dvid_P": [1e-4 if dvid_type == 2 else 5e-4], "mort_reg": [1.0]}
    if dvid_type == 2:
        over["dvid_reg"] = [1.0]
    else:
        over.update(dvid_reg=[1.6], dvid_reg2=[1.0])
    spec = pi.ModelSpec(g, max_individuals=6000, seed=43, phyto=pi.phyto_setup(over, [2500], 1),
                        bgc_params={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    uvw = pi.workloads.cellular_flow(g, U=1.0, L=4e3)
    gm, om = _drive(pi, orc, spec, uvw, 300.0, 20, parf_dt=1200.0)
    assert om.diag(0, "dvid").sum() > 0
    assert gm._native.count_active(0) == om.count_active(0)
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_parity_four_species_3d(pi, orc):
    """Four species as in the large single-instance workload (512x512x128,
    4 species, dt = 30 s), at 16x16x12 with a 24-mode Fourier flow: the
    species loops of time_step.jl:44-79, species-major ordered sums, and the
    per-species division slots."""
    spec, (u, v, w) = pi.workloads.config_3d(n=16, nz=12, N=900, nsp=4, cap=2400, seed=47, nmodes=24,
                                             overrides={"dvid_P": [1e-3, 2e-3, 1.5e-3, 5e-4],
                                                        "mort_P": [5e-4, 1e-3, 2e-4, 8e-4]},
                                             kappa=(1.0, 1e-4))
    gm, om = _drive(pi, orc, spec, (u * 20, v * 20, w * 20), 30.0, 16, diag_tracers=("DIC", "Dust"))
    for s in range(4):
        assert gm._native.count_active(s) == om.count_active(s)
    bad = compare_models(gm._native, om, 4, diag_tracers=[pi._abi.TRACER_ID["DIC"], pi._abi.TRACER_ID["Dust"]])
    assert not bad, "\n".join(bad[:20])


def test_parity_plain_2d_workload(pi, orc):
    """The plain 2-D run at its full size: 128x128x1 periodic, Δ = 1 km, one
    species of 32,768 individuals, the cellular flow U = 0.1 m/s, L = 16 km,
    dt = 60 s, division and death on, NO3/NH4/PO4/DOC uniform plus 5 % noise.
    The reference's diel surface PAR and temperature series
    (param_default.jl:5-38) are staged by update! every hour; the run covers
    four hours from 05:00 (PAR switching on at dawn, five PAR and temperature
    slices).  240 steps, identical to the oracle."""
    spec, (u, v, w) = pi.workloads.config_2d(overrides={"mort_reg": [1.0], "grz_P": [1.0e-6], "dvid_reg": [1.8]},
                                             bgc={"κhP": 1.0, "κvP": 0.0, "κh": 10.0, "κv": 0.0})
    g = spec.grid
    n, dt = 240, 60.0
    PARF = np.roll(pi.default_PARF(g, dt, 1440), -5, axis=2).astype(np.float32)
    gm, om = _drive(pi, orc, spec, (u, v, w), dt, n, parf_dt=3600.0, PARF=PARF, diag_tracers=("NH4",))
    assert gm._native.count_active(0) == om.count_active(0)
    assert om.diag(0, "dvid").sum() > 0 and om.diag(0, "mort").sum() > 0
    bad = compare_models(gm._native, om, 1, diag_tracers=[pi._abi.TRACER_ID["NH4"]])
    assert not bad, "\n".join(bad[:20])
