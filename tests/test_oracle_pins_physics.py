"""More pins of the oracle (the CPU restatement the GPU is checked against),
each chosen so that a plausible mistake in the restatement fails it:

  halos       the exact identities of test/field_test.jl:21-98 (tracer, velocity
              and Gc fills, replayed statement by statement), the flux_sink
              fills (halo_regions.jl:34-35) and the composed velocity fill of
              update!'s staging (halo_regions.jl:64-99)
  PAR         the Beer-Lambert closed form down each column (utils.jl:55-68)
              on a stretched grid, with and without Chl attenuation
  sinking     a single layer: the flux difference over the unpadded dzC[k]
              index, downward, the bottom flux zeroed (tracer_sinking.jl:1-43)
  vertical    uniform and linear w (sign of z - wΔT, the z-face offset of
              w_itpl), the CFL = 1 exact upward shift of a tracer column, the
              decay of a Neumann cosine mode under κv, the z random walk
              variance with κvP > 0
  forcing     each term of tracer_forcings.jl:2-50 alone: its source and sink
              tracers, its sign, the max(0, c) clamp, iron conserved
  budgets     carbon, nitrogen and phosphorus closed with super-individuals
              large enough that the individuals' exchange is 100x the tolerance
  decisions   every grazing / mortality / division decision of a step against
              the shared stream draw and the probability of
              division_death_probability.jl:7-12,56-89 (Sizer, Timer,
              Sizer_Timer; shared and per-species pop)
"""
import ctypes as C
import math

import numpy as np
import pytest


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------------- halos
def test_field_test_halo_statements(orc):
    """test/field_test.jl:21-98 on a (4,6,2) grid with halo 2, statement by
    statement (periodic and bounded tracer fills, the bounded *_vel! and *_Gc!
    fills), each followed by the identity the reference asserts, plus the
    planes each statement must leave alone."""
    lib = orc.load()
    Nx, Ny, Nz, H = 4, 6, 2, 2
    rng = np.random.default_rng(7)
    c = np.asfortranarray(rng.random((Nx + 2 * H, Ny + 2 * H, Nz + 2 * H)).astype(np.float32))
    P, B = 0, 1

    def op(o, dim, topo):
        before = c.copy(order="F")
        assert lib.orc_fill_halo_op(o, dim, topo, Nx, Ny, Nz, H, H, H, _ptr(c)) == 0
        return before

    def planes(a, dim, lo, hi):          # 0-based planes [lo, hi) of axis dim
        s = [slice(None)] * 3
        s[dim] = slice(lo, hi)
        return a[tuple(s)]

    N = (Nx, Ny, Nz)
    for dim in range(3):                 # periodic east/north/bottom, then west/south/top
        n = N[dim]
        b = op(1, dim, P)
        assert np.array_equal(planes(c, dim, n + H, n + 2 * H), planes(c, dim, H, 2 * H))
        assert np.array_equal(planes(c, dim, 0, n + H), planes(b, dim, 0, n + H))
        b = op(0, dim, P)
        assert np.array_equal(planes(c, dim, 0, H), planes(c, dim, n, n + H))
        assert np.array_equal(planes(c, dim, H, n + 2 * H), planes(b, dim, H, n + 2 * H))
    for dim in range(3):                 # bounded east/west, north/south, bottom/top
        n = N[dim]
        c[...] = rng.random(c.shape).astype(np.float32)
        op(1, dim, B)
        for i in range(1, H + 1):
            assert np.array_equal(planes(c, dim, n + H + i - 1, n + H + i), planes(c, dim, n + H - 1, n + H))
        op(0, dim, B)
        for i in range(1, H + 1):
            assert np.array_equal(planes(c, dim, i - 1, i), planes(c, dim, H, H + 1))
    for dim in range(3):                 # bounded *_vel!: planes N+H+2:N+2H from plane N+H+1
        n = N[dim]
        c[...] = rng.random(c.shape).astype(np.float32)
        b = op(2, dim, B)
        for i in range(1, H):
            assert np.array_equal(planes(c, dim, n + H + i, n + H + i + 1),
                                  planes(c, dim, n + H, n + H + 1))
        assert np.array_equal(planes(c, dim, 0, n + H + 1), planes(b, dim, 0, n + H + 1))
        b = op(2, dim, P)                # periodic *_vel! = the periodic east/north/bottom fill
        assert np.array_equal(planes(c, dim, n + H, n + 2 * H), planes(c, dim, H, 2 * H))
    for dim in range(3):                 # bounded *_Gc!: zeros; periodic: the periodic fill
        n = N[dim]
        c[...] = rng.random(c.shape).astype(np.float32) + np.float32(1)
        b = op(3, dim, B)
        assert np.all(planes(c, dim, n + H, n + 2 * H) == 0)
        assert np.array_equal(planes(c, dim, 0, n + H), planes(b, dim, 0, n + H))
        op(3, dim, P)
        assert np.array_equal(planes(c, dim, n + H, n + 2 * H), planes(c, dim, H, 2 * H))
    # flux_sink: top planes 1:H and N+H:N+2H (the last interior level too) zeroed
    c[...] = rng.random(c.shape).astype(np.float32) + np.float32(1)
    b = op(4, 2, B)
    assert np.all(c[:, :, :H] == 0) and np.array_equal(c[:, :, H:], b[:, :, H:])
    c[...] = rng.random(c.shape).astype(np.float32) + np.float32(1)
    b = op(5, 2, B)
    assert np.all(c[:, :, Nz + H - 1:] == 0) and np.array_equal(c[:, :, :Nz + H - 1], b[:, :, :Nz + H - 1])


def test_velocity_staging_halos(pi, orc):
    """update!'s velocity staging (vel_copy!, Simulation/utils.jl:5-11, then
    fill_halo_u!/v!/w!, halo_regions.jl:64-99) on a grid bounded in x and z and
    periodic in y: the face arrays of size N+1 land on faces 1..N+1 of the
    bounded directions, the outer bounded face halo copies face N+1, the other
    halos follow the tracer rules."""
    Nx, Ny, Nz, H = 5, 4, 3, 2
    g = pi.RectilinearGrid(size=(Nx, Ny, Nz), x=(0, 5e3), y=(0, 4e3), z=(0, -30),
                           topology=(pi.Bounded, pi.Periodic, pi.Bounded))
    spec = pi.ModelSpec(g, max_individuals=4, phyto=pi.phyto_setup(None, [0], 1))
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(2)
    u = rng.random((Nx + 1, Ny, Nz)).astype(np.float32)
    v = rng.random((Nx, Ny, Nz)).astype(np.float32)
    w = rng.random((Nx, Ny, Nz + 1)).astype(np.float32)
    om.set_velocity(u, v, w)
    U, V, W = om.grid_field("u"), om.grid_field("v"), om.grid_field("w")
    assert np.array_equal(U[H:H + Nx + 1, H:H + Ny, H:H + Nz], u)
    assert np.array_equal(V[H:H + Nx, H:H + Ny, H:H + Nz], v)
    assert np.array_equal(W[H:H + Nx, H:H + Ny, H:H + Nz + 1], w)
    # u: east_vel (bounded): planes beyond face N+1 copy it; west (bounded): copies face 1
    for i in range(Nx + H + 1, Nx + 2 * H):
        assert np.array_equal(U[i], U[Nx + H])
    for i in range(H):
        assert np.array_equal(U[i], U[H])
    # v, w: periodic y; w's bottom_vel copies face N+1, its top copies face 1
    assert np.array_equal(V[:, :H], V[:, Ny:Ny + H]) and np.array_equal(V[:, Ny + H:], V[:, H:2 * H])
    for k in range(Nz + H + 1, Nz + 2 * H):
        assert np.array_equal(W[:, :, k], W[:, :, Nz + H])
    for k in range(H):
        assert np.array_equal(W[:, :, k], W[:, :, H])
    # u and v: bounded bottom (z) halo copies the last level, not the face
    for k in range(Nz + H, Nz + 2 * H):
        assert np.array_equal(U[:, :, k], U[:, :, Nz + H - 1])
    # temperature: copy_interior! (Simulation/utils.jl:13-34), interior only
    T = rng.random((Nx, Ny, Nz)).astype(np.float32)
    om.set_temperature(T)
    assert np.array_equal(g.interior(om.grid_field("temp")), T)


# --------------------------------------------------------------------- PAR
@pytest.mark.parametrize("kc", [None, 0.0])
def test_par_beer_lambert_closed_form(pi, orc, kc):
    """calc_par! (Individuals/utils.jl:55-68) level by level from the top: with
    a_k = (Chl_k / V_k · kc + kw) · ΔzF_k, par_k = PARF_k (1 - e^-a_k) / a_k and
    PARF_{k+1} = PARF_k e^-a_k.  A stretched grid (ΔzF ≠ ΔzC), random surface
    PAR; kc = 0 reduces the attenuation to kw ΔzF.  The accumulated PAR
    diagnostic equals par after one step."""
    g = pi.RectilinearGrid(size=(6, 5, 4), x=(0, 6e3), y=(0, 5e3), z=[0.0, -2.0, -5.0, -9.0, -14.0])
    bgc = {} if kc is None else {"kc": kc}
    spec = pi.ModelSpec(g, max_individuals=900, seed=5, phyto=pi.phyto_setup(None, [600], 1), bgc_params=bgc)
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(3)
    parf = (100 + 800 * rng.random((6, 5))).astype(np.float32)
    z3 = (np.zeros((6, 5, 4), np.float32), np.zeros((6, 5, 4), np.float32), np.zeros((6, 5, 5), np.float32))
    om.update(60.0, 1, vel=z3, parf=parf, temp=np.full((6, 5, 4), 20.0, np.float32))
    chl = om.grid_field("Chl").astype(np.float64)
    par = om.grid_field("par").astype(np.float64)
    b = spec.bgc_params
    kcv, kw = float(b["kc"]), float(b["kw"])
    H = g.Hz
    P = parf.astype(np.float64)
    assert chl[2:8, 2:7, H:H + 4].sum() > 0
    for k in range(4):
        kk = H + k
        dzF = float(g.dzF[kk])
        vol = float(g.Δx) * float(g.Δy) * dzF
        a = (chl[2:8, 2:7, kk] / vol * kcv + kw) * dzF
        expect = P * (1 - np.exp(-a)) / a
        assert np.allclose(par[2:8, 2:7, kk], expect, rtol=2e-6, atol=0)
        P = P * np.exp(-a)
    assert np.allclose(om.grid_field("PARF"), P, rtol=2e-6, atol=0)
    assert np.array_equal(om.grid_field("diag_PAR"), om.grid_field("par"))


# ----------------------------------------------------------------- sinking
def _column_spec(pi, g, bgc, layers, N=0, seed=1):
    init = {n: np.zeros((g.Nx, g.Ny, g.Nz), np.float32) for n in pi.TRACER_NAMES}
    for name, arr in layers.items():
        init[name] = np.asarray(arr, np.float32)
    base = {"kDOC": 0.0, "Nit": 0.0, "kDON": 0.0, "kDOP": 0.0, "kPOC": 0.0, "kPON": 0.0, "kPOP": 0.0,
            "lambda_POC": 0.0, "lambda_min": 0.0, "kdiss": 0.0, "lambda_dust": 0.0, "lambda_Fe": 0.0,
            "w_sink_org": 0.0, "w_sink_inorg": 0.0}
    base.update(bgc)
    return pi.ModelSpec(g, max_individuals=max(4, 2 * N), seed=seed, phyto=pi.phyto_setup(None, [N], 1),
                        bgc_params=base, tracer_initial=init)


def _still(g):
    return (np.zeros((g.Nx, g.Ny, g.Nz), np.float32), np.zeros((g.Nx, g.Ny, g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny, g.Nz + 1), np.float32))


@pytest.mark.parametrize("zf", [(0, -60), [0.0, -3.0, -7.0, -12.0, -20.0, -30.0, -45.0]])
def test_sinking_single_layer(pi, orc, zf):
    """Dust (w_sink_inorg) and PFe_bio (w_sink_org) each one non-zero level:
    level k loses w c ΔT / dzC[k] and level k+1 gains w c ΔT / dzC[k+1], with
    dzC indexed by the unpadded level (tracer_sinking.jl:13); nothing sinks
    out of the bottom level (its flux is zeroed with the halo,
    halo_regions.jl:35); a uniform grid conserves Σ c·dz."""
    g = pi.RectilinearGrid(size=(3, 2, 6), x=(0, 3e3), y=(0, 2e3), z=zf)
    ws_in, ws_org, dt = 1.0e-3, 2.0e-3, 100.0
    dust = np.zeros((3, 2, 6), np.float32)
    dust[:, :, 2] = 1.5
    dust[:, :, 5] = 0.75                  # bottom level: stays
    pfe = np.zeros((3, 2, 6), np.float32)
    pfe[:, :, 4] = 2.0                    # sinks into the bottom level
    spec = _column_spec(pi, g, {"w_sink_inorg": ws_in, "w_sink_org": ws_org}, {"Dust": dust, "PFe_bio": pfe})
    om = orc.OracleModel(spec)
    om.update(dt, 1, vel=_still(g), parf=np.zeros((3, 2), np.float32), temp=np.zeros((3, 2, 6), np.float32))
    dzC = g.dzC.astype(np.float64)        # padded array, index k-1 = Julia's g.dzC[k]
    H = g.Hz
    D = g.interior(om.tracer(pi._abi.TRACER_ID["Dust"])).astype(np.float64)[0, 0]
    F = g.interior(om.tracer(pi._abi.TRACER_ID["PFe_bio"])).astype(np.float64)[0, 0]
    eD = np.zeros(6)
    eD[2] = 1.5 - ws_in * 1.5 * dt / dzC[2]        # level 3 (1-based) divides by g.dzC[3]
    eD[3] = ws_in * 1.5 * dt / dzC[3]
    eD[5] = 0.75
    eF = np.zeros(6)
    eF[4] = 2.0 - ws_org * 2.0 * dt / dzC[4]
    eF[5] = ws_org * 2.0 * dt / dzC[5]
    assert np.allclose(D, eD, rtol=1e-6, atol=1e-12)
    assert np.allclose(F, eF, rtol=1e-6, atol=1e-12)
    if isinstance(zf, tuple):
        dz = g.dzF[H:H + 6].astype(np.float64)
        assert abs((D * dz).sum() - (1.5 + 0.75) * 10.0) < 1e-5
    # every other tracer untouched
    for name in ("DIC", "PFe_inorg", "DFe", "POC"):
        assert np.all(g.interior(om.tracer(pi._abi.TRACER_ID[name])) == 0)


# ---------------------------------------------------------------- vertical
@pytest.mark.parametrize("w0", [0.01, -0.01])
def test_uniform_w_moves_individuals(pi, orc, w0):
    """Constant w (positive upward: the grid's k grows downward): z changes by
    -w ΔT / Δz per step (calc_coord_kernel!: z - w·ΔT, particle_advection_kernels.jl:40-50),
    with the first step's RK4 stage 1 at vel₀ = 0 (half a step after step 1);
    individuals that cross the surface or the bottom are reflected, one shot
    (particle_boundary_condition, Bounded: z < 0 -> -z, z > N -> 2N - z)."""
    g = pi.RectilinearGrid(size=(4, 4, 32), x=(0, 4.0), y=(0, 4.0), z=(0, -32.0))
    spec = pi.ModelSpec(g, max_individuals=4000, seed=3, phyto=pi.phyto_setup(None, [3000], 1))
    om = orc.OracleModel(spec)
    w = np.full((4, 4, 33), w0, np.float32)
    z3 = (np.zeros((4, 4, 32), np.float32), np.zeros((4, 4, 32), np.float32), w)
    z0 = om.field(0, "z").astype(np.float64)
    ac = om.field(0, "ac")
    om.update(10.0, 2, vel=z3, parf=np.zeros((4, 4), np.float32), temp=np.zeros((4, 4, 32), np.float32))
    z2 = om.field(0, "z").astype(np.float64)
    refl = lambda z: np.where(z < 0, -z, np.where(z > 32, 64 - z, z))
    expect = refl(refl(z0 - 0.5 * w0 * 10.0) - w0 * 10.0)
    assert np.allclose(z2[ac], expect[ac], atol=2e-6)
    crossed = (z0 - 1.5 * w0 * 10.0 < 0) | (z0 - 1.5 * w0 * 10.0 > 32)
    assert (crossed & ac).sum() > 5


def test_linear_w_interpolation(pi, orc):
    """w linear in depth on the z faces, w(k) = a + b k: w_itpl puts w on the
    faces (z + 1) and divides by ΔzC, so dz/dt = -(a + b z) / Δz, solved exactly
    by z(t) = (z0 + a/b) e^{-b t / Δz} - a/b; a wrong face offset or sign fails."""
    g = pi.RectilinearGrid(size=(2, 2, 16), x=(0, 2.0), y=(0, 2.0), z=(0, -16.0))
    spec = pi.ModelSpec(g, max_individuals=600, seed=9, phyto=pi.phyto_setup(None, [300], 1))
    om = orc.OracleModel(spec)
    a, b = -1e-3, 2e-4
    kf = np.arange(17, dtype=np.float64)
    w = np.broadcast_to((a + b * kf)[None, None, :], (2, 2, 17)).astype(np.float32)
    z3 = (np.zeros((2, 2, 16), np.float32), np.zeros((2, 2, 16), np.float32), np.asfortranarray(w))
    zp, zt = np.zeros((2, 2), np.float32), np.zeros((2, 2, 16), np.float32)
    om.update(10.0, 1, vel=z3, parf=zp, temp=zt)
    z1 = om.field(0, "z").astype(np.float64)
    om.update(10.0, 5, vel=z3, parf=zp, temp=zt)
    z6 = om.field(0, "z").astype(np.float64)
    ac = om.field(0, "ac") & (z1 > 0.5) & (z1 < 14.5)
    exact = (z1 + a / b) * np.exp(-b * 50.0) - a / b
    assert ac.sum() > 150
    assert np.max(np.abs(z6[ac] - exact[ac])) < 5e-5


def test_dst3fl_cfl1_is_exact_upward_shift(pi, orc):
    """w ΔT / Δz = 1 upward: the limiter vanishes (d0 = d1 = 0) and the z sweep
    (DST3FL.jl:63-73, multi_dim_adv.jl) moves every level's value one level up;
    the bottom level empties (the bounded bottom Gc halo is zero, nothing comes
    in from below) and the top level's content leaves through the surface.
    Powers of two keep every product exact."""
    g = pi.RectilinearGrid(size=(2, 2, 8), x=(0, 2.0), y=(0, 2.0), z=(0, -32.0))
    rng = np.random.default_rng(5)
    pop = (np.round(rng.random((2, 2, 8)) * 64) / 64 + 0.25).astype(np.float32)
    spec = _column_spec(pi, g, {}, {"POP": pop})
    om = orc.OracleModel(spec)
    w = np.ones((2, 2, 9), np.float32)
    z3 = (np.zeros((2, 2, 8), np.float32), np.zeros((2, 2, 8), np.float32), w)
    om.update(4.0, 1, vel=z3, parf=np.zeros((2, 2), np.float32), temp=np.zeros((2, 2, 8), np.float32))
    out = g.interior(om.tracer(pi._abi.TRACER_ID["POP"]))
    expect = np.zeros_like(pop)
    expect[:, :, :7] = pop[:, :, 1:]
    assert np.array_equal(out, expect)


def test_vertical_diffusion_neumann_mode(pi, orc):
    """tracer_diffusion! in z (tracer_diffusion.jl:1-30): the bounded column has
    no flux through the top and bottom faces (the halo copies the edge level),
    so cos(π m (k - ½) / N) is an eigenmode and decays by
    1 - 4 κv ΔT / Δz² sin²(π m / 2N) per step."""
    n, m, kap, dt = 32, 3, 0.05, 1.0
    g = pi.RectilinearGrid(size=(2, 2, n), x=(0, 2.0), y=(0, 2.0), z=(0, -float(n)))
    k = np.arange(1, n + 1)
    mode = np.cos(np.pi * m * (k - 0.5) / n)
    poc = np.broadcast_to((10.0 + mode)[None, None, :], (2, 2, n)).astype(np.float32)
    spec = _column_spec(pi, g, {"κv": kap, "κh": 0.0}, {"POC": poc})
    om = orc.OracleModel(spec)
    om.update(dt, 1, vel=_still(g), parf=np.zeros((2, 2), np.float32), temp=np.zeros((2, 2, n), np.float32))
    out = g.interior(om.tracer(pi._abi.TRACER_ID["POC"]))[0, 0].astype(np.float64) - 10.0
    fac = 1 - 4 * kap * dt * np.sin(np.pi * m / (2 * n)) ** 2
    assert np.allclose(out, fac * mode, atol=3e-6)
    assert abs(fac - 1) > 1e-3


def test_vertical_random_walk_statistics(pi, orc):
    """Zero flow, κvP > 0: the z walk is U(-1,1)·√(κvP ΔT)/ΔzC per step
    (particle_diffusion.jl:2-33), zero mean and variance κvP ΔT / (3 Δz²)."""
    g = pi.RectilinearGrid(size=(8, 8, 64), x=(0, 8e3), y=(0, 8e3), z=(0, -64.0))
    kap, dt = 1e-3, 60.0
    spec = pi.ModelSpec(g, max_individuals=20000, seed=13, phyto=pi.phyto_setup(None, [20000], 1),
                        bgc_params={"κhP": 0.0, "κvP": kap})
    om = orc.OracleModel(spec)
    z0 = om.field(0, "z").astype(np.float64)
    om.update(dt, 1, vel=_still(g), parf=np.zeros((8, 8), np.float32), temp=np.zeros((8, 8, 64), np.float32))
    d = om.field(0, "z").astype(np.float64) - z0
    keep = (z0 > 0.5) & (z0 < 63.5)
    d = d[keep]
    s2 = kap * dt / 3.0
    assert abs(d.mean()) < 4 * math.sqrt(s2 / len(d))
    assert abs(d.var() / s2 - 1.0) < 0.05
    assert np.all(np.abs(d) <= math.sqrt(kap * dt) * (1 + 1e-6))


# ----------------------------------------------------------------- forcing
def _forcing_run(pi, orc, bgc, values, dt=100.0):
    g = pi.RectilinearGrid(size=(4, 3, 2), x=(0, 4e3), y=(0, 3e3), z=(0, -20))
    spec = _column_spec(pi, g, bgc, {})
    om = orc.OracleModel(spec)
    for name, arr in values.items():
        c = np.zeros(g.padded_shape, np.float32, order="F")
        g.interior(c)[...] = arr
        pi.fill_halo_tracer(c, g)
        om.upload_tracer(pi._abi.TRACER_ID[name], c)
    om.update(dt, 1, vel=_still(g), parf=np.zeros((4, 3), np.float32), temp=np.zeros((4, 3, 2), np.float32))
    after = {n: g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64) for n in pi.TRACER_NAMES}
    return after, g


def _rand(seed, lo, hi, neg=False):
    rng = np.random.default_rng(seed)
    a = (lo + (hi - lo) * rng.random((4, 3, 2))).astype(np.float32)
    if neg:
        a[0, :, 0] = -a[0, :, 0]        # negative values force nothing (max(0, c))
    return a


def test_forcing_nitrification(pi, orc):
    """NH4 -> NO3 at rate Nit (tracer_forcings.jl:16-19) on max(0, NH4); N conserved."""
    nh4, no3 = _rand(1, 0.1, 1.0, neg=True), _rand(2, 0.1, 1.0)
    Nit, dt = 2e-5, 100.0
    out, _ = _forcing_run(pi, orc, {"Nit": Nit}, {"NH4": nh4, "NO3": no3}, dt)
    f = np.maximum(0.0, nh4.astype(np.float64)) * Nit * dt
    assert np.allclose(out["NH4"], nh4 - f, rtol=1e-6, atol=1e-9)
    assert np.allclose(out["NO3"], no3 + f, rtol=1e-6, atol=1e-9)
    assert np.array_equal(out["NH4"][0, :, 0], nh4[0, :, 0].astype(np.float64))
    for n in ("DON", "PON", "DIC", "DOC"):
        assert np.all(out[n] == 0)


@pytest.mark.parametrize("term", ["kdiss", "lambda_min", "lambda_dust", "lambda_POC", "lambda_Fe", "kPOC"])
def test_forcing_iron_terms(pi, orc, term):
    """Each iron term of tracer_forcings.jl:35-50 alone: dissolution
    PFe_inorg -> DFe (kdiss), scavenging DFe -> PFe_inorg (lambda_min, and
    lambda_dust with Dust), biological scavenging DFe -> PFe_bio (lambda_POC
    with POC), coagulation of DFe above the ligand -> PFe_inorg (lambda_Fe),
    remineralisation PFe_bio -> DFe and POC -> DOC (kPOC).  Iron
    (DFe + PFe_inorg + PFe_bio) is conserved."""
    dt = 100.0
    vals = {"DFe": _rand(3, 2e-4, 1.2e-3, neg=True), "PFe_inorg": _rand(4, 1e-4, 1e-3),
            "PFe_bio": _rand(5, 1e-4, 1e-3), "POC": _rand(6, 0.5, 2.0), "Dust": _rand(7, 0.1, 1.0)}
    rates = {"kdiss": 4e-5, "lambda_min": 3e-5, "lambda_dust": 2e-3, "lambda_POC": 5e-3, "lambda_Fe": 40.0,
             "kPOC": 2e-5}
    bgc = {term: rates[term], "DFeFrac": 0.5, "ligand": 6.0e-4}
    out, _ = _forcing_run(pi, orc, bgc, vals, dt)
    T = {k: np.maximum(0.0, v.astype(np.float64)) for k, v in vals.items()}
    r, fr = rates[term], 0.5
    if term == "kdiss":
        src, dst, f = "PFe_inorg", "DFe", T["PFe_inorg"] * r * dt
    elif term == "lambda_min":
        src, dst, f = "DFe", "PFe_inorg", r * T["DFe"] * fr * dt
    elif term == "lambda_dust":
        src, dst, f = "DFe", "PFe_inorg", r * T["Dust"] * T["DFe"] * fr * dt
    elif term == "lambda_POC":
        src, dst, f = "DFe", "PFe_bio", r * T["POC"] * T["DFe"] * fr * dt
    elif term == "lambda_Fe":
        src, dst, f = "DFe", "PFe_inorg", np.maximum(T["DFe"] - 6.0e-4, 0.0) * T["DFe"] * fr * r * dt
        assert np.any(f == 0) and np.any(f > 0)
    else:
        src, dst, f = "PFe_bio", "DFe", T["PFe_bio"] * r * dt
    assert np.max(f) > 1e-3 * np.max(T[src])
    assert np.allclose(out[src], vals[src] - f, rtol=1e-5, atol=1e-12)
    assert np.allclose(out[dst], vals[dst] + f, rtol=1e-5, atol=1e-12)
    fe0 = sum(vals[k].astype(np.float64) for k in ("DFe", "PFe_inorg", "PFe_bio"))
    fe1 = sum(out[k] for k in ("DFe", "PFe_inorg", "PFe_bio"))
    assert np.allclose(fe1, fe0, rtol=1e-6, atol=1e-12)
    if term == "kPOC":
        fc = T["POC"] * r * dt
        assert np.allclose(out["POC"], vals["POC"] - fc, rtol=1e-6)
        assert np.allclose(out["DOC"], fc, rtol=1e-6)
    else:
        assert np.array_equal(out["POC"], vals["POC"].astype(np.float64))


# ----------------------------------------------------------------- budgets
def _totals(pi, om, g, p):
    vol = g.Δx * g.Δy * g.dzF[g.Hz:g.Hz + g.Nz].astype(np.float64)
    tr = {n: float((g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64).sum(axis=(0, 1)) * vol).sum())
          for n in pi.TRACER_NAMES}
    ac = om.field(0, "ac")
    f = {n: om.field(0, n)[ac].astype(np.float64) for n in ("Bm", "Cq", "Nq", "Pq")}
    names = pi._abi.PARAM_NAMES
    R_NC, R_PC = float(p[names.index("R_NC")]), float(p[names.index("R_PC")])
    ind = {"C": float((f["Bm"] + f["Cq"]).sum()), "N": float((f["Nq"] + f["Bm"] * R_NC).sum()),
           "P": float((f["Pq"] + f["Bm"] * R_PC).sum())}
    tot = {"C": tr["DIC"] + tr["DOC"] + tr["POC"] + ind["C"],
           "N": tr["NH4"] + tr["NO3"] + tr["DON"] + tr["PON"] + ind["N"],
           "P": tr["PO4"] + tr["DOP"] + tr["POP"] + ind["P"]}
    return tot, ind


def test_carbon_nitrogen_phosphorus_budgets(pi, orc):
    """QuotaMode with light, uptake, respiration, exudation, grazing, mortality
    and division, no flow: total C (DIC+DOC+POC+Σ(Cq+Bm)), N and P are each
    conserved (the reference's own test, test/example_2D_test.jl:19-47, checks
    P with atol 1e1).  Super-individuals of 1e12 cells make the individuals'
    exchange with the tracers at least 100x the tolerance, so a dropped or
    mis-signed term of photosynthesis, respiration, uptake, exudation, the
    losses or division fails."""
    g = pi.RectilinearGrid(size=(16, 16, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DIC=2.0, DOC=1.0)
    spec = pi.ModelSpec(g, max_individuals=4096, seed=19, tracer_initial=init,
                        phyto=pi.phyto_setup({"Nsuper": [1e12], "dvid_P": [2e-3], "mort_reg": [1.0],
                                              "grz_P": [1e-4], "VDOCmax": [1e-6], "KsatDOC": [0.2]}, [1024], 1))
    om = orc.OracleModel(spec)
    p = spec.species_params[0]
    tot0, ind0 = _totals(pi, om, g, p)
    z3 = _still(g)
    om.update(60.0, 30, vel=z3, parf=np.full((16, 16), 900.0, np.float32), temp=np.full((16, 16, 1), 26.0, np.float32))
    tot1, ind1 = _totals(pi, om, g, p)
    assert om.diag(0, "dvid").sum() > 0 and om.diag(0, "graz").sum() > 0 and om.diag(0, "mort").sum() > 0
    for e, tol in (("C", 1e-6), ("N", 1e-6), ("P", 1e-6)):
        assert abs(tot1[e] - tot0[e]) <= tol * tot0[e], (e, tot0[e], tot1[e])
        assert abs(ind1[e] - ind0[e]) > 100 * tol * tot0[e], (e, ind0[e], ind1[e])


# --------------------------------------------------------------- decisions
def _shape_dec(x, xmax, k, pw):
    fx = np.maximum(0.0, np.minimum(1.0, 1.0 - x / xmax))
    return fx ** pw / (k + fx ** pw)


def _decisions_case(pi, orc, over, nsp=1, dt=60.0, shared=True, N=6000):
    g = pi.RectilinearGrid(size=(6, 6, 1), x=(0, 6e3), y=(0, 6e3), z=(0, -10))
    bgc = {} if shared else {"shared_graz": 0.0}
    spec = pi.ModelSpec(g, max_individuals=N + 100, seed=23,
                        phyto=pi.phyto_setup(over, [N] * nsp, nsp), bgc_params=bgc)
    om = orc.OracleModel(spec)
    om.update(dt, 1, vel=_still(g), parf=np.full((6, 6), 500.0, np.float32), temp=np.full((6, 6, 1), 24.0, np.float32))
    # brute-force per-cell counts of the individuals alive at the gather (all, in step 1)
    cnt = np.zeros((nsp,) + g.padded_shape)
    data = []
    for s in range(nsp):
        ac = om.field(s, "ac")
        xi, yi, zi = (om.field(s, f) for f in ("xi", "yi", "zi"))
        np.add.at(cnt[s], (xi[ac] - 1, yi[ac] - 1, zi[ac] - 1), 1.0)
        data.append((ac, xi, yi, zi))
    out = []
    for s in range(nsp):
        ac, xi, yi, zi = data[s]
        popc = cnt.sum(axis=0) if shared else cnt[s]
        pop = popc[xi - 1, yi - 1, zi - 1]
        slots = np.nonzero(ac)[0]
        u = np.stack(pi.stream.uniform3(spec.seed, 1, s, pi.stream.KIND_DECISION, slots)).astype(np.float64)
        q = dict(zip(pi._abi.PARAM_NAMES, spec.species_params[s].astype(np.float64)))
        Sz = om.field(s, "Sz")[slots].astype(np.float64)
        dec = {k: om.field(s, k)[slots] for k in ("graz", "mort", "dvid")}
        out.append((slots, pop[slots], u, q, Sz, dec))
    return out


def _check(dec, u, pdt, what, min_count=50):
    """every decision is U < P·ΔT, except within rounding of the threshold"""
    expect = u < pdt
    near = np.abs(u - pdt) <= 1e-6 * np.maximum(pdt, 1e-30)
    bad = (dec.astype(bool) != expect) & ~near
    assert not bad.any(), f"{what}: {bad.sum()} decisions differ"
    n, mu = float(expect.sum()), float(np.clip(pdt, 0, 1).sum())
    assert n >= min_count, f"{what}: only {n} events"
    assert abs(n - mu) < 5 * math.sqrt(mu) + 1, (what, n, mu)


@pytest.mark.parametrize("shared", [True, False])
def test_grazing_mortality_sizer_decisions(pi, orc, shared):
    """Grazing P = pop·grz_P (quadratic grazing, pop of every species when
    shared_graz = 1, of the individual's own species otherwise,
    time_step.jl:21-118), mortality P = mort_P·shape_func_dec(Sz, mort_reg, 1e-5)
    (pow 4), Sizer division P = dvid_P·shape_func_inc(Sz, dvid_reg, 2e-3; pow = 2)
    (division_death_probability.jl:7,56-89, Individuals/utils.jl:109-121)."""
    over = {"grz_P": [2e-6, 4e-6], "mort_P": [2e-3, 1e-3], "mort_reg": [1.3, 1.4], "dvid_P": [2e-3, 3e-3],
            "dvid_reg": [1.3, 1.25]}
    for slots, pop, u, q, Sz, dec in _decisions_case(pi, orc, over, nsp=2, shared=shared):
        dt = 60.0
        _check(dec["graz"], u[0], pop * q["grz_P"] * dt, "graz")
        _check(dec["mort"], u[1], q["mort_P"] * _shape_dec(Sz, q["mort_reg"], 1e-5, 4) * dt, "mort")
        _check(dec["dvid"], u[2], q["dvid_P"] * (1 - _shape_dec(Sz, q["dvid_reg"], 2e-3, 2)) * dt, "dvid")


@pytest.mark.parametrize("dvid_type", [2, 3])
def test_timer_division_decisions(pi, orc, dvid_type):
    """Timer: P = dvid_P·shape_func_inc(t % 86400 / 3600, dvid_reg, 1e-5) and
    Sizer_Timer: P = dvid_P·shape_func_inc(Sz, dvid_reg, 1e-3)·shape_func_inc(hour,
    dvid_reg2, 1e-5) (division_death_probability.jl:9-12, pow 4), at the
    model time of the step (t = iteration·ΔT, time_step.jl:15): hour 0.8333."""
    dt = 3000.0
    over = {"dvid_type": [dvid_type], "dvid_P": [2e-4], "dvid_reg": [1.0 if dvid_type == 2 else 1.3],
            "dvid_reg2": [1.0]}
    hour = (dt % 86400) / 3600
    for slots, pop, u, q, Sz, dec in _decisions_case(pi, orc, over, dt=dt, N=20000):
        th = 1 - _shape_dec(hour, q["dvid_reg"] if dvid_type == 2 else q["dvid_reg2"], 1e-5, 4)
        if dvid_type == 2:
            P = q["dvid_P"] * th * np.ones_like(Sz)
        else:
            P = q["dvid_P"] * (1 - _shape_dec(Sz, q["dvid_reg"], 1e-3, 4)) * th
        _check(dec["dvid"], u[2], P * dt, f"dvid type {dvid_type}", min_count=20)


# ----------------------------------------------------- growth, transcribed
def _f64_growth(q, t, s, dt):
    """plankton_growth! of one step for individuals without decisions, written
    in float64 from QuotaMode/growth_kernels.jl:1-171 and plankton_growth.jl
    (an independent transcription: numpy, vectorised, no shared code)."""
    e30 = 1e-30
    ex = lambda v: np.exp(v)
    Topt, Tmax, Ea = q["Topt"], q["Tmax"], q["Ea"]
    OGT = ex(-Ea / (8.3145 * (Topt + 273.15)))
    kT = ex(-Ea / (8.3145 * (t["T"] + 273.15)))
    tf_ps = np.minimum(1.0, np.maximum(0.0, kT * _shape_dec(t["T"] - Topt, Tmax - Topt, 4e-2, 4)) / OGT)
    tf = np.minimum(1.0, np.maximum(0.0, kT) / OGT)
    Bm, Cq, Nq, Pq, Chl = (s[k].copy() for k in ("Bm", "Cq", "Nq", "Pq", "Chl"))
    aI = t["par"] * q["α"] * q["Φ"]
    PCm = q["PCmax"] * tf_ps
    PS = PCm * (1 - ex(-aI / np.maximum(e30, PCm) * Chl / np.maximum(e30, Bm))) * Bm
    Qn = (Nq + Bm * q["R_NC"]) / np.maximum(e30, Bm + Cq)
    Qp = (Pq + Bm * q["R_PC"]) / np.maximum(e30, Bm + Cq)
    rN, rP = _shape_dec(Qn, q["Nqmax"], 1e-4, 4), _shape_dec(Qp, q["Pqmax"], 1e-4, 4)
    pp = np.maximum(1.0, t["pop"])
    V = {}
    ncap = 0
    for n, r in (("NH4", rN), ("NO3", rN), ("PO4", rP)):
        v = q[f"V{n}max"] * r * t[n] / np.maximum(e30, t[n] + q[f"Ksat{n}"]) * tf * Bm
        V[n] = np.minimum(v, t[n] / dt / pp)
        ncap += int(np.sum(t[n] / dt / pp < 0.5 * v))
    Cq = Cq + PS * dt
    Pq = Pq + V["PO4"] * dt
    Nq = Nq + (V["NH4"] + V["NO3"]) * dt
    Qc = Cq / np.maximum(e30, Bm + Cq)
    VN = q["VDOCmax"] * _shape_dec(Qc, q["Cqmax"], 1e-4, 4) * t["DOC"] / np.maximum(e30, t["DOC"] + q["KsatDOC"]) * tf * Bm
    VDOC = np.minimum(VN, t["DOC"] / dt / pp)
    with np.errstate(invalid="ignore", divide="ignore"):
        VDOC = np.where(PS / (VDOC + PS) > 1e-2, VDOC, 0.0)
    rho = PS / np.maximum(e30, Bm) * q["Chl2N"] / np.maximum(e30, t["par"] * q["α"] * q["Φ"] * Chl / Bm)
    rho = np.where(t["par"] > 0.1, rho, 0.0)
    resp = q["respir"] * Bm * tf
    Cq = Cq + (VDOC - resp) * dt
    borrow = np.maximum(0.0, -Cq)
    Bm, Nq, Pq, Cq = Bm - borrow, Nq + borrow * q["R_NC"], Pq + borrow * q["R_PC"], Cq + borrow
    mn = np.minimum(np.minimum(Cq, Nq / q["R_NC"]), Pq / q["R_PC"])
    BS = mn * q["k_mtb"]
    exu = np.maximum(0.0, Cq - mn) * q["k_mtb"]
    Bm, Cq = Bm + dt * BS, Cq - dt * (BS + exu)
    Nq, Pq = Nq - dt * BS * q["R_NC"], Pq - dt * BS * q["R_PC"]
    Chl = Chl + dt * BS * q["R_NC"] * rho
    out = dict(Bm=Bm, Cq=Cq, Nq=Nq, Pq=Pq, Chl=Chl, Sz=Bm / (q["Cquota"] * q["Nsuper"]), age=s["age"] + dt / 3600.0)
    cons = dict(DIC=(resp - PS) * dt, DOC=(exu - VDOC) * dt, NH4=-V["NH4"] * dt, NO3=-V["NO3"] * dt, PO4=-V["PO4"] * dt)
    return out, cons, ncap


def test_growth_and_consumption_transcription(pi, orc):
    """One step of QuotaMode growth and the consumption scatter against an
    independent float64 transcription of growth_kernels.jl, fed by a
    brute-force find_NPT! gather (utils.jl:14-30): dark, dim and bright columns
    (the PS, ρChl and DOC-uptake gates), temperatures on both sides of Topt (both
    temperature functions), DOC uptake on, nutrients low enough in some cells
    that the c/ΔT/pop caps bind for some individuals.  Individual state and the per-cell tracer
    change (calc_consume!, consume_loss.jl:2-14, divided by the cell volume)
    agree to 1e-4."""
    g = pi.RectilinearGrid(size=(4, 4, 3), x=(0, 4.0), y=(0, 4.0), z=[0.0, -4.0, -10.0, -18.0])
    rng = np.random.default_rng(31)
    init = {n: np.zeros((4, 4, 3), np.float32) for n in pi.TRACER_NAMES}
    init.update(NH4=(0.5 * rng.random((4, 4, 3))).astype(np.float32),
                NO3=(0.8 * rng.random((4, 4, 3))).astype(np.float32),
                PO4=(0.1 * rng.random((4, 4, 3))).astype(np.float32),
                DOC=(5.0 * rng.random((4, 4, 3))).astype(np.float32),
                DIC=np.full((4, 4, 3), 20.0, np.float32))
    init["NH4"][0, 0, :] = 1e-9
    init["PO4"][1, 1, :] = 2e-10
    bgc = {k: 0.0 for k in ("kDOC", "Nit", "kDON", "kDOP", "kPOC", "kPON", "kPOP")}
    spec = pi.ModelSpec(g, max_individuals=500, seed=37, tracer_initial=init,
                        phyto=pi.phyto_setup({"Nsuper": [1e12], "VDOCmax": [4e-6], "KsatDOC": [0.5]}, [400], 1),
                        bgc_params=bgc)
    om = orc.OracleModel(spec)
    ini = spec.initial_individuals[0]
    parf = np.zeros((4, 4))
    parf[1] = (0.15, 0.5, 1.5, 5.0)       # dim columns: par across the 0.1 gate, PS ~ VDOC
    parf[2:] = 1500.0
    temp = (20.0 + 9.0 * rng.random((4, 4, 3))).astype(np.float32)
    dt = 600.0
    om.update(dt, 1, vel=_still(g), parf=parf.astype(np.float32), temp=temp)
    ac = om.field(0, "ac")
    sl = np.nonzero(ac)[0]
    xi, yi, zi = (om.field(0, f)[sl] - 1 for f in ("xi", "yi", "zi"))
    pop = np.zeros(g.padded_shape)
    np.add.at(pop, (xi, yi, zi), 1.0)
    par = om.grid_field("par").astype(np.float64)
    H = g.Hx
    t = {n: np.maximum(0.0, spec.initial_tracers[pi._abi.TRACER_ID[n]][xi, yi, zi].astype(np.float64))
         for n in ("NH4", "NO3", "PO4", "DOC")}
    t["par"] = par[xi, yi, zi]
    t["T"] = temp[xi - H, yi - H, zi - H].astype(np.float64)
    t["pop"] = pop[xi, yi, zi]
    assert (t["par"] == 0).any() and (t["par"] > 1).any() and ((t["par"] > 0.1) & (t["par"] < 1)).any()
    q = dict(zip(pi._abi.PARAM_NAMES, spec.species_params[0].astype(np.float64)))
    s = {k: ini[k][sl].astype(np.float64) for k in ("Bm", "Cq", "Nq", "Pq", "Chl", "age")}
    exp_state, cons, ncap = _f64_growth(q, t, s, dt)
    assert ncap > 0                       # the c/ΔT/pop caps bind for some individuals
    scale = np.abs(s["Bm"]) + np.abs(s["Cq"])
    for k, v in exp_state.items():
        got = om.field(0, k)[sl].astype(np.float64)
        ref = scale * (q["R_NC"] if k == "Nq" else q["R_PC"] if k == "Pq" else 1.0) if k in ("Cq", "Nq", "Pq") else np.abs(v)
        assert np.all(np.abs(got - v) <= 1e-4 * ref + 1e-30), k
    vol = (g.Δx * g.Δy * g.dzF.astype(np.float64))[None, None, :]
    for n, c in cons.items():
        d = np.zeros(g.padded_shape)
        np.add.at(d, (xi, yi, zi), c)
        before = spec.initial_tracers[pi._abi.TRACER_ID[n]].astype(np.float64)
        after = om.tracer(pi._abi.TRACER_ID[n]).astype(np.float64)
        I = (slice(H, H + 4), slice(H, H + 4), slice(g.Hz, g.Hz + 3))
        expect = before[I] + (d / vol)[I]
        assert np.allclose(after[I], expect, rtol=1e-5, atol=1e-4 * np.abs((d / vol)[I]).max() + 1e-12), n
        assert np.abs((d / vol)[I]).max() > 1e-5 * np.abs(before[I]).max(), n     # a visible exchange


def test_find_inds_brute_force(pi, orc):
    """find_inds! (Individuals/utils.jl:2-12): xi = trunc(x + 1) + H for active
    individuals (so x = N exactly lands in the first halo cell), H for inactive
    ones; positions placed on cell edges, inside cells and at 0 and N."""
    g = pi.RectilinearGrid(size=(5, 4, 3), x=(0, 5e3), y=(0, 4e3), z=(0, -30))
    spec = pi.ModelSpec(g, max_individuals=40, seed=2, phyto=pi.phyto_setup(None, [30], 1))
    ini = dict(spec.initial_individuals[0])
    xs = np.array([0.0, 0.5, 1.0, 2.999999, 3.0, 4.5, 5.0] * 6)[:40].astype(np.float32)
    ys = np.array([0.0, 4.0, 1.25, 3.999, 2.0] * 8)[:40].astype(np.float32)
    zs = np.array([0.0, 3.0, 0.001, 2.5] * 10)[:40].astype(np.float32)
    ini.update(x=xs, y=ys, z=zs)
    ac = np.ones(40, bool)
    ac[35:] = False
    ini["ac"] = ac
    om = orc.OracleModel(spec)
    om.upload_individuals(0, ini)
    om.update(1.0, 1, vel=_still(g), parf=np.zeros((5, 4), np.float32), temp=np.zeros((5, 4, 3), np.float32))
    H = g.Hx
    for f, pos in (("xi", xs), ("yi", ys), ("zi", zs)):
        got = om.field(0, f)
        expect = np.where(ac, np.trunc(pos.astype(np.float64) + 1).astype(np.int64) + H, H)
        assert np.array_equal(got, expect), f
