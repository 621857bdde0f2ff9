"""Pins of LatLonGrid (Grids/lat_lon_grid.jl) and of the oracle's use of its
metric arrays (Grids/areas_volumes_spacings.jl:24-38): every spacing, area and
volume taken from the array at the reference's index.

  grid        LatLonGrid: the sizes and topologies of test/grid_test.jl:71-84,
              dyC = R Δφ, dxC = R cos φ Δλ per row, the zonal sum of Az over
              the latitudes = R² Δλ (sin φN − sin φS), Ax = dzF dyF and
              Vol = Az dzF; LoadLatLonGrid: the MITgcm arrays placed in the
              interior, halos by topology, the first dzC doubled, hFac scaling
  advection   a uniform zonal (meridional) flow over spacings that vary in y
              (x): the displacement is ΔT u ((1 − yd)/dxC[., yi] + yd/dxC[., yi+1])
              — the divisor of each corner from its own column
  walk        the random walk divides by dxC, dyC at the stale (xi + Hx, yi + Hy)
              of the previous step, exactly
  tracers     Σ Vol·c of every tracer conserved by DST3FL advection under a
              divergence-free transport and by diffusion, with arbitrary
              (MITgcm-like) spacings, areas and hFac; the one-step diffusion of
              a zonally uniform field against κ Ay δc / dyF differences over Vol
  PAR         the Beer–Lambert attenuation with Chl / Vol(i, j, k)
"""
import numpy as np
import pytest

from test_oracle_pins_bcs import ZERO_RATES

R = 6370.0e3


# -------------------------------------------------------------------- grids
def test_lat_lon_grid_shape_and_metrics(pi):
    g = pi.LatLonGrid(size=(360, 160, 10), lat=(-80, 80), lon=(-180, 180), z=(0, -20))
    assert (g.Nx, g.Ny, g.Nz) == (360, 160, 10)
    for a, n in ((g.xC, 364), (g.yC, 164), (g.zC, 14), (g.xF, 364), (g.yF, 164), (g.zF, 14)):
        assert len(a) == n
    assert g.topology == (pi.Periodic, pi.Bounded, pi.Bounded)
    H = 2
    d2r = np.pi / 180
    assert np.all(g.dyC == np.float32(R * 1.0 * d2r)) and np.array_equal(g.dyF, g.dyC)
    exact = R * np.cos(g.yC.astype(np.float64) * d2r) * 1.0 * d2r
    # the reference forms the angle yC·π/180 in Float32: up to ~6e-7 relative near ±80°
    assert np.allclose(g.dxC[0], exact, rtol=1e-6, atol=0)
    assert np.all(g.dxC == g.dxC[0][None, :]) and np.array_equal(g.dxF, g.dxC)
    band = g.Az[H, H:H + 160].astype(np.float64).sum()
    assert abs(band / (R * R * d2r * (np.sin(80 * d2r) - np.sin(-80 * d2r))) - 1) < 1e-5
    assert np.array_equal(g.Ax, (g.dzF[None, None, :] * g.dyF[:, :, None]).astype(np.float32))
    assert np.array_equal(g.Ay, (g.dzF[None, None, :] * g.dxF[:, :, None]).astype(np.float32))
    assert np.array_equal(g.Vol, (g.Az[:, :, None] * g.dzF[None, None, :]).astype(np.float32))
    assert g.Az[-1, -1] == 0                 # the corner halo the reference's copies leave at zero
    assert pi.grids.Ax(1, 1, 1, g) == g.Ax[0, 0, 0] and pi.grids.volume(3, 4, 5, g) == g.Vol[2, 3, 4]
    b = pi.LatLonGrid(size=(10, 8, 2), lat=(-20, 20), lon=(0, 50), z=(0, -10))
    assert b.topology == (pi.Bounded, pi.Bounded, pi.Bounded)
    with pytest.raises(AssertionError):
        pi.LatLonGrid(size=(10, 8, 2), lat=(-90, 90), lon=(0, 50), z=(0, -10))


def _grid_info(Nx, Ny, Nz, seed, dxc_rows=None, dyc_cols=None, rac_scale=1.0):
    """Synthetic MITgcm grid information with distinct spacings, areas and hFac."""
    rng = np.random.default_rng(seed)
    RF = -np.concatenate([[0.0], np.cumsum(5.0 + 10.0 * rng.random(Nz))])
    RC = 0.5 * (RF[:-1] + RF[1:])
    DRF = RF[:-1] - RF[1:]
    DRC = np.concatenate([[RF[0] - RC[0]], RC[:-1] - RC[1:]])
    r2 = lambda: 4e3 + 6e3 * rng.random((Nx, Ny))
    DXC = r2() if dxc_rows is None else np.broadcast_to(np.asarray(dxc_rows)[None, :], (Nx, Ny)).copy()
    DYC = r2() if dyc_cols is None else np.broadcast_to(np.asarray(dyc_cols)[:, None], (Nx, Ny)).copy()
    h = lambda: 0.5 + 0.5 * rng.random((Nx, Ny, Nz))
    return dict(RF=RF, RC=RC, DRF=DRF, DRC=DRC, DXC=DXC, DXG=r2(), DYC=DYC, DYG=r2(),
                RAC=(2e7 + 5e7 * rng.random((Nx, Ny))) * rac_scale, hFacC=h(), hFacW=h(), hFacS=h())


def test_load_lat_lon_grid(pi):
    Nx, Ny, Nz, H = 12, 6, 4, 2
    gi = _grid_info(Nx, Ny, Nz, 1)
    g = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-60, 60), lon=(-180, 180))
    f32 = np.float32
    assert g.topology[0] is pi.Periodic
    assert np.array_equal(g.dxC[H:H + Nx, H:H + Ny], gi["DXC"].astype(f32))
    assert np.array_equal(g.dxF[H:H + Nx, H:H + Ny], gi["DXG"].astype(f32))
    assert np.array_equal(g.dxC[:H], g.dxC[Nx:Nx + H]) and np.array_equal(g.dxC[Nx + H:], g.dxC[H:2 * H])
    assert np.all(g.dyF[:, :H] == g.dyF[:, H:H + 1]) and np.all(g.Az[:, Ny + H:] == g.Az[:, Ny + H - 1:Ny + H])
    assert g.dzC[H] == f32(2) * f32(gi["DRC"][0]) and g.dzF.min() == f32(gi["DRF"].min())
    hW = np.ones(g.padded_shape, f32)
    hW[H:H + Nx, H:H + Ny, H:H + Nz] = gi["hFacW"]
    c = (H + 3, H + 2, H + 1)
    assert g.Ax[c] == (g.dzF[c[2]] * g.dyF[c[0], c[1]]) * hW[c]
    assert g.Vol[c] == (g.dzF[c[2]] * g.Az[c[0], c[1]]) * f32(gi["hFacC"][3, 2, 1])
    assert g.Ax[0, c[1], c[2]] == g.Ax[Nx, c[1], c[2]]        # periodic x halo of hFacW too
    b = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-60, 60), lon=(0, 60))
    assert b.topology[0] is pi.Bounded and np.all(b.dxC[:H] == b.dxC[H:H + 1])


# ------------------------------------------------------------------ helpers
def _model(pi, orc, g, N=0, cap=4, bgc=None, seed=3, init=None):
    kw = {} if init is None else {"tracer_initial": init}
    spec = pi.ModelSpec(g, max_individuals=max(cap, 2 * N), seed=seed, phyto=pi.phyto_setup(None, [N], 1),
                        bgc_params=dict(ZERO_RATES, **(bgc or {})), **kw)
    return spec, orc.OracleModel(spec)


def _inputs(g):
    return (np.zeros((g.Nx, g.Ny), np.float32) + np.float32(500), np.full((g.Nx, g.Ny, g.Nz), 20.0, np.float32))


# --------------------------------------------------------------- particles
@pytest.mark.parametrize("axis", [0, 1])
def test_uniform_flow_over_varying_spacing(pi, orc, axis):
    """A uniform u (v) over a LoadLatLonGrid whose dxC varies with y (dyC with
    x): each individual stays on its row (column) and moves, in step 2 (vel₀ =
    vel₁), by ΔT u ((1 − yd)/dxC[yi] + yd/dxC[yi+1]) with yd, yi of y + ½."""
    Nx, Ny, Nz, H = 10, 8, 1, 2
    rng = np.random.default_rng(4)
    rows = 3e3 + 4e3 * rng.random(Ny)
    cols = 3e3 + 4e3 * rng.random(Nx)
    gi = _grid_info(Nx, Ny, Nz, 2, dxc_rows=rows if axis == 0 else None, dyc_cols=cols if axis == 1 else None)
    g = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-40, 40), lon=(-180, 180))
    spec, om = _model(pi, orc, g, N=300, cap=400)
    U = np.float32(0.7)
    u = np.zeros((Nx, Ny, Nz), np.float32)
    v = np.zeros((Nx, Ny + 1, Nz), np.float32)
    (u if axis == 0 else v)[...] = U
    w = np.zeros((Nx, Ny, Nz + 1), np.float32)
    parf, temp = _inputs(g)
    dt = 400.0
    om.update(dt, 1, vel=(u, v, w), parf=parf, temp=temp)
    x1, y1 = om.field(0, "x").astype(np.float64), om.field(0, "y").astype(np.float64)
    om.update(dt, 1, vel=(u, v, w), parf=parf, temp=temp)
    x2, y2 = om.field(0, "x").astype(np.float64), om.field(0, "y").astype(np.float64)
    ac = om.field(0, "ac")
    if axis == 0:
        s = y1 + 0.5
        i0 = np.floor(s).astype(int) + H             # padded 1-based corner row yi
        d = s - np.floor(s)
        D = g.dxC[H, :].astype(np.float64)           # rows of dxC (no x dependence)
        rate = U * ((1 - d) / D[i0 - 1] + d / D[i0])
        move = np.mod(x2 - x1, Nx)
        assert np.allclose(y2[ac], y1[ac]) and np.allclose(move[ac], (dt * rate)[ac], rtol=0, atol=4e-6)
    else:
        s = x1 + 0.5
        i0 = np.floor(s).astype(int) + H
        d = s - np.floor(s)
        D = g.dyC[:, H].astype(np.float64)
        rate = U * ((1 - d) / D[i0 - 1] + d / D[i0])
        inner = ac & (y1 > 0.2) & (y1 + dt * rate < Ny - 0.2)
        assert np.allclose(x2[ac], x1[ac]) and np.allclose((y2 - y1)[inner], (dt * rate)[inner], rtol=0, atol=4e-6)


def test_random_walk_divides_by_the_stale_columns_spacing(pi, orc):
    """No flow, κhP > 0: step 2 moves x by (2(U − ½)·√(κ ΔT)) / dxC[xi₁ + Hx, yi₁ + Hy]
    (and y with dyC), the stale indices of step 1's find_inds offset by the halo
    again (particle_diffusion.jl:4-6), with U the shared stream's draw."""
    Nx, Ny, Nz, H = 9, 7, 1, 2
    gi = _grid_info(Nx, Ny, Nz, 5)
    g = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-40, 40), lon=(-180, 180))
    kappa = 2.0e4
    spec, om = _model(pi, orc, g, N=200, cap=300, bgc={"κhP": kappa})
    z3 = (np.zeros((Nx, Ny, Nz), np.float32), np.zeros((Nx, Ny + 1, Nz), np.float32),
          np.zeros((Nx, Ny, Nz + 1), np.float32))
    parf, temp = _inputs(g)
    dt = np.float32(60.0)
    om.update(dt, 1, vel=z3, parf=parf, temp=temp)
    x1, y1 = om.field(0, "x"), om.field(0, "y")
    xi, yi = om.field(0, "xi"), om.field(0, "yi")
    ac = om.field(0, "ac")
    om.update(dt, 1, vel=z3, parf=parf, temp=temp)
    x2, y2 = om.field(0, "x"), om.field(0, "y")
    lib = orc.load()
    import ctypes as C
    sq = np.sqrt(np.float32(kappa) * dt, dtype=np.float32)
    f32 = np.float32
    checked = 0
    for s in np.nonzero(ac)[0][:60]:
        out = (C.c_float * 3)()
        lib.orc_uniform3(spec.seed, 2, 0, pi.stream.KIND_DIFFUSION, int(s), out)
        rx = (f32(out[0]) - f32(0.5)) * f32(2)
        ry = (f32(out[1]) - f32(0.5)) * f32(2)
        ci = min(max(int(xi[s]) + H, 1), g.Nx + 2 * H)
        cj = min(max(int(yi[s]) + H, 1), g.Ny + 2 * H)
        ex = x1[s] + (rx * sq) / g.dxC[ci - 1, cj - 1]
        ey = y1[s] + (ry * sq) / g.dyC[ci - 1, cj - 1]
        if not (0 < ex < Nx and 0 < ey < Ny):
            continue                       # the boundary conditions act; covered elsewhere
        assert x2[s] == ex and y2[s] == ey, s
        checked += 1
    assert checked > 30


# ----------------------------------------------------------------- tracers
def _streamfunction_flow(g, seed):
    """u, v with Ax u and Ay v the discrete curl of a corner streamfunction that
    vanishes on the bounded y walls: zero divergence of the transport up to
    the rounding of u = (Ax u) / Ax."""
    rng = np.random.default_rng(seed)
    Nx, Ny, Nz, H = g.Nx, g.Ny, g.Nz, g.Hx
    psi = rng.standard_normal((Nx, Ny + 1)) * 4e3          # transports of ~0.1 m/s
    psi[:, 0] = psi[:, -1] = 0.0
    u = np.zeros((Nx, Ny, Nz)); v = np.zeros((Nx, Ny + 1, Nz))
    for k in range(Nz):
        Ax = g.Ax[H:H + Nx, H:H + Ny, H + k].astype(np.float64)
        Ay = g.Ay[H:H + Nx, H:H + Ny + 1, H + k].astype(np.float64)
        u[:, :, k] = (psi[:, 1:] - psi[:, :-1]) / Ax
        v[:, :, k] = -(np.roll(psi, -1, axis=0) - psi) / Ay
    return u.astype(np.float32), v.astype(np.float32), np.zeros((Nx, Ny, Nz + 1), np.float32)


def test_tracer_mass_conserved_on_latlon(pi, orc):
    """Σ Vol·c of every tracer is conserved (to 2e-6) by DST3FL advection with a
    divergence-free transport and by horizontal and vertical diffusion, on a
    periodic-x, bounded-y LoadLatLonGrid with arbitrary spacings, areas and
    hFac (no individuals, no forcing, no sinking)."""
    Nx, Ny, Nz, H = 12, 8, 3, 2
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 7), size=(Nx, Ny, Nz), lat=(-40, 40),
                          lon=(-180, 180))
    rng = np.random.default_rng(8)
    init = {n: (1.0 + rng.random((Nx, Ny, Nz))).astype(np.float32) for n in pi.TRACER_NAMES}
    spec, om = _model(pi, orc, g, bgc={"κh": 300.0, "κv": 1e-3}, init=init)
    uvw = _streamfunction_flow(g, 9)
    parf, temp = _inputs(g)
    vol = g.Vol[H:H + Nx, H:H + Ny, H:H + Nz].astype(np.float64)
    m0 = [(g.interior(om.tracer(q)).astype(np.float64) * vol).sum() for q in range(14)]
    om.update(2000.0, 4, vel=uvw, parf=parf, temp=temp)
    for q in range(14):
        c = g.interior(om.tracer(q)).astype(np.float64)
        assert abs((c * vol).sum() / m0[q] - 1) < 2e-6, q
        assert not np.array_equal(c, g.interior(spec.initial_tracers[q]))


def test_meridional_diffusion_one_step(pi, orc):
    """A zonally uniform field c(j), no flow, κh > 0, κv = 0: one step changes
    cell (i, j) by ΔT/Vol[i,j]·(κAy[j+1](c[j+1] − c[j])/dyF[j+1] − κAy[j](c[j] − c[j−1])/dyF[j])
    with the bounded walls' halo copies (no flux through them)."""
    Nx, Ny, Nz, H = 6, 7, 1, 2
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 11), size=(Nx, Ny, Nz), lat=(-40, 40),
                          lon=(-180, 180))
    prof = (1.0 + np.arange(Ny) ** 2 * 0.1).astype(np.float32)
    init = {n: np.zeros((Nx, Ny, Nz), np.float32) for n in pi.TRACER_NAMES}
    init["NO3"] = np.broadcast_to(prof[None, :, None], (Nx, Ny, Nz)).astype(np.float32)
    kh = 500.0
    spec, om = _model(pi, orc, g, bgc={"κh": kh, "κv": 0.0}, init=init)
    z3 = (np.zeros((Nx, Ny, Nz), np.float32), np.zeros((Nx, Ny + 1, Nz), np.float32),
          np.zeros((Nx, Ny, Nz + 1), np.float32))
    parf, temp = _inputs(g)
    c0 = spec.initial_tracers[pi._abi.TRACER_ID["NO3"]].astype(np.float64)
    dt = 100.0
    om.update(dt, 1, vel=z3, parf=parf, temp=temp)
    c1 = om.tracer(pi._abi.TRACER_ID["NO3"]).astype(np.float64)
    k = H
    for i in range(H, H + Nx):
        for j in range(H, H + Ny):
            fs = kh * g.Ay[i, j, k] * (c0[i, j, k] - c0[i, j - 1, k]) / g.dyF[i, j]
            fn = kh * g.Ay[i, j + 1, k] * (c0[i, j + 1, k] - c0[i, j, k]) / g.dyF[i, j + 1]
            fw = kh * g.Ax[i, j, k] * (c0[i, j, k] - c0[i - 1, j, k]) / g.dxF[i, j]
            fe = kh * g.Ax[i + 1, j, k] * (c0[i + 1, j, k] - c0[i, j, k]) / g.dxF[i + 1, j]
            exp = c0[i, j, k] + dt / g.Vol[i, j, k] * ((fe - fw) + (fn - fs))
            assert abs(c1[i, j, k] - exp) <= 2e-6 * abs(exp), (i, j)


def test_par_attenuation_with_cell_volumes(pi, orc):
    """calc_par! on a LatLonGrid: a_k = (Chl / Vol[i,j,k] · kc + kw) ΔzF_k per
    column, par_k = PARF (1 − e^−a)/a down the column."""
    Nx, Ny, Nz, H = 6, 5, 3, 2
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 13, rac_scale=1e-8), size=(Nx, Ny, Nz), lat=(-40, 40),
                          lon=(-180, 180))
    spec = pi.ModelSpec(g, max_individuals=900, seed=5, phyto=pi.phyto_setup({"Nsuper": [1e12]}, [600], 1))
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(3)
    parf = (100 + 800 * rng.random((Nx, Ny))).astype(np.float32)
    z3 = (np.zeros((Nx, Ny, Nz), np.float32), np.zeros((Nx, Ny + 1, Nz), np.float32),
          np.zeros((Nx, Ny, Nz + 1), np.float32))
    om.update(60.0, 1, vel=z3, parf=parf, temp=np.full((Nx, Ny, Nz), 20.0, np.float32))
    chl = om.grid_field("Chl").astype(np.float64)
    par = om.grid_field("par").astype(np.float64)
    kc, kw = float(spec.bgc_params["kc"]), float(spec.bgc_params["kw"])
    P = parf.astype(np.float64)
    sl = (slice(H, H + Nx), slice(H, H + Ny))
    assert chl[sl + (slice(H, H + Nz),)].sum() > 0
    # the self-shading term is a sizeable part of the attenuation
    assert np.median(chl[sl + (H,)] / g.Vol[sl + (H,)] * kc) > 0.1 * kw
    for k in range(Nz):
        kk = H + k
        dzF = float(g.dzF[kk])
        a = (chl[sl + (kk,)] / g.Vol[sl + (kk,)].astype(np.float64) * kc + kw) * dzF
        assert np.allclose(par[sl + (kk,)], P * (1 - np.exp(-a)) / a, rtol=2e-6, atol=0)
        P = P * np.exp(-a)


def test_budgets_on_latlon(pi, orc):
    """Total C, N and P (Σ Vol·tracers + the individuals' quotas) conserved to
    1e-6 through uptake, respiration, exudation, grazing, mortality and
    division with no flow on a LoadLatLonGrid: the exchange with the individuals
    is added per cell volume Vol[i, j, k] (plk / volume in apply_tendency!)."""
    Nx, Ny, Nz, H = 8, 6, 2, 2
    g = pi.LoadLatLonGrid(grid_info=_grid_info(Nx, Ny, Nz, 17, rac_scale=1e-5), size=(Nx, Ny, Nz), lat=(-40, 40),
                          lon=(-180, 180))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DIC=2.0, DOC=1.0)
    spec = pi.ModelSpec(g, max_individuals=3000, seed=19, tracer_initial=init,
                        phyto=pi.phyto_setup({"Nsuper": [1e12], "dvid_P": [2e-3], "mort_reg": [1.0],
                                              "grz_P": [1e-4], "VDOCmax": [1e-6], "KsatDOC": [0.2]}, [1000], 1))
    om = orc.OracleModel(spec)
    p = spec.species_params[0]
    vol = g.Vol[H:H + Nx, H:H + Ny, H:H + Nz].astype(np.float64)

    def totals():
        tr = {n: float((g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64) * vol).sum())
              for n in pi.TRACER_NAMES}
        ac = om.field(0, "ac")
        f = {n: om.field(0, n)[ac].astype(np.float64) for n in ("Bm", "Cq", "Nq", "Pq")}
        names = pi._abi.PARAM_NAMES
        R_NC, R_PC = float(p[names.index("R_NC")]), float(p[names.index("R_PC")])
        ind = {"C": float((f["Bm"] + f["Cq"]).sum()), "N": float((f["Nq"] + f["Bm"] * R_NC).sum()),
               "P": float((f["Pq"] + f["Bm"] * R_PC).sum())}
        return {"C": tr["DIC"] + tr["DOC"] + tr["POC"] + ind["C"],
                "N": tr["NH4"] + tr["NO3"] + tr["DON"] + tr["PON"] + ind["N"],
                "P": tr["PO4"] + tr["DOP"] + tr["POP"] + ind["P"]}, ind

    tot0, ind0 = totals()
    z3 = (np.zeros((Nx, Ny, Nz), np.float32), np.zeros((Nx, Ny + 1, Nz), np.float32),
          np.zeros((Nx, Ny, Nz + 1), np.float32))
    om.update(60.0, 30, vel=z3, parf=np.full((Nx, Ny), 900.0, np.float32), temp=np.full((Nx, Ny, Nz), 26.0, np.float32))
    tot1, ind1 = totals()
    assert om.diag(0, "dvid").sum() > 0 and om.diag(0, "graz").sum() > 0
    for e in ("C", "N", "P"):
        assert abs(tot1[e] - tot0[e]) <= 1e-6 * tot0[e], (e, tot0[e], tot1[e])
        assert abs(ind1[e] - ind0[e]) > 1e-4 * tot0[e], (e, ind0[e], ind1[e])


@pytest.mark.parametrize("axis", [0, 1])
def test_unit_courant_shift_on_latlon(pi, orc, axis):
    """DST3FL at a Courant number of one is the exact one-cell upwind shift:
    with u = dxC[j]/ΔT on each row (v = dyC[i]/ΔT on each column) and cell areas
    Az = dyF·dxC (Az = dxF·dyC), so that ΔT·Ax·u / Vol = 1 too, every tracer moves
    by one cell bit for bit.  Spacings, areas and ΔT are powers of two (exact
    products), dxF ≠ dxC and dyF ≠ dyC, so a transport with the wrong area or a
    Courant number with the wrong spacing breaks the shift."""
    Nx, Ny, Nz, H = 8, 6, 2, 2
    rng = np.random.default_rng(21 + axis)
    p2 = lambda lo, hi, shape: 2.0 ** rng.integers(lo, hi, shape)
    gi = _grid_info(Nx, Ny, Nz, 23)
    gi["RF"] = -np.array([0.0, 4.0, 12.0])
    gi["RC"] = 0.5 * (gi["RF"][:-1] + gi["RF"][1:])
    gi["DRF"] = np.array([4.0, 8.0])
    gi["DRC"] = np.array([2.0, 6.0])
    for n in ("hFacC", "hFacW", "hFacS"):
        gi[n] = np.ones((Nx, Ny, Nz))
    if axis == 0:
        gi["DXC"] = np.broadcast_to(p2(11, 14, Ny)[None, :], (Nx, Ny)).copy()
        gi["DXG"] = p2(10, 15, (Nx, Ny))
        gi["DYG"] = np.full((Nx, Ny), 2.0 ** 12)
        gi["DYC"] = p2(10, 15, (Nx, Ny))
        gi["RAC"] = gi["DYG"] * gi["DXC"]
        lon = (-180, 180)
    else:
        gi["DYC"] = np.broadcast_to(p2(11, 14, Nx)[:, None], (Nx, Ny)).copy()
        gi["DYG"] = p2(10, 15, (Nx, Ny))
        gi["DXG"] = np.full((Nx, Ny), 2.0 ** 12)
        gi["DXC"] = p2(10, 15, (Nx, Ny))
        gi["RAC"] = gi["DXG"] * gi["DYC"]
        lon = (0, 60)
    g = pi.LoadLatLonGrid(grid_info=gi, size=(Nx, Ny, Nz), lat=(-40, 40), lon=lon)
    rng2 = np.random.default_rng(5)
    init = {n: (1.0 + rng2.random((Nx, Ny, Nz))).astype(np.float32) for n in pi.TRACER_NAMES}
    spec, om = _model(pi, orc, g, init=init)
    dt = 2.0 ** 8
    u = np.zeros((Nx + axis, Ny, Nz), np.float32)       # x is Bounded for axis 1: Nx + 1 faces
    v = np.zeros((Nx, Ny + 1, Nz), np.float32)
    w = np.zeros((Nx, Ny, Nz + 1), np.float32)
    if axis == 0:
        u[...] = (gi["DXC"][:, :, None] / dt).astype(np.float32)
    else:
        v[:, 1:Ny] = (gi["DYC"][:, 1:Ny, None] / dt).astype(np.float32)   # walls still
    parf, temp = _inputs(g)
    om.update(dt, 1, vel=(u, v, w), parf=parf, temp=temp)
    for q in range(14):
        c0 = g.interior(spec.initial_tracers[q])
        c1 = g.interior(om.tracer(q))
        if axis == 0:
            assert np.array_equal(c1, np.roll(c0, 1, axis=0)), q
        else:       # each row takes the row south of it; the first keeps its value (nothing enters
                    # through the south wall, and the divergence term returns what left)
            assert np.array_equal(c1[:, 1:], c0[:, :-1]) and np.array_equal(c1[:, 0], c0[:, 0]), q
