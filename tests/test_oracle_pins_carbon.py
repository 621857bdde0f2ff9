"""Pins of the oracle's CarbonMode (CPU): an independent float64 transcription
of the CarbonMode growth kernels, the carbon budget, every decision against its
formula, and the division rule's copy list (the daughter keeps the slot's Bd
and age), so that a plausible mistake in the restatement fails here."""
import math

import numpy as np
import pytest

from test_oracle_pins_physics import _still, _shape_dec, _check


def _carbon_spec(pi, g, over, N, cap, seed, init=None, bgc=None):
    return pi.ModelSpec(g, mode=pi.CarbonMode(), max_individuals=cap, seed=seed, tracer_initial=init,
                        phyto=pi.phyto_setup(over, [N], 1), bgc_params=bgc)


def _f64_carbon(q, t, s, dt):
    """plankton_growth! of one step (CarbonMode/growth_kernels.jl:1-120,
    plankton_growth.jl:2-29) for living individuals, in float64 numpy."""
    e30 = 1e-30
    Topt, Tmax, Ea = q["Topt"], q["Tmax"], q["Ea"]
    OGT = np.exp(-Ea / (8.3145 * (Topt + 273.15)))
    kT = np.exp(-Ea / (8.3145 * (t["T"] + 273.15)))
    tf_ps = np.minimum(1.0, np.maximum(0.0, kT * _shape_dec(t["T"] - Topt, Tmax - Topt, 0.1, 4)) / OGT)
    tf = np.minimum(1.0, np.maximum(0.0, kT) / OGT)
    Bm, Bd, Chl = s["Bm"], s["Bd"], s["Chl"]
    th = q["thermal"]
    aI = t["par"] * q["α"] * q["Φ"]
    PCm = q["PCmax"] * (tf * th + tf_ps * (1.0 - th))
    light = 1.0 - np.exp(-aI / np.maximum(e30, PCm) * Chl / np.maximum(e30, Bm))
    PS = PCm * np.maximum(0.0, Bm - Bd) * (light * (1.0 - q["is_bact"]) + q["is_bact"])
    gam = np.where(Bd > 0, Bd / np.maximum(e30, Bm), 0.0)
    gam = np.minimum(1.0, gam) * (1.0 if th == 1.0 else 0.0)
    BS, RP = PS * (1.0 - gam), PS * gam
    TD = (t["T"] - Topt) * q["f_T2B"] * (Bm - Bd) * (Topt < t["T"]) * (Bd < Bm) * (1.0 if th == 1.0 else 0.0)
    TD = np.maximum(0.0, np.minimum(TD, (Bm - Bd) / dt))
    RS = q["respir"] * Bm * tf
    Bm2 = Bm + (BS - RS) * dt
    Bd2 = Bd + (TD - RP) * dt
    out = dict(Bm=Bm2, Bd=Bd2, Sz=Bm2 / (q["Cquota"] * q["Nsuper"]), Chl=Bm2 * q["Chl2C"], age=s["age"] + dt / 3600.0)
    return out, (RS - PS) * dt, dict(PS=PS, RP=RP, TD=TD)


@pytest.mark.parametrize("case", ["thermal", "no_thermal", "bacteria"])
def test_carbon_growth_transcription(pi, orc, case):
    """One CarbonMode step of individuals with damaged biomass (Bd between 0 and
    Bm), dark/dim/bright columns and temperatures on both sides of Topt (thermal
    damage where T > Topt, repair where Bd > 0): state and DIC consumption
    against the float64 transcription to 1e-4."""
    g = pi.RectilinearGrid(size=(4, 4, 3), x=(0, 4.0), y=(0, 4.0), z=[0.0, -4.0, -10.0, -18.0])
    rng = np.random.default_rng(43)
    init = {n: np.zeros((4, 4, 3), np.float32) for n in pi.TRACER_NAMES}
    init["DIC"] = np.full((4, 4, 3), 20.0, np.float32)
    bgc = {k: 0.0 for k in ("kDOC", "Nit", "kDON", "kDOP", "kPOC", "kPON", "kPOP")}
    over = {"Nsuper": [1e12], "grz_P": [0.0], "mort_P": [0.0], "dvid_P": [0.0], "f_T2B": [2e-4],
            "thermal": [0.0 if case == "no_thermal" else 1.0], "is_bact": [1.0 if case == "bacteria" else 0.0]}
    spec = _carbon_spec(pi, g, over, 400, 500, 41, init, bgc)
    ini = spec.initial_individuals[0]
    n = 400
    ini["Cq"][:n] = (ini["Bm"][:n] * rng.random(n) * np.where(rng.random(n) < 0.3, 0.0, 0.8)).astype(np.float32)
    om = orc.OracleModel(spec)
    parf = np.zeros((4, 4))
    parf[1] = (0.15, 0.5, 1.5, 5.0)
    parf[2:] = 1500.0
    temp = (22.0 + 7.5 * rng.random((4, 4, 3))).astype(np.float32)     # Topt 27, Tmax 30
    dt = 600.0
    om.update(dt, 1, vel=_still(g), parf=parf.astype(np.float32), temp=temp)
    ac = om.field(0, "ac")
    sl = np.nonzero(ac)[0]
    assert len(sl) == n
    xi, yi, zi = (om.field(0, f)[sl] - 1 for f in ("xi", "yi", "zi"))
    H = g.Hx
    par = om.grid_field("par").astype(np.float64)
    t = {"par": par[xi, yi, zi], "T": temp[xi - H, yi - H, zi - H].astype(np.float64)}
    q = dict(zip(pi._abi.CARBON_PARAM_NAMES, spec.species_params[0].astype(np.float64)))
    s = {"Bm": ini["Bm"][sl].astype(np.float64), "Bd": ini["Cq"][sl].astype(np.float64),
         "Chl": ini["Chl"][sl].astype(np.float64), "age": ini["age"][sl].astype(np.float64)}
    exp_state, dic, rates = _f64_carbon(q, t, s, dt)
    if case != "no_thermal":
        assert (rates["TD"] > 0).sum() > 20 and (rates["RP"] > 0).sum() > 20
    assert (rates["PS"] > 0).sum() > 50
    if case != "bacteria":                 # bacteria photosynthesise without light
        assert (rates["PS"] == 0).sum() > 20
    for k, v in exp_state.items():
        got = om.field(0, "Cq" if k == "Bd" else k)[sl].astype(np.float64)
        ref = s["Bm"] if k == "Bd" else np.abs(v)
        assert np.all(np.abs(got - v) <= 1e-4 * ref + 1e-30), k
    vol = (g.Δx * g.Δy * g.dzF.astype(np.float64))[None, None, :]
    d = np.zeros(g.padded_shape)
    np.add.at(d, (xi, yi, zi), dic)
    I = (slice(H, H + 4), slice(H, H + 4), slice(g.Hz, g.Hz + 3))
    before = spec.initial_tracers[pi._abi.TRACER_ID["DIC"]].astype(np.float64)
    after = om.tracer(pi._abi.TRACER_ID["DIC"]).astype(np.float64)
    assert np.allclose(after[I], before[I] + (d / vol)[I], rtol=1e-6, atol=1e-4 * np.abs((d / vol)[I]).max())
    assert np.abs((d / vol)[I]).max() > 1e-4
    for other in ("NH4", "NO3", "PO4", "DOC"):      # CarbonMode exchanges only DIC
        assert not np.any(om.tracer(pi._abi.TRACER_ID[other]))


def test_carbon_budget_without_thermal_damage(pi, orc):
    """thermal = 0: no repair or damage, so DIC + DOC + POC + ΣBm is conserved
    through photosynthesis, respiration, grazing and mortality losses (DOC and
    POC, consume_loss.jl:11-21) and division (biomass halved between parent and
    daughter); a dropped or mis-signed term fails."""
    g = pi.RectilinearGrid(size=(16, 16, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DIC=2.0, DOC=1.0)
    over = {"Nsuper": [1e12], "dvid_P": [2e-3], "dvid_reg": [1.0], "mort_P": [1e-3], "mort_reg": [1.0],
            "grz_P": [1e-4], "thermal": [0.0]}
    spec = _carbon_spec(pi, g, over, 1024, 4096, 19, init, {"kDOC": 0.0, "kPOC": 0.0})
    om = orc.OracleModel(spec)
    vol = float(g.Δx) * float(g.Δy) * float(g.dzF[g.Hz])

    def total():
        tr = sum(float(g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64).sum()) * vol
                 for n in ("DIC", "DOC", "POC"))
        ac = om.field(0, "ac")
        return tr + float(om.field(0, "Bm")[ac].astype(np.float64).sum()), float(om.field(0, "Bm")[ac].sum())

    t0, b0 = total()
    om.update(60.0, 30, vel=_still(g), parf=np.full((16, 16), 900.0, np.float32),
              temp=np.full((16, 16, 1), 26.0, np.float32))
    t1, b1 = total()
    assert om.diag(0, "dvid").sum() > 0 and om.diag(0, "graz").sum() > 0 and om.diag(0, "mort").sum() > 0
    assert abs(t1 - t0) <= 1e-6 * t0, (t0, t1)
    assert abs(b1 - b0) > 1e-4 * t0


def _carbon_decisions(pi, orc, over, bd_frac, dt=60.0, N=8000):
    g = pi.RectilinearGrid(size=(6, 6, 1), x=(0, 6e3), y=(0, 6e3), z=(0, -10))
    spec = _carbon_spec(pi, g, over, N, N + 100, 23)
    ini = spec.initial_individuals[0]
    ini["Cq"][:N] = (ini["Bm"][:N] * bd_frac(N)).astype(np.float32)
    om = orc.OracleModel(spec)
    om.update(dt, 1, vel=_still(g), parf=np.full((6, 6), 500.0, np.float32), temp=np.full((6, 6, 1), 24.0, np.float32))
    ac = om.field(0, "ac")
    xi, yi, zi = (om.field(0, f) for f in ("xi", "yi", "zi"))
    cnt = np.zeros(g.padded_shape)
    np.add.at(cnt, (xi[ac] - 1, yi[ac] - 1, zi[ac] - 1), 1.0)
    slots = np.nonzero(ac)[0]
    u = np.stack(pi.stream.uniform3(spec.seed, 1, 0, pi.stream.KIND_DECISION, slots)).astype(np.float64)
    q = dict(zip(pi._abi.CARBON_PARAM_NAMES, spec.species_params[0].astype(np.float64)))
    st = {k: om.field(0, k)[slots].astype(np.float64) for k in ("Sz", "Bm", "Cq")}
    dec = {k: om.field(0, k)[slots] for k in ("graz", "mort", "dvid")}
    return cnt[xi[slots] - 1, yi[slots] - 1, zi[slots] - 1], u, q, st, dec


def test_carbon_decisions_thermal_mortality(pi, orc):
    """thermal = 1: mortality P = mort_P·shape_func_inc(Bd/max(1e-30, Bm), 0.99,
    1e-6) (division_death_probability.jl:126-136) with Bd/Bm spread over
    0.85-1.0; grazing P = pop·grz_P; Sizer division (pow 2)."""
    rng = np.random.default_rng(5)
    over = {"grz_P": [3e-6], "mort_P": [3e-3], "dvid_P": [2e-3], "dvid_reg": [1.3], "thermal": [1.0]}
    pop, u, q, st, dec = _carbon_decisions(pi, orc, over, lambda n: 0.85 + 0.15 * rng.random(n))
    dt = 60.0
    x = st["Cq"] / np.maximum(1e-30, st["Bm"])
    _check(dec["mort"], u[1], q["mort_P"] * (1 - _shape_dec(x, 0.99, 1e-6, 4)) * dt, "thermal mort")
    _check(dec["graz"], u[0], pop * q["grz_P"] * dt, "graz")
    _check(dec["dvid"], u[2], q["dvid_P"] * (1 - _shape_dec(st["Sz"], q["dvid_reg"], 2e-3, 2)) * dt, "dvid")


def test_carbon_decisions_size_mortality(pi, orc):
    """thermal = 0: mortality P = mort_P·shape_func_dec(Sz, mort_reg, 1e-5)."""
    over = {"mort_P": [3e-3], "mort_reg": [1.35], "thermal": [0.0], "dvid_P": [1e-4]}
    pop, u, q, st, dec = _carbon_decisions(pi, orc, over, lambda n: np.zeros(n))
    _check(dec["mort"], u[1], q["mort_P"] * _shape_dec(st["Sz"], q["mort_reg"], 1e-5, 4) * 60.0, "size mort")


def test_carbon_division_copies(pi, orc):
    """divide! (CarbonMode/division_death.jl:28-70) on a tiny array, brute force:
    the k-th divider fills the k-th inactive slot; the daughter copies x, y, z,
    Sz, Bm, Chl, gen (and ac) but keeps the slot's own Bd and age; parent and
    daughter halve Sz, Bm, Bd, Chl, add one generation and zero the age.  In the
    dark, below Topt and without respiration, growth then only recomputes Sz,
    Chl and adds ΔT/3600 to the age."""
    g = pi.RectilinearGrid(size=(2, 2, 1), x=(0, 2e3), y=(0, 2e3), z=(0, -10))
    over = {"respir": [0.0], "mort_P": [0.0], "dvid_P": [0.0], "grz_P": [0.0]}
    spec = _carbon_spec(pi, g, over, 6, 12, 3)
    ini = spec.initial_individuals[0]
    # slots 0..5 alive, 6..11 inactive with leftovers of dead individuals
    ini["Cq"][:6] = (ini["Bm"][:6] * np.float32(0.25)).astype(np.float32)
    ini["Cq"][6:] = np.arange(1, 7, dtype=np.float32) * np.float32(1e-12)
    ini["age"][:] = np.arange(12, dtype=np.float32) + np.float32(5.0)
    ini["gen"][:6] = np.arange(1, 7, dtype=np.float32)
    ini["dvid"][:] = 0
    ini["dvid"][[1, 4]] = 1.0            # dividers (the previous step's decision)
    ini["graz"][:] = 0
    ini["mort"][:] = 0
    before = {k: ini[k].copy() for k in ini}
    om = orc.OracleModel(spec)
    dt = 60.0
    om.update(dt, 1, vel=_still(g), parf=np.zeros((2, 2), np.float32), temp=np.full((2, 2, 1), 10.0, np.float32))
    q = dict(zip(pi._abi.CARBON_PARAM_NAMES, spec.species_params[0]))
    f32 = np.float32
    h = f32(0.5)
    age_add = f32(f32(dt) / f32(3600.0))
    got = {k: om.field(0, k) for k in ("x", "y", "z", "Bm", "Cq", "Chl", "Sz", "gen", "age", "ac")}
    assert got["ac"][:8].all() and not got["ac"][8:].any()
    for parent, daughter in ((1, 6), (4, 7)):
        for i, src_bd, src_age in ((parent, before["Cq"][parent], before["age"][parent]),
                                   (daughter, before["Cq"][daughter], before["age"][daughter])):
            assert got["x"][i] == before["x"][parent] and got["y"][i] == before["y"][parent]
            Bm = f32(before["Bm"][parent] * h)
            assert got["Bm"][i] == Bm
            assert got["Cq"][i] == f32(src_bd * h)
            assert got["gen"][i] == before["gen"][parent] + f32(1)
            assert got["age"][i] == f32(f32(src_age * f32(0.0)) + age_add)
            assert got["Chl"][i] == f32(Bm * q["Chl2C"]) and got["Sz"][i] == f32(Bm / f32(q["Cquota"] * q["Nsuper"]))
    for i in (0, 2, 3, 5):               # the others: unchanged but for age, Sz and Chl
        assert got["Bm"][i] == before["Bm"][i] and got["Cq"][i] == before["Cq"][i]
        assert got["age"][i] == f32(before["age"][i] + age_add)
