"""Pins of the oracle's IronEnergyMode (Individuals/Plankton/IronEnergyMode) and of
the host generator of its initial state.

  initial state  CH, qNO3, qNH4, qP and the iron total below a tenth of their
                 maxima times Bm (+ CH), the iron split 0.4 / 0.3·is_nr /
                 0.3·(is_croc + is_tric) into the photosynthesis, nitrate
                 reduction and N-fixation pools (plankton_generation.jl:48-94)
  budgets        N (NH4+NO3+DON+PON+Σ(qNH4+qNO3+Bm·R_NC)), P and Fe (with no
                 dust) conserved through uptake, nitrate reduction, biosynthesis,
                 grazing, mortality, division, the Fe cycle and sinking; C
                 conserved in the dark; in the light DIC loses ΔT·Σ PS, the
                 light reaction's energy, not the carbon fixed
                 (consume_loss.jl:5 against growth_kernels.jl:88-96)
  iron uptake    ΔqFe = ΔT·KSAFe·SA·Sz^(2/3)·DFe·regQFe·Nsuper at T = Topt
  iron pools     rising light moves k_Fe_ST2PS·qFe from storage to photosynthesis;
                 at sunset (par = 0, dpar < 0) the photosynthesis pool returns at
                 twice k_Fe_PS2ST (the Int sum of two Bools); in the dark tdark
                 counts ΔT per step and light resets it
  dead slots     an individual that died keeps moving its iron to storage
                 and counting tdark (par = 0 for it), and exchanges nothing
"""
import math

import numpy as np
import pytest


def _still(g):
    return (np.zeros((g.Nx, g.Ny, g.Nz), np.float32), np.zeros((g.Nx, g.Ny, g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny, g.Nz + 1), np.float32))


def _spec(pi, g, over, N, cap, seed=5, bgc=None, init=None):
    return pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=cap, seed=seed, tracer_initial=init,
                        phyto=pi.phyto_setup(over, [N], 1), bgc_params=bgc or {})


def _q(pi, spec):
    return dict(zip(pi._abi.IRON_PARAM_NAMES, spec.species_params[0].astype(np.float64)))


FE = ("qFe", "qFePS", "qFeNR", "qFeNF")


def test_initial_state(pi):
    g = pi.RectilinearGrid(size=(8, 8, 4), x=(0, 8e3), y=(0, 8e3), z=(0, -40))
    spec = _spec(pi, g, {"is_nr": [1.0], "is_croc": [1.0]}, 20000, 20100)
    q = _q(pi, spec)
    d = spec.initial_individuals[0]
    ac = d["ac"]
    f = {k: d[pi._abi.FIELD_ALIAS.get(k, k)][ac].astype(np.float64) for k in ("Bm", "CH", "qNO3", "qNH4", "qP") + FE}
    BC = f["Bm"] + f["CH"]
    Fe = f["qFe"] + f["qFePS"] + f["qFeNR"] + f["qFeNF"]
    for r, hi in ((f["CH"] / f["Bm"], q["CHmax"]), (f["qNO3"] / BC, q["qNO3max"]), (f["qNH4"] / BC, q["qNH4max"]),
                  (f["qP"] / BC, q["qPmax"]), (Fe / BC, q["qFemax"])):
        hi = 0.1 * hi
        assert r.min() >= 0 and r.max() <= hi * (1 + 1e-6) and r.max() > 0.98 * hi and r.min() < 0.02 * hi
    assert np.allclose(f["qFePS"], 0.4 * Fe, rtol=1e-5, atol=0)
    assert np.allclose(f["qFeNR"], 0.3 * Fe, rtol=1e-5, atol=0)
    assert np.allclose(f["qFeNF"], 0.3 * Fe, rtol=1e-5, atol=0)
    assert np.allclose(f["qFe"], 0.0, atol=1e-6 * Fe.max())           # 1 − 0.4 − 0.3 − 0.3
    for k in ("ex5", "ex6", "ex7"):
        assert not d[k].any()
    assert np.allclose(d["Chl"][ac], d["Bm"][ac] * np.float32(q["Chl2Cint"]), rtol=1e-6, atol=0)


def _totals(pi, om, g):
    vol = g.Δx * g.Δy * g.dzF[g.Hz:g.Hz + g.Nz].astype(np.float64)
    tr = {n: float((g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64).sum(axis=(0, 1)) * vol).sum())
          for n in pi.TRACER_NAMES}
    ac = om.field(0, "ac")
    f = {n: om.field(0, n)[ac].astype(np.float64) for n in ("Bm", "CH", "qNO3", "qNH4", "qP") + FE}
    return tr, f


def _budget(tr, f, q):
    return {"C": tr["DIC"] + tr["DOC"] + tr["POC"] + (f["Bm"] + f["CH"]).sum(),
            "N": tr["NH4"] + tr["NO3"] + tr["DON"] + tr["PON"] + (f["qNH4"] + f["qNO3"] + f["Bm"] * q["R_NC"]).sum(),
            "P": tr["PO4"] + tr["DOP"] + tr["POP"] + (f["qP"] + f["Bm"] * q["R_PC"]).sum(),
            "Fe": tr["DFe"] + tr["PFe_inorg"] + tr["PFe_bio"] + sum(f[k].sum() for k in FE)}


BUDGET_OVER = {"Nsuper": [1e12], "dvid_P": [5e-4], "dvid_reg": [1.0], "grz_P": [1e-4], "mort_P": [2e-3], "mort_reg": [3.0],
               "k_rs": [5e-5], "k_nr": [5e-4], "KSAFe": [2.77e-12]}


@pytest.mark.parametrize("light", [False, True])
def test_budgets(pi, orc, light):
    """N, P, Fe conserved to 1e-6 (and C in the dark) with uptake, nitrate
    reduction, biosynthesis, grazing, mortality and division on a slab with no
    flow (the light case without the C balance: DIC pays the light reaction's
    energy, the next test)."""
    g = pi.RectilinearGrid(size=(12, 12, 1), x=(0, 24), y=(0, 24), z=(0, -24))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DIC=2.0, DOC=1.0, DFe=5e-3, PFe_inorg=1e-3)
    spec = _spec(pi, g, BUDGET_OVER, 600, 2400, seed=19, init=init)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    b0 = _budget(*_totals(pi, om, g), q)
    parf = np.full((12, 12), 900.0 if light else 0.0, np.float32)
    n, dt = 20, 60.0
    om.update(dt, n, vel=_still(g), parf=parf, temp=np.full((12, 12, 1), 25.0, np.float32))
    tr1, f1 = _totals(pi, om, g)
    b1 = _budget(tr1, f1, q)
    assert om.diag(0, "dvid").sum() > 5 and om.diag(0, "graz").sum() > 5 and om.diag(0, "mort").sum() > 5
    for e in ("N", "P", "Fe"):
        assert abs(b1[e] - b0[e]) <= 1e-6 * b0[e], (e, b0[e], b1[e])
    if not light:                                   # in the light DIC pays PS (next test)
        assert abs(b1["C"] - b0["C"]) <= 1e-6 * b0["C"], (b0["C"], b1["C"])


def test_dic_pays_for_the_light_reaction(pi, orc):
    """One step in the light with respiration, nitrate reduction and losses off:
    ΔDIC (amount) = −ΔT·Σ PS with PS = PCmax·(1 − e^{−par·α·Chl/Bm})·Bm·Ksat(qFePS)
    from the state before the step, while CH gains ΔT·CF."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4), y=(0, 4), z=(0, -4))
    zero_rm = {k: 0.0 for k in ("kDOC", "kDON", "kDOP", "kPOC", "kPON", "kPOP", "Nit")}
    over = {"Nsuper": [1e12], "k_rs": [0.0], "k_nr": [0.0], "mort_P": [0.0], "dvid_P": [0.0], "PCmax": [7.8e-6]}
    spec = _spec(pi, g, over, 40, 64, seed=3, bgc=zero_rm)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    ac = om.field(0, "ac")
    Bm, CH, Chl, qFePS = (om.field(0, k)[ac].astype(np.float64) for k in ("Bm", "CH", "Chl", "qFePS"))
    tr0, _ = _totals(pi, om, g)
    om.update(60.0, 1, vel=_still(g), parf=np.full((4, 4), 500.0, np.float32), temp=np.full((4, 4, 1), 25.0, np.float32))
    par = om.grid_field("par")[om.field(0, "xi")[ac] - 1, om.field(0, "yi")[ac] - 1, om.field(0, "zi")[ac] - 1]
    Qps = qFePS / (Bm + CH)
    PS = q["PCmax"] * (1 - np.exp(-par * q["α"] * Chl / Bm)) * Bm * (Qps / (Qps + q["KfePS"]))
    tr1, _ = _totals(pi, om, g)
    vol = float(g.Δx * g.Δy * g.dzF[g.Hz])
    assert PS.min() > 0
    assert math.isclose(tr1["DIC"] - tr0["DIC"], -60.0 * PS.sum(), rel_tol=1e-4)


def test_iron_uptake_surface_area(pi, orc):
    """T = Topt, internal iron transfers off, plenty of DFe: one step raises the
    storage pool by ΔT·KSAFe·SA·Sz^(2/3)·DFe·regQFe·Nsuper, regQFe the square
    shape function of the iron quota."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    over = {"KSAFe": [2.77e-15], "mort_P": [0.0], "dvid_P": [0.0]}
    for k in ("k_Fe_ST2PS", "k_Fe_PS2ST", "k_Fe_ST2NR", "k_Fe_NR2ST", "k_Fe_ST2NF", "k_Fe_NF2ST"):
        over[k] = [0.0]
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DFe=1e-3)
    init["rand_noise"] = {k: 0.0 for k in init["rand_noise"]}
    spec = _spec(pi, g, over, 60, 64, seed=4, init=init)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    ac = om.field(0, "ac")
    Sz, Bm, CH = (om.field(0, k)[ac].astype(np.float64) for k in ("Sz", "Bm", "CH"))
    Fe0 = sum(om.field(0, k)[ac].astype(np.float64) for k in FE)
    qFe0 = om.field(0, "qFe")[ac].astype(np.float64)
    om.update(60.0, 1, vel=_still(g), parf=np.full((4, 4), 300.0, np.float32),
              temp=np.full((4, 4, 1), np.float32(q["Topt"]), np.float32))
    fx = np.clip(1 - (Fe0 / (Bm + CH)) / q["qFemax"], 0, 1)
    reg = fx ** 2 / (1e-4 + fx ** 2)
    want = 60.0 * q["KSAFe"] * q["SA"] * Sz ** (2 / 3) * np.float32(1e-3) * reg * q["Nsuper"]
    got = om.field(0, "qFe")[ac].astype(np.float64) - qFe0
    assert np.allclose(got, want, rtol=1e-4, atol=0) and want.min() > 0


def _iron_model(pi, orc, over=None, N=30):
    g = pi.RectilinearGrid(size=(3, 3, 1), x=(0, 3e3), y=(0, 3e3), z=(0, -10))
    o = {"KSAFe": [0.0], "mort_P": [0.0], "dvid_P": [0.0], "k_Fe_ST2PS": [2e-3], "k_Fe_PS2ST": [1e-3],
         "is_croc": [0.0], "is_tric": [0.0]}
    o.update(over or {})
    spec = _spec(pi, g, o, N, N + 10, seed=6)
    return g, spec, _q(pi, spec), orc.OracleModel(spec)


def test_iron_pools_with_light(pi, orc):
    """Rising light (dpar > 0): qFePS gains ΔT·k_Fe_ST2PS·qFe; at sunset (par = 0,
    dpar < 0) qFePS loses 2·ΔT·k_Fe_PS2ST·qFePS; with constant light nothing
    moves between storage and photosynthesis; tdark counts ΔT per dark step
    and light resets it."""
    g, spec, q, om = _iron_model(pi, orc)
    n = 30
    still, temp = _still(g), np.full((3, 3, 1), 20.0, np.float32)
    get = lambda k: om.field(0, k)[:n].astype(np.float64)
    om.update(60.0, 1, vel=still, parf=np.full((3, 3), 400.0, np.float32), temp=temp)    # par rises from 0
    qFe1, ps1 = get("qFe"), get("qFePS")
    om.update(60.0, 1, vel=still, parf=np.full((3, 3), 400.0, np.float32), temp=temp)    # steady light
    assert np.array_equal(get("qFePS"), ps1)
    assert not get("tdark").any()
    ps2, qFe2 = get("qFePS"), get("qFe")
    om.update(60.0, 1, vel=still, parf=np.zeros((3, 3), np.float32), temp=temp)          # sunset
    assert np.allclose(get("qFePS"), ps2 - 60.0 * 2 * q["k_Fe_PS2ST"] * ps2, rtol=1e-5, atol=0)
    assert np.all(get("tdark") == 60.0)
    om.update(60.0, 1, vel=still, parf=np.zeros((3, 3), np.float32), temp=temp)          # dark, par0 = 0
    ps3 = get("qFePS")
    om.update(45.0, 1, vel=still, parf=np.zeros((3, 3), np.float32), temp=temp)
    assert np.allclose(get("qFePS"), ps3 - 45.0 * q["k_Fe_PS2ST"] * ps3, rtol=1e-5, atol=0)
    assert np.all(get("tdark") == 60.0 + 60.0 + 45.0)
    om.update(60.0, 1, vel=still, parf=np.full((3, 3), 400.0, np.float32), temp=temp)    # sunrise
    assert not get("tdark").any()
    # the first step's rise: storage to photosynthesis at k_Fe_ST2PS
    g2, spec2, q2, om2 = _iron_model(pi, orc)
    get2 = lambda k: om2.field(0, k)[:n].astype(np.float64)
    qFe0, ps0 = get2("qFe"), get2("qFePS")
    om2.update(60.0, 1, vel=still, parf=np.full((3, 3), 400.0, np.float32), temp=temp)
    moved = np.minimum(60.0 * q2["k_Fe_ST2PS"] * qFe0, qFe0)
    assert np.allclose(get2("qFePS") - ps0, moved, rtol=1e-4, atol=1e-6 * ps0.max()) and moved.max() > 0


def test_dead_slots_move_iron_and_count_dark(pi, orc):
    """Every individual dies in step 2; in steps 3-4 (bright cells) the dead slots
    see par = 0: qFePS returns to storage at k_Fe_PS2ST, tdark counts ΔT, the
    iron total stays, and no tracer changes."""
    zero_rm = {k: 0.0 for k in ("kDOC", "kDON", "kDOP", "kPOC", "kPON", "kPOP", "Nit", "lambda_min", "lambda_dust",
                                "lambda_Fe", "kdiss", "lambda_POC", "w_sink_org", "w_sink_inorg")}
    g = pi.RectilinearGrid(size=(3, 3, 1), x=(0, 3e3), y=(0, 3e3), z=(0, -10))
    o = {"KSAFe": [0.0], "mort_P": [1.0], "mort_reg": [1e3], "dvid_P": [0.0], "k_Fe_PS2ST": [1e-3]}
    spec = _spec(pi, g, o, 30, 40, seed=6, bgc=zero_rm)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    still, temp, parf = _still(g), np.full((3, 3, 1), 20.0, np.float32), np.full((3, 3), 400.0, np.float32)
    om.update(60.0, 2, vel=still, parf=parf, temp=temp)
    assert not om.field(0, "ac").any()
    get = lambda k: om.field(0, k)[:30].astype(np.float64)
    ps, td, Fe = get("qFePS"), get("tdark"), sum(get(k) for k in FE)
    tr0 = [om.tracer(t).copy() for t in range(14)]
    om.update(60.0, 1, vel=still, parf=parf, temp=temp)
    assert np.allclose(get("qFePS"), ps - 60.0 * q["k_Fe_PS2ST"] * ps, rtol=1e-5, atol=0) and ps.min() > 0
    assert np.array_equal(get("tdark"), td + 60.0)
    assert np.allclose(sum(get(k) for k in FE), Fe, rtol=1e-6, atol=0)
    for t in range(14):
        assert np.array_equal(om.tracer(t), tr0[t]), t
