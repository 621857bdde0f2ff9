"""Pins of the oracle's MacroMolecularMode (Individuals/Plankton/MacroMolecularMode,
Plankton/division_death_probability.jl:39-44,96-101) and of the host generator
of its initial state.

  initial state  DNA, RNA, protein, CH, NST and PST in their documented ranges
                 and ratios to DNA and CH (plankton_generation.jl:31-78)
  budgets        C (DIC+DOC+POC+Σ total_C_biomass), N and P conserved with
                 light, uptake, respiration, synthesis, grazing, mortality and
                 division; exudation moves ΔT·exu into DOC while CH loses exu
                 (consume_loss.jl:5 against growth_kernels.jl:190-200)
  respiration    beyond CH it takes protein and returns its N to NST
  synthesis      with saturating limits at T = Topt one step changes protein
                 by ΔT·k_pro·RNA, RNA by ΔT·k_rna·PRO, DNA by ΔT·k_dna·PRO
                 (none once DNA ≥ 2·C_DNA·Nsuper), and CH, NST, PST by the
                 C, N and P those syntheses use
  decisions      division only above 2 C_DNA·Nsuper of DNA at P = dvid_P, mortality
                 by age at mort_P·shape_func_inc(age, mort_reg, 1e-4), each
                 decision U < P·ΔT with U the shared stream's draw
  dead slots     an individual that died keeps synthesising: the next step
                 changes its protein by ΔT·k_pro·RNA·tempFunc(0) (find_NPT!
                 gathers T·ac = 0 for it), and exchanges nothing with the tracers
"""
import math

import numpy as np
import pytest


def _still(g):
    return (np.zeros((g.Nx, g.Ny, g.Nz), np.float32), np.zeros((g.Nx, g.Ny, g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny, g.Nz + 1), np.float32))


def _spec(pi, g, over, N, cap, seed=5, bgc=None, init=None):
    return pi.ModelSpec(g, mode=pi.MacroMolecularMode(), max_individuals=cap, seed=seed, tracer_initial=init,
                        phyto=pi.phyto_setup(over, [N], 1), bgc_params=bgc or {})


def _q(pi, spec):
    return dict(zip(pi._abi.MM_PARAM_NAMES, spec.species_params[0].astype(np.float64)))


def test_initial_state(pi):
    """initialize_plankton! of MacroMolecularMode: DNA/(C_DNA·Nsuper) in [1, 1+var),
    RNA, protein, CH, NST and PST each 1 ± var times their ratio to DNA (RNA2DNA,
    PRO2DNA, CH2DNA) or to CH (16/106, 1/106), Chl = DNA·Chl2DNA·893.49/55, the
    draws spread over the whole range; inactive slots all zero."""
    g = pi.RectilinearGrid(size=(8, 8, 4), x=(0, 8e3), y=(0, 8e3), z=(0, -40))
    spec = _spec(pi, g, None, 20000, 20100)
    q = _q(pi, spec)
    d = spec.initial_individuals[0]
    ac = d["ac"]
    assert ac[:20000].all() and not ac[20000:].any()
    unit = q["C_DNA"] * q["Nsuper"]
    DNA, RNA, PRO = (d[k][ac].astype(np.float64) for k in ("Sz", "ex0", "Bm"))
    CH, NST, PST, Chl = (d[k][ac].astype(np.float64) for k in ("Cq", "Nq", "Pq", "Chl"))
    v = q["var"]
    for r, lo, hi in ((DNA / unit, 1, 1 + v), (RNA / (unit * q["RNA2DNA"]), 1 - v, 1 + v),
                      (PRO / (unit * q["PRO2DNA"]), 1 - v, 1 + v), (CH / (DNA * q["CH2DNA"]), 1 - v, 1 + v),
                      (NST / (CH * 16 / 106), 1 - v, 1 + v), (PST / (CH / 106), 1 - v, 1 + v)):
        assert r.min() >= lo - 1e-6 and r.max() <= hi + 1e-6
        assert r.min() < lo + 0.01 * (hi - lo) and r.max() > hi - 0.01 * (hi - lo)
    assert np.allclose(Chl, DNA * q["Chl2DNA"] * 893.49 / 55, rtol=1e-6, atol=0)
    for k in ("Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "ex0", "gen", "x"):
        assert not d[k][~ac].any(), k
    # independent draws: the ratios are uncorrelated
    assert abs(np.corrcoef(RNA, PRO)[0, 1]) < 0.05 and abs(np.corrcoef(DNA, RNA)[0, 1]) < 0.05


def _totals(pi, om, g, q):
    vol = g.Δx * g.Δy * g.dzF[g.Hz:g.Hz + g.Nz].astype(np.float64)
    tr = {n: float((g.interior(om.tracer(pi._abi.TRACER_ID[n])).astype(np.float64).sum(axis=(0, 1)) * vol).sum())
          for n in pi.TRACER_NAMES}
    ac = om.field(0, "ac")
    f = {n: om.field(0, n)[ac].astype(np.float64) for n in ("PRO", "DNA", "RNA", "CH", "NST", "PST", "Chl")}
    ind = {"C": float((f["PRO"] + f["DNA"] + f["RNA"] + f["CH"] + f["Chl"] / 893.49 * 55).sum()),
           "N": float((f["PRO"] * q["R_NC_PRO"] + f["DNA"] * q["R_NC_DNA"] + f["RNA"] * q["R_NC_RNA"] + f["NST"] +
                       f["Chl"] / 893.49 * 4).sum()),
           "P": float((f["DNA"] * q["R_PC_DNA"] + f["RNA"] * q["R_PC_RNA"] + f["PST"]).sum())}
    tot = {"C": tr["DIC"] + tr["DOC"] + tr["POC"] + ind["C"],
           "N": tr["NH4"] + tr["NO3"] + tr["DON"] + tr["PON"] + ind["N"],
           "P": tr["PO4"] + tr["DOP"] + tr["POP"] + ind["P"]}
    return tot, ind, tr


def test_budgets(pi, orc):
    """Light, uptake, respiration, protein/DNA/RNA/Chl synthesis, grazing, mortality
    (by age) and division on a 16x16 slab, no flow: total C, N and P are each
    conserved to 1e-6; the individuals exchange at least 100x that with the
    tracers (super-individuals of 1e12 cells), so a dropped or mis-signed term,
    a wrong stoichiometric ratio or a wrong pool in the losses fails.  CHmax = 1
    keeps the exudation at zero (the next test covers it)."""
    g = pi.RectilinearGrid(size=(16, 16, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    init = pi.fields.default_tracer_init()
    init["initial_condition"].update(DIC=2.0, DOC=1.0)
    spec = _spec(pi, g, {"Nsuper": [1e12], "k_dna": [2e-4], "dvid_P": [2e-3], "grz_P": [1e-4], "mort_P": [2e-3],
                         "mort_reg": [0.2], "CHmax": [1.0], "VDOCmax": [1e-6], "KsatDOC": [0.2]},
                 1024, 4096, seed=19, init=init)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    tot0, ind0, _ = _totals(pi, om, g, q)
    om.update(60.0, 30, vel=_still(g), parf=np.full((16, 16), 900.0, np.float32),
              temp=np.full((16, 16, 1), 26.0, np.float32))
    tot1, ind1, _ = _totals(pi, om, g, q)
    assert om.diag(0, "dvid").sum() > 20 and om.diag(0, "graz").sum() > 20 and om.diag(0, "mort").sum() > 20
    for e in ("C", "N", "P"):
        assert abs(tot1[e] - tot0[e]) <= 1e-6 * tot0[e], (e, tot0[e], tot1[e])
        assert abs(ind1[e] - ind0[e]) > 1e-4 * tot0[e], (e, ind0[e], ind1[e])


def test_exudation_goes_to_doc_times_dt(pi, orc):
    """No light, no respiration, no losses, no remineralisation, CHmax = 0.05:
    the only carbon leaving the individuals is exu (CH -= exu), and DOC gains
    ΔT·exu (calc_consume!: (exu − VDOC)·ΔT), so ΔDOC = −ΔT·ΔC_individuals."""
    g = pi.RectilinearGrid(size=(8, 8, 1), x=(0, 16), y=(0, 16), z=(0, -16))
    zero_rm = {k: 0.0 for k in ("kDOC", "kDON", "kDOP", "kPOC", "kPON", "kPOP", "Nit")}
    spec = _spec(pi, g, {"Nsuper": [1e12], "respir": [0.0], "CHmax": [0.05], "mort_P": [0.0], "dvid_P": [0.0]},
                 300, 400, seed=7, bgc=zero_rm)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    _, ind0, tr0 = _totals(pi, om, g, q)
    dt = 60.0
    om.update(dt, 1, vel=_still(g), parf=np.zeros((8, 8), np.float32), temp=np.full((8, 8, 1), 20.0, np.float32))
    _, ind1, tr1 = _totals(pi, om, g, q)
    lost = ind0["C"] - ind1["C"]
    assert lost > 0.05 * ind0["C"]
    assert math.isclose(tr1["DOC"] - tr0["DOC"], dt * lost, rel_tol=1e-4)
    assert math.isclose(tr1["DIC"], tr0["DIC"], rel_tol=1e-9)


def _upload(pi, om, spec, cols):
    cap = spec.max_individuals
    arrays = {k: np.zeros(cap, np.float32) for k in ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "gen", "age",
                                                     "ex0")}
    arrays.update({k: np.zeros(cap, np.int64) for k in ("xi", "yi", "zi")})
    arrays["ac"] = np.zeros(cap, bool)
    n = len(cols["Sz"])
    for k, v in cols.items():
        arrays[k][:n] = v
    arrays["ac"][:n] = True
    arrays["gen"][:n] = 1
    om.upload_individuals(0, arrays)


def test_synthesis_at_saturation(pi, orc):
    """T = Topt (tempFunc = 1), half-saturation constants ~0 (every limit = 1),
    no light, uptake, respiration or exudation: one step changes PRO by
    ΔT·k_pro·RNA, RNA by ΔT·k_rna·PRO, DNA by ΔT·k_dna·PRO below 2 C_DNA·Nsuper
    and not at all above, CH by −(the three), NST by −Σ R_NC·(each), PST by
    −Σ R_PC·(DNA, RNA)."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    over = {"Nsuper": [1.0], "k_sat_pro": [1e-30], "k_sat_dna": [1e-30], "k_sat_rna": [1e-30], "respir": [0.0],
            "VNH4max": [0.0], "VNO3max": [0.0], "VPO4max": [0.0], "CHmax": [1.0], "mort_P": [0.0], "dvid_P": [0.0],
            "k_pro": [3e-4], "k_rna": [2e-5], "k_dna": [7e-6]}
    spec = _spec(pi, g, over, 0, 64)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(3)
    n = 40
    unit = q["C_DNA"]
    DNA = (unit * np.where(np.arange(n) < 20, 1.5, 2.5) * (1 + 0.1 * rng.random(n))).astype(np.float32)
    cols = {"x": rng.random(n).astype(np.float32) * 4, "y": rng.random(n).astype(np.float32) * 4,
            "z": np.full(n, 0.5, np.float32), "Sz": DNA,
            "ex0": (DNA * 1.7 * (1 + 0.2 * rng.random(n))).astype(np.float32),
            "Bm": (DNA * 35 * (1 + 0.2 * rng.random(n))).astype(np.float32),
            "Cq": (DNA * 20).astype(np.float32), "Nq": (DNA * 3).astype(np.float32), "Pq": (DNA * 0.5).astype(np.float32),
            "Chl": (DNA * 0.1).astype(np.float32)}
    _upload(pi, om, spec, cols)
    dt = 30.0
    om.update(dt, 1, vel=_still(g), parf=np.zeros((4, 4), np.float32),
              temp=np.full((4, 4, 1), np.float32(q["Topt"]), np.float32))
    got = {k: om.field(0, k)[:n].astype(np.float64) for k in ("Bm", "Sz", "ex0", "Cq", "Nq", "Pq")}
    c = {k: v.astype(np.float64) for k, v in cols.items()}
    sP, sR = dt * q["k_pro"] * c["ex0"], dt * q["k_rna"] * c["Bm"]
    sD = np.where(c["Sz"] / unit < 2, dt * q["k_dna"] * c["Bm"], 0.0)
    assert np.allclose(got["Bm"] - c["Bm"], sP, rtol=1e-3, atol=0) and (sP > 1e-4 * c["Bm"]).all()
    assert np.allclose(got["ex0"] - c["ex0"], sR, rtol=1e-3, atol=0)
    assert np.allclose(got["Sz"] - c["Sz"], sD, rtol=1e-3, atol=1e-9 * unit) and (sD[20:] == 0).all()
    assert np.allclose(got["Cq"] - c["Cq"], -(sP + sD + sR), rtol=1e-3, atol=0)
    assert np.allclose(got["Nq"] - c["Nq"], -(sP * q["R_NC_PRO"] + sD * q["R_NC_DNA"] + sR * q["R_NC_RNA"]), rtol=1e-3, atol=0)
    assert np.allclose(got["Pq"] - c["Pq"], -(sD * q["R_PC_DNA"] + sR * q["R_PC_RNA"]), rtol=1e-3, atol=0)


def test_respiration_beyond_carbohydrate_takes_protein(pi, orc):
    """update_quotas_2_kernel!: respiration larger than the carbohydrate pool
    leaves CH = 0, takes the rest from protein and returns its N to NST
    (R_NC_PRO); no light, no uptake, T = Topt, synthesis off (k = 0)."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    over = {"Nsuper": [1.0], "respir": [2e-3], "VNH4max": [0.0], "VNO3max": [0.0], "VPO4max": [0.0],
            "k_pro": [0.0], "k_rna": [0.0], "k_dna": [0.0], "CHmax": [1.0], "mort_P": [0.0], "dvid_P": [0.0]}
    spec = _spec(pi, g, over, 0, 64)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    n = 30
    rng = np.random.default_rng(9)
    DNA = (q["C_DNA"] * (1.2 + 0.3 * rng.random(n))).astype(np.float32)
    cols = {"x": (rng.random(n) * 4).astype(np.float32), "y": (rng.random(n) * 4).astype(np.float32),
            "z": np.full(n, 0.5, np.float32), "Sz": DNA, "ex0": DNA * 2, "Bm": DNA * 30,
            "Cq": (DNA * (0.5 + rng.random(n))).astype(np.float32), "Nq": DNA * 3, "Pq": DNA * 0.5, "Chl": DNA * 0.1}
    _upload(pi, om, spec, cols)
    dt = 60.0
    om.update(dt, 1, vel=_still(g), parf=np.zeros((4, 4), np.float32),
              temp=np.full((4, 4, 1), np.float32(q["Topt"]), np.float32))
    c = {k: v.astype(np.float64) for k, v in cols.items()}
    resp = q["respir"] * c["Bm"] * dt
    short = resp - c["Cq"]
    assert (short > 0.1 * resp).all()
    assert not om.field(0, "Cq")[:n].any()
    assert np.allclose(om.field(0, "Bm")[:n], c["Bm"] - short, rtol=1e-5, atol=0)
    assert np.allclose(om.field(0, "Nq")[:n], c["Nq"] + short * q["R_NC_PRO"], rtol=1e-5, atol=0)


def _tf(q, T):
    k = math.exp(-q["Ea"] / (8.3145 * (T + 273.15)))
    return min(1.0, k / math.exp(-q["Ea"] / (8.3145 * (q["Topt"] + 273.15))))


def test_decisions_division_by_dna_mortality_by_age(pi, orc):
    """calc_MM_dvid!: P = dvid_P above 2 C_DNA·Nsuper of DNA, 0 below; calc_MM_mort!:
    P = mort_P·shape_func_inc(age, mort_reg, 1e-4); each decision U < P·ΔT with U
    the shared stream's draw for the slot, in the counts P·ΔT predicts."""
    g = pi.RectilinearGrid(size=(6, 6, 1), x=(0, 6e3), y=(0, 6e3), z=(0, -10))
    over = {"dvid_P": [2e-3], "mort_P": [3e-3], "mort_reg": [4.0], "k_dna": [0.0], "grz_P": [0.0]}
    n = 8000
    spec = _spec(pi, g, over, 0, n + 10, seed=29)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(5)
    unit = q["C_DNA"] * q["Nsuper"]
    DNA = (unit * (1.0 + 2.0 * rng.random(n))).astype(np.float32)
    age = (rng.random(n) * 8.0).astype(np.float32)
    cols = {"x": (rng.random(n) * 6).astype(np.float32), "y": (rng.random(n) * 6).astype(np.float32),
            "z": np.full(n, 0.5, np.float32), "Sz": DNA, "ex0": DNA * 2, "Bm": DNA * 30, "Cq": DNA * 20,
            "Nq": DNA * 3, "Pq": DNA * 0.5, "Chl": DNA * 0.1, "age": age}
    _upload(pi, om, spec, cols)
    dt = 60.0
    om.update(dt, 1, vel=_still(g), parf=np.full((6, 6), 300.0, np.float32), temp=np.full((6, 6, 1), 22.0, np.float32))
    u = np.stack(pi.stream.uniform3(spec.seed, 1, 0, pi.stream.KIND_DECISION, np.arange(n))).astype(np.float64)
    age1 = (age + np.float32(dt / 3600)).astype(np.float64)
    dna1 = om.field(0, "Sz")[:n].astype(np.float64)
    fx = np.clip(1 - age1 / q["mort_reg"], 0, 1)
    p_mort = q["mort_P"] * (1 - fx ** 4 / (1e-4 + fx ** 4))
    p_dvid = np.where(dna1 / unit > 2, q["dvid_P"], 0.0)
    for k, p, row in (("mort", p_mort, 1), ("dvid", p_dvid, 2)):
        dec = om.field(0, k)[:n].astype(bool)
        pdt = p * dt
        near = np.abs(u[row] - pdt) <= 1e-6 * np.maximum(pdt, 1e-30)
        assert not ((dec != (u[row] < pdt)) & ~near).any(), k
        mu = float(np.clip(pdt, 0, 1).sum())
        assert dec.sum() > 50 and abs(dec.sum() - mu) < 5 * math.sqrt(mu) + 1, (k, dec.sum(), mu)
    assert not om.field(0, "dvid")[:n][dna1 / unit <= 2].any()


def test_dead_individuals_keep_synthesising(pi, orc):
    """Every individual dies in step 1 (mort_P·ΔT > 1 at any age once age >
    mort_reg); in step 2 find_NPT! gives each dead slot T = 0, so its protein
    grows by ΔT·k_pro·RNA·tempFunc(0) (saturating limits) whatever the cell's
    temperature, its C, N and P totals stay, and the tracers receive nothing
    (remineralisation and nitrification off, no flow: they stay bit for bit)."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    over = {"Nsuper": [1.0], "k_sat_pro": [1e-30], "k_sat_dna": [1e-30], "k_sat_rna": [1e-30], "respir": [0.0],
            "CHmax": [1.0], "mort_P": [1.0], "mort_reg": [1e-3], "dvid_P": [0.0], "k_pro": [3e-4]}
    zero_rm = {k: 0.0 for k in ("kDOC", "kDON", "kDOP", "kPOC", "kPON", "kPOP", "Nit")}
    spec = _spec(pi, g, over, 0, 64, bgc=zero_rm)
    q = _q(pi, spec)
    om = orc.OracleModel(spec)
    n = 30
    rng = np.random.default_rng(8)
    DNA = (q["C_DNA"] * (1.2 + 0.3 * rng.random(n))).astype(np.float32)
    cols = {"x": (rng.random(n) * 4).astype(np.float32), "y": (rng.random(n) * 4).astype(np.float32),
            "z": np.full(n, 0.5, np.float32), "Sz": DNA, "ex0": DNA * 2, "Bm": DNA * 30, "Cq": DNA * 20,
            "Nq": DNA * 3, "Pq": DNA * 0.5, "Chl": DNA * 0.1, "age": np.full(n, 5.0, np.float32)}
    _upload(pi, om, spec, cols)
    still, parf, temp = _still(g), np.full((4, 4), 300.0, np.float32), np.full((4, 4, 1), 26.0, np.float32)
    om.update(60.0, 1, vel=still, parf=parf, temp=temp)          # the mortality decisions
    om.update(60.0, 1, vel=still, parf=parf, temp=temp)          # all die, their last step with gathered T
    assert not om.field(0, "ac").any()
    before = {k: om.field(0, k)[:n].astype(np.float64) for k in ("Bm", "ex0", "Sz", "Cq", "Nq", "Pq", "Chl")}
    tr0 = [om.tracer(t).copy() for t in range(14)]
    dt = 45.0
    om.update(dt, 1, vel=still, parf=parf, temp=temp)
    after = {k: om.field(0, k)[:n].astype(np.float64) for k in before}
    tf0 = _tf(q, 0.0)
    assert 0.05 < tf0 < 0.5 < _tf(q, 26.0)
    assert np.allclose(after["Bm"] - before["Bm"], dt * q["k_pro"] * before["ex0"] * tf0, rtol=1e-3, atol=0)
    C = lambda d: d["Bm"] + d["Sz"] + d["ex0"] + d["Cq"] + d["Chl"] / 893.49 * 55
    P = lambda d: d["Sz"] * q["R_PC_DNA"] + d["ex0"] * q["R_PC_RNA"] + d["Pq"]
    assert np.allclose(C(after), C(before), rtol=1e-6, atol=0) and np.allclose(P(after), P(before), rtol=1e-6, atol=0)
    for t in (0, 1, 2, 3, 5, 6, 7, 9, 10, 11):     # DIC NH4 NO3 PO4 DOC DON DOP POC PON POP
        assert np.array_equal(om.tracer(t), tr0[t]), t
