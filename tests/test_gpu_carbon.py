"""CarbonMode on the GPU against the oracle (same bar as QuotaMode: every field,
tracer and diagnostic identical).  CarbonMode (src/Individuals/Plankton/CarbonMode)
keeps Bm and the damaged biomass Bd (the Cq column), exchanges only DIC by
consumption and DOC/POC by losses, damages biomass above Topt and repairs it,
and its daughters keep their slot's Bd."""
import numpy as np
import pytest

from parity import compare_models
from test_gpu_parity import _drive
from test_gpu_batch import _inputs

pytestmark = pytest.mark.gpu


def _carbon(pi, g, over, N, cap, seed, nsp=1, bgc=None):
    return pi.ModelSpec(g, mode=pi.CarbonMode(), max_individuals=cap, seed=seed,
                        phyto=pi.phyto_setup(over, N, nsp), bgc_params=bgc)


def test_parity_carbon_3d_two_species(pi, orc):
    """Two species, one with thermal damage (surface above Topt = 26.5 on the
    default temperature profile) and one without; division, grazing and
    mortality (thermal and size-dependent); Fourier flow."""
    g = pi.RectilinearGrid(size=(16, 16, 8), x=(0, 16e3), y=(0, 16e3), z=(0, -80))
    over = {"Topt": [26.5, 26.5], "f_T2B": [2e-3, 2e-4], "thermal": [1.0, 0.0], "dvid_P": [2e-3, 1.5e-3],
            "dvid_reg": [1.1, 1.2], "mort_P": [2e-3, 1e-3], "mort_reg": [1.3, 1.3], "grz_P": [3e-5, 1e-5]}
    spec = _carbon(pi, g, over, [1500, 1200], 4000, 51, nsp=2, bgc={"κhP": 1.0, "κvP": 1e-4, "κh": 10.0, "κv": 1e-4})
    # species 1 starts with damaged biomass up to 99.5 % of Bm (thermal mortality from step 1)
    ini = spec.initial_individuals[0]
    frac = np.random.default_rng(3).random(1500) * np.float32(0.995)
    ini["Cq"][:1500] = (ini["Bm"][:1500] * frac).astype(np.float32)
    uvw = pi.workloads.fourier_flow(g, nmodes=10, seed=5, rms=1.0)
    gm, om = _drive(pi, orc, spec, uvw, 60.0, 16, diag_tracers=("DIC", "DOC", "POC"))
    for s in range(2):
        assert gm._native.count_active(s) == om.count_active(s)
    for s in range(2):
        assert om.diag(s, "dvid").sum() > 0 and om.diag(s, "mort").sum() > 0, s
    assert np.any(om.field(0, "Cq")[om.field(0, "ac")] > 0)            # damaged biomass present
    ids = [pi._abi.TRACER_ID[t] for t in ("DIC", "DOC", "POC")]
    bad = compare_models(gm._native, om, 2, diag_tracers=ids)
    assert not bad, "\n".join(bad[:20])


def test_parity_carbon_reborn_daughters(pi, orc):
    """Dense division and death in a 2-D run, bacteria on: daughters placed in
    slots whose occupant dies in the same step (they inherit that occupant's Bd
    and gathered cell) and in slots free before; Timer division."""
    g = pi.RectilinearGrid(size=(12, 12, 1), x=(0, 12e3), y=(0, 12e3), z=(0, -10))
    over = {"dvid_P": [1e-3], "dvid_type": [3], "dvid_reg": [1.0], "dvid_reg2": [0.2], "mort_P": [4e-3],
            "Topt": [20.0], "f_T2B": [5e-4], "is_bact": [1.0], "grz_P": [2e-5]}
    spec = _carbon(pi, g, over, [2500], 12000, 53, bgc={"κhP": 1.0, "κh": 10.0})
    ini = spec.initial_individuals[0]
    frac = np.random.default_rng(4).random(2500) * np.float32(0.999)
    ini["Cq"][:2500] = (ini["Bm"][:2500] * frac).astype(np.float32)
    uvw = pi.workloads.cellular_flow(g, U=2.0, L=4e3)
    gm, om = _drive(pi, orc, spec, uvw, 120.0, 14)
    assert om.diag(0, "dvid").sum() > 50 and om.diag(0, "mort").sum() > 50
    bad = compare_models(gm._native, om, 1)
    assert not bad, "\n".join(bad[:20])


def test_carbon_and_quota_members_in_one_batch(pi, orc):
    """A CarbonMode and a QuotaMode member stepped by the same batched launches
    (each kernel position holds both physiology kernels, each member launching
    its own): both identical to their oracle runs."""
    g = pi.RectilinearGrid(size=(14, 10, 3), x=(0, 14e3), y=(0, 10e3), z=(0, -30))
    cs = _carbon(pi, g, {"dvid_P": [2e-3], "mort_P": [1e-3], "Topt": [20.0], "f_T2B": [3e-4]}, [1200], 3000, 61,
                 bgc={"κhP": 1.0, "κh": 10.0})
    qs = pi.ModelSpec(g, max_individuals=3000, seed=62, phyto=pi.phyto_setup({"dvid_P": [2e-3]}, [1200], 1),
                      bgc_params={"κhP": 1.0, "κh": 10.0})
    uvw = pi.workloads.cellular_flow(g, U=1.0, L=5e3)
    hs = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s in (cs, qs)]
    oms = [orc.OracleModel(s) for s in (cs, qs)]
    batch = pi._native.NativeBatch()
    try:
        for s, h in zip((cs, qs), hs):
            _inputs(pi, h, s, uvw)
            batch.add(h)
        batch.run(60.0, 8)
        batch.sync()
        for om, s in zip(oms, (cs, qs)):
            for _ in range(8):                   # update!: the inputs staged before every step
                _inputs(pi, om, s, uvw)
                om.step(np.float32(60.0))
        for k, (h, om) in enumerate(zip(hs, oms)):
            bad = compare_models(h, om, 1)
            assert not bad, f"member {k}: " + "\n".join(bad[:20])
    finally:
        batch.close()
        for h in hs:
            h.close()


def test_carbon_device_generated_state_matches_host(pi):
    """pi_generate_state in CarbonMode (Bm and Chl from the size, Bd = 0)
    against the host generator, before and after stepping."""
    g = pi.RectilinearGrid(size=(10, 8, 4), x=(0, 10e3), y=(0, 8e3), z=(0, -40))
    kw = dict(mode=pi.CarbonMode(), max_individuals=2500, seed=71,
              phyto=pi.phyto_setup({"dvid_P": [2e-3, 1e-3]}, [1500, 700], 2))
    A = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, **kw))._native
    B = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, host_state=False, **kw))._native
    try:
        bad = compare_models(A, B, 2, diags=False)
        assert not bad, "\n".join(bad[:20])
        spec = pi.ModelSpec(g, **dict(kw, phyto=pi.phyto_setup(None, [1, 1], 2)))
        for h in (A, B):
            _inputs(pi, h, spec, None)
            h.run(60.0, 3)
            h.sync()
        bad = compare_models(A, B, 2)
        assert not bad, "\n".join(bad[:20])
    finally:
        A.close()
        B.close()


def test_carbon_model_api(pi):
    """The mirror of the CarbonMode StructArray: `Bd` is a column, Cq/Nq/Pq are not."""
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 4e3), y=(0, 4e3), z=(0, -20))
    m = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, mode=pi.CarbonMode(), max_individuals=100,
                                                         phyto=pi.phyto_setup(None, [50], 1)))
    d = m.individuals.phytos.sp1.data
    assert d.Bd.shape == (100,) and not np.any(d.Bd)
    with pytest.raises(AttributeError):
        d.Cq
    assert np.array_equal(d.Chl[:50], d.Bm[:50] * np.float32(m.species_params[0][pi._abi.CARBON_PARAM_NAMES.index("Chl2C")]))
