"""CPU tests of the host side: the C ABI library exports what include/pi_hip.h
declares (no compute call without a GPU), host-side logic mirrors the
reference's constructors and validators, and the oracle wrapper is wired."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(REPO, "include", "pi_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:pi_status|const char\*)\s+(pi_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol(pi):
    path = pi.lib_path()
    if not os.path.exists(path):
        pytest.skip("libpi_hip.so not built (run __graft_entry__.build())")
    syms = declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = C.CDLL(path)          # loads without touching a device
    for s in syms:
        assert hasattr(lib, s)


def test_oracle_exports_mirror_abi(orc):
    lib = orc.load()
    for s in declared_symbols():
        name = "orc_" + s[3:]
        if s in ("pi_init", "pi_version", "pi_model_create", "pi_run", "pi_sync", "pi_enable_timing",
                 "pi_kernel_times", "pi_probe_copy", "pi_footprint", "pi_model_footprint", "pi_generate_state",
                     "pi_set_launch_shape", "pi_get_launch_shape"):
            continue
        if s.startswith("pi_batch_"):       # batches of GPU models: no CPU counterpart
            continue
        assert hasattr(lib, name), name


def test_descriptor_layout_matches_header(pi):
    A = pi._abi
    assert C.sizeof(A.GridDesc) == 9 * 4 + 2 * 4 + 4 + 3 * 8    # 9 ints, 2 floats, pad, 3 pointers
    assert A.ModelDesc.grid.offset == 0
    assert A.ModelDesc.capacity.offset % 8 == 0
    assert len(A.PARAM_NAMES) == 40 and len(A.BGC_NAMES) == 23 and len(A.TRACER_NAMES) == 14


def test_param_defaults_and_update(pi):
    b = pi.bgc_params_default()
    assert b["kDOC"] == np.float32(1 / 30 / 86400) and b["shared_graz"] == 1.0
    with pytest.raises(ValueError, match="bgc parameter not found"):
        pi.update_bgc_params({"nope": 1.0})
    p = pi.update_phyt_params({"dvid_P": [2e-4]}, N=1)
    assert p["dvid_P"] == [np.float32(2e-4)]
    with pytest.raises(ValueError, match="phyt parameter not found"):
        pi.update_phyt_params({"nope": [1]})
    m = pi.params.species_param_matrix(pi.phyt_params_default(2), 2)
    assert m.shape == (2, 64) and m[0, pi._abi.PARAM_NAMES.index("R_NC")] == np.float32(16 / 106)
    assert not m[:, 40:].any()
    with pytest.raises(ValueError, match="division type"):
        pi.params.species_param_matrix(pi.update_phyt_params({"dvid_type": [7]}), 1)


def test_modelspec_validation(pi):
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 4), y=(0, 4), z=(0, -2))
    with pytest.raises(AssertionError):
        pi.ModelSpec(g, max_individuals=10, phyto=pi.phyto_setup(None, [20], 1))
    with pytest.raises(ValueError, match="length of `N`"):
        pi.ModelSpec(g, max_individuals=100, phyto=pi.phyto_setup(None, [20, 20], 1))
    with pytest.raises(NotImplementedError):
        pi.ModelSpec(g, mode=object())
    with pytest.raises(ValueError, match="parameter not found"):
        pi.ModelSpec(g, mode=pi.CarbonMode(), phyto=pi.phyto_setup({"Cqmax": [0.4]}, [10], 1))
    s = pi.ModelSpec(g, max_individuals=64, phyto=pi.phyto_setup(None, [40], 1), seed=3)
    ind = s.initial_individuals[0]
    assert ind["ac"][:40].all() and not ind["ac"][40:].any()
    assert (ind["Sz"][:40] >= 1.0).all() and (ind["x"][:40] < 4).all()
    assert np.all(ind["gen"][:40] == 1.0)
    # QuotaMode/plankton_generation.jl:63-69
    q = dict(zip(pi._abi.PARAM_NAMES, s.species_params[0]))
    assert np.array_equal(ind["Bm"], (q["Cquota"] * ind["Sz"]) * q["Nsuper"])
    assert np.array_equal(ind["Chl"], ind["Bm"] * q["Chl2Cint"])
    tr = s.initial_tracers[pi._abi.TRACER_ID["DIC"]]
    assert tr.max() < 23.0 and tr.min() > 17.0      # test/field_test.jl:12-14


def test_landmask_masks_individuals_and_tracers(pi):
    lm = np.ones((4, 4, 1), bool)
    lm[:2] = False
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4), y=(0, 4), z=(0, -1), landmask=lm)
    s = pi.ModelSpec(g, max_individuals=400, phyto=pi.phyto_setup(None, [300], 1), seed=1)
    ind = s.initial_individuals[0]
    on_land = ind["x"][:300] < 2.0
    assert not ind["ac"][:300][on_land].any() and ind["ac"][:300][~on_land].all()
    assert np.all(g.interior(s.initial_tracers[0])[:2] == 0.0)


def test_simulation_validators(pi):
    class Fake:
        pass
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 4), y=(0, 4), z=(0, -2))
    m = Fake()
    m.grid = g
    m.n_species = 1
    m.tracers = {}

    class N:
        def set_diag_tracers(self, ids):
            pass
    m._native = N()
    with pytest.raises(ValueError, match="integer times"):
        pi.PlanktonSimulation(m, ΔT=60, iterations=10, ΔT_PAR=90.0)
    with pytest.raises(ValueError, match="size of PARF"):
        pi.PlanktonSimulation(m, ΔT=60, iterations=10, PARF=np.zeros((3, 3, 2)))
    with pytest.raises(ValueError, match="size of u"):
        pi.PlanktonSimulation(m, ΔT=60, iterations=2, vels=dict(u=np.zeros((3, 4, 2, 3)), v=np.zeros((4, 4, 2, 3)),
                                                               w=np.zeros((4, 4, 3, 3))))
    s = pi.PlanktonSimulation(m, ΔT=60, iterations=10)
    assert s.PARF.shape == (4, 4, 24) and s.temp.shape == (4, 4, 2, 24)
    with pytest.raises(ValueError, match="not one of the diagnostics"):
        pi.PlanktonDiagnostics(m, tracer=("nope",))


def test_default_forcing_values(pi):
    g = pi.RectilinearGrid(size=(2, 2, 3), x=(0, 2), y=(0, 2), z=(0, -30))
    par = pi.default_PARF(g, 60, 1440)
    assert par.shape == (2, 2, 48) and par[0, 0, 12] == 1823.7921666666664 and par[0, 0, 0] == 0.0
    t = pi.default_temperature(g, 60, 10)
    # bottom layer = the surface series; upper layers add 4e-4 * zC[j] with the padded zC (param_default.jl:33-35)
    assert t[0, 0, 2, 0] == 26.646446609406727
    assert t[0, 0, 1, 0] == 26.646446609406727 + 4.0e-4 * float(g.zC[1])


def test_product_refuses_to_run_without_library(pi, monkeypatch, tmp_path):
    monkeypatch.setattr(pi._native, "_lib", None)
    monkeypatch.setattr(pi._native, "lib_path", lambda: str(tmp_path / "libpi_hip.so"))
    with pytest.raises(pi.NativeError, match="no CPU fallback"):
        pi._native.load_library()


def test_footprint_is_a_dry_run(pi):
    """pi_footprint sums the allocations of pi_model_create without a device:
    positive, growing with capacity and grid, and linear in the capacity up to
    the 256-byte rounding of each allocation."""
    if not os.path.exists(pi.lib_path()):
        pytest.skip("libpi_hip.so not built (run __graft_entry__.build())")
    from importlib import import_module
    nat = import_module(pi.__name__ + "._native")
    P = import_module(pi.__name__ + ".params")
    g = pi.RectilinearGrid(size=(16, 16, 4), x=(0, 16e3), y=(0, 16e3), z=(0, -40))
    pp = P.update_phyt_params(P.phyt_params_default(1, pi.QuotaMode()), np.float32, N=1, mode=pi.QuotaMode())
    prm = P.species_param_matrix(pp, 1)
    bgc = P.bgc_vector(pi.bgc_params_default())
    f = [nat.footprint(g, 1, cap, prm, bgc) for cap in (10000, 20000, 40000)]
    assert 0 < f[0] < f[1] < f[2]
    assert abs((f[2] - f[1]) - 2 * (f[1] - f[0])) < 64 * 256
    g2 = pi.RectilinearGrid(size=(32, 32, 4), x=(0, 32e3), y=(0, 32e3), z=(0, -40))
    assert nat.footprint(g2, 1, 10000, prm, bgc) > f[0]


def test_ctypes_mirror_matches_header(pi, tmp_path):
    """The Python mirror of the C ABI's structs (size and every field's offset)
    against include/pi_hip.h compiled by gcc."""
    A = pi._abi
    structs = {"pi_grid_desc": A.GridDesc, "pi_metrics": A.Metrics, "pi_model_desc": A.ModelDesc,
               "pi_individuals_host": A.IndividualsHost}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pi_hip.h"', "int main(void) {"]
    for cname, st in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in st._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for cname, st in structs.items():
        assert int(got[cname]) == C.sizeof(st), cname
        for f in st._fields_:
            assert int(got[f"{cname}.{f[0]}"]) == getattr(st, f[0]).offset, f"{cname}.{f[0]}"


def test_binding_test_struct_matches_abi(pi):
    """The reference-layout mirror used by the binding test (written apart from
    the package, as a Julia binding would be) has the ABI's size and offsets."""
    import test_gpu_binding as B
    for mine, ref in ((B.ModelDesc, pi._abi.ModelDesc), (B.GridDesc, pi._abi.GridDesc),
                      (B.IndividualsHost, pi._abi.IndividualsHost)):
        assert C.sizeof(mine) == C.sizeof(ref), mine.__name__
        assert [getattr(mine, f[0]).offset for f in mine._fields_] == \
               [getattr(ref, f[0]).offset for f in ref._fields_], mine.__name__
