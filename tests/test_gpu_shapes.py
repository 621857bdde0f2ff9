"""Launch shapes (pi_set_launch_shape) and the autotuner.

Every launch shape only changes how a step's work is split over workgroups:
a model stepped with extreme shapes (one row per x/y work item, three levels
per column run, a handful of workgroups for advection, physiology and the
per-cell sums) must end bit-identical to the same model with the defaults,
alone and in a batch.  The tuner's choices are therefore free to be anything;
its tests check that it searches, caches and re-applies them."""
import numpy as np
import pytest

from parity import compare_models
from test_gpu_batch import _members, _inputs

pytestmark = pytest.mark.gpu

EXTREME = dict(xy_run=1, final_run=3, adv_blocks=7, phys_blocks=5, cellsum_per_cu=1, plk_blocks=3)


def _specs(pi):
    out = []
    spec, (u, v, w) = pi.workloads.config_3d(n=40, nz=20, N=6000, nsp=2, cap=16000, seed=7, nmodes=8,
                                             overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    out.append((spec, (u * 20, v * 20, w * 20)))
    lm = np.ones((24, 18, 9), dtype=bool)
    lm[5:9, 4:8, 2:6] = False
    g = pi.RectilinearGrid(size=(24, 18, 9), x=(0, 24e3), y=(0, 18e3), z=(0, -90),
                           topology=(pi.Periodic, pi.Bounded, pi.Bounded), landmask=lm)
    rng = np.random.default_rng(8)
    out.append((pi.ModelSpec(g, max_individuals=9000, seed=9,
                             phyto=pi.phyto_setup({"dvid_P": [2e-3]}, [4000], 1),
                             bgc_params={"κhP": 5.0, "κvP": 1e-3, "κh": 20.0, "κv": 1e-3}),
                ((rng.standard_normal((24, 18, 9)) * 2).astype(np.float32),
                 (rng.standard_normal((24, 19, 9)) * 2).astype(np.float32),
                 (rng.standard_normal((24, 18, 10)) * 1e-3).astype(np.float32))))
    # rows wider than 256 cells: the wide fused sweep
    spec, (u, v, w) = pi.workloads.config_3d(n=300, nz=3, N=20000, nsp=1, cap=50000, seed=10, nmodes=6,
                                             kappa=(1.0, 1e-4))
    out.append((spec, (u * 10, v * 10, w * 10)))
    return out


def test_shape_roundtrip_and_range(pi):
    spec, _ = _specs(pi)[0]
    h = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    try:
        assert h.shape() == {k: d for k, (_, d) in pi._abi.SHAPES.items()}
        h.set_shape(**EXTREME)
        assert h.shape() == EXTREME
        for bad in (dict(xy_run=0), dict(final_run=5000), dict(cellsum_per_cu=65), dict(adv_blocks=-1)):
            with pytest.raises(ValueError):
                h.set_shape(**bad)
        assert h.shape() == EXTREME
    finally:
        h.close()


@pytest.mark.parametrize("case", [0, 1, 2])
def test_shapes_do_not_change_results(pi, case):
    spec, uvw = _specs(pi)[case]
    a = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    b = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    try:
        b.set_shape(**EXTREME)
        for h in (a, b):
            _inputs(pi, h, spec, uvw)
            h.run(60.0, 3)
            if h is b:
                h.set_shape(final_run=5, xy_run=3)              # a change between runs
            h.run(60.0, 3)
            h.sync()
        bad = compare_models(a, b, spec.n_species)
        assert not bad, "\n".join(bad[:20])
    finally:
        a.close()
        b.close()


def test_batched_member_shapes_do_not_change_results(pi):
    members = _members(pi)[:5]
    A = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    B = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    batch = pi._native.NativeBatch()
    try:
        for (s, uvw), a, b in zip(members, A, B):
            _inputs(pi, a, s, uvw)
            _inputs(pi, b, s, uvw)
            a.run(60.0, 5)
            a.sync()
        for k, b in enumerate(B):
            b.set_shape(**(EXTREME if k % 2 == 0 else dict(xy_run=2, final_run=1, adv_blocks=1)))
            batch.add(b)
        batch.run(60.0, 5)
        batch.sync()
        for k, ((s, _), a, b) in enumerate(zip(members, A, B)):
            bad = compare_models(a, b, s.n_species)
            assert not bad, f"member {k}: " + "\n".join(bad[:20])
    finally:
        batch.close()
        for h in A + B:
            h.close()


def test_autotune_model_caches_and_applies(pi, tmp_path, monkeypatch):
    monkeypatch.setenv("PI_TUNE_CACHE", str(tmp_path / "shapes.json"))
    spec, uvw = _specs(pi)[0]
    h = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    h2 = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    try:
        _inputs(pi, h, spec, uvw)
        res = pi.autotune.tune_model(h, 60.0, steps=2, reps=1, warm=1)
        for name, cands in pi.autotune.CANDIDATES.items():
            assert name in res["trials"]
            assert res["shapes"][name] in set(cands) | {pi._abi.SHAPES[name][1]} | {spec.grid.Nz, spec.grid.Ny}
        assert h.shape() == res["shapes"]
        assert pi.autotune.apply_cached(h2)
        assert h2.shape() == res["shapes"]
    finally:
        h.close()
        h2.close()


def test_autotune_batch_per_class(pi, tmp_path, monkeypatch):
    monkeypatch.setenv("PI_TUNE_CACHE", str(tmp_path / "shapes.json"))
    members = _members(pi)[:5]
    hs = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    batch = pi._native.NativeBatch()
    try:
        for (s, uvw), h in zip(members, hs):
            _inputs(pi, h, s, uvw)
            batch.add(h)
        res = pi.autotune.tune_batch(batch, 60.0, steps=1, reps=1, warm=1,
                                     shapes=("xy_run", "final_run"), budgets=(512, 16384))
        assert set(res["classes"]) == {"0d", "1d", "2d", "3d"}
        assert res["budget"] in (512, 16384) and batch.budget == res["budget"]
        for h in hs:
            assert h.shape() == res["classes"][pi.autotune.grid_class(h.grid)]
        assert "budget" in res["trials"] and "3d:final_run" in res["trials"]
    finally:
        batch.close()
        for h in hs:
            h.close()
