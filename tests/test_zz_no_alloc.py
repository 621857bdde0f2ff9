"""pi_batch_run_noalloc (launch tables in the library's static device scratch,
no device allocation) against pi_batch_run on the same members: identical
states after a first call, and again after a second call on a smaller batch
(one member removed from each), both through the C ABI."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu


def _specs(pi):
    out = []
    g = pi.RectilinearGrid(size=(1, 1, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    out.append((pi.ModelSpec(g, max_individuals=4000, seed=61,
                             phyto=pi.phyto_setup({"dvid_P": [3e-3], "mort_P": [2e-3], "grz_P": [2e-8]}, [1500], 1)),
                None))
    g = pi.RectilinearGrid(size=(24, 20, 1), x=(0, 24e3), y=(0, 20e3), z=(0, -10))
    out.append((pi.ModelSpec(g, max_individuals=6000, seed=62,
                             phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                             bgc_params={"κhP": 1.0, "κh": 20.0}),
                pi.workloads.cellular_flow(g, U=2.0, L=5e3)))
    spec, (u, v, w) = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=63, nmodes=12,
                                             overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    out.append((spec, (u * 20, v * 20, w * 20)))
    return out


def _start(pi, specs):
    hs = []
    for s, uvw in specs:
        h = pi.PlanktonModel(pi.HIPArch(), spec=s)._native
        g = s.grid
        if uvw is not None:
            h.set_velocity(*uvw)
        h.set_parf(np.full((g.Nx, g.Ny), 650.0, np.float32))
        h.set_temperature(np.full((g.Nx, g.Ny, g.Nz), 23.0, np.float32))
        hs.append(h)
    return hs


def test_batch_run_noalloc_matches_batch_run(pi):
    specs = _specs(pi)
    A, B = _start(pi, specs), _start(pi, specs)
    ba, bb = pi._native.NativeBatch(), pi._native.NativeBatch()
    try:
        for a, b in zip(A, B):
            ba.add(a)
            bb.add(b)
        ba.run(60.0, 5)
        bb.run_noalloc(60.0, 5)
        ba.sync()
        bb.sync()
        for k, ((s, _), a, b) in enumerate(zip(specs, A, B)):
            assert a.iteration() == b.iteration() == 5
            bad = compare_models(a, b, s.n_species)
            assert not bad, f"first call, member {k}: " + "\n".join(bad[:20])
        # second call: a smaller batch (smaller tables), one velocity change
        ba.remove(A[2])
        bb.remove(B[2])
        s1, uvw1 = specs[1]
        A[1].set_velocity(*(np.asarray(x) * np.float32(1.5) for x in uvw1))
        B[1].set_velocity(*(np.asarray(x) * np.float32(1.5) for x in uvw1))
        ba.run(60.0, 4)
        bb.run_noalloc(60.0, 4)
        ba.sync()
        bb.sync()
        for k, ((s, _), a, b) in enumerate(zip(specs[:2], A[:2], B[:2])):
            assert a.iteration() == b.iteration() == 9
            bad = compare_models(a, b, s.n_species)
            assert not bad, f"second call, member {k}: " + "\n".join(bad[:20])
    finally:
        ba.close()
        bb.close()
        for h in A + B:
            h.close()
