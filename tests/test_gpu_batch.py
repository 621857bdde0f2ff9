"""Batched stepping (pi_batch_*): every member of a batch must end bit-identical
to the same model stepped alone (which the parity tests pin to the oracle), for
a mix of 0-D, 1-D, 2-D and 3-D members, with land, diagnostics, a velocity change
mid-run, a member leaving the batch, and a member that overflows its capacity
without disturbing the others."""
import numpy as np
import pytest

from parity import compare_models

pytestmark = pytest.mark.gpu


def _members(pi):
    out = []
    g = pi.RectilinearGrid(size=(1, 1, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    out.append((pi.ModelSpec(g, max_individuals=4000, seed=41,
                             phyto=pi.phyto_setup({"dvid_P": [3e-3], "mort_P": [2e-3], "grz_P": [2e-8]}, [1500], 1)),
                None))
    g = pi.RectilinearGrid(size=(1, 1, 12), x=(0, 32), y=(0, 32), z=(0, -120))
    rng = np.random.default_rng(1)
    w = np.zeros((1, 1, 13), np.float32)
    w[0, 0, 1:12] = (rng.standard_normal(11) * 1e-3).astype(np.float32)
    out.append((pi.ModelSpec(g, max_individuals=3000, seed=42,
                             phyto=pi.phyto_setup({"dvid_P": [2e-3, 2e-3]}, [800, 600], 2),
                             bgc_params={"κvP": 1e-3, "κv": 1e-3}),
                (np.zeros((1, 1, 12), np.float32), np.zeros((1, 1, 12), np.float32), w)))
    g = pi.RectilinearGrid(size=(20, 20, 1), x=(0, 20e3), y=(0, 20e3), z=(0, -10))
    out.append((pi.ModelSpec(g, max_individuals=6000, seed=43,
                             phyto=pi.phyto_setup({"dvid_P": [2e-3], "mort_P": [1e-3]}, [2500], 1),
                             bgc_params={"κhP": 1.0, "κh": 20.0}),
                pi.workloads.cellular_flow(g, U=2.0, L=5e3)))
    lm = np.ones((12, 10, 4), dtype=bool)
    lm[3:6, 2:5, 1:3] = False
    g = pi.RectilinearGrid(size=(12, 10, 4), x=(0, 12e3), y=(0, 10e3), z=(0, -40),
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded), landmask=lm)
    rng = np.random.default_rng(4)
    out.append((pi.ModelSpec(g, max_individuals=2500, seed=44,
                             phyto=pi.phyto_setup({"dvid_P": [2e-3]}, [1000], 1),
                             bgc_params={"κhP": 5.0, "κvP": 1e-3, "κh": 20.0, "κv": 1e-3}),
                ((rng.standard_normal((13, 10, 4)) * 2).astype(np.float32),
                 (rng.standard_normal((12, 11, 4)) * 2).astype(np.float32),
                 (rng.standard_normal((12, 10, 5)) * 1e-3).astype(np.float32))))
    spec, (u, v, w) = pi.workloads.config_3d(n=16, nz=8, N=1500, nsp=2, cap=4000, seed=45, nmodes=12,
                                             overrides={"dvid_P": [1e-3], "mort_P": [5e-4]}, kappa=(1.0, 1e-4))
    out.append((spec, (u * 20, v * 20, w * 20)))
    g = pi.RectilinearGrid(size=(8, 8, 1), x=(0, 8e3), y=(0, 8e3), z=(0, -10))
    out.append((pi.ModelSpec(g, max_individuals=120, seed=2,
                             phyto=pi.phyto_setup({"dvid_P": [0.05], "dvid_reg": [1.0e-3]}, [100], 1)), None))
    return out


def _inputs(pi, h, spec, uvw, scale=1.0):
    g = spec.grid
    if uvw is not None:
        h.set_velocity(*(np.asarray(a, np.float32) * np.float32(scale) for a in uvw))
    h.set_parf(np.full((g.Nx, g.Ny), 550.0, np.float32))
    h.set_temperature(np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))


def test_batch_matches_alone(pi):
    members = _members(pi)
    A = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    B = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    DIAG = [pi.TRACER_NAMES.index("NO3"), pi.TRACER_NAMES.index("DOC")]
    for (s, uvw), a, b in zip(members, A, B):
        _inputs(pi, a, s, uvw)
        _inputs(pi, b, s, uvw)
    A[2].set_diag_tracers(DIAG)
    B[2].set_diag_tracers(DIAG)
    # alone: 5 steps, a velocity change on members 2 and 3, 4 more steps
    for a in A:
        a.run(60.0, 5)
    for k in (2, 3):
        _inputs(pi, A[k], members[k][0], members[k][1], scale=1.5)
    for a in A:
        a.run(60.0, 4)
    # batched: all members for 5 steps; member 0 then leaves and runs alone
    batch = pi._native.NativeBatch()
    for b in B:
        batch.add(b)
    batch.run(60.0, 5)
    batch.remove(B[0])
    for k in (2, 3):
        _inputs(pi, B[k], members[k][0], members[k][1], scale=1.5)
    batch.run(60.0, 4)
    B[0].run(60.0, 4)
    batch.sync()
    for a in A[:-1]:
        a.sync()
    for b in B[:-1]:
        b.sync()
    # the member that overflows its capacity fails alone and in the batch
    with pytest.raises(pi.CapacityError):
        A[-1].sync()
    with pytest.raises(pi.CapacityError):
        B[-1].sync()
    for k, ((s, _), a, b) in enumerate(zip(members[:-1], A[:-1], B[:-1])):
        assert a.iteration() == b.iteration() == 9
        bad = compare_models(a, b, s.n_species, grid_fields=("par", "Chl", "pop"))
        assert not bad, f"member {k}: " + "\n".join(bad[:20])
    for q in DIAG:
        assert np.array_equal(A[2].diag_tracer(q), B[2].diag_tracer(q))
    batch.close()


def test_batched_members_match_oracle(pi, orc):
    """Members stepped in one batch against the oracle directly: a 0-D box, a
    1-D column with two species and vertical flow, a 2-D slab under the
    cellular flow, a bounded 3-D box with land and a 3-D Fourier-flow grid with
    two species, 8 steps in one batch; every member identical to its own oracle
    run."""
    members = _members(pi)[:5]
    hs = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
    for (s, uvw), h in zip(members, hs):
        _inputs(pi, h, s, uvw)
    batch = pi._native.NativeBatch()
    for h in hs:
        batch.add(h)
    batch.run(60.0, 8)
    batch.sync()
    for k, ((s, uvw), h) in enumerate(zip(members, hs)):
        g = s.grid
        om = orc.OracleModel(s)
        om.update(60.0, 8, vel=uvw, parf=np.full((g.Nx, g.Ny), 550.0, np.float32),
                  temp=np.full((g.Nx, g.Ny, g.Nz), 21.0, np.float32))
        assert h.iteration() == 8
        bad = compare_models(h, om, s.n_species)
        assert not bad, f"member {k}: " + "\n".join(bad[:20])
    batch.close()


def test_ensemble_update_matches_update(pi):
    """Ensemble.update (batched update! of several simulations with different
    lengths and forcing) against update on each simulation alone."""
    def sims():
        out = []
        for k, (spec, uvw) in enumerate(_members(pi)[:5]):
            g = spec.grid
            n = 6 + 2 * k
            m = pi.PlanktonModel(pi.HIPArch(), spec=spec)
            temp = pi.default_temperature(g, 60.0, n).astype(np.float32)
            kw = {}
            if uvw is not None:
                kw = dict(vels=dict(u=uvw[0][..., None], v=uvw[1][..., None], w=uvw[2][..., None]), ΔT_vel=60.0 * 1e6)
            out.append(pi.PlanktonSimulation(m, ΔT=60.0, iterations=n, temp=temp, **kw))
        return out
    alone = sims()
    for s in alone:
        pi.update(s)
    together = sims()
    ens = pi.Ensemble(together)
    ens.update()
    assert not ens.failed
    for k, (a, b) in enumerate(zip(alone, together)):
        na, nb = a.model._native, b.model._native
        assert na.iteration() == nb.iteration() == a.iterations
        bad = compare_models(na, nb, a.model.n_species)
        assert not bad, f"member {k}: " + "\n".join(bad[:20])
    ens.close()


@pytest.mark.parametrize("land", [False, True])
def test_device_generated_state_matches_host(pi, land):
    """pi_generate_state (the initial state drawn on the device) against the
    host-generated arrays of ModelSpec uploaded: identical before and after
    stepping (two species, N < capacity, periodic and bounded grids, land)."""
    if land:
        lm = np.ones((10, 9, 3), dtype=bool)
        lm[2:5, 3:6, 1:3] = False
        g = pi.RectilinearGrid(size=(10, 9, 3), x=(0, 10e3), y=(0, 9e3), z=(0, -30),
                               topology=(pi.Bounded, pi.Periodic, pi.Bounded), landmask=lm)
    else:
        g = pi.RectilinearGrid(size=(14, 12, 5), x=(0, 14e3), y=(0, 12e3), z=(0, -50))
    kw = dict(max_individuals=3000, seed=77, phyto=pi.phyto_setup({"dvid_P": [2e-3, 1e-3]}, [1800, 900], 2))
    A = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, **kw))._native
    B = pi.PlanktonModel(pi.HIPArch(), spec=pi.ModelSpec(g, host_state=False, **kw))._native
    bad = compare_models(A, B, 2, diags=False)
    assert not bad, "\n".join(bad[:20])
    for h in (A, B):
        _inputs(pi, h, pi.ModelSpec(g, **dict(kw, phyto=pi.phyto_setup(None, [1, 1], 2))), None)
        h.run(60.0, 3)
        h.sync()
    bad = compare_models(A, B, 2)
    assert not bad, "\n".join(bad[:20])


def test_served_members_match_alone(pi):
    """The serving loop (members admitted by footprint while others run, joining
    and leaving the running batch) against each member run alone through
    update: identical final state.  The budget admits only part of the stream
    at a time, so members also wait in the queue."""
    from importlib import import_module
    sv = import_module(pi.__name__ + ".serving")
    table = [dict(id=i, kind=k, size=sz, N=n, params={"dvid_P": [2e-3 * (1 + 0.1 * i)]}, seed=900 + i, steps=st)
             for i, (k, sz, n, st) in enumerate([(0, (1, 1, 1), 1500, 30), (1, (1, 1, 16), 4000, 45),
                                                   (2, (24, 24, 1), 3000, 20), (0, (1, 1, 1), 800, 50),
                                                   (2, (16, 16, 1), 2000, 35), (1, (1, 1, 8), 2500, 25)])]
    states = {}

    def grab(req, sim):
        nat = sim.model._native
        states[req.rid] = ({f: nat.field(0, f) for f in ("x", "y", "z", "Bm", "Cq", "Nq", "Pq", "Chl", "ac")},
                           [nat.tracer(q) for q in range(14)])

    reqs = [sv.Request(e["id"], (lambda e=e: pi.workloads.ensemble_member(e)), e["steps"]) for e in table]
    one = sv.member_footprint(pi.workloads.ensemble_member(table[1])[0])
    srv = sv.EnsembleServer(budget=3 * one, workers=2, on_finish=grab)
    try:
        for r in reqs[:3]:
            srv.submit(r)
        for _ in range(10):
            srv.step()
        for r in reqs[3:]:
            srv.submit(r)
        srv.drain()
    finally:
        srv.close()
    assert not [r.error for r in srv.finished if r.error], [r.error for r in srv.finished]
    assert sorted(states) == list(range(len(table)))
    for e in table:
        spec, vel = pi.workloads.ensemble_member(e)
        sim = sv.simulation_for(spec, vel, e["steps"])
        pi.update(sim)
        nat = sim.model._native
        f, t = states[e["id"]]
        for k, v in f.items():
            assert np.array_equal(nat.field(0, k), v), (e["id"], k)
        for q in range(14):
            assert np.array_equal(nat.tracer(q), t[q]), (e["id"], q)
        nat.close()


def test_block_budget_does_not_change_results(pi):
    """The batch's block budget only redistributes blocks: a tiny budget (one
    block per member for most kernels) and the default give identical states."""
    members = _members(pi)[:5]
    sets = []
    for budget in (256, None):
        hs = [pi.PlanktonModel(pi.HIPArch(), spec=s)._native for s, _ in members]
        for (s, uvw), h in zip(members, hs):
            _inputs(pi, h, s, uvw)
        b = pi._native.NativeBatch()
        if budget:
            b.set_budget(budget)
        for h in hs:
            b.add(h)
        b.run(60.0, 6)
        b.sync()
        sets.append((b, hs))
    for k, ((s, _), x, y) in enumerate(zip(members, sets[0][1], sets[1][1])):
        bad = compare_models(x, y, s.n_species)
        assert not bad, f"member {k}: " + "\n".join(bad[:20])
    for b, _ in sets:
        b.close()


def _alone_vs_budget(pi, spec, uvw, dt, steps, budget=512):
    """The model stepped alone (every kernel on its own full grid) and as a
    one-member batch on a budget of `budget` blocks per launch (each thread then
    walks many entries, cells and columns): returns both handles after `steps`
    steps.  The caller closes them."""
    a = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
    b = None
    try:
        b = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
        _inputs(pi, a, spec, uvw)
        _inputs(pi, b, spec, uvw)
        a.run(dt, steps)
        a.sync()
        batch = pi._native.NativeBatch()
        try:
            batch.set_budget(budget)
            batch.add(b)
            batch.run(dt, steps)
            batch.sync()
        finally:
            batch.close()
    except BaseException:
        a.close()
        if b is not None:
            b.close()
        raise
    return a, b


@pytest.mark.timeout(300)
def test_full_size_batched_matches_alone(pi):
    """The throughput workload at its full size: 256 x 256 x 64, periodic x/y,
    2 species x 4,194,304 individuals, capacity 2 x that.  Stepped alone and as
    a one-member batch on a block budget of 512, the states after 3 steps are
    identical.  Both runs use the same kernels, so this shows that the launch
    shape (grid-stride loops, look-ahead loads across iterations) does not change
    results at full scale; correctness itself is pinned by the oracle tests at
    smaller sizes."""
    spec, uvw = pi.workloads.config_3d(n=256, nz=64, N=4194304, nsp=2, cap=2 * 4194304, seed=2, host_state=False)
    a, b = _alone_vs_budget(pi, spec, uvw, 60.0, 3)
    try:
        n = [a.count_active(s) for s in range(2)]
        assert all(k > 4_000_000 for k in n), n
        bad = compare_models(a, b, 2)
        assert not bad, "\n".join(bad[:20])
    finally:
        a.close()
        b.close()


def _p_budget(pi, h, spec):
    """Total phosphorus (mol): PO4 + DOP + POP over the interior cells times
    their volume, plus Σ (Pq + Bm·R_PC) over the active individuals
    (the quantity test/example_2D_test.jl:19-47 checks)."""
    g = spec.grid
    vol = g.Δx * g.Δy * g.dzF[g.Hz:g.Hz + g.Nz].astype(np.float64)
    tot = 0.0
    for name in ("PO4", "DOP", "POP"):
        c = g.interior(h.tracer(pi._abi.TRACER_ID[name])).astype(np.float64)
        tot += float((c.sum(axis=(0, 1)) * vol).sum())
    R_PC = pi._abi.PARAM_NAMES.index("R_PC")
    for s in range(spec.n_species):
        ac = h.field(s, "ac")
        tot += float(h.field(s, "Pq")[ac].astype(np.float64).sum())
        tot += float(h.field(s, "Bm")[ac].astype(np.float64).sum()) * float(spec.species_params[s][R_PC])
    return tot


@pytest.mark.timeout(600)
def test_large_instance_invariants(pi):
    """The large single instance at its full size: 512 x 512 x 128, 4 species x
    8,388,608 individuals (capacity 2 x that), a 128-mode Fourier flow, dt = 30 s,
    state drawn on the device.  After 3 steps: per species, live individuals =
    initial - grazed - dead + divided (the per-species diagnostics, summed over
    the grid); total phosphorus conserved to 1e-6 relative; and the run alone
    equals the run as a one-member batch on 512 blocks."""
    N = 8388608
    spec, uvw = pi.workloads.config_3d(n=512, nz=128, N=N, nsp=4, cap=2 * N, seed=3, nmodes=128, host_state=False)
    a, b = _alone_vs_budget(pi, spec, uvw, 30.0, 3)
    try:
        # the initial budget, from a fresh copy of the same device-drawn state
        c = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
        try:
            P0 = _p_budget(pi, c, spec)
        finally:
            c.close()
        for s in range(4):
            d = {k: int(a.diag(s, k).astype(np.float64).sum()) for k in ("graz", "mort", "dvid")}
            assert d["dvid"] > 0 and d["mort"] > 0, d
            assert a.count_active(s) == N - d["graz"] - d["mort"] + d["dvid"], (s, d, a.count_active(s))
        P3 = _p_budget(pi, a, spec)
        assert abs(P3 - P0) <= 1e-6 * P0, (P0, P3)
        bad = compare_models(a, b, 4)
        assert not bad, "\n".join(bad[:20])
    finally:
        a.close()
        b.close()
