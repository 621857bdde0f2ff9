"""Pins of the oracle's phytoplankton-particle interaction and release
(Individuals/Abiotic/particle_interaction.jl:1-138; Model/time_step.jl:119-159).

  interaction  hand-built cells whose outcome follows from the rules alone: a
               particle joins an individual of its own cell closer than Rd (in
               metres, through the grid spacings), particles in slot order,
               candidates in residue order (i mod max_candidates) + 1 with a
               later individual of the same residue replacing an earlier one,
               and only an individual below max_ptc takes one; on a random
               cloud, no particle is left beside an individual with room and
               none is taken from afar
  release      ptc 0 never releases, ptc·release_P·ΔT ≥ 1 always does; a
               released particle sits at the individual's x plus the draw that
               decided the release, and at y, z plus draws that split exactly
               like the last species' mortality and division decisions of the
               same step; the k-th releaser in slot order fills the k-th free
               slot; more releases than free slots is the capacity error
  diagnostics  ptc of the living individuals after the interaction, and the
               particle count after it (released particles keep the stale
               halo indices of their inactive slot)
  budget       ptc + active particles is conserved without creation
"""
import numpy as np
import pytest

HX = 2


def _still(g):
    return (np.zeros((g.Nx, g.Ny, g.Nz), np.float32), np.zeros((g.Nx, g.Ny, g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny, g.Nz + 1), np.float32))


def _forcing(g):
    return np.full((g.Nx, g.Ny), 400.0, np.float32), np.full((g.Nx, g.Ny, g.Nz), 20.0, np.float32)


def _spec(pi, g, N, Na, cap, intac=(), release=(), mc=25, over=None, aover=None, seed=5):
    nsp = len(N)
    ph = {"grz_P": [0.0] * nsp, "mort_P": [0.0] * nsp, "dvid_P": [0.0] * nsp}
    ph.update(over or {})
    return pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=cap, seed=seed, max_candidates=mc,
                        phyto=pi.phyto_setup(ph, list(N), nsp),
                        abiotic=pi.abiotic_setup(aover, list(Na), len(Na), palat=pi.Palat(intac=list(intac),
                                                                                       release=list(release))),
                        bgc_params={"κhP": 0.0, "κvP": 0.0})


def _place(ind, slots, pos, ptc=None):
    ind["ac"][:] = False
    for s, (x, y, z) in zip(slots, pos):
        ind["ac"][s] = True
        ind["x"][s], ind["y"][s], ind["z"][s] = x, y, z
    if ptc is not None:
        ind["ex7"][:] = 0.0
        for s, v in zip(slots, ptc):
            ind["ex7"][s] = v


# individuals (slots 1..4, 1-based) and particles (slots 1..5) of the hand-built
# cell: a 2x1x1 grid with Δx = 4 m, Δy = 2 m, Δz = 1 m and Rd = 0.9 m
IND = [(0.20, 0.50, 0.5), (0.20, 0.90, 0.5), (0.20, 0.10, 0.5), (0.95, 0.50, 0.5)]
PTC = [(0.20, 0.70, 0.5),   # A: 0.4 m from i1 and i2, 1.2 m from i3
       (0.20, 0.30, 0.5),   # B: 0.4 m from i1 and i3
       (0.60, 0.95, 0.5),   # C: 1.4 m or more from every individual
       (1.05, 0.50, 0.5),   # D: 0.4 m from i4, but in the next cell
       (0.45, 0.50, 0.5)]   # E: 0.25 grid units (1.0 m) from i1


@pytest.mark.parametrize("mc,maxptc,ptc0,want_ptc,want_ac", [
    # residues 2, 3, 4, 5: A and B both meet i1 first
    (25, 5.0, [0, 0, 0, 0], [2, 0, 0, 0], [0, 0, 1, 1, 1]),
    # i1 is full after A: B goes on to i3 (ptc < max_ptc is strict)
    (25, 1.0, [0, 0, 0, 0], [1, 0, 1, 0], [0, 0, 1, 1, 1]),
    # residues 2, 1, 2, 1: A meets i2 first; i3 replaces i1 in residue 2, so B joins i3
    (2, 5.0, [0, 0, 0, 0], [0, 1, 1, 0], [0, 0, 1, 1, 1]),
    # i3 full: B stays although i1, replaced in its residue, has room
    (2, 1.0, [0, 0, 1, 0], [0, 1, 1, 0], [0, 1, 1, 1, 1]),
])
def test_interaction_known_answers(pi, orc, mc, maxptc, ptc0, want_ptc, want_ac):
    g = pi.RectilinearGrid(size=(2, 1, 1), x=(0, 8.0), y=(0, 2.0), z=(0, -1.0))
    spec = _spec(pi, g, [4], [5], 16, intac=[("sp1", "sa1")], mc=mc, over={"max_ptc": [maxptc]},
                 aover={"Rd": [0.9]})
    _place(spec.initial_individuals[0], range(4), IND, ptc0)
    spec.initial_individuals[0]["ex7"][9] = 3.0           # an inactive slot's ptc: not in the diagnostic
    _place(spec.initial_abiotics[0], range(5), PTC)
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
    assert np.array_equal(om.field(0, "ptc")[:4], np.array(want_ptc, np.float32))
    assert np.array_equal(om.abiotic_field(0, "ac")[:5], np.array(want_ac, bool))
    assert np.array_equal(om.field(0, "x")[:4], np.array([p[0] for p in IND], np.float32))   # nothing moved
    d = om.diag(0, "ptc")
    assert d[HX, HX, HX] == sum(want_ptc) and d.sum() == sum(want_ptc)
    n = om.abiotic_diag(0)
    assert n[HX, HX, HX] == sum(want_ac[:3]) + want_ac[4] and n[HX + 1, HX, HX] == 1 and n.sum() == sum(want_ac)


def test_interaction_cloud(pi, orc):
    """400 particles and 60 individuals of max_ptc 3 scattered over a 4x3x2 grid
    with Δx ≠ Δy ≠ Δz and Rd of about a cell, three steps without motion (so the
    positions of the step are the initial ones): every particle taken had an
    individual of its own cell within Rd; every particle left has none with
    room; ptc stays in [0, max_ptc] and integral; ptc + particles is conserved.
    60 individuals ≤ max_candidates = 64: no residue is shared."""
    g = pi.RectilinearGrid(size=(4, 3, 2), x=(0, 4 * 3.0), y=(0, 3 * 2.0), z=(0, -2 * 1.5))
    spec = _spec(pi, g, [60], [400], 512, intac=[("sp1", "sa1")], mc=64, over={"max_ptc": [3.0]},
                 aover={"Rd": [0.8]}, seed=11)
    rng = np.random.default_rng(3)
    f32 = np.float32
    ind, ab = spec.initial_individuals[0], spec.initial_abiotics[0]
    _place(ind, range(60), [(f32(rng.random() * 4), f32(rng.random() * 3), f32(rng.random() * 2)) for _ in range(60)],
           [0.0] * 60)
    _place(ab, range(400), [(f32(rng.random() * 4), f32(rng.random() * 3), f32(rng.random() * 2)) for _ in range(400)])
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    X = np.stack([ind["x"][:60], ind["y"][:60], ind["z"][:60]], 1).astype(np.float64)
    P = np.stack([ab["x"][:400], ab["y"][:400], ab["z"][:400]], 1).astype(np.float64)
    cell = lambda A: np.floor(A).astype(int) @ np.array([1, 4, 12])
    ci, cp = cell(X), cell(P)
    dist = np.sqrt(((((P[:, None, :] - X[None, :, :]) * np.array([3.0, 2.0, 1.5]))) ** 2).sum(-1))
    near = (dist < 0.8) & (cp[:, None] == ci[None, :])
    total = 400.0
    prev = np.ones(400, bool)
    for step in range(3):
        om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
        ptc = om.field(0, "ptc")[:60].astype(np.float64)
        ac = om.abiotic_field(0, "ac")[:400]
        assert np.all(ptc == np.round(ptc)) and ptc.min() >= 0 and ptc.max() <= 3
        assert ptc.sum() + ac.sum() == total
        taken = prev & ~ac
        assert near[taken].any(1).all()
        room = ptc < 3
        assert not (near[ac] & room[None, :]).any()
        prev = ac
    assert 0 < ptc.sum() < 180 and (ptc == 3).sum() > 5


def _release_case(pi, orc, dvid=True):
    """32 individuals of each of two species in every other cell of a 64x1x1
    grid (species 1 at x = 2i + 0.3, species 2 at 2i + 0.6), alike in every column so
    that each species' mortality and division probabilities are the same for
    all of them (≈ 0.5 per step for species 2, zero for species 1); species 1
    holds ptc 0, 1 or 2 and releases into an abiotic species with
    release_P·ΔT = 0.5; particles already active in slots 0, 2 and 5."""
    g = pi.RectilinearGrid(size=(64, 1, 1), x=(0, 64.0), y=(0, 1.0), z=(0, -1.0))
    dt = 60.0
    over = {"mort_P": [0.0, 0.5 / dt], "dvid_P": [0.0, 0.5 / dt if dvid else 0.0], "mort_reg": [1.0, 1000.0],
            "dvid_type": [1, 1], "dvid_reg": [1.0, 1e-3]}
    spec = _spec(pi, g, [32, 32], [3], 64, release=[("sp1", "sa1")], over=over, aover={"release_P": [0.5 / dt]})
    for s in range(2):
        ind = spec.initial_individuals[s]
        for k, v in ind.items():
            if k != "ac":
                v[:32] = v[0]
        ind["x"][:32] = np.arange(0, 64, 2, dtype=np.float32) + np.float32(0.3 + 0.3 * s)
        ind["y"][:32] = 0.5
        ind["z"][:32] = 0.5
    ptc = np.ones(32, np.float32)
    ptc[3::4] = 0.0
    ptc[2::4] = 2.0
    spec.initial_individuals[0]["ex7"][:32] = ptc
    _place(spec.initial_abiotics[0], [0, 2, 5], [(63.9, 0.9, 0.9)] * 3)
    return g, dt, spec, ptc


def test_release_known_answers(pi, orc):
    g, dt, spec, ptc0 = _release_case(pi, orc)
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    om.update(dt, 1, vel=_still(g), parf=parf, temp=temp)
    ac = om.abiotic_field(0, "ac")
    x, y, z = (om.abiotic_field(0, n).astype(np.float64) for n in "xyz")
    new = np.flatnonzero(ac)
    new = new[~np.isin(new, [0, 2, 5])]
    who = np.floor((x[new] - 0.3) / 2).astype(int)           # the releaser: x offset in [0, 1)
    free = np.setdiff1d(np.arange(64), [0, 2, 5])
    assert np.array_equal(new, free[:len(new)])              # the k-th releaser fills the k-th free slot
    assert np.all(np.diff(who) > 0)
    xoff = x[new] - (2 * who + 0.3)
    assert np.all(xoff >= -1e-6) and np.all(xoff < 0.5 * ptc0[who])   # the deciding draw, not a fresh one
    ptc = om.field(0, "ptc")[:32]
    rel = np.zeros(32, bool)
    rel[who] = True
    assert not rel[ptc0 == 0].any() and rel[ptc0 == 2].all() and 2 <= rel[ptc0 == 1].sum() <= 14
    assert np.array_equal(ptc, ptc0 - rel)
    n = om.abiotic_diag(0)
    assert n.sum() == len(new) + 3 and n[HX + 63, HX, HX] == 3     # released ones counted at the halo cell
    yoff, zoff = y[new] - 0.5, z[new] - 0.5
    assert np.all((yoff >= 0) & (yoff < 1)) and np.all((zoff >= 0) & (zoff < 1))
    # the next step applies species 2's decisions drawn in this one
    m1, d1 = om.diag(1, "mort").copy(), om.diag(1, "dvid").copy()
    om.update(dt, 1, vel=_still(g), parf=parf, temp=temp)
    died = (om.diag(1, "mort") - m1)[HX + 2 * who, HX, HX] > 0
    split = (om.diag(1, "dvid") - d1)[HX + 2 * who, HX, HX] > 0
    assert died.any() and (~died).any()
    assert yoff[died].max() < yoff[~died].min()              # rnd.y = the last species' mortality draw
    alive = ~died
    assert split[alive].any() and (~split[alive]).any()
    assert zoff[alive & split].max() < zoff[alive & ~split].min()   # rnd.z = its division draw


def test_release_capacity_error(pi, orc):
    """62 of 64 particle slots taken and every individual sure to release."""
    g, dt, spec, _ = _release_case(pi, orc, dvid=False)
    spec.initial_individuals[0]["ex7"][:32] = 2.0
    ab = spec.initial_abiotics[0]
    ab["ac"][:] = True
    ab["ac"][[7, 40]] = False
    ab["x"][:], ab["y"][:], ab["z"][:] = 63.9, 0.9, 0.9
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    with pytest.raises(pi.CapacityError, match="number of abiotic particles exceeds the capacity at timestep"):
        om.update(dt, 1, vel=_still(g), parf=parf, temp=temp)


def test_budget_with_interaction_and_release(pi, orc):
    """Interaction and release together on a cloud, six steps: ptc + active
    particles conserved; released particles land within a cell width of
    their individual in x."""
    g = pi.RectilinearGrid(size=(4, 4, 2), x=(0, 8.0), y=(0, 8.0), z=(0, -4.0))
    spec = _spec(pi, g, [80], [200], 400, intac=[("sp1", "sa1")], release=[("sp1", "sa1")],
                 over={"max_ptc": [4.0]}, aover={"Rd": [1.5], "release_P": [0.1 / 60.0]}, seed=21)
    rng = np.random.default_rng(9)
    f32 = np.float32
    _place(spec.initial_individuals[0], range(80),
           [(f32(rng.random() * 4), f32(rng.random() * 4), f32(rng.random() * 2)) for _ in range(80)], [0.0] * 80)
    _place(spec.initial_abiotics[0], range(200),
           [(f32(rng.random() * 4), f32(rng.random() * 4), f32(rng.random() * 2)) for _ in range(200)])
    om = orc.OracleModel(spec)
    parf, temp = _forcing(g)
    for _ in range(6):
        om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
        assert om.field(0, "ptc")[:80].sum() + om.abiotic_field(0, "ac").sum() == 200
    assert om.field(0, "ptc")[:80].sum() > 0


def test_interaction_needs_ironenergymode(pi):
    g = pi.RectilinearGrid(size=(2, 2, 1), x=(0, 2.0), y=(0, 2.0), z=(0, -1.0))
    with pytest.raises(ValueError, match="IronEnergyMode"):
        pi.ModelSpec(g, max_individuals=16, phyto=pi.phyto_setup(None, [4], 1),
                     abiotic=pi.abiotic_setup(None, [4], 1, palat=pi.Palat(intac=[("sp1", "sa1")])))
    with pytest.raises(ValueError, match="sa2 is not generated"):
        pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=16, phyto=pi.phyto_setup(None, [4], 1),
                     abiotic=pi.abiotic_setup(None, [4], 1, palat=pi.Palat(release=[("sp1", "sa2")])))
    # release alone is accepted in every mode (ptc stays 0 outside IronEnergyMode)
    pi.ModelSpec(g, max_individuals=16, phyto=pi.phyto_setup(None, [4], 1),
                 abiotic=pi.abiotic_setup(None, [4], 1, palat=pi.Palat(release=[("sp1", "sa1")])))
