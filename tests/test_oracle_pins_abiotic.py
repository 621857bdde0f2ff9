"""Pins of the oracle's abiotic particles (Individuals/Abiotic/particle_generation.jl,
particle_interaction.jl:158-267; Model/time_step.jl:21-38, 152-159).

  creation   a face whose flux makes every draw pass creates one particle per
             cell next to it, the k-th cell in column-major order filling the
             k-th free slot at (i, j, k) − H − 0.4; each cell's decision is
             isless(u·sz_min·Nsuper, bc·ΔT·A) with u the cell's draw of the
             shared stream and A the face's area; at an edge the later face of
             top, bottom, east, west, north, south decides; a per-iteration
             flux is read at the model's iteration; more particles than free
             slots is the capacity error
  transport  the RK4 displacement of a uniform flow and the random walk with
             the shared stream's draws of the particle species and the stale
             (xi, yi, zi), exactly
  num        diags.abiotics.saN.num is the per-cell count of active particles,
             summed over the steps
"""
import ctypes as C

import numpy as np
import pytest


def _model(pi, orc, g, N=(0,), cap=400, bgc=None, seed=4, nphy=0):
    spec = pi.ModelSpec(g, max_individuals=cap, seed=seed, phyto=pi.phyto_setup(None, [nphy], 1),
                        abiotic=pi.abiotic_setup(None, list(N), len(N)), bgc_params=bgc or {})
    return spec, orc.OracleModel(spec)


def _still(g):
    return (np.zeros((g.Nx + (g.topology_ids[0] == 1), g.Ny, g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny + (g.topology_ids[1] == 1), g.Nz), np.float32),
            np.zeros((g.Nx, g.Ny, g.Nz + 1), np.float32))


def _forcing(g):
    return np.full((g.Nx, g.Ny), 400.0, np.float32), np.full((g.Nx, g.Ny, g.Nz), 20.0, np.float32)


def _draw(orc, seed, it, sa, cell):
    out = (C.c_float * 3)()
    orc.load().orc_uniform3(seed, it, 64 + sa, 2, int(cell), out)
    return np.float32(out[0])


def test_creation_order_and_positions(pi, orc):
    """A huge flux through the top face of a 5x4x3 grid: one particle per top
    cell (20), the k-th cell in column-major order in the k-th free slot (after
    the 7 initial particles), at (i − 0.4, j − 0.4, 0.6) in grid units; no flow
    and no walk, so they stay there."""
    g = pi.RectilinearGrid(size=(5, 4, 3), x=(0, 5e3), y=(0, 4e3), z=(0, -30))
    spec, om = _model(pi, orc, g, N=(7,), cap=40)
    om.set_particle_bc(0, "top", np.float32(1e6))
    parf, temp = _forcing(g)
    om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
    ac = om.abiotic_field(0, "ac")
    assert ac.sum() == 7 + 20 and ac[:27].all() and not ac[27:].any()
    x, y, z = (om.abiotic_field(0, n) for n in "xyz")
    k = np.arange(20)
    assert np.array_equal(x[7:27], (k % 5 + 1).astype(np.float32) - np.float32(0.4))
    assert np.array_equal(y[7:27], (k // 5 + 1).astype(np.float32) - np.float32(0.4))
    assert np.all(z[7:27] == np.float32(1) - np.float32(0.4))
    assert np.array_equal(om.abiotic_field(0, "xi")[7:27], k % 5 + 1 + 2)


@pytest.mark.parametrize("face", ["top", "bottom", "east", "west", "north", "south"])
def test_creation_decision_per_cell(pi, orc, face):
    """A flux such that bc·ΔT·A / (sz_min·Nsuper) ≈ 0.4 on one face of a grid
    with Δx ≠ Δy ≠ Δz: the set of cells that create a particle is exactly
    those whose shared-stream draw u has isless(u·sz_min·Nsuper, bc·ΔT·A)
    (A = Az, Ax or Ay of the face), and the flux of a second species on the
    same face draws from its own sub-stream."""
    Nx, Ny, Nz, H = 9, 7, 5, 2
    g = pi.RectilinearGrid(size=(Nx, Ny, Nz), x=(0, 4.0 * Nx), y=(0, 2.0 * Ny), z=(0, -1.0 * Nz),
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded))
    spec, om = _model(pi, orc, g, N=(0, 0), cap=600)
    A = {"top": 8.0, "bottom": 8.0, "east": 2.0, "west": 2.0, "north": 4.0, "south": 4.0}[face]
    sz = float(spec.abiotic_matrix[0, 3]) * float(spec.abiotic_matrix[0, 0])
    dt = 60.0
    bc = np.float32(0.4 * sz / (dt * A))
    om.set_particle_bc(0, face, bc)
    om.set_particle_bc(1, face, bc)
    parf, temp = _forcing(g)
    om.update(dt, 1, vel=_still(g), parf=parf, temp=temp)
    cells = {"top": [(i, j, 1) for j in range(1, Ny + 1) for i in range(1, Nx + 1)],
             "bottom": [(i, j, Nz) for j in range(1, Ny + 1) for i in range(1, Nx + 1)],
             "east": [(Nx, j, k) for k in range(1, Nz + 1) for j in range(1, Ny + 1)],
             "west": [(1, j, k) for k in range(1, Nz + 1) for j in range(1, Ny + 1)],
             "north": [(i, Ny, k) for k in range(1, Nz + 1) for i in range(1, Nx + 1)],
             "south": [(i, 1, k) for k in range(1, Nz + 1) for i in range(1, Nx + 1)]}[face]
    f32 = np.float32
    rhs = (bc * f32(dt)) * f32(A)
    made = []
    for sa in range(2):
        want = []
        for (i, j, k) in cells:
            c = (i + H - 1) + (Nx + 2 * H) * ((j + H - 1) + (Ny + 2 * H) * (k + H - 1))
            u = _draw(orc, spec.seed, 1, sa, c)
            if (u * f32(spec.abiotic_matrix[sa, 3])) * f32(spec.abiotic_matrix[sa, 0]) < rhs:
                want.append((c, (i, j, k)))
        want.sort()
        ac = om.abiotic_field(sa, "ac")
        x, y, z = (om.abiotic_field(sa, n) for n in "xyz")
        got = sorted(zip(x[ac], y[ac], z[ac]))
        exp = sorted((f32(i) - f32(0.4), f32(j) - f32(0.4), f32(k) - f32(0.4)) for _, (i, j, k) in want)
        assert got == exp and 0.2 * len(cells) < len(exp) < 0.6 * len(cells), (sa, len(exp))
        made.append(got)
    assert made[0] != made[1]


def test_later_face_decides_at_an_edge(pi, orc):
    """Top creates everywhere, east never (bc = 0): on the top-east edge the east
    kernel runs after the top one, so those Ny cells create nothing."""
    g = pi.RectilinearGrid(size=(6, 5, 3), x=(0, 6e3), y=(0, 5e3), z=(0, -30),
                           topology=(pi.Bounded, pi.Bounded, pi.Bounded))
    spec, om = _model(pi, orc, g, N=(0,), cap=100)
    om.set_particle_bc(0, "top", np.float32(1e6))
    om.set_particle_bc(0, "east", np.float32(0))
    parf, temp = _forcing(g)
    om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
    ac = om.abiotic_field(0, "ac")
    x = om.abiotic_field(0, "x")
    assert ac.sum() == 6 * 5 - 5 and np.all(x[ac] < 5.0)


def test_series_flux_and_capacity(pi, orc):
    """A (Nx, Ny, nt) flux read at the model's iteration: zero in step 1, huge in
    step 2 (every top cell creates one), zero in step 3; a step that would create
    more particles than free slots is the capacity error."""
    g = pi.RectilinearGrid(size=(4, 3, 2), x=(0, 4e3), y=(0, 3e3), z=(0, -20))
    spec, om = _model(pi, orc, g, N=(0,), cap=20)
    series = np.zeros((4, 3, 5), np.float32)
    series[:, :, 1] = 1e6
    series[:, :, 3] = 1e6
    om.set_particle_bc(0, "top", series)
    parf, temp = _forcing(g)
    counts = []
    for _ in range(3):
        om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
        counts.append(int(om.abiotic_field(0, "ac").sum()))
    assert counts == [0, 12, 12]
    with pytest.raises(pi.CapacityError):
        om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)


def test_transport_and_walk(pi, orc):
    """Uniform u: step 2 moves every particle by u ΔT/Δx (RK4 of a constant field,
    vel₀ = vel₁); with κhP > 0 and no flow, the walk of step 2 adds
    (2(U − ½)·√(κ ΔT))/Δ with U the draw of the particle species' sub-stream,
    bit for bit."""
    g = pi.RectilinearGrid(size=(8, 8, 2), x=(0, 8e3), y=(0, 8e3), z=(0, -20))
    spec, om = _model(pi, orc, g, N=(200,), cap=300)
    u = np.full((8, 8, 2), 2.0, np.float32)
    v = np.zeros((8, 8, 2), np.float32)
    w = np.zeros((8, 8, 3), np.float32)
    parf, temp = _forcing(g)
    om.update(100.0, 1, vel=(u, v, w), parf=parf, temp=temp)
    x1 = om.abiotic_field(0, "x").astype(np.float64)
    om.update(100.0, 1, vel=(u, v, w), parf=parf, temp=temp)
    x2 = om.abiotic_field(0, "x").astype(np.float64)
    ac = om.abiotic_field(0, "ac")
    assert np.allclose(np.mod(x2 - x1, 8.0)[ac], 0.2, atol=2e-6)
    kappa = 40.0
    spec, om = _model(pi, orc, g, N=(200,), cap=300, bgc={"κhP": kappa}, seed=9)
    om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
    xa, ya = om.abiotic_field(0, "x"), om.abiotic_field(0, "y")
    om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
    xb, yb = om.abiotic_field(0, "x"), om.abiotic_field(0, "y")
    f32 = np.float32
    sq = np.sqrt(f32(kappa) * f32(60.0), dtype=np.float32)
    lib = orc.load()
    n = 0
    for s in np.nonzero(om.abiotic_field(0, "ac"))[0][:50]:
        out = (C.c_float * 3)()
        lib.orc_uniform3(spec.seed, 2, 64, 0, int(s), out)
        ex = xa[s] + ((f32(out[0]) - f32(0.5)) * f32(2) * sq) / f32(1000)
        ey = ya[s] + ((f32(out[1]) - f32(0.5)) * f32(2) * sq) / f32(1000)
        if 0 < ex < 8 and 0 < ey < 8:
            assert xb[s] == ex and yb[s] == ey, s
            n += 1
    assert n > 30


def test_num_counts_active_particles(pi, orc):
    """num accumulates, every step, one per active particle at its cell
    (find_inds! after the move), and pi_reset_diags zeroes it."""
    g = pi.RectilinearGrid(size=(6, 6, 2), x=(0, 6e3), y=(0, 6e3), z=(0, -20))
    spec, om = _model(pi, orc, g, N=(150,), cap=300, bgc={"κhP": 30.0})
    om.set_particle_bc(0, "west", np.float32(1e6))
    parf, temp = _forcing(g)
    acc = np.zeros(g.padded_shape)
    for _ in range(3):
        om.update(60.0, 1, vel=_still(g), parf=parf, temp=temp)
        ac = om.abiotic_field(0, "ac")
        xi, yi, zi = (om.abiotic_field(0, f) for f in ("xi", "yi", "zi"))
        np.add.at(acc, (xi[ac] - 1, yi[ac] - 1, zi[ac] - 1), 1.0)
    assert np.array_equal(om.abiotic_diag(0), acc.astype(np.float32))
    assert om.abiotic_field(0, "ac").sum() == 150 + 3 * 12
    om.reset_diags()
    assert not om.abiotic_diag(0).any()
