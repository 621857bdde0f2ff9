"""Pins of the oracle's tracer flux boundary conditions (set_bc!,
HaloRegions/boundary_conditions.jl:12-40; apply_bcs!, HaloRegions/apply_bcs.jl).

Every case runs whole oracle steps in which nothing but the boundary
conditions changes a tracer: zero velocity, zero diffusivities, no forcing
rates, no individuals, tracers starting at zero.  A tracer then gains exactly
the apply_bcs! term of each step, Gc ± bc·ΔT·A/V at the cells next to a face,
which the tests compute from the grid spacings alone:

  reference   test/field_test.jl:108-122: bc = 1 on the west face of a
              4x4x4 grid with Δ = 8 gives Gc = 1/8 on the first interior x plane
  faces       each face on its own tracer, with Δx = 8, Δy = 4, Δz = 2: the
              area over volume of west/east is 1/Δx, of south/north 1/Δy, of
              bottom/top 1/Δz; west/south/bottom add, east/north/top subtract;
              face arrays indexed (j, k), (i, k), (i, j)
  order       all six faces on the one cell of a 1x1x1 grid, with values whose
              float sum depends on the order west, east, south, north, bottom, top
  time        a (face..., nt) array read at the model's iteration (step n reads
              bc[:, :, n]), and a step past nt refused
"""
import numpy as np
import pytest

ZERO_RATES = {"kDOC": 0.0, "Nit": 0.0, "kDON": 0.0, "kDOP": 0.0, "kPOC": 0.0, "kPON": 0.0, "kPOP": 0.0,
              "κh": 0.0, "κv": 0.0, "κhP": 0.0, "κvP": 0.0, "w_sink_org": 0.0, "w_sink_inorg": 0.0,
              "lambda_POC": 0.0, "lambda_min": 0.0, "kdiss": 0.0, "lambda_dust": 0.0, "lambda_Fe": 0.0}


def _quiet_model(pi, orc, size, x, y, z, topology=None):
    kw = {} if topology is None else {"topology": topology}
    g = pi.RectilinearGrid(size=size, x=x, y=y, z=z, **kw)
    spec = pi.ModelSpec(g, max_individuals=4, phyto=pi.phyto_setup(None, [0], 1), bgc_params=dict(ZERO_RATES))
    om = orc.OracleModel(spec)
    zero = np.zeros(g.padded_shape, np.float32, order="F")
    for q in range(14):
        om.upload_tracer(q, zero)
    return g, om


def _interior(g, c):
    return c[g.Hx:g.Hx + g.Nx, g.Hy:g.Hy + g.Ny, g.Hz:g.Hz + g.Nz]


def test_field_test_west_flux(pi, orc):
    """test/field_test.jl:108-122: west = ones(4,4,10), ΔT = 1 on a 4x4x4 grid
    with Δx = Δy = Δz = 8: Gc = 1/8 on the first interior x plane, 0 elsewhere.
    Ten steps read bc[:, :, 1..10] and add 1/8 each."""
    g, om = _quiet_model(pi, orc, (4, 4, 4), (0, 32), (0, 32), (0, -32))
    om.set_tracer_bc(pi._abi.TRACER_ID["DIC"], "west", np.ones((4, 4, 10), np.float32))
    om.step(1.0)
    d = _interior(g, om.tracer(0))
    assert np.all(d[0] == np.float32(1 / 8)) and np.all(d[1:] == 0)
    for _ in range(9):
        om.step(1.0)
    d = _interior(g, om.tracer(0))
    assert np.all(d[0] == np.float32(10 / 8)) and np.all(d[1:] == 0)
    for q in range(1, 14):
        assert not np.any(om.tracer(q))


def test_each_face_area_sign_and_index(pi, orc):
    """One face per tracer on a 3x5x2 grid with Δx = 8, Δy = 4, Δz = 2, ΔT = 1.
    West (NH4) adds bc[j, k]/8 at i = 1, east (NO3) subtracts 3/8 at i = Nx,
    south (PO4) adds bc[i, k]/4 at j = 1, north (DOC) subtracts 5/4 at j = Ny,
    bottom (DON) adds bc[i, j]/2 at the deepest level k = Nz, top (DOP)
    subtracts 7/2 at k = 1."""
    Nx, Ny, Nz = 3, 5, 2
    g, om = _quiet_model(pi, orc, (Nx, Ny, Nz), (0, 8 * Nx), (0, 4 * Ny), (0, -2 * Nz),
                         topology=(pi.Bounded, pi.Bounded, pi.Bounded))
    T = pi._abi.TRACER_ID
    jj, kk = np.meshgrid(np.arange(1, Ny + 1), np.arange(1, Nz + 1), indexing="ij")
    west = (10 * jj + kk).astype(np.float32)                     # (Ny, Nz)
    ii, kk2 = np.meshgrid(np.arange(1, Nx + 1), np.arange(1, Nz + 1), indexing="ij")
    south = (100 * ii + kk2).astype(np.float32)                  # (Nx, Nz)
    ii3, jj3 = np.meshgrid(np.arange(1, Nx + 1), np.arange(1, Ny + 1), indexing="ij")
    bottom = (10 * ii3 + jj3).astype(np.float32)                 # (Nx, Ny)
    om.set_tracer_bc(T["NH4"], "west", west)
    om.set_tracer_bc(T["NO3"], "east", np.float32(3))
    om.set_tracer_bc(T["PO4"], "south", south)
    om.set_tracer_bc(T["DOC"], "north", np.float32(5))
    om.set_tracer_bc(T["DON"], "bottom", bottom)
    om.set_tracer_bc(T["DOP"], "top", np.float32(7))
    om.step(1.0)
    exp = {q: np.zeros((Nx, Ny, Nz), np.float32) for q in ("NH4", "NO3", "PO4", "DOC", "DON", "DOP")}
    exp["NH4"][0] = west / np.float32(8)
    exp["NO3"][Nx - 1] = -np.float32(3) / np.float32(8)
    exp["PO4"][:, 0, :] = south / np.float32(4)
    exp["DOC"][:, Ny - 1, :] = -np.float32(5) / np.float32(4)
    exp["DON"][:, :, Nz - 1] = bottom / np.float32(2)
    exp["DOP"][:, :, 0] = -np.float32(7) / np.float32(2)
    for q, e in exp.items():
        assert np.array_equal(_interior(g, om.tracer(T[q])), e), q
    for q in set(pi.TRACER_NAMES) - set(exp):
        assert not np.any(om.tracer(T[q])), q


def test_six_faces_applied_in_order(pi, orc):
    """All faces of DIC on a 1x1x1 grid (Δ = 1, ΔT = 1, A/V = 1): the cell gains
    ((((((0 + w) - e) + s) - n) + b) - t) in float32.  The values are chosen so
    that the float sum taken in any other order of the six faces differs, but
    for 3 of the 719 other orders."""
    import itertools
    g, om = _quiet_model(pi, orc, (1, 1, 1), (0, 1), (0, 1), (0, -1))
    f32 = np.float32
    faces = ("west", "east", "south", "north", "bottom", "top")
    v = (f32(-3 * 2 ** 26), f32(-7 * 2 ** 25), f32(-1.25), f32(20), f32(1792), f32(28672))
    for pos, val in zip(faces, v):
        om.set_tracer_bc(0, pos, val)
    om.step(1.0)

    def total(order):
        G = f32(0)
        for f in order:
            G = G + v[f] if f % 2 == 0 else G - v[f]
        return G

    G = total(range(6))
    assert sum(total(p) == G for p in itertools.permutations(range(6))) == 4
    assert _interior(g, om.tracer(0))[0, 0, 0] == G


def test_time_dependent_flux_read_at_the_iteration(pi, orc):
    """bc[:, :, t] = t on the south face: step n adds n/Δy, so after n steps the
    cells next to the face hold (1 + ... + n)/Δy; a step past nt is refused."""
    Nx, Ny, Nz, nt = 4, 3, 2, 5
    g, om = _quiet_model(pi, orc, (Nx, Ny, Nz), (0, 4 * Nx), (0, 2 * Ny), (0, -2 * Nz))
    series = np.broadcast_to(np.arange(1, nt + 1, dtype=np.float32), (Nx, Nz, nt)).copy()
    om.set_tracer_bc(pi._abi.TRACER_ID["DIC"], "south", series)
    acc = np.float32(0)
    for n in range(1, nt + 1):
        om.step(1.0)
        acc = acc + np.float32(n) / np.float32(2)
        d = _interior(g, om.tracer(0))
        assert np.all(d[:, 0, :] == acc) and not np.any(d[:, 1:, :]), n
    with pytest.raises(ValueError):
        om.step(1.0)


def test_set_bc_host_surface(pi):
    """set_bc!(model; tracer, pos, bc_value) keeps FT.(bc_value) in
    model.tracers.<tracer>.bc.<pos> (test/field_test.jl:112-117), and
    validate_bcs (boundary_conditions.jl:62-83) rejects an array of the wrong
    face size or length in time."""
    g = pi.RectilinearGrid(size=(4, 4, 4), x=(0, 32), y=(0, 32), z=(0, -32))

    class _Stub:                         # a model without a device: the host mirror only
        pass

    m = _Stub()
    m.tracers = {n: type("F", (), {"bc": pi.BoundaryConditions()})() for n in pi.TRACER_NAMES}
    m._native = type("N", (), {"set_tracer_bc": lambda self, *a: None})()
    pi.set_bc(m, tracer="DIC", pos="west", bc_value=0.1)
    assert m.tracers["DIC"].bc.west == np.float32(0.1)
    pi.set_bc(m, tracer=":DIC", pos=":west", bc_value=np.ones((4, 4)))
    assert m.tracers["DIC"].bc.west.dtype == np.float32 and m.tracers["DIC"].bc.west.shape == (4, 4)
    pi.set_bc(m, tracer="DIC", pos="west", bc_value=np.ones((4, 4, 10)))
    pi.validate_bcs(m.tracers, g, 10)
    with pytest.raises(ValueError):
        pi.validate_bcs(m.tracers, g, 11)
    pi.set_bc(m, tracer="NO3", pos="top", bc_value=np.ones((4, 3)))
    with pytest.raises(ValueError):
        pi.validate_bcs(m.tracers, g, 10)
    with pytest.raises(AssertionError):
        pi.set_bc(m, tracer="Fe", pos="top", bc_value=1.0)
