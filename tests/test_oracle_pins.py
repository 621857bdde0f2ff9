"""Pins of the oracle (the CPU restatement used as the checker) and of the host
logic, against the reference's own tests and against what the mathematics fixes.

Reference-held pins: test/grid_test.jl:5-68 (metrics, exact vectors) and
test/field_test.jl:21-106 (halo identities), test/example_2D_test.jl:19-47
(QuotaMode mass budget).  Published known-answer vectors: Philox4x32-10
(Random123 kat_vectors).  The rest are closed forms / invariants chosen so a
dropped term, a wrong sign or an off-by-one index fails.
"""
import math

import numpy as np
import pytest

from parity import same


# ------------------------------------------------------------ grid (grid_test.jl)
def test_rectilinear_grid_uniform(pi):
    g = pi.RectilinearGrid(size=(4, 6, 2), x=(0, 12), y=(0, 12), z=(0, -8), halo=(2, 2, 2))
    assert (g.Nx, g.Ny, g.Nz) == (4, 6, 2)
    assert g.Δx == 3.0 and g.Δy == 2.0
    assert list(g.dzC) == [4.0] * 6
    assert list(g.dzF) == [4.0] * 6
    assert len(g.xC) == 8 and len(g.yC) == 10 and len(g.zC) == 6
    assert len(g.xF) == 8 and len(g.yF) == 10 and len(g.zF) == 6


def test_rectilinear_grid_stretched(pi):
    g = pi.RectilinearGrid(size=(4, 6, 4), x=(0, 12), y=(0, 12), z=[0.0, -1.0, -3.0, -5.0, -10.0], halo=(2, 2, 2))
    assert list(g.dzC) == [1.0, 1.0, 1.5, 2.0, 3.5, 5.0, 5.0, 5.0]
    assert list(g.dzF) == [1.0, 1.0, 1.0, 2.0, 2.0, 5.0, 5.0, 5.0]
    assert list(g.zF) == [2.0, 1.0, -0.0, -1.0, -3.0, -5.0, -10.0, -15.0]
    assert list(g.zC) == [1.5, 0.5, -0.5, -2.0, -4.0, -7.5, -12.5, -17.5]


def test_rectilinear_areas_volumes(pi):
    g = pi.RectilinearGrid(size=(4, 6, 4), x=(0, 12), y=(0, 12), z=[0.0, -1.0, -3.0, -5.0, -10.0])
    assert g.ΔxC(1, 1, 1) == g.Δx and g.ΔyC(1, 1, 1) == g.Δy and g.ΔzC(1, 1, 1) == g.dzC[0]
    assert g.ΔxF(1, 1, 1) == g.Δx and g.ΔyF(1, 1, 1) == g.Δy and g.ΔzF(1, 1, 1) == g.dzF[0]
    assert g.Ax(1, 1, 1) == g.Δy * g.dzF[0]
    assert g.Ay(1, 1, 1) == g.Δx * g.dzF[0]
    assert g.Az(1, 1, 1) == g.Δx * g.Δy
    assert g.volume(1, 1, 1) == g.Δx * g.Δy * g.dzF[0]


# ------------------------------------------------------- halos (field_test.jl)
def _halo_identities(c, Nx, Ny, Nz, Hx, Hy, Hz, periodic):
    if periodic:
        assert np.array_equal(c[Nx + Hx:Nx + 2 * Hx], c[Hx:2 * Hx])
        assert np.array_equal(c[0:Hx], c[Nx:Nx + Hx])
        assert np.array_equal(c[:, Ny + Hy:Ny + 2 * Hy], c[:, Hy:2 * Hy])
        assert np.array_equal(c[:, 0:Hy], c[:, Ny:Ny + Hy])
    else:
        for i in range(Hx):
            assert np.array_equal(c[Nx + Hx + i], c[Nx + Hx - 1])
            assert np.array_equal(c[i], c[Hx])
        for i in range(Hy):
            assert np.array_equal(c[:, Ny + Hy + i], c[:, Ny + Hy - 1])
            assert np.array_equal(c[:, i], c[:, Hy])
    for i in range(Hz):
        assert np.array_equal(c[:, :, Nz + Hz + i], c[:, :, Nz + Hz - 1])
        assert np.array_equal(c[:, :, i], c[:, :, Hz])


@pytest.mark.parametrize("periodic", [True, False])
def test_fill_halos_host_and_oracle(pi, orc, periodic):
    T = pi.Periodic if periodic else pi.Bounded
    g = pi.RectilinearGrid(size=(4, 6, 2), x=(0, 12), y=(0, 12), z=(0, -8), topology=(T, T, pi.Bounded))
    rng = np.random.default_rng(0)
    c0 = np.asfortranarray(rng.random(g.padded_shape).astype(np.float32))
    c1 = pi.fill_halo_tracer(c0.copy(order="F"), g)
    _halo_identities(c1, 4, 6, 2, 2, 2, 2, periodic)
    import ctypes as C
    c2 = c0.copy(order="F")
    lib = orc.load()
    t = 0 if periodic else 1
    lib.orc_fill_halo(4, 6, 2, 2, 2, 2, t, t, 1, c2.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(c1, c2)
    assert np.array_equal(g.interior(c1), g.interior(c0))


def test_fill_halo_aliasing_small_n(pi, orc):
    """N < H on a periodic axis: each copy reads its source before writing (Q17)."""
    import ctypes as C
    c = np.asfortranarray(np.arange(5 * 5 * 5, dtype=np.float32).reshape((5, 5, 5), order="F"))
    ref = c.copy(order="F")
    # literal statements for x: west c[1:2] = c[2:3]; east c[4:5] = c[3:4] (1-based)
    w = ref[1:3].copy(); ref[0:2] = w
    e = ref[2:4].copy(); ref[3:5] = e
    w = ref[:, 1:3].copy(); ref[:, 0:2] = w
    e = ref[:, 2:4].copy(); ref[:, 3:5] = e
    ref[:, :, 0:2] = ref[:, :, 2:3].copy()
    ref[:, :, 3:5] = ref[:, :, 2:3].copy()
    lib = orc.load()
    lib.orc_fill_halo(1, 1, 1, 2, 2, 2, 0, 0, 1, c.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(c, ref)


# --------------------------------------------------- shared stream (Philox KAT)
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(pi, orc, ctr, key, expect):
    import ctypes as C
    lib = orc.load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox(c, k, o)
    assert tuple(o) == expect
    h = pi.stream.philox4x32_10(*[np.uint32(v) for v in ctr], np.uint32(key[0]), np.uint32(key[1]))
    assert tuple(int(v) for v in h) == expect


def test_stream_host_matches_oracle_and_is_uniform(pi, orc):
    import ctypes as C
    lib = orc.load()
    slots = np.arange(0, 5000, 7)
    a, b, c = pi.stream.uniform3(123456789012, 17, 1, 1, slots)
    out = (C.c_float * 3)()
    for n, s in enumerate(slots[:50]):
        lib.orc_uniform3(123456789012, 17, 1, 1, int(s), out)
        assert (out[0], out[1], out[2]) == (a[n], b[n], c[n])
    u, _, _ = pi.stream.uniform3(5, 3, 0, 0, np.arange(200000))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3
    assert np.all((u * 2 ** 23) == np.round(u * 2 ** 23))          # 23-bit grid like rand(Float32)


# -------------------------------------------------------------- transcendentals
def test_exp_f64_accuracy(orc):
    lib = orc.load()
    xs = np.concatenate([np.linspace(-740, 709, 2001), np.linspace(-1, 1, 2001), [0.0, -0.0, 1e-300, -1e-300]])
    for x in xs:
        e = lib.orc_exp_f64(float(x))
        r = math.exp(float(x))
        assert abs(e - r) <= 2 * math.ulp(r) + 1e-320, (x, e, r)


def test_init_transcendentals_host_mirror_and_accuracy(pi, orc):
    """pi_log_f64 / pi_cos2pi_f64 / pi_init_normal (include/pi_math.h, compiled
    by the oracle's gcc) against their numpy mirrors: bit-identical; and against
    libm: log within 2 ulp, cos(2 pi b) within 2e-15 absolute (the libm
    reference itself carries the rounding of 2 pi b); the normal has
    mean 0 and variance 1."""
    import ctypes as C
    from importlib import import_module
    stream = import_module(pi.__name__ + ".stream")
    lib = orc.load()
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(0, 1, 50000), np.exp(rng.uniform(-700, 700, 50000)),
                        1.0 - np.arange(1, 2000) * 2.0 ** -24, [1.0, 0.5, 2.0, 5e-324, 1e308]])
    y = np.empty_like(x)
    lib.orc_log_f64_v(x.ctypes.data_as(C.POINTER(C.c_double)), y.ctypes.data_as(C.POINTER(C.c_double)), len(x))
    assert np.array_equal(y, stream.log_f64(x))
    ref = np.log(x)
    assert np.all(np.abs(y - ref) <= 2 * np.spacing(np.abs(ref)) + 1e-300)
    b = np.concatenate([rng.uniform(0, 1, 100000), np.arange(0, 4096) / 4096.0]).astype(np.float32).astype(np.float64)
    c = np.empty_like(b)
    lib.orc_cos2pi_f64_v(b.ctypes.data_as(C.POINTER(C.c_double)), c.ctypes.data_as(C.POINTER(C.c_double)), len(b))
    assert np.array_equal(c, stream.cos2pi_f64(b))
    assert np.max(np.abs(c - np.cos(2.0 * np.pi * b))) < 2e-15
    ua = rng.random(200000, dtype=np.float32)
    ub = rng.random(200000, dtype=np.float32)
    z = np.empty(200000, np.float32)
    fp = C.POINTER(C.c_float)
    lib.orc_init_normal_v(ua.ctypes.data_as(fp), ub.ctypes.data_as(fp), z.ctypes.data_as(fp), len(z))
    assert np.array_equal(z, stream.init_normal(ua, ub))
    assert abs(float(z.mean())) < 0.01 and abs(float(z.var()) - 1.0) < 0.01


def test_expf_correctly_rounded_sample(orc):
    lib = orc.load()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-87, 88, 20000), rng.uniform(-1, 1, 20000)]).astype(np.float32)
    bad = 0
    for x in xs:
        e = lib.orc_expf(float(x))
        r = np.float32(math.exp(float(x)))
        bad += int(e != r)
    assert bad == 0


def test_tempfunc_closed_forms(pi, orc):
    import ctypes as C
    lib = orc.load()
    p = pi.params.species_param_matrix(pi.update_phyt_params(pi.phyt_params_default(1)), 1)[0]
    P = p.ctypes.data_as(C.POINTER(C.c_float))
    Topt = float(p[pi._abi.PARAM_NAMES.index("Topt")])
    assert lib.orc_tempfunc(P, Topt) == 1.0                      # k/OGT = 1 at Topt
    r = np.float32(1.0) / np.float32(1.04)                                 # fx = 1: reg = 1/(k+1)
    assert abs(lib.orc_tempfunc_ps(P, Topt) - r) <= 2 * np.spacing(r)     # (e·reg)/e rounds twice
    assert lib.orc_tempfunc(P, Topt + 2.0) == 1.0                 # clamped by min(1, ·) above Topt
    assert 0.0 < lib.orc_tempfunc(P, Topt - 10.0) < 1.0


# -------------------------------------------------------------- dynamics pins
def _static(pi, g, N, cap, params=None, bgc=None, seed=1):
    return pi.ModelSpec(g, max_individuals=cap, seed=seed, phyto=pi.phyto_setup(params, [N], 1), bgc_params=bgc)


def test_uniform_flow_displacement_and_wrap(pi, orc):
    """Constant u: every individual moves by u ΔT/Δx grid units per step, wrapping
    periodically (RK4 of a constant field; trilinear weights sum to one)."""
    g = pi.RectilinearGrid(size=(8, 8, 1), x=(0, 8.0), y=(0, 8.0), z=(0, -1))
    spec = _static(pi, g, 200, 400)
    om = orc.OracleModel(spec)
    u = np.full((8, 8, 1), 0.125, np.float32)
    v = np.full((8, 8, 1), -0.0625, np.float32)
    w = np.zeros((8, 8, 2), np.float32)
    x0, y0 = om.field(0, "x").copy(), om.field(0, "y").copy()
    om.update(4.0, 1, vel=(u, v, w), parf=np.zeros((8, 8), np.float32), temp=np.zeros((8, 8, 1), np.float32))
    om.update(4.0, 1, vel=(u, v, w), parf=np.zeros((8, 8), np.float32), temp=np.zeros((8, 8, 1), np.float32))
    ac = om.field(0, "ac")
    # step 1 uses vel₀ = 0 for the first RK4 stage (vel½ = u/2): displacement (0 + 2·½ + 2·½ + 1)/6 · uΔT = ½ uΔT
    dx = (0.5 + 1.0) * 0.125 * 4.0
    dy = (0.5 + 1.0) * -0.0625 * 4.0
    ex = np.mod(x0 + dx, 8.0)
    ey = np.mod(y0 + dy, 8.0)
    assert np.allclose(om.field(0, "x")[ac], ex[ac], atol=2e-6)
    assert np.allclose(om.field(0, "y")[ac], ey[ac], atol=2e-6)


def test_linear_field_interpolation(pi, orc):
    """u linear in x: dx/dt = a + b x has the exact solution x(t); RK4 with C-grid
    trilinear interpolation reproduces it to roundoff (a wrong face offset fails)."""
    g = pi.RectilinearGrid(size=(16, 4, 1), x=(0, 16.0), y=(0, 4.0), z=(0, -1),
                           topology=(pi.Bounded, pi.Periodic, pi.Bounded))
    spec = _static(pi, g, 300, 600)
    om = orc.OracleModel(spec)
    a, b = 1e-3, 2e-4
    xf = np.arange(17, dtype=np.float64)           # u on x faces 0..16
    u = np.repeat((a + b * xf)[:, None, None], 4, axis=1).astype(np.float32)
    v = np.zeros((16, 4, 1), np.float32)
    w = np.zeros((16, 4, 2), np.float32)
    zp, zt = np.zeros((16, 4), np.float32), np.zeros((16, 4, 1), np.float32)
    om.update(10.0, 1, vel=(u, v, w), parf=zp, temp=zt)   # step 1 starts from vel₀ = 0
    x1 = om.field(0, "x").astype(np.float64)
    om.update(10.0, 5, vel=(u, v, w), parf=zp, temp=zt)
    x6 = om.field(0, "x").astype(np.float64)
    ac = om.field(0, "ac") & (x1 > 0.5) & (x1 < 14.5)
    exact = (x1 + a / b) * np.exp(b * 50.0) - a / b
    assert np.max(np.abs(x6[ac] - exact[ac])) < 5e-5


def test_diffusion_statistics(pi, orc):
    """Zero flow: the walk is U(-1,1)·√(κΔT)/Δ per step (variance κΔT/(3Δ²)), zero mean."""
    g = pi.RectilinearGrid(size=(64, 64, 1), x=(0, 64e3), y=(0, 64e3), z=(0, -10))
    kap, dt = 400.0, 60.0
    spec = _static(pi, g, 20000, 20000, bgc={"κhP": kap, "κvP": 0.0})
    om = orc.OracleModel(spec)
    x0 = om.field(0, "x").astype(np.float64)
    z3 = (np.zeros((64, 64, 1), np.float32), np.zeros((64, 64, 1), np.float32), np.zeros((64, 64, 2), np.float32))
    om.update(dt, 1, vel=z3, parf=np.zeros((64, 64), np.float32), temp=np.zeros((64, 64, 1), np.float32))
    d = om.field(0, "x").astype(np.float64) - x0
    d = d[np.abs(d) < 1.0]
    s2 = kap * dt / 1e6 / 3.0
    assert abs(d.mean()) < 4 * math.sqrt(s2 / len(d))
    assert abs(d.var() / s2 - 1.0) < 0.05


def test_dst3fl_cfl1_is_exact_upwind_shift(pi, orc):
    """u ΔT/Δx = 1: d0 = d1 = 0, Ψ = 0, the scheme is first-order upwind and shifts a
    periodic tracer by exactly one cell (powers of two keep every product exact)."""
    g = pi.RectilinearGrid(size=(16, 4, 1), x=(0, 16.0), y=(0, 4.0), z=(0, -1.0))
    spec = _static(pi, g, 1, 4)
    om = orc.OracleModel(spec)
    rng = np.random.default_rng(0)
    c = np.zeros(g.padded_shape, np.float32, order="F")
    g.interior(c)[...] = np.round(rng.random((16, 4, 1)) * 64) / 64
    pi.fill_halo_tracer(c, g)
    om.upload_tracer(0, c)
    u = np.full((16, 4, 1), 0.25, np.float32)
    z3 = (u, np.zeros((16, 4, 1), np.float32), np.zeros((16, 4, 2), np.float32))
    om.update(4.0, 1, vel=z3, parf=np.zeros((16, 4), np.float32), temp=np.zeros((16, 4, 1), np.float32))
    out = om.tracer(0)
    # first step: vel₁ = u for the tracers (time_step.jl:163) -> exact shift by one cell,
    # plus the DIC remineralisation source kDOC·DOC·ΔT (tracer_forcings.jl:7)
    doc = g.interior(np.asarray(spec.initial_tracers[5]))
    kdoc = np.float32(1 / 30 / 86400)
    expect = np.roll(g.interior(c), 1, axis=0) + np.maximum(np.float32(0), doc) * kdoc * np.float32(4.0)
    assert np.allclose(g.interior(out), expect, rtol=0, atol=1e-6)


def test_dst3fl_conserves_mass_and_bounds(pi, orc):
    """Divergence-free periodic flow: Σ c·V conserved to roundoff, no new extrema."""
    g = pi.RectilinearGrid(size=(24, 24, 1), x=(0, 24e3), y=(0, 24e3), z=(0, -10))
    spec = _static(pi, g, 1, 4, bgc={"kDOC": 0.0, "kPOC": 0.0, "Nit": 0.0, "kDON": 0.0, "kDOP": 0.0,
                                     "kPON": 0.0, "kPOP": 0.0})
    om = orc.OracleModel(spec)
    u, v, w = pi.workloads.cellular_flow(g, U=5.0, L=6e3)
    c = np.zeros(g.padded_shape, np.float32, order="F")
    xx, yy = np.meshgrid(np.arange(24), np.arange(24), indexing="ij")
    g.interior(c)[..., 0] = (1.0 + np.exp(-((xx - 8) ** 2 + (yy - 12) ** 2) / 10.0)).astype(np.float32)
    pi.fill_halo_tracer(c, g)
    om.upload_tracer(pi._abi.TRACER_ID["POP"], c)
    m0 = g.interior(c).astype(np.float64).sum()
    om.update(60.0, 20, vel=(u, v, w), parf=np.zeros((24, 24), np.float32), temp=np.zeros((24, 24, 1), np.float32))
    out = g.interior(om.tracer(pi._abi.TRACER_ID["POP"])).astype(np.float64)
    assert abs(out.sum() - m0) / m0 < 2e-6
    assert out.max() <= g.interior(c).max() + 1e-5 and out.min() >= g.interior(c).min() - 1e-5


def test_tracer_diffusion_fourier_mode(pi, orc):
    """A cosine in x decays by 1 - 4κΔT/Δx² sin²(kΔx/2) per step."""
    n = 32
    g = pi.RectilinearGrid(size=(n, 2, 1), x=(0, float(n)), y=(0, 2.0), z=(0, -1.0))
    kap, dt = 0.05, 1.0
    spec = _static(pi, g, 1, 4, bgc={"κh": kap, "κv": 0.0, "kDOC": 0.0, "kPOC": 0.0, "Nit": 0.0,
                                     "kDON": 0.0, "kDOP": 0.0, "kPON": 0.0, "kPOP": 0.0})
    om = orc.OracleModel(spec)
    k = 2 * np.pi * 3 / n
    c = np.zeros(g.padded_shape, np.float32, order="F")
    xs = np.arange(n) + 0.5
    g.interior(c)[...] = (10.0 + np.cos(k * xs))[:, None, None].astype(np.float32)
    pi.fill_halo_tracer(c, g)
    q = pi._abi.TRACER_ID["POC"]
    om.upload_tracer(q, c)
    z3 = (np.zeros((n, 2, 1), np.float32), np.zeros((n, 2, 1), np.float32), np.zeros((n, 2, 2), np.float32))
    om.update(dt, 1, vel=z3, parf=np.zeros((n, 2), np.float32), temp=np.zeros((n, 2, 1), np.float32))
    out = g.interior(om.tracer(q))[:, 0, 0].astype(np.float64) - 10.0
    fac = 1 - 4 * kap * dt * np.sin(k / 2) ** 2
    assert np.allclose(out, fac * np.cos(k * xs), atol=3e-6)


def _budget(om, g, sp_params, nut, quota, ratio_idx):
    vol = float(g.volume(1, 1, g.Hz + 1))
    m = sum(float(g.interior(om.tracer(t)).astype(np.float64).sum()) * vol for t in nut)
    q = om.field(0, quota).astype(np.float64)
    bm = om.field(0, "Bm").astype(np.float64)
    return m + float(np.sum(q + bm * float(sp_params[ratio_idx])))


def test_phosphorus_and_nitrogen_budgets_close(pi, orc):
    """QuotaMode with light, uptake, grazing, mortality and division and no flow:
    total P (PO4+DOP+POP+Σ(Pq+Bm·R_PC)) and N are conserved (reference:
    test/example_2D_test.jl:19-47 checks total P with atol 1e1)."""
    g = pi.RectilinearGrid(size=(16, 16, 1), x=(0, 32), y=(0, 32), z=(0, -32))
    spec = _static(pi, g, 1024, 4096, params={"dvid_P": [2e-3], "mort_reg": [1.0], "grz_P": [1e-4]})
    om = orc.OracleModel(spec)
    I = pi._abi.TRACER_ID
    p = spec.species_params[0]
    names = pi._abi.PARAM_NAMES
    P0 = _budget(om, g, p, (I["PO4"], I["DOP"], I["POP"]), "Pq", names.index("R_PC"))
    N0 = _budget(om, g, p, (I["NH4"], I["NO3"], I["DON"], I["PON"]), "Nq", names.index("R_NC"))
    z3 = (np.zeros((16, 16, 1), np.float32), np.zeros((16, 16, 1), np.float32), np.zeros((16, 16, 2), np.float32))
    om.update(60.0, 30, vel=z3, parf=np.full((16, 16), 900.0, np.float32), temp=np.full((16, 16, 1), 26.0, np.float32))
    P1 = _budget(om, g, p, (I["PO4"], I["DOP"], I["POP"]), "Pq", names.index("R_PC"))
    N1 = _budget(om, g, p, (I["NH4"], I["NO3"], I["DON"], I["PON"]), "Nq", names.index("R_NC"))
    assert abs(P1 - P0) <= 1e-5 * P0
    assert abs(N1 - N0) <= 1e-5 * N0
    assert om.diag(0, "dvid").sum() > 0 and om.diag(0, "graz").sum() > 0


def test_division_slot_rule_brute_force(pi, orc):
    """k-th surviving divider (ascending slot) -> k-th inactive slot after losses
    (QuotaMode/division_death.jl:28-77): gen + 1 on both, same position, age reset."""
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    cap = 24
    spec = _static(pi, g, 14, cap, params={"mort_reg": [1e-9]})
    ini = spec.initial_individuals[0]
    rng = np.random.default_rng(3)
    ac = np.zeros(cap, bool)
    ac[rng.choice(cap, 14, replace=False)] = True
    ini["ac"] = ac
    for f, p in (("graz", 0.15), ("mort", 0.15), ("dvid", 0.5)):
        ini[f] = ((rng.random(cap) < p) & ac).astype(np.float32)
    om = orc.OracleModel(spec)
    z3 = (np.zeros((4, 4, 1), np.float32), np.zeros((4, 4, 1), np.float32), np.zeros((4, 4, 2), np.float32))
    x0 = om.field(0, "x").copy()
    gen0 = om.field(0, "gen").copy()
    om.update(1e-6, 1, vel=z3, parf=np.zeros((4, 4), np.float32), temp=np.zeros((4, 4, 1), np.float32))
    surv = ac & (ini["graz"] == 0) & (ini["mort"] == 0)
    divs = np.nonzero(surv & (ini["dvid"] == 1))[0]
    free = np.nonzero(~surv)[0]
    expect_ac = surv.copy()
    expect_ac[free[:len(divs)]] = True
    assert np.array_equal(om.field(0, "ac"), expect_ac)
    gen, x, age = om.field(0, "gen"), om.field(0, "x"), om.field(0, "age")
    for k, d in enumerate(divs):
        s = free[k]
        assert gen[s] == gen0[d] + 1 and gen[d] == gen0[d] + 1
        assert x[s] == x[d]
        assert age[s] == age[d]


def test_counts_and_chl_scatter_match_brute_force(pi, orc):
    """pop = individuals per cell and Chl = Σ Chl per cell (slot order) after a step."""
    g = pi.RectilinearGrid(size=(6, 5, 3), x=(0, 6e3), y=(0, 5e3), z=(0, -30))
    spec = _static(pi, g, 500, 800)
    om = orc.OracleModel(spec)
    z3 = (np.zeros((6, 5, 3), np.float32), np.zeros((6, 5, 3), np.float32), np.zeros((6, 5, 4), np.float32))
    chl0 = om.field(0, "Chl").copy()
    om.update(60.0, 1, vel=z3, parf=np.zeros((6, 5), np.float32), temp=np.zeros((6, 5, 3), np.float32))
    xi, yi, zi = (om.field(0, f) for f in ("xi", "yi", "zi"))
    ac0 = spec.initial_individuals[0]["ac"]
    pop = np.zeros(g.padded_shape, np.float32)
    chl = np.zeros(g.padded_shape, np.float32)
    for s in range(800):
        if ac0[s]:
            pop[xi[s] - 1, yi[s] - 1, zi[s] - 1] += 1
            chl[xi[s] - 1, yi[s] - 1, zi[s] - 1] += chl0[s]
    assert np.array_equal(om.grid_field("pop"), pop)
    assert same(om.grid_field("Chl"), chl)
    assert om.grid_field("pop").sum() == ac0.sum()


def test_capacity_error_in_oracle(pi, orc):
    g = pi.RectilinearGrid(size=(4, 4, 1), x=(0, 4e3), y=(0, 4e3), z=(0, -10))
    spec = _static(pi, g, 100, 110, params={"dvid_P": [0.05], "dvid_reg": [1e-3]})
    om = orc.OracleModel(spec)
    z3 = (np.zeros((4, 4, 1), np.float32), np.zeros((4, 4, 1), np.float32), np.zeros((4, 4, 2), np.float32))
    with pytest.raises(pi.CapacityError):
        om.update(60.0, 40, vel=z3, parf=np.zeros((4, 4), np.float32), temp=np.zeros((4, 4, 1), np.float32))


def test_fp64_oracle_close_to_fp32(pi, orc):
    """The FP64 build of the same restatement stays within FP32 roundoff over a few steps."""
    g = pi.RectilinearGrid(size=(8, 8, 2), x=(0, 8e3), y=(0, 8e3), z=(0, -20))
    spec = _static(pi, g, 300, 400)
    a, b = orc.OracleModel(spec, 32), orc.OracleModel(spec, 64)
    u, v, w = pi.workloads.cellular_flow(g, U=1.0, L=2e3)
    for m in (a, b):
        m.update(60.0, 3, vel=(u, v, w), parf=np.full((8, 8), 600.0, np.float32), temp=np.full((8, 8, 2), 25.0, np.float32))
    ac = a.field(0, "ac")
    assert np.array_equal(ac, b.field(0, "ac"))
    for f in ("x", "Bm", "Cq", "Nq", "Pq", "Chl"):
        x32 = a.field(0, f)[ac].astype(np.float64)
        x64 = b.field_f64(0, f)[ac]
        assert np.allclose(x32, x64, rtol=1e-5, atol=1e-12), f


def test_oracle_threads_bit_identical(pi, orc):
    """The oracle's element-wise loops run on OpenMP threads (the CPU baseline
    over all host cores); every accumulation stays serial, so 1 and 4 threads
    give the same bits."""
    from parity import compare_models
    lib = orc.load()
    spec, (u, v, w) = pi.workloads.config_3d(n=12, nz=6, N=900, nsp=2, cap=2400, seed=7, nmodes=10,
                                             overrides={"dvid_P": [2e-3], "mort_P": [1e-3], "grz_P": [2e-5]},
                                             kappa=(1.0, 1e-4))
    g = spec.grid
    runs = []
    for nt in (1, 4):
        assert lib.orc_set_threads(nt) == nt
        om = orc.OracleModel(spec, diag_tracers=[3])
        om.update(60.0, 6, vel=(u * 20, v * 20, w * 20), parf=np.full((g.Nx, g.Ny), 800.0, np.float32),
                  temp=np.full((g.Nx, g.Ny, g.Nz), 24.0, np.float32))
        runs.append(om)
    lib.orc_set_threads(1)
    assert runs[0].diag(0, "dvid").sum() > 0
    bad = compare_models(runs[1], runs[0], 2, diag_tracers=[3])
    assert not bad, "\n".join(bad[:10])
