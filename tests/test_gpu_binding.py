"""The Julia binding's call sequence, replayed through ctypes.

A `HIPArch` extension of the reference binds libpi_hip.so with `ccall`: on the
model's first step it creates the model from the grid and parameters and
uploads the state it built on the host (StructArray columns with Int64 indices
and Bool activity, padded column-major tracers); every step of `update!`
forwards the velocity / PAR / temperature slices whose index changed (the
interior of the padded vel₁ as vel_copy! wrote it, N+1 faces on bounded axes,
Simulation/utils.jl:5-34; update.jl:18-26), calls pi_step and pi_sync; before
an output writer runs it downloads the host mirrors and zeroes the device
diagnostics at the diagnostics interval (write_outputs.jl:186-201).

This test declares the prototypes itself, from the C header, and drives the
library with arrays in that layout only; the oracle runs the same update!
literally.  Bar: every downloaded array identical, at every writer call.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
VP = C.c_void_p


class GridDesc(C.Structure):          # pi_grid_desc
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
                ("Hx", C.c_int32), ("Hy", C.c_int32), ("Hz", C.c_int32),
                ("tx", C.c_int32), ("ty", C.c_int32), ("tz", C.c_int32),
                ("dx", C.c_float), ("dy", C.c_float), ("dzC", fp), ("dzF", fp), ("landmask", u8p)]


class ModelDesc(C.Structure):         # pi_model_desc
    _fields_ = [("grid", GridDesc), ("n_species", C.c_int32), ("capacity", C.c_int64),
                ("params", fp), ("bgc", fp), ("seed", C.c_uint64),
                ("n_diag", C.c_int32), ("diag", C.POINTER(C.c_int32)), ("mode", C.c_int32),
                ("metrics", C.c_void_p),   # NULL: RectilinearGrid
                ("n_abiotic", C.c_int32), ("abiotic_params", C.c_void_p), ("n_interactions", C.c_int32),
                ("interactions", C.c_void_p), ("n_releases", C.c_int32), ("releases", C.c_void_p),
                ("max_candidates", C.c_int32)]


FLOAT_COLS = ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "gen", "age", "graz", "mort", "dvid")


class IndividualsHost(C.Structure):   # pi_individuals_host (no extra columns in QuotaMode)
    _fields_ = ([(n, fp) for n in FLOAT_COLS] + [("xi", i64p), ("yi", i64p), ("zi", i64p), ("ac", u8p)] +
                [("ex", fp * 8)])


def _bind(path):
    """The prototypes of include/pi_hip.h that the binding calls."""
    lib = C.CDLL(path)
    protos = {
        "pi_init": [C.c_int32], "pi_last_error": [],
        "pi_model_create": [C.POINTER(ModelDesc), C.POINTER(VP)], "pi_model_destroy": [VP],
        "pi_upload_individuals": [VP, C.c_int32, C.POINTER(IndividualsHost)],
        "pi_upload_tracer": [VP, C.c_int32, fp],
        "pi_set_velocity": [VP, fp, fp, fp], "pi_set_parf": [VP, fp], "pi_set_temperature": [VP, fp],
        "pi_step": [VP, C.c_float], "pi_sync": [VP],
        "pi_download_field": [VP, C.c_int32, C.c_int32, VP, C.c_int64],
        "pi_download_tracer": [VP, C.c_int32, fp],
        "pi_download_grid": [VP, C.c_int32, fp, C.c_int64],
        "pi_download_diag": [VP, C.c_int32, C.c_int32, fp],
        "pi_download_diag_tracer": [VP, C.c_int32, fp],
        "pi_reset_diags": [VP],
    }
    for name, args in protos.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_char_p if name == "pi_last_error" else C.c_int
    return lib


def _ok(lib, st):
    assert st == 0, f"status {st}: {lib.pi_last_error().decode()}"


def _p(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t))


class Binding:
    """HIPArch's handle!, forward_inputs!, TimeStep! and sync_host! over ctypes."""

    def __init__(self, lib, g, params, bgc, seed, cap, columns, tracers, diag_tracers):
        self.lib, self.g, self.cap, self.nsp = lib, g, cap, len(columns)
        self.keep = [np.ascontiguousarray(g["dzC"], np.float32), np.ascontiguousarray(g["dzF"], np.float32),
                     np.asfortranarray(g["landmask"]).view(np.uint8),          # Julia Bool = one byte
                     np.ascontiguousarray(params, np.float32), np.ascontiguousarray(bgc, np.float32),
                     np.ascontiguousarray(diag_tracers, np.int32)]
        dzC, dzF, lm, prm, bg, dtr = self.keep
        gd = GridDesc(*g["N"], *g["H"], *g["topo"], g["dx"], g["dy"], _p(dzC), _p(dzF), _p(lm, C.c_uint8))
        d = ModelDesc(gd, self.nsp, cap, _p(prm), _p(bg), seed, len(dtr),
                      _p(dtr, C.c_int32) if len(dtr) else None, 0, None,    # PI_MODE_QUOTA, no metrics
                      0, None, 0)                                            # no abiotic particles
        self.h = VP()
        _ok(lib, lib.pi_model_create(C.byref(d), C.byref(self.h)))
        for s, col in enumerate(columns):
            cols = [np.ascontiguousarray(col[f], np.float32) for f in FLOAT_COLS]
            ints = [np.ascontiguousarray(col[f], np.int64) for f in ("xi", "yi", "zi")]
            ac = np.ascontiguousarray(col["ac"], np.bool_).view(np.uint8)
            hv = IndividualsHost(*[_p(c) for c in cols], *[_p(c, C.c_int64) for c in ints], _p(ac, C.c_uint8))
            _ok(lib, lib.pi_upload_individuals(self.h, s, C.byref(hv)))
        for q, t in enumerate(tracers):
            _ok(lib, lib.pi_upload_tracer(self.h, q, _p(np.asfortranarray(t, np.float32))))
        self.staged = [-1, -1, -1]
        self.diag_tracers = list(diag_tracers)

    def forward_inputs(self, idx, vel1, parf, temp):
        g, lib = self.g, self.lib
        (Nx, Ny, Nz), (Hx, Hy, Hz) = g["N"], g["H"]
        ex = [1 if t == 1 else 0 for t in g["topo"]]
        if idx[0] != self.staged[0]:
            u = np.asfortranarray(vel1["u"][Hx:Hx + Nx + ex[0], Hy:Hy + Ny, Hz:Hz + Nz])
            v = np.asfortranarray(vel1["v"][Hx:Hx + Nx, Hy:Hy + Ny + ex[1], Hz:Hz + Nz])
            w = np.asfortranarray(vel1["w"][Hx:Hx + Nx, Hy:Hy + Ny, Hz:Hz + Nz + ex[2]])
            _ok(lib, lib.pi_set_velocity(self.h, _p(u), _p(v), _p(w)))
        if idx[1] != self.staged[1]:
            _ok(lib, lib.pi_set_parf(self.h, _p(np.asfortranarray(parf, np.float32))))
        if idx[2] != self.staged[2]:
            t = np.asfortranarray(temp[Hx:Hx + Nx, Hy:Hy + Ny, Hz:Hz + Nz])
            _ok(lib, lib.pi_set_temperature(self.h, _p(t)))
        self.staged = list(idx)

    def time_step(self, dt):
        _ok(self.lib, self.lib.pi_step(self.h, C.c_float(dt)))
        _ok(self.lib, self.lib.pi_sync(self.h))

    def sync_host(self):
        """The host mirrors the output writers read."""
        lib, g = self.lib, self.g
        shape = tuple(n + 2 * h for n, h in zip(g["N"], g["H"]))
        out = {}
        for s in range(self.nsp):
            for f, name in enumerate(FLOAT_COLS):
                a = np.empty(self.cap, np.float32)
                _ok(lib, lib.pi_download_field(self.h, s, f, a.ctypes.data_as(VP), self.cap))
                out[f"sp{s}.{name}"] = a
            for f, name in ((14, "xi"), (15, "yi"), (16, "zi")):
                a = np.empty(self.cap, np.int64)
                _ok(lib, lib.pi_download_field(self.h, s, f, a.ctypes.data_as(VP), self.cap))
                out[f"sp{s}.{name}"] = a
            a = np.empty(self.cap, np.uint8)
            _ok(lib, lib.pi_download_field(self.h, s, 17, a.ctypes.data_as(VP), self.cap))
            out[f"sp{s}.ac"] = a != 0
            for k, name in enumerate(("num", "graz", "mort", "dvid")):
                a = np.empty(shape, np.float32, order="F")
                _ok(lib, lib.pi_download_diag(self.h, s, k, _p(a)))
                out[f"diag sp{s}.{name}"] = a
        for q in range(14):
            a = np.empty(shape, np.float32, order="F")
            _ok(lib, lib.pi_download_tracer(self.h, q, _p(a)))
            out[f"tracer {q}"] = a
        for q in self.diag_tracers:
            a = np.empty(shape, np.float32, order="F")
            _ok(lib, lib.pi_download_diag_tracer(self.h, q, _p(a)))
            out[f"diag tracer {q}"] = a
        a = np.empty(shape, np.float32, order="F")
        _ok(lib, lib.pi_download_grid(self.h, 9, _p(a), a.size))        # diags.tracer.PAR
        out["diag PAR"] = a
        for gid, name in ((6, "u"), (7, "v"), (8, "w")):                  # vel₁ with its halos
            a = np.empty(shape, np.float32, order="F")
            _ok(lib, lib.pi_download_grid(self.h, gid, _p(a), a.size))
            out[name] = a
        return out

    def reset_diags(self):
        _ok(self.lib, self.lib.pi_reset_diags(self.h))

    def close(self):
        if self.h:
            self.lib.pi_model_destroy(self.h)
            self.h = VP()


def _oracle_mirror(om, nsp, diag_tracers):
    out = {}
    for s in range(nsp):
        for name in FLOAT_COLS + ("xi", "yi", "zi", "ac"):
            out[f"sp{s}.{name}"] = om.field(s, name)
        for name in ("num", "graz", "mort", "dvid"):
            out[f"diag sp{s}.{name}"] = om.diag(s, name)
    for q in range(14):
        out[f"tracer {q}"] = om.tracer(q)
    for q in diag_tracers:
        out[f"diag tracer {q}"] = om.diag_tracer(q)
    out["diag PAR"] = om.grid_field("diag_PAR")
    for name in ("u", "v", "w"):
        out[name] = om.grid_field(name)
    return out


def _padded_vel(g, u, v, w):
    """vel₁ as vel_copy! leaves it before fill_halo: interior written, halos zero."""
    (Nx, Ny, Nz), (Hx, Hy, Hz) = g["N"], g["H"]
    shape = (Nx + 2 * Hx, Ny + 2 * Hy, Nz + 2 * Hz)
    vel = {k: np.zeros(shape, np.float32, order="F") for k in "uvw"}
    vel["u"][Hx:Hx + u.shape[0], Hy:Hy + Ny, Hz:Hz + Nz] = u
    vel["v"][Hx:Hx + Nx, Hy:Hy + v.shape[1], Hz:Hz + Nz] = v
    vel["w"][Hx:Hx + Nx, Hy:Hy + Ny, Hz:Hz + w.shape[2]] = w
    return vel


def test_binding_call_sequence(pi, orc):
    from parity import same, first_diff

    Nx, Ny, Nz = 20, 16, 4
    lm = np.ones((Nx, Ny, Nz), bool)
    lm[14:18, 3:9, :] = False                        # an island
    lm[2:5, 10:12, 2:] = False                       # a seamount
    grid = pi.RectilinearGrid(size=(Nx, Ny, Nz), x=(0, 20e3), y=(0, 16e3), z=[0.0, -3.0, -7.0, -12.0, -20.0],
                              topology=(pi.Bounded, pi.Periodic, pi.Bounded), landmask=lm)
    spec = pi.ModelSpec(grid, max_individuals=5000, seed=31,
                        phyto=pi.phyto_setup({"dvid_P": [2e-3, 1e-3], "mort_P": [5e-4, 5e-4],
                                              "grz_P": [2e-4, 0.0]}, [1800, 1200], 2),
                        bgc_params={"κhP": 2.0, "κvP": 1e-4, "κh": 20.0, "κv": 1e-4})
    g = dict(N=(Nx, Ny, Nz), H=(grid.Hx, grid.Hy, grid.Hz), topo=tuple(grid.topology_ids),
             dx=np.float32(grid.Δx), dy=np.float32(grid.Δy), dzC=grid.dzC, dzF=grid.dzF,
             landmask=np.asfortranarray(grid.landmask, bool))
    assert g["landmask"].shape == grid.padded_shape

    # the StructArray columns and padded tracers generate_individuals / generate_tracers built
    columns = []
    for s in range(spec.n_species):
        src = spec.initial_individuals[s]
        col = {f: np.asarray(src[f], np.float32) for f in FLOAT_COLS}
        col.update({f: np.asarray(src[f], np.int64) for f in ("xi", "yi", "zi")})
        col["ac"] = np.asarray(src["ac"]).astype(np.bool_)
        columns.append(col)
    tracers = [np.asfortranarray(t, np.float32) for t in spec.initial_tracers]

    # inputs of update!: two velocity slices, hourly PAR slices, two temperature slices
    dt, n, vel_dt, par_dt, temp_dt = 60.0, 14, 300.0, 180.0, 420.0
    rng = np.random.default_rng(9)
    vslices = []
    for _ in range(2):
        vslices.append(((rng.standard_normal((Nx + 1, Ny, Nz)) * 2).astype(np.float32),
                        (rng.standard_normal((Nx, Ny, Nz)) * 2).astype(np.float32),
                        (rng.standard_normal((Nx, Ny, Nz + 1)) * 1e-3).astype(np.float32)))
    nt_par = int(n * dt / par_dt) + 2
    PARF = (300.0 + 400.0 * rng.random((Nx, Ny, nt_par))).astype(np.float32)
    temps = [(15.0 + 10.0 * rng.random((Nx, Ny, Nz))).astype(np.float32) for _ in range(3)]
    diag_tracers = [pi._abi.TRACER_ID["NH4"], pi._abi.TRACER_ID["DOC"]]
    writer_every, diag_interval = 3, 6

    lib = _bind(pi._native.lib_path())
    _ok(lib, lib.pi_init(0))                         # isfunctional(HIPArch())
    b = Binding(lib, g, spec.species_params, spec.bgc_vector, spec.seed, spec.max_individuals,
                columns, tracers, diag_tracers)
    om = orc.OracleModel(spec, diag_tracers=diag_tracers)
    H = g["H"]
    try:
        t = np.float32(0.0)
        for i in range(1, n + 1):
            tv, tp, tt = (int(np.floor(t / np.float32(s))) for s in (vel_dt, par_dt, temp_dt))
            vel1 = _padded_vel(g, *vslices[tv % 2])
            temp_pad = np.zeros(grid.padded_shape, np.float32, order="F")
            temp_pad[H[0]:H[0] + Nx, H[1]:H[1] + Ny, H[2]:H[2] + Nz] = temps[tt % 3]
            b.forward_inputs((tv, tp, tt), vel1, PARF[:, :, tp], temp_pad)
            b.time_step(dt)
            # the oracle's update!: every slice copied in before every step
            om.set_velocity(*vslices[tv % 2])
            om.set_parf(PARF[:, :, tp])
            om.set_temperature(temps[tt % 3])
            om.step(np.float32(dt))
            t = np.float32(np.float32(i) * np.float32(dt))
            if i % writer_every == 0:
                got, want = b.sync_host(), _oracle_mirror(om, spec.n_species, diag_tracers)
                bad = [f"step {i} {k}: {first_diff(got[k], want[k])}" for k in want if not same(got[k], want[k])]
                assert not bad, "\n".join(bad[:20])
                if i % diag_interval == 0:
                    b.reset_diags()
                    om.reset_diags()
        assert b.staged[0] == 2 and b.staged[2] == 1      # the slices did change mid-run
        n_live = [int(b.sync_host()[f"sp{s}.ac"].sum()) for s in range(spec.n_species)]
        assert n_live == [om.count_active(s) for s in range(spec.n_species)]
    finally:
        b.close()
