"""Loader of libpi_hip.so and a thin handle class over its C ABI.

The product has exactly one compute path: the HIP library.  If the library is
missing, every entry point raises — there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import PI_OK, PI_ERR_CAPACITY, PI_ERR_ARG, PI_ERR_UNSUPPORTED

LIB_NAME = "libpi_hip.so"
_lib = None


class CapacityError(ValueError):
    """ArgumentError("number of individual exceeds the capacity …") — QuotaMode/plankton_update.jl:20-22."""


class NativeError(RuntimeError):
    pass


def lib_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


def load_library():
    """Load libpi_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not os.path.exists(p):
        raise NativeError(f"{LIB_NAME} is not built ({p}); run __graft_entry__.build() — "
                          "there is no CPU fallback for the hot path")
    lib = C.CDLL(p)
    _abi.declare(lib, "pi_")
    lib.pi_init.restype = C.c_int
    lib.pi_init.argtypes = [C.c_int32]
    lib.pi_version.restype = C.c_char_p
    lib.pi_model_create.restype = C.c_int
    lib.pi_model_create.argtypes = [C.POINTER(_abi.ModelDesc), C.POINTER(C.c_void_p)]
    lib.pi_run.restype = C.c_int
    lib.pi_run.argtypes = [C.c_void_p, C.c_float, C.c_int64]
    lib.pi_sync.restype = C.c_int
    lib.pi_sync.argtypes = [C.c_void_p]
    lib.pi_set_diag_tracers.restype = C.c_int
    lib.pi_set_diag_tracers.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
    lib.pi_enable_timing.restype = C.c_int
    lib.pi_enable_timing.argtypes = [C.c_void_p, C.c_int32]
    lib.pi_kernel_times.restype = C.c_int
    lib.pi_kernel_times.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_char_p),
                                    C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int32]
    lib.pi_probe_copy.restype = C.c_int
    lib.pi_probe_copy.argtypes = [C.c_int64]
    lib.pi_generate_state.restype = C.c_int
    lib.pi_generate_state.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_float),
                                      C.POINTER(C.c_float), C.POINTER(C.c_int32)]
    lib.pi_footprint.restype = C.c_int
    lib.pi_footprint.argtypes = [C.POINTER(_abi.ModelDesc), C.POINTER(C.c_int64)]
    lib.pi_model_footprint.restype = C.c_int
    lib.pi_model_footprint.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.pi_set_launch_shape.restype = C.c_int
    lib.pi_set_launch_shape.argtypes = [C.c_void_p, C.c_int32, C.c_int64]
    lib.pi_get_launch_shape.restype = C.c_int
    lib.pi_get_launch_shape.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
    for f, a in (("create", [C.POINTER(C.c_void_p)]), ("add", [C.c_void_p, C.c_void_p]),
                 ("remove", [C.c_void_p, C.c_void_p]), ("run", [C.c_void_p, C.c_float, C.c_int64]),
                 ("run_noalloc", [C.c_void_p, C.c_float, C.c_int64]),
                 ("sync", [C.c_void_p]), ("size", [C.c_void_p, C.POINTER(C.c_int32)]),
                 ("graphs", [C.c_void_p, C.c_int32]), ("set_budget", [C.c_void_p, C.c_int64]),
                 ("destroy", [C.c_void_p])):
        fn = getattr(lib, "pi_batch_" + f)
        fn.restype = C.c_int
        fn.argtypes = a
    _lib = lib
    return lib


def check(lib, prefix, status):
    if status == PI_OK:
        return
    msg = getattr(lib, prefix + "last_error")().decode(errors="replace")
    if status == PI_ERR_CAPACITY:
        raise CapacityError(msg)
    if status in (PI_ERR_ARG, PI_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise NativeError(f"status {status}: {msg}")


class Handle:
    """One model instance behind an ABI with prefix `prefix` (shared by the product and
    by test-side wrappers of other implementations of the same ABI)."""

    def __init__(self, lib, prefix, raw, grid, nsp, cap, diag_tracers=()):
        self.lib, self.prefix, self.raw = lib, prefix, C.c_void_p(raw)
        self.grid, self.nsp, self.cap = grid, nsp, cap
        self.diag_tracers = tuple(diag_tracers)

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _call(self, name, *args):
        check(self.lib, self.prefix, self._f(name)(self.raw, *args))

    def close(self):
        if self.raw is not None and self.raw.value:
            self._f("model_destroy")(self.raw)
            self.raw = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- uploads
    def upload_individuals(self, sp, arrays):
        arrays = {_abi.FIELD_ALIAS.get(k, k): v for k, v in arrays.items()}
        keep = {}
        for n in _abi.FLOAT_FIELDS + _abi.EX_FIELDS:
            if arrays.get(n) is not None:
                keep[n] = np.ascontiguousarray(arrays[n], dtype=np.float32)
                assert keep[n].shape == (self.cap,)
        for n in _abi.INT_FIELDS:
            if arrays.get(n) is not None:
                keep[n] = np.ascontiguousarray(arrays[n], dtype=np.int64)
        if arrays.get("ac") is not None:
            keep["ac"] = np.ascontiguousarray(arrays["ac"], dtype=np.uint8)
        h = _abi.individuals_view(keep)
        self._call("upload_individuals", sp, C.byref(h))

    def upload_tracer(self, tid, padded):
        a = np.asfortranarray(padded, dtype=np.float32)
        assert a.shape == self.grid.padded_shape
        self._call("upload_tracer", tid, a.ctypes.data_as(_abi.fp))

    def set_velocity(self, u, v, w):
        u, v, w = (np.asfortranarray(a, dtype=np.float32) for a in (u, v, w))
        g = self.grid
        b = [t == _abi.BOUNDED for t in g.topology_ids]
        want = ((g.Nx + b[0], g.Ny, g.Nz), (g.Nx, g.Ny + b[1], g.Nz), (g.Nx, g.Ny, g.Nz + b[2]))
        for a, s, n in zip((u, v, w), want, "uvw"):
            if a.shape != s:       # the library reads exactly this many values (vel_copy!)
                raise ValueError(f"Dimension mismatch: the size of {n} must be {s}, got {a.shape}.")
        self._call("set_velocity", u.ctypes.data_as(_abi.fp), v.ctypes.data_as(_abi.fp),
                   w.ctypes.data_as(_abi.fp))

    def set_parf(self, parf):
        a = np.asfortranarray(parf, dtype=np.float32)
        assert a.shape == (self.grid.Nx, self.grid.Ny)
        self._call("set_parf", a.ctypes.data_as(_abi.fp))

    def set_temperature(self, temp):
        a = np.asfortranarray(temp, dtype=np.float32)
        assert a.shape == (self.grid.Nx, self.grid.Ny, self.grid.Nz)
        self._call("set_temperature", a.ctypes.data_as(_abi.fp))

    def step(self, dt):
        self._call("step", C.c_float(dt))

    def set_tracer_bc(self, tid, pos, bc_value):
        """Flux boundary condition of tracer `tid` on face `pos` (set_bc!,
        boundary_conditions.jl:12-29): None, a number, a face array (2-D) or a
        face array per iteration (3-D), converted to Float32 like FT.(bc_value)."""
        self._set_bc("set_tracer_bc", tid, pos, bc_value)

    def _set_bc(self, fn, who, pos, bc_value):
        f = _abi.BC_POS[pos]
        if bc_value is None:
            self._call(fn, who, f, _abi.BCK_NONE, None, C.c_int64(0))
            return
        a = np.asarray(bc_value, dtype=np.float32)
        if a.ndim == 0:
            v = np.ascontiguousarray(a.reshape(1))
            self._call(fn, who, f, _abi.BCK_CONST, v.ctypes.data_as(_abi.fp), C.c_int64(0))
            return
        g = self.grid
        face = (g.Ny, g.Nz) if pos in ("west", "east") else (g.Nx, g.Nz) if pos in ("south", "north") else (g.Nx, g.Ny)
        if a.ndim not in (2, 3) or a.shape[:2] != face:
            raise ValueError(f"BC: grid mismatch, size(bc) must equal to {face} for a constant flux boundary "
                             f"condition or {face + ('nΔT',)} for a time-dependent one.")
        v = np.asfortranarray(a)
        kind, nt = (_abi.BCK_FIELD, 0) if a.ndim == 2 else (_abi.BCK_SERIES, a.shape[2])
        self._call(fn, who, f, kind, v.ctypes.data_as(_abi.fp), C.c_int64(nt))

    # ---- abiotic particles
    def upload_abiotic(self, sa, arrays):
        x, y, z = (np.ascontiguousarray(arrays[n], dtype=np.float32) for n in ("x", "y", "z"))
        ac = np.ascontiguousarray(arrays["ac"], dtype=np.uint8)
        assert x.shape == y.shape == z.shape == ac.shape == (self.cap,)
        self._call("upload_abiotic", sa, x.ctypes.data_as(_abi.fp), y.ctypes.data_as(_abi.fp),
                   z.ctypes.data_as(_abi.fp), ac.ctypes.data_as(_abi.u8p))

    def abiotic_field(self, sa, name):
        fid = _abi.ABIO_FIELD_ID[name]
        out = np.empty(self.cap, np.float32 if fid <= 2 else np.int64 if fid <= 5 else np.uint8)
        self._call("download_abiotic", sa, fid, out.ctypes.data_as(C.c_void_p), self.cap)
        return out.astype(bool) if name == "ac" else out

    def abiotic_diag(self, sa):
        out = np.empty(self.grid.padded_shape, np.float32, order="F")
        self._call("download_abiotic_diag", sa, out.ctypes.data_as(_abi.fp))
        return out

    def set_particle_bc(self, sa, pos, bc_value):
        """Particle flux boundary condition of abiotic species `sa` on face
        `pos` (set_bc_particle!, boundary_conditions.jl:31-39)."""
        self._set_bc("set_particle_bc", sa, pos, bc_value)

    # ---- downloads
    def field(self, sp, name):
        name = _abi.FIELD_ALIAS.get(name, name)
        fid = _abi.FIELD_ID[name]
        if name in _abi.INT_FIELDS:
            out = np.empty(self.cap, np.int64)
        elif name == "ac":
            out = np.empty(self.cap, np.uint8)
        else:
            out = np.empty(self.cap, np.float32)
        self._call("download_field", sp, fid, out.ctypes.data_as(C.c_void_p), self.cap)
        return out.astype(bool) if name == "ac" else out

    def tracer(self, tid):
        out = np.empty(self.grid.padded_shape, np.float32, order="F")
        self._call("download_tracer", tid, out.ctypes.data_as(_abi.fp))
        return out

    def grid_field(self, name):
        if name == "PARF":
            out = np.empty((self.grid.Nx, self.grid.Ny), np.float32, order="F")
            wid = _abi.GRID_ID[name]
        elif name.startswith("plk_"):
            out = np.empty(self.grid.padded_shape, np.float32, order="F")
            wid = _abi.PLK_BASE + _abi.TRACER_ID[name[4:]]
        else:
            out = np.empty(self.grid.padded_shape, np.float32, order="F")
            wid = _abi.GRID_ID[name]
        self._call("download_grid", wid, out.ctypes.data_as(_abi.fp), out.size)
        return out

    def diag(self, sp, name):
        out = np.empty(self.grid.padded_shape, np.float32, order="F")
        self._call("download_diag", sp, _abi.DIAG_ID[name], out.ctypes.data_as(_abi.fp))
        return out

    def diag_tracer(self, tid):
        out = np.empty(self.grid.padded_shape, np.float32, order="F")
        self._call("download_diag_tracer", tid, out.ctypes.data_as(_abi.fp))
        return out

    def reset_diags(self):
        self._call("reset_diags")

    def iteration(self):
        v = C.c_int64(0)
        self._call("get_iteration", C.byref(v))
        return v.value

    def count_active(self, sp):
        v = C.c_int64(0)
        self._call("count_active", sp, C.byref(v))
        return v.value

    def individual_steps(self):
        """Σ over the steps taken of the individuals stepped (pi_individual_steps)."""
        v = C.c_int64(0)
        self._call("individual_steps", C.byref(v))
        return v.value


def footprint(grid, nsp, cap, params, bgc, seed=0, diag_tracers=(), mode=_abi.MODE_QUOTA, abiotic_params=None,
              palat=None):
    """Device bytes a model of this shape allocates (pi_footprint: a dry run of
    pi_model_create, no device needed)."""
    lib = load_library()
    desc, keep = _abi.make_desc(grid, nsp, cap, params, bgc, seed, diag_tracers, mode, abiotic_params, palat)
    v = C.c_int64(0)
    check(lib, "pi_", lib.pi_footprint(C.byref(desc), C.byref(v)))
    del keep
    return v.value


class NativeModel(Handle):
    """A model living on the GPU (libpi_hip.so)."""

    def __init__(self, grid, nsp, cap, params, bgc, seed, diag_tracers=(), device=0, mode=_abi.MODE_QUOTA,
                 abiotic_params=None, palat=None):
        lib = load_library()
        check(lib, "pi_", lib.pi_init(device))
        desc, keep = _abi.make_desc(grid, nsp, cap, params, bgc, seed, diag_tracers, mode, abiotic_params, palat)
        raw = C.c_void_p()
        check(lib, "pi_", lib.pi_model_create(C.byref(desc), C.byref(raw)))
        del keep
        super().__init__(lib, "pi_", raw.value, grid, nsp, cap, diag_tracers)
        self.batch = None

    def generate_state(self, N, c0, lower, n):
        """pi_generate_state: the initial state drawn on the device."""
        Na = np.ascontiguousarray(np.array(N, dtype=np.int64))
        assert len(Na) == self.nsp
        c0 = np.ascontiguousarray(c0, dtype=np.float32)
        lower = np.ascontiguousarray(lower, dtype=np.float32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        self._call("generate_state", Na.ctypes.data_as(C.POINTER(C.c_int64)), c0.ctypes.data_as(_abi.fp),
                   lower.ctypes.data_as(_abi.fp), n.ctypes.data_as(C.POINTER(C.c_int32)))

    def footprint(self):
        v = C.c_int64(0)
        self._call("model_footprint", C.byref(v))
        return v.value

    def set_shape(self, **values):
        """pi_set_launch_shape: set launch shapes by name (_abi.SHAPES); the results
        do not depend on them."""
        for name, v in values.items():
            self._call("set_launch_shape", _abi.SHAPES[name][0], C.c_int64(int(v)))

    def shape(self):
        """{name: value} of every launch shape."""
        out = {}
        for name, (i, _) in _abi.SHAPES.items():
            v = C.c_int64(0)
            self._call("get_launch_shape", i, C.byref(v))
            out[name] = v.value
        return out

    def close(self):
        if self.batch is not None:
            self.batch.remove(self)
        super().close()

    def run(self, dt, n):
        self._call("run", C.c_float(dt), C.c_int64(n))

    def sync(self):
        self._call("sync")

    def set_diag_tracers(self, ids):
        a = np.ascontiguousarray(np.array(ids, dtype=np.int32))
        self._call("set_diag_tracers", len(a), a.ctypes.data_as(C.POINTER(C.c_int32)) if len(a) else None)
        self.diag_tracers = tuple(ids)

    def enable_timing(self, on=True):
        self._call("enable_timing", 1 if on else 0)

    def kernel_times(self):
        """{group name: (total ms, launches)} since timing was enabled or last read."""
        cap = 64
        n = C.c_int32(0)
        names = (C.c_char_p * cap)()
        ms = (C.c_float * cap)()
        cnt = (C.c_int32 * cap)()
        self._call("kernel_times", C.byref(n), names, ms, cnt, cap)
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(n.value)}


class NativeBatch:
    """Models stepped together by batched launches (pi_batch_*): one launch per
    kernel of the step covers every member."""

    def __init__(self, device=0):
        self.lib = load_library()
        check(self.lib, "pi_", self.lib.pi_init(device))
        raw = C.c_void_p()
        check(self.lib, "pi_", self.lib.pi_batch_create(C.byref(raw)))
        self.raw = raw
        self.members = []
        self.budget = 16384                   # the library's default block budget

    def add(self, model):
        check(self.lib, "pi_", self.lib.pi_batch_add(self.raw, model.raw))
        model.batch = self
        self.members.append(model)

    def remove(self, model):
        check(self.lib, "pi_", self.lib.pi_batch_remove(self.raw, model.raw))
        model.batch = None
        self.members.remove(model)

    def run(self, dt, n=1):
        check(self.lib, "pi_", self.lib.pi_batch_run(self.raw, C.c_float(dt), C.c_int64(n)))

    def run_noalloc(self, dt, n=1):
        """pi_batch_run_noalloc: the same steps with the launch tables in the
        library's static device scratch (no device allocation, no graphs)."""
        check(self.lib, "pi_", self.lib.pi_batch_run_noalloc(self.raw, C.c_float(dt), C.c_int64(n)))

    def sync(self):
        check(self.lib, "pi_", self.lib.pi_batch_sync(self.raw))

    def set_budget(self, blocks):
        check(self.lib, "pi_", self.lib.pi_batch_set_budget(self.raw, C.c_int64(int(blocks))))
        self.budget = int(blocks)

    def autotune(self, dt, steps=4, reps=3, shapes=None, log=None):
        """The block budget and the launch shapes of every member-shape class,
        chosen by timing real steps of every member (autotune.tune_batch; no
        choice changes any result).  Returns its report."""
        from .autotune import tune_batch
        return tune_batch(self, dt, steps=steps, reps=reps, shapes=shapes, log=log)

    def graphs(self, on):
        check(self.lib, "pi_", self.lib.pi_batch_graphs(self.raw, 1 if on else 0))

    def __len__(self):
        return len(self.members)

    def close(self):
        if self.raw is not None and self.raw.value:
            for m in list(self.members):
                m.batch = None
            self.members = []
            self.lib.pi_batch_destroy(self.raw)
            self.raw = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
