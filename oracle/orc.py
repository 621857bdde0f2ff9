"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of oracle/liboracle_pi.so.

`OracleModel(spec)` builds the CPU restatement from the same host-side ModelSpec
the product uses, so both start from identical bits.  `update(sim_inputs)`
mirrors the reference's update! literally (inputs re-staged every step).
"""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
# PI_ORACLE_LIB: another build of the restatement (oracle/mutation_check.py
# points the pin tests at copies with planted mistakes)
LIB = os.environ.get("PI_ORACLE_LIB") or os.path.join(HERE, "liboracle_pi.so")


def _pkg():
    name = "pi_amd"
    if name in sys.modules:
        return sys.modules[name]
    d = os.path.join(REPO, "planktonindividuals.jl_amd")
    spec = importlib.util.spec_from_file_location(name, os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build as b
            b.build()
        pi = _pkg()
        lib = C.CDLL(LIB)
        pi._abi.declare(lib, "orc_")
        lib.orc_model_create.restype = C.c_int
        lib.orc_model_create.argtypes = [C.POINTER(pi._abi.ModelDesc), C.c_int32, C.POINTER(C.c_void_p)]
        lib.orc_set_diag_tracers.restype = C.c_int
        lib.orc_set_diag_tracers.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        lib.orc_download_field_f64.restype = C.c_int
        lib.orc_download_field_f64.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_int64]
        lib.orc_download_tracer_f64.restype = C.c_int
        lib.orc_download_tracer_f64.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        lib.orc_fill_halo.restype = C.c_int
        lib.orc_fill_halo.argtypes = [C.c_int32] * 9 + [C.POINTER(C.c_float)]
        lib.orc_set_threads.restype = C.c_int32
        lib.orc_set_threads.argtypes = [C.c_int32]
        lib.orc_fill_halo_op.restype = C.c_int
        lib.orc_fill_halo_op.argtypes = [C.c_int32] * 9 + [C.POINTER(C.c_float)]
        lib.orc_expf.restype = C.c_float
        lib.orc_expf.argtypes = [C.c_float]
        lib.orc_exp_f64.restype = C.c_double
        lib.orc_exp_f64.argtypes = [C.c_double]
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        lib.orc_log_f64_v.restype = None
        lib.orc_log_f64_v.argtypes = [dp, dp, C.c_int64]
        lib.orc_cos2pi_f64_v.restype = None
        lib.orc_cos2pi_f64_v.argtypes = [dp, dp, C.c_int64]
        lib.orc_init_normal_v.restype = None
        lib.orc_init_normal_v.argtypes = [fp, fp, fp, C.c_int64]
        lib.orc_uniform3.restype = None
        lib.orc_uniform3.argtypes = [C.c_uint64, C.c_int64, C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_float)]
        lib.orc_philox.restype = None
        lib.orc_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.orc_tempfunc.restype = C.c_float
        lib.orc_tempfunc.argtypes = [C.POINTER(C.c_float), C.c_float]
        lib.orc_tempfunc_ps.restype = C.c_float
        lib.orc_tempfunc_ps.argtypes = [C.POINTER(C.c_float), C.c_float]
        _lib = lib
    return _lib


class OracleModel:
    """The oracle instance; same methods as the product's native handle."""

    def __init__(self, spec, precision=32, diag_tracers=()):
        pi = _pkg()
        lib = load()
        desc, keep = pi._abi.make_desc(spec.grid, spec.n_species, spec.max_individuals,
                                       spec.species_params, spec.bgc_vector, spec.seed, diag_tracers,
                                       getattr(spec, "mode_id", pi._abi.MODE_QUOTA),
                                       getattr(spec, "abiotic_matrix", None), getattr(spec, "palat", None))
        raw = C.c_void_p()
        pi._native.check(lib, "orc_", lib.orc_model_create(C.byref(desc), precision, C.byref(raw)))
        del keep
        self.h = pi._native.Handle(lib, "orc_", raw.value, spec.grid, spec.n_species,
                                   spec.max_individuals, diag_tracers)
        self.spec = spec
        self.precision = precision
        for s in range(spec.n_species):
            self.h.upload_individuals(s, spec.initial_individuals[s])
        for tid, c in enumerate(spec.initial_tracers):
            self.h.upload_tracer(tid, c)
        for a, arrays in enumerate(getattr(spec, "initial_abiotics", [])):
            self.h.upload_abiotic(a, arrays)

    def __getattr__(self, k):
        return getattr(self.h, k)

    def set_diag_tracers(self, ids):
        a = np.ascontiguousarray(np.array(ids, dtype=np.int32))
        pi = _pkg()
        pi._native.check(self.h.lib, "orc_", self.h.lib.orc_set_diag_tracers(
            self.h.raw, len(a), a.ctypes.data_as(C.POINTER(C.c_int32)) if len(a) else None))

    def field_f64(self, sp, name):
        pi = _pkg()
        out = np.empty(self.spec.max_individuals, np.float64)
        pi._native.check(self.h.lib, "orc_", self.h.lib.orc_download_field_f64(
            self.h.raw, sp, pi._abi.FIELD_ID[name], out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def tracer_f64(self, tid):
        pi = _pkg()
        out = np.empty(self.spec.grid.padded_shape, np.float64, order="F")
        pi._native.check(self.h.lib, "orc_", self.h.lib.orc_download_tracer_f64(
            self.h.raw, tid, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def update(self, dt, n, vel=None, parf=None, temp=None, vel_fn=None, parf_fn=None, temp_fn=None):
        """n steps of update!: before every step the velocity / PARF / temperature
        slices are copied in (update.jl:18-26), from constants or from fn(model_t)."""
        t = np.float32(0.0)
        for _ in range(n):
            v = vel_fn(t) if vel_fn else vel
            if v is not None:
                self.h.set_velocity(*v)
            p = parf_fn(t) if parf_fn else parf
            if p is not None:
                self.h.set_parf(p)
            tt = temp_fn(t) if temp_fn else temp
            if tt is not None:
                self.h.set_temperature(tt)
            self.h.step(np.float32(dt))
            t = np.float32(np.float32(self.h.iteration()) * np.float32(dt))
