"""Output writers: PlanktonOutputWriter and write_output! (src/Output/output_writers.jl,
src/Output/write_outputs.jl).

* Text logs (`write_log`): per species and call, one line of
  `dynamic_species00N.txt` with the day, hour, population and the population
  averages of the QuotaMode state (write_outputs.jl:98-117).  The averages are
  the reference's `dot(field, ac) / pop`: a Float32 sum in slot order over every
  slot (an inactive slot adds `field * false`, a signed zero), divided by
  Float32(pop); the line uses the same printf conversions.
* Diagnostics (`save_diags`) and individuals (`save_phytoplankton`): the
  groups and datasets of write_diags_to_jld2 / write_individuals_to_jld2
  (write_outputs.jl:159-202) — `timeseries/t/<iter>`, `timeseries/<tracer>/<iter>`,
  `timeseries/phyto/<sp>/<proc>/<iter>` (interior ./ iteration_interval),
  `timeseries/<sp>/<att>/<iter>` — in HDF5 files (JLD2's container format),
  written through the HDF5 C library (`PI_HDF5_LIB`, else libhdf5 from the
  library path or /opt/conda/lib).  Arrays keep Julia's column-major layout: a
  (Nx, Ny, Nz) array is stored with HDF5 dimensions (Nz, Ny, Nx), as JLD2 stores
  it; Bool columns are an HDF5 enum over int8 (FALSE = 0, TRUE = 1).  The
  diagnostics are zeroed after they are written (on the device).
* File rotation (`max_filesize`): as start_next_diags_file /
  start_next_phytoplankton_file / start_next_abiotic_particle_file
  (write_outputs.jl:42-75): when the current file
  has reached the limit, the first file is renamed `<prefix>_part1.jld2` and
  writing continues in `<prefix>_part2.jld2`, `_part3`, ...

The reference's per-individual rate diagnostics are not on this path; the
`ptc` diagnostic is the device's (zero unless an IronEnergyMode model has
abiotic particles).
"""
import ctypes as C
import ctypes.util
import math
import os
import re
import shutil

import numpy as np


# ---------------------------------------------------------------- HDF5 (C library)
class _HDF5:
    """The few HDF5 C entry points the writers use (HDF5 1.10 ABI: hid_t is int64)."""

    def __init__(self):
        path = os.environ.get("PI_HDF5_LIB")
        cands = [path] if path else []
        found = ctypes.util.find_library("hdf5")
        if found:
            cands.append(found)
        cands += ["/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so"]
        lib, err = None, None
        for p in cands:
            try:
                lib = C.CDLL(p)
                break
            except OSError as e:
                err = e
        if lib is None:
            raise RuntimeError(f"the HDF5 C library is needed for save_diags / save_phytoplankton ({err}); "
                               "set PI_HDF5_LIB to libhdf5.so")
        hid, herr, hsize = C.c_int64, C.c_int, C.c_ulonglong
        proto = {
            "H5open": (herr, []), "H5Fcreate": (hid, [C.c_char_p, C.c_uint, hid, hid]),
            "H5Fopen": (hid, [C.c_char_p, C.c_uint, hid]), "H5Fclose": (herr, [hid]),
            "H5Pcreate": (hid, [hid]), "H5Pclose": (herr, [hid]),
            "H5Pset_create_intermediate_group": (herr, [hid, C.c_uint]),
            "H5Screate": (hid, [C.c_int]), "H5Screate_simple": (hid, [C.c_int, C.POINTER(hsize), C.c_void_p]),
            "H5Sclose": (herr, [hid]), "H5Sget_simple_extent_ndims": (C.c_int, [hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid, C.POINTER(hsize), C.c_void_p]),
            "H5Dcreate2": (hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
            "H5Dopen2": (hid, [hid, C.c_char_p, hid]), "H5Dclose": (herr, [hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dread": (herr, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dget_space": (hid, [hid]), "H5Dget_type": (hid, [hid]),
            "H5Tget_class": (C.c_int, [hid]), "H5Tget_size": (C.c_size_t, [hid]), "H5Tclose": (herr, [hid]),
            "H5Tenum_create": (hid, [hid]), "H5Tenum_insert": (herr, [hid, C.c_char_p, C.c_void_p]),
            "H5Lexists": (C.c_int, [hid, C.c_char_p, hid]),
        }
        for name, (res, args) in proto.items():
            f = getattr(lib, name)
            f.restype, f.argtypes = res, args
        if lib.H5open() < 0:
            raise RuntimeError("H5open failed")
        g = lambda n: C.c_int64.in_dll(lib, n).value
        self.lib = lib
        self.T = {np.dtype(np.float32): g("H5T_NATIVE_FLOAT_g"), np.dtype(np.float64): g("H5T_NATIVE_DOUBLE_g"),
                  np.dtype(np.int64): g("H5T_NATIVE_INT64_g"), np.dtype(np.int8): g("H5T_NATIVE_INT8_g"),
                  np.dtype(np.uint8): g("H5T_NATIVE_UINT8_g")}
        self.lcpl = lib.H5Pcreate(g("H5P_CLS_LINK_CREATE_ID_g"))
        lib.H5Pset_create_intermediate_group(self.lcpl, 1)
        self.bool_t = lib.H5Tenum_create(self.T[np.dtype(np.int8)])
        for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
            x = C.c_int8(v)
            lib.H5Tenum_insert(self.bool_t, name, C.byref(x))

    def _ok(self, v, what):
        if v < 0:
            raise RuntimeError(f"HDF5: {what} failed")
        return v

    def open(self, path):
        """The file opened for appending ("a+": created when absent)."""
        p = path.encode()
        if os.path.exists(path):
            return self._ok(self.lib.H5Fopen(p, 1, 0), f"open {path}")
        return self._ok(self.lib.H5Fcreate(p, 2, 0, 0), f"create {path}")

    def close(self, fid):
        self._ok(self.lib.H5Fclose(fid), "close")

    def write(self, fid, key, value):
        """Dataset `key` (intermediate groups created) holding `value`: a scalar
        or an array in Julia's column-major layout."""
        a = np.asarray(value)
        is_bool = a.dtype == np.bool_
        if is_bool:
            a = a.astype(np.int8)
        if a.dtype not in self.T:
            raise TypeError(f"HDF5 writer: unsupported dtype {a.dtype}")
        t = self.bool_t if is_bool else self.T[a.dtype]
        lib = self.lib
        if a.ndim == 0:
            sid = lib.H5Screate(0)
            buf = np.ascontiguousarray(a)
        else:
            buf = np.asfortranarray(a)
            dims = (C.c_ulonglong * a.ndim)(*reversed(a.shape))
            sid = lib.H5Screate_simple(a.ndim, dims, None)
        self._ok(sid, "dataspace")
        try:
            did = self._ok(lib.H5Dcreate2(fid, key.encode(), t, sid, self.lcpl, 0, 0), f"create {key}")
            try:
                self._ok(lib.H5Dwrite(did, t, 0, 0, 0, buf.ctypes.data), f"write {key}")
            finally:
                lib.H5Dclose(did)
        finally:
            lib.H5Sclose(sid)

    def read(self, fid, key):
        """Dataset `key` as a numpy value (column-major arrays, Julia's shape)."""
        lib = self.lib
        did = self._ok(lib.H5Dopen2(fid, key.encode(), 0), f"open {key}")
        try:
            sid = lib.H5Dget_space(did)
            n = lib.H5Sget_simple_extent_ndims(sid)
            dims = (C.c_ulonglong * max(n, 1))()
            lib.H5Sget_simple_extent_dims(sid, dims, None)
            lib.H5Sclose(sid)
            tid = lib.H5Dget_type(did)
            cls, size = lib.H5Tget_class(tid), lib.H5Tget_size(tid)
            lib.H5Tclose(tid)
            if cls == 1:
                dt, mt = (np.float32, self.T[np.dtype(np.float32)]) if size == 4 else (np.float64, self.T[np.dtype(np.float64)])
            elif cls == 8:
                dt, mt = np.int8, self.bool_t
            else:
                dt, mt = np.int64, self.T[np.dtype(np.int64)]
            shape = tuple(int(d) for d in dims[:n])
            out = np.empty(shape[::-1] if n else (), dt, order="F")
            self._ok(lib.H5Dread(did, mt, 0, 0, 0, out.ctypes.data), f"read {key}")
            return out.astype(bool) if cls == 8 else out
        finally:
            lib.H5Dclose(did)

    def exists(self, fid, key):
        """True if every link of the path `key` exists."""
        parts = key.strip("/").split("/")
        for i in range(1, len(parts) + 1):
            if self.lib.H5Lexists(fid, "/".join(parts[:i]).encode(), 0) <= 0:
                return False
        return True


_h5 = None


def hdf5():
    global _h5
    if _h5 is None:
        _h5 = _HDF5()
    return _h5


def read_dataset(path, key):
    """One dataset of an output file (the reader counterpart of the writers)."""
    h = hdf5()
    fid = h._ok(h.lib.H5Fopen(path.encode(), 0, 0), path)
    try:
        return h.read(fid, key)
    finally:
        h.close(fid)


def has_dataset(path, key):
    h = hdf5()
    fid = h._ok(h.lib.H5Fopen(path.encode(), 0, 0), path)
    try:
        return h.exists(fid, key)
    finally:
        h.close(fid)


# ---------------------------------------------------------------- the writer
class PlanktonOutputWriter:
    """PlanktonOutputWriter(; dir="./results", diags_prefix="diags",
    phytoplankton_prefix="phytoplankton", write_log=false, save_diags=false,
    save_phytoplankton=false, phytoplankton_include=(:x, :y, :z),
    phytoplankton_iteration_interval=1, abiotic_particle_prefix="abiotic_particle",
    save_abiotic_particle=false, abiotic_particle_include=(:x, :y, :z),
    abiotic_particle_iteration_interval=1, max_filesize=Inf) — output_writers.jl:19-80.
    Like the reference, an existing `dir` is removed and created again."""

    def __init__(self, *, dir="./results", diags_prefix="diags", phytoplankton_prefix="phytoplankton",
                 abiotic_particle_prefix="abiotic_particle", write_log=False, save_diags=False,
                 save_phytoplankton=False, save_abiotic_particle=False, phytoplankton_include=("x", "y", "z"),
                 abiotic_particle_include=("x", "y", "z"), phytoplankton_iteration_interval=1,
                 abiotic_particle_iteration_interval=1, max_filesize=math.inf):
        if os.path.isdir(dir):
            shutil.rmtree(dir)
        os.mkdir(dir)
        self.filepath = dir
        self.write_log = bool(write_log)
        self.save_diags = bool(save_diags)
        self.save_phytoplankton = bool(save_phytoplankton)
        self.save_abiotic_particle = bool(save_abiotic_particle)
        self.diags_file = os.path.join(dir, diags_prefix + ".jld2") if save_diags else ""
        self.phytoplankton_file = os.path.join(dir, phytoplankton_prefix + ".jld2") if save_phytoplankton else ""
        self.phytoplankton_include = tuple(phytoplankton_include)
        self.phytoplankton_iteration_interval = int(phytoplankton_iteration_interval)
        self.abiotic_particle_file = (os.path.join(dir, abiotic_particle_prefix + ".jld2")
                                      if save_abiotic_particle else "")
        self.abiotic_particle_include = tuple(abiotic_particle_include)
        self.abiotic_particle_iteration_interval = int(abiotic_particle_iteration_interval)
        self.max_filesize = max_filesize
        self.part_diags = 1
        self.part_phytoplankton = 1
        self.part_abiotic_particle = 1

    def __repr__(self):
        """show(io, writer) — output_writers.jl:93-128."""
        d = f"diagnostics are saved as {self.diags_file}" if self.save_diags else "diagnostics are not saved"
        p = (f"phytoplankton is saved as {self.phytoplankton_file}\n"
             f"│   ├── every {self.phytoplankton_iteration_interval} seconds\n"
             f"│   └── including: {_julia_tuple(self.phytoplankton_include)}"
             if self.save_phytoplankton else "phytoplankton are not saved")
        a = (f"abiotic particle is saved as {self.abiotic_particle_file}\n"
             f"│   ├── every {self.abiotic_particle_iteration_interval} seconds\n"
             f"│   └── including: {_julia_tuple(self.abiotic_particle_include)}"
             if self.save_abiotic_particle else "abiotic particle are not saved")
        return (f"PlanktonOutputWriter:\n├── files are saved at {self.filepath}\n├── {d}\n├── {p}\n"
                f"├── {a}\n├── write log: {str(self.write_log).lower()}\n"
                f"└── Maximum file size: {humanize_filesize(self.max_filesize)}")


def _julia_tuple(names):
    """How Julia prints a tuple of Symbols: (:x, :y, :z), (:x,) for one."""
    body = ", ".join(":" + n for n in names)
    return "(" + body + ("," if len(names) == 1 else "") + ")"


def humanize_filesize(s):
    """humanize_filesize (output_writers.jl:129-142): the first power of 1024
    above the size names the unit, printed as %3.1f of 1024·s/unit.  The unit
    is Julia's Int 1024^power, which wraps to 0 from 1024^7 on, so sizes of
    1024^6 B and more (and Inf) end in YiB with unit 0: "Inf YiB"."""
    suffix = ("B", "KiB", "MiB", "GiB", "TiB", "PiB", "EiB", "ZiB", "YiB")
    big, unit = suffix[0], 1024
    for power in range(1, len(suffix) + 1):
        unit = 1024 ** power if power <= 6 else 0
        big = suffix[power - 1]
        if s < unit:
            break
    v = math.inf if unit == 0 else 1024 * s / unit
    return _fmt("%3.1f", v) + " " + big


def _filesize(path):
    """Julia's filesize: 0 for a missing file."""
    try:
        return os.path.getsize(path)
    except OSError:
        return 0


def _next_part(path, part):
    """start_next_*_file (write_outputs.jl:42-75): returns (new path, new part)."""
    if part == 1:
        p1 = re.sub(r"\.jld2$", "_part1.jld2", path)
        if os.path.exists(path):
            shutil.move(path, p1)
        path = p1
    part += 1
    return re.sub(r"part\d+.jld2$", "part" + str(part) + ".jld2", path), part


def write_output(writer, model, diags):
    """write_output!(writer, model, diags) — write_outputs.jl:2-40; called by
    `update` after every step when the simulation has a writer."""
    if writer is None:
        return
    it = model.iteration
    if writer.write_log:
        write_species_dynamics(model.t, model, writer.filepath)
    if writer.save_diags and it % diags.iteration_interval == 0:
        if _filesize(writer.diags_file) >= writer.max_filesize:
            writer.diags_file, writer.part_diags = _next_part(writer.diags_file, writer.part_diags)
        write_diags(diags, writer.diags_file, model.t, it, diags.iteration_interval, model.grid)
    if writer.save_phytoplankton and it % writer.phytoplankton_iteration_interval == 0:
        if _filesize(writer.phytoplankton_file) >= writer.max_filesize:
            writer.phytoplankton_file, writer.part_phytoplankton = _next_part(writer.phytoplankton_file,
                                                                              writer.part_phytoplankton)
        write_individuals(model.individuals.phytos, writer.phytoplankton_file, model.t, it,
                          writer.phytoplankton_include)
    if writer.save_abiotic_particle and it % writer.abiotic_particle_iteration_interval == 0:
        if _filesize(writer.abiotic_particle_file) >= writer.max_filesize:
            writer.abiotic_particle_file, writer.part_abiotic_particle = _next_part(
                writer.abiotic_particle_file, writer.part_abiotic_particle)
        write_individuals(model.individuals.abiotics, writer.abiotic_particle_file, model.t, it,
                          writer.abiotic_particle_include)


# ---------------------------------------------------------------- text logs
def dot_ac(x, ac):
    """Julia's generic dot(x::Vector{Float32}, ac::Vector{Bool}): s = 0f0, then
    s += x[i] * ac[i] in slot order (x * false = copysign(0, x)), in Float32."""
    x = np.asarray(x, np.float32)
    ac = np.asarray(ac, bool)
    terms = np.where(ac, x, np.copysign(np.float32(0.0), x)).astype(np.float32)
    if terms.size == 0:
        return np.float32(0.0)
    return np.cumsum(terms, dtype=np.float32)[-1] + np.float32(0.0)


def _fmt(spec, x):
    """One printf conversion of Julia's Printf (C conversions; NaN and Inf spelled
    as Julia does)."""
    v = float(x)
    if math.isnan(v) or math.isinf(v):
        s = "NaN" if math.isnan(v) else ("Inf" if v > 0 else "-Inf")
        w = re.match(r"%(\d*)", spec).group(1)
        return s.rjust(int(w)) if w else s
    return spec % v


DYNAMICS_FIELDS = ("gen", "age", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl")
DYNAMICS_SPECS = ("%3.0f", "%2.2f", "%6.0f", "%1.2f", "%1.2f", "%1.2f", "%.8E", "%.8E", "%.8E", "%.8E", "%.8E")
# CarbonMode (write_outputs.jl:142-157): day, hour, pop, gen, age, size and Bm averages
CARBON_DYNAMICS_FIELDS = ("gen", "age", "Sz", "Bm")
CARBON_DYNAMICS_SPECS = ("%3.0f", "%2.2f", "%6.0f", "%1.2f", "%1.2f", "%1.2f", "%.8E")
# MacroMolecularMode (write_outputs.jl:77-96): day, hour, pop, gen, age, then the
# protein, DNA, RNA, CH, NST, PST and Chl averages
MM_DYNAMICS_FIELDS = ("gen", "age", "PRO", "DNA", "RNA", "CH", "NST", "PST", "Chl")
MM_DYNAMICS_SPECS = ("%3.0f", "%2.2f", "%6.0f", "%1.2f", "%1.2f") + ("%.8E",) * 7
# IronEnergyMode (write_outputs.jl:118-140) averages a column `En` that its
# StructArray does not have (it has exEn), so the reference fails there before
# writing; reading `En` from the species data raises the same way here
IRON_DYNAMICS_FIELDS = ("gen", "age", "Sz", "Bm", "En", "CH", "qNO3", "qNH4", "qP", "qFe", "Chl")
IRON_DYNAMICS_SPECS = ("%3.0f", "%2.2f", "%6.0f", "%1.2f", "%1.2f", "%1.2f") + ("%.8E",) * 8
_DYNAMICS = {0: (DYNAMICS_FIELDS, DYNAMICS_SPECS), 1: (CARBON_DYNAMICS_FIELDS, CARBON_DYNAMICS_SPECS),
             2: (MM_DYNAMICS_FIELDS, MM_DYNAMICS_SPECS), 3: (IRON_DYNAMICS_FIELDS, IRON_DYNAMICS_SPECS)}


def dynamics_line(t, columns, mode=0):
    """The log line of QuotaMode (write_outputs.jl:98-117), CarbonMode
    (:142-157) or MacroMolecularMode (:77-96; `mode` is the ABI's mode id, True
    for CarbonMode) from Float32 model time t and the species' columns {field:
    array, "ac": Bool array}."""
    fields, specs = _DYNAMICS[int(mode)]
    t = np.float32(t)
    ac = np.asarray(columns["ac"], bool)
    pop = int(np.count_nonzero(ac))                        # dot(ac, ac)
    with np.errstate(divide="ignore", invalid="ignore"):
        aves = [dot_ac(columns[f], ac) / np.float32(pop) for f in fields]
    day = np.float32(math.trunc(float(t) / 86400.0))       # t ÷ 86400
    hour = np.float32(np.fmod(t, np.float32(86400.0))) / np.float32(3600.0)
    vals = [day, hour, pop] + aves
    return "  ".join(_fmt(s, v) for s, v in zip(specs, vals))


def write_species_dynamics(t, model, filepath):
    """write_species_dynamics(t, phytos, filepath, mode): one line per species."""
    mode = getattr(getattr(model, "spec", None), "mode_id", 0)
    fields = _DYNAMICS[mode][0]
    for s in range(model.n_species):
        d = model.individuals.phytos[f"sp{s + 1}"].data
        cols = {f: getattr(d, f) for f in fields + ("ac",)}
        with open(os.path.join(filepath, "dynamic_species" + str(s + 1).zfill(3) + ".txt"), "a") as io:
            io.write(dynamics_line(t, cols, mode) + "\n")


# ---------------------------------------------------------------- HDF5 files
def _interior(a, grid):
    return a[grid.Hx:grid.Hx + grid.Nx, grid.Hy:grid.Hy + grid.Ny, grid.Hz:grid.Hz + grid.Nz]


def write_individuals(particles, filepath, t, it, atts):
    """write_individuals_to_jld2 (write_outputs.jl:159-169): `particles` is
    model.individuals.phytos or model.individuals.abiotics (name → .data)."""
    h = hdf5()
    fid = h.open(filepath)
    try:
        h.write(fid, f"timeseries/t/{it}", np.float32(t))
        for sp, x in particles.items():
            for att in atts:
                h.write(fid, f"timeseries/{sp}/{att}/{it}", getattr(x.data, att))
    finally:
        h.close(fid)


def write_diags(diags, filepath, t, it, ncounts, grid):
    """write_diags_to_jld2 (write_outputs.jl:171-202): interior ./ ncounts, then
    every accumulator zeroed."""
    h = hdf5()
    n = np.float32(ncounts)
    fid = h.open(filepath)
    try:
        h.write(fid, f"timeseries/t/{it}", np.float32(t))
        tr = diags.tracer
        for key in diags.tracer_keys():
            h.write(fid, f"timeseries/{key}/{it}", (_interior(tr[key], grid) / n).astype(np.float32))
        for sp, procs in diags.phytos.items():
            for proc in diags.phyto_keys():
                h.write(fid, f"timeseries/phyto/{sp}/{proc}/{it}", (_interior(procs[proc], grid) / n).astype(np.float32))
        for sa, procs in getattr(diags, "abiotics", {}).items():
            for proc in procs:
                h.write(fid, f"timeseries/abiotic/{sa}/{proc}/{it}", (_interior(procs[proc], grid) / n).astype(np.float32))
    finally:
        h.close(fid)
    diags.reset()
