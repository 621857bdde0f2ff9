"""Output writers (CPU): the log line's arithmetic and printf conversions, the
HDF5 container (datasets read back through the HDF5 library, layout checked with
h5dump when present), intervals, file rotation and directory handling of
write_output! (write_outputs.jl) on stand-in models."""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest


@pytest.fixture(scope="module")
def out(pi):
    from importlib import import_module
    return import_module("pi_amd.output")


def _h5(out):
    try:
        return out.hdf5()
    except RuntimeError as e:
        pytest.skip(str(e))


# ---------------------------------------------------------------- log arithmetic
def test_dot_ac_is_julias_sequential_float32_dot(out):
    """dot(x::Vector{Float32}, ac::Vector{Bool}) is `s = 0f0; s += x[i]*ac[i]` in
    slot order: the result differs from a pairwise sum on these values, and
    inactive slots add signed zeros (an inactive -x never turns the sum to -0)."""
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000)).astype(np.float32)
    ac = rng.random(5000) < 0.7
    s = np.float32(0.0)
    for a, b in zip(x, ac):                 # the reference's loop, literally
        s = np.float32(s + (a if b else np.copysign(np.float32(0.0), a)))
    got = out.dot_ac(x, ac)
    assert got.dtype == np.float32 and got == s
    assert np.float32(np.sum(np.where(ac, x, 0), dtype=np.float32)) != s      # order matters here
    neg = np.array([-1.0, -2.0], np.float32)
    assert math.copysign(1.0, out.dot_ac(neg, np.array([False, False]))) == 1.0
    assert out.dot_ac(np.zeros(0, np.float32), np.zeros(0, bool)) == 0.0


def _cfmt(spec, v):
    libc = C.CDLL(None)
    buf = C.create_string_buffer(64)
    libc.snprintf(buf, 64, spec.encode(), C.c_double(v))
    return buf.value.decode()


def test_printf_conversions_match_c(out):
    """Each conversion of the log line against C's snprintf (the conversions
    Julia's Printf implements), on values that round at the last digit, tiny and
    huge values and signed zeros; NaN and Inf are spelled as Julia prints them."""
    vals = [0.0, -0.0, 0.005, 0.015, 0.125, 2.675, 1.0 / 3.0, 123456.5, 1e-30, 3.4028235e38, 7.0e-45,
            np.float32(0.1), np.float32(2.5e-12), np.float32(-1.75e-9)]
    for spec in ("%3.0f", "%2.2f", "%6.0f", "%1.2f", "%.8E"):
        for v in vals:
            assert out._fmt(spec, v) == _cfmt(spec, float(v)), (spec, v)
    assert out._fmt("%1.2f", float("nan")) == "NaN"
    assert out._fmt("%3.0f", float("inf")) == "Inf"
    assert out._fmt("%.8E", float("-inf")) == "-Inf"


def test_dynamics_line(out):
    """A log line from hand-made columns: day and hour of t = 1.5 days + 2.5 h,
    the population, the averages in Float32 (sum / Float32(pop))."""
    n = 6
    ac = np.array([True, False, True, True, False, True])
    cols = {f: np.arange(1, n + 1, dtype=np.float32) * np.float32(k + 1) for k, f in enumerate(out.DYNAMICS_FIELDS)}
    cols["ac"] = ac
    t = np.float32(1.5 * 86400 + 2.5 * 3600)
    line = out.dynamics_line(t, cols)
    means = [np.float32(np.float32((k + 1) * (1 + 3 + 4 + 6)) / np.float32(4)) for k in range(8)]
    want = ["  1", "14.50", "     4"] + ["%1.2f" % m for m in means[:3]] + ["%.8E" % m for m in means[3:]]
    assert line == "  ".join(want)
    assert want[3] == "3.50" and want[6] == "1.40000000E+01"
    empty = dict(cols, ac=np.zeros(n, bool))
    assert out.dynamics_line(t, empty) == "  ".join(["  1", "14.50", "     0"] + ["NaN"] * 8)
    # CarbonMode: gen, age and size with %1.2f, Bm with %.8E and nothing else
    cline = out.dynamics_line(t, cols, 1)
    assert cline == "  ".join(want[:6] + ["%.8E" % means[3]])
    assert cline.endswith("1.40000000E+01") and len(cline.split()) == 7
    # MacroMolecularMode: gen and age with %1.2f, then PRO, DNA, RNA, CH, NST, PST, Chl with %.8E
    mcols = {f: np.arange(1, n + 1, dtype=np.float32) * np.float32(k + 1)
             for k, f in enumerate(out.MM_DYNAMICS_FIELDS)}
    mcols["ac"] = ac
    mm = [np.float32(np.float32((k + 1) * 14) / np.float32(4)) for k in range(9)]
    mline = out.dynamics_line(t, mcols, 2)
    assert mline == "  ".join(["  1", "14.50", "     4", "%1.2f" % mm[0], "%1.2f" % mm[1]] + ["%.8E" % m for m in mm[2:]])
    assert len(mline.split()) == 12 and mline.split()[5] == "1.05000000E+01"


# ---------------------------------------------------------------- HDF5 container
def test_hdf5_datasets_round_trip(out, tmp_path):
    h = _h5(out)
    path = str(tmp_path / "a.jld2")
    a3 = np.asfortranarray(np.arange(24, dtype=np.float32).reshape((2, 3, 4), order="F"))
    vals = {"timeseries/t/10": np.float32(600.0), "timeseries/sp1/x/10": np.linspace(0, 1, 7, dtype=np.float32),
            "timeseries/NO3/10": a3, "timeseries/sp1/xi/10": np.arange(5, dtype=np.int64) - 2,
            "timeseries/sp1/ac/10": np.array([True, False, True])}
    fid = h.open(path)
    for k, v in vals.items():
        h.write(fid, k, v)
    h.close(fid)
    fid = h.open(path)                       # appending to the existing file
    h.write(fid, "timeseries/t/20", np.float32(1200.0))
    h.close(fid)
    for k, v in list(vals.items()) + [("timeseries/t/20", np.float32(1200.0))]:
        r = out.read_dataset(path, k)
        assert r.dtype == np.asarray(v).dtype and r.shape == np.asarray(v).shape, k
        assert np.array_equal(r, v), k
    assert out.has_dataset(path, "timeseries/sp1/x/10") and not out.has_dataset(path, "timeseries/sp2/x/10")
    dump = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)
    if dump:
        txt = subprocess.run([dump, "-H", path], capture_output=True, text=True, check=True).stdout
        assert 'DATASET "10"' in txt and "H5T_IEEE_F32LE" in txt and "H5T_STD_I64LE" in txt
        assert "( 4, 3, 2 )" in txt                 # column-major (2, 3, 4) stored as JLD2 stores it
        assert 'GROUP "timeseries"' in txt and 'GROUP "NO3"' in txt


# ---------------------------------------------------------------- write_output! on stand-ins
class _Grid:
    Nx, Ny, Nz, Hx, Hy, Hz = 3, 2, 2, 2, 2, 2
    padded_shape = (7, 6, 6)


class _Data:
    def __init__(self, cols):
        self._c = cols

    def __getattr__(self, k):
        if k.startswith("_"):
            raise AttributeError(k)
        return self._c[k]


class _Sp:
    def __init__(self, cols):
        self.data = _Data(cols)


class _Model:
    def __init__(self, n, nsp=2):
        rng = np.random.default_rng(n)
        self.grid, self.n_species, self.iteration, self.t = _Grid(), nsp, 0, np.float32(0.0)
        cols = []
        for s in range(nsp):
            c = {f: rng.random(n).astype(np.float32) for f in ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl",
                                                               "gen", "age")}
            c["ac"] = rng.random(n) < 0.8
            c["xi"] = rng.integers(0, 5, n).astype(np.int64)
            cols.append(c)
        self.individuals = type("I", (), {})()
        self.individuals.phytos = {f"sp{s + 1}": _Sp(cols[s]) for s in range(nsp)}


class _Diags:
    def __init__(self, model, interval):
        self.iteration_interval = interval
        self.model = model
        self.resets = 0
        self.acc = np.ones(model.grid.padded_shape, np.float32, order="F")

    def tracer_keys(self):
        return ("NO3", "PAR")

    def phyto_keys(self):
        return ("num", "graz", "mort", "dvid", "ptc")

    @property
    def tracer(self):
        return {"NO3": self.acc * 3, "PAR": self.acc * 5}

    @property
    def phytos(self):
        return {f"sp{s + 1}": {k: self.acc * (s + 1) for k in self.phyto_keys()} for s in range(self.model.n_species)}

    def reset(self):
        self.resets += 1


def test_write_output_intervals_rotation_and_logs(out, tmp_path):
    _h5(out)
    d = str(tmp_path / "results")
    os.mkdir(d)
    open(os.path.join(d, "stale.txt"), "w").close()
    w = out.PlanktonOutputWriter(dir=d, write_log=True, save_diags=True, save_phytoplankton=True,
                                 phytoplankton_include=("x", "ac", "xi"), phytoplankton_iteration_interval=2,
                                 max_filesize=1)
    assert not os.path.exists(os.path.join(d, "stale.txt"))        # isdir(dir) && rm(dir, recursive=true)
    m = _Model(40)
    diags = _Diags(m, 3)
    for it in range(1, 7):
        m.iteration, m.t = it, np.float32(it * 60.0)
        out.write_output(w, m, diags)
    # diagnostics at 3 and 6, the second in a new part (every file exceeds 1 byte)
    assert sorted(os.listdir(d)) == ["diags_part1.jld2", "diags_part2.jld2", "dynamic_species001.txt",
                                     "dynamic_species002.txt", "phytoplankton_part1.jld2",
                                     "phytoplankton_part2.jld2", "phytoplankton_part3.jld2"]
    assert diags.resets == 2
    p1 = os.path.join(d, "diags_part1.jld2")
    assert out.has_dataset(p1, "timeseries/t/3") and not out.has_dataset(p1, "timeseries/t/6")
    no3 = out.read_dataset(os.path.join(d, "diags_part2.jld2"), "timeseries/NO3/6")
    assert no3.shape == (3, 2, 2) and np.all(no3 == np.float32(3.0) / np.float32(3.0))
    num = out.read_dataset(p1, "timeseries/phyto/sp2/num/3")
    assert np.all(num == np.float32(2.0) / np.float32(3.0))
    ph = os.path.join(d, "phytoplankton_part2.jld2")
    assert np.array_equal(out.read_dataset(ph, "timeseries/sp1/ac/4"), m.individuals.phytos["sp1"].data.ac)
    assert np.array_equal(out.read_dataset(ph, "timeseries/sp2/xi/4"), m.individuals.phytos["sp2"].data.xi)
    assert out.read_dataset(ph, "timeseries/t/4") == np.float32(240.0)
    lines = open(os.path.join(d, "dynamic_species001.txt")).read().splitlines()
    assert len(lines) == 6
    cols = {f: getattr(m.individuals.phytos["sp1"].data, f) for f in out.DYNAMICS_FIELDS + ("ac",)}
    assert lines[-1] == out.dynamics_line(np.float32(360.0), cols)


def test_writer_without_limits_keeps_one_file(out, tmp_path):
    _h5(out)
    d = str(tmp_path / "r")
    w = out.PlanktonOutputWriter(dir=d, save_diags=True)
    m = _Model(10, nsp=1)
    diags = _Diags(m, 1)
    for it in range(1, 4):
        m.iteration, m.t = it, np.float32(it * 30.0)
        out.write_output(w, m, diags)
    assert os.listdir(d) == ["diags.jld2"]
    for it in (1, 2, 3):
        assert out.read_dataset(os.path.join(d, "diags.jld2"), f"timeseries/t/{it}") == np.float32(it * 30.0)


class _AbioticModel(_Model):
    """The stand-in with abiotic particles (model.individuals.abiotics.saN.data)."""

    def __init__(self, n, nsa=2):
        super().__init__(n, nsp=1)
        rng = np.random.default_rng(n + 1)
        ab = {}
        for a in range(nsa):
            c = {f: rng.random(n).astype(np.float32) for f in ("x", "y", "z")}
            c["ac"] = rng.random(n) < 0.5
            c["zi"] = rng.integers(0, 5, n).astype(np.int64)
            ab[f"sa{a + 1}"] = _Sp(c)
        self.individuals.abiotics = ab


class _AbioticDiags(_Diags):
    @property
    def abiotics(self):
        return {f"sa{a + 1}": {"num": self.acc * (10 + a)} for a in range(2)}


def test_abiotic_particles_and_their_diagnostics(out, tmp_path):
    """save_abiotic_particle: every `abiotic_particle_iteration_interval` steps the
    included columns of every abiotic species go to timeseries/saN/<att>/<iter>
    of abiotic_particle.jld2, rotated like the other files (write_outputs.jl:31-38,
    65-74, 159-169); the diagnostics file carries timeseries/abiotic/saN/num/<iter>
    as interior ./ ncounts (write_outputs.jl:182-186)."""
    _h5(out)
    d = str(tmp_path / "res")
    w = out.PlanktonOutputWriter(dir=d, save_diags=True, save_abiotic_particle=True,
                                 abiotic_particle_prefix="ptc", abiotic_particle_include=("x", "ac", "zi"),
                                 abiotic_particle_iteration_interval=3, max_filesize=1)
    m = _AbioticModel(30)
    diags = _AbioticDiags(m, 2)
    for it in range(1, 7):
        m.iteration, m.t = it, np.float32(it * 10.0)
        out.write_output(w, m, diags)
    files = sorted(os.listdir(d))
    assert files == ["diags_part1.jld2", "diags_part2.jld2", "diags_part3.jld2", "ptc_part1.jld2", "ptc_part2.jld2"]
    assert w.abiotic_particle_file == os.path.join(d, "ptc_part2.jld2") and w.part_abiotic_particle == 2
    p1, p2 = os.path.join(d, "ptc_part1.jld2"), os.path.join(d, "ptc_part2.jld2")
    assert out.read_dataset(p1, "timeseries/t/3") == np.float32(30.0)
    assert out.has_dataset(p2, "timeseries/t/6") and not out.has_dataset(p2, "timeseries/t/3")
    for sa in ("sa1", "sa2"):
        data = m.individuals.abiotics[sa].data
        for att in ("x", "ac", "zi"):
            got = out.read_dataset(p2, f"timeseries/{sa}/{att}/6")
            assert got.dtype == getattr(data, att).dtype and np.array_equal(got, getattr(data, att)), (sa, att)
        assert not out.has_dataset(p2, f"timeseries/{sa}/y/6")
    num = out.read_dataset(os.path.join(d, "diags_part2.jld2"), "timeseries/abiotic/sa2/num/4")
    assert num.shape == (3, 2, 2) and np.all(num == np.float32(11.0) / np.float32(2.0))


def test_writer_show_and_humanize_filesize(out, tmp_path):
    """show(io, writer) and humanize_filesize (output_writers.jl:93-142): the
    included columns as a Julia tuple of Symbols; sizes in the first power of
    1024 above them, with Julia's Int 1024^power wrapping to 0 from 1024^7 on
    (so Inf prints as "Inf YiB")."""
    assert [out.humanize_filesize(v) for v in (0, 1023, 1024, 1536, 5e6, 3 * 1024 ** 3, math.inf)] == \
        ["0.0 B", "1023.0 B", "1.0 KiB", "1.5 KiB", "4.8 MiB", "3.0 GiB", "Inf YiB"]
    d = str(tmp_path / "s")
    w = out.PlanktonOutputWriter(dir=d, save_abiotic_particle=True, abiotic_particle_include=("z",),
                                 abiotic_particle_iteration_interval=5, max_filesize=2048)
    assert str(w).splitlines() == [
        "PlanktonOutputWriter:", f"├── files are saved at {d}", "├── diagnostics are not saved",
        "├── phytoplankton are not saved", f"├── abiotic particle is saved as {d}/abiotic_particle.jld2",
        "│   ├── every 5 seconds", "│   └── including: (:z,)", "├── write log: false",
        "└── Maximum file size: 2.0 KiB"]
