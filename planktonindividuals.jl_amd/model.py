"""PlanktonModel on the MI355X architecture (reference: src/Model/models.jl,
src/Architectures.jl, src/Individuals/Individuals.jl,
src/Individuals/Plankton/QuotaMode/plankton_generation.jl)."""
import numpy as np

from . import _abi, params as P, stream
from .fields import default_tracer_init, generate_tracers, tracer_ranges
from ._native import NativeModel


# --------------------------------------------------------------- architectures
class Architecture:
    """abstract type Architecture (src/Architectures.jl:16)."""


class HIPArch(Architecture):
    """Run the per-timestep hot path on one AMD Instinct MI355X (gfx950) through
    libpi_hip.so.  The MI355X counterpart of `GPU()` (src/Architectures.jl:24)."""

    def __init__(self, device=0):
        self.device = device

    def __repr__(self):
        return f"HIPArch({self.device})"


GPU = HIPArch


def isfunctional(arch):
    """isfunctional(arch) (src/Architectures.jl:34): the HIP library loads and a device opens."""
    if not isinstance(arch, HIPArch):
        return False
    from ._native import load_library
    lib = load_library()
    return lib.pi_init(arch.device) == 0


class phyto_setup:
    """phyto_setup(params, N, Nsp) (src/PlanktonIndividuals.jl:112-116)."""

    def __init__(self, params=None, N=(2 ** 10,), Nsp=1):
        self.params = params
        self.N = list(N)
        self.Nsp = Nsp


class Palat:
    """Palat(intac, release) (src/PlanktonIndividuals.jl:118-121): (phytoplankton,
    abiotic) species pairs that interact, and pairs that release particles."""

    def __init__(self, intac=(), release=()):
        self.intac = list(intac)
        self.release = list(release)


class abiotic_setup:
    """abiotic_setup(params, N, Nsa, palat) (src/PlanktonIndividuals.jl:122-127)."""

    def __init__(self, params=None, N=(2 ** 10,), Nsa=1, palat=None):
        self.params = params
        self.N = list(N)
        self.Nsa = Nsa
        self.palat = palat if palat is not None else Palat()


def initialize_abiotic_particle(N, cap, grid, seed, sa):
    """initialize_abiotic_particle! (Abiotic/particle_generation.jl:24-35): the
    first N slots active, positions U·N per axis (the draws from the shared
    stream, species PI_ABIO_SPECIES + sa), then mask_individuals! (land)."""
    f32 = np.float32
    slots = np.arange(cap, dtype=np.int64)
    ux, uy, uz = stream.uniform3(seed, 0, stream.ABIO_SPECIES + sa, stream.KIND_INIT_POS, slots)
    ac = np.zeros(cap, dtype=bool)
    ac[:N] = True
    x = _mask(ac, ux * f32(grid.Nx))
    y = _mask(ac, uy * f32(grid.Ny))
    z = _mask(ac, uz * f32(grid.Nz))
    n = slice(0, N)
    xi = np.trunc(_mask(ac[n], x[n] + f32(1))).astype(np.int64) + grid.Hx
    yi = np.trunc(_mask(ac[n], y[n] + f32(1))).astype(np.int64) + grid.Hy
    zi = np.trunc(_mask(ac[n], z[n] + f32(1))).astype(np.int64) + grid.Hz
    ac[n] = grid.landmask[xi - 1, yi - 1, zi - 1] & ac[n]
    return dict(x=x, y=y, z=z, ac=ac)


# ------------------------------------------------------ initial individuals
def _mask(ac, v):
    """Bool * x  ==  ifelse(b, x, copysign(0, x))."""
    return np.where(ac, v, np.copysign(np.float32(0), v)).astype(np.float32)


def initialize_plankton(p, N, cap, grid, seed, sp, mode=_abi.MODE_QUOTA):
    """initialize_plankton! (QuotaMode/plankton_generation.jl:35-72,
    CarbonMode/plankton_generation.jl:34-56) with the draws of `randn!`/`rand!`
    taken from the shared stream (the reference's generator is entropy-seeded).
    `p` is the species' parameter row.  CarbonMode: Bm and Chl from the size,
    the damaged biomass Bd (in the Cq column) zero, no N or P quotas.
    MacroMolecularMode (MacroMolecularMode/plankton_generation.jl:31-78): DNA,
    RNA and protein from the third draw triple, CH, NST and PST from the second;
    returned in the ABI's columns (Sz = DNA, Bm = PRO, Cq = CH, Nq = NST,
    Pq = PST, ex0 = RNA)."""
    f32 = np.float32
    carbon = mode == _abi.MODE_CARBON
    mm = mode == _abi.MODE_MACROMOLECULAR
    iron = mode == _abi.MODE_IRONENERGY
    names = (_abi.CARBON_PARAM_NAMES if carbon else _abi.MM_PARAM_NAMES if mm else
             _abi.IRON_PARAM_NAMES if iron else _abi.PARAM_NAMES)
    q = {n: p[i] for i, n in enumerate(names)}
    slots = np.arange(cap, dtype=np.int64)
    ux, uy, uz = stream.uniform3(seed, 0, sp, stream.KIND_INIT_POS, slots)
    uc, un, up = stream.uniform3(seed, 0, sp, stream.KIND_INIT_QUOTA, slots)
    a, b, u3 = stream.uniform3(seed, 0, sp, stream.KIND_INIT_SIZE, slots)
    # randn via Box-Muller on (0,1] x [0,1), with the binary64 log and cos of
    # include/pi_math.h (the device generates the same values)
    Sz = stream.init_normal(a, b)
    ac = np.zeros(cap, dtype=bool)
    ac[:N] = True
    gen = np.zeros(cap, f32)
    gen[:N] = f32(1.0)
    age = np.zeros(cap, f32)
    x = _mask(ac, ux * f32(grid.Nx))
    y = _mask(ac, uy * f32(grid.Ny))
    z = _mask(ac, uz * f32(grid.Nz))
    extra = {}
    if mm:
        var, cN, Ns = q["var"], q["C_DNA"], q["Nsuper"]
        spread = lambda u: u * var * f32(2) + f32(1) - var      # rand * var * 2 + 1 - var
        Sz = _mask(ac, (a * var + f32(1)) * cN * Ns)             # DNA
        extra["ex0"] = (_mask(ac, spread(b) * cN * Ns) * q["RNA2DNA"]).astype(f32)
        Bm = _mask(ac, spread(u3) * cN * Ns) * q["PRO2DNA"]     # PRO
        Cq = spread(uc) * Sz * q["CH2DNA"]                      # CH
        Nq = spread(un) * Cq * f32(16) / f32(106)               # NST
        Pq = spread(up) * Cq / f32(106)                         # PST
        Chl = Sz * q["Chl2DNA"] * f32(893.49) / f32(55)
    elif iron:
        # IronEnergyMode/plankton_generation.jl:48-94: the maxima at a tenth, qP from
        # the size triple's third draw, qFe from its own triple; columns Cq = CH,
        # Nq = qNH4, Pq = qP, ex0..ex7 = qNO3, qFe, qFePS, qFeNR, qFeNF, tdark, exEn, ptc
        uf, _, _ = stream.uniform3(seed, 0, sp, stream.KIND_INIT_FE, slots)
        Sz = _mask(ac, np.maximum(f32(1.0), Sz * q["var"] + q["mean"]))
        Bm = q["Cquota"] * Sz * q["Nsuper"]
        Cq = uc * (q["CHmax"] * f32(0.1)) * Bm
        BC = Bm + Cq
        extra["ex0"] = (un * ((q["qNO3max"] * f32(0.1)) * BC)).astype(f32)
        Nq = up * ((q["qNH4max"] * f32(0.1)) * BC)
        Pq = u3 * ((q["qPmax"] * f32(0.1)) * BC)
        qFe = uf * ((q["qFemax"] * f32(0.1)) * BC)
        PSf = qFe * f32(0.4)
        NRf = qFe * f32(0.3) * q["is_nr"]
        NFf = qFe * f32(0.3) * (q["is_croc"] + q["is_tric"])
        extra.update(ex1=(qFe - (PSf + NRf + NFf)).astype(f32), ex2=PSf.astype(f32), ex3=NRf.astype(f32),
                     ex4=NFf.astype(f32), ex5=np.zeros(cap, f32), ex6=np.zeros(cap, f32), ex7=np.zeros(cap, f32))
        Chl = Bm * q["Chl2Cint"]
    else:
        Sz = _mask(ac, np.maximum(f32(1.0), Sz * q["var"] + q["mean"]))
        Bm = (q["Cquota"] * Sz) * q["Nsuper"]
        if carbon:
            Cq = np.zeros(cap, f32)
            Nq = np.zeros(cap, f32)
            Pq = np.zeros(cap, f32)
            Chl = Bm * q["Chl2C"]
        else:
            Cq = (uc * q["Cqmax"]) * Bm
            Nq = un * (q["Nqmax"] * (Bm + Cq) - Bm * q["R_NC"])
            Pq = up * (q["Pqmax"] * (Bm + Cq) - Bm * q["R_PC"])
            Chl = Bm * q["Chl2Cint"]
    # mask_individuals! over the first N slots (Individuals/utils.jl:89-100)
    Hx, Hy, Hz = grid.Hx, grid.Hy, grid.Hz
    n = slice(0, N)
    xi = np.trunc(_mask(ac[n], x[n] + f32(1))).astype(np.int64) + Hx
    yi = np.trunc(_mask(ac[n], y[n] + f32(1))).astype(np.int64) + Hy
    zi = np.trunc(_mask(ac[n], z[n] + f32(1))).astype(np.int64) + Hz
    ac[n] = grid.landmask[xi - 1, yi - 1, zi - 1] & ac[n]
    zeros = np.zeros(cap, f32)
    return dict(x=x, y=y, z=z, Sz=Sz.astype(f32), Bm=Bm.astype(f32), Cq=Cq.astype(f32), Nq=Nq.astype(f32),
                Pq=Pq.astype(f32), Chl=Chl.astype(f32), gen=gen, age=age, graz=zeros.copy(),
                mort=zeros.copy(), dvid=zeros.copy(), xi=np.zeros(cap, np.int64),
                yi=np.zeros(cap, np.int64), zi=np.zeros(cap, np.int64), ac=ac, **extra)


# ------------------------------------------------------------------- views
class _SpeciesData:
    """`model.individuals.phytos.spN.data`: attribute access downloads one column."""

    def __init__(self, model, sp):
        self._model, self._sp = model, sp

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if name == "idx":
            return np.zeros(self._model.max_individuals, np.int64)
        mode = self._model.spec.mode_id
        if mode == _abi.MODE_CARBON:
            # CarbonMode columns (CarbonMode/plankton_generation.jl:2-12): Bd is the Cq column
            if name in ("Cq", "Nq", "Pq"):
                raise AttributeError(name)
            if name == "Bd":
                name = "Cq"
        elif mode in (_abi.MODE_MACROMOLECULAR, _abi.MODE_IRONENERGY):
            # MacroMolecularMode / IronEnergyMode columns (plankton_generation.jl:1-12 of
            # each): their own names only
            state = ("Sz", "Bm", "Cq", "Nq", "Pq") + _abi.EX_FIELDS + tuple(_abi.FIELD_ALIAS)
            if name in state and name not in _abi.MODE_NAMES[mode]:
                raise AttributeError(name)
        elif name in _abi.FIELD_ALIAS or name in _abi.EX_FIELDS:
            raise AttributeError(name)
        return self._model._native.field(self._sp, name)


class _Species:
    def __init__(self, model, sp):
        self.data = _SpeciesData(model, sp)
        self.p = dict(zip(P.param_names(model.mode), model.species_params[sp]))


class _Phytos(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k)


class _AbioticData:
    """`model.individuals.abiotics.saN.data`: x, y, z, xi, yi, zi, ac (merg and idx stay 0)."""

    def __init__(self, model, sa):
        self._model, self._sa = model, sa

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if name in ("merg", "idx"):
            return np.zeros(self._model.max_individuals, np.int64)
        return self._model._native.abiotic_field(self._sa, name)


class _Abiotic:
    def __init__(self, model, sa):
        self.data = _AbioticData(model, sa)
        self.p = dict(zip(_abi.ABIOTIC_PARAM_NAMES, model.spec.abiotic_matrix[sa]))
        self.bc = BoundaryConditions()


class _Individuals:
    def __init__(self, model):
        self.phytos = _Phytos({f"sp{s + 1}": _Species(model, s) for s in range(model.n_species)})
        self.abiotics = _Phytos({f"sa{a + 1}": _Abiotic(model, a) for a in range(model.spec.n_abiotic)})


class BoundaryConditions:
    """BoundaryConditions (src/PlanktonIndividuals.jl:27-34): the flux boundary
    condition on each face, `None` (nothing), a number or an array."""

    def __init__(self):
        self.west = self.east = self.south = self.north = self.bottom = self.top = None


class _Field:
    def __init__(self, model, tid):
        self._m, self._t = model, tid
        self.bc = BoundaryConditions()

    @property
    def data(self):
        return self._m._native.tracer(self._t)


class _Tracers(dict):
    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k)


class _Timestepper:
    """Read access to the scratch fields of `model.timestepper` (Model/timestepper.jl:1-21)."""

    def __init__(self, model):
        self._m = model

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        if name == "plk":
            return _Tracers({n: type("F", (), {"data": self._m._native.grid_field("plk_" + n)})()
                             for n in _abi.TRACER_NAMES})
        key = {"par₀": "par0"}.get(name, name)
        return self._m._native.grid_field(key)


# ------------------------------------------------------------------- model
class ModelSpec:
    """Everything `PlanktonModel(arch, grid; ...)` decides on the host before any
    device work: validated parameters (models.jl:65-97) and the initial state
    (generate_individuals, Individuals.jl:42-72; generate_tracers,
    tracer_fields.jl:93-134).  Pure numpy; any implementation of the ABI can be
    loaded from it."""

    def __init__(self, grid, *, FT=np.float32, mode=None, max_individuals=8 * 1024,
                 bgc_params=None, tracer_initial=None, phyto=None, abiotic=None, t=0.0,
                 max_candidates=25, seed=0, host_state=True):
        mode = mode if mode is not None else P.QuotaMode()
        if not isinstance(mode, (P.QuotaMode, P.CarbonMode, P.MacroMolecularMode, P.IronEnergyMode)):
            raise NotImplementedError(f"{mode}: not a physiology mode")
        self.n_abiotic, self.abiotic_params, self.abiotic_matrix, self.abiotic_N = 0, None, None, []
        self.palat = ((), (), int(max_candidates))
        if abiotic is not None:                       # models.jl:99-117
            if not isinstance(abiotic, abiotic_setup):
                raise ValueError("PlanktonModel:`abiotic` must be either Nothing or `abiotic_setup`!")
            if max(abiotic.N) > max_individuals:
                raise AssertionError("maximum(abiotic.N) ≤ max_individuals")
            if len(abiotic.N) != abiotic.Nsa:
                raise ValueError(f"PlanktonModel: `abiotic`: The length of `N` must be {abiotic.Nsa}, the same as "
                                 "`Nsa`, each species has its own initial condition")
            if abiotic.Nsa > _abi.MAX_ABIOTIC:
                raise NotImplementedError(f"at most {_abi.MAX_ABIOTIC} abiotic species per model on the MI355X path")
            if abiotic.params is None:
                ap = P.update_abiotic_params(P.abiotic_params_default(abiotic.Nsa), FT, N=abiotic.Nsa)
            elif isinstance(abiotic.params, dict):
                ap = P.update_abiotic_params(abiotic.params, FT, N=abiotic.Nsa)
            else:
                raise ValueError("Abiotic particle parameters must be either Nothing or Dict!")
            nphy = phyto.Nsp if phyto is not None else 1
            pairs = {"intac": [], "release": []}
            for kind in ("intac", "release"):                                       # models.jl:121-142
                for pair in getattr(abiotic.palat, kind):
                    sp, sa = (str(v).lstrip(":") for v in pair)
                    if not (sp.startswith("sp") and sp[2:].isdigit() and 1 <= int(sp[2:]) <= nphy):
                        raise ValueError(f"Abiotic: {sp} is not generated")
                    if not (sa.startswith("sa") and sa[2:].isdigit() and 1 <= int(sa[2:]) <= abiotic.Nsa):
                        raise ValueError(f"Abiotic: {sa} is not generated")
                    pairs[kind].append((int(sp[2:]) - 1, int(sa[2:]) - 1))
            if pairs["intac"] and not isinstance(mode, P.IronEnergyMode):
                # particle_interaction! reads plank_p.max_ptc (particle_interaction.jl:60), an
                # IronEnergyMode parameter: the other modes fail there
                raise ValueError("phytoplankton-particle interaction needs IronEnergyMode (max_ptc)")
            if len(pairs["intac"]) > 4 or len(pairs["release"]) > 4:
                raise NotImplementedError("at most 4 interaction and 4 release pairs per model on the MI355X path")
            if (pairs["intac"] or pairs["release"]) and not 1 <= int(max_candidates) <= 64:
                raise NotImplementedError("max_candidates in [1, 64] on the MI355X path")
            self.palat = (pairs["intac"], pairs["release"], int(max_candidates))
            self.n_abiotic = abiotic.Nsa
            self.abiotic_params = ap
            self.abiotic_matrix = P.abiotic_param_matrix(ap, abiotic.Nsa)
            self.abiotic_N = list(abiotic.N)
        if FT is not np.float32:
            raise ValueError("FT must be Float32")
        if bgc_params is None:
            bgc = P.update_bgc_params(P.bgc_params_default(FT), FT)
        elif isinstance(bgc_params, dict):
            bgc = P.update_bgc_params(bgc_params, FT)
        else:
            raise ValueError("Phytoplankton parameters must be either Nothing or Dict!")
        if phyto is None:
            phyto = phyto_setup(None, [2 ** 10], 1)
        if not isinstance(phyto, phyto_setup):
            raise ValueError("PlanktonModel:`phyto` must be either Nothing or `phyto_setup`!")
        if max(phyto.N) > max_individuals:
            raise AssertionError("maximum(phyto.N) ≤ max_individuals")
        if len(phyto.N) != phyto.Nsp:
            raise ValueError(f"PlanktonModel: `phyto`: The length of `N` must be {phyto.Nsp}, the same as `Nsp`, "
                             "each species has its own initial condition")
        if phyto.params is None:
            pp = P.update_phyt_params(P.phyt_params_default(phyto.Nsp, mode), FT, N=phyto.Nsp, mode=mode)
        elif isinstance(phyto.params, dict):
            pp = P.update_phyt_params(phyto.params, FT, N=phyto.Nsp, mode=mode)
        else:
            raise ValueError("Phytoplankton parameters must be either Nothing or Dict!")
        self.grid, self.mode, self.FT = grid, mode, FT
        self.max_individuals = int(max_individuals)
        self.max_candidates = max_candidates
        self.bgc_params = bgc
        self.phyt_params = pp
        self.n_species = phyto.Nsp
        self.N = list(phyto.N)
        self.mode_id = P.mode_id(mode)
        self.species_params = P.species_param_matrix(pp, phyto.Nsp, mode)
        self.bgc_vector = P.bgc_vector(bgc)
        self.seed = int(seed)
        self.t = np.float32(t)
        source = tracer_initial if tracer_initial is not None else default_tracer_init()
        # host_state=False: the device draws the same initial state itself
        # (pi_generate_state); only the tracer value ranges are kept
        self.host_state = bool(host_state)
        self.initial_individuals = self.initial_tracers = None
        self.initial_abiotics = []
        self.tracer_ranges = None
        if self.host_state:
            self.initial_individuals = [initialize_plankton(self.species_params[s], int(phyto.N[s]),
                                                            self.max_individuals, grid, self.seed, s,
                                                            self.mode_id)
                                        for s in range(self.n_species)]
            self.initial_tracers = generate_tracers(grid, source, self.seed)
            self.initial_abiotics = [initialize_abiotic_particle(int(self.abiotic_N[a]), self.max_individuals, grid,
                                                                 self.seed, a) for a in range(self.n_abiotic)]
        else:
            self.tracer_ranges = tracer_ranges(source)


class PlanktonModel:
    """PlanktonModel(arch, grid; FT=Float32, mode=QuotaMode(), max_individuals=8*1024,
    bgc_params=nothing, tracer_initial=default_tracer_init(), phyto=nothing,
    abiotic=nothing, t=0.0f0, max_candidates=25) — Model/models.jl:51-155.

    Extra keyword: `seed`, the key of the shared random stream (initial state and
    every per-step draw).  `spec=` builds from an existing ModelSpec instead.
    """

    def __init__(self, arch, grid=None, *, spec=None, **kw):
        if not isinstance(arch, HIPArch):
            raise ValueError(f"{arch}: this package implements the HIPArch architecture only")
        if spec is None:
            spec = ModelSpec(grid, **kw)
        self.spec = spec
        self.arch, self.grid, self.mode, self.FT = arch, spec.grid, spec.mode, spec.FT
        self.max_individuals = spec.max_individuals
        self.max_candidates = spec.max_candidates
        self.bgc_params = spec.bgc_params
        self.phyt_params = spec.phyt_params
        self.n_species = spec.n_species
        self.species_params = spec.species_params
        self.t = spec.t
        self._native = NativeModel(self.grid, self.n_species, self.max_individuals, self.species_params,
                                   spec.bgc_vector, spec.seed, (), device=arch.device, mode=spec.mode_id,
                                   abiotic_params=spec.abiotic_matrix, palat=spec.palat)
        if spec.host_state:
            for s in range(self.n_species):
                self._native.upload_individuals(s, spec.initial_individuals[s])
            for tid, c in enumerate(spec.initial_tracers):
                self._native.upload_tracer(tid, c)
        else:
            self._native.generate_state(spec.N, *spec.tracer_ranges)
            for a in range(spec.n_abiotic):       # abiotic positions are drawn on the host
                self._native.upload_abiotic(a, initialize_abiotic_particle(
                    int(spec.abiotic_N[a]), spec.max_individuals, spec.grid, spec.seed, a))
        for a, arrays in enumerate(spec.initial_abiotics):
            self._native.upload_abiotic(a, arrays)
        self.individuals = _Individuals(self)
        self.tracers = _Tracers({n: _Field(self, i) for i, n in enumerate(_abi.TRACER_NAMES)})
        self.timestepper = _Timestepper(self)

    @property
    def iteration(self):
        return self._native.iteration()

    def calc_active_individuals(self):
        return [self._native.count_active(s) for s in range(self.n_species)]

    def __repr__(self):
        N = self.calc_active_individuals()
        return ("PlanktonModel:\n"
                f"├── floating point data type: Float32\n"
                f"├── grid: {self.grid!r}\n"
                f"├── {self.mode} is selected for phytoplankton physiology\n"
                f"├── phytoplankton: {self.n_species} species with {N} individuals for each species\n"
                + ("├── No abiotic particles available\n" if not self.spec.n_abiotic else
                   f"├── abiotic particles: {self.spec.n_abiotic} species with "
                   f"{[int(self._native.abiotic_field(a, 'ac').sum()) for a in range(self.spec.n_abiotic)]} "
                   "particles for each species\n") +
                f"└── maximum number of individuals: {self.max_individuals} per species")


def set_bc(model, *, tracer, pos, bc_value):
    """set_bc!(model; tracer, pos, bc_value) — HaloRegions/boundary_conditions.jl:12-29:
    the flux boundary condition of `tracer` on `pos` (:west, :east, :south,
    :north, :bottom, :top) becomes FT.(bc_value): a number, a face array (2-D:
    (Ny, Nz) west/east, (Nx, Nz) south/north, (Nx, Ny) bottom/top) or one face
    array per iteration (3-D, read at the model's iteration).  The value is
    applied to the tracer's tendency every step by apply_bcs! (apply_bcs.jl)."""
    tracer, pos = str(tracer).lstrip(":"), str(pos).lstrip(":")
    if tracer not in _abi.TRACER_ID:
        raise AssertionError("tracer in tracer_names")
    if pos not in _abi.BC_POS:
        raise ValueError(f"type BoundaryConditions has no field {pos}")
    if bc_value is None:                       # back to `nothing`
        v = None
    else:
        v = np.float32(bc_value) if np.ndim(bc_value) == 0 else np.asarray(bc_value, dtype=np.float32)
    model._native.set_tracer_bc(_abi.TRACER_ID[tracer], pos, v)
    setattr(model.tracers[tracer].bc, pos, v)


def set_bc_particle(model, *, sa, pos, bc_value):
    """set_bc_particle!(model; sa, pos, bc_value) — boundary_conditions.jl:31-39:
    the flux of abiotic particles of species `sa` through face `pos` (a number,
    a face array or one per iteration), turned into particles at the cells
    next to the face by particles_from_bcs! every step."""
    sa, pos = str(sa).lstrip(":"), str(pos).lstrip(":")
    if sa not in model.individuals.abiotics:
        raise KeyError(sa)
    if pos not in _abi.BC_POS:
        raise ValueError(f"type BoundaryConditions has no field {pos}")
    if bc_value is None:
        v = None
    else:
        v = np.float32(bc_value) if np.ndim(bc_value) == 0 else np.asarray(bc_value, dtype=np.float32)
    model._native.set_particle_bc(int(sa[2:]) - 1, pos, v)
    setattr(model.individuals.abiotics[sa].bc, pos, v)


def validate_bcs(tracers, grid, nΔT):
    """validate_bcs (boundary_conditions.jl:62-83): every array boundary condition
    has the face's size, and a time-dependent one nΔT entries."""
    sizes = {"west": (grid.Ny, grid.Nz), "east": (grid.Ny, grid.Nz), "south": (grid.Nx, grid.Nz),
             "north": (grid.Nx, grid.Nz), "bottom": (grid.Nx, grid.Ny), "top": (grid.Nx, grid.Ny)}
    for name in tracers:
        for pos, size in sizes.items():
            bc = getattr(tracers[name].bc, pos)
            if bc is None or np.ndim(bc) == 0:
                continue
            if np.ndim(bc) == 2 and np.shape(bc) != size:
                raise ValueError(f"BC: grid mismatch, size(bc) must equal to {size} for a constant flux "
                                 "boundary condition.")
            if np.ndim(bc) == 3 and np.shape(bc) != size + (nΔT,):
                raise ValueError(f"BC: grid mismatch, size(bc) must equal to {size + (nΔT,)} for a "
                                 "time-dependent flux boundary condition.")


def TimeStep(model, ΔT, diags=None):
    """TimeStep!(model, ΔT, diags) — Model/time_step.jl:12-180, one device step."""
    model._native.step(np.float32(ΔT))
    model.t = np.float32(np.float32(model.iteration) * np.float32(ΔT))
