"""ctypes mirror of include/pi_hip.h (descriptor structs, ids, prototypes).

The same prototypes describe any library exposing the ABI under a prefix, so
one table serves libpi_hip.so ("pi_").
"""
import ctypes as C
import numpy as np

# ---- status codes (pi_status) ------------------------------------------------
PI_OK, PI_ERR_ARG, PI_ERR_CAPACITY, PI_ERR_HIP, PI_ERR_UNSUPPORTED, PI_ERR_NOMEM = range(6)
PERIODIC, BOUNDED = 0, 1

TRACER_NAMES = ("DIC", "NH4", "NO3", "PO4", "DFe", "DOC", "DON", "DOP",
                "PFe_inorg", "POC", "PON", "POP", "PFe_bio", "Dust")
TRACER_ID = {n: i for i, n in enumerate(TRACER_NAMES)}

# QuotaMode parameter order (QuotaMode/plankton_generation.jl:16-20)
PARAM_NAMES = ("Nsuper", "Cquota", "mean", "var", "Chl2Cint", "α", "Φ", "Topt", "Tmax", "Ea",
               "PCmax", "VDOCmax", "VNO3max", "VNH4max", "VPO4max",
               "KsatDOC", "KsatNH4", "KsatNO3", "KsatPO4", "Cqmax", "Nqmax", "Pqmax",
               "Chl2N", "R_NC", "R_PC", "k_mtb", "respir", "grz_P", "dvid_type", "dvid_P",
               "dvid_reg", "dvid_reg2", "mort_P", "mort_reg", "grazFracC", "grazFracN",
               "grazFracP", "mortFracC", "mortFracN", "mortFracP")
# CarbonMode parameter order (CarbonMode/plankton_generation.jl:14-16)
CARBON_PARAM_NAMES = ("Nsuper", "Cquota", "mean", "var", "α", "Φ", "Topt", "Tmax", "Ea", "PCmax", "Chl2C",
                      "respir", "f_T2B", "grz_P", "dvid_type", "dvid_P", "dvid_reg", "dvid_reg2", "mort_P",
                      "mort_reg", "grazFracC", "mortFracC", "thermal", "is_bact")
# MacroMolecularMode parameter order (MacroMolecularMode/plankton_generation.jl:15-21)
MM_PARAM_NAMES = ("Nsuper", "C_DNA", "var", "PRO2DNA", "RNA2DNA", "CH2DNA", "Chl2DNA",
                  "α", "Φ", "Topt", "Tmax", "Ea",
                  "PCmax", "VDOCmax", "VNO3max", "VNH4max", "VPO4max", "KsatDOC", "KsatNH4", "KsatNO3",
                  "KsatPO4", "CHmax", "NSTmax", "PSTmax",
                  "Chl2N", "R_NC_PRO", "R_NC_DNA", "R_NC_RNA", "R_PC_DNA", "R_PC_RNA", "respir",
                  "k_pro", "k_sat_pro", "k_rna", "k_sat_rna", "k_dna", "k_sat_dna",
                  "dvid_P", "grz_P", "mort_P", "mort_reg", "grazFracC", "grazFracN", "grazFracP",
                  "mortFracC", "mortFracN", "mortFracP")
# IronEnergyMode parameter order (IronEnergyMode/plankton_generation.jl:28-40)
IRON_PARAM_NAMES = ("Nsuper", "Cquota", "SA", "mean", "var", "Chl2Cint",
                    "α", "Topt", "Tmax", "Ea", "is_nr", "is_croc", "is_tric",
                    "PCmax", "VNO3max", "VNH4max", "VPO4max",
                    "k_cf", "k_rs", "k_nr", "k_nf", "k_mtb",
                    "e_cf", "e_rs", "e_nr", "e_nf",
                    "k_Fe_ST2PS", "k_Fe_PS2ST", "k_Fe_ST2NR", "k_Fe_NR2ST", "k_Fe_ST2NF", "k_Fe_NF2ST",
                    "KfePS", "KfeNR", "KfeNF", "KsatNH4", "KsatNO3", "KsatPO4", "KSAFe",
                    "CHmax", "qNH4max", "qNO3max", "qPmax", "qFemax",
                    "Chl2N", "R_NC", "R_PC", "NF_clock",
                    "grz_P", "dvid_type", "dvid_P", "dvid_reg", "dvid_reg2", "mort_P", "mort_reg",
                    "grazFracC", "grazFracN", "grazFracP", "grazFracFe",
                    "mortFracC", "mortFracN", "mortFracP", "mortFracFe", "max_ptc")
MODE_QUOTA, MODE_CARBON, MODE_MACROMOLECULAR, MODE_IRONENERGY = 0, 1, 2, 3
NPARAMS = 64                                     # floats per species row (PI_NPARAMS)
# abiotic particle parameter order (Abiotic/particle_generation.jl:8)
ABIOTIC_PARAM_NAMES = ("Nsuper", "Rd", "release_P", "sz_min", "Ktr")
MAX_ABIOTIC = 4                                  # PI_MAXSA
ABIO_FIELD_ID = {"x": 0, "y": 1, "z": 2, "xi": 3, "yi": 4, "zi": 5, "ac": 6}
BGC_NAMES = ("kw", "kc", "kDOC", "Nit", "kDON", "kDOP", "kPOC", "kPON", "kPOP", "κh", "κv",
             "κhP", "κvP", "shared_graz", "w_sink_org", "w_sink_inorg", "lambda_POC",
             "lambda_min", "kdiss", "lambda_dust", "lambda_Fe", "ligand", "DFeFrac")

FLOAT_FIELDS = ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "gen", "age",
                "graz", "mort", "dvid")
INT_FIELDS = ("xi", "yi", "zi")
NEX = 8                                          # extra float columns (PI_NEX)
EX_FIELDS = tuple(f"ex{k}" for k in range(NEX))
FIELD_ID = {n: i for i, n in enumerate(FLOAT_FIELDS + INT_FIELDS + ("ac",) + EX_FIELDS)}
# the columns of the other modes in the ABI's columns (PI_F_Bd, PI_F_PRO, ... of pi_hip.h)
FIELD_ALIAS = {"Bd": "Cq", "PRO": "Bm", "DNA": "Sz", "RNA": "ex0", "CH": "Cq", "NST": "Nq", "PST": "Pq",
               "qNH4": "Nq", "qP": "Pq", "qNO3": "ex0", "qFe": "ex1", "qFePS": "ex2", "qFeNR": "ex3",
               "qFeNF": "ex4", "tdark": "ex5", "exEn": "ex6", "ptc": "ex7"}
# the state columns of each mode's StructArray, in the ABI's names
MODE_COLUMNS = {MODE_QUOTA: ("Sz", "Bm", "Cq", "Nq", "Pq", "Chl"),
                MODE_CARBON: ("Sz", "Bm", "Cq", "Chl"),
                MODE_MACROMOLECULAR: ("Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "ex0"),
                MODE_IRONENERGY: ("Sz", "Bm", "Cq", "Nq", "Pq", "Chl") + EX_FIELDS}
# the names of each mode's StructArray columns (the ABI's column names where they differ)
MODE_NAMES = {MODE_QUOTA: ("Sz", "Bm", "Cq", "Nq", "Pq", "Chl"), MODE_CARBON: ("Sz", "Bm", "Bd", "Chl"),
              MODE_MACROMOLECULAR: ("DNA", "PRO", "CH", "NST", "PST", "Chl", "RNA"),
              MODE_IRONENERGY: ("Sz", "Bm", "CH", "qNH4", "qP", "Chl", "qNO3", "qFe", "qFePS", "qFeNR", "qFeNF",
                                "tdark", "exEn", "ptc")}

GRID_ID = {"par": 0, "par₀": 1, "par0": 1, "Chl": 2, "pop": 3, "temp": 4, "PARF": 5,
           "u": 6, "v": 7, "w": 8, "diag_PAR": 9}
PLK_BASE = 10
DIAG_ID = {"num": 0, "graz": 1, "mort": 2, "dvid": 3, "ptc": 4}
# flux boundary conditions (pi_set_tracer_bc): positions and getbc's kinds
BC_POS = {"west": 0, "east": 1, "south": 2, "north": 3, "bottom": 4, "top": 5}
BCK_NONE, BCK_CONST, BCK_FIELD, BCK_SERIES = range(4)
# launch shapes (pi_set_launch_shape): name -> (id, default)
SHAPES = {"xy_run": (0, 16), "final_run": (1, 16), "adv_blocks": (2, 16384), "phys_blocks": (3, 4096),
          "cellsum_per_cu": (4, 8), "plk_blocks": (5, 1 << 22)}

fp = C.POINTER(C.c_float)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)


class GridDesc(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
                ("Hx", C.c_int32), ("Hy", C.c_int32), ("Hz", C.c_int32),
                ("topo_x", C.c_int32), ("topo_y", C.c_int32), ("topo_z", C.c_int32),
                ("dx", C.c_float), ("dy", C.c_float),
                ("dzC", fp), ("dzF", fp), ("landmask", u8p)]


class Metrics(C.Structure):
    """pi_metrics: LatLonGrid spacings, areas and volumes (padded, column-major)."""
    _fields_ = [(n, fp) for n in ("dxC", "dyC", "dxF", "dyF", "Az", "Ax", "Ay", "Vol")]


class ModelDesc(C.Structure):
    _fields_ = [("grid", GridDesc), ("n_species", C.c_int32), ("capacity", C.c_int64),
                ("params", fp), ("bgc", fp), ("rng_seed", C.c_uint64),
                ("n_diag_tracers", C.c_int32), ("diag_tracers", C.POINTER(C.c_int32)),
                ("mode", C.c_int32), ("metrics", C.POINTER(Metrics)),
                ("n_abiotic", C.c_int32), ("abiotic_params", fp), ("n_interactions", C.c_int32),
                ("interactions", C.POINTER(C.c_int32)), ("n_releases", C.c_int32),
                ("releases", C.POINTER(C.c_int32)), ("max_candidates", C.c_int32)]


class IndividualsHost(C.Structure):
    _fields_ = [(n, fp) for n in FLOAT_FIELDS] + [(n, i64p) for n in INT_FIELDS] + [("ac", u8p), ("ex", fp * NEX)]


def ptr(a, ctype):
    """Pointer into a contiguous numpy array (caller keeps `a` alive)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def make_desc(grid, n_species, capacity, params, bgc, seed, diag_tracers=(), mode=MODE_QUOTA,
              abiotic_params=None, palat=None):
    """`palat` = (interaction pairs, release pairs, max_candidates), the pairs
    0-based (species, abiotic species)."""
    """Build a ModelDesc; returns (desc, keepalive) — keep the second alive while desc is used."""
    dzC = np.ascontiguousarray(grid.dzC, dtype=np.float32)
    dzF = np.ascontiguousarray(grid.dzF, dtype=np.float32)
    lm = None if grid.no_land else np.asfortranarray(grid.landmask, dtype=np.uint8)
    params = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
    bgc = np.ascontiguousarray(bgc, dtype=np.float32)
    dt = np.ascontiguousarray(np.array(diag_tracers, dtype=np.int32))
    g = GridDesc(grid.Nx, grid.Ny, grid.Nz, grid.Hx, grid.Hy, grid.Hz,
                 grid.topology_ids[0], grid.topology_ids[1], grid.topology_ids[2],
                 np.float32(grid.Δx), np.float32(grid.Δy),
                 ptr(dzC, C.c_float), ptr(dzF, C.c_float),
                 None if lm is None else lm.ctypes.data_as(u8p))
    marrs, mt = (), None
    if hasattr(grid, "metric_arrays"):                 # LatLonGrid
        marrs = grid.metric_arrays()
        mt = Metrics(*(ptr(a, C.c_float) for a in marrs))
    ap = None
    if abiotic_params is not None and len(abiotic_params):
        ap = np.ascontiguousarray(abiotic_params, dtype=np.float32).reshape(-1)
    d = ModelDesc(g, n_species, capacity, ptr(params, C.c_float), ptr(bgc, C.c_float),
                  C.c_uint64(seed), len(dt), ptr(dt, C.c_int32) if len(dt) else None, int(mode),
                  C.pointer(mt) if mt is not None else None,
                  0 if ap is None else len(ap) // len(ABIOTIC_PARAM_NAMES), ptr(ap, C.c_float))
    intac, release, mc = palat if palat is not None else ((), (), 25)
    pi_ = np.ascontiguousarray(np.array(intac, dtype=np.int32).reshape(-1))
    pr = np.ascontiguousarray(np.array(release, dtype=np.int32).reshape(-1))
    d.n_interactions, d.n_releases, d.max_candidates = len(pi_) // 2, len(pr) // 2, int(mc)
    d.interactions = ptr(pi_, C.c_int32) if len(pi_) else None
    d.releases = ptr(pr, C.c_int32) if len(pr) else None
    return d, (dzC, dzF, lm, params, bgc, dt, marrs, mt, ap, pi_, pr)


def individuals_view(arrays):
    """IndividualsHost over a dict of full-capacity numpy arrays (keep `arrays` alive)."""
    h = IndividualsHost()
    for n in FLOAT_FIELDS:
        a = arrays.get(n)
        setattr(h, n, ptr(a, C.c_float) if a is not None else None)
    for n in INT_FIELDS:
        a = arrays.get(n)
        setattr(h, n, ptr(a, C.c_int64) if a is not None else None)
    a = arrays.get("ac")
    h.ac = ptr(a, C.c_uint8) if a is not None else None
    for k, n in enumerate(EX_FIELDS):
        a = arrays.get(n)
        h.ex[k] = ptr(a, C.c_float) if a is not None else None
    return h


def declare(lib, prefix, model_t=C.c_void_p):
    """Attach argtypes/restype for the ABI functions exported under `prefix`."""
    M = model_t
    proto = {
        "last_error": (C.c_char_p, []),
        "model_destroy": (C.c_int, [M]),
        "upload_individuals": (C.c_int, [M, C.c_int32, C.POINTER(IndividualsHost)]),
        "upload_tracer": (C.c_int, [M, C.c_int32, fp]),
        "set_velocity": (C.c_int, [M, fp, fp, fp]),
        "set_parf": (C.c_int, [M, fp]),
        "set_temperature": (C.c_int, [M, fp]),
        "step": (C.c_int, [M, C.c_float]),
        "download_field": (C.c_int, [M, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]),
        "download_tracer": (C.c_int, [M, C.c_int32, fp]),
        "download_grid": (C.c_int, [M, C.c_int32, fp, C.c_int64]),
        "download_diag": (C.c_int, [M, C.c_int32, C.c_int32, fp]),
        "download_diag_tracer": (C.c_int, [M, C.c_int32, fp]),
        "reset_diags": (C.c_int, [M]),
        "get_iteration": (C.c_int, [M, i64p]),
        "count_active": (C.c_int, [M, C.c_int32, i64p]),
        "individual_steps": (C.c_int, [M, i64p]),
        "set_tracer_bc": (C.c_int, [M, C.c_int32, C.c_int32, C.c_int32, fp, C.c_int64]),
        "set_particle_bc": (C.c_int, [M, C.c_int32, C.c_int32, C.c_int32, fp, C.c_int64]),
        "upload_abiotic": (C.c_int, [M, C.c_int32, fp, fp, fp, u8p]),
        "download_abiotic": (C.c_int, [M, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]),
        "download_abiotic_diag": (C.c_int, [M, C.c_int32, fp]),
    }
    for name, (res, args) in proto.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
    return lib
