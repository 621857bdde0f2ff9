"""Parameter sets (reference: src/Parameters/param_default.jl, param_update.jl).

Values are the reference's Float64 literals; conversion to Float32 happens where
the reference converts (`Dict{String, FT}` for bgc parameters, `FT.(tmp)` when a
species' NamedTuple is built, QuotaMode/plankton_generation.jl:22-31).
"""
import numpy as np

from ._abi import (PARAM_NAMES, BGC_NAMES, CARBON_PARAM_NAMES, MM_PARAM_NAMES, IRON_PARAM_NAMES, ABIOTIC_PARAM_NAMES,
                   MODE_QUOTA, MODE_CARBON, MODE_MACROMOLECULAR, MODE_IRONENERGY, NPARAMS)


class QuotaMode:
    """QuotaMode <: AbstractMode (src/PlanktonIndividuals.jl:73)."""

    def __repr__(self):
        return "QuotaMode()"


class CarbonMode:
    """CarbonMode <: AbstractMode (src/PlanktonIndividuals.jl:65)."""

    def __repr__(self):
        return "CarbonMode()"


class MacroMolecularMode:
    """MacroMolecularMode <: AbstractMode (src/PlanktonIndividuals.jl:77)."""

    def __repr__(self):
        return "MacroMolecularMode()"


class IronEnergyMode:
    """IronEnergyMode <: AbstractMode (src/PlanktonIndividuals.jl:81)."""

    def __repr__(self):
        return "IronEnergyMode()"


def bgc_params_default(FT=np.float32):
    """bgc_params_default(FT) — param_default.jl:45-71."""
    d = {
        "kw": 0.046, "kc": 0.04,
        "kDOC": 1 / 30 / 86400, "Nit": 1 / 30 / 86400, "kDON": 1 / 30 / 86400,
        "kDOP": 1 / 30 / 86400, "kPOC": 1 / 30 / 86400, "kPON": 1 / 30 / 86400,
        "kPOP": 1 / 30 / 86400,
        "κh": 0.0e-6, "κv": 0.0e-6, "κhP": 0.0e-6, "κvP": 0.0e-6,
        "shared_graz": 1.0,
        "w_sink_org": 2.31e-8, "w_sink_inorg": 2.31e-5,
        "lambda_POC": 5.8e-8, "lambda_min": 3.5e-10, "kdiss": 4.4e-7,
        "lambda_dust": 1.7e-3, "lambda_Fe": 1.7, "ligand": 6.0e-4, "DFeFrac": 0.01,
    }
    return {k: FT(v) for k, v in d.items()}


def mode_id(mode):
    """The ABI's physiology mode of a mode object (QuotaMode when None)."""
    if mode is None or isinstance(mode, QuotaMode):
        return MODE_QUOTA
    if isinstance(mode, CarbonMode):
        return MODE_CARBON
    if isinstance(mode, MacroMolecularMode):
        return MODE_MACROMOLECULAR
    if isinstance(mode, IronEnergyMode):
        return MODE_IRONENERGY
    raise NotImplementedError(f"{mode} physiology is not part of the MI355X hot path")


def param_names(mode):
    return {MODE_CARBON: CARBON_PARAM_NAMES, MODE_MACROMOLECULAR: MM_PARAM_NAMES,
            MODE_IRONENERGY: IRON_PARAM_NAMES}.get(mode_id(mode), PARAM_NAMES)


def phyt_params_default(N=1, mode=None):
    """phyt_params_default(N, ::QuotaMode) — param_default.jl:234-284,
    phyt_params_default(N, ::CarbonMode) — param_default.jl:290-323 and
    phyt_params_default(N, ::MacroMolecularMode) — param_default.jl:94-146 and
    phyt_params_default(N, ::IronEnergyMode) — param_default.jl:156-228
    (Float64 vectors)."""
    if mode_id(mode) == MODE_IRONENERGY:
        params = {
            "Nsuper": [1], "Cquota": [1.8e-11], "SA": [0.18], "mean": [1.2], "var": [0.3], "Chl2Cint": [0.10],
            "α": [4.5e-2], "Topt": [27.0], "Tmax": [30.0], "Ea": [5.3e4], "is_nr": [1.0], "is_croc": [0.0],
            "is_tric": [0.0], "PCmax": [7.8e-8], "VNH4max": [6.9e-6], "VNO3max": [6.9e-6], "VPO4max": [1.2e-6],
            "k_cf": [1.0e-5], "k_rs": [1.5e-6], "k_nr": [2.8e-6], "k_nf": [2.8e-6], "k_mtb": [3.5e-5],
            "e_cf": [9.0], "e_rs": [5.0], "e_nf": [8.0], "e_nr": [10.0], "k_Fe_ST2PS": [2.4e-5],
            "k_Fe_PS2ST": [1.2e-6], "k_Fe_ST2NR": [1.2e-5], "k_Fe_NR2ST": [1.2e-5], "k_Fe_ST2NF": [1.2e-5],
            "k_Fe_NF2ST": [1.2e-5], "KfePS": [3.0e-6], "KfeNR": [2.0e-6], "KfeNF": [5.0e-6], "KsatNH4": [0.005],
            "KsatNO3": [0.010], "KsatPO4": [0.003], "KSAFe": [2.77e-19], "qNO3max": [0.25], "qNH4max": [0.25],
            "qPmax": [0.02], "qFemax": [2.0e-5], "CHmax": [0.4], "Chl2N": [3.0], "R_NC": [16 / 106],
            "R_PC": [1 / 106], "NF_clock": [21600.0], "grz_P": [0.0], "dvid_P": [1e-4], "dvid_type": [1],
            "dvid_reg": [2.5], "dvid_reg2": [12.0], "mort_P": [5e-5], "mort_reg": [0.5], "grazFracC": [0.7],
            "grazFracN": [0.7], "grazFracP": [0.7], "grazFracFe": [0.1], "mortFracC": [0.5], "mortFracN": [0.5],
            "mortFracP": [0.5], "mortFracFe": [0.1], "max_ptc": [25],
        }
        return params if N == 1 else generate_n_species_params(N, params)
    if mode_id(mode) == MODE_MACROMOLECULAR:
        params = {
            "Nsuper": [1], "C_DNA": [1.8e-13], "var": [0.3], "RNA2DNA": [1.73], "PRO2DNA": [34.83],
            "CH2DNA": [23.3], "Chl2DNA": [3.45], "α": [2.0e-2], "Φ": [4.0e-5], "Topt": [27.0],
            "Tmax": [30.0], "Ea": [5.3e4], "PCmax": [6.2e-5], "VDOCmax": [0.0], "VNH4max": [6.9e-6],
            "VNO3max": [6.9e-6], "VPO4max": [1.2e-6], "KsatNH4": [0.005], "KsatNO3": [0.010],
            "KsatPO4": [0.003], "KsatDOC": [0.0], "NSTmax": [0.12], "PSTmax": [0.70], "CHmax": [0.4],
            "respir": [1.2e-6], "k_pro": [6.0e-5], "k_sat_pro": [4.5e-13], "k_dna": [1.6e-7],
            "k_sat_dna": [1.0e-15], "k_rna": [3.0e-7], "k_sat_rna": [1.0e-12], "Chl2N": [3.0],
            "R_NC_PRO": [1 / 4.5], "R_NC_DNA": [1 / 2.9], "R_PC_DNA": [1 / 11.1], "R_NC_RNA": [1 / 2.8],
            "R_PC_RNA": [1 / 10.7], "dvid_P": [1.0e-5], "grz_P": [0.0], "mort_P": [5e-5], "mort_reg": [0.5],
            "grazFracC": [0.7], "grazFracN": [0.7], "grazFracP": [0.7],
            "mortFracC": [0.5], "mortFracN": [0.5], "mortFracP": [0.5],
        }
        return params if N == 1 else generate_n_species_params(N, params)
    if mode_id(mode) == MODE_CARBON:
        params = {
            "Nsuper": [1], "Cquota": [1.8e-11], "mean": [1.2], "var": [0.3], "Chl2C": [0.10],
            "α": [2.0e-2], "Φ": [4.0e-5], "Topt": [27.0], "Tmax": [30.0], "Ea": [5.3e4],
            "PCmax": [4.2e-5], "respir": [1.2e-6], "f_T2B": [2.7e-6], "grz_P": [0.0], "dvid_P": [5e-5],
            "dvid_type": [1], "dvid_reg": [1.9], "dvid_reg2": [12.0], "mort_P": [5e-5], "mort_reg": [0.5],
            "grazFracC": [0.7], "mortFracC": [0.5], "thermal": [1.0], "is_bact": [0.0],
        }
        return params if N == 1 else generate_n_species_params(N, params)
    params = {
        "Nsuper": [1], "Cquota": [1.8e-11], "mean": [1.2], "var": [0.3], "Chl2Cint": [0.10],
        "α": [2.0e-2], "Φ": [4.0e-5], "Topt": [27.0], "Tmax": [30.0], "Ea": [5.3e4],
        "PCmax": [4.2e-5], "VDOCmax": [0.0], "VNH4max": [6.9e-6], "VNO3max": [6.9e-6],
        "VPO4max": [1.2e-6], "KsatNH4": [0.005], "KsatNO3": [0.010], "KsatPO4": [0.003],
        "KsatDOC": [0.0], "Nqmax": [0.25], "Pqmax": [0.02], "Cqmax": [0.4], "k_mtb": [3.5e-5],
        "respir": [1.2e-6], "Chl2N": [3.0], "R_NC": [16 / 106], "R_PC": [1 / 106],
        "grz_P": [0.0], "dvid_P": [1e-4], "dvid_type": [1], "dvid_reg": [2.5],
        "dvid_reg2": [12.0], "mort_P": [5e-5], "mort_reg": [0.5],
        "grazFracC": [0.7], "grazFracN": [0.7], "grazFracP": [0.7],
        "mortFracC": [0.5], "mortFracN": [0.5], "mortFracP": [0.5],
    }
    if N == 1:
        return params
    return generate_n_species_params(N, params)


def generate_n_species_params(N, params):
    """param_default.jl:345-352: replicate species 1's values N times."""
    return {k: [v[0]] * N for k, v in params.items()}


def update_bgc_params(tmp, FT=np.float32):
    """update_bgc_params(tmp::Dict, FT) — param_update.jl:10-22."""
    parameters = bgc_params_default(FT)
    for key, val in tmp.items():
        if key not in parameters:
            raise ValueError(f"PARAM: bgc parameter not found {key}")
        parameters[key] = FT(val)
    return parameters


def update_phyt_params(tmp, FT=np.float32, *, N=1, mode=None):
    """update_phyt_params(tmp::Dict, FT; N, mode) — param_update.jl:34-46."""
    parameters = phyt_params_default(N, mode if mode is not None else QuotaMode())
    for key, val in tmp.items():
        if key not in parameters:
            raise ValueError(f"PARAM: phyt parameter not found {key}")
        v = list(np.atleast_1d(np.asarray(val, dtype=np.float64)))
        parameters[key] = [FT(x) for x in v]
    return parameters


def species_param_matrix(phyt_params, nsp, mode=None):
    """(nsp, 64) float32 in the mode's param_names order (QuotaMode: 40 names;
    CarbonMode: 24; MacroMolecularMode: 47; IronEnergyMode: 64; the rest zero):
    `p = NamedTuple{param_names}(FT.(tmp))`."""
    names = param_names(mode)
    out = np.zeros((nsp, NPARAMS), dtype=np.float32)
    for s in range(nsp):
        for q, name in enumerate(names):
            if name not in phyt_params:
                raise ValueError(f"PARAM: parameter not found {name}")
            vals = phyt_params[name]
            out[s, q] = np.float32(np.float64(vals[s]))
    if "dvid_type" in names:                      # MacroMolecularMode divides by DNA content
        dt = out[:, names.index("dvid_type")]
        if np.any((dt < 1) | (dt > 3)):
            raise ValueError("Wrong cell division type, must be in 1 to 3")
    return out


def bgc_vector(bgc):
    return np.array([bgc[k] for k in BGC_NAMES], dtype=np.float32)


def abiotic_params_default(N=1):
    """abiotic_params_default(N) — param_default.jl:325-343."""
    params = {"Nsuper": [1], "Rd": [1.0e-4], "release_P": [1.0e-6], "sz_min": [1.0e-3], "Ktr": [0.0e-5]}
    if N == 1:
        return params
    return generate_n_species_params(N, params)


def update_abiotic_params(tmp, FT=np.float32, *, N=1):
    """update_abiotic_params(tmp::Dict, FT; N) — param_update.jl:48-69."""
    parameters = abiotic_params_default(N)
    for key, val in tmp.items():
        if key not in parameters:
            raise ValueError(f"PARAM: abiotic parameter not found {key}")
        v = list(np.atleast_1d(np.asarray(val, dtype=np.float64)))
        parameters[key] = [FT(x) for x in v]
    return parameters


def abiotic_param_matrix(params, nsa):
    """(nsa, 5) float32 in param_names order (Abiotic/particle_generation.jl:8-20):
    `p = NamedTuple{param_names}(FT.(tmp))`, tmp[i] = params[name][sa]."""
    out = np.zeros((nsa, len(ABIOTIC_PARAM_NAMES)), dtype=np.float32)
    for s in range(nsa):
        for q, name in enumerate(ABIOTIC_PARAM_NAMES):
            if name not in params:
                raise ValueError(f"PARAM: parameter not found {name}")
            out[s, q] = np.float32(np.float64(params[name][s]))
    return out
