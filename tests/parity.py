"""Shared helpers of the parity tests: run the GPU library and the oracle from one
ModelSpec with identical inputs and compare every observable."""
import numpy as np

FLOAT_FIELDS = ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "gen", "age", "graz", "mort", "dvid")
INT_FIELDS = ("xi", "yi", "zi", "ac")


def same(a, b):
    """Bitwise equality up to the sign of zero and NaN payloads."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.kind == "f":
        return bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
    return bool(np.array_equal(a, b))


def first_diff(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    bad = np.nonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))[0] if a.dtype.kind == "f" else np.nonzero(a != b)[0]
    if len(bad) == 0:
        return None
    i = bad[0]
    return f"{len(bad)} differ; first at {i}: {a[i]!r} vs {b[i]!r}"


def compare_models(gpu, ref, nsp, tracer_ids=range(14), grid_fields=("par", "Chl", "pop", "diag_PAR", "PARF"),
                   diags=True, diag_tracers=(), extra_fields=()):
    """Returns a list of mismatch descriptions (empty = identical).  Besides the
    individuals and tracers: par, Chl, pop, the accumulated PAR diagnostic
    (diags.tracer.PAR, time_step.jl:167), PARF after the column scan, the
    per-species diagnostics (ptc included), and the tracer diagnostics of `diag_tracers`
    (time_step.jl:168-172); `extra_fields` adds the extra state columns of a
    mode (ex0 = RNA in MacroMolecularMode)."""
    bad = []
    if gpu.individual_steps() != ref.individual_steps():
        bad.append(f"individual steps: {gpu.individual_steps()} vs {ref.individual_steps()}")
    for q in diag_tracers:
        a, b = gpu.diag_tracer(q), ref.diag_tracer(q)
        if not same(a, b):
            bad.append(f"diag tracer {q}: {first_diff(a, b)}")
    for s in range(nsp):
        for f in FLOAT_FIELDS + INT_FIELDS + tuple(extra_fields):
            a, b = gpu.field(s, f), ref.field(s, f)
            if not same(a, b):
                bad.append(f"sp{s + 1}.{f}: {first_diff(a, b)}")
    for q in tracer_ids:
        a, b = gpu.tracer(q), ref.tracer(q)
        if not same(a, b):
            bad.append(f"tracer {q}: {first_diff(a, b)}")
    for g in grid_fields:
        a, b = gpu.grid_field(g), ref.grid_field(g)
        if not same(a, b):
            bad.append(f"{g}: {first_diff(a, b)}")
    if diags:
        for s in range(nsp):
            for d in ("num", "graz", "mort", "dvid", "ptc"):
                a, b = gpu.diag(s, d), ref.diag(s, d)
                if not same(a, b):
                    bad.append(f"diag sp{s + 1}.{d}: {first_diff(a, b)}")
    return bad
