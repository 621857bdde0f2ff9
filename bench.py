#!/usr/bin/env python3
"""Benchmark of the PlanktonIndividuals.jl per-timestep hot path on one MI355X.

Workload (default): the plain 3-D run — grid 256x256x64 (periodic x/y, bounded
z, Δx = Δy = 1 km, Δz = 10 m), two QuotaMode species with 4,194,304 individuals
each (capacity 8,388,608 each), a steady divergence-free velocity from 64
seeded Fourier modes (k^-5/3, rms 0.05 m/s), κh = κhP = 1e-2 and κv = κvP = 1e-5
m²/s, ΔT = 60 s, midday PAR and the default temperature profile.  Initial state
from the shared random stream (seeded); inputs are resident in HBM before the
timed region.

One step = one TimeStep! (RK4 + diffusion + binning + PAR + physiology with
division/death + consumption + DST3FL tracer update of 14 tracers).

Prints ONE JSON line.  value = individual-steps/s (Σ live individuals per step
÷ step time); tracer cell-steps/s, the roofline of the dominant kernel (HIP
events around each launch during the timed region) and the CPU baseline (the
oracle, a CPU restatement of the reference step, on a bounded sample of the same
workload) are reported beside it.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK = 8.0e12    # MI355X HBM3E, bytes/s


def load_pkg():
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(name, n_live, g):
    """Compulsory HBM bytes of one launch (each logical array counted once per pass
    its semantics need; float32 = 4 B, flag = 1 B, key/cell/index = 4 B)."""
    ncell = g.ncells_padded
    nint = g.Nx * g.Ny * g.Nz
    if name == "adv":       # R key,x,y,z,flags,cell  W x,y,z,cell,flags  + velocity (vel₀ = vel₁: 3 fields)
        return 38 * n_live + 12 * ncell
    if name == "phys":      # R key,src + record 53 B, W record 53 B + contrib 25 B, + 9 gathered fields/cell
        return 139 * n_live + 36 * nint
    if name == "plk":       # R seg, key, contrib 25 B/ind; W 10 plk fields per cell
        return 29 * n_live + 44 * ncell
    if name == "sweep_x":   # R 14 tracers + u, W 14 temps
        return (56 + 4 + 56) * nint
    if name == "sweep_y":   # R 14 temps + 14 tracers + v, W 14 temps
        return (56 + 56 + 4 + 56) * nint
    if name == "final":     # R 14 temps + 14 tracers + 10 plk + w, W 14 tracers
        return (56 + 56 + 40 + 4 + 56) * nint
    if name == "par":       # R Chl, W par, RMW diag PAR
        return 16 * nint
    return None


def cpu_baseline(pi, seconds_hint=20.0):
    """The oracle (FP32 reference order, one thread) on a bounded sample of the same
    workload: the 128x128x64 quarter of the domain at the same density
    (2 x 1,048,576 individuals), one step."""
    sys.path.insert(0, REPO)
    from oracle import orc
    spec, uvw = pi.workloads.config_3d(n=128, nz=64, N=1048576, nsp=2, cap=2097152, seed=2)
    g = spec.grid
    om = orc.OracleModel(spec)
    parf = np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32)
    temp = pi.default_temperature(g, 60.0, 1)[..., 12].astype(np.float32)
    n0 = om.count_active(0) + om.count_active(1)
    t0 = time.perf_counter()
    om.update(60.0, 1, vel=uvw, parf=parf, temp=temp)
    dt = time.perf_counter() - t0
    return {"value": n0 / dt, "unit": "individual-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle (CPU restatement of the reference CPU() step, FP32, 1 thread): one step of the "
                      "128x128x64 quarter domain with 2 x 1,048,576 individuals (same density and forcing)",
            "seconds": dt, "tracer_cell_steps_per_s": g.Nx * g.Ny * g.Nz / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="horizontal grid size")
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--individuals", type=int, default=4194304, help="per species at start")
    ap.add_argument("--species", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: a plain device copy of 256 MiB (k_probe_copy, known traffic, to "
                         "scale the counters), then --warmup + --steps steps without timing; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    pi = load_pkg()
    t0 = time.perf_counter()
    spec, (u, v, w) = pi.workloads.config_3d(n=args.n, nz=args.nz, N=args.individuals, nsp=args.species,
                                             cap=2 * args.individuals, seed=2)
    g = spec.grid
    log(f"[bench] workload built in {time.perf_counter() - t0:.1f} s")
    model = pi.PlanktonModel(pi.HIPArch(0), spec=spec)
    nat = model._native
    dt = 60.0
    parf = np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32)      # default_PARF, 11:00-12:00
    temp = pi.default_temperature(g, dt, 1)[..., 12].astype(np.float32)
    nat.set_velocity(u, v, w)
    nat.set_parf(parf)
    nat.set_temperature(temp)
    log(f"[bench] model resident on the GPU after {time.perf_counter() - t0:.1f} s")

    if args.pmc_probe:
        lib = pi._native.load_library()
        pi._native.check(lib, "pi_", lib.pi_probe_copy(1 << 28))      # k_probe_copy: 256 MiB read + 256 MiB written
        log("[bench] copy probe: 268435456 bytes read, 268435456 bytes written")
        nat.run(dt, args.warmup + args.steps)
        nat.sync()
        log(f"[bench] {args.warmup + args.steps} steps done")
        return
    nat.run(dt, args.warmup)
    nat.sync()
    live0 = sum(nat.count_active(s) for s in range(args.species))
    nat.enable_timing(True)
    t1 = time.perf_counter()
    nat.run(dt, args.steps)
    nat.sync()
    t2 = time.perf_counter()
    kt = nat.kernel_times()
    nat.enable_timing(False)
    live1 = sum(nat.count_active(s) for s in range(args.species))
    wall = t2 - t1
    ms_step = 1e3 * wall / args.steps
    live = 0.5 * (live0 + live1)
    ind_steps = live * args.steps / wall
    cell_steps = g.Nx * g.Ny * g.Nz * args.steps / wall
    log(f"[bench] {args.steps} steps in {wall:.3f} s ({ms_step:.3f} ms/step), live {live0} -> {live1}")

    kernels = {}
    for name, (ms, cnt) in kt.items():
        b = algorithmic_bytes(name, live, g)
        avg = ms / max(cnt, 1)
        kernels[name] = {"ms_per_launch": avg, "launches": cnt,
                         "share": ms / max(1e-9, sum(x[0] for x in kt.values())),
                         "gbps_algorithmic": (b / (avg * 1e-3) / 1e9) if b else None}
    dom = max((k for k in kernels if kernels[k]["gbps_algorithmic"] is not None),
              key=lambda k: kernels[k]["ms_per_launch"] * kernels[k]["launches"])
    ach = kernels[dom]["gbps_algorithmic"] * 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": ach / HBM_PEAK, "traffic": None,
            "bytes_per_launch": algorithmic_bytes(dom, live, g)}
    base = None
    if not args.no_cpu:
        del model, nat
        log("[bench] CPU baseline (oracle) ...")
        base = cpu_baseline(pi)
        log(f"[bench] CPU baseline {base['value']:.4g} individual-steps/s in {base['seconds']:.1f} s")

    out = {
        "metric": "individual-steps/s & tracer cell-steps/s (1 GPU); ensemble p50/p99 step latency",
        "value": ind_steps,
        "unit": "individual-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Fourier velocity (64 modes, k^-5/3, rms 0.05 m/s); initial state from the shared stream",
        "config": {"workload": f"plain 3-D run {args.n}x{args.n}x{args.nz}, {args.species} species x "
                               f"{args.individuals} individuals (capacity {2 * args.individuals}), dt 60 s",
                   "grid": [args.n, args.n, args.nz], "species": args.species,
                   "individuals_per_species": args.individuals, "dt": dt},
        "tracer_cell_steps_per_s": cell_steps,
        "live_individuals": [live0, live1],
        "roofline": roof,
        "cpu_baseline": base,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
