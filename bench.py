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


# The logical kernels of the step with their compulsory HBM bytes per step
# (each logical array counted once per pass its semantics need; float32 = 4 B,
# flag = 1 B, cell id = 4 B, counter-based uniforms = 0 B), per individual
# (N = live individuals), per interior cell and per capacity slot, and the
# launch groups (pi_kernel_times names) and rocprof kernels that implement them.
LOGICAL = {
    "K_adv":    dict(groups=["adv"], per_ind=29, per_cell=24, per_slot=0,
                     kernels=["k_adv"]),
    "K_phys":   dict(groups=["phys"], per_ind=76, per_cell=36, per_slot=0,
                     kernels=["k_phys"]),
    "K_bin":    dict(groups=["bin", "plk"], per_ind=4, per_cell=48, per_slot=0,
                     kernels=["k_cellcount", "k_place_stay", "k_place_small", "k_mixsort", "k_bigcells",
                              "k_cellsum", "k_plk", "k_scan_partials", "k_scan_final"]),
    "K_div":    dict(groups=["marks+scan"], per_ind=0, per_cell=0, per_slot=2,
                     kernels=["k_marks", "k_capcheck"]),
    "K_dst3fl": dict(groups=["sweep_xy", "sweep_x", "fill", "sweep_y", "final"], per_ind=0, per_cell=164,
                     per_slot=0, kernels=["k_xy", "k_sweep_x", "k_sweep_y", "k_fill_line", "k_final"]),
    "K_par":    dict(groups=["par"], per_ind=0, per_cell=8, per_slot=0, kernels=["k_par"]),
}


ALL_KERNELS = {k for L in LOGICAL.values() for k in L["kernels"]}


def logical_bytes(name, n_live, n_cells, n_slots):
    L = LOGICAL[name]
    return L["per_ind"] * n_live + L["per_cell"] * n_cells + L["per_slot"] * n_slots


def step_bytes(n_live, n_cells, n_slots):
    """B_step = 109 N + 2 N_cap + 280 N_cells: the sum of the logical kernels."""
    return sum(logical_bytes(k, n_live, n_cells, n_slots) for k in LOGICAL)


def rocprof_base(name):
    """Kernel name without arguments, template arguments and the landmask
    variant digit (k_final0 / k_final<false> -> k_final)."""
    n = name.split("(")[0].replace("void ", "").strip()
    n = n.split("<")[0]
    return n[:-1] if n[-1:] in ("0", "1") and n[:-1] in ALL_KERNELS else n


VALU_PEAK = 256 * 2.4e9   # MI355X: 256 CUs x 2.4 GHz, one wave64 VALU instruction per CU per clock (non-packed)


def measured_valu(logical):
    """VALU wave-instructions per step of a logical kernel (SQ_INSTS_VALU from the
    committed counter summary), or None."""
    path = os.path.join(REPO, "profiles", "traffic_3d_256x256x64_2sp.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    tot = [v["valu_insts_per_step"] for k, v in t["kernels"].items()
           if rocprof_base(k) in LOGICAL[logical]["kernels"] and "valu_insts_per_step" in v]
    return sum(tot) if tot else None


def measured_traffic(logical):
    """HBM bytes per step of a logical kernel from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy probe), or None."""
    path = os.path.join(REPO, "profiles", "traffic_3d_256x256x64_2sp.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    fs, ws = t["scale"]["fetch"], t["scale"]["write"]
    tot, seen = 0.0, False
    for k, v in t["kernels"].items():
        if rocprof_base(k) in LOGICAL[logical]["kernels"]:
            tot += (v["fetch_kib_per_step"] * fs + v["write_kib_per_step"] * ws) * 1024.0
            seen = True
    return tot if seen else None


def rocpd_summary(db, title, cmd, skip_steps=0):
    """Markdown table of a rocprofv3 kernel trace (rocpd SQLite output).  With
    skip_steps > 0 only the dispatches that start after the end of the
    skip_steps-th step (a step ends with its k_finish) are counted: model
    creation, the first step after an upload (which sorts every individual from
    slot order) and the warm-up are left out, so the averages are those of
    steady steps."""
    import sqlite3
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)")]
    t0c = next(k for k in ("start", "start_ts", "begin") if k in cols)
    t1c = next(k for k in ("end", "end_ts", "stop") if k in cols)
    where = ""
    if skip_steps > 0:
        fin = [r[0] for r in c.execute(f'select "{t1c}" from kernels where name like "%k_finish%" order by "{t0c}"')]
        if len(fin) > skip_steps:
            where = f'where "{t0c}" > {fin[skip_steps - 1]}'
    rows = c.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                     f"from kernels {where} group by name order by sum(duration) desc").fetchall()
    tot = sum(r[2] for r in rows)
    steady = (f"Steady steps only: dispatches after the end of step {skip_steps} (model creation, the first step "
              "after the upload, which sorts every individual, and the warm-up are not counted).") if where else ""
    out = [f"# {title}", "", "Command (from the repository root, one MI355X, ROCm 7.2):", "", "```", cmd, "```", "",
           steady, "", "Kernel trace statistics (durations in microseconds):", "",
           "| kernel | calls | total us | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for n, k, sm, a, mn, mx in rows:
        out.append(f"| {n.split('(')[0]} | {k} | {sm / 1e3:.1f} | {a / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
                   f"{100 * sm / tot:.1f} |")
    return "\n".join(out)


def pmc_summary(fetch_db, write_db, steps, valu_db=None):
    """Per-kernel FETCH_SIZE / WRITE_SIZE (KiB) (and SQ_INSTS_VALU) per step of
    --pmc-probe runs: the dispatches of the last `steps` steps, a step being the
    dispatches up to and including a k_finish (the first step after an upload
    sorts from slot order and is not typical; model creation precedes the
    steps), and the scale factors of the copy probe (k_probe_copy reads and
    writes 262144 KiB)."""
    import collections
    import sqlite3

    def load(db, cname):
        c = sqlite3.connect(db)
        per = collections.defaultdict(float)
        names = {}
        for d, k, n, v in c.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection"):
            if n.startswith(cname):
                per[d] += v
                names[d] = k.split("(")[0]
        return per, names

    def per_step(db, cname):
        per, names = load(db, cname)
        ids = sorted(per)
        fin = [d for d in ids if rocprof_base(names[d]) == "k_finish"]
        lo = fin[-steps - 1] if len(fin) > steps else -1
        out = collections.defaultdict(lambda: [0.0, 0])
        for d in ids:
            if lo < d <= fin[-1]:
                out[names[d]][0] += per[d] / steps
                out[names[d]][1] += 1
        probe = [d for d in ids if "k_probe_copy" in names[d]]
        return out, (per[probe[0]] if probe else None)

    f, fp = per_step(fetch_db, "FETCH_SIZE")
    w, wp = per_step(write_db, "WRITE_SIZE")
    v, _ = per_step(valu_db, "SQ_INSTS_VALU") if valu_db else ({}, None)
    scale = {"fetch": 262144.0 / fp, "write": 262144.0 / wp}
    kernels = {}
    for k, (fv, cnt) in f.items():
        kernels[k] = {"fetch_kib_per_step": fv, "write_kib_per_step": w[k][0] if k in w else 0.0,
                      "dispatches_per_step": cnt // steps}
        if k in v:
            kernels[k]["valu_insts_per_step"] = v[k][0]
    return {"scale": scale, "steps_averaged": steps, "kernels": kernels}


def cpu_threads():
    """Host threads for the all-cores CPU baseline: the CPU share of this process
    (affinity), capped by OMP_NUM_THREADS when it is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(pi, seconds_hint=10.0):
    """The oracle (FP32 reference order) on a bounded sample of the same
    workload: the 128x128x64 quarter of the domain at the same density
    (2 x 1,048,576 individuals), one untimed step (every individual sorted from
    scratch), then as many timed steps as fit in about `seconds_hint` s; first
    on one thread, then on every host core of this process (the element-wise
    loops on OpenMP threads; the ordered per-cell sums stay serial, so the
    results are the same bits)."""
    sys.path.insert(0, REPO)
    from oracle import orc
    lib = orc.load()
    spec, uvw = pi.workloads.config_3d(n=128, nz=64, N=1048576, nsp=2, cap=2097152, seed=2)
    g = spec.grid
    parf = np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32)
    temp = pi.default_temperature(g, 60.0, 1)[..., 12].astype(np.float32)
    res = {}
    for nt in (1, cpu_threads()):
        lib.orc_set_threads(nt)
        om = orc.OracleModel(spec)
        t0 = time.perf_counter()
        om.update(60.0, 1, vel=uvw, parf=parf, temp=temp)
        first = time.perf_counter() - t0
        k = max(1, min(8, int(round(seconds_hint / max(first, 1e-3)))))
        n0 = om.count_active(0) + om.count_active(1)
        t0 = time.perf_counter()
        om.update(60.0, k, vel=uvw, parf=parf, temp=temp)
        dt = time.perf_counter() - t0
        n1 = om.count_active(0) + om.count_active(1)
        res[nt] = {"value": 0.5 * (n0 + n1) * k / dt, "steps": k, "seconds": dt,
                   "tracer_cell_steps_per_s": g.Nx * g.Ny * g.Nz * k / dt}
        del om
        log(f"[bench] CPU baseline, {nt} thread(s): {res[nt]['value']:.4g} individual-steps/s ({k} steps, {dt:.1f} s)")
    lib.orc_set_threads(1)
    nt = max(res)
    one = res[1]
    return {"value": res[nt]["value"], "unit": "individual-steps/s", "cores": nt, "kind": "port",
            "sample": f"oracle (CPU restatement of the reference CPU() step, FP32, {nt} OpenMP threads): "
                      f"{res[nt]['steps']} steps of the 128x128x64 quarter domain with 2 x 1,048,576 individuals "
                      "(same density and forcing), after one untimed step",
            "seconds": res[nt]["seconds"], "tracer_cell_steps_per_s": res[nt]["tracer_cell_steps_per_s"],
            "single_thread": {"value": one["value"], "cores": 1, "steps": one["steps"], "seconds": one["seconds"],
                              "tracer_cell_steps_per_s": one["tracer_cell_steps_per_s"]}}


def ensemble_leg(pi, members=256, steps=20, seq_steps=4, seed=4, batched_only=False, tune=True):
    """Batched ensemble: `members` drawn from the ensemble table, stepped
    one-by-one through the plain path (pi_run + sync per member) and then all
    together (pi_batch_run); ms per ensemble step of each and the ratio."""
    table = pi.workloads.ensemble_table(members, seed=seed)
    t0 = time.perf_counter()
    nats = []
    for e in table:
        spec, vel = pi.workloads.ensemble_member(e)
        g = spec.grid
        nat = pi.PlanktonModel(pi.HIPArch(0), spec=spec)._native
        if vel is not None:
            nat.set_velocity(*vel)
        nat.set_parf(np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32))
        nat.set_temperature(pi.default_temperature(g, 60.0, 1)[..., 12].astype(np.float32))
        nats.append(nat)
    log(f"[bench] ensemble: {members} members created in {time.perf_counter() - t0:.1f} s")
    t_seq = float("nan")
    if not batched_only:
        for nat in nats:
            nat.run(60.0, 2)
            nat.sync()
        t0 = time.perf_counter()
        for nat in nats:
            nat.run(60.0, seq_steps)
            nat.sync()
        t_seq = (time.perf_counter() - t0) / seq_steps
    batch = pi._native.NativeBatch()
    for nat in nats:
        batch.add(nat)
    batch.run(60.0, 2)
    batch.sync()
    t_tune = time.perf_counter()
    tuned = batch.autotune(60.0, steps=3, reps=2) if tune else None
    t_tune = time.perf_counter() - t_tune
    if tuned:
        log(f"[bench] ensemble: budget {tuned['budget']}, shapes per class {tuned['classes']} "
            f"(searched in {t_tune:.1f} s)")
    c0 = sum(nat.individual_steps() for nat in nats)
    t0 = time.perf_counter()
    batch.run(60.0, steps)
    batch.sync()
    t_b = (time.perf_counter() - t0) / steps
    live = (sum(nat.individual_steps() for nat in nats) - c0) / steps   # individuals stepped per ensemble step
    batch_budget = batch.budget
    batch.close()
    for nat in nats:
        nat.close()
    kinds = [sum(1 for e in table if e["kind"] == k) for k in range(3)]
    log(f"[bench] ensemble: one-by-one {t_seq * 1e3:.2f} ms, batched {t_b * 1e3:.2f} ms per ensemble step")
    return {"workload": f"{members} members (0-D boxes 1x1x1 with 1k-64k individuals, 1-D columns 1x1x64 with "
                        f"8k-256k, 2-D slabs 64-256 squared with 16k-1M; VNH4max, VNO3max, VPO4max, dvid_P x "
                        f"U[0.8,1.2]), dt 60 s, seed {seed}",
            "members": members, "members_by_kind": kinds, "live_individuals": live,
            "batched_ms_per_step": t_b * 1e3, "one_by_one_ms_per_step": t_seq * 1e3,
            "speedup_vs_one_by_one": t_seq / t_b,
            "batched_individual_steps_per_s": live / t_b, "one_by_one_individual_steps_per_s": live / t_seq,
            "batched_steps": steps, "one_by_one_steps": seq_steps,
            "block_budget": batch_budget,
            "launch_shapes_by_class": None if tuned is None else tuned["classes"],
            "autotune": None if tuned is None else {"trials_ms_per_step": tuned["trials"], "seconds": t_tune}}


def _oracle_member(pi, entry, steps, dt=60.0):
    """A served member re-run by the oracle with update!'s staging (the
    reference's diel PAR and temperature slices, the steady velocity)."""
    sys.path.insert(0, REPO)
    from oracle import orc
    spec, vel = pi.workloads.ensemble_member(entry, host_state=True, cap_factor=4)
    g = spec.grid
    PARF = pi.default_PARF(g, dt, steps).astype(np.float32)
    temp = pi.default_temperature(g, dt, steps).astype(np.float32)
    idx = lambda t, s: int(np.floor(np.float32(t) / np.float32(s)))
    om = orc.OracleModel(spec)
    om.update(dt, steps, vel=vel, parf_fn=lambda t: PARF[:, :, idx(t, 3600.0)],
              temp_fn=lambda t: temp[:, :, :, idx(t, 3600.0)])
    return om


SERVE_FIELDS = ("x", "y", "z", "Sz", "Bm", "Cq", "Nq", "Pq", "Chl", "gen", "age", "ac")


def serving_leg(pi, rate=50.0, window=6.0, cap_s=150.0, budget_gib=200, seed=5, n_oracle=4, n_alone=8):
    """Served ensemble stream: requests from the ensemble table (lengths
    100-2000 steps) arrive as a Poisson process at `rate`/s for `window` s,
    are admitted by device footprint within `budget_gib` and stepped in one
    running batch; per-step and end-to-end latency and throughput (live
    individuals counted on the device at every step).  Validation of finished
    members: the `n_oracle` smallest (individuals x steps) against the oracle
    (the CPU restatement of the reference step) and `n_alone` more, spread over
    the sizes and including the largest, against the same member run alone on
    the GPU; bit for bit."""
    from importlib import import_module
    sv = import_module(pi.__name__ + ".serving")
    table = pi.workloads.ensemble_table(4096, seed=seed, lengths=(100, 2000))
    rng = np.random.default_rng(seed + 6)
    arr = np.cumsum(rng.exponential(1.0 / rate, 4096))
    arr = arr[arr < window]
    make = lambda e: (lambda: pi.workloads.ensemble_member(e, cap_factor=4))
    order = sorted(range(len(arr)), key=lambda i: table[i]["N"] * table[i]["steps"])
    check_o = order[:n_oracle]
    rest = order[n_oracle:]
    by_n = sorted(rest, key=lambda i: table[i]["N"])
    pick = {by_n[-1]}
    for q in np.linspace(0, len(by_n) - 1, n_alone - 1):
        pick.add(by_n[int(q)])
    check_a = sorted(pick)[:n_alone]
    check = set(check_o) | set(check_a)
    states = {}

    def grab(req, sim):
        if req.rid in check:
            nat = sim.model._native
            states[req.rid] = ([nat.field(0, f) for f in SERVE_FIELDS], [nat.tracer(q) for q in range(14)])

    srv = sv.EnsembleServer(budget=budget_gib * 2 ** 30, workers=8, on_finish=grab)
    t0 = time.perf_counter()
    k = 0
    last = t0
    try:
        while True:
            now = time.perf_counter() - t0
            while k < len(arr) and arr[k] <= now:
                srv.submit(sv.Request(k, make(table[k]), table[k]["steps"], arrival=t0 + arr[k]))
                k += 1
            busy = srv.step()
            if not busy and k >= len(arr):
                break
            if not busy:
                time.sleep(0.001)
            if time.perf_counter() - last > 10:
                last = time.perf_counter()
                log(f"[bench] serving t={last - t0:.0f}s arrived={k} active={len(srv.ens.sims)} "
                    f"finished={len(srv.finished)} used={srv.used / 2 ** 30:.1f} GiB")
            if time.perf_counter() - t0 > cap_s:
                break
        srv.drain(deadline=t0 + cap_s)
    finally:
        wall = time.perf_counter() - t0
        st = srv.stats()
        failed = [r for r in srv.finished if r.error]
        srv.close()

    def same(f, t, ref):
        ok = all(np.array_equal(ref.field(0, n), a) for n, a in zip(SERVE_FIELDS, f))
        return ok and all(np.array_equal(ref.tracer(q), t[q]) for q in range(14))

    ok_o = ok_a = 0
    for i in check_o:
        if i in states:
            om = _oracle_member(pi, table[i], table[i]["steps"])
            ok_o += int(same(*states[i], om))
            del om
    for i in check_a:
        if i in states:
            spec, vel = pi.workloads.ensemble_member(table[i], cap_factor=4)
            sim = sv.simulation_for(spec, vel, table[i]["steps"])
            pi.update(sim)
            ok_a += int(same(*states[i], sim.model._native))
            sim.model._native.close()
    n_o = sum(1 for i in check_o if i in states)
    n_a = sum(1 for i in check_a if i in states)
    log(f"[bench] serving: {k} requests, {st['finished']} finished in {wall:.1f} s, step p50 {st['p50_step_ms']:.2f} ms "
        f"p99 {st['p99_step_ms']:.2f} ms; {ok_o}/{n_o} identical to the oracle, {ok_a}/{n_a} identical alone")
    return {"workload": f"Poisson arrivals {rate:g}/s for {window:g} s from the ensemble member table (seed {seed}), "
                        f"lengths U[100,2000] steps, capacity 4 N, dt 60 s, admission by device footprint within "
                        f"{budget_gib} GiB",
            "requests": k, "finished": st["finished"], "failed": len(failed), "wall_s": wall,
            "p50_step_ms": st["p50_step_ms"], "p99_step_ms": st["p99_step_ms"], "mean_step_ms": st["mean_step_ms"],
            "steps": st["steps"], "p50_end_to_end_s": st["p50_e2e_s"], "p99_end_to_end_s": st["p99_e2e_s"],
            "member_steps": st["member_steps"], "individual_steps": st["individual_steps"],
            "individual_steps_per_s": st["individual_steps"] / wall,
            "validated": {"against_oracle": n_o, "identical_to_oracle": ok_o,
                          "oracle_members": [dict(N=table[i]["N"], steps=table[i]["steps"], size=table[i]["size"])
                                             for i in check_o],
                          "against_alone": n_a, "identical_alone": ok_a,
                          "alone_members": [dict(N=table[i]["N"], steps=table[i]["steps"], size=table[i]["size"])
                                            for i in check_a]}}


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
    ap.add_argument("--no-large", action="store_true", help="skip the large single-instance leg")
    ap.add_argument("--workload", choices=("plain", "large"), default="plain",
                    help="large: only the large single instance (for profiling; --steps/--warmup apply to it)")
    ap.add_argument("--full-steps", type=int, default=500,
                    help="after the timed and per-kernel steps, run the plain workload to this many steps")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain 3-D run (for profiling the other legs); prints their results only")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-ensemble leg")
    ap.add_argument("--no-interaction", action="store_true", help="skip the particle-interaction leg")
    ap.add_argument("--ens-members", type=int, default=256)
    ap.add_argument("--ens-batched-only", action="store_true",
                    help="ensemble leg without the one-by-one pass (profiling the batched launches)")
    ap.add_argument("--serve-window", type=float, default=6.0, help="seconds of arrivals")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="arrivals per second")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: a plain device copy of 256 MiB (k_probe_copy, known traffic, to "
                         "scale the counters), then --warmup + --steps steps without timing; prints no JSON")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (no autotune search before the timed steps)")
    ap.add_argument("--shapes", default=None,
                    help='launch shapes as JSON, e.g. \'{"xy_run": 32}\', set after any search (profiling runs)')
    ap.add_argument("--rocpd-summary", nargs="+", metavar="ARG",
                    help="DB TITLE CMD [SKIP_STEPS]: print the markdown kernel table of a rocprofv3 kernel-trace "
                         "database (steady steps only with SKIP_STEPS) and exit")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DB",
                    help="FETCH_DB WRITE_DB OUT_JSON [VALU_DB]: write the per-step FETCH/WRITE (and SQ_INSTS_VALU) "
                         "summary of --pmc-probe runs and exit")
    args = ap.parse_args()
    if args.rocpd_summary:
        a = args.rocpd_summary
        print(rocpd_summary(a[0], a[1], a[2], int(a[3]) if len(a) > 3 else 0))
        return
    if args.pmc_summary:
        res = pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.steps,
                          args.pmc_summary[3] if len(args.pmc_summary) > 3 else None)
        with open(args.pmc_summary[2], "w") as f:
            json.dump(res, f, indent=1)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    pi = load_pkg()
    if args.no_plain:
        out = {"ensemble": None if args.no_ensemble else ensemble_leg(pi, members=args.ens_members,
                                                                      batched_only=args.ens_batched_only,
                                                                      tune=not args.no_tune),
               "serving": None if args.no_serve else serving_leg(pi, rate=args.serve_rate, window=args.serve_window)}
        print(json.dumps(out), flush=True)
        return
    if args.workload == "large":
        if args.pmc_probe:
            large_leg(pi, steps=args.steps, warmup=args.warmup, probe=True, shapes=parse_shapes(args.shapes))
            return
        print(json.dumps({"large_instance": large_leg(pi, steps=args.steps, warmup=args.warmup,
                                                      tune=not args.no_tune, shapes=parse_shapes(args.shapes))}),
              flush=True)
        return

    t0 = time.perf_counter()
    spec, (u, v, w) = pi.workloads.config_3d(n=args.n, nz=args.nz, N=args.individuals, nsp=args.species,
                                             cap=2 * args.individuals, seed=2)
    g = spec.grid
    log(f"[bench] workload built in {time.perf_counter() - t0:.1f} s")
    dt = 60.0
    parf = np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32)      # default_PARF, 11:00-12:00
    temp = pi.default_temperature(g, dt, 1)[..., 12].astype(np.float32)
    tuned = None
    if not args.no_tune and not args.pmc_probe:
        tuned = tune_clone(pi, spec, (u, v, w), parf, temp, dt)
    model = pi.PlanktonModel(pi.HIPArch(0), spec=spec)
    nat = model._native
    nat.set_velocity(u, v, w)
    nat.set_parf(parf)
    nat.set_temperature(temp)
    if tuned:
        nat.set_shape(**tuned["shapes"])
    if args.shapes:
        nat.set_shape(**parse_shapes(args.shapes))
    log(f"[bench] model resident on the GPU after {time.perf_counter() - t0:.1f} s")

    if args.pmc_probe:
        probe_copy(pi)
        nat.run(dt, args.warmup + args.steps)
        nat.sync()
        log(f"[bench] {args.warmup + args.steps} steps done")
        return
    nat.run(dt, args.warmup)
    nat.sync()
    live0 = sum(nat.count_active(s) for s in range(args.species))
    c0 = nat.individual_steps()
    t1 = time.perf_counter()
    nat.run(dt, args.steps)                # steady steps replay the captured step graph
    nat.sync()
    t2 = time.perf_counter()
    c1 = nat.individual_steps()
    live1 = sum(nat.count_active(s) for s in range(args.species))
    wall = t2 - t1
    ms_step = 1e3 * wall / args.steps
    live = (c1 - c0) / args.steps          # individuals stepped per step, counted on the device
    ind_steps = (c1 - c0) / wall
    cell_steps = g.Nx * g.Ny * g.Nz * args.steps / wall
    log(f"[bench] {args.steps} steps in {wall:.3f} s ({ms_step:.3f} ms/step), live {live0} -> {live1}")
    # per-kernel breakdown: HIP events around each launch group, eager launches
    nkt = max(5, min(args.steps, 20))
    nat.enable_timing(True)
    nat.run(dt, nkt)
    nat.sync()
    kt = nat.kernel_times()
    nat.enable_timing(False)
    # the rest of the workload's 500 steps (the population keeps evolving)
    done = args.warmup + args.steps + nkt
    full = None
    if args.full_steps > done:
        c2 = nat.individual_steps()
        t3 = time.perf_counter()
        nat.run(dt, args.full_steps - done)
        nat.sync()
        t4 = time.perf_counter()
        c3 = nat.individual_steps()
        nrest = args.full_steps - done
        full = {"steps": args.full_steps, "timed_steps": nrest, "ms_per_step": 1e3 * (t4 - t3) / nrest,
                "individual_steps_per_s": (c3 - c2) / (t4 - t3),
                "live_at_end": sum(nat.count_active(s) for s in range(args.species))}
        log(f"[bench] steps {done + 1}-{args.full_steps}: {full['ms_per_step']:.3f} ms/step, "
            f"live at the end {full['live_at_end']}")

    n_cells = g.Nx * g.Ny * g.Nz
    n_slots = args.species * 2 * args.individuals
    kernels = {name: {"ms_per_launch": ms / max(cnt, 1), "launches": cnt} for name, (ms, cnt) in kt.items()}
    logical = logical_table(kt, nkt, live, n_cells, n_slots)
    dom = max(logical, key=lambda k: logical[k]["ms_per_step"])
    ach = logical[dom]["gbps_algorithmic"] * 1e9
    traffic = measured_traffic(dom)
    roof = {"bound": "hbm", "kernel": dom, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": ach / HBM_PEAK, "traffic": traffic, "bytes_per_launch": logical[dom]["bytes_per_step"],
            "launch_ms": logical[dom]["ms_per_step"]}
    valu = measured_valu(dom)
    issue = None
    if valu:
        ms = logical[dom]["ms_per_step"]
        issue = {"bound": "valu-issue", "achieved": valu / (ms * 1e-3), "peak": VALU_PEAK,
                 "unit": "wave64 VALU instructions/s", "frac": valu / (ms * 1e-3) / VALU_PEAK,
                 "valu_insts_per_step": valu}
    bstep = step_bytes(live, n_cells, n_slots)
    step_roof = {"bytes_per_step": bstep, "achieved": bstep / (ms_step * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                 "unit": "GB/s", "frac": bstep / (ms_step * 1e-3) / HBM_PEAK}
    shapes_used = nat.shape()
    del model, nat
    large = None
    if not args.no_large:
        large = large_leg(pi, tune=not args.no_tune)
    ens = None
    if not args.no_ensemble:
        ens = ensemble_leg(pi, members=args.ens_members, tune=not args.no_tune)
    serve = None
    if not args.no_serve:
        serve = serving_leg(pi, rate=args.serve_rate, window=args.serve_window)
    inter = None
    if not args.no_interaction:
        inter = interaction_leg(pi)
    base = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        base = cpu_baseline(pi)

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
        "data": "synthetic: seeded Fourier velocity (64 modes, k^-5/3, rms 0.05 m/s); tracers uniform plus 5 % noise; "
                "initial state from the shared stream",
        "config": {"workload": f"plain 3-D run {args.n}x{args.n}x{args.nz}, {args.species} species x "
                               f"{args.individuals} individuals (capacity {2 * args.individuals}), dt 60 s",
                   "grid": [args.n, args.n, args.nz], "species": args.species,
                   "individuals_per_species": args.individuals, "dt": dt},
        "tracer_cell_steps_per_s": cell_steps,
        "live_individuals": [live0, live1],
        "launch_shapes": shapes_used,
        "autotune": None if tuned is None else {"trials_ms_per_step": tuned["trials"], "seconds": tuned["seconds"],
                                                "searched_on": "a second instance of the same workload"},
        "roofline": roof,
        "issue_roofline": issue,
        "step_roofline": step_roof,
        "full_run": full,
        "cpu_baseline": base,
        "large_instance": large,
        "ensemble": ens,
        "interaction": inter,
        "serving": serve,
        "logical_kernels": logical,
        "kernels": kernels,
    }
    print(json.dumps(out), flush=True)


def probe_copy(pi):
    lib = pi._native.load_library()
    pi._native.check(lib, "pi_", lib.pi_probe_copy(1 << 28))      # k_probe_copy: 256 MiB read + 256 MiB written
    log("[bench] copy probe: 268435456 bytes read, 268435456 bytes written")


def logical_table(kt, nkt, live, n_cells, n_slots):
    """Per logical kernel: ms per step (HIP events around its launch groups),
    algorithmic bytes per step and the algorithmic GB/s."""
    logical = {}
    for lk, L in LOGICAL.items():
        ms = sum(kt[gname][0] for gname in L["groups"] if gname in kt) / nkt     # per step
        b = logical_bytes(lk, live, n_cells, n_slots)
        logical[lk] = {"ms_per_step": ms, "bytes_per_step": b, "gbps_algorithmic": b / (ms * 1e-3) / 1e9 if ms else None,
                       "launch_groups": L["groups"]}
    return logical


def tune_clone(pi, spec, uvw, parf, temp, dt, warmup=3, steps=10, reps=3):
    """Launch shapes for the workload, searched on a second instance of the same
    spec (same initial state and inputs) so the measured model's state does not
    advance: autotune.tune_model after `warmup` steps.  Returns its report."""
    t0 = time.perf_counter()
    clone = pi.PlanktonModel(pi.HIPArch(0), spec=spec)._native
    try:
        clone.set_velocity(*uvw)
        clone.set_parf(parf)
        clone.set_temperature(temp)
        clone.run(dt, warmup)
        clone.sync()
        res = pi.autotune.tune_model(clone, dt, steps=steps, reps=reps, log=log)
    finally:
        clone.close()
    res["seconds"] = time.perf_counter() - t0
    log(f"[bench] launch shapes {res['shapes']} (searched in {res['seconds']:.1f} s)")
    return res


def parse_shapes(text):
    """--shapes: JSON {name: value} of launch shapes, or None."""
    return json.loads(text) if text else None


def large_leg(pi, steps=200, warmup=2, probe=False, tune=True, shapes=None):
    """The large single instance: 512 x 512 x 128 (periodic x/y, bounded z),
    4 species x 8,388,608 individuals (capacity 2 x that), a 128-mode Fourier
    flow, dt = 30 s, 200 steps; initial state drawn on the device.  Steps per
    second, individual- and cell-steps/s, and per logical kernel the time (HIP
    events, eager launches) and algorithmic GB/s."""
    N, nsp = 8388608, 4
    t0 = time.perf_counter()
    spec, (u, v, w) = pi.workloads.config_3d(n=512, nz=128, N=N, nsp=nsp, cap=2 * N, seed=3, nmodes=128,
                                             host_state=False)
    g = spec.grid
    dt = 30.0
    parf = np.full((g.Nx, g.Ny), 1823.7921666666664, np.float32)
    temp = pi.default_temperature(g, dt, 1)[..., 12].astype(np.float32)
    tuned = tune_clone(pi, spec, (u, v, w), parf, temp, dt, steps=4, reps=2) if tune and not probe else None
    model = pi.PlanktonModel(pi.HIPArch(0), spec=spec)
    nat = model._native
    nat.set_velocity(u, v, w)
    nat.set_parf(parf)
    nat.set_temperature(temp)
    if tuned:
        nat.set_shape(**tuned["shapes"])
    if shapes:
        nat.set_shape(**shapes)
    log(f"[bench] large instance resident on the GPU after {time.perf_counter() - t0:.1f} s")
    if probe:
        probe_copy(pi)
        nat.run(dt, warmup + steps)
        nat.sync()
        log(f"[bench] large instance: {warmup + steps} steps done")
        return None
    nat.run(dt, warmup)
    nat.sync()
    c0 = nat.individual_steps()
    t1 = time.perf_counter()
    nat.run(dt, steps)
    nat.sync()
    wall = time.perf_counter() - t1
    c1 = nat.individual_steps()
    nkt = 5
    nat.enable_timing(True)
    nat.run(dt, nkt)
    nat.sync()
    kt = nat.kernel_times()
    nat.enable_timing(False)
    live = (c1 - c0) / steps
    n_cells = g.Nx * g.Ny * g.Nz
    logical = logical_table(kt, nkt, live, n_cells, nsp * 2 * N)
    bstep = step_bytes(live, n_cells, nsp * 2 * N)
    ms = 1e3 * wall / steps
    out = {"workload": "large single instance 512x512x128, 4 species x 8388608 individuals (capacity 16777216), "
                       "128-mode Fourier flow, dt 30 s",
           "steps": steps, "ms_per_step": ms, "individual_steps_per_s": (c1 - c0) / wall,
           "tracer_cell_steps_per_s": n_cells * steps / wall, "live_individuals_per_step": live,
           "step_roofline": {"bytes_per_step": bstep, "achieved": bstep / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                             "unit": "GB/s", "frac": bstep / (ms * 1e-3) / HBM_PEAK},
           "logical_kernels": logical,
           "launch_shapes": nat.shape(),
           "autotune": None if tuned is None else {"trials_ms_per_step": tuned["trials"], "seconds": tuned["seconds"]},
           "kernels": {name: {"ms_per_launch": t / max(cnt, 1), "launches": cnt} for name, (t, cnt) in kt.items()}}
    log(f"[bench] large instance: {steps} steps in {wall:.3f} s ({ms:.3f} ms/step), "
        f"{out['individual_steps_per_s']:.4g} individual-steps/s")
    del model, nat
    return out


def interaction_leg(pi, steps=20):
    """Phytoplankton-particle interaction and release (particle_interaction.jl,
    time_step.jl:119-159) on an IronEnergyMode model: 128x128x32, 1,048,576
    individuals and 524,288 abiotic particles in a cellular flow, with one
    interaction and one release pair against the same model without pairs;
    the `interact` launch group timed with HIP events."""
    g = pi.RectilinearGrid(size=(128, 128, 32), x=(0, 128e3), y=(0, 128e3), z=(0, -320))
    out = {"workload": "IronEnergyMode 128x128x32, 1048576 individuals (capacity 2097152), 524288 abiotic particles "
                       "(Rd 300 m, release_P 1e-4), cellular flow, dt 60 s", "steps": steps}
    for pairs in (False, True):
        pal = pi.Palat(intac=[("sp1", "sa1")], release=[("sp1", "sa1")]) if pairs else pi.Palat()
        spec = pi.ModelSpec(g, mode=pi.IronEnergyMode(), max_individuals=2 ** 21, seed=7,
                            phyto=pi.phyto_setup({"max_ptc": [3.0]}, [2 ** 20], 1),
                            abiotic=pi.abiotic_setup({"Rd": [300.0], "release_P": [1e-4]}, [2 ** 19], 1, palat=pal),
                            bgc_params={"κhP": 1.0, "κh": 1.0})
        nat = pi.PlanktonModel(pi.HIPArch(), spec=spec)._native
        nat.set_velocity(*pi.workloads.cellular_flow(g, U=0.5, L=32e3))
        nat.set_parf(np.full((128, 128), 500.0, np.float32))
        nat.set_temperature(np.full((128, 128, 32), 20.0, np.float32))
        nat.run(60.0, 3)
        nat.sync()
        t0 = time.perf_counter()
        nat.run(60.0, steps)
        nat.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        nat.enable_timing(True)
        nat.run(60.0, 5)
        nat.sync()
        kt = nat.kernel_times()
        nat.enable_timing(False)
        key = "with_pairs" if pairs else "without_pairs"
        out[key] = {"ms_per_step": ms, "interact_ms_per_step": kt.get("interact", (0.0, 0))[0] / 5,
                    "particles_active": int(nat.abiotic_field(0, "ac").sum()),
                    "ptc_held": float(nat.field(0, "ptc").sum())}
        log(f"[bench] interaction ({key}): {ms:.3f} ms/step, interact group "
            f"{out[key]['interact_ms_per_step']:.3f} ms/step")
        del nat
    return out


if __name__ == "__main__":
    main()
