"""Launch-shape autotuner.

The kernels of a step split their work over workgroups in ways that change the
speed but never the results (pi_set_launch_shape): the rows per work item of
the fused x/y tracer sweep, the levels per column run of the z sweep, and the
workgroup counts of advection, physiology, the per-cell Chl/pop sums and the
per-cell exchange sums.  The best values depend on the grid's shape and on the
individuals per cell, so they are searched by timing real steps:

* coordinate descent, one shape at a time, over a short list of candidates;
* each candidate is timed `reps` times over `steps` steps (after warm-up
  steps that rebuild the step's launches or its captured graph) and its
  minimum kept;
* a candidate replaces the incumbent only when it is faster by more than
  `margin` (a fraction), so a choice is never made on timing noise;
* a batch is searched per member-shape class (0-D boxes, 1-D columns, 2-D
  slabs, 3-D grids): a candidate is applied to every member of the class and
  the whole batched step is timed; its block budget is searched first.

The steps taken while tuning are real steps of the model (or of every member):
tune a separate instance when the state must not advance.  Results are cached
per device and shape class in a JSON file (the `PI_TUNE_CACHE` environment
variable, default ~/.cache/pi_amd/launch_shapes.json); `apply_cached` sets a
cached choice on a new model of the same class without timing anything.
"""
import json
import os
import time

from . import _abi

CANDIDATES = {
    "xy_run": (8, 16, 32, 64),
    "final_run": (8, 16, 32, 64),
    "adv_blocks": (4096, 8192, 16384, 32768),
    "phys_blocks": (2048, 4096, 8192, 16384),
    "cellsum_per_cu": (4, 8, 16),
    "plk_blocks": (4096, 16384, 65536, 1 << 22),
}
BUDGETS = (4096, 8192, 16384, 32768, 65536)


def _cache_path():
    return os.environ.get("PI_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "pi_amd",
                                                          "launch_shapes.json")


def _load_cache():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, value):
    path = _cache_path()
    data = _load_cache()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass                                  # a read-only home: the result is still returned


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0).replace(" ", "_")
    except Exception:
        return "gpu"


def grid_class(grid):
    """0d (one cell), 1d (one axis longer than 1), 2d, 3d."""
    return f"{sum(1 for n in (grid.Nx, grid.Ny, grid.Nz) if n > 1)}d"


def model_key(nat):
    """Cache key of a single model: class, sizes, species, capacity, land."""
    g = nat.grid
    land = 0 if g.no_land else 1
    return (f"{_device_name()}|model|{grid_class(g)}|{g.Nx}x{g.Ny}x{g.Nz}|sp{nat.nsp}|cap{nat.cap}|"
            f"land{land}")


def _candidates(name, grids):
    """Candidates of one shape, without run lengths that exceed every grid's
    extent (they are the same launch as the extent itself)."""
    c = CANDIDATES[name]
    if name == "xy_run":
        top = max(g.Ny for g in grids)
    elif name == "final_run":
        top = max(g.Nz for g in grids)
    else:
        return c
    keep = sorted({min(v, top) for v in c})
    return tuple(keep)


def _time(run_steps, sync, steps, reps, warm):
    run_steps(warm)
    sync()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        run_steps(steps)
        sync()
        best = min(best, (time.perf_counter() - t0) * 1e3 / steps)
    return best


def _descend(knobs, get, put, timer, margin, log):
    """Coordinate descent; knobs = [(label, candidates)].  Returns the trials."""
    trials = {}
    base = timer()
    for label, cands in knobs:
        cur = get(label)
        res = {cur: base}
        for v in cands:
            if v == cur:
                continue
            put(label, v)
            res[v] = timer()
        best = min(res, key=res.get)
        if best != cur and res[best] < res[cur] * (1.0 - margin):
            put(label, best)
            base = res[best]
        else:
            put(label, cur)
            base = res[cur]
        trials[label] = {str(k): round(v, 4) for k, v in res.items()}
        if log:
            log(f"[autotune] {label}: " + ", ".join(f"{k}: {v:.3f}" for k, v in res.items())
                + f" ms -> {get(label)}")
    return trials


def tune_model(nat, dt, steps=10, reps=3, margin=0.015, warm=2, shapes=None, cache=True, log=None):
    """Searches the launch shapes of one model (its state advances by the steps
    timed).  Returns {"shapes": {...}, "trials": {shape: {value: ms/step}},
    "ms_per_step": best}."""
    names = list(shapes or CANDIDATES)
    knobs = [(n, _candidates(n, [nat.grid])) for n in names]

    def timer():
        return _time(lambda k: nat.run(dt, k), nat.sync, steps, reps, warm)

    trials = _descend(knobs, lambda n: nat.shape()[n], lambda n, v: nat.set_shape(**{n: v}), timer, margin, log)
    out = {"shapes": nat.shape(), "trials": trials,
           "ms_per_step": min(min(t.values()) for t in trials.values()) if trials else None}
    if cache:
        _store_cache(model_key(nat), out["shapes"])
    return out


def apply_cached(nat):
    """Sets the cached shapes of this model's class; False if none is cached."""
    v = _load_cache().get(model_key(nat))
    if not v:
        return False
    nat.set_shape(**{k: x for k, x in v.items() if k in _abi.SHAPES})
    return True


def batch_key(batch):
    classes = {}
    for m in batch.members:
        classes[grid_class(m.grid)] = classes.get(grid_class(m.grid), 0) + 1
    return f"{_device_name()}|batch|" + ",".join(f"{k}x{v}" for k, v in sorted(classes.items()))


def tune_batch(batch, dt, steps=4, reps=3, margin=0.015, warm=1, shapes=None, budgets=BUDGETS, cache=True,
               log=None):
    """Searches a batch: its block budget, then every launch shape per
    member-shape class (applied to all members of the class).  Every member's
    state advances by the steps timed.  Returns {"budget", "classes": {class:
    shapes}, "trials", "ms_per_step"}."""
    by_class = {}
    for m in batch.members:
        by_class.setdefault(grid_class(m.grid), []).append(m)
    names = list(shapes or CANDIDATES)
    state = {"budget": getattr(batch, "budget", 16384)}

    def get(label):
        if label == "budget":
            return state["budget"]
        cls, name = label.split(":")
        return by_class[cls][0].shape()[name]

    def put(label, v):
        if label == "budget":
            batch.set_budget(v)
            state["budget"] = v
            return
        cls, name = label.split(":")
        for m in by_class[cls]:
            m.set_shape(**{name: v})

    def timer():
        return _time(lambda k: batch.run(dt, k), batch.sync, steps, reps, warm)

    knobs = [("budget", tuple(budgets))]
    for cls, ms in sorted(by_class.items()):
        knobs += [(f"{cls}:{n}", _candidates(n, [m.grid for m in ms])) for n in names]
    batch.set_budget(state["budget"])
    trials = _descend(knobs, get, put, timer, margin, log)
    batch.budget = state["budget"]
    out = {"budget": state["budget"],
           "classes": {cls: ms[0].shape() for cls, ms in sorted(by_class.items())},
           "trials": trials,
           "ms_per_step": min(min(t.values()) for t in trials.values()) if trials else None}
    if cache:
        _store_cache(batch_key(batch), {"budget": out["budget"], "classes": out["classes"]})
    return out
