"""Autotuner for the tiles engine's launch parameters.

Knobs (cmi_tuning): the LDS tile edge (cells), threads per workgroup, the
packets a wave takes per refill ("packets per wave" of a refill batch), and
the fixed chunk size of the per-round work descriptors.  The search is
coordinate descent from the defaults: each pass sweeps one knob at a time
with the others held at the best values so far, for `passes` passes, and
every configuration is timed `repeats` times (the median counts) from the
same field state, with device-event timing of the transport.  Results are
cached in a JSON file keyed by the device name and a workload signature.
"""
import json
import os
import statistics

DEFAULT = {"tile": 16, "block_threads": 512, "refill_batch": 32, "chunk": 8192}
SPACE = {
    "tile": [8, 16, 32],
    "block_threads": [256, 512],
    "refill_batch": [8, 16, 32, 64],
    "chunk": [4096, 8192, 16384, 32768],
}


def _key(cfg):
    return tuple(sorted(cfg.items()))


def coordinate_descent(evaluate, space=None, start=None, passes=2,
                       repeats=3, log=None):
    """Minimise evaluate(cfg) -> seconds over the space; returns
    (best cfg, best seconds, every timing taken as a list of dicts)."""
    space = dict(space or SPACE)
    best = dict(start or DEFAULT)
    for k in space:
        if best[k] not in space[k]:
            best[k] = space[k][0]
    seen = {}
    trials = []

    def measure(cfg):
        k = _key(cfg)
        if k not in seen:
            ts = [evaluate(dict(cfg)) for _ in range(repeats)]
            seen[k] = statistics.median(ts)
            trials.append({**cfg, "seconds": seen[k], "samples": ts})
            if log:
                log("autotune %s: %.3f ms" % (cfg, 1e3 * seen[k]))
        return seen[k]

    best_t = measure(best)
    for _ in range(passes):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cfg = dict(best, **{knob: v})
                t = measure(cfg)
                if t < best_t:
                    best, best_t, changed = cfg, t, True
        if not changed:
            break
    return best, best_t, trials


def load_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_cache(path, cache):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def cache_key(device, workload):
    return "%s|%s" % (device, workload)


def tune_context(prob, npackets, seed=1, iteration=3, space=None,
                 passes=2, repeats=3, log=None):
    """Tune the transport of one context (a warmstart.Problem) from its
    current field state, which is restored before every timed transport."""
    import numpy as np
    from . import ENGINE_TILES
    from . import workloads as W

    _, x = prob.ctx.get_fields()
    sc = [n // s for n, s in zip(prob.ncell, prob.nsub)]
    sp = dict(space or SPACE)
    sp["tile"] = [t for t in sp["tile"] if all(c % t == 0 for c in sc)]

    def evaluate(cfg):
        prob.ctx.set_engine(ENGINE_TILES, tile=[cfg["tile"]] * 3,
                            block_threads=cfg["block_threads"],
                            refill_batch=cfg["refill_batch"],
                            chunk_min=cfg["chunk"])
        prob.ctx.set_fields(xn=x)
        prob.ctx.reset_intensity()
        s = prob.ctx.transport(npackets, seed, iteration)
        return s["seconds"]

    best, t, trials = coordinate_descent(evaluate, sp, passes=passes,
                                         repeats=repeats, log=log)
    # leave the context on the best configuration and its original state
    evaluate(best)
    prob.ctx.set_fields(xn=x)
    prob.ctx.reset_intensity()
    return best, t, trials


def tune_problem_set(make_set, npackets, seed=1, iteration=3, space=None,
                     passes=2, repeats=3, log=None):
    """Tune a problem set.  make_set(cfg) returns a cmi.Batch built with that
    configuration and its fields and sources set (the tile edge is the
    subgrid of every problem, so it is fixed at creation); one set per tile
    edge is kept and its fields are re-set before each timed iteration."""
    sets = {}
    sp = dict(space or SPACE)

    def evaluate(cfg):
        t = cfg["tile"]
        if t not in sets:
            sets[t] = make_set(cfg)
        b = sets[t]
        b.retune(cfg)
        b.restore()
        r = b.iteration(npackets, seed, iteration, b.alpha, form=1)
        return r["seconds"]

    try:
        best, t, trials = coordinate_descent(evaluate, sp, passes=passes,
                                             repeats=repeats, log=log)
    finally:
        for b in sets.values():
            b.close()
    return best, t, trials
