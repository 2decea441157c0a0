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
    "tile": [8, 16],      # a 32^3 tile needs 384 KiB of LDS
    "block_threads": [256, 512],
    "refill_batch": [8, 16, 24, 28, 32, 36, 40, 48, 64],
    "chunk": [4096, 8192, 12288, 14336, 16384, 24576, 32768],
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
            try:
                ts = [evaluate(dict(cfg)) for _ in range(repeats)]
            except Exception as e:  # an infeasible configuration
                if log:
                    log("autotune %s: rejected (%s)" % (cfg, e))
                ts = [float("inf")]
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
    from . import ENGINE_TILES

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


def main(argv=None):
    """python -m parallelcmacionize_amd.autotune --workload batched64 \\
           --out parallelcmacionize_amd/data/tuning.json"""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["batched64", "clumpy256"],
                    default="batched64")
    ap.add_argument("--out", required=True)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args(argv)

    def log(*m):
        print(*m, flush=True)

    t0 = time.time()
    if a.workload == "batched64":
        from . import batched as B
        spec = B.problem_set_spec()
        fields = B.build_fields(spec)
        ps = B.ProblemSet(spec, fields)
        for it in range(3):                      # a mid-convergence state
            ps.iteration(seed=1, iteration=it)
        ps.snapshot()
        state = ps.xn_state
        npk = ps.npackets
        ps.close()

        def make(cfg):
            p = B.ProblemSet(spec, fields, tile=cfg["tile"])
            p.xn_state = state
            return p
        space = dict(SPACE)
        space["tile"] = [8, 16]                  # must divide 48
        best, t, trials = tune_problem_set(make, npk, seed=3, iteration=7,
                                           space=space, passes=a.passes,
                                           repeats=a.repeats, log=log)
    else:
        from . import warmstart as WS
        box, n, pos, L = WS.clumpy_problem()
        prob = WS.Problem(box, n, pos, L)
        prob.iterate(100_000_000, seed=1, tol=0.0, max_it=3)
        best, t, trials = tune_context(prob, 100_000_000, seed=3, iteration=7,
                                       passes=a.passes, repeats=a.repeats,
                                       log=log)
        prob.close()
    dev = "MI355X (gfx950)"
    cache = load_cache(a.out)
    cache[a.workload] = dict(best)
    cache[a.workload + "_record"] = {"device": dev, "seconds": t,
                                     "trials": trials,
                                     "tuning_seconds": time.time() - t0}
    save_cache(a.out, cache)
    log("best", best, "%.3f ms" % (1e3 * t))


if __name__ == "__main__":
    main()
