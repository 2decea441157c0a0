"""The batched workload: 64 independent problems of uneven size in one problem
set (cmi_batch_*), each iteration of all of them one replay of a captured
hipGraph.

Problem q (seeded): edge n_q drawn from {32, 48, 64, 96, 128} cells, a
log-normal medium with sigma_ln uniform in [0.5, 1.5] (mean 1e8 m^-3,
P(k) ~ k^-11/3), 1-16 sources at uniform-random positions, packets
clamp(2 * n_q^3, 2e4, 4e6).  Cells have the 128^3-in-10-pc size (0.078 pc),
so a problem's box side is n_q / 128 * 10 pc, and source luminosities are
log-uniform in [1e48, 1e49] s^-1 scaled by (n_q / 128)^3, which keeps the
ionized volume fraction comparable across sizes.
"""
import numpy as np

from . import Batch
from . import workloads as W

SIZES = (32, 48, 64, 96, 128)
CELL = 10.0 * W.PC / 128


def problem_set_spec(nprob=64, seed=5, packets_per_cell=2.0):
    rng = np.random.default_rng(seed)
    sizes = rng.choice(SIZES, nprob)
    sig = rng.uniform(0.5, 1.5, nprob)
    nsrc = rng.integers(1, 17, nprob)
    spec = []
    for q in range(nprob):
        n = int(sizes[q])
        side = n * CELL
        box = np.array([0.0, 0.0, 0.0, side, side, side])
        scale = (n / 128.0) ** 3
        spec.append({
            "n": n, "box": box, "sigma_ln": float(sig[q]), "nsrc": int(nsrc[q]),
            "field_seed": 10000 * seed + q, "src_seed": 20000 * seed + q,
            "lum_seed": 30000 * seed + q, "lmin": 1e48 * scale,
            "lmax": 1e49 * scale,
            "packets": int(np.clip(packets_per_cell * n ** 3, 2e4, 4e6)),
        })
    return spec


def build_fields(spec):
    """Densities and sources of every problem (host arrays)."""
    out = []
    for s in spec:
        n = W.lognormal_density(s["n"], sigma_ln=s["sigma_ln"],
                                seed=s["field_seed"])
        pos, L = W.random_sources(s["nsrc"], s["box"], seed=s["src_seed"],
                                  lum_seed=s["lum_seed"], lmin=s["lmin"],
                                  lmax=s["lmax"])
        out.append((n, pos, L))
    return out


class ProblemSet:
    """A cmi.Batch holding the workload, with its initial state kept so that
    timed iterations can restart from it."""

    def __init__(self, spec, fields, tile=16, block_threads=0, refill_batch=0,
                 chunk=0, alpha=W.ALPHA_B):
        self.spec = spec
        self.fields = fields
        self.alpha = alpha
        self.npackets = np.array([s["packets"] for s in spec], np.uint64)
        self.batch = Batch([s["box"] for s in spec],
                           [[s["n"]] * 3 for s in spec], tile=tile,
                           block_threads=block_threads,
                           refill_batch=refill_batch, chunk=chunk)
        self.xn_state = None
        for q, (n, pos, L) in enumerate(fields):
            self.batch.set_sources(q, pos, L)
        self.restore()

    def snapshot(self):
        """Keep the current x_n of every problem as the restart state."""
        self.xn_state = [self.batch.get_fields(q)[1] for q in range(len(self.spec))]

    def restore(self):
        for q, (n, _, _) in enumerate(self.fields):
            x = (np.full(n.size, W.XN0) if self.xn_state is None
                 else self.xn_state[q])
            self.batch.set_fields(q, n, x)

    def retune(self, cfg):
        self.batch.set_tuning(block_threads=cfg.get("block_threads", 0),
                              refill_batch=cfg.get("refill_batch", 0),
                              chunk=cfg.get("chunk", 0))

    def iteration(self, npackets=None, seed=1, iteration=0, alpha=None,
                  form=1):
        return self.batch.iteration(self.npackets if npackets is None else npackets,
                                    seed, iteration,
                                    self.alpha if alpha is None else alpha, form)

    def close(self):
        self.batch.close()


def run(ps, iterations, seed=1, it0=0):
    """Iterate the whole set; per iteration the set's device time and every
    problem's completion time.  Returns the summary of the bench line."""
    secs, per_prob, cross, packets = [], [], 0, 0
    for it in range(iterations):
        r = ps.iteration(seed=seed, iteration=it0 + it)
        if r["errors"]:
            raise RuntimeError("problem set: walk errors")
        secs.append(r["seconds"])
        per_prob.append(r["problem_seconds"].copy())
        cross += int(r["crossings"])
        packets += int(ps.npackets.sum())
    pp = np.concatenate(per_prob)
    total = float(np.sum(secs))
    return {
        "iterations": iterations,
        "problems": len(ps.spec),
        "cells": int(sum(s["n"] ** 3 for s in ps.spec)),
        "packets_per_iteration": int(ps.npackets.sum()),
        "seconds_per_iteration": total / iterations,
        "crossings_per_s": cross / total,
        "packets_per_s": packets / total,
        "problem_seconds_p50": float(np.percentile(pp, 50)),
        "problem_seconds_p99": float(np.percentile(pp, 99)),
        "problem_seconds_max": float(pp.max()),
    }
