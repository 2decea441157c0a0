"""Learned warm start of the ionization iteration.

The plain method (testDensitySubGrid_omp.cpp:2274-2434) starts every cell at
x_n = 1e-6 (DensitySubGrid.hpp:549) and alternates a photon transport with the
ionization-equilibrium update until the neutral fraction stops changing.  The
warm start replaces the starting x_n only; the plain iteration then runs from
there and still decides convergence (mean |x_n change| per cell < tol).

The shipped network (input mode 1, start "passes") reads log10 n_H and the
logit of x_n after PASSES cheap iterations of PASS_PACKETS packets on the
full grid and adds a correction to that logit.  The other start ("coarse")
reads the logit of a coarse solution instead: the same medium averaged onto
a grid COARSE times coarser per axis and converged there by the plain
iteration (COARSE_PACKETS packets, about 1/64 of a full iteration's work per
iteration), interpolated back to the full grid on the device.  It places
the fronts of the whole box, but the clumps it averages away leave errors
the full iteration removes slowly: on the 256^3 clumpy workload it starts
10-12 iterations from convergence (the cheap passes: 7), 8-9 with a network
trained on it.  Input mode 0, the first design, reads one
cheap transport through the x_n = 1e-6 medium instead (log10 n_H and log10 of
the unattenuated photoionization rate jH) and predicts the logit itself; it
cannot see the shadows that decide where the fronts are.

Training data are fields the method itself converges: clumpy log-normal media
at the 256^3 clumpy workload's cell size, cut into 64^3 crops (target: the
logit of the converged x_n, or in mode 1 its difference from the cheap
iterate's).
"""
import json
import time

import numpy as np

from . import (ENGINE_TILES, PREC_FP32, Context, WarmStartNet)
from . import workloads as W

FIRST_PASS_PACKETS = 1_000_000
# the trained network's inputs come from PASSES cheap transports of
# PASS_PACKETS packets (the same packets each time, the ionization update in
# between): the J of the last one carries the shadowing of the ionized
# regions, which a single pass through the x_n = 1e-6 medium cannot
PASSES = 5
PASS_PACKETS = 10_000_000
# coarse start: grid factor, packets per coarse iteration (about the full
# run's crossings per cell: 1e8 packets on 256^3 cross a cell ~680 times),
# convergence bar of the coarse plain iteration
COARSE = 4
COARSE_PACKETS = 3_000_000
COARSE_TOL = 1e-3
COARSE_MAX_IT = 40
XN_CLAMP = 1e-7          # targets use logit of x_n clamped to [1e-7, 1 - 1/150]
XN_HIGH = 1.0 - 1.0 / 150.0   # above it every cell is opaque (tau ~ 76 per cell)
LOGIT_SCALE = 8.0        # the network outputs logit / 8
FEATURE_MEAN = (8.0, -7.0)
FEATURE_SCALE = (2.0, 0.5)
# dilations of the four 3x3x3 layers: a receptive field of 31^3 cells
DILATION = (1, 2, 4, 8)
# held-out problems (field, source and luminosity seeds) of the 256^3 clumpy
# kind: neither the training problems (seeds 1000+, 2000+, 3000+) nor the
# 256^3 clumpy workload (2, 3, 4)
HELDOUT = ((999, 998, 997), (995, 994, 993), (990, 989, 988), (985, 984, 983))


def logit(x):
    x = np.clip(np.asarray(x, dtype=np.float64), XN_CLAMP, XN_HIGH)
    return np.log(x / (1.0 - x))


def clumpy_problem(ncell=256, box_side_pc=10.0, nsrc=8, field_seed=2,
                   src_seed=3, lum_seed=4, sigma_ln=1.0, lmin=1e48, lmax=1e49):
    """Log-normal medium (P(k) ~ k^-11/3, mean 1e8 m^-3) with point sources
    at uniform-random positions and log-uniform luminosities.  The defaults
    are the 256^3 clumpy workload; returns (box, n_H lattice, pos, L)."""
    half = 0.5 * box_side_pc * W.PC
    box = np.array([-half, -half, -half, 2 * half, 2 * half, 2 * half])
    n = W.lognormal_density(ncell, sigma_ln=sigma_ln, seed=field_seed)
    pos, L = W.random_sources(nsrc, box, seed=src_seed, lum_seed=lum_seed,
                              lmin=lmin, lmax=lmax)
    return box, n, pos, L


class Problem:
    """A FP32-traversal context on the tiles engine holding one medium, with
    its density kept for resets."""

    def __init__(self, box, n_lattice, pos, L, nsub=4, alpha=W.ALPHA_B):
        ncell = list(n_lattice.shape)
        self.ncell = ncell
        self.box, self.pos, self.Lsrc = box, pos, L
        self.coarse = None
        self.nsub = [nsub] * 3
        self.ctx = Context(box, ncell, self.nsub, PREC_FP32)
        self.ctx.set_engine(ENGINE_TILES)
        # convergence runs do not read the per-launch kernel timing
        self.ctx.set_kernel_timing(False)
        self.n_ref = W.global_to_subgrid_major(n_lattice, self.nsub)
        self.xn0 = np.full(self.n_ref.size, W.XN0)
        self.ctx.set_sources(pos, L)
        self.L = self.ctx.total_luminosity()
        self.alpha = alpha
        self.reset()

    def reset(self):
        """x_n = 1e-6 everywhere, J = 0."""
        self.ctx.set_fields(self.n_ref, self.xn0)
        self.ctx.reset_intensity()
        self.ctx.synchronize()

    def coarse_context(self, factor=COARSE, npackets=COARSE_PACKETS):
        """The coarse grid of the coarse start (built once, with its
        transport buffers for npackets, then kept)."""
        if self.coarse is None or self.coarse[0] != factor or self.coarse[2] < npackets:
            if self.coarse is not None:
                self.coarse[1].close()
            nc = [c // factor for c in self.ncell]
            ns = [max(1, min(4, c // 16)) for c in nc]
            ctx = Context(self.box, nc, ns, PREC_FP32)
            ctx.set_engine(ENGINE_TILES)
            ctx.set_kernel_timing(False)
            ctx.set_sources(self.pos, self.Lsrc)
            ctx.transport(npackets, 0, 0)   # builds the engine's buffers
            ctx.reset_intensity()
            self.coarse = (factor, ctx, npackets)
        return self.coarse[1]

    def coarse_start(self, seed, factor=COARSE, npackets=COARSE_PACKETS,
                     tol=COARSE_TOL, max_it=COARSE_MAX_IT):
        """x_n of this context from a coarse solve, all on the device: the
        density averaged onto the coarse grid, the plain iteration there with
        common random numbers until mean |dx_n| < tol (from the second
        iteration on), the coarse logit interpolated back.  Returns the
        coarse iterations."""
        c = self.coarse_context(factor, npackets)
        self.ctx.restrict_density_to(c)
        it = 0
        for it in range(1, max_it + 1):
            c.transport(npackets, seed, 1)
            l1 = c.update_ionization(self.L, npackets, self.alpha, form=1)
            if it >= MIN_IT and l1 < tol:
                break
        self.ctx.prolong_neutral_fraction_from(c)
        return it

    def xn_lattice(self):
        _, x = self.ctx.get_fields()
        return W.subgrid_major_to_global(x, self.ncell, self.nsub)

    def first_pass(self, seed, npackets=FIRST_PASS_PACKETS, passes=1,
                   update_last=False):
        """`passes` cheap transports of npackets packets (the same packets
        each time), with the ionization update between them; the J of the
        last one is left in place (the input of mode 0), or with
        update_last the last update is done too (x_n is the input of
        mode 1)."""
        for k in range(passes):
            s = self.ctx.transport(npackets, seed, 0)
            if k + 1 < passes or update_last:
                self.ctx.update_ionization(self.L, npackets, self.alpha, form=1)
        return s

    def iterate(self, npackets, seed, tol=1e-3, max_it=40, it0=1,
                min_it=1, frozen=False):
        """Plain iterations (transport + FP64 update) until the mean
        |x_n change| < tol; returns the per-iteration L1 changes and the wall
        time (host-synchronised).  frozen: every iteration draws the same
        packets (common random numbers), so the L1 change measures the
        iteration's convergence rather than Monte Carlo noise."""
        l1s = []
        t0 = time.perf_counter()
        for it in range(max_it):
            self.ctx.transport(npackets, seed, it0 if frozen else it0 + it)
            l1 = self.ctx.update_ionization(self.L, npackets, self.alpha, form=1)
            l1s.append(l1)
            if it + 1 >= min_it and l1 < tol:
                break
        self.ctx.synchronize()
        return {"l1": l1s, "iterations": len(l1s),
                "seconds": time.perf_counter() - t0,
                "converged": l1s[-1] < tol}

    def close(self):
        if self.coarse is not None:
            self.coarse[1].close()
            self.coarse = None
        self.ctx.close()


# Convergence protocol of a run: the plain iteration with common random
# numbers (every iteration of the run draws the same packets; with fresh
# packets the mean |dx_n| of successive iterations has a Monte Carlo floor of
# 2.3e-3 at 1e8 packets on 256^3, above the 1e-3 criterion), converged when
# the mean |dx_n| < tol at the second iteration or later (the first iteration
# from x_n = 1e-6 changes little only because the medium is transparent).
MIN_IT = 2


def converge_plain(prob, npackets, seed, tol=1e-3, max_it=40, frozen=True):
    prob.reset()
    r = prob.iterate(npackets, seed, tol=tol, max_it=max_it, frozen=frozen,
                     min_it=MIN_IT)
    r["total_seconds"] = r["seconds"]
    return r


def net_start(prob, net, seed):
    """The network's inputs in the context (its coarse solve or its cheap
    passes); returns the packet count the inputs were made with."""
    if getattr(net, "start", "passes") == "coarse":
        prob.coarse_start(seed, net.coarse_factor, net.coarse_packets,
                          net.coarse_tol)
        return net.coarse_packets
    first_packets = getattr(net, "pass_packets", FIRST_PASS_PACKETS)
    prob.first_pass(seed, first_packets, getattr(net, "passes", 1),
                    update_last=net.input_mode() == 1)
    return first_packets


def converge_warm(prob, net, npackets, seed, tol=1e-3, max_it=40, frozen=True):
    """The network's start (coarse solve or cheap passes) + inference + plain
    iterations; the wall time includes the start and the inference."""
    prob.reset()
    if getattr(net, "start", "passes") == "coarse":
        # set-up, like the full grid's context: not part of the timed start
        prob.coarse_context(net.coarse_factor, net.coarse_packets)
    t0 = time.perf_counter()
    npk = net_start(prob, net, seed ^ 0x5A5A5A5A)
    net.warm_start(prob.ctx, prob.L, npk)
    t1 = time.perf_counter()
    r = prob.iterate(npackets, seed, tol=tol, max_it=max_it, frozen=frozen,
                     min_it=MIN_IT)
    r["warm_seconds"] = t1 - t0
    r["total_seconds"] = r["seconds"] + (t1 - t0)
    return r


def crops(a, size):
    """All non-overlapping size^3 crops of a lattice (leading axes)."""
    n0, n1, n2 = a.shape[:3]
    out = []
    for i in range(0, n0 - size + 1, size):
        for j in range(0, n1 - size + 1, size):
            for k in range(0, n2 - size + 1, size):
                out.append(a[i:i + size, j:j + size, k:k + size])
    return out


def make_training_set(seeds, npackets, crop=64, ncell=256, box_side_pc=10.0,
                      nsrc=8, tol=2e-4, max_it=40, extra_it=0, log=None,
                      feature_net=None, passes=1,
                      first_packets=FIRST_PASS_PACKETS, mode=0, start="passes",
                      coarse=(COARSE, COARSE_PACKETS, COARSE_TOL)):
    """Converge one clumpy problem per seed with the plain method and cut
    (features, target logits) crops.  Returns float32 arrays
    (S, crop, crop, crop, 2) and (S, crop, crop, crop) plus per-problem
    records."""
    net = feature_net or WarmStartNet(seed=1)
    net.set_input_mode(mode)
    feats, tgts, recs = [], [], []
    for s in seeds:
        box, n, pos, L = clumpy_problem(ncell, box_side_pc, nsrc,
                                        field_seed=1000 + s, src_seed=2000 + s,
                                        lum_seed=3000 + s)
        prob = Problem(box, n, pos, L)
        if start == "coarse":
            prob.coarse_start(7000 + s, *coarse)
            f = net.context_features(prob.ctx, prob.L, coarse[1])
        else:
            prob.first_pass(seed=7000 + s, npackets=first_packets, passes=passes,
                            update_last=mode == 1)
            f = net.context_features(prob.ctx, prob.L, first_packets)
        prob.reset()
        r = prob.iterate(npackets, seed=5000 + s, tol=tol, max_it=max_it,
                         frozen=True, min_it=MIN_IT)
        if extra_it:
            prob.iterate(npackets, seed=6000 + s, tol=0.0, max_it=extra_it,
                         it0=100)
        t = logit(prob.xn_lattice()).astype(np.float32)
        if mode == 1:
            t = t - f[..., 1]   # the correction to the cheap iterate's logit
        feats += crops(f, crop)
        tgts += crops(t, crop)
        recs.append({"seed": int(s), "iterations": r["iterations"],
                     "l1": r["l1"][-1]})
        if log:
            log("training problem %d: %d iterations, last L1 %.2e" %
                (s, r["iterations"], r["l1"][-1]))
        prob.close()
    if feature_net is None:
        net.close()
    return (np.ascontiguousarray(np.stack(feats)),
            np.ascontiguousarray(np.stack(tgts)), recs)


# offset of the last layer's bias in the master parameter vector (layers of
# W[27][NOUT][32] + b[NOUT], NOUT = 32, 32, 32, 16)
LAST_LAYER = 3 * (27 * 32 * 32 + 32)
LAST_BIAS = LAST_LAYER + 27 * 16 * 32


def shift_logits(net, delta):
    """Add delta to every predicted logit (through the last bias)."""
    p = net.params()
    p[LAST_BIAS] += delta / LOGIT_SCALE
    net.set_params(p)


def new_net(seed=1, mean=FEATURE_MEAN, scale=FEATURE_SCALE, dilation=DILATION):
    net = WarmStartNet(seed=seed)
    net.set_normalisation(mean, scale, LOGIT_SCALE)
    net.set_dilation(dilation)
    net.norm = (np.asarray(mean, np.float32), np.asarray(scale, np.float32))
    return net


def save_net(net, path, meta=None, passes=1, pass_packets=FIRST_PASS_PACKETS):
    mean, scale = getattr(net, "norm", (FEATURE_MEAN, FEATURE_SCALE))
    np.savez(path, params=net.params(), feature_mean=np.asarray(mean),
             passes=passes, pass_packets=pass_packets,
             start=np.array(getattr(net, "start", "passes")),
             coarse_factor=getattr(net, "coarse_factor", COARSE),
             coarse_packets=getattr(net, "coarse_packets", COARSE_PACKETS),
             coarse_tol=getattr(net, "coarse_tol", COARSE_TOL),
             input_mode=net.input_mode(),
             feature_scale=np.asarray(scale), logit_scale=LOGIT_SCALE,
             dilation=net.dilation(),
             start_rule=np.array(net.start_rule()),
             meta=np.array(json.dumps(meta or {})))


def load_net(path):
    d = np.load(path, allow_pickle=False)
    net = WarmStartNet(seed=1)
    net.set_normalisation(d["feature_mean"], d["feature_scale"],
                          float(d["logit_scale"]))
    if "dilation" in d:
        net.set_dilation(d["dilation"])
    if "start_rule" in d:
        net.set_start_rule(float(d["start_rule"][0]), float(d["start_rule"][1]))
    if "input_mode" in d:
        net.set_input_mode(int(d["input_mode"]))
    net.passes = int(d["passes"]) if "passes" in d else 1
    net.pass_packets = int(d["pass_packets"]) if "pass_packets" in d else FIRST_PASS_PACKETS
    net.start = str(d["start"]) if "start" in d else "passes"
    net.coarse_factor = int(d["coarse_factor"]) if "coarse_factor" in d else COARSE
    net.coarse_packets = int(d["coarse_packets"]) if "coarse_packets" in d else COARSE_PACKETS
    net.coarse_tol = float(d["coarse_tol"]) if "coarse_tol" in d else COARSE_TOL
    net.set_params(d["params"])
    return net
