"""Reemission (do_reemission, testDensitySubGrid_omp.cpp:675-702; the INSIDE
buffer, 1188-1191; the reemit task, 1348-) on the GPU against the oracle.

- FP64 walk engine, explicit packet lists: bit for bit against the oracle's
  orc_transport_list_reemit (status, crossings, visits, final positions,
  directions and optical depths; J up to the order of the atomic sums).  The
  draws are exact in IEEE arithmetic on both sides.
- FP32 walk engine: the same draws, FP32 geometry (fates and counts against
  the oracle within the FP32 bars).
- The emitted transport draws what the list transport of the same emitted
  packets draws.
- FP32 tiles engine (its own FP32 draws keyed by the absorption point):
  statistically against the oracle on the same emitted packets, and against
  the closed forms of an opaque medium (geometric reemission count, mean
  deposit per leg).
- With probability p each source photon is absorbed 1/(1-p) times on average,
  so the Stromgren radius grows to R = (3 Q / ((1 - p) 4 pi n^2 alpha))^(1/3)
  (scripts/plot_neutral_fraction_profile.py:68)."""
import numpy as np
import pytest

import oracle_lib as O
import parallelcmacionize_amd as cmi
from parallelcmacionize_amd import workloads as W

pytestmark = pytest.mark.gpu

SOURCES = ([[0.3e16, -1.1e16, 0.7e16], [-2.0e16, 1.0e16, 0.0]], [3e49, 1e49])


def _clumpy(ncell, seed, xlo=-4, xhi=-1):
    rng = np.random.default_rng(seed)
    n = 1e8 * np.exp(rng.normal(0.0, 1.0, ncell ** 3))
    xn = 10 ** rng.uniform(xlo, xhi, ncell ** 3)
    return n, xn


def _oracle_grid(box, ncell, nsub, n, xn):
    og = O.OracleGrid(O.load_oracle(), box, [ncell] * 3, nsub)
    og.n[:] = n
    og.xn[:] = xn
    og.J[:] = 0
    return og


def test_fp64_walk_list_bit_exact_against_oracle():
    box, ncell, nsub = W.box_10pc(), 32, [2, 2, 2]
    n, xn = _clumpy(ncell, 21)
    N, p, seed, it = 20000, 0.6, 11, 2
    with cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP64) as ctx:
        ctx.set_fields(n, xn)
        ctx.set_sources(*SOURCES)
        pk, ss = ctx.emit(N, 5, 0)
        ctx.set_reemission(p, seed=seed, iteration=it)
        ctx.reset_intensity()
        q, st, cr, hp, stats = ctx.transport_packets(pk, ss)
        J, _ = ctx.get_fields()
    og = _oracle_grid(box, ncell, nsub, n, xn)
    qo = pk.copy()
    sto, cro, hpo, reo = og.transport_reemit(qo, ss, p, seed, it, nthreads=8)
    print("reemissions", int(reo.sum()), "max per packet", int(reo.max()))
    assert reo.sum() > N and stats["reemitted"] == int(reo.sum())
    assert np.array_equal(st, sto)
    assert np.array_equal(cr, cro)
    assert np.array_equal(hp, hpo)
    for f in ("pos", "dir", "tau"):
        assert np.array_equal(q[f], qo[f]), f
    assert stats["absorbed"] == int((sto == 0).sum())
    assert stats["escaped"] == int((sto == 1).sum())
    assert stats["crossings"] == int(cro.sum())
    scale = np.abs(og.J).max()
    assert np.abs(J - og.J).max() <= 1e-12 * scale


def test_fp32_walk_list_against_oracle():
    box, ncell, nsub = W.box_10pc(), 32, [2, 2, 2]
    n, xn = _clumpy(ncell, 22)
    N, p = 20000, 0.5
    with cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP64) as emitter, \
            cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP32) as ctx:
        emitter.set_sources(*SOURCES)
        pk, ss = emitter.emit(N, 6, 0)
        ctx.set_fields(n, xn)
        ctx.set_reemission(p, seed=3, iteration=1)
        ctx.reset_intensity()
        q, st, cr, hp, stats = ctx.transport_packets(pk, ss)
        J, _ = ctx.get_fields()
    og = _oracle_grid(box, ncell, nsub, n, xn)
    sto, cro, hpo, reo = og.transport_reemit(pk.copy(), ss, p, 3, 1, nthreads=8)
    # the draws are keyed by packet and absorption, not by position: FP32
    # geometry changes fates only for packets that graze the box boundary
    assert (st == sto).mean() >= 0.999
    assert abs(stats["reemitted"] / reo.sum() - 1) < 2e-3
    assert abs(J.sum() / og.J.sum() - 1) < 1e-3
    both = st == sto
    assert np.abs(cr.astype(np.int64) - cro)[both].max() <= 2 + 2 * reo[both].max()


def test_emitted_walk_transport_draws_like_the_list_transport():
    box, ncell, nsub = W.box_10pc(), 32, [2, 2, 2]
    n, xn = _clumpy(ncell, 23)
    N, p, seed, it = 30000, 0.7, 8, 4
    with cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP64) as ctx:
        ctx.set_fields(n, xn)
        ctx.set_sources(*SOURCES)
        ctx.set_reemission(p, seed=seed, iteration=it)
        ctx.reset_intensity()
        s1 = ctx.transport(N, seed, it)
        J1, _ = ctx.get_fields()
        pk, ss = ctx.emit(N, seed, it)
        ctx.reset_intensity()
        _, _, _, _, s2 = ctx.transport_packets(pk, ss)
        J2, _ = ctx.get_fields()
    assert s1["reemitted"] > N
    for k in ("crossings", "absorbed", "escaped", "reemitted", "errors"):
        assert s1[k] == s2[k], k
    assert np.abs(J1 - J2).max() <= 1e-12 * np.abs(J1).max()


@pytest.mark.parametrize("p", [0.3, 0.75])
def test_tiles_reemission_closed_forms_in_opaque_medium(p):
    """Opaque medium (sigma n Delta ~ 600 per cell, x_n = 1): every packet is
    absorbed in its source cell, reemitted r times with P(r) = (1 - p) p^r,
    and every leg deposits tau / n (E[tau] = 1)."""
    box, ncell = W.box_10pc(), 32
    N = 1_000_000
    with cmi.Context(box, [ncell] * 3, [2, 2, 2], cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_fields(np.full(ncell ** 3, 1e8), np.ones(ncell ** 3))
        ctx.set_sources(*SOURCES)
        ctx.set_reemission(p)
        ctx.reset_intensity()
        s = ctx.transport(N, 12, 0)
        J, _ = ctx.get_fields()
    assert s["absorbed"] == N and s["escaped"] == 0 and s["errors"] == 0
    mean, var = p / (1 - p), p / (1 - p) ** 2
    assert abs(s["reemitted"] / N - mean) < 5 * np.sqrt(var / N)
    legs = N + s["reemitted"]
    # sum of tau over the legs: E = legs, sd = sqrt(legs) given the legs
    assert abs(J.sum() * 1e8 / legs - 1) < 5 / np.sqrt(legs) + 1e-5


def test_tiles_reemission_against_oracle():
    box, ncell, nsub = W.box_10pc(), 32, [2, 2, 2]
    n, xn = _clumpy(ncell, 24, -3, -1)
    N, p, seed = 200000, 0.5, 14
    with cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP64) as emitter, \
            cmi.Context(box, [ncell] * 3, nsub, cmi.PREC_FP32) as ctx:
        emitter.set_sources(*SOURCES)
        pk, ss = emitter.emit(N, seed, 0)
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_fields(n, xn)
        ctx.set_sources(*SOURCES)
        ctx.set_reemission(p)
        ctx.reset_intensity()
        s = ctx.transport(N, seed, 0)  # emits the same packets as emitter.emit
        J, _ = ctx.get_fields()
    og = _oracle_grid(box, ncell, nsub, n, xn)
    sto, cro, hpo, reo = og.transport_reemit(pk.copy(), ss, p, 99, 0, nthreads=16)
    # independent draws: compare as two samples of the same distributions
    fa = (sto == 0).mean()
    sd_abs = np.sqrt(2 * N * fa * (1 - fa))
    print("absorbed", s["absorbed"], int((sto == 0).sum()), "reemitted",
          s["reemitted"], int(reo.sum()))
    assert abs(s["absorbed"] - (sto == 0).sum()) < 5 * sd_abs + 1
    sd_re = np.sqrt(2 * N * reo.var())
    assert abs(s["reemitted"] - int(reo.sum())) < 5 * sd_re
    assert abs(J.sum() / og.J.sum() - 1) < 0.02


def test_reemission_contract():
    box = W.box_10pc()
    n, xn = _clumpy(32, 25)
    with cmi.Context(box, [32] * 3, [2, 2, 2], cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_fields(n, xn)
        ctx.set_reemission(0.3)
        pk = np.zeros(4, cmi.PHOTON_DTYPE)
        pk["dir"][:, 0] = 1.0
        pk["tau"] = 1.0
        pk["sigma"] = 6.3e-22
        pk["weight"] = 1.0
        with pytest.raises(cmi.CMIError) as e:  # tiles: explicit lists refuse
            ctx.transport_packets(pk, np.zeros(4, np.uint32))
        assert e.value.code == -1
        with pytest.raises(cmi.CMIError):
            ctx.set_reemission(1.0)
    # per-subgrid walks are DensitySubGrid::interact: absorbed packets come
    # back INSIDE whatever the probability (reemitting is the caller's step)
    with cmi.Context(box, [16] * 3, [1, 1, 1], cmi.PREC_FP64) as ctx:
        ctx.set_fields(np.full(16 ** 3, 1e8), np.full(16 ** 3, 1e-2))
        rng = np.random.default_rng(4)
        pk = np.zeros(500, cmi.PHOTON_DTYPE)
        pk["pos"] = rng.uniform(-4e16, 4e16, (500, 3))
        d = rng.normal(size=(500, 3))
        pk["dir"] = d / np.linalg.norm(d, axis=1)[:, None]
        pk["tau"] = rng.exponential(1.0, 500)
        pk["sigma"] = 6.3e-22
        pk["weight"] = 1.0
        a = ctx.interact_batch(0, 0, pk)
        ctx.set_reemission(0.9, seed=1)
        b = ctx.interact_batch(0, 0, pk)
        assert (a[1] == 0).any()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("engine,p", [(cmi.ENGINE_TILES, 0.0), (cmi.ENGINE_TILES, 0.5),
                                      (cmi.ENGINE_WALK, 0.5)])
def test_reemission_stromgren_radius(engine, p):
    box = W.box_10pc()
    Rs0 = 3.0 * W.PC
    Q = 4 / 3 * np.pi * Rs0 ** 3 * W.N_H ** 2 * W.ALPHA_B
    N = 1_000_000
    with cmi.Context(box, [64] * 3, [4, 4, 4], cmi.PREC_FP32) as ctx:
        ctx.set_engine(engine)
        ctx.set_fields(np.full(64 ** 3, W.N_H), np.full(64 ** 3, W.XN0))
        ctx.set_sources([[0.0, 0.0, 0.0]], [Q])
        ctx.set_reemission(p)
        for it in range(14):
            s = ctx.transport(N, 3, it)
            assert s["errors"] == 0
            assert s["absorbed"] + s["escaped"] == N
            ctx.update_ionization(Q, N, W.ALPHA_B, form=1)
        x = ctx.get_fields()[1]
    R = W.ionized_radius(x, box, [64] * 3)
    Rp = Rs0 / (1.0 - p) ** (1.0 / 3.0)
    print("engine", engine, "p", p, "R/Rp", R / Rp)
    assert abs(R / Rp - 1) < 0.02
