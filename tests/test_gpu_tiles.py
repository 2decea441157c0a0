"""GPU checks of the tiles engine (FP32 DDA inside LDS tiles, packets binned
per tile between rounds) against the all-FP64 oracle.  The FP32 path is an
approximation of the reference arithmetic, so the bars are statistical: the
fate of each packet (absorbed / escaped) agrees for >= 99.9 % of packets,
per-packet crossing counts within 2 (ties at cell edges are walked one axis at
a time), and the per-cell path sums within FP32 rounding (median relative
error < 1e-5, total < 1e-5)."""
import numpy as np
import pytest

import oracle_lib as O
import parallelcmacionize_amd as cmi
from parallelcmacionize_amd import workloads as W

pytestmark = pytest.mark.gpu


def _field(seed, ncell, nsub, xlo=-6, xhi=-3):
    rng = np.random.default_rng(seed)
    nt = int(np.prod(ncell))
    n = 1e8 * np.exp(rng.normal(0, 1, nt))
    xn = 10 ** rng.uniform(xlo, xhi, nt)
    return n, xn


@pytest.mark.parametrize("tile", [(16, 16, 16), (8, 8, 8), (4, 8, 16),
                                  (1, 16, 16), (16, 2, 4), (2, 16, 1)])
def test_tiles_list_matches_fp64_oracle(oracle, tile):
    ncell, nsub = [32, 32, 32], [2, 2, 2]
    box = W.box_10pc()
    n, xn = _field(3, ncell, nsub)
    pos, L = W.random_sources(3, box, seed=5)
    N = 50000
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP64) as c64:
        c64.set_sources(pos, L)
        pk, ss = c64.emit(N, 8, 0)
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES, tile=tile)
        ctx.set_fields(n, xn)
        p, st, cr, hp, stats = ctx.transport_packets(pk, ss)
        J, _ = ctx.get_fields()
    og = O.OracleGrid(oracle, box, ncell, nsub)
    og.n[:] = n
    og.xn[:] = xn
    og.J[:] = 0
    q = pk.copy()
    sto, cro, hpo = og.transport(q, ss, nthreads=4)
    assert stats["errors"] == 0
    assert stats["absorbed"] + stats["escaped"] == N
    assert (st == sto).mean() > 0.999
    same = st == sto
    assert np.abs(cr[same].astype(int) - cro[same].astype(int)).max() <= 2
    assert abs(J.sum() / og.J.sum() - 1) < 1e-5
    big = og.J > 1e-3 * og.J.max()
    rel = np.abs(J[big] - og.J[big]) / og.J[big]
    assert np.median(rel) < 1e-5 and rel.max() < 1e-2
    # final positions: the same point up to FP32 rounding (cells of 0.3 pc)
    d = np.linalg.norm(p["pos"][same] - q["pos"][same], axis=1)
    assert np.median(d) < 1e-4 * (box[3] / 32)


@pytest.mark.parametrize("thick", [False, True])
def test_tiles_tied_walks(oracle, thick):
    """The tied walks of the oracle's pinning (axes, face and body diagonals
    through exact cell faces, edges and corners) through the FP32 tiles
    engine: it steps tied axes one at a time, so its crossing counts may be
    larger, but the packets' fates and the deposited path lengths agree."""
    from test_oracle import degenerate_case
    box, ncell, n, xn, pk = degenerate_case(oracle, 0, 0, thick)
    pk["weight"] = 1.0          # the FP32 engine folds sigma into kappa, weight 1
    ss = np.zeros(len(pk), np.uint32)
    og = O.OracleGrid(oracle, box, ncell, [1, 1, 1])
    og.n[:] = n
    og.xn[:] = xn
    og.J[:] = 0
    q = pk.copy()
    sto, cro, hpo = og.transport(q, ss, nthreads=4)
    with cmi.Context(box, ncell, [1, 1, 1], cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES, tile=(8, 8, 8))
        ctx.set_fields(n, xn)
        p, st, cr, hp, stats = ctx.transport_packets(pk, ss)
        J, _ = ctx.get_fields()
    assert stats["errors"] == 0
    assert (st == sto).mean() >= 0.999
    same = st == sto
    assert (cr[same].astype(int) >= cro[same].astype(int) - 2).all()
    assert abs(J.sum() / og.J.sum() - 1) < 1e-5
    big = og.J > 1e-3 * og.J.max()
    rel = np.abs(J[big] - og.J[big]) / og.J[big]
    assert np.median(rel) < 1e-5


def test_tiles_transport_matches_walk(oracle):
    """Emission + transport through the tiles engine and through the walk
    engine (both FP32) deposit the same intensity field."""
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    box = W.box_10pc()
    n, xn = _field(4, ncell, nsub, -5, -2)
    pos, L = W.random_sources(8, box, seed=3)
    N = 400000
    out = {}
    for eng in (cmi.ENGINE_WALK, cmi.ENGINE_TILES):
        with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
            ctx.set_engine(eng)
            ctx.set_fields(n, xn)
            ctx.set_sources(pos, L)
            s = ctx.transport(N, 3, 1)
            J, _ = ctx.get_fields()
            out[eng] = (s, J)
    sw, Jw = out[cmi.ENGINE_WALK]
    st, Jt = out[cmi.ENGINE_TILES]
    assert st["errors"] == 0 and st["absorbed"] + st["escaped"] == N
    assert abs(st["absorbed"] - sw["absorbed"]) <= max(5, 1e-4 * N)
    assert abs(st["crossings"] / sw["crossings"] - 1) < 2e-3
    assert abs(Jt.sum() / Jw.sum() - 1) < 1e-5
    big = Jw > 1e-3 * Jw.max()
    rel = np.abs(Jt[big] - Jw[big]) / Jw[big]
    assert np.median(rel) < 1e-5


@pytest.mark.parametrize("tile", [(64, 8, 4), (4, 64, 16), (2, 4, 64), (64, 16, 8),
                                  (2, 64, 64)])
def test_tiles_wide_tiles_match_walk(tile):
    """Tiles with a 64-cell axis (the widest step counter of the packed walk
    state) on a grid of 64^3 subgrids: the same intensity field and fates
    as the walk engine, within the FP32 bars.  (64, 16, 8) and (2, 64, 64)
    have the largest allowed 8192 cells; the latter, with a 4096-cell
    plane as bias, fills all 32 bits of the walk state."""
    ncell, nsub = [128, 128, 128], [2, 2, 2]
    box = W.box_10pc()
    n, xn = _field(9, ncell, nsub, -5, -2)
    pos, L = W.random_sources(4, box, seed=13)
    N = 300000
    out = {}
    for eng in (cmi.ENGINE_WALK, cmi.ENGINE_TILES):
        with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
            if eng == cmi.ENGINE_TILES:
                ctx.set_engine(eng, tile=tile)
            else:
                ctx.set_engine(eng)
            ctx.set_fields(n, xn)
            ctx.set_sources(pos, L)
            s = ctx.transport(N, 5, 1)
            J, _ = ctx.get_fields()
            out[eng] = (s, J)
    sw, Jw = out[cmi.ENGINE_WALK]
    st, Jt = out[cmi.ENGINE_TILES]
    assert st["errors"] == 0 and st["absorbed"] + st["escaped"] == N
    assert abs(st["absorbed"] - sw["absorbed"]) <= max(5, 1e-4 * N)
    assert abs(st["crossings"] / sw["crossings"] - 1) < 2e-3
    assert abs(Jt.sum() / Jw.sum() - 1) < 1e-5
    big = Jw > 1e-3 * Jw.max()
    rel = np.abs(Jt[big] - Jw[big]) / Jw[big]
    assert np.median(rel) < 1e-5


def test_tiles_round_count_and_timing_switch():
    """The host checks the packets left only every few rounds; the reported
    round count is still the number of rounds that had packets (= the most
    tile visits of any packet, when no round is walked out in one launch),
    and switching the per-launch timing off changes nothing but the timing
    record."""
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    box = W.box_10pc()
    n, xn = _field(6, ncell, nsub, -6, -3)
    pos, L = W.random_sources(4, box, seed=9)
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP64) as c64:
        c64.set_sources(pos, L)
        pk, ss = c64.emit(20000, 2, 0)
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES, tile=(8, 8, 8), finish_packets=-1)
        ctx.set_fields(n, xn)
        _, st, cr, hp, stats = ctx.transport_packets(pk, ss)
        assert stats["errors"] == 0
        assert stats["rounds"] == int(hp.max())
        ctx.set_sources(pos, L)
        ctx.reset_intensity()
        ctx.kernel_timing(reset=True)
        s_on = ctx.transport(300000, 4, 2)
        J_on, _ = ctx.get_fields()
        avg, calls = ctx.kernel_timing(reset=True)
        assert calls >= s_on["rounds"] and avg > 0
        ctx.set_kernel_timing(False)
        ctx.reset_intensity()
        s_off = ctx.transport(300000, 4, 2)
        J_off, _ = ctx.get_fields()
        assert ctx.kernel_timing()[1] == 0
        for k in ("crossings", "absorbed", "escaped", "rounds", "hops"):
            assert s_on[k] == s_off[k], k
        assert np.allclose(J_on, J_off, rtol=1e-12, atol=0)


def test_tiles_walkout_matches_visits():
    """When the packets left (counted at the start and every 8 rounds) are at
    most finish_packets, one walk-out launch walks them all to the end, with
    the opacities read from memory.  Each tile leg repeats a visit's
    arithmetic, so every packet's crossings, tile visits, fate and final
    position agree exactly with walking it visit by visit, and J agrees to
    the order of its FP64 sums.  Thresholds: never, the default and 1e9 (the
    whole transport walked out at once), and N - 1 (walked out at the first
    check, after 8 rounds)."""
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    box = W.box_10pc()
    n, xn = _field(7, ncell, nsub, -6, -3)
    pos, L = W.random_sources(4, box, seed=11)
    N, NE = 200000, 300000
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP64) as c64:
        c64.set_sources(pos, L)
        pk, ss = c64.emit(N, 2, 0)
    out = {}
    for fp in (-1, 0, N - 1, 10 ** 9):
        with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
            ctx.set_engine(cmi.ENGINE_TILES, tile=(8, 8, 8), finish_packets=fp)
            ctx.set_fields(n, xn)
            p, st, cr, hp, stats = ctx.transport_packets(pk, ss)
            J, _ = ctx.get_fields()
            ctx.set_sources(pos, L)
            ctx.reset_intensity()
            s2 = ctx.transport(NE, 4, 2)
            J2, _ = ctx.get_fields()
        out[fp] = (p, st, cr, hp, stats, J, s2, J2)
    p0, st0, cr0, hp0, s0, J0, e0, Je0 = out[-1]
    assert s0["errors"] == 0 and s0["rounds"] == int(hp0.max())
    assert s0["walkout_crossings"] == 0 and e0["walkout_crossings"] == 0
    # the first check comes after 8 rounds: this medium needs more
    assert s0["rounds"] > 9 and e0["rounds"] > 9
    for fp in (0, N - 1, 10 ** 9):
        p, st, cr, hp, stats, J, s2, J2 = out[fp]
        assert stats["errors"] == 0, fp
        assert np.array_equal(st, st0), fp
        assert np.array_equal(cr, cr0), fp
        assert np.array_equal(hp, hp0), fp
        assert np.array_equal(p["pos"], p0["pos"]), fp
        for k in ("crossings", "absorbed", "escaped", "hops"):
            assert stats[k] == s0[k], (fp, k)
            assert s2[k] == e0[k], (fp, k)
        assert np.allclose(J, J0, rtol=1e-10, atol=0), fp
        assert np.allclose(J2, Je0, rtol=1e-10, atol=0), fp
    for fp in (0, 10 ** 9):  # under the threshold from the start
        stats, s2 = out[fp][4], out[fp][6]
        assert stats["rounds"] == 1 and s2["rounds"] == 1
        assert stats["walkout_crossings"] == stats["crossings"]
        assert stats["walkout_hops"] == stats["hops"]
    stats, s2 = out[N - 1][4], out[N - 1][6]
    for s in (stats, s2):  # rounds 0-7, then round 8 walked out
        assert s["rounds"] == 9
        assert 0 < s["walkout_crossings"] < s["crossings"]
        assert 0 < s["walkout_hops"] < s["hops"]


def test_tiles_larger_than_8192_cells_refused():
    """A tile of more than 8192 cells does not fit the packed walk state (its
    byte offset and step counters): the engine refuses it and keeps the
    previous tile."""
    box = W.box_10pc()
    with cmi.Context(box, [64] * 3, [1] * 3, cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES, tile=(16, 16, 16))
        for tile in ((32, 32, 16), (64, 16, 16), (64, 64, 4)):
            with pytest.raises(cmi.CMIError):
                ctx.set_engine(cmi.ENGINE_TILES, tile=tile)
        ctx.set_engine(cmi.ENGINE_TILES, tile=(32, 16, 16))  # 8192: allowed


def test_tiles_16_kernel_matches_any_shape_kernel():
    """Emission-driven transports of 16^3 tiles at 512 threads run a process
    kernel specialised to that shape (a static LDS block, the walk layout as
    constants).  The same emitted packets through it, through the kernel for
    any shape (256 threads) and entirely through the walk-out give the same
    crossings, tile visits and fates, and J to the order of its FP64 sums;
    an explicit packet list (the kernel for any shape at both thread counts)
    agrees packet by packet."""
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    box = W.box_10pc()
    n, xn = _field(9, ncell, nsub, -6, -3)
    pos, L = W.random_sources(4, box, seed=13)
    N, NE = 200000, 300000
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP64) as c64:
        c64.set_sources(pos, L)
        pk, ss = c64.emit(N, 5, 0)
    out = {}
    for name, threads, fp in (("t16", 512, -1), ("any", 256, -1), ("walkout", 512, 10 ** 9)):
        with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
            ctx.set_engine(cmi.ENGINE_TILES, tile=(16, 16, 16), block_threads=threads,
                           finish_packets=fp)
            ctx.set_fields(n, xn)
            p, st, cr, hp, stats = ctx.transport_packets(pk, ss)
            J, _ = ctx.get_fields()
            ctx.set_sources(pos, L)
            ctx.reset_intensity()
            s2 = ctx.transport(NE, 6, 1)
            J2, _ = ctx.get_fields()
        out[name] = (p, st, cr, hp, stats, J, s2, J2)
    p0, st0, cr0, hp0, s0, J0, e0, Je0 = out["t16"]
    assert s0["errors"] == 0 and s0["absorbed"] + s0["escaped"] == N
    assert s0["walkout_crossings"] == 0 and s0["rounds"] > 1
    for name in ("any", "walkout"):
        p, st, cr, hp, stats, J, s2, J2 = out[name]
        assert np.array_equal(st, st0), name
        assert np.array_equal(cr, cr0), name
        assert np.array_equal(hp, hp0), name
        assert np.array_equal(p["pos"], p0["pos"]), name
        for k in ("crossings", "absorbed", "escaped", "hops"):
            assert stats[k] == s0[k], (name, k)
            assert s2[k] == e0[k], (name, k)
        assert np.allclose(J, J0, rtol=1e-10, atol=0), name
        assert np.allclose(J2, Je0, rtol=1e-10, atol=0), name
    assert out["walkout"][4]["walkout_crossings"] == s0["crossings"]


def test_tiles_results_independent_of_chunking():
    """How a round's packets are cut into workgroup visits (the chunk, which
    the plan also shrinks for rounds with few packets) changes only the order
    of the FP64 sums: every packet walks the same cells, so the crossings,
    fates, tile visits and rounds agree exactly and J to FP64 rounding.
    2e6 packets: the first rounds hold more than a quarter of the resident
    workgroups' worth of full chunks (192 x 8192 on 256 CUs), so the 8192 and
    16384 chunks are compared uncut there and cut in the later rounds (no
    round is walked out in one launch here, so every round is cut)."""
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    box = W.box_10pc()
    n, xn = _field(8, ncell, nsub, -6, -3)
    pos, L = W.random_sources(6, box, seed=12)
    out = []
    N = 2_000_000
    for chunk in (256, 8192, 16384):
        with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
            ctx.set_engine(cmi.ENGINE_TILES, chunk_min=chunk, finish_packets=-1)
            ctx.set_fields(n, xn)
            ctx.set_sources(pos, L)
            s = ctx.transport(N, 5, 1)
            J, _ = ctx.get_fields()
            assert s["errors"] == 0 and s["absorbed"] + s["escaped"] == N
            out.append((s, J))
    sa, Ja = out[0]
    for sb, Jb in out[1:]:
        for k in ("crossings", "absorbed", "escaped", "rounds", "hops"):
            assert sa[k] == sb[k], k
        assert np.allclose(Ja, Jb, rtol=1e-9, atol=0)


@pytest.mark.parametrize("prec", [cmi.PREC_FP32, cmi.PREC_FP64])
def test_reset_intensity_clears_nonfinite(prec):
    """A context's J comes from freshly allocated device memory, which may
    hold NaN or Inf bit patterns: the reset stores zeros rather than scaling
    by zero (which would keep them and make every later update of those cells
    fully neutral)."""
    ncell = [32, 32, 32]
    with cmi.Context(W.box_10pc(), ncell, [2, 2, 2], prec) as ctx:
        J = np.full(32 ** 3, np.nan)
        J[::3] = np.inf
        J[1::3] = -np.inf
        ctx.set_intensity(J)
        ctx.reset_intensity()
        Jg, _ = ctx.get_fields()
        assert np.all(Jg == 0.0)


def test_tiles_stromgren_radius():
    box = W.box_10pc()
    ncell, nsub = [64, 64, 64], [4, 4, 4]
    N = 2_000_000
    with cmi.Context(box, ncell, nsub, cmi.PREC_FP32) as ctx:
        ctx.set_engine(cmi.ENGINE_TILES)
        ctx.set_sources([[0.0, 0.0, 0.0]], [W.Q_STROMGREN])
        for it in range(16):
            s = ctx.transport(N, 1, it)
            assert s["errors"] == 0
            assert s["absorbed"] + s["escaped"] == N
            l1 = ctx.update_ionization(W.Q_STROMGREN, N, W.ALPHA_B, form=1)
            print("it %d: absorbed %d crossings %d rounds %d l1 %.3e"
                  % (it, s["absorbed"], s["crossings"], s["rounds"], l1))
        _, x = ctx.get_fields()
    R = W.ionized_radius(x, box, ncell)
    Rs = W.stromgren_radius(W.Q_STROMGREN, W.N_H, W.ALPHA_B)
    print("tiles R/Rs", R / Rs, "rounds", s["rounds"])
    assert abs(R / Rs - 1) < 0.02
